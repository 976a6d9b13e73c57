// Constraint elimination by QL decomposition (pede's default constraint mode, icelim = 1,
// src/mpmod.f90:104), host-side set-up.
//
// The constraint matrix C (ncon x nvar, rows over variable parameters) is decomposed as
// C^T = Q [0; L] with Q = H_m ... H_1 a product of Householder reflections H_k = I - 2 v_k v_k^T
// (src/mpqldec.f90:219-393, qldecb).  Constraint k's reflector acts on the parameter range of its
// constraint block plus the position kn = nvar - ncon + k, so the last ncon parameters are the
// eliminated ones.  The solve then works on the first nfit = nvar - ncon components of Q^T x.
//
// The constraints are laid out as pede does (src/pede.f90:840-1003 grpcon, 1077-1371 prpcon,
// 1423-1460 feasma): sorted by their first and last parameter, split into groups that share no
// parameter, groups merged into blocks of overlapping ranges; each constraint's row is stored
// over [first parameter of its group's chain, last parameter of its block].  Ties of the sort are
// kept in input order (pede's quicksort leaves their order to the pivoting; any order gives the
// same constrained solution).
//
// For the device the reflectors are turned into the compact WY form H_1 H_2 ... H_m = I - V T V^T
// (T upper triangular, T_jj = 2, T_{1:j-1,j} = -2 T_{1:j-1,1:j-1} V_{:,1:j-1}^T v_j), so that
// Q x = x - V T^T V^T x and Q^T x = x - V T V^T x take three parallel kernels instead of a chain
// of ncon dependent reflections.
#include "mpgpu_internal.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace mpgpu {

void QlElim::build(int32_t nvar_, const std::vector<std::vector<std::pair<int32_t, double>>> &con_tot,
                   const std::vector<int32_t> &var_index)
{
    nvar = nvar_;
    ncon = (int32_t)con_tot.size();
    nfit = nvar - ncon;
    if (ncon <= 0) throw MpError(30, "constraint elimination without constraints");
    if (nfit <= 0) throw MpError(27, "Aborted, more constraints than variable parameters");
    if (ncon > 2048) throw MpError(30, "constraint elimination supports at most 2048 constraints");
    const int m = ncon;

    // ---- grpcon: ranges over all labels (total index), sort, disjoint groups ----
    std::vector<int32_t> f_all(m), l_all(m);
    for (int j = 0; j < m; j++) {
        int32_t lo = INT32_MAX, hi = -1;
        for (auto &p : con_tot[j]) { lo = std::min(lo, p.first); hi = std::max(hi, p.first); }
        f_all[j] = lo;
        l_all[j] = hi;
    }
    std::vector<int> srt(m);
    std::iota(srt.begin(), srt.end(), 0);
    std::stable_sort(srt.begin(), srt.end(), [&](int a, int b) {
        return f_all[a] != f_all[b] ? f_all[a] < f_all[b] : l_all[a] < l_all[b];
    });
    // connected components over shared labels, numbered by first appearance in sorted order
    std::vector<int> parent(m);
    std::iota(parent.begin(), parent.end(), 0);
    auto root = [&](int i) { while (parent[i] != i) i = parent[i] = parent[parent[i]]; return i; };
    {
        std::vector<std::pair<int32_t, int>> lab;
        for (int j = 0; j < m; j++)
            for (auto &p : con_tot[j]) lab.push_back({p.first, j});
        std::sort(lab.begin(), lab.end());
        for (size_t q = 1; q < lab.size(); q++)
            if (lab[q].first == lab[q - 1].first) {
                int a = root(lab[q].second), b = root(lab[q - 1].second);
                if (a != b) parent[std::max(a, b)] = std::min(a, b);
            }
    }
    std::vector<int> gid_of_root(m, -1), grp(m);
    int ngrp = 0;
    for (int q = 0; q < m; q++) {
        int r = root(srt[q]);
        if (gid_of_root[r] < 0) gid_of_root[r] = ngrp++;
        grp[srt[q]] = gid_of_root[r];
    }
    // order: by group, then by sorted position
    std::vector<int> pos(m);
    for (int q = 0; q < m; q++) pos[srt[q]] = q;
    order.resize(m);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return grp[a] != grp[b] ? grp[a] < grp[b] : pos[a] < pos[b]; });

    // ---- prpcon: ranges over variable parameters (variable index, 1-based) ----
    std::vector<int32_t> vf(m), vl(m);
    for (int q = 0; q < m; q++) {
        int j = order[q];
        int32_t lo = INT32_MAX, hi = -1;
        for (auto &p : con_tot[j]) {
            int32_t v = var_index[p.first];
            if (v > 0) { lo = std::min(lo, v); hi = std::max(hi, v); }
        }
        if (hi < 0) throw MpError(27, "Aborted, empty constraint (no variable parameter)");
        vf[q] = lo;
        vl[q] = hi;
    }
    // group ranges, in order
    std::vector<int> gstart;     // first sorted constraint of each group
    std::vector<int32_t> gfirst_v, glast_v;
    for (int q = 0; q < m; q++) {
        if (q == 0 || grp[order[q]] != grp[order[q - 1]]) {
            gstart.push_back(q);
            gfirst_v.push_back(vf[q]);
            glast_v.push_back(vl[q]);
        } else {
            gfirst_v.back() = std::min(gfirst_v.back(), vf[q]);
            glast_v.back() = std::max(glast_v.back(), vl[q]);
        }
    }
    const int ng = (int)gstart.size();
    gstart.push_back(m);
    rcon.assign(4 * (size_t)m, 0);
    bcon.clear();
    int32_t ilast = 0;
    int isblck = 0;
    for (int g = 0; g < ng; g++) {
        ilast = std::max(ilast, glast_v[g]);
        const int32_t next_first = g + 1 < ng ? gfirst_v[g + 1] : INT32_MAX;
        if (next_first > ilast) {
            bcon.push_back(gstart[isblck] + 1);      // first constraint (1-based)
            bcon.push_back(gfirst_v[isblck]);        // first parameter of the block
            bcon.push_back(ilast);                   // last parameter of the block
            int32_t jfrst = gfirst_v[g];
            for (int i = g; i >= isblck; i--) {
                jfrst = std::min(jfrst, gfirst_v[i]);
                for (int q = gstart[i]; q < gstart[i + 1]; q++) {
                    rcon[4 * q + 0] = gfirst_v[i];
                    rcon[4 * q + 1] = glast_v[i];
                    rcon[4 * q + 2] = jfrst;
                    rcon[4 * q + 3] = ilast;
                }
            }
            isblck = g + 1;
        }
    }
    nblk = (int32_t)(bcon.size() / 3);
    bcon.push_back(m + 1);
    bcon.push_back(0);
    bcon.push_back(0);

    // ---- feasma: the block-compressed matrix (column per constraint over [rcon3, rcon4]) ----
    ioffRow.assign(m + 1, 0);
    ioffPar.assign(m, 0);
    for (int q = 0; q < m; q++) {
        ioffPar[q] = rcon[4 * q + 2] - 1;
        ioffRow[q + 1] = ioffRow[q] + (rcon[4 * q + 3] - ioffPar[q]);
    }
    matV.assign(ioffRow[m], 0.0);
    for (int q = 0; q < m; q++)
        for (auto &p : con_tot[order[q]]) {
            int32_t v = var_index[p.first];
            if (v > 0) matV[ioffRow[q] + (v - 1 - ioffPar[q])] += p.second;
        }
    decompose();
    wy();
}

// qldecb for one parameter block covering all variable parameters (src/mpqldec.f90:219-393);
// arrays are 0-based here, the index arithmetic follows the reference's 1-based one
void QlElim::decompose()
{
    const int m = ncon;
    vecVk.assign(m, 0.0);
    matL.assign((size_t)m * m, 0.0);
    irng.assign(2 * (size_t)m, 0);
    std::vector<double> vecN(nvar + 1, 0.0);   // 1-based
    auto V = [&](int64_t i) -> double & { return matV[i - 1]; };
    auto Lm = [&](int64_t i) -> double & { return matL[i - 1]; };
    auto N = [&](int64_t i) -> double & { return vecN[i]; };
    auto IOR = [&](int k) { return ioffRow[k - 1]; };
    auto IOP = [&](int k) { return ioffPar[k - 1]; };
    auto IRN = [&](int a, int k) -> int32_t & { return irng[2 * (size_t)(k - 1) + a - 1]; };
    auto BC = [&](int a, int b) { return bcon[3 * (size_t)(b - 1) + a - 1]; };
    for (int k = 1; k <= m; k++) { IRN(1, k) = rcon[4 * (k - 1)]; IRN(2, k) = rcon[4 * (k - 1) + 1]; }
    const int iblast = nvar, icoff = 0, iclast = m, iboff = 0;
    // matL from the part of each column at or below its diagonal position kn: parameter p goes to
    // row p - (nvar - ncon), the rows the Householder updates below use.  (The reference starts
    // the copy at row iclast - ncol, the same row only when the constraint block's storage ends
    // at the last parameter; L enters only the solve for the constraint residuals.)
    for (int k = 1; k <= m; k++) {
        const int kn = iblast + k - iclast;
        const int64_t ioff1 = IOR(k);
        const int npb = (int)(IOR(k + 1) - ioff1);
        const int ncol = IOP(k) + npb - kn;
        if (ncol >= 0) {
            const int64_t ioff3 = (int64_t)(k - 1) * m;
            for (int t = 0; t <= ncol; t++) Lm(ioff3 + k - icoff + t) = V(ioff1 + npb - ncol + t);
        }
    }
    int ibcon = nblk;
    int k1 = BC(1, ibcon);
    for (int k = iclast; k >= icoff + 1; k--) {
        const int kn = iblast + k - iclast;
        if (k < k1) { ibcon--; k1 = BC(1, ibcon); }
        const int ipoff = IOP(k);
        int ifirst = ipoff + 1;
        if (ifirst > kn) continue;
        const int64_t ioff1 = IOR(k);
        const int npb = (int)(IOR(k + 1) - ioff1);
        const int iplast = IOP(k) + npb;
        int ilast = std::min(iplast, kn);
        int ncol = ilast - ipoff;
        const int in0 = iblast + k1 - iclast;
        for (int t = in0; t <= kn; t++) N(t) = 0.0;
        for (int t = 0; t < ncol; t++) N(ifirst + t) = V(ioff1 + 1 + t);
        double nrm = 0.0;
        for (int t = ifirst; t <= ilast; t++) nrm += N(t) * N(t);
        nrm = std::sqrt(nrm);
        if (nrm == 0.0) continue;
        if (N(kn) >= 0.0) N(kn) += nrm; else N(kn) -= nrm;
        if (ilast < kn) {
            double s2 = 0.0;
            for (int t = ifirst; t <= ilast; t++) s2 += N(t) * N(t);
            nrm = std::sqrt(s2 + N(kn) * N(kn));
            for (int t = ifirst; t <= ilast; t++) N(t) /= nrm;
            N(kn) /= nrm;
        } else {
            double s2 = 0.0;
            for (int t = ifirst; t <= ilast; t++) s2 += N(t) * N(t);
            nrm = std::sqrt(s2);
            for (int t = ifirst; t <= ilast; t++) N(t) /= nrm;
        }
        int64_t ioff3 = (int64_t)(k1 - 2) * m;
        for (int i = k1; i <= k; i++) {
            ioff3 += m;
            if (IRN(1, k) > IRN(2, i)) continue;
            const int64_t ioff2 = IOR(i) + IOP(k) - IOP(i);
            double sp = 0.0;
            for (int t = 0; t < ncol; t++) sp += N(ifirst + t) * V(ioff2 + 1 + t);
            if (sp != 0.0) {
                for (int t = 0; t < ncol; t++) V(ioff2 + 1 + t) -= 2.0 * N(ifirst + t) * sp;
                const int in1 = iblast + i - iclast;
                for (int t = 0; t <= k - i; t++) Lm(ioff3 + i - icoff + t) -= 2.0 * N(in1 + t) * sp;
                IRN(1, i) = std::min(IRN(1, i), IRN(1, k));
                IRN(2, i) = std::max(IRN(2, i), IRN(2, k));
            }
        }
        vecVk[icoff + k - 1] = N(kn);
        ifirst = IRN(1, k);
        ilast = std::min(IRN(2, k), kn - 1);
        ncol = ilast - ifirst + 1;
        for (int t = 0; t < ncol; t++) V(ioff1 + 1 + t) = N(ifirst + t);
        for (int t = std::max(ncol, 0); t < npb; t++) V(ioff1 + 1 + t) = 0.0;
        IRN(1, k) = ifirst - iboff;
        IRN(2, k) = ilast - iboff;
    }
    // qlgete (src/mpqldec.f90:955-981): a vanishing diagonal element of L means dependent
    // constraints (src/pede.f90:1532-1535)
    double emin = matL[0], emax = matL[0];
    for (int i = 0; i < m; i++) {
        double d = matL[(size_t)i * m + i];
        if (std::fabs(emax) < std::fabs(d)) emax = d;
        if (std::fabs(emin) > std::fabs(d)) emin = d;
    }
    evmin = emin;
    evmax = emax;
    if (emin == 0.0) throw MpError(27, "Aborted, singular QL decomposition of constraints matrix");
}

// Reflectors as sparse rows (range + the element at kn), their transpose per parameter, the
// Gram matrix V^T V and the WY factor T.
void QlElim::wy()
{
    const int m = ncon;
    rf_ptr.assign(m + 1, 0);
    rf_lo.assign(m, 0);
    rf_kn.assign(m, 0);
    rf_vk.assign(m, 0.0);
    rf_val.clear();
    for (int k = 1; k <= m; k++) {
        const int32_t lo = irng[2 * (k - 1)], hi = irng[2 * (k - 1) + 1];
        rf_lo[k - 1] = lo - 1;
        rf_kn[k - 1] = nfit + k - 1;
        rf_vk[k - 1] = vecVk[k - 1];
        for (int32_t p = lo; p <= hi; p++) rf_val.push_back(matV[ioffRow[k - 1] + (p - lo)]);
        rf_ptr[k] = (int64_t)rf_val.size();
    }
    // transpose: per parameter the (reflector, value) pairs, exact zeros dropped
    std::vector<int64_t> cnt(nvar + 1, 0);
    auto each = [&](auto &&fn) {
        for (int k = 0; k < m; k++) {
            for (int64_t q = rf_ptr[k]; q < rf_ptr[k + 1]; q++) {
                double v = rf_val[q];
                if (v != 0.0) fn(rf_lo[k] + (int32_t)(q - rf_ptr[k]), k, v);
            }
            if (rf_vk[k] != 0.0) fn(rf_kn[k], k, rf_vk[k]);
        }
    };
    each([&](int32_t p, int, double) { cnt[p + 1]++; });
    for (int32_t p = 0; p < nvar; p++) cnt[p + 1] += cnt[p];
    tr_ptr = cnt;
    tr_k.assign(cnt[nvar], 0);
    tr_v.assign(cnt[nvar], 0.0);
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    each([&](int32_t p, int k, double v) { tr_k[fill[p]] = k; tr_v[fill[p]++] = v; });
    // Gram matrix of the reflectors
    std::vector<double> G((size_t)m * m, 0.0);
    for (int32_t p = 0; p < nvar; p++)
        for (int64_t a = tr_ptr[p]; a < tr_ptr[p + 1]; a++)
            for (int64_t b = tr_ptr[p]; b < tr_ptr[p + 1]; b++) G[(size_t)tr_k[a] * m + tr_k[b]] += tr_v[a] * tr_v[b];
    // products v_k^T v_k' as qlpssq keeps them (src/mpqldec.f90:826-848)
    vtvp.assign((size_t)m * m, 0.0);
    for (int a = 0; a < m; a++)
        for (int b = 0; b < m; b++)
            vtvp[(size_t)a * m + b] = a == b ? 1.0 : (std::fabs(G[(size_t)a * m + b]) > 16.0 * 2.220446049250313e-16
                                                          ? G[(size_t)a * m + b] : 0.0);
    // T (row-major, upper triangular)
    T.assign((size_t)m * m, 0.0);
    std::vector<double> g(m);
    for (int j = 0; j < m; j++) {
        for (int i = 0; i < j; i++) {
            double s = 0.0;
            for (int l = i; l < j; l++) s += T[(size_t)i * m + l] * G[(size_t)l * m + j];
            g[i] = s;
        }
        for (int i = 0; i < j; i++) T[(size_t)i * m + j] = -2.0 * g[i];
        T[(size_t)j * m + j] = 2.0;
    }
}

void QlElim::upload(hipStream_t s)
{
    d_rf_ptr.upload(rf_ptr, s);
    d_rf_lo.upload(rf_lo, s);
    d_rf_kn.upload(rf_kn, s);
    d_rf_vk.upload(rf_vk, s);
    d_rf_val.upload(rf_val.empty() ? std::vector<double>(1, 0.0) : rf_val, s);
    d_tr_ptr.upload(tr_ptr, s);
    d_tr_k.upload(tr_k.empty() ? std::vector<int32_t>(1, 0) : tr_k, s);
    d_tr_v.upload(tr_v.empty() ? std::vector<double>(1, 0.0) : tr_v, s);
    d_T.upload(T, s);
    d_w.alloc(ncon);
    d_u.alloc(ncon);
    dev.n = nvar;
    dev.m = ncon;
    dev.rf_ptr = d_rf_ptr.p;
    dev.rf_lo = d_rf_lo.p;
    dev.rf_kn = d_rf_kn.p;
    dev.rf_val = d_rf_val.p;
    dev.rf_vk = d_rf_vk.p;
    dev.tr_ptr = d_tr_ptr.p;
    dev.tr_k = d_tr_k.p;
    dev.tr_v = d_tr_v.p;
    dev.T = d_T.p;
    dev.w = d_w.p;
    dev.u = d_u.p;
}

// pede's band transform (qlpssq, src/mpqldec.f90:711-945) with t = .true. (reflectors from the
// last to the first): every step adds, for reflector k with A v_k and its current correction
// c_k = sum_q Cf[k][q] v_q, the band of 4 (v^T A v) v v^T - 2 (A v) v^T - 2 v (A v)^T once with
// the initial A v and once with the correction, v^T A v being the initial value the first time
// and the corrected one the second.  Summed: band of sum_k D_k v_k v_k^T - 2 Y_k v_k^T - 2 v_k Y_k^T
// with Y_k = A v_k + c_k and D_k = 4 (v^T A v + v^T A_k v).  The corrections follow the
// reference's recurrence (lines 878-935); X[c * m + j] = v_j^T A v_c.
void QlElim::band_coef(const std::vector<double> &X, std::vector<double> &Cf, std::vector<double> &D) const
{
    const int m = ncon;
    Cf.assign((size_t)m * m, 0.0);
    D.assign(m, 0.0);
    std::vector<int> lo(m, m - 1), hi(m, 0);   // non-zero range of each coefficient row (0-based)
    auto XA = [&](int j, int c) { return X[(size_t)c * m + j]; };   // v_j^T A v_c
    for (int jj = 0; jj < m; jj++) {
        const int k = m - 1 - jj;
        double dd = 0.0;
        for (int q = lo[k]; q <= hi[k]; q++) dd += Cf[(size_t)k * m + q] * vtvp[(size_t)k * m + q];
        const double vtAv = XA(k, k) + dd;
        D[k] = 4.0 * XA(k, k) + 4.0 * vtAv;
        for (int jj2 = jj + 1; jj2 < m; jj2++) {
            const int k2 = m - 1 - jj2;
            const double sv = vtvp[(size_t)k * m + k2];
            int m1 = lo[k2], m2 = hi[k2];
            double vta = XA(k, k2);
            if (m1 <= m2) {
                double d2 = 0.0;
                for (int q = m1; q <= m2; q++) d2 += Cf[(size_t)k2 * m + q] * vtvp[(size_t)k * m + q];
                vta = vta + d2;
            }
            m1 = std::min(m1, k);
            m2 = std::max(m2, k);
            Cf[(size_t)k2 * m + k] += 2.0 * (2.0 * vtAv * sv - vta);
            if (sv != 0.0) {
                m1 = std::min(m1, lo[k]);
                m2 = std::max(m2, hi[k]);
                for (int q = m1; q <= m2; q++) Cf[(size_t)k2 * m + q] -= 2.0 * Cf[(size_t)k * m + q] * sv;
            }
            lo[k2] = m1;
            hi[k2] = m2;
        }
    }
}

// qlbsub (src/mpqldec.f90:991-1014): y from L^T y = d by backward substitution
std::vector<double> QlElim::bsub(const std::vector<double> &d) const
{
    const int m = ncon;
    std::vector<double> y(m, 0.0);
    for (int k = m - 1; k >= 0; k--) {
        double s = d[k];
        for (int r = k + 1; r < m; r++) s -= matL[(size_t)k * m + r] * y[r];
        y[k] = s / matL[(size_t)k * m + k];
    }
    return y;
}

} // namespace mpgpu
