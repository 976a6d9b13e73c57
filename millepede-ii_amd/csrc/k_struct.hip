// Index maps and the sparsity structure of the matrix, built on the device (loop1 / loop2's
// label and pair bookkeeping, src/pede.f90:2311-2404, 5920-6062, 6600-6920, and the mpbits
// index, src/mpbits.f90:73-566, 1220-1350, in this library's group-block layout):
//   label table      a bitmap over [lmin, lmax] of every label in use; the rank of a label's bit
//                    is its total index (0-based), so labels ascending <-> indices ascending
//   record lists     one wavefront per record sorts the total indices of all its global words
//                    (bitonic sort in LDS); the runs give the record's sorted distinct list, the
//                    position of every global word in it, the entry counters and the group
//                    boundaries of pargrp
//   record stats     one wavefront per record: parameter groups (segments) of its variables,
//                    local-fit sizes, the size of its matrix update, and the group pairs it
//                    couples, marked in a bitmap over the lower triangle of group pairs
//   blocks           the rank of a pair's bit is its block index (rows ascending, columns
//                    ascending within a row); offsets by a scan of the block sizes; column lists
//                    and per-record block ids of every segment pair
//   incidences       per batch of records, (block, update-buffer offset) of every segment pair;
//                    a stable radix sort by block keeps processing order within each block, so the
//                    gather sums every block in a fixed order
#include "mpgpu_internal.h"
#include <hipcub/hipcub.hpp>
#include <climits>

namespace mpgpu {

#define SW 64
#define STRUCT_WAVES 4

__device__ __forceinline__ void wsync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ int lane_id() { return threadIdx.x & (SW - 1); }
template <typename T> __device__ __forceinline__ T wmax(T v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, SW));
    return v;
}
template <typename T> __device__ __forceinline__ T wmin(T v)
{
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, SW));
    return v;
}
template <typename T> __device__ __forceinline__ T wsum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, SW);
    return v;
}
__device__ __forceinline__ int wscan_incl(int v)
{
    const int lane = lane_id();
    for (int o = 1; o < SW; o <<= 1) {
        const int u = __shfl_up(v, o, SW);
        if (lane >= o) v += u;
    }
    return v;
}
// set a bit (read first: most bits are set many times over)
__device__ __forceinline__ void set_bit(uint32_t *bits, int64_t o)
{
    const uint32_t m = 1u << (o & 31);
    if ((bits[o >> 5] & m) == 0u) atomicOr(&bits[o >> 5], m);
}
__device__ __forceinline__ void set_flag(uint8_t *f, int64_t i)
{
    if (f[i] == 0) f[i] = 1;
}
__device__ __forceinline__ int64_t bit_rank(int64_t o, const uint32_t *bits, const int64_t *wrank)
{
    return wrank[o >> 5] + __popc(bits[o >> 5] & ((1u << (o & 31)) - 1u));
}
struct LabelMapD {
    int32_t lmin;
    const uint32_t *bits;
    const int64_t *wrank;
};
__device__ __forceinline__ int32_t tix_of(const LabelMapD &L, int32_t l)
{
    return (int32_t)bit_rank((int64_t)l - L.lmin, L.bits, L.wrank);
}
__device__ __forceinline__ int64_t pair_bit(int64_t a, int64_t b) { return a * (a + 1) / 2 + b; }
// pair index t -> (a, b), b <= a, in the order a = 0, 1, ...; b = 0 .. a
__device__ __forceinline__ void pair_of(int64_t t, int &a, int &b)
{
    int64_t x = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (x * (x + 1) / 2 > t) x--;
    while ((x + 1) * (x + 2) / 2 <= t) x++;
    a = (int)x;
    b = (int)(t - x * (x + 1) / 2);
}
static unsigned grid_for(int64_t n, int per)
{
    const int64_t g = (n + per - 1) / per;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 1 << 20));
}
static int key_bits(int64_t maxkey)
{
    int b = 1;
    while (b < 63 && (int64_t(1) << b) <= maxkey) b++;
    return b;
}

// ---------------------------------------------------------------------------------------------
// record decoding (decodeNextMeasurement, src/pede.f90:2428-2468)
// ---------------------------------------------------------------------------------------------
// One thread walks one record.  Words are 1-based within the record (word k at w0 + k - 1); a
// zero index word starts a measurement (its value) and another one its local block ends with the
// error; the global words run to the next zero index word.  A measurement word followed at once
// by a zero word with a negative value opens special data: -value words are skipped.  Returns the
// number of equations, or -1 for a record that breaks the format; FILL: stores each equation's
// measurement word and (local words | global words << 16).
template <typename T, bool FILL>
__device__ int64_t decode_record_dev(const int32_t *I, const T *V, int64_t w0, int64_t w1, int64_t *sl, uint32_t *nn)
{
    const int64_t nw = w1 - w0, end = nw;
    int64_t last = 1, ne = 0;
    for (;;) {
        if (last >= end) break;
        do {
            last++;
            if (last > nw) return -1;
        } while (I[w0 + last - 1] != 0);
        const int64_t sL = last;
        do {
            last++;
            if (last > nw) return -1;
        } while (I[w0 + last - 1] != 0);
        const int64_t sG = last;
        const double vg = (double)V[w0 + sG - 1];
        if (sL + 1 == sG && vg < 0.0) {
            last = last + (int64_t)llround(-vg);
            continue;
        }
        while (last < end && I[w0 + last] != 0) last++;
        const int64_t nl = sG - sL - 1, ng = last - sG;
        for (int64_t k = sL + 1; k < sG; k++)
            if (I[w0 + k - 1] < 1) return -1;      // local parameter indices are 1-based
        if (nl > 0xffff || ng > 0xffff) return -1;
        if (FILL) {
            sl[ne] = w0 + sL - 1;
            nn[ne] = (uint32_t)nl | ((uint32_t)ng << 16);
        }
        ne++;
    }
    return ne;
}
template <typename T>
__global__ void k_decode_count(int64_t nrec, const int64_t *rec_w, const int32_t *I, const T *V, int64_t *neq,
                               unsigned long long *bad)
{
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r < nrec) {
        const int64_t ne = decode_record_dev<T, false>(I, V, rec_w[r], rec_w[r + 1], nullptr, nullptr);
        neq[r] = ne < 0 ? 0 : ne;
        if (ne < 0) atomicMin(bad, (unsigned long long)r);
    } else if (r == nrec) {
        neq[r] = 0;
    }
}
template <typename T>
__global__ void k_decode_fill(int64_t nrec, const int64_t *rec_w, const int32_t *I, const T *V, const int64_t *rec_eq,
                              int64_t *eq_sl, uint32_t *eq_n)
{
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r < nrec) decode_record_dev<T, true>(I, V, rec_w[r], rec_w[r + 1], eq_sl + rec_eq[r], eq_n + rec_eq[r]);
}

// ---------------------------------------------------------------------------------------------
// label table
// ---------------------------------------------------------------------------------------------
// label range, and the equation counters of pede's first data loop (prepareRecords,
// src/mppar.f90:345-351): equations with global derivatives and the number of global words
__global__ void k_label_range(int64_t neq, const int64_t *eq_sl, const uint32_t *eq_n, const int32_t *I, int32_t *mm,
                              unsigned long long *cnt)
{
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    unsigned long long ne = 0, nd = 0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < neq; e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t n = eq_n[e];
        const int64_t sg = eq_sl[e] + (n & 0xffff) + 1;
        ne += (n >> 16) > 0;
        nd += n >> 16;
        for (int j = 1; j <= (int)(n >> 16); j++) {
            const int32_t l = I[sg + j];
            lo = min(lo, l);
            hi = max(hi, l);
        }
    }
    lo = wmin(lo);
    hi = wmax(hi);
    ne = wsum(ne);
    nd = wsum(nd);
    if (lane_id() == 0) {
        atomicMin(&mm[0], lo);
        atomicMax(&mm[1], hi);
        atomicAdd(&cnt[0], ne);
        atomicAdd(&cnt[1], nd);
    }
}
__global__ void k_label_mark(int64_t neq, const int64_t *eq_sl, const uint32_t *eq_n, const int32_t *I, int32_t lmin,
                             uint32_t *bits)
{
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < neq; e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t n = eq_n[e];
        const int64_t sg = eq_sl[e] + (n & 0xffff) + 1;
        for (int j = 1; j <= (int)(n >> 16); j++) set_bit(bits, (int64_t)I[sg + j] - lmin);
    }
}
__global__ void k_label_mark_list(int64_t n, const int32_t *lab, int32_t lmin, uint32_t *bits)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) set_bit(bits, (int64_t)lab[i] - lmin);
}
__global__ void k_word_pop(int64_t nw, const uint32_t *bits, int64_t *pc)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= nw; i += (int64_t)gridDim.x * blockDim.x)
        pc[i] = i < nw ? __popc(bits[i]) : 0;
}
__global__ void k_label_enum(int64_t nw, const uint32_t *bits, const int64_t *wrank, int32_t lmin, int32_t *labels)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nw; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t m = bits[i];
        int64_t k = wrank[i];
        while (m) {
            const int j = __ffs(m) - 1;
            m &= m - 1u;
            labels[k++] = (int32_t)((int64_t)lmin + 32 * i + j);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// record lists
// ---------------------------------------------------------------------------------------------
// number of global words of every record (m[nrec] = 0 for the scan)
__global__ void k_rec_m(int64_t nrec, const int64_t *rec_eq, const uint32_t *eq_n, int64_t *m, int32_t *mmax)
{
    int64_t s = 0;
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r < nrec) {
        for (int64_t e = rec_eq[r]; e < rec_eq[r + 1]; e++) s += eq_n[e] >> 16;
        m[r] = s;
    } else if (r == nrec) {
        m[r] = 0;
    }
    const int32_t mx = wmax((int32_t)min(s, (int64_t)INT32_MAX));
    if (lane_id() == 0) atomicMax(mmax, mx);
}

// One wavefront per record.  LDS per wavefront (cap >= the most global words of any record, a
// power of two): keys[cap] (u64), run labels U[cap] (i32), run flags unz[cap] (u8).
// Key of a global word: total index << 32 | word offset in the record << 1 | nonzero derivative.
template <typename T, bool PER_RECORD>
__global__ __launch_bounds__(SW *STRUCT_WAVES) void k_rec_labels(int64_t nrec, int cap, const int64_t *rec_eq,
                                                                const int64_t *rec_w, const int64_t *eq_sl,
                                                                const uint32_t *eq_n, const int32_t *I, const T *val,
                                                                LabelMapD L, const int64_t *mofs, int32_t *widx,
                                                                int32_t *tmp_gl, int64_t *ngl, unsigned long long *cnt,
                                                                uint8_t *bnd)
{
    extern __shared__ __align__(16) unsigned char dyn[];
    const int nw = blockDim.x / SW, wv = threadIdx.x / SW, lane = lane_id();
    unsigned char *base = dyn + (size_t)wv * 13 * (size_t)cap;
    uint64_t *keys = (uint64_t *)base;
    int32_t *U = (int32_t *)(base + 8 * (size_t)cap);
    uint8_t *unz = base + 12 * (size_t)cap;
    for (int64_t r = blockIdx.x * (int64_t)nw + wv; r < nrec; r += (int64_t)gridDim.x * nw) {
        const int64_t e0 = rec_eq[r], e1 = rec_eq[r + 1], w0 = rec_w[r];
        const int m = (int)(mofs[r + 1] - mofs[r]);
        int P = 1;
        while (P < m) P <<= 1;
        // gather: one lane per equation, its global words at the equation's running offset
        int off = 0;
        for (int64_t eb = e0; eb < e1; eb += SW) {
            const int64_t e = eb + lane;
            uint32_t n = 0;
            int64_t sg = 0;
            if (e < e1) {
                n = eq_n[e];
                sg = eq_sl[e] + (n & 0xffff) + 1;
            }
            const int ng = (int)(n >> 16);
            const int incl = wscan_incl(ng);
            const int at = off + incl - ng;
            for (int j = 1; j <= ng; j++) {
                const int64_t wd = sg + j;
                const uint32_t t = (uint32_t)tix_of(L, I[wd]);
                const uint64_t nz = val[wd] != T(0) ? 1ull : 0ull;
                keys[at + j - 1] = ((uint64_t)t << 32) | ((uint64_t)(wd - w0) << 1) | nz;
            }
            off += __shfl(incl, SW - 1, SW);
        }
        for (int i = m + lane; i < P; i += SW) keys[i] = ~0ull;
        for (int i = lane; i < m; i += SW) unz[i] = 0;
        wsync();
        // bitonic sort of keys[0, P)
        for (int k = 2; k <= P; k <<= 1)
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = lane; i < P; i += SW) {
                    const int l = i ^ j;
                    if (l > i) {
                        const uint64_t a = keys[i], b = keys[l];
                        if ((a > b) == ((i & k) == 0)) {
                            keys[i] = b;
                            keys[l] = a;
                        }
                    }
                }
                wsync();
            }
        // runs of equal total index: the record's distinct labels
        int nrun = 0;
        for (int i0 = 0; i0 < m; i0 += SW) {
            const int i = i0 + lane;
            uint64_t key = 0;
            bool st = false;
            if (i < m) {
                key = keys[i];
                st = i == 0 || (keys[i - 1] >> 32) != (key >> 32);
            }
            const unsigned long long sm = __ballot(st);
            const int k = nrun + __popcll(sm & ((2ull << lane) - 1ull)) - 1;
            if (i < m) {
                if (st) U[k] = (int32_t)(key >> 32);
                widx[w0 + (int64_t)((key >> 1) & 0x7fffffffull)] = k;
                if (key & 1ull) unz[k] = 1;
            }
            nrun += __popcll(sm);
        }
        wsync();
        if (lane == 0) ngl[r] = nrun;
        for (int k = lane; k < nrun; k += SW) tmp_gl[mofs[r] + k] = U[k];
        if (PER_RECORD) {
            // entries counted once per record; group boundaries of the record's sorted list of
            // labels with a nonzero derivative
            for (int k = lane; k < nrun; k += SW) {
                if (!unz[k]) continue;
                const int32_t x = U[k];
                atomicAdd(&cnt[x], 1ull);
                const bool m1 = k > 0 && unz[k - 1] && U[k - 1] == x - 1;
                const bool p1 = k + 1 < nrun && unz[k + 1] && U[k + 1] == x + 1;
                if (!m1) set_flag(bnd, x);
                if (!p1) set_flag(bnd, (int64_t)x + 1);
            }
        } else {
            // per equation: every nonzero derivative counts; boundaries of the equation's sorted
            // list with repeats kept: in front of x unless x-1 is in the list, behind x unless x+1
            // is, and on both sides of a repeated x
            int32_t *sc = (int32_t *)keys;
            for (int64_t e = e0; e < e1; e++) {
                const uint32_t n = eq_n[e];
                const int ng = (int)(n >> 16);
                const int64_t sg = eq_sl[e] + (n & 0xffff) + 1;
                wsync();
                for (int j = lane; j < ng; j += SW) {
                    const int64_t wd = sg + 1 + j;
                    sc[j] = val[wd] != T(0) ? tix_of(L, I[wd]) : -1;
                }
                wsync();
                for (int j = lane; j < ng; j += SW) {
                    const int32_t x = sc[j];
                    if (x < 0) continue;
                    atomicAdd(&cnt[x], 1ull);
                    bool dup = false, m1 = false, p1 = false;
                    for (int q = 0; q < ng; q++) {
                        const int32_t y = sc[q];
                        dup |= y == x && q != j;
                        m1 |= y >= 0 && y == x - 1;
                        p1 |= y == x + 1;
                    }
                    if (dup || !m1) set_flag(bnd, x);
                    if (dup || !p1) set_flag(bnd, (int64_t)x + 1);
                }
            }
        }
        wsync();
    }
}

__global__ void k_compact_gl(int64_t nrec, const int64_t *mofs, const int64_t *rec_gl, const int32_t *tmp, int32_t *gl)
{
    const int nw = blockDim.x / SW, wv = threadIdx.x / SW, lane = lane_id();
    for (int64_t r = blockIdx.x * (int64_t)nw + wv; r < nrec; r += (int64_t)gridDim.x * nw) {
        const int64_t n = rec_gl[r + 1] - rec_gl[r];
        for (int64_t k = lane; k < n; k += SW) gl[rec_gl[r] + k] = tmp[mofs[r] + k];
    }
}

// ---------------------------------------------------------------------------------------------
// record stats, group pairs
// ---------------------------------------------------------------------------------------------
// Segments of a record (runs of its variables in one parameter group; the record's list is
// sorted, and variables and groups ascend with it): group sg[s] and size sz[s]; returns their
// number, nv = number of variables.
__device__ int record_segments(int64_t g0, int64_t g1, const int32_t *gl_tix, const int32_t *var, const int32_t *group_of,
                               int32_t *sg, int32_t *sz, int &nv)
{
    const int lane = lane_id();
    int nseg = 0, lastg = -1, nlive = 0;
    for (int64_t b0 = g0; b0 < g1; b0 += SW) {
        const int64_t q = b0 + lane;
        const int v = q < g1 ? var[gl_tix[q]] : 0;
        const int g = v > 0 ? group_of[v] : -1;
        const unsigned long long live = __ballot(g >= 0);
        const unsigned long long below = live & ((1ull << lane) - 1ull);
        const int src = below ? 63 - __clzll(below) : lane;
        const int gsrc = __shfl(g, src, SW);
        const int prev = below ? gsrc : lastg;
        const bool st = g >= 0 && g != prev;
        const unsigned long long sm = __ballot(st);
        if (st) {
            const int s = nseg + __popcll(sm & ((1ull << lane) - 1ull));
            sg[s] = g;
            sz[s] = nlive + __popcll(below);    // rank of the segment's first variable
        }
        nseg += __popcll(sm);
        nlive += __popcll(live);
        const int glast = __shfl(g, live ? 63 - __clzll(live) : 0, SW);
        if (live) lastg = glast;
    }
    wsync();
    for (int s0 = 0; s0 < nseg; s0 += SW) {
        const int s = s0 + lane;
        int a = 0, b = 0;
        if (s < nseg) {
            a = sz[s];
            b = s + 1 < nseg ? sz[s + 1] : nlive;
        }
        wsync();
        if (s < nseg) sz[s] = b - a;
        wsync();
    }
    nv = nlive;
    return nseg;
}

// per record: st[0] variables, st[1] segments, st[2] largest local index, st[3] sum over
// equations of (local words x global words), st[4] most equations with a nonzero derivative of
// one record variable, st[5] first parameter group (sort key of the processing order; ngroup if
// none); cnt_r: values of its update (segment a holds sz[a] x (variables up to a's end));
// MARK: its group pairs into the bitmap.  LDS per wavefront: s_g, s_z, qc [cap] i32.
template <typename T, bool MARK>
__global__ __launch_bounds__(SW *STRUCT_WAVES) void k_rec_stats(int64_t nrec, int cap, int32_t ngroup,
                                                               const int64_t *rec_eq, const int64_t *eq_sl,
                                                               const uint32_t *eq_n, const int32_t *widx, const T *val,
                                                               const int64_t *rec_gl, const int32_t *gl_tix,
                                                               const int32_t *var, const int32_t *group_of, int32_t *st,
                                                               int64_t *cnt_r, uint32_t *pbits)
{
    extern __shared__ __align__(16) unsigned char dyn[];
    const int nw = blockDim.x / SW, wv = threadIdx.x / SW, lane = lane_id();
    int32_t *s_g = (int32_t *)(dyn + (size_t)wv * 12 * (size_t)cap);
    int32_t *s_z = s_g + cap;
    int32_t *qc = s_z + cap;
    for (int64_t r = blockIdx.x * (int64_t)nw + wv; r < nrec; r += (int64_t)gridDim.x * nw) {
        const int64_t g0 = rec_gl[r], g1 = rec_gl[r + 1], e0 = rec_eq[r], e1 = rec_eq[r + 1];
        int nv = 0;
        const int nseg = record_segments(g0, g1, gl_tix, var, group_of, s_g, s_z, nv);
        const int ngl = (int)(g1 - g0);
        for (int k = lane; k < ngl; k += SW) qc[k] = 0;
        wsync();
        int nalc = 0;
        int64_t sup = 0;
        for (int64_t eb = e0; eb < e1; eb += SW) {
            const int64_t e = eb + lane;
            if (e >= e1) continue;
            const uint32_t n = eq_n[e];
            const int nl = (int)(n & 0xffff), ng = (int)(n >> 16);
            const int64_t sl = eq_sl[e], sg = sl + nl + 1;
            for (int j = 1; j <= nl; j++) nalc = max(nalc, widx[sl + j]);
            sup += (int64_t)nl * ng;
            for (int j = 1; j <= ng; j++) {
                if (val[sg + j] == T(0)) continue;
                const int pos = widx[sg + j];
                if (var[gl_tix[g0 + pos]] > 0) atomicAdd(&qc[pos], 1);
            }
        }
        nalc = wmax(nalc);
        sup = wsum(sup);
        wsync();
        int qe = 0;
        for (int k = lane; k < ngl; k += SW) qe = max(qe, qc[k]);
        qe = wmax(qe);
        if (lane == 0) {
            // segment a's region: the columns of all variables up to its end, then the column of
            // its rows' vector values, padded to a multiple of four values (the regions keep the
            // 32-byte alignment that the columns alone give them)
            int64_t n = 0, upto = 0;
            for (int a = 0; a < nseg; a++) {
                upto += s_z[a];
                n += (int64_t)s_z[a] * upto + ((s_z[a] + 3) & ~3);
            }
            cnt_r[r] = n;
            st[r] = nv;
            st[nrec + r] = nseg;
            st[2 * nrec + r] = nalc;
            st[3 * nrec + r] = (int32_t)min(sup, (int64_t)INT32_MAX);
            st[4 * nrec + r] = qe;
            st[5 * nrec + r] = nseg > 0 ? s_g[0] : ngroup;
        }
        if (MARK) {
            const int64_t np = (int64_t)nseg * (nseg + 1) / 2;
            for (int64_t t = lane; t < np; t += SW) {
                int a, b;
                pair_of(t, a, b);
                set_bit(pbits, pair_bit(s_g[a], s_g[b]));
            }
        }
        wsync();
    }
}

// ---------------------------------------------------------------------------------------------
// blocks
// ---------------------------------------------------------------------------------------------
__global__ void k_set_diag(int32_t ng, uint32_t *bits)
{
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g < ng) set_bit(bits, pair_bit(g, g));
}
__global__ void k_set_bits(int64_t n, const int64_t *o, uint32_t *bits)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) set_bit(bits, o[i]);
}
__global__ void k_brow(int32_t ng, const uint32_t *bits, const int64_t *wrank, int64_t nw, int64_t *brow)
{
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g < ng) brow[g] = bit_rank(pair_bit(g, 0), bits, wrank);
    if (g == ng) brow[g] = wrank[nw];
}
// row and column group of every block, its size, and its column group as (first index, size)
__global__ void k_blocks(int64_t nw, const uint32_t *bits, const int64_t *wrank, const int32_t *gfirst, int32_t *brg,
                         int32_t *bcol, int64_t *bsz, int32_t *bcf)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nw; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t m = bits[i];
        int64_t k = wrank[i];
        while (m) {
            const int j = __ffs(m) - 1;
            m &= m - 1u;
            int a, b;
            pair_of(32 * i + j, a, b);
            const int32_t na = gfirst[a + 1] - gfirst[a], nb = gfirst[b + 1] - gfirst[b];
            brg[k] = a;
            bcol[k] = b;
            bsz[k] = (int64_t)na * nb;
            bcf[2 * k] = gfirst[b] - 1;
            bcf[2 * k + 1] = nb;
            k++;
        }
    }
}
__global__ void k_tail_zero(int64_t *p) { *p = 0; }
// column lists: sort key of every block (its column group; diagonal blocks last)
__global__ void k_ccol_keys(int64_t nb, int32_t ng, const int32_t *brg, const int32_t *bcol, uint32_t *key, int32_t *val)
{
    for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nb; b += (int64_t)gridDim.x * blockDim.x) {
        key[b] = brg[b] != bcol[b] ? (uint32_t)bcol[b] : (uint32_t)ng;
        val[b] = (int32_t)b;
    }
}
__global__ void k_ccol_fill(int64_t n, const int32_t *blk, const int64_t *boff, const int32_t *brg, const int32_t *gfirst,
                            int64_t *ccoff, int32_t *ccfn)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int32_t b = blk[q], g = brg[b];
        ccoff[q] = boff[b];
        ccfn[2 * q] = gfirst[g] - 1;
        ccfn[2 * q + 1] = gfirst[g + 1] - gfirst[g];
    }
}
// pointers of sorted keys: ptr[k] = first position with key >= k, k = 0 .. nk
__global__ void k_key_ptr(int64_t n, const uint32_t *key, int64_t nk, int64_t *ptr)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t lo = i == 0 ? -1 : (int64_t)key[i - 1];
        const int64_t hi = i == n ? nk : (int64_t)key[i];
        for (int64_t k = lo + 1; k <= hi; k++) ptr[k] = i;
    }
}
__global__ void k_npair(int64_t nrec, const int32_t *nseg, int64_t *np)
{
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r < nrec) np[r] = (int64_t)nseg[r] * (nseg[r] + 1) / 2;
    if (r == nrec) np[r] = 0;
}
// block index of every segment pair of every record (pair order a = 0.., b = 0..a)
__global__ __launch_bounds__(SW *STRUCT_WAVES) void k_pair_blocks(int64_t nrec, int cap, const int64_t *rec_gl,
                                                                 const int32_t *gl_tix, const int32_t *var,
                                                                 const int32_t *group_of, const int64_t *rec_pb,
                                                                 const uint32_t *bits, const int64_t *wrank,
                                                                 int32_t *pair_blk)
{
    extern __shared__ __align__(16) unsigned char dyn[];
    const int nw = blockDim.x / SW, wv = threadIdx.x / SW, lane = lane_id();
    int32_t *s_g = (int32_t *)(dyn + (size_t)wv * 8 * (size_t)cap);
    int32_t *s_z = s_g + cap;
    for (int64_t r = blockIdx.x * (int64_t)nw + wv; r < nrec; r += (int64_t)gridDim.x * nw) {
        int nv = 0;
        const int nseg = record_segments(rec_gl[r], rec_gl[r + 1], gl_tix, var, group_of, s_g, s_z, nv);
        const int64_t np = (int64_t)nseg * (nseg + 1) / 2;
        for (int64_t t = lane; t < np; t += SW) {
            int a, b;
            pair_of(t, a, b);
            pair_blk[rec_pb[r] + t] = (int32_t)bit_rank(pair_bit(s_g[a], s_g[b]), bits, wrank);
        }
        wsync();
    }
}

// ---------------------------------------------------------------------------------------------
// processing order, incidences
// ---------------------------------------------------------------------------------------------
__global__ void k_iota64(int64_t n, int64_t *v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) v[i] = i;
}
__global__ void k_u32_of(int64_t n, const int32_t *a, uint32_t *b)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (uint32_t)a[i];
}
// pairs of the batch's records in processing order (0 for records without an update)
__global__ void k_batch_npair(int64_t q0, int64_t q1, const int64_t *order, const int64_t *rec_pb, const int64_t *rec_cb,
                              int64_t *np)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t q = q0 + i;
    if (q < q1) {
        const int64_t r = order[q];
        np[i] = rec_cb[r] >= 0 ? rec_pb[r + 1] - rec_pb[r] : 0;
    } else if (q == q1) {
        np[i] = 0;
    }
}
// (block, buffer offset) of every segment pair of the batch's records: the record's update lies
// at rec_cb[r]; segment a's region holds the columns of all variables up to its end (sz[a]
// values each) and then the column of its vector values (padded to a multiple of four), so
// pair (a, b) starts at region(a) + (first column of b) * sz[a] and the vector column follows
// the diagonal pair (a, a).  LDS per
// wavefront: s_reg [cap] i64, s_g, s_z, s_col [cap] i32.
__global__ __launch_bounds__(SW *STRUCT_WAVES) void k_incidence(int64_t q0, int64_t q1, int cap, const int64_t *order,
                                                               const int64_t *rec_gl, const int32_t *gl_tix,
                                                               const int32_t *var, const int32_t *group_of,
                                                               const int64_t *rec_pb, const int32_t *pair_blk,
                                                               const int64_t *rec_cb, const int64_t *pos, uint32_t *key,
                                                               int64_t *val)
{
    extern __shared__ __align__(16) unsigned char dyn[];
    const int nw = blockDim.x / SW, wv = threadIdx.x / SW, lane = lane_id();
    unsigned char *base = dyn + (size_t)wv * 20 * (size_t)cap;
    int64_t *s_reg = (int64_t *)base;
    int32_t *s_g = (int32_t *)(base + 8 * (size_t)cap);
    int32_t *s_z = s_g + cap;
    int32_t *s_col = s_z + cap;
    for (int64_t q = q0 + blockIdx.x * (int64_t)nw + wv; q < q1; q += (int64_t)gridDim.x * nw) {
        const int64_t r = order[q];
        if (rec_cb[r] < 0) continue;
        int nv = 0;
        const int nseg = record_segments(rec_gl[r], rec_gl[r + 1], gl_tix, var, group_of, s_g, s_z, nv);
        if (lane == 0) {
            int64_t reg = 0, upto = 0;
            int32_t col = 0;
            for (int a = 0; a < nseg; a++) {
                s_reg[a] = reg;
                s_col[a] = col;
                col += s_z[a];
                upto += s_z[a];
                reg += (int64_t)s_z[a] * upto + ((s_z[a] + 3) & ~3);
            }
        }
        wsync();
        const int64_t out = pos[q - q0], np = (int64_t)nseg * (nseg + 1) / 2;
        for (int64_t t = lane; t < np; t += SW) {
            int a, b;
            pair_of(t, a, b);
            key[out + t] = (uint32_t)pair_blk[rec_pb[r] + t];
            val[out + t] = rec_cb[r] + s_reg[a] + (int64_t)s_col[b] * s_z[a];
        }
        wsync();
    }
}

// ---------------------------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------------------------
// temporary storage of the library scans and sorts: grows only, and only after the stream has
// finished the work that used it
static void *tmp_bytes(DBuf<uint8_t> &tmp, size_t bytes, hipStream_t s)
{
    if (tmp.n < std::max<size_t>(bytes, 1)) {
        HIPCHK(hipStreamSynchronize(s));
        tmp.alloc(std::max<size_t>(bytes, 1));
    }
    return tmp.p;
}
static void scan_excl(const int64_t *in, int64_t *out, int64_t n, hipStream_t s, DBuf<uint8_t> &tmp)
{
    if (n > INT_MAX) throw MpError(30, "structure scan too long");
    size_t bytes = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, s));
    void *t = tmp_bytes(tmp, bytes, s);
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(t, bytes, in, out, (int)n, s));
}
template <typename V>
static void sort_pairs(const uint32_t *kin, uint32_t *kout, const V *vin, V *vout, int64_t n, int bits, hipStream_t s,
                       DBuf<uint8_t> &tmp)
{
    if (n > INT_MAX) throw MpError(30, "structure sort too long");
    size_t bytes = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, (int)n, 0, bits, s));
    void *t = tmp_bytes(tmp, bytes, s);
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(t, bytes, kin, kout, vin, vout, (int)n, 0, bits, s));
}
// LDS per wavefront of `per` bytes per entry: capacity rounded up to a multiple of 64 entries,
// wavefronts per workgroup so that a workgroup stays within 64 KB
static void wave_lds(int64_t entries, int per, int &cap, int &waves, size_t &bytes, const char *what)
{
    cap = (int)std::max<int64_t>(SW, (entries + SW - 1) / SW * SW);
    if ((int64_t)cap * per > 160 * 1024)
        throw MpError(30, std::string("record too large for the structure kernels (") + what + ")");
    waves = (int)std::max<int64_t>(1, std::min<int64_t>(STRUCT_WAVES, (64 * 1024) / ((int64_t)cap * per)));
    bytes = (size_t)waves * cap * per;
}

int64_t st_decode(StructIn &in, DBuf<int64_t> &rec_eq, DBuf<int64_t> &eq_sl, DBuf<uint32_t> &eq_n, hipStream_t s)
{
    const int64_t nrec = in.nrec;
    DBuf<uint8_t> tmp;
    DBuf<int64_t> cnt;
    DBuf<unsigned long long> bad;
    cnt.alloc(nrec + 1);
    rec_eq.alloc(nrec + 1);
    const unsigned long long none = ~0ull;
    bad.upload(&none, 1, s);
    const unsigned gr = grid_for(nrec + 1, 256);
    if (in.vd) hipLaunchKernelGGL(k_decode_count<double>, dim3(gr), dim3(256), 0, s, nrec, in.rec_w, in.lab, in.vd, cnt.p, bad.p);
    else hipLaunchKernelGGL(k_decode_count<float>, dim3(gr), dim3(256), 0, s, nrec, in.rec_w, in.lab, in.vf, cnt.p, bad.p);
    scan_excl(cnt.p, rec_eq.p, nrec + 1, s, tmp);
    unsigned long long hb = 0;
    int64_t neq = 0;
    bad.download(&hb, 1, s);
    HIPCHK(hipMemcpyAsync(&neq, rec_eq.p + nrec, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (hb != none) throw MpError(20, "bad record structure in record " + std::to_string(hb + 1));
    eq_sl.alloc(std::max<int64_t>(neq, 1));
    eq_n.alloc(std::max<int64_t>(neq, 1));
    if (in.vd) hipLaunchKernelGGL(k_decode_fill<double>, dim3(gr), dim3(256), 0, s, nrec, in.rec_w, in.lab, in.vd, rec_eq.p, eq_sl.p, eq_n.p);
    else hipLaunchKernelGGL(k_decode_fill<float>, dim3(gr), dim3(256), 0, s, nrec, in.rec_w, in.lab, in.vf, rec_eq.p, eq_sl.p, eq_n.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    in.neq = neq;
    in.rec_eq = rec_eq.p;
    in.eq_sl = eq_sl.p;
    in.eq_n = eq_n.p;
    return neq;
}

void st_label_table(const StructIn &in, const std::vector<int32_t> &extra, LabelMap &map, std::vector<int32_t> &labels,
                    hipStream_t s)
{
    DBuf<int32_t> mm;
    DBuf<unsigned long long> cnt;
    const int32_t init[2] = {INT32_MAX, 0};
    const unsigned long long zero2[2] = {0, 0};
    mm.upload(init, 2, s);
    cnt.upload(zero2, 2, s);
    // a few thousand workgroups stride over the equations: every wavefront ends with four
    // atomics on the same words, which one wavefront per 256 equations would queue up
    if (in.neq > 0)
        hipLaunchKernelGGL(k_label_range, dim3(std::min<unsigned>(grid_for(in.neq, 256), 4096)), dim3(256), 0, s, in.neq,
                           in.eq_sl, in.eq_n, in.lab, mm.p, cnt.p);
    int32_t h[2];
    unsigned long long hc[2];
    mm.download(h, 2, s);
    cnt.download(hc, 2, s);
    HIPCHK(hipStreamSynchronize(s));
    map.neq_glob = (int64_t)hc[0];
    map.nglob_words = (int64_t)hc[1];
    for (int32_t l : extra) {
        h[0] = std::min(h[0], l);
        h[1] = std::max(h[1], l);
    }
    if (h[1] == 0) throw MpError(22, "no global parameters");
    if (h[0] <= 0) throw MpError(20, "label <= 0 in input");
    map.lmin = h[0];
    const int64_t span = (int64_t)h[1] - h[0] + 1;
    map.nw = (span + 31) / 32;
    map.bits.alloc(map.nw);
    map.bits.zero(s);
    if (in.neq > 0)
        hipLaunchKernelGGL(k_label_mark, dim3(grid_for(in.neq, 256)), dim3(256), 0, s, in.neq, in.eq_sl, in.eq_n, in.lab,
                           map.lmin, map.bits.p);
    DBuf<int32_t> ex;
    if (!extra.empty()) {
        ex.upload(extra, s);
        hipLaunchKernelGGL(k_label_mark_list, dim3(grid_for((int64_t)extra.size(), 256)), dim3(256), 0, s,
                           (int64_t)extra.size(), ex.p, map.lmin, map.bits.p);
    }
    DBuf<int64_t> pc;
    DBuf<uint8_t> tmp;
    pc.alloc(map.nw + 1);
    map.wrank.alloc(map.nw + 1);
    hipLaunchKernelGGL(k_word_pop, dim3(grid_for(map.nw + 1, 256)), dim3(256), 0, s, map.nw, map.bits.p, pc.p);
    scan_excl(pc.p, map.wrank.p, map.nw + 1, s, tmp);
    int64_t ntot = 0;
    HIPCHK(hipMemcpyAsync(&ntot, map.wrank.p + map.nw, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (ntot > INT32_MAX / 2) throw MpError(30, "too many global parameters");
    DBuf<int32_t> lab;
    lab.alloc(std::max<int64_t>(ntot, 1));
    hipLaunchKernelGGL(k_label_enum, dim3(grid_for(map.nw, 256)), dim3(256), 0, s, map.nw, map.bits.p, map.wrank.p,
                       map.lmin, lab.p);
    labels.resize(ntot);
    lab.download(labels.data(), ntot, s);
    HIPCHK(hipStreamSynchronize(s));
}

// FP32 record values widened to FP64 on the device (an input that mixes FP32 and FP64 records)
__global__ void k_widen(int64_t n, const float *src, double *dst)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (double)src[i];
}
void st_widen(const float *src, double *dst, int64_t n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_widen, dim3(grid_for(n, 256)), dim3(256), 0, s, n, src, dst);
    HIPCHK(hipGetLastError());
}

// Residual monitoring: per equation the label of its first global word (0: none) and its
// measurement error, in equation order
template <typename T>
__global__ void k_eq_first(int64_t neq, const int64_t *eq_sl, const uint32_t *eq_n, const int32_t *I, const T *V,
                           int32_t *lab, double *err)
{
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < neq; e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t n = eq_n[e];
        const int64_t sg = eq_sl[e] + (n & 0xffff) + 1;
        lab[e] = (n >> 16) > 0 ? I[sg + 1] : 0;
        err[e] = (double)V[sg];
    }
}
void st_eq_first(const StructIn &in, std::vector<int32_t> &lab, std::vector<double> &err, hipStream_t s)
{
    DBuf<int32_t> dl;
    DBuf<double> de;
    dl.alloc(std::max<int64_t>(in.neq, 1));
    de.alloc(std::max<int64_t>(in.neq, 1));
    if (in.neq > 0) {
        if (in.vd)
            hipLaunchKernelGGL(k_eq_first<double>, dim3(grid_for(in.neq, 256)), dim3(256), 0, s, in.neq, in.eq_sl, in.eq_n,
                               in.lab, in.vd, dl.p, de.p);
        else
            hipLaunchKernelGGL(k_eq_first<float>, dim3(grid_for(in.neq, 256)), dim3(256), 0, s, in.neq, in.eq_sl, in.eq_n,
                               in.lab, in.vf, dl.p, de.p);
    }
    HIPCHK(hipGetLastError());
    lab.resize(in.neq);
    err.resize(in.neq);
    dl.download(lab.data(), in.neq, s);
    de.download(err.data(), in.neq, s);
    HIPCHK(hipStreamSynchronize(s));
}

// Iterated entries cut (loop1i, src/pede.f90:6196-6255): the global words of every equation
// none of whose global words belongs to a parameter fixed by the entries cut (low[t] = 1) are
// counted again, zero derivatives included.  One thread per equation; integer atomics.
__global__ void k_entries_recount(int64_t neq, const int64_t *eq_sl, const uint32_t *eq_n, const int32_t *I,
                                  LabelMapD L, const uint8_t *low, unsigned long long *cnt)
{
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < neq; e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t n = eq_n[e];
        const int ng = (int)(n >> 16);
        const int64_t sg = eq_sl[e] + (n & 0xffff) + 1;
        bool any_low = false;
        for (int j = 1; j <= ng; j++) any_low |= low[tix_of(L, I[sg + j])] != 0;
        if (any_low) continue;
        for (int j = 1; j <= ng; j++) atomicAdd(&cnt[tix_of(L, I[sg + j])], 1ull);
    }
}

void st_entries_recount(const StructIn &in, const LabelMap &map, const std::vector<uint8_t> &low,
                        std::vector<int64_t> &counts, hipStream_t s)
{
    const int64_t ntot = (int64_t)low.size();
    DBuf<uint8_t> d_low;
    d_low.upload(low, s);
    DBuf<unsigned long long> cnt;
    cnt.alloc(std::max<int64_t>(ntot, 1));
    cnt.zero(s);
    const LabelMapD L{map.lmin, map.bits.p, map.wrank.p};
    if (in.neq > 0)
        hipLaunchKernelGGL(k_entries_recount, dim3(grid_for(in.neq, 256)), dim3(256), 0, s, in.neq, in.eq_sl, in.eq_n,
                           in.lab, L, d_low.p, cnt.p);
    HIPCHK(hipGetLastError());
    counts.resize(ntot);
    cnt.download((unsigned long long *)counts.data(), ntot, s);
    HIPCHK(hipStreamSynchronize(s));
}

void st_record_lists(const StructIn &in, const LabelMap &map, int32_t ntot, bool per_record, DBuf<int32_t> &widx,
                     DBuf<int64_t> &rec_gl, DBuf<int32_t> &gl_tix, std::vector<int32_t> &ngl_rec,
                     std::vector<int64_t> &counters, std::vector<uint8_t> &bnd, hipStream_t s)
{
    const int64_t nrec = in.nrec;
    DBuf<uint8_t> tmp;
    DBuf<int64_t> m, mofs, ngl;
    DBuf<int32_t> mmax;
    m.alloc(nrec + 1);
    mofs.alloc(nrec + 1);
    ngl.alloc(nrec + 1);
    const int32_t zero = 0;
    mmax.upload(&zero, 1, s);
    hipLaunchKernelGGL(k_rec_m, dim3(grid_for(nrec + 1, 256)), dim3(256), 0, s, nrec, in.rec_eq, in.eq_n, m.p, mmax.p);
    scan_excl(m.p, mofs.p, nrec + 1, s, tmp);
    int32_t mx = 0;
    int64_t mtot = 0;
    mmax.download(&mx, 1, s);
    HIPCHK(hipMemcpyAsync(&mtot, mofs.p + nrec, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    int64_t p2 = 1;
    while (p2 < mx) p2 <<= 1;
    int cap, waves;
    size_t lds;
    wave_lds(p2, 13, cap, waves, lds, "global words of a record");
    DBuf<int32_t> tgl;
    tgl.alloc(std::max<int64_t>(mtot, 1));
    DBuf<unsigned long long> cnt;
    cnt.alloc(std::max<int32_t>(ntot, 1));
    cnt.zero(s);
    DBuf<uint8_t> dbnd;
    dbnd.alloc((size_t)ntot + 1);
    dbnd.zero(s);
    const LabelMapD L{map.lmin, map.bits.p, map.wrank.p};
    const unsigned gr = grid_for(nrec, waves);
#define ST_LAUNCH(T, PR, V)                                                                                              \
    hipLaunchKernelGGL((k_rec_labels<T, PR>), dim3(gr), dim3(SW * waves), lds, s, nrec, cap, in.rec_eq, in.rec_w,        \
                       in.eq_sl, in.eq_n, in.lab, V, L, mofs.p, widx.p, tgl.p, ngl.p, cnt.p, dbnd.p)
    if (nrec > 0) {
        if (in.vd) {
            if (per_record) ST_LAUNCH(double, true, in.vd); else ST_LAUNCH(double, false, in.vd);
        } else {
            if (per_record) ST_LAUNCH(float, true, in.vf); else ST_LAUNCH(float, false, in.vf);
        }
    }
#undef ST_LAUNCH
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_tail_zero, dim3(1), dim3(1), 0, s, ngl.p + nrec);
    rec_gl.alloc(nrec + 1);
    scan_excl(ngl.p, rec_gl.p, nrec + 1, s, tmp);
    int64_t gtot = 0;
    HIPCHK(hipMemcpyAsync(&gtot, rec_gl.p + nrec, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    gl_tix.alloc(std::max<int64_t>(gtot, 1));
    if (nrec > 0)
        hipLaunchKernelGGL(k_compact_gl, dim3(grid_for(nrec, STRUCT_WAVES)), dim3(SW * STRUCT_WAVES), 0, s, nrec, mofs.p,
                           rec_gl.p, tgl.p, gl_tix.p);
    std::vector<int64_t> h(nrec);
    ngl.download(h.data(), nrec, s);
    counters.resize(ntot);
    static_assert(sizeof(unsigned long long) == sizeof(int64_t), "counter width");
    cnt.download((unsigned long long *)counters.data(), ntot, s);
    bnd.resize((size_t)ntot + 1);
    dbnd.download(bnd.data(), (size_t)ntot + 1, s);
    HIPCHK(hipStreamSynchronize(s));
    ngl_rec.resize(nrec);
    for (int64_t r = 0; r < nrec; r++) ngl_rec[r] = (int32_t)h[r];
}

void st_record_stats(const StructIn &in, const DBuf<int32_t> &widx, const DBuf<int64_t> &rec_gl,
                     const DBuf<int32_t> &gl_tix, const int32_t *var, const int32_t *group_of, int32_t ngroup,
                     int32_t max_ngl, bool sparse, PairMap &pm, std::vector<int32_t> &st, std::vector<int64_t> &cnt_r,
                     DBuf<int32_t> &d_st, hipStream_t s)
{
    const int64_t nrec = in.nrec;
    int cap, waves;
    size_t lds;
    wave_lds(max_ngl, 12, cap, waves, lds, "distinct global labels of a record");
    d_st.alloc(std::max<int64_t>(6 * nrec, 1));
    DBuf<int64_t> dcnt;
    dcnt.alloc(std::max<int64_t>(nrec, 1));
    if (sparse) {
        const int64_t npairs = (int64_t)ngroup * (ngroup + 1) / 2;
        pm.nw = (npairs + 31) / 32;
        if (pm.nw > ((int64_t)1 << 33))
            throw MpError(30, "too many parameter groups for the group-pair bitmap of sparse storage");
        pm.bits.alloc(pm.nw);
        pm.bits.zero(s);
    }
    const unsigned gr = grid_for(nrec, waves);
#define ST_LAUNCH(T, MK, V)                                                                                              \
    hipLaunchKernelGGL((k_rec_stats<T, MK>), dim3(gr), dim3(SW * waves), lds, s, nrec, cap, ngroup, in.rec_eq,          \
                       in.eq_sl, in.eq_n, widx.p, V, rec_gl.p, gl_tix.p, var, group_of, d_st.p, dcnt.p, pm.bits.p)
    if (nrec > 0) {
        if (in.vd) {
            if (sparse) ST_LAUNCH(double, true, in.vd); else ST_LAUNCH(double, false, in.vd);
        } else {
            if (sparse) ST_LAUNCH(float, true, in.vf); else ST_LAUNCH(float, false, in.vf);
        }
    }
#undef ST_LAUNCH
    HIPCHK(hipGetLastError());
    st.resize((size_t)6 * nrec);
    cnt_r.resize(nrec);
    d_st.download(st.data(), (size_t)6 * nrec, s);
    dcnt.download(cnt_r.data(), nrec, s);
    HIPCHK(hipStreamSynchronize(s));
}

void st_blocks(PairMap &pm, int32_t ngroup, const int32_t *gfirst, const std::vector<int64_t> &extra_pairs, BlockDev &B,
               hipStream_t s)
{
    DBuf<uint8_t> tmp;
    const int64_t ng = ngroup;
    if (ng > 0) hipLaunchKernelGGL(k_set_diag, dim3(grid_for(ng, 256)), dim3(256), 0, s, ngroup, pm.bits.p);
    DBuf<int64_t> ex;
    if (!extra_pairs.empty()) {
        ex.upload(extra_pairs, s);
        hipLaunchKernelGGL(k_set_bits, dim3(grid_for((int64_t)extra_pairs.size(), 256)), dim3(256), 0, s,
                           (int64_t)extra_pairs.size(), ex.p, pm.bits.p);
    }
    DBuf<int64_t> pc;
    pc.alloc(pm.nw + 1);
    pm.wrank.alloc(pm.nw + 1);
    hipLaunchKernelGGL(k_word_pop, dim3(grid_for(pm.nw + 1, 256)), dim3(256), 0, s, pm.nw, pm.bits.p, pc.p);
    scan_excl(pc.p, pm.wrank.p, pm.nw + 1, s, tmp);
    int64_t nb = 0;
    HIPCHK(hipMemcpyAsync(&nb, pm.wrank.p + pm.nw, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (nb >= ((int64_t)1 << 31)) throw MpError(30, "too many matrix blocks");
    B.nblock = nb;
    B.brow.alloc(ng + 1);
    B.brg.alloc(std::max<int64_t>(nb, 1));
    B.bcol.alloc(std::max<int64_t>(nb, 1));
    B.bcf.alloc(std::max<int64_t>(2 * nb, 1));
    B.boff.alloc(nb + 1);
    DBuf<int64_t> bsz;
    bsz.alloc(nb + 1);
    hipLaunchKernelGGL(k_brow, dim3(grid_for(ng + 1, 256)), dim3(256), 0, s, ngroup, pm.bits.p, pm.wrank.p, pm.nw,
                       B.brow.p);
    hipLaunchKernelGGL(k_blocks, dim3(grid_for(pm.nw, 256)), dim3(256), 0, s, pm.nw, pm.bits.p, pm.wrank.p, gfirst,
                       B.brg.p, B.bcol.p, bsz.p, B.bcf.p);
    hipLaunchKernelGGL(k_tail_zero, dim3(1), dim3(1), 0, s, bsz.p + nb);
    scan_excl(bsz.p, B.boff.p, nb + 1, s, tmp);
    HIPCHK(hipMemcpyAsync(&B.nstored, B.boff.p + nb, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    // column lists of the off-diagonal blocks (ascending block row within a column): a stable
    // sort of the blocks by column group, the diagonal blocks (key ngroup) cut off at the end
    const int64_t noff = nb - ng;
    DBuf<uint32_t> k0, k1;
    DBuf<int32_t> v0, v1;
    k0.alloc(std::max<int64_t>(nb, 1));
    k1.alloc(std::max<int64_t>(nb, 1));
    v0.alloc(std::max<int64_t>(nb, 1));
    v1.alloc(std::max<int64_t>(nb, 1));
    if (nb > 0) {
        hipLaunchKernelGGL(k_ccol_keys, dim3(grid_for(nb, 256)), dim3(256), 0, s, nb, ngroup, B.brg.p, B.bcol.p, k0.p, v0.p);
        sort_pairs(k0.p, k1.p, v0.p, v1.p, nb, key_bits(ng), s, tmp);
    }
    B.ccol_ptr.alloc(ng + 1);
    hipLaunchKernelGGL(k_key_ptr, dim3(grid_for(noff + 1, 256)), dim3(256), 0, s, noff, k1.p, ng, B.ccol_ptr.p);
    B.ncc = noff;
    B.cc_off.alloc(std::max<int64_t>(noff, 1));
    B.cc_fn.alloc(std::max<int64_t>(2 * noff, 1));
    B.cc_blk.alloc(std::max<int64_t>(noff, 1));
    if (noff > 0) {
        HIPCHK(hipMemcpyAsync(B.cc_blk.p, v1.p, sizeof(int32_t) * noff, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_ccol_fill, dim3(grid_for(noff, 256)), dim3(256), 0, s, noff, v1.p, B.boff.p, B.brg.p, gfirst,
                           B.cc_off.p, B.cc_fn.p);
    }
    HIPCHK(hipStreamSynchronize(s));
}

void st_pair_blocks(const StructIn &in, const DBuf<int64_t> &rec_gl, const DBuf<int32_t> &gl_tix, const int32_t *var,
                    const int32_t *group_of, int32_t max_nseg, const DBuf<int32_t> &d_st, const PairMap &pm,
                    DBuf<int64_t> &rec_pb, DBuf<int32_t> &pair_blk, hipStream_t s)
{
    const int64_t nrec = in.nrec;
    DBuf<uint8_t> tmp;
    DBuf<int64_t> np;
    np.alloc(nrec + 1);
    rec_pb.alloc(nrec + 1);
    hipLaunchKernelGGL(k_npair, dim3(grid_for(nrec + 1, 256)), dim3(256), 0, s, nrec, d_st.p + nrec, np.p);
    scan_excl(np.p, rec_pb.p, nrec + 1, s, tmp);
    int64_t tot = 0;
    HIPCHK(hipMemcpyAsync(&tot, rec_pb.p + nrec, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    pair_blk.alloc(std::max<int64_t>(tot, 1));
    int cap, waves;
    size_t lds;
    wave_lds(max_nseg, 8, cap, waves, lds, "parameter groups of a record");
    if (nrec > 0)
        hipLaunchKernelGGL(k_pair_blocks, dim3(grid_for(nrec, waves)), dim3(SW * waves), lds, s, nrec, cap, rec_gl.p,
                           gl_tix.p, var, group_of, rec_pb.p, pm.bits.p, pm.wrank.p, pair_blk.p);
    HIPCHK(hipGetLastError());
}

void st_order(int64_t nrec, int32_t ngroup, const DBuf<int32_t> &d_st, DBuf<int64_t> &order, std::vector<int64_t> &h_order,
              hipStream_t s)
{
    DBuf<uint8_t> tmp;
    DBuf<uint32_t> k0, k1;
    DBuf<int64_t> v0;
    k0.alloc(std::max<int64_t>(nrec, 1));
    k1.alloc(std::max<int64_t>(nrec, 1));
    v0.alloc(std::max<int64_t>(nrec, 1));
    order.alloc(std::max<int64_t>(nrec, 1));
    if (nrec > 0) {
        hipLaunchKernelGGL(k_u32_of, dim3(grid_for(nrec, 256)), dim3(256), 0, s, nrec, d_st.p + 5 * nrec, k0.p);
        hipLaunchKernelGGL(k_iota64, dim3(grid_for(nrec, 256)), dim3(256), 0, s, nrec, v0.p);
        sort_pairs(k0.p, k1.p, v0.p, order.p, nrec, key_bits(ngroup), s, tmp);
    }
    h_order.resize(nrec);
    order.download(h_order.data(), nrec, s);
    HIPCHK(hipStreamSynchronize(s));
}

void st_incidence(const StructIn &in, const DBuf<int64_t> &rec_gl, const DBuf<int32_t> &gl_tix, const int32_t *var,
                  const int32_t *group_of, int32_t max_nseg, const DBuf<int64_t> &order, const DBuf<int64_t> &rec_pb,
                  const DBuf<int32_t> &pair_blk, const DBuf<int64_t> &rec_cb, int64_t q0, int64_t q1, int64_t nblock,
                  int64_t *ptr, int64_t *src, int64_t nsrc, hipStream_t s)
{
    (void)in;
    DBuf<uint8_t> tmp;
    const int64_t nq = q1 - q0;
    DBuf<int64_t> np, pos;
    np.alloc(nq + 1);
    pos.alloc(nq + 1);
    hipLaunchKernelGGL(k_batch_npair, dim3(grid_for(nq + 1, 256)), dim3(256), 0, s, q0, q1, order.p, rec_pb.p, rec_cb.p,
                       np.p);
    scan_excl(np.p, pos.p, nq + 1, s, tmp);
    int64_t n = 0;
    HIPCHK(hipMemcpyAsync(&n, pos.p + nq, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (n != nsrc) throw MpError(30, "incidence count mismatch in the structure build");
    DBuf<uint32_t> k0, k1;
    DBuf<int64_t> v0;
    k0.alloc(std::max<int64_t>(n, 1));
    k1.alloc(std::max<int64_t>(n, 1));
    v0.alloc(std::max<int64_t>(n, 1));
    int cap, waves;
    size_t lds;
    wave_lds(max_nseg, 20, cap, waves, lds, "parameter groups of a record");
    if (nq > 0)
        hipLaunchKernelGGL(k_incidence, dim3(grid_for(nq, waves)), dim3(SW * waves), lds, s, q0, q1, cap, order.p, rec_gl.p,
                           gl_tix.p, var, group_of, rec_pb.p, pair_blk.p, rec_cb.p, pos.p, k0.p, v0.p);
    HIPCHK(hipGetLastError());
    if (n > 0) sort_pairs(k0.p, k1.p, v0.p, src, n, key_bits(nblock), s, tmp);
    hipLaunchKernelGGL(k_key_ptr, dim3(grid_for(n + 1, 256)), dim3(256), 0, s, n, k1.p, nblock, ptr);
    HIPCHK(hipStreamSynchronize(s));
}

} // namespace mpgpu
