// Seeded synthetic Mille-record generators for the benchmark and test workloads.
// Records follow the Mille layout (src/mille.f90:27-59): word 1 = (0.0, 0), then per
// measurement (rmeas,0), local (derivative, index)..., (sigma,0), global (derivative, label)...,
// zero derivatives omitted as MILLE does (src/mille.f90:129-146).  Every track uses its own
// counter-based random stream, so the output is independent of the number of threads.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <omp.h>
#include <vector>

namespace {

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x)
    {
        uint64_t z = (x += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    Rng(uint64_t seed, uint64_t stream)
    {
        uint64_t x = seed * 0x632be59bd9b4e019ull + stream * 0x9e3779b97f4a7c15ull + 0x1234567ull;
        for (int i = 0; i < 4; i++) s[i] = splitmix(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double gauss()
    {
        double u1 = uni(), u2 = uni();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    int range(int lo, int hi) { return lo + (int)(uni() * (double)(hi - lo + 1)); }
};

struct Rec {
    std::vector<float> v;
    std::vector<int32_t> i;
    void begin() { v.assign(1, 0.0f); i.assign(1, 0); }
    void meas(float m, const float *dl, const int *il, int nl, float sigma, const float *dg, const int *lg, int ng)
    {
        v.push_back(m); i.push_back(0);
        for (int k = 0; k < nl; k++) if (dl[k] != 0.0f) { v.push_back(dl[k]); i.push_back(il[k]); }
        v.push_back(sigma); i.push_back(0);
        for (int k = 0; k < ng; k++) if (dg[k] != 0.0f && lg[k] > 0) { v.push_back(dg[k]); i.push_back(lg[k]); }
    }
};

} // namespace

typedef struct mpgen_records {
    int64_t nrec, nwords;
    int64_t *rec_ptr;
    int32_t *inder;
    float *glder;
} mpgen_records;

// Generic driver: builds ntracks records with a per-track callback, in parallel, in order.
template <typename F> static int make_records(int64_t ntracks, mpgen_records *out, F gen)
{
    const int64_t CH = 2048;
    const int64_t nch = (ntracks + CH - 1) / CH;
    std::vector<std::vector<float>> cv(nch);
    std::vector<std::vector<int32_t>> ci(nch);
    std::vector<std::vector<int64_t>> cl(nch);
#pragma omp parallel
    {
        Rec rec;
#pragma omp for schedule(dynamic, 1)
        for (int64_t ch = 0; ch < nch; ch++) {
            for (int64_t t = ch * CH; t < std::min(ntracks, (ch + 1) * CH); t++) {
                rec.begin();
                gen(t, rec);
                cv[ch].insert(cv[ch].end(), rec.v.begin(), rec.v.end());
                ci[ch].insert(ci[ch].end(), rec.i.begin(), rec.i.end());
                cl[ch].push_back((int64_t)rec.v.size());
            }
        }
    }
    std::vector<int64_t> woff(nch + 1, 0);
    for (int64_t ch = 0; ch < nch; ch++) woff[ch + 1] = woff[ch] + (int64_t)cv[ch].size();
    out->nrec = ntracks;
    out->nwords = woff[nch];
    out->rec_ptr = (int64_t *)std::malloc(sizeof(int64_t) * (ntracks + 1));
    out->inder = (int32_t *)std::malloc(sizeof(int32_t) * (out->nwords > 0 ? out->nwords : 1));
    out->glder = (float *)std::malloc(sizeof(float) * (out->nwords > 0 ? out->nwords : 1));
    if (!out->rec_ptr || !out->inder || !out->glder) return 30;
    out->rec_ptr[0] = 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t ch = 0; ch < nch; ch++) {
        std::memcpy(out->glder + woff[ch], cv[ch].data(), sizeof(float) * cv[ch].size());
        std::memcpy(out->inder + woff[ch], ci[ch].data(), sizeof(int32_t) * ci[ch].size());
        int64_t p = woff[ch];
        for (size_t k = 0; k < cl[ch].size(); k++) { p += cl[ch][k]; out->rec_ptr[ch * CH + k + 1] = p; }
        std::vector<float>().swap(cv[ch]);
        std::vector<int32_t>().swap(ci[ch]);
    }
    return 0;
}

extern "C" {

void mpgen_free(mpgen_records *r)
{
    std::free(r->rec_ptr); std::free(r->inder); std::free(r->glder);
    r->rec_ptr = nullptr; r->inder = nullptr; r->glder = nullptr; r->nrec = r->nwords = 0;
}

// ---- toy planar detector (10 planes x 5 modules, 2 globals per module) ---------------------
// labels: 100*plane + 10*module + 1 (shift), +2 (scale); plane 1..10, module 0..4 -> 100 labels
// truth[100] in label order (shift, scale per module)
int mpgen_toy(uint64_t seed, int64_t ntracks, mpgen_records *out, double *truth)
{
    double tr[10][5][2];
    Rng rt(seed, 0xfffffffffull);
    for (int p = 0; p < 10; p++)
        for (int m = 0; m < 5; m++) { tr[p][m][0] = 0.01 * rt.gauss(); tr[p][m][1] = 1e-3 * rt.gauss(); }
    if (truth)
        for (int p = 0; p < 10; p++) for (int m = 0; m < 5; m++) { truth[(p * 5 + m) * 2] = tr[p][m][0]; truth[(p * 5 + m) * 2 + 1] = tr[p][m][1]; }
    const double sigma = 0.015;
    return make_records(ntracks, out, [&](int64_t t, Rec &rec) {
        Rng r(seed, (uint64_t)t + 1);
        double x0 = 50.0 * r.uni(), sl = 0.02 * r.gauss();
        for (int p = 0; p < 10; p++) {
            double z = 10.0 * (p + 1);
            double x = x0 + sl * z;
            if (x < 0.0 || x >= 50.0) continue;
            int m = (int)(x / 10.0);
            double xl = x - (10.0 * m + 5.0);
            double meas = x + tr[p][m][0] + tr[p][m][1] * xl + sigma * r.gauss();
            float dl[2] = {1.0f, (float)z};
            int il[2] = {1, 2};
            float dg[2] = {1.0f, (float)xl};
            int lg[2] = {100 * (p + 1) + 10 * m + 1, 100 * (p + 1) + 10 * m + 2};
            rec.meas((float)meas, dl, il, 2, (float)sigma, dg, lg, 2);
        }
    });
}

// ---- layered pixel detector with 6 rigid-body globals per module ---------------------------
// layer l at z = z0 + dz*l, a grid of nx * ny modules of size (wx, wy) centred on the beam
// axis; labels (l+1)*100000 + (module+1)*10 + k, k = 1..6 (du, dv, dw, alpha, beta, gamma).
// A track crosses a contiguous run of min(nh, nlayers) layers; hits beyond nlayers are
// overlap hits in the neighbouring module of the first layers of the run.  Each hit gives
// two measurements (u, v).  truth: 6*nlayers*nx*ny values in label order (may be NULL).
int mpgen_pixel(uint64_t seed, int nlayers, int nx, int ny, double wx, double wy, double z0, double dz,
                int64_t ntracks, int hmin, int hmax, double slope_sigma, double sigma, mpgen_records *out,
                double *truth)
{
    const int nmod = nx * ny;
    std::vector<double> tr((size_t)nlayers * nmod * 6);
    {
        Rng rt(seed, 0xfffffffffull);
        for (size_t k = 0; k < tr.size(); k++) {
            int c = (int)(k % 6);
            tr[k] = (c < 3 ? 0.01 : 1e-4) * rt.gauss();
        }
    }
    if (truth) std::memcpy(truth, tr.data(), sizeof(double) * tr.size());
    const double xmin = -0.5 * nx * wx, ymin = -0.5 * ny * wy;
    return make_records(ntracks, out, [&](int64_t t, Rec &rec) {
        Rng r(seed, (uint64_t)t + 1);
        int nh = r.range(hmin, hmax);
        int nl = nh < nlayers ? nh : nlayers;
        int extra = std::min(nh - nl, nl);
        int l0 = r.range(0, nlayers - nl);
        // track through the middle of the run of layers
        double zmid = z0 + dz * (l0 + 0.5 * (nl - 1));
        double tx = slope_sigma * r.gauss(), ty = slope_sigma * r.gauss();
        double xm = xmin + (nx * wx) * r.uni(), ym = ymin + (ny * wy) * r.uni();
        double x0 = xm - tx * zmid, y0 = ym - ty * zmid;
        for (int q = 0; q < nl + extra; q++) {
            bool ovl = q >= nl;
            int l = ovl ? l0 + (q - nl) : l0 + q;
            double z = z0 + dz * l;
            double x = x0 + tx * z, y = y0 + ty * z;
            int ix = (int)std::floor((x - xmin) / wx), iy = (int)std::floor((y - ymin) / wy);
            if (ovl) ix += 1;
            if (ix < 0 || ix >= nx || iy < 0 || iy >= ny) continue;
            int m = iy * nx + ix;
            double xc = xmin + (ix + 0.5) * wx, yc = ymin + (iy + 0.5) * wy;
            double u = x - xc, v = y - yc;
            const double *p = &tr[((size_t)l * nmod + m) * 6];
            // linearised rigid-body derivatives for a module normal to z
            double gu[6] = {1.0, 0.0, -tx, -tx * v, tx * u, -v};
            double gv[6] = {0.0, 1.0, -ty, -ty * v, ty * u, u};
            double mu = x, mv = y;
            for (int k = 0; k < 6; k++) { mu += gu[k] * p[k]; mv += gv[k] * p[k]; }
            mu += sigma * r.gauss();
            mv += sigma * r.gauss();
            int lab[6];
            for (int k = 0; k < 6; k++) lab[k] = (l + 1) * 100000 + (m + 1) * 10 + k + 1;
            float dgu[6], dgv[6];
            for (int k = 0; k < 6; k++) { dgu[k] = (float)gu[k]; dgv[k] = (float)gv[k]; }
            float dlu[2] = {1.0f, (float)z}, dlv[2] = {1.0f, (float)z};
            int ilu[2] = {1, 3}, ilv[2] = {2, 4};
            rec.meas((float)mu, dlu, ilu, 2, (float)sigma, dgu, lab, 6);
            rec.meas((float)mv, dlv, ilv, 2, (float)sigma, dgv, lab, 6);
        }
    });
}


// ---- broken-lines tracks through a stack of planes (bordered-band local fits) -------------
// nlayers planes at z = z0 + dz*l, each a grid of nx*ny modules with 6 rigid-body globals
// (labels as mpgen_pixel).  A track crosses np consecutive planes (np uniform in
// [pmin, pmax]); its local parameters are 4 border parameters (quadratic and cubic bending
// in x and y: derivatives (z-zc)^2/2 and (z-zc)^3/6) followed by 2 offsets (u, v) per plane.
// Per plane two hit measurements (offset + border terms + 6 globals), and per interior plane
// two kink pseudo-measurements with value 0 on the offsets of the plane and its neighbours
// (the broken-lines curvature terms, src/mptest2.f90:379-395), sigma theta0.  Local indices
// are ordered border first, so the local matrix is a band of half-width 4 with a border of 4.
int mpgen_brokenlines(uint64_t seed, int nlayers, int nx, int ny, double wx, double wy, double z0, double dz,
                      int64_t ntracks, int pmin, int pmax, double slope_sigma, double sigma, double theta0,
                      mpgen_records *out)
{
    const int nmod = nx * ny;
    const double xmin = -0.5 * nx * wx, ymin = -0.5 * ny * wy;
    std::vector<double> tr((size_t)nlayers * nmod * 6);
    {
        Rng rt(seed, 0xfffffffffull);
        for (size_t k = 0; k < tr.size(); k++) tr[k] = ((k % 6) < 3 ? 0.01 : 1e-4) * rt.gauss();
    }
    return make_records(ntracks, out, [&](int64_t t, Rec &rec) {
        Rng r(seed, (uint64_t)t + 1);
        int np = r.range(pmin, std::min(pmax, nlayers));
        int l0 = r.range(0, nlayers - np);
        double zc = z0 + dz * (l0 + 0.5 * (np - 1));
        double tx = slope_sigma * r.gauss(), ty = slope_sigma * r.gauss();
        double xm = xmin + (nx * wx) * (0.1 + 0.8 * r.uni()), ym = ymin + (ny * wy) * (0.1 + 0.8 * r.uni());
        // true trajectory: straight line plus random kinks (offsets accumulate)
        std::vector<double> ux(np), uy(np);
        double ax = tx, ay = ty, x = xm - tx * (zc - z0 - dz * l0), y = ym - ty * (zc - z0 - dz * l0);
        for (int q = 0; q < np; q++) {
            ux[q] = x; uy[q] = y;
            ax += theta0 * r.gauss(); ay += theta0 * r.gauss();
            x += ax * dz; y += ay * dz;
        }
        const int nb = 4;
        int nalc = nb + 2 * np;
        std::vector<float> dl(nalc);
        std::vector<int> il(nalc);
        for (int k = 0; k < nalc; k++) il[k] = k + 1;
        for (int q = 0; q < np; q++) {
            int l = l0 + q;
            double z = z0 + dz * l, s = z - zc;
            int ix = (int)std::floor((ux[q] - xmin) / wx), iy = (int)std::floor((uy[q] - ymin) / wy);
            if (ix < 0 || ix >= nx || iy < 0 || iy >= ny) continue;
            int m = iy * nx + ix;
            double xcen = xmin + (ix + 0.5) * wx, ycen = ymin + (iy + 0.5) * wy;
            double u = ux[q] - xcen, v = uy[q] - ycen;
            const double *p = &tr[((size_t)l * nmod + m) * 6];
            double gu[6] = {1.0, 0.0, -tx, -tx * v, tx * u, -v};
            double gv[6] = {0.0, 1.0, -ty, -ty * v, ty * u, u};
            double mu = ux[q], mv = uy[q];
            for (int k = 0; k < 6; k++) { mu += gu[k] * p[k]; mv += gv[k] * p[k]; }
            mu += sigma * r.gauss();
            mv += sigma * r.gauss();
            int lab[6];
            for (int k = 0; k < 6; k++) lab[k] = (l + 1) * 100000 + (m + 1) * 10 + k + 1;
            float dgu[6], dgv[6];
            for (int k = 0; k < 6; k++) { dgu[k] = (float)gu[k]; dgv[k] = (float)gv[k]; }
            // u: border (s^2/2, 0, s^3/6, 0), offset u_q
            std::fill(dl.begin(), dl.end(), 0.0f);
            dl[0] = (float)(0.5 * s * s); dl[2] = (float)(s * s * s / 6.0); dl[nb + 2 * q] = 1.0f;
            rec.meas((float)mu, dl.data(), il.data(), nalc, (float)sigma, dgu, lab, 6);
            std::fill(dl.begin(), dl.end(), 0.0f);
            dl[1] = (float)(0.5 * s * s); dl[3] = (float)(s * s * s / 6.0); dl[nb + 2 * q + 1] = 1.0f;
            rec.meas((float)mv, dl.data(), il.data(), nalc, (float)sigma, dgv, lab, 6);
        }
        // kinks at interior planes (value 0, no globals)
        const float dn = (float)(1.0 / dz), dp = (float)(1.0 / dz);
        for (int q = 1; q < np - 1; q++)
            for (int c = 0; c < 2; c++) {
                std::fill(dl.begin(), dl.end(), 0.0f);
                dl[nb + 2 * (q - 1) + c] = dp;
                dl[nb + 2 * q + c] = -dp - dn;
                dl[nb + 2 * (q + 1) + c] = dn;
                rec.meas(0.0f, dl.data(), il.data(), nalc, (float)theta0, nullptr, nullptr, 0);
            }
    });
}

// ---- dense detector: every track crosses every layer ----------------------------------------
// nlayers layers at z = 0 .. nlayers-1, each a grid of nx * ny unit modules that is periodic in
// both directions (a barrel unrolled in phi and, for the synthetic case, also closed along the
// beam), 6 rigid-body globals per module (labels as mpgen_pixel).  A track is a straight line
// through (x0, y0) at the middle layer with slopes uniform in +-slope_frac * nx / (nlayers - 1)
// (and the same for y), so that across the whole stack it sweeps slope_frac of the ring: each
// track gives nlayers hits (2 measurements each, 4 local parameters) and the module pairs it
// couples fill a large part of the normal matrix.  truth: 6*nlayers*nx*ny values (may be NULL).
int mpgen_dense(uint64_t seed, int nlayers, int nx, int ny, int64_t ntracks, double slope_frac, double sigma,
                mpgen_records *out, double *truth)
{
    const int nmod = nx * ny;
    std::vector<double> tr((size_t)nlayers * nmod * 6);
    {
        Rng rt(seed, 0xfffffffffull);
        for (size_t k = 0; k < tr.size(); k++) tr[k] = ((k % 6) < 3 ? 0.01 : 1e-3) * rt.gauss();
    }
    if (truth) std::memcpy(truth, tr.data(), sizeof(double) * tr.size());
    const double zmid = 0.5 * (nlayers - 1);
    const double sx = nlayers > 1 ? slope_frac * nx / (nlayers - 1) : 0.0;
    const double sy = nlayers > 1 ? slope_frac * ny / (nlayers - 1) : 0.0;
    return make_records(ntracks, out, [&](int64_t t, Rec &rec) {
        Rng r(seed, (uint64_t)t + 1);
        const double x0 = nx * r.uni(), y0 = ny * r.uni();
        const double tx = sx * (2.0 * r.uni() - 1.0), ty = sy * (2.0 * r.uni() - 1.0);
        for (int l = 0; l < nlayers; l++) {
            const double dz = l - zmid;
            const double x = x0 + tx * dz, y = y0 + ty * dz;   // unwrapped track position
            const double xw = x - nx * std::floor(x / nx), yw = y - ny * std::floor(y / ny);
            int ix = std::min((int)xw, nx - 1), iy = std::min((int)yw, ny - 1);
            const int m = iy * nx + ix;
            const double u = xw - (ix + 0.5), v = yw - (iy + 0.5);
            const double *p = &tr[((size_t)l * nmod + m) * 6];
            double gu[6] = {1.0, 0.0, -tx, -tx * v, tx * u, -v};
            double gv[6] = {0.0, 1.0, -ty, -ty * v, ty * u, u};
            double mu = x, mv = y;
            for (int k = 0; k < 6; k++) { mu += gu[k] * p[k]; mv += gv[k] * p[k]; }
            mu += sigma * r.gauss();
            mv += sigma * r.gauss();
            int lab[6];
            for (int k = 0; k < 6; k++) lab[k] = (l + 1) * 100000 + (m + 1) * 10 + k + 1;
            float dgu[6], dgv[6];
            for (int k = 0; k < 6; k++) { dgu[k] = (float)gu[k]; dgv[k] = (float)gv[k]; }
            float dlu[2] = {1.0f, (float)dz}, dlv[2] = {1.0f, (float)dz};
            int ilu[2] = {1, 3}, ilv[2] = {2, 4};
            rec.meas((float)mu, dlu, ilu, 2, (float)sigma, dgu, lab, 6);
            rec.meas((float)mv, dlv, ilv, 2, (float)sigma, dgv, lab, 6);
        }
    });
}

// write records as a Mille C-binary (fortran = 0) or Fortran sequential (fortran != 0) file
int mpgen_write(const char *path, const mpgen_records *r, int fortran)
{
    FILE *f = std::fopen(path, "wb");
    if (!f) return 18;
    for (int64_t k = 0; k < r->nrec; k++) {
        int64_t w0 = r->rec_ptr[k], w1 = r->rec_ptr[k + 1];
        int32_t nw = (int32_t)(w1 - w0), nr = 2 * nw;
        if (fortran) {
            int32_t len = 4 + 8 * nw;
            std::fwrite(&len, 4, 1, f);
            std::fwrite(&nr, 4, 1, f);
            std::fwrite(r->glder + w0, 4, nw, f);
            std::fwrite(r->inder + w0, 4, nw, f);
            std::fwrite(&len, 4, 1, f);
        } else {
            std::fwrite(&nr, 4, 1, f);
            std::fwrite(r->glder + w0, 4, nw, f);
            std::fwrite(r->inder + w0, 4, nw, f);
        }
    }
    std::fclose(f);
    return 0;
}

} // extern "C"
