// Pede-compatible driver on top of the C-ABI: steering file -> Mille files -> structure ->
// iterations (data passes, global solve, line search) -> millepede.res.
//
// Restates the parts of pede that decide what the GPU path computes:
//   steering keywords and file lists (filetc/filetx/intext, src/pede.f90:9406-11121),
//   the iteration driver xloopn (src/pede.f90:8501-9009) with the line search ptline
//   (src/linesrch.f90:89-231), and the result file writer prtglo (src/pede.f90:4477-4651).
// Keywords are matched case-insensitively on the whole word or an unambiguous prefix of at
// least 80% of its length (pede's matint accepts 80% fuzzy matches).
#include "mpgpu_internal.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <ctime>
#include <fstream>
#include <functional>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>
#include <sys/stat.h>
#include <unistd.h>

namespace {

struct ParLine { int32_t label; double value, presigma; };
struct EqLine { double rhs, sigma; std::vector<int32_t> labels; std::vector<double> f; };

struct Steering {
    std::vector<std::pair<std::string, int>> binfiles; // name, 1 = C binary, 3 = Fortran
    std::vector<std::pair<std::string, int>> files;    // as named in the steering file: 1 C binary, 2 text, 3 Fortran
    bool scale_given = false;                          // 'scaleerrors' present (scaleErrors = 1)
    std::vector<ParLine> params;
    std::vector<EqLine> cons, meas;
    mpgpu_options opt;
    int mitera = 1, matrit = 0, lsearch = 2, ipcntr = 0, igcorr = 0, isubit = 0, nofeas = 0, iforce = 0;
    std::vector<std::string> unknown;      // keywords pede does not know (exit 13)
    std::vector<std::string> missing;      // file names (with an extension) that do not exist (exit 15/16)
    std::vector<std::string> unsupported;  // pede keywords that would change the result but are not built (exit 40)
    std::vector<int32_t> errlabels;   // 'errlabels': MINRES error estimates (at most 100 labels)
    int imonit = 0;                    // 'monitorresiduals'/'monitorpulls': bit 0 first, bit 1 last pass, < 0 all
    float dflim = 0.0f;
    bool method_given = false;
    std::string err;
};

std::string lower(std::string s)
{
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

bool key_match(const std::string &tok, const char *key)
{
    std::string t = lower(tok), k = lower(key);
    if (t == k) return true;
    if (t.size() > k.size()) return false;
    return k.compare(0, t.size(), t) == 0 && 10 * t.size() >= 8 * k.size();
}

bool file_exists(const std::string &f)
{
    struct stat st;
    return stat(f.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

std::vector<std::string> tokens(const std::string &line)
{
    std::string l = line;
    size_t c = l.find('!');
    if (c != std::string::npos) l = l.substr(0, c);
    std::vector<std::string> t;
    std::istringstream is(l);
    std::string w;
    while (is >> w) t.push_back(w);
    return t;
}

bool is_number(const std::string &s, double &v)
{
    std::string t = s;
    for (auto &ch : t) if (ch == 'd' || ch == 'D') ch = 'e';
    char *end = nullptr;
    v = std::strtod(t.c_str(), &end);
    return end && *end == '\0' && !t.empty();
}

// parse one text file (steering or constraint/parameter file)
bool parse_text(Steering &S, const std::string &path, const std::string &dir, bool top, int &cur_ftype)
{
    std::ifstream in(path);
    if (!in) { S.err = "cannot open text file " + path; return false; }
    enum { NONE, PARAM, CONS, MEAS } mode = NONE;
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && line[0] == '*') continue;
        auto t = tokens(line);
        if (t.empty()) continue;
        double v0;
        if (is_number(t[0], v0)) {
            std::vector<double> nums;
            for (auto &w : t) { double v; if (is_number(w, v)) nums.push_back(v); }
            if (mode == PARAM && nums.size() >= 1) {
                ParLine p{(int32_t)std::llround(nums[0]), nums.size() > 1 ? nums[1] : 0.0, nums.size() > 2 ? nums[2] : 0.0};
                S.params.push_back(p);
            } else if ((mode == CONS || mode == MEAS) && nums.size() >= 2) {
                auto &E = mode == CONS ? S.cons.back() : S.meas.back();
                for (size_t k = 0; k + 1 < nums.size(); k += 2) {
                    E.labels.push_back((int32_t)std::llround(nums[k]));
                    E.f.push_back(nums[k + 1]);
                }
            }
            continue;
        }
        const std::string &k = t[0];
        std::vector<double> nums;
        for (size_t q = 1; q < t.size(); q++) { double v; if (is_number(t[q], v)) nums.push_back(v); }
        mode = NONE;
        if (key_match(k, "parameter") || key_match(k, "parameters")) { mode = PARAM; continue; }
        if (key_match(k, "constraint")) {
            mode = CONS;
            S.cons.push_back(EqLine{nums.size() > 0 ? nums[0] : 0.0, nums.size() > 1 ? nums[1] : 0.0, {}, {}});
            continue;
        }
        if (key_match(k, "measurement")) {
            mode = MEAS;
            S.meas.push_back(EqLine{nums.size() > 0 ? nums[0] : 0.0, nums.size() > 1 ? nums[1] : 0.0, {}, {}});
            continue;
        }
        if (key_match(k, "end")) { if (top) break; continue; }
        if (key_match(k, "method")) {
            std::string m = t.size() > 1 ? lower(t[1]) : "";
            S.method_given = true;
            S.err.clear();          // the last method line decides (CI/mp2str.txt lists several)
            if (m == "inversion") { S.opt.metsol = MPGPU_INVERSION; S.opt.matsto = MPGPU_FULL; }
            else if (m == "decomposition" || m == "cholesky") { S.opt.metsol = MPGPU_DECOMPOSITION; S.opt.matsto = MPGPU_FULL; }
            else if (m == "fullminres") { S.opt.metsol = MPGPU_MINRES; S.opt.matsto = MPGPU_FULL; }
            else if (m == "sparseminres") { S.opt.metsol = MPGPU_MINRES; S.opt.matsto = MPGPU_SPARSE; }
            else if (m == "fullminres-qlp") { S.opt.metsol = MPGPU_MINRES_QLP; S.opt.matsto = MPGPU_FULL; }
            else if (m == "sparseminres-qlp") { S.opt.metsol = MPGPU_MINRES_QLP; S.opt.matsto = MPGPU_SPARSE; }
            else if (m == "cholesky-fp32") { S.opt.metsol = MPGPU_CHOLESKY_FP32_IR; S.opt.matsto = MPGPU_FULL; }
            else if (m == "fulllapack") { S.opt.metsol = MPGPU_LAPACK_PACKED; S.opt.matsto = MPGPU_FULL; }
            else if (m == "unpackedlapack") { S.opt.metsol = MPGPU_LAPACK_UNPACKED; S.opt.matsto = MPGPU_FULL; }
            else { S.opt.metsol = -1; S.err = "Unsupported solution method " + m; }
            std::vector<double> mn;
            for (size_t q = 2; q < t.size(); q++) { double v; if (is_number(t[q], v)) mn.push_back(v); }
            if (mn.size() > 0) S.mitera = std::max(1, (int)std::llround(mn[0]));
            if (mn.size() > 1) S.dflim = (float)mn[1];
            continue;
        }
        if (key_match(k, "entries")) {
            if (nums.size() > 0 && nums[0] > 0.5) S.opt.mreqenf = (int)std::llround(nums[0]);
            if (nums.size() > 1 && nums[1] > 0.5) S.opt.mreqena = (int)std::llround(nums[1]);
            if (nums.size() > 2 && nums[2] > 0.5) S.opt.iteren = (int)std::llround(nums[0] * nums[2]);
            continue;
        }
        if (key_match(k, "chisqcut")) {
            if (nums.empty()) { S.opt.chicut = 1.0f; S.opt.chirem = 1.0f; }
            else {
                S.opt.chicut = (float)nums[0];
                if (S.opt.chicut < 1.0f) S.opt.chicut = -1.0f;
                if (nums.size() == 1) S.opt.chirem = 1.0f;
                else {
                    S.opt.chirem = (float)nums[1];
                    if (S.opt.chirem < 1.0f) S.opt.chirem = 1.0f;
                    if (S.opt.chicut >= 1.0f) S.opt.chirem = std::min(S.opt.chirem, S.opt.chicut);
                }
            }
            continue;
        }
        if (key_match(k, "hugecut")) { if (!nums.empty()) S.opt.chhuge = std::max(1.0f, (float)nums[0]); continue; }
        if (key_match(k, "linesearch")) { if (!nums.empty()) S.lsearch = (int)std::llround(nums[0]); continue; }
        if (key_match(k, "localfit")) { if (nums.size() > 1) S.opt.lfitbb = (int)std::llround(nums[1]); continue; }
        if (key_match(k, "regularisation") || key_match(k, "regularization")) {
            S.opt.nregul = 1;
            if (!nums.empty()) S.opt.regula = (float)nums[0];
            if (nums.size() >= 2) S.opt.default_presigma = (float)nums[1];
            continue;
        }
        if (key_match(k, "presigma")) { if (!nums.empty()) S.opt.default_presigma = (float)nums[0]; continue; }
        if (key_match(k, "matiter")) { if (!nums.empty()) S.matrit = (int)std::llround(nums[0]); continue; }
        if (key_match(k, "bandwidth")) {
            if (!nums.empty()) S.opt.mbandw = std::max(-1, (int)std::llround(nums[0]));
            if (nums.size() > 1) S.opt.lprecm = (int)std::llround(nums[1]);
            continue;
        }
        if (key_match(k, "outlierdownweighting")) {
            if (!nums.empty()) { int l = (int)std::llround(nums[0]); if (l > 0 && l <= 2) l = 2; S.opt.lhuber = l; }
            continue;
        }
        if (key_match(k, "dwfractioncut")) {
            if (!nums.empty()) S.opt.dwcut = (float)std::min(0.5, std::max(0.0, nums[0]));
            continue;
        }
        if (key_match(k, "maxlocalcond")) { if (!nums.empty()) S.opt.cndlmx = (float)nums[0]; continue; }
        if (key_match(k, "lapackwitherrors")) { S.opt.lapack_errors = 1; continue; }
        if (key_match(k, "printcounts")) { S.ipcntr = nums.empty() ? 1 : (int)std::llround(nums[0]); continue; }
        if (key_match(k, "globalcorr")) { S.igcorr = 1; continue; }
        if (key_match(k, "errlabels")) {
            for (double v : nums)
                if (S.errlabels.size() < 100) S.errlabels.push_back((int32_t)std::llround(v));
            continue;
        }
        if (key_match(k, "subito")) { S.isubit = 1; continue; }
        if (key_match(k, "nofeasiblestart")) { S.nofeas = 1; continue; }
        if (key_match(k, "force")) { S.iforce = 1; continue; }
        if (key_match(k, "scaleerrors")) {
            // src/pede.f90:10697-10704: one factor for both kinds, or one each
            S.scale_given = true;
            if (nums.size() > 0) S.opt.scale_errors[0] = S.opt.scale_errors[1] = nums[0];
            if (nums.size() > 1) S.opt.scale_errors[1] = nums[1];
            continue;
        }
        if (key_match(k, "iterateentries")) {
            // src/pede.f90:10706-10712
            S.opt.iteren = nums.empty() ? INT32_MAX : (int)std::llround(nums[0]);
            continue;
        }
        if (key_match(k, "skipemptycons")) { S.opt.skip_empty_cons = 1; continue; }
        if (key_match(k, "resolveredundancycons")) { S.opt.resolve_redundancy_cons = 1; continue; }
        if (key_match(k, "pairentries")) {
            // mreqpe > 1 drops rarely filled matrix blocks (src/pede.f90:10762-10772): not built
            if (!nums.empty() && nums[0] > 1.5) S.unsupported.push_back("pairentries");
            continue;
        }
        if (key_match(k, "withmultipliers")) { S.opt.withmultipliers = 1; continue; }
        if (key_match(k, "withelimination")) { S.opt.withmultipliers = 0; continue; }
        if (key_match(k, "countrecords")) { S.opt.mcount = 1; continue; }
        if (key_match(k, "mrestol")) {
            if (!nums.empty() && nums[0] >= 1e-10 && nums[0] <= 1e-4) S.opt.mrestl = nums[0];
            continue;
        }
        if (key_match(k, "mrestranscond")) { if (!nums.empty()) S.opt.mrtcnd = nums[0]; continue; }
        if (key_match(k, "mresmode")) { if (!nums.empty()) S.opt.mrmode = (int)std::llround(nums[0]); continue; }
        if (key_match(k, "monitorresiduals") || key_match(k, "monitorpulls")) {
            // src/pede.f90:10668-10684
            S.imonit = nums.empty() ? 3 : (int)std::llround(nums[0]);
            if (nums.size() > 1) S.opt.monitor_bins = std::max(S.opt.monitor_bins, (int)std::llround(nums[1]));
            S.opt.monitor = S.imonit == 0 ? 0 : (key_match(k, "monitorpulls") ? 2 : 1);
            continue;
        }
        if (key_match(k, "fortranfiles")) { cur_ftype = 3; continue; }
        if (key_match(k, "cfiles")) { cur_ftype = 1; continue; }
        // file names (steering file only)
        if (top) {
            std::string f = t[0];
            std::string full = file_exists(f) ? f : dir + "/" + f;
            if (file_exists(full)) {
                std::string tail = lower(f.size() >= 2 ? f.substr(f.size() - 2) : f);
                if (tail == "xt" || tail == "tx") {
                    int ft = cur_ftype;
                    S.files.push_back({f, 2});
                    if (!parse_text(S, full, dir, false, ft)) return false;
                } else {
                    S.files.push_back({f, cur_ftype});
                    S.binfiles.push_back({full, cur_ftype});
                }
                continue;
            }
            if (f.find('.') != std::string::npos && t.size() == 1) { S.missing.push_back(f); continue; }
        }
        {
            // pede keywords that change the results but are not part of this path
            static const char *changes[] = {"weightedcons", "checkinput", "postprocessing", "maxrecord"};
            // pede keywords without effect on the build and solve (threads, caches, printout,
            // monitoring of progress, PARDISO tuning; 'wolfe' sets only the c1 constant, which
            // the line search does not use, src/pede.f90:7389-7406)
            static const char *neutral[] = {"threads", "cache", "compress", "print", "debug", "printrecord",
                                            "closeandreopen", "matmoni", "outlierfracwarnthreshold", "pullrange",
                                            "memorydebug", "checkparametergroups", "monitorprogress", "histprint",
                                            "readerroraseof", "debugpardiso", "blocksizepardiso",
                                            "withlapackelimination", "wolfe"};
            bool known = false;
            for (const char *w : changes)
                if (key_match(k, w)) { S.unsupported.push_back(w); known = true; break; }
            if (!known)
                for (const char *w : neutral)
                    if (key_match(k, w)) { known = true; break; }
            if (!known) S.unknown.push_back(k);
        }
    }
    return true;
}

// NaN and infinities as gfortran writes them (right-justified; "Inf" where "Infinity" does not fit)
static std::string nonfinite(double x, int w)
{
    std::string t = x != x ? "NaN" : (w >= 9 || (x > 0 && w >= 8) ? "Infinity" : "Inf");
    if (x == x && x < 0) t = "-" + t;
    if ((int)t.size() > w) return std::string(w, '*');
    return std::string(w - t.size(), ' ') + t;
}

// Fortran Gw.d edit descriptor (gfortran rendering)
std::string fmt_gd(double x, int w, int d)
{
    if (!std::isfinite(x)) return nonfinite(x, w);
    char buf[64];
    if (x == 0.0) {
        std::snprintf(buf, sizeof buf, "%*.*f%*s", w - 4, d - 1, 0.0, 4, "");
        return buf;
    }
    char e[64];
    std::snprintf(e, sizeof e, "%.*e", d - 1, std::fabs(x));
    int ex = std::atoi(std::strchr(e, 'e') + 1);
    int k = ex + 1;
    if (k >= 0 && k <= d) {
        std::snprintf(buf, sizeof buf, "%*.*f%*s", w - 4, d - k, x, 4, "");
        return buf;
    }
    std::string digits;
    for (char c : std::string(e)) { if (c == 'e') break; if (c >= '0' && c <= '9') digits += c; }
    int eo = ex + 1;
    char s2[64];
    std::snprintf(s2, sizeof s2, "%s0.%sE%c%02d", x < 0 ? "-" : "", digits.c_str(), eo < 0 ? '-' : '+', std::abs(eo));
    std::snprintf(buf, sizeof buf, "%*s", w, s2);
    return buf;
}
std::string fmt_g(float xf, int w = 14, int d = 5) { return fmt_gd((double)xf, w, d); }

static std::string rjust(const std::string &s, int w)
{
    if ((int)s.size() > w) return std::string(w, '*');
    return std::string(w - s.size(), ' ') + s;
}
// Ew.d: [-]0.ddddE+xx, the leading zero dropped when the width is short
std::string fmt_e(double x, int w, int d)
{
    if (!std::isfinite(x)) return nonfinite(x, w);
    char e[64];
    std::snprintf(e, sizeof e, "%.*e", d - 1, std::fabs(x));
    std::string digits;
    for (const char *q = e; *q && *q != 'e'; q++) if (*q >= '0' && *q <= '9') digits += *q;
    const int eo = x == 0.0 ? 0 : std::atoi(std::strchr(e, 'e') + 1) + 1;
    char ex[16];
    if (std::abs(eo) <= 99) std::snprintf(ex, sizeof ex, "E%c%02d", eo < 0 ? '-' : '+', std::abs(eo));
    else std::snprintf(ex, sizeof ex, "%c%03d", eo < 0 ? '-' : '+', std::abs(eo));
    std::string body = std::string(x < 0 ? "-" : "") + "0." + digits + ex;
    if ((int)body.size() > w) body = std::string(x < 0 ? "-" : "") + "." + digits + ex;
    return rjust(body, w);
}
// Fw.d
std::string fmt_f(double x, int w, int d)
{
    if (!std::isfinite(x)) return nonfinite(x, w);
    char b[64];
    std::snprintf(b, sizeof b, "%.*f", d, x);
    std::string s = b;
    if ((int)s.size() > w && s.rfind("0.", 0) == 0) s = s.substr(1);
    if ((int)s.size() > w && s.rfind("-0.", 0) == 0) s = "-" + s.substr(2);
    return rjust(s, w);
}
// list-directed output of REAL(8), REAL(4) and integers after a character item (gfortran)
std::string ld_r8(double x)
{
    if (!std::isfinite(x)) return nonfinite(x, 21) + "     ";
    char e[64], b[64];
    std::snprintf(e, sizeof e, "%.16e", std::fabs(x));
    const int k = x == 0.0 ? 1 : std::atoi(std::strchr(e, 'e') + 1) + 1;
    if (k >= 0 && k <= 17) {
        std::snprintf(b, sizeof b, "%21.*f     ", 17 - k, x);
        return b;
    }
    std::snprintf(e, sizeof e, "%.16E", x);
    const char *q = std::strchr(e, 'E');
    const int eo = std::atoi(q + 1);
    std::snprintf(b, sizeof b, "%.*sE%c%03d", (int)(q - e), e, eo < 0 ? '-' : '+', std::abs(eo));
    return rjust(b, 26);
}
std::string ld_r4(float xf)
{
    const double x = (double)xf;
    if (!std::isfinite(x)) return nonfinite(x, 13) + "    ";
    char e[64], b[64];
    std::snprintf(e, sizeof e, "%.8e", std::fabs(x));
    const int k = x == 0.0 ? 1 : std::atoi(std::strchr(e, 'e') + 1) + 1;
    if (k >= 0 && k <= 9) {
        std::snprintf(b, sizeof b, "%13.*f    ", 9 - k, x);
        return b;
    }
    std::snprintf(e, sizeof e, "%.8E", x);
    const char *q = std::strchr(e, 'E');
    const int eo = std::atoi(q + 1);
    std::snprintf(b, sizeof b, "%.*sE%c%02d", (int)(q - e), e, eo < 0 ? '-' : '+', std::abs(eo));
    return rjust(b, 17);
}
std::string ld_i(long long v, int w = 12) { return rjust(std::to_string(v), w); }

// Fortran Aw output: a shorter string right-justified, a longer one cut to its first w characters
std::string fa(const std::string &t, int w)
{
    return (int)t.size() >= w ? t.substr(0, w) : std::string(w - t.size(), ' ') + t;
}
// FORMAT(1X,a8,' =',i14,' = ',a) of the size tables (src/pede.f90:6143, 7527)
std::string f101(const char *name, long long v, const char *text)
{
    return " " + fa(name, 8) + " =" + rjust(std::to_string(v), 14) + " = " + text;
}
// the Solution algorithm table, FORMATs 121-124 (src/pede.f90:8679-8682)
std::string f121(const char *a, const char *b = "") { return " " + fa(a, 40) + "   " + b; }
std::string f122(const char *a, long long v, const char *b = "") { return " " + fa(a, 40) + "   " + std::to_string(v) + b; }
std::string f123(const char *a, double v) { return " " + fa(a, 40) + "  " + fmt_e(v, 9, 2); }

// Chi^2/ndf cut for 2 or 3 standard deviations (chindl, src/mpnum.f90:2078-2115): the quantiles
// of the chi^2 distribution at the 2 and 3 sigma probabilities, divided by ndf, up to ndf 30;
// beyond, the normal approximation (z + sqrt(2 ndf - 1))^2 / (2 ndf)
float chi2_ndf_cut(int nsig, int nd)
{
    static const float q2[30] = {4.0000f, 3.0900f, 2.6750f, 2.4290f, 2.2628f, 2.1415f, 2.0481f, 1.9736f, 1.9124f, 1.8610f,
                                 1.8171f, 1.7791f, 1.7457f, 1.7161f, 1.6897f, 1.6658f, 1.6442f, 1.6246f, 1.6065f, 1.5899f,
                                 1.5745f, 1.5603f, 1.5470f, 1.5346f, 1.5230f, 1.5120f, 1.5017f, 1.4920f, 1.4829f, 1.4742f};
    static const float q3[30] = {9.0000f, 5.9146f, 4.7184f, 4.0628f, 3.6410f, 3.3436f, 3.1209f, 2.9468f, 2.8063f, 2.6902f,
                                 2.5922f, 2.5082f, 2.4352f, 2.3711f, 2.3143f, 2.2635f, 2.2178f, 2.1764f, 2.1386f, 2.1040f,
                                 2.0722f, 2.0428f, 2.0155f, 1.9901f, 1.9665f, 1.9443f, 1.9235f, 1.9040f, 1.8855f, 1.8681f};
    if (nd < 1) return 0.0f;
    if (nd <= 30) return nsig == 2 ? q2[nd - 1] : q3[nd - 1];
    const float z = nsig == 2 ? 1.690140f : 2.782170f;
    const float t = z + std::sqrt((float)(nd + nd - 1));
    return t * t / (float)(nd + nd);
}

// ptline (src/linesrch.f90:89-231) with ptldef (232-257)
struct LineSearch {
    int nsfd = 0, idgl = 0, idgr = 0, idgm = 0, minf = 1, maxf = 5, im = 1;
    double sfd[4][21];
    double stmx = 0.9, gtol = 0.9, dginit = 0, alpha = 0, tot = 0, fsaved = 0;
    void define(float gtole, float stmax, int minfe, int maxfe)
    {
        gtol = std::max(1.0e-4, (double)std::min(gtole, 0.9f));
        if (gtole == 0.0f) gtol = 0.9;
        stmx = stmax;
        if (stmx == 0.0) stmx = 10.0;
        minf = std::max(1, std::min(minfe, 18));
        maxf = std::max(2, std::min(maxfe, 19));
        if (maxfe == 0) maxf = 5;
        nsfd = 0;
    }
    // x, g, s of length n; returns info; f in/out; step out
    int line(int n, double *x, double &f, const double *g, const double *s, double &step)
    {
        int info = 0;
        double dg = 0.0;
        for (int i = 0; i < n; i++) dg = dg - g[i] * s[i];
        auto S = [&](int r, int c) -> double & { return sfd[r - 1][c]; };
        if (nsfd == 0) {
            dginit = dg;
            if (dginit >= 0.0) { step = 0.0; return 0; }
            step = 1.0;
            alpha = step;
            tot = 0.0;
            idgl = 1;
            idgr = 0;
            fsaved = f;
            nsfd = 1;
            S(1, 1) = 0.0; S(2, 1) = 0.0; S(3, 1) = dginit; S(4, 1) = 0.0;
            im = 1;
        } else {
            nsfd++;
            S(1, nsfd) = tot; S(2, nsfd) = f - fsaved; S(3, nsfd) = dg; S(4, nsfd) = 0.0;
            if (dg < S(3, im)) im = nsfd;
            if (dg <= 0.0 && dg >= S(3, idgl)) idgl = nsfd;
            if (dg >= 0.0) { if (idgr == 0) idgr = nsfd; if (dg <= S(3, idgr)) idgr = nsfd; }
            int i1, i2;
            if (idgr == 0) { i1 = nsfd - 1; i2 = nsfd; } else { i1 = idgl; i2 = idgr; }
            double fp1 = S(3, i1), fp2 = S(3, i2);
            S(4, nsfd) = (S(1, i1) * fp2 - S(1, i2) * fp1) / (fp2 - fp1);
            if (nsfd >= minf && std::fabs(dg) <= std::fabs(dginit) * gtol) {
                alpha = tot + alpha;
                step = alpha;
                idgm = idgl;
                if (idgr != 0 && S(3, idgr) + S(3, idgl) < 0.0) idgm = idgr;
                info = 1;
            } else if (nsfd >= maxf) info = 2;
            else {
                alpha = std::min(S(4, nsfd), stmx) - tot;
                if (std::fabs(alpha) < 1.0e-3 && S(4, nsfd) > stmx) info = 3;
                else if (std::fabs(alpha) < 1.0e-3) info = 4;
            }
            if (info > 0) {
                im = 1;
                for (int i = 1; i <= nsfd; i++) if (std::fabs(S(3, i)) < std::fabs(S(3, im))) im = i;
                alpha = S(1, im) - S(1, nsfd);
                if (im == nsfd) return info;
                for (int i = 0; i < n; i++) x[i] = x[i] + alpha * s[i];
                f = S(2, im) + fsaved;
                step = S(1, im);
                if (im != 1) return info;
                step = 0.0;
                return 5;
            }
        }
        for (int i = 0; i < n; i++) x[i] = x[i] + alpha * s[i];
        tot = tot + alpha;
        step = tot;
        return -1;
    }
    // ptlopt (src/linesrch.f90:258-288): slopes and steps of the first and the last call
    void opt(float slopes[3], float steps[3]) const
    {
        if (nsfd == 0) { for (int i = 0; i < 3; i++) slopes[i] = steps[i] = 0.0f; return; }
        int m = 1;
        for (int i = 1; i <= nsfd; i++) if (std::fabs(sfd[2][i]) < std::fabs(sfd[2][m])) m = i;
        slopes[0] = (float)sfd[2][1]; slopes[1] = (float)sfd[2][nsfd]; slopes[2] = (float)sfd[2][m];
        steps[0] = (float)sfd[0][1]; steps[1] = (float)sfd[0][nsfd]; steps[2] = (float)sfd[0][m];
    }
    // ptlprt (src/linesrch.f90:294-343): the table of this line search
    void print(FILE *f, int lsinfo) const
    {
        if (!f || nsfd <= 0) return;
        std::fprintf(f, "  \n PTLINE: line-search method based on slopes with sufficient slope-decrease\n");
        std::fprintf(f, " PTLDEF: slope ratio limit=%s\n", ld_r8(gtol).c_str());
        std::fprintf(f, " PTLDEF: maximum step =%s\n", ld_r8(stmx).c_str());
        std::fprintf(f, " PTLDEF:%s  <= nr of calls <=%s\n", ld_i(minf).c_str(), ld_i(maxf).c_str());
        std::fprintf(f, "  i      x              F(x)         F'(X)          minimum     F'(X)\n");
        int m = 1;
        for (int i = 1; i <= nsfd; i++) if (std::fabs(sfd[2][i]) < std::fabs(sfd[2][m])) m = i;
        for (int i = 1; i <= nsfd; i++) {
            char tlr[3] = "  ";
            if (i == m) tlr[0] = tlr[1] = '*';
            if (i == idgl) tlr[0] = 'L';
            if (i == idgr) tlr[1] = 'R';
            std::string l = rjust(std::to_string(i - 1), 3) + fmt_f(sfd[0][i], 12, 6) + " " + tlr +
                            fmt_gd(sfd[1][i], 15, 6) + fmt_gd(sfd[2][i], 14, 6) + fmt_f(sfd[3][i], 12, 6);
            if (i == 1) l += "     ratio";
            else l += fmt_f((float)std::fabs(sfd[2][i] / sfd[2][1]), 10, 3);
            std::fprintf(f, "%s\n", l.c_str());
        }
        static const char *msg[5] = {"PTLINE: INFO=0  input error (e.g. gradient not negative)",
                                     "PTLINE: INFO=1  convergence reached", "PTLINE: INFO=2  too many function calls",
                                     "PTLINE: INFO=3  maximum step reached", "PTLINE: INFO=4  step too small (< 0.001)"};
        if (lsinfo >= 0 && lsinfo <= 4) std::fprintf(f, " %s\n", msg[lsinfo]);
        std::fprintf(f, "  \n");
    }
};

// Feasibility of the linear constraints (feasma/feasib, src/pede.f90:1377-1680): the
// constraints are split into groups that share no parameter; per group the product matrix
// A A^T (A = constraint factors on the variable parameters) is inverted once, and a correction
// A^T (A A^T)^-1 r moves the parameters onto the constraint surface (two passes, the first
// skipped when the rms discrepancy is below the cut).
struct Feasibility {
    struct Group { std::vector<int> cons; std::vector<double> inv; };
    std::vector<Group> groups;
    const std::vector<EqLine> *cons = nullptr;
    std::vector<int32_t> tot_of_con_term;   // total index of each constraint term (-1: unknown label)
    std::vector<int64_t> term_ptr;
    int rank = 0;                            // rank of the product matrix A A^T (sum over groups)

    // symmetric Gauss-Jordan with diagonal pivoting on a dense n x n matrix (the sqminv
    // algorithm, src/mpnum.f90:97-213): rows without a pivot above 16 eps of their original
    // diagonal are cleared, so dependent constraints get zero weight; returns the rank
    static int invert(std::vector<double> &V, int n)
    {
        int nrank = 0;
        std::vector<double> diag(n);
        std::vector<char> done(n, 0);
        for (int i = 0; i < n; i++) diag[i] = std::fabs(V[(size_t)i * n + i]);
        const double eps = 16.0 * 2.220446049250313e-16;
        for (int it = 0; it < n; it++) {
            int k = -1;
            double vkk = 0.0;
            for (int j = 0; j < n; j++)
                if (!done[j] && std::fabs(V[(size_t)j * n + j]) > std::max(std::fabs(vkk), eps * diag[j])) {
                    vkk = V[(size_t)j * n + j];
                    k = j;
                }
            if (k < 0) break;
            done[k] = 1;
            nrank++;
            double piv = 1.0 / V[(size_t)k * n + k];
            for (int j = 0; j < n; j++)
                for (int l = 0; l < n; l++)
                    if (j != k && l != k) V[(size_t)j * n + l] -= V[(size_t)j * n + k] * piv * V[(size_t)k * n + l];
            for (int j = 0; j < n; j++)
                if (j != k) { V[(size_t)j * n + k] *= piv; V[(size_t)k * n + j] *= piv; }
            V[(size_t)k * n + k] = -piv;
        }
        for (int j = 0; j < n; j++)
            for (int l = 0; l < n; l++) V[(size_t)j * n + l] = (done[j] && done[l]) ? -V[(size_t)j * n + l] : 0.0;
        return nrank;
    }

    void setup(const std::vector<EqLine> &C, const std::vector<int32_t> &labels, const std::vector<int32_t> &var)
    {
        cons = &C;
        const int nc = (int)C.size();
        term_ptr.assign(1, 0);
        tot_of_con_term.clear();
        for (auto &e : C) {
            for (int32_t l : e.labels) {
                auto it = std::lower_bound(labels.begin(), labels.end(), l);
                tot_of_con_term.push_back(it != labels.end() && *it == l ? (int32_t)(it - labels.begin()) : -1);
            }
            term_ptr.push_back((int64_t)tot_of_con_term.size());
        }
        // connected components of constraints over shared variable parameters
        std::vector<int> parent(nc);
        for (int i = 0; i < nc; i++) parent[i] = i;
        std::function<int(int)> root = [&](int i) { while (parent[i] != i) i = parent[i] = parent[parent[i]]; return i; };
        std::map<int32_t, int> owner;
        for (int i = 0; i < nc; i++)
            for (int64_t q = term_ptr[i]; q < term_ptr[i + 1]; q++) {
                int32_t t = tot_of_con_term[q];
                if (t < 0 || var[t] <= 0) continue;
                auto ins = owner.emplace(var[t], i);
                if (!ins.second) parent[root(i)] = root(ins.first->second);
            }
        std::map<int, int> gid;
        groups.clear();
        rank = 0;
        for (int i = 0; i < nc; i++) {
            auto ins = gid.emplace(root(i), (int)groups.size());
            if (ins.second) groups.emplace_back();
            groups[ins.first->second].cons.push_back(i);
        }
        for (auto &g : groups) {
            int n = (int)g.cons.size();
            std::map<int32_t, std::vector<std::pair<int, double>>> cols;   // var -> (row, factor)
            for (int a = 0; a < n; a++) {
                int i = g.cons[a];
                for (int64_t q = term_ptr[i]; q < term_ptr[i + 1]; q++) {
                    int32_t t = tot_of_con_term[q];
                    if (t >= 0 && var[t] > 0) cols[var[t]].push_back({a, C[i].f[q - term_ptr[i]]});
                }
            }
            g.inv.assign((size_t)n * n, 0.0);
            for (auto &cv : cols) {
                // merge repeated labels of one constraint first
                std::map<int, double> f;
                for (auto &rf : cv.second) f[rf.first] += rf.second;
                for (auto &x : f)
                    for (auto &y : f) g.inv[(size_t)x.first * n + y.first] += x.second * y.second;
            }
            rank += invert(g.inv, n);
        }
    }

    // returns 1 if a correction was applied; par is indexed by total index
    int apply(double concut, std::vector<double> &par, const std::vector<int32_t> &var) const
    {
        const auto &C = *cons;
        const int nc = (int)C.size();
        if (nc == 0) return 0;
        int iact = 0;
        std::vector<double> r(nc);
        for (int pass = 1; pass <= 2; pass++) {
            double s1 = 0;
            for (int i = 0; i < nc; i++) {
                double rhs = C[i].rhs;
                for (int64_t q = term_ptr[i]; q < term_ptr[i + 1]; q++) {
                    int32_t t = tot_of_con_term[q];
                    if (t >= 0) rhs -= C[i].f[q - term_ptr[i]] * par[t];
                }
                r[i] = rhs;
                s1 += rhs * rhs;
            }
            s1 = std::sqrt(s1 / nc);
            if (pass == 1 && s1 < concut) return 0;
            iact = 1;
            std::vector<double> lam(nc, 0.0);
            for (auto &g : groups) {
                int n = (int)g.cons.size();
                for (int a = 0; a < n; a++) {
                    double s = 0;
                    for (int b2 = 0; b2 < n; b2++) s += g.inv[(size_t)a * n + b2] * r[g.cons[b2]];
                    lam[g.cons[a]] = s;
                }
            }
            for (int i = 0; i < nc; i++)
                for (int64_t q = term_ptr[i]; q < term_ptr[i + 1]; q++) {
                    int32_t t = tot_of_con_term[q];
                    if (t >= 0 && var[t] > 0) par[t] += lam[i] * C[i].f[q - term_ptr[i]];
                }
        }
        return iact;
    }
};

// millepede.mon rows of one pass (monres, src/pede.f90:7533-7633): per measurement group the
// entries, the median from the cumulative histogram (interpolated in the bin), the median of the
// absolute deviations from it times 1.4826, and the mean error; REAL(4) arithmetic as pede
void write_monres(FILE *f, int nloopn, const std::vector<int32_t> &lab, const std::vector<float> &res,
                  const std::vector<int64_t> &hist, int bins)
{
    const float bsz = 0.1f;
    for (size_t g = 0; g < lab.size(); g++) {
        const int64_t *h = hist.data() + g * bins;
        std::vector<int64_t> isuml(bins + 1, 0), isums(bins + 1, 0);   // 1-based
        isuml[1] = h[0];
        for (int j = 2; j <= bins; j++) isuml[j] = isuml[j - 1] + h[j - 1];
        const int64_t nent = isuml[bins];
        float amed = 0.0f, amad = 0.0f;
        if (nent > 0) {
            int j = 2;
            for (; j <= bins; j++) if (2 * isuml[j] > nent) break;
            const int imed = j;
            amed = (float)j;
            if (isuml[j] > isuml[j - 1])
                amed = amed + (float)(nent - 2 * isuml[j - 1]) / (float)(2 * isuml[j] - 2 * isuml[j - 1]);
            amed = bsz * (amed - (float)(bins / 2));
            for (int q = imed; q <= bins; q++) isums[q - imed + 1] += h[q - 1];
            for (int q = imed - 1; q >= 1; q--) isums[imed - q] += h[q - 1];
            for (int q = 2; q <= bins; q++) isums[q] += isums[q - 1];
            for (j = 2; j <= bins; j++) if (2 * isums[j] > nent) break;
            amad = (float)(j - 1);
            if (isums[j] > isums[j - 1])
                amad = amad + (float)(nent - 2 * isums[j - 1]) / (float)(2 * isums[j] - 2 * isums[j - 1]);
            amad = bsz * amad;
        }
        std::fprintf(f, "%5d%10d%10lld%s%s%s\n", nloopn, lab[g], (long long)nent, fmt_g(amed).c_str(),
                     fmt_g(amad * 1.4826f).c_str(), fmt_g(res[g]).c_str());
    }
}

int write_res(const std::string &path, const std::vector<int32_t> &labels, const std::vector<int32_t> &var,
              const std::vector<double> &par, const std::vector<double> &start, const std::vector<double> &presig,
              const std::vector<int64_t> &cnt_var, const std::vector<int64_t> &cnt_tot, const std::vector<double> *err,
              const std::vector<double> *gcor, int ipcntr, int mreqena)
{
    // keep an existing file as name~ (mvopen, src/pede.f90:11302-11333)
    if (file_exists(path)) std::rename(path.c_str(), (path + "~").c_str());
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return 30;
    std::fprintf(f, " Parameter   ! first 3 elements per line are significant (if used as input)\n");
    for (size_t t = 0; t < labels.size(); t++) {
        int v = var[t];
        float p = (float)par[t];
        float ps = (float)presig[t];
        int64_t icount = 0;
        bool lowstat = false;
        if (v > 0) {
            icount = cnt_var[v - 1];
            lowstat = icount < mreqena;
        }
        if (ipcntr > 1) icount = cnt_tot[t];
        if (lowstat) icount = -(icount + 1);
        std::string line;
        char lab[32];
        std::snprintf(lab, sizeof lab, "%10d  ", labels[t]);
        line = lab;
        if (v <= 0) {
            line += fmt_g(p) + fmt_g(ps);
            if (ipcntr != 0) {
                char c[32];
                std::snprintf(c, sizeof c, "%28s%12lld", "", (long long)icount);
                line += c;
            }
        } else {
            float dpa = (float)(par[t] - start[t]);
            if (err) {
                line += fmt_g(p) + fmt_g(ps) + fmt_g(dpa) + fmt_g((float)(*err)[v - 1]);
                if (ipcntr != 0) { char c[32]; std::snprintf(c, sizeof c, "%12lld", (long long)icount); line += c; }
                else if (gcor) { char c[32]; std::snprintf(c, sizeof c, "%8.3f", (float)(*gcor)[v - 1]); line += c; }
            } else {
                line += fmt_g(p) + fmt_g(ps) + fmt_g(dpa);
                if (ipcntr != 0) { char c[48]; std::snprintf(c, sizeof c, "%14s%12lld", "", (long long)icount); line += c; }
            }
        }
        while (!line.empty() && line.back() == ' ') line.pop_back();
        std::fprintf(f, "%s\n", line.c_str());
    }
    std::fclose(f);
    return 0;
}

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

} // namespace

extern "C" int mpgpu_run_steering(const char *steering_file, const char *output_dir, int32_t device, mpgpu_run_stats *st)
{
    mpgpu_run_stats loc{};
    auto finish = [&](int code) {
        loc.exit_code = code;
        if (st) *st = loc;
        std::string od = output_dir ? output_dir : ".";
        FILE *e = std::fopen((od + "/millepede.end").c_str(), "w");
        if (e) {
            std::fprintf(e, "%5d   %s\n", code, code == 0 ? "Ended normally" : "Aborted");
            std::fclose(e);
        }
        return code;
    };
    std::string od = output_dir ? output_dir : ".";
    const double t_start = now_ms();
    {
        // crash sentinel first, as pede does (src/pede.f90:134)
        FILE *e = std::fopen((od + "/millepede.end").c_str(), "w");
        if (e) { std::fprintf(e, "%5d   %s\n", -1, "Still running or crashed"); std::fclose(e); }
    }
    Steering S;
    mpgpu_default_options(&S.opt);
    std::string sf = steering_file;
    std::string dir = sf.find('/') == std::string::npos ? "." : sf.substr(0, sf.rfind('/'));
    int ftype = 1;
    if (!parse_text(S, sf, dir, true, ftype)) { std::fprintf(stderr, "%s\n", S.err.c_str()); return finish(10); }
    if (S.opt.metsol < 0) { std::fprintf(stderr, "%s\n", S.err.c_str()); return finish(36); }
    for (auto &f : S.missing) {
        // open errors for text (16) or binary (15) files (src/pede.f90:9700-9842)
        std::string tail = lower(f.size() >= 3 ? f.substr(f.size() - 3) : f);
        std::fprintf(stderr, " Open error for file %s\n", f.c_str());
        return finish(tail == "txt" ? 16 : 15);
    }
    if (!S.unknown.empty()) {
        // pede: 'Unknown keyword(s)' for each, then exit 13 (src/pede.f90:10886-10896, 10015-10020)
        for (auto &u : S.unknown) std::fprintf(stderr, " Unknown keyword(s): %s\n", u.c_str());
        return finish(13);
    }
    if (!S.unsupported.empty()) {
        std::fprintf(stderr, "pede(gpu): steering keyword '%s' changes the result and is not supported\n",
                     S.unsupported[0].c_str());
        return finish(40);
    }
    // millepede.log: header and file table (src/pede.f90:129-132, 9725-9735), the sizes of the
    // two structure loops, the solution algorithm, the iteration table (ploopa-ploopd,
    // src/pede.f90:2891-3040), the line search tables (ptlprt, src/linesrch.f90:294-343), the
    // final chi^2 summary and the loop counts and times (src/pede.f90:455-480)
    std::unique_ptr<FILE, int (*)(FILE *)> flog_keep(std::fopen((od + "/millepede.log").c_str(), "w"), std::fclose);
    FILE *flog = flog_keep.get();
    auto logl = [&](const std::string &l) {
        if (flog) std::fprintf(flog, "%s\n", l.c_str());
    };
    auto both = [&](const std::string &l) {
        std::printf("%s\n", l.c_str());
        logl(l);
    };
    {
        const std::time_t now = std::time(nullptr);
        char date[32] = "", host[256] = "";
        std::strftime(date, sizeof date, "%a %b %d %H:%M:%S %Y", std::localtime(&now));
        if (gethostname(host, sizeof host - 1) != 0) host[0] = '\0';
        std::string h = host;
        if (h.size() < 24) h += std::string(24 - h.size(), ' ');
        logl(" Using mpgpu: pede's global matrix build and global solve on HIP (gfx950)");
        logl("  ");
        logl(std::string(" Log-file Millepede II-P                        ") + date);
        logl("                                                " + h);
        logl("  ");
        logl(" Text and data files:");
        static const char *kind[4] = {"", "C_binary", "text", "Fortran_binary"};
        for (size_t i = 0; i < S.files.size(); i++) {
            std::string k = kind[S.files[i].second];
            logl(ld_i((long long)i + 1, 3) + "  " + k + std::string(14 - k.size(), ' ') + "   " + S.files[i].first);
        }
    }
    const double t_text = now_ms() - t_start;
    if (S.binfiles.empty()) return finish(11);
    mpgpu_ctx *c = nullptr;
    int rc = mpgpu_create(&S.opt, device, &c);
    if (rc) return finish(rc);
    auto bail = [&](int code) {
        std::fprintf(stderr, "pede(gpu): %s\n", mpgpu_last_error(c));
        mpgpu_destroy(c);
        return finish(code);
    };
    double t0 = now_ms();
    for (auto &bf : S.binfiles) {
        int64_t n = 0;
        rc = mpgpu_read_mille_file(c, bf.first.c_str(), bf.second == 3 ? 1 : 0, &n);
        if (rc) return bail(rc);
        loc.nrec += n;
    }
    const double t_read = now_ms() - t0;
    std::vector<int32_t> pl;
    std::vector<double> pv, ps;
    for (auto &p : S.params) { pl.push_back(p.label); pv.push_back(p.value); ps.push_back(p.presigma); }
    if (!pl.empty() && (rc = mpgpu_set_steering_parameters(c, (int32_t)pl.size(), pl.data(), pv.data(), ps.data()))) return bail(rc);
    for (int kind = 0; kind < 2; kind++) {
        auto &L = kind == 0 ? S.cons : S.meas;
        if (L.empty()) continue;
        std::vector<int64_t> ptr{0};
        std::vector<int32_t> lab;
        std::vector<double> fac, rhs, sig;
        for (auto &e : L) {
            lab.insert(lab.end(), e.labels.begin(), e.labels.end());
            fac.insert(fac.end(), e.f.begin(), e.f.end());
            ptr.push_back((int64_t)lab.size());
            rhs.push_back(e.rhs);
            sig.push_back(e.sigma);
        }
        rc = mpgpu_add_equations(c, kind, (int32_t)L.size(), ptr.data(), lab.data(), fac.data(), rhs.data(), sig.data());
        if (rc) return bail(rc);
    }
    int32_t ntot, nvar, ncon;
    const double t_build0 = now_ms();
    if ((rc = mpgpu_build_structure(c, &ntot, &nvar, &ncon))) return bail(rc);
    loc.ntot = ntot; loc.nvar = nvar; loc.ncon = ncon;
    loc.ms_structure = now_ms() - t0;
    const double t_build = now_ms() - t_build0;
    const int nall = nvar + ncon;
    std::vector<int32_t> labels(ntot), var(ntot), grp(ntot);
    std::vector<int64_t> cnt_tot(ntot);
    mpgpu_get_labels(c, labels.data(), var.data(), cnt_tot.data(), grp.data());
    mpgpu_record_summary sm{};
    mpgpu_get_record_summary(c, &sm);
    // pede's parameter counts: with elimination (its default, and always for decomposition,
    // src/pede.f90:6476-6502) nConstraints fewer fit parameters, else nConstraints multipliers
    const bool elim_count = !S.opt.withmultipliers || S.opt.metsol == MPGPU_DECOMPOSITION;
    const int64_t nall_active = elim_count ? (int64_t)sm.nvar : (int64_t)sm.nvar + sm.ncon;
    const int64_t nfitpar = elim_count ? (int64_t)sm.nvar - sm.ncon : (int64_t)sm.nvar + sm.ncon;
    {
        // loop1 (src/pede.f90:5802-6141): labels of the text files, parameter and group counts,
        // record and equation counters
        std::vector<int32_t> txt;
        for (auto &p : S.params) txt.push_back(p.label);
        for (auto &e : S.cons) txt.insert(txt.end(), e.labels.begin(), e.labels.end());
        for (auto &e : S.meas) txt.insert(txt.end(), e.labels.begin(), e.labels.end());
        std::sort(txt.begin(), txt.end());
        txt.erase(std::unique(txt.begin(), txt.end()), txt.end());
        long long npresg = 0, nfix3 = 0;
        for (auto &p : S.params) npresg += p.presigma > 0.0;
        for (int32_t t = 0; t < ntot; t++) nfix3 += var[t] == -3;
        const bool iterate = S.opt.iteren > S.opt.mreqenf;
        logl("  ");
        logl(" LOOP1: starting");
        if (!txt.empty()) logl(" LOOP1:" + ld_i((long long)txt.size()) + "  labels from txt data stored");
        logl(" LOOP1: reading data files");
        logl(" LOOP1:" + ld_i(sm.ntot) + "  is total number nTotalGlobalPar of labels/parameters");
        logl(" LOOP1:" + ld_i(npresg) + "  is number of pre-sigmas");
        logl(" LOOP1:" + ld_i(sm.nvar + nfix3) + "  is number nVarGlobalPar of variable parameters");
        if (iterate && S.opt.mcount == 0) {
            logl("  ");
            logl(" LOOP1: iterating");
            logl(" LOOP1:" + ld_i(sm.nvar) + "  is number nVarGlobalPar of variable parameters");
        } else if (iterate) {
            logl(" LOOP1: counting records, NO iteration of entries cut !");
        }
        logl(" LOOP1:" + ld_i(sm.ngroup_tot) + "  is total number nTotalParGroups of label/parameter groups");
        logl("  ");
        logl(f101("  NREC", sm.nrec, "number of records"));
        if (sm.fp64_payload) logl(f101(" NRECD", sm.nrec, "number of records containing doubles"));
        logl(f101("  nbEqRead", sm.neq, "number of equations (measurements)"));
        logl(f101("  nbEqWithGlobPar", sm.neq_glob, "number of equations with global parameters"));
        logl(f101("  nbGlobDeriv", sm.nglob_deriv, "number of global derivatives"));
        logl(f101("MREQENF", S.opt.mreqenf, S.opt.mcount == 0 ? "required number of entries (eqns in binary files)"
                                                               : "required number of entries (recs in binary files)"));
        if (iterate && S.opt.mcount == 0)
            logl(f101("ITEREN", S.opt.iteren, "iterate cut for parameters with less entries"));
        logl(f101("MREQENA", S.opt.mreqena, "required number of entries (from accepted fits)"));
        logl(" LOOP1: ending");
        logl("  ");
        // loop2 (src/pede.f90:6456-7493): sizes, solution method, chi^2 cut table
        logl("  ");
        logl(" LOOP2: starting");
        logl(" LOOP2: start event reading");
        logl(" LOOP2: event reading ended - end of data");
        logl("  ");
        logl(f101("nTotalGlobalPar", sm.ntot, "total number of parameters"));
        logl(std::string(22, ' ') + "(all parameters, appearing in binary files)");
        logl(f101("nVarGlobalPar", sm.nvar, "number of variable parameters"));
        logl(std::string(22, ' ') + "(appearing in fit matrix/vectors)");
        logl(f101("nAllActivePar", nall_active, "number of all parameters"));
        logl(std::string(22, ' ') + "(including Lagrange multiplier or reduced)");
        logl(f101("nTotalParGroups", sm.ngroup_tot, "total number of parameter groups"));
        logl(f101("nVarParGroups", sm.ngroup_var, "number of variable parameter groups"));
        logl(f101("nFitPar", nfitpar, "number of fit parameters"));
        if (S.opt.metsol == MPGPU_MINRES || S.opt.metsol == MPGPU_MINRES_QLP) {
            logl(f101("MBANDW", S.opt.mbandw, "band width of preconditioner matrix"));
            logl(std::string(22, ' ') + "(if <0, no preconditioner matrix)");
        }
        const long long noff = nall_active * (nall_active - 1) / 2;
        if (nall_active >= 65536) logl(f101("NOFF/K", noff / 1000, "max number of off-diagonal elements"));
        else logl(f101("NOFF", noff, "max number of off-diagonal elements"));
        logl(f101("nConstraints", sm.ncon, "number of constraints"));
        logl(f101("nGlobalMaxPerRec", sm.max_glob_rec, "max number of global parameters in an event"));
        logl(f101("nLocalMaxPerRec", sm.max_loc_rec, "max number of local parameters in an event"));
        logl(f101("nEqMaxPerRec", sm.max_eq_rec, "max number of equations in an event"));
        if (sm.nempty_rec > 0) logl(f101("nEmptyRecords", sm.nempty_rec, "number of records without variable parameters"));
        logl("  ");
        logl("  ");
        logl(" Solution method and matrix-storage mode:");
        const int ms = S.opt.metsol;
        char rtol[64];
        std::snprintf(rtol, sizeof rtol, "%s)", ld_r8(S.opt.mrestl).c_str());
        if (ms == MPGPU_INVERSION) logl("      METSOL = 1:  matrix inversion");
        else if (ms == MPGPU_DECOMPOSITION) logl("      METSOL = 3:  decomposition");
        else if (ms == MPGPU_MINRES) logl(std::string("      METSOL = 4:  MINRES (rtol") + rtol);
        else if (ms == MPGPU_MINRES_QLP) logl(std::string("      METSOL = 5:  MINRES-QLP (rtol") + rtol);
        else if (ms == MPGPU_LAPACK_PACKED) logl("      METSOL = 7:  LAPACK factorization");
        else if (ms == MPGPU_LAPACK_UNPACKED) logl("      METSOL = 8:  LAPACK factorization");
        else if (ms == MPGPU_CHOLESKY_FP32_IR) logl("      METSOL = 11:  Cholesky, FP32 factor with FP64 refinement");
        logl("                   with" + ld_i(S.mitera) + "  iterations");
        if (ms == MPGPU_LAPACK_UNPACKED) logl("      MATSTO = 0:  unpacked symmetric matrix, n*n elements");
        else if (S.opt.matsto == MPGPU_FULL) logl("      MATSTO = 1:  full symmetric matrix, (n*n+n)/2 elements");
        else logl("      MATSTO = 2:  sparse matrix (custom)");
        if (S.dflim != 0.0f) logl(" Convergence assumed, if expected dF <" + fmt_gd(S.dflim, 12, 4));
        if (sm.ncon > 0)
            logl(nfitpar < sm.nvar ? " Constraints handled by elimination" : " Constraints handled by Lagrange multipliers");
    }
    const double t_loop1 = t_read, t_loop2 = t_build;
    std::vector<double> par(ntot), start(ntot), presig(ntot, 0.0);
    mpgpu_get_parameters(c, par.data());
    start = par;
    for (auto &p : S.params) {
        auto it = std::lower_bound(labels.begin(), labels.end(), p.label);
        if (it != labels.end() && *it == p.label) presig[it - labels.begin()] = p.presigma;
    }
    std::vector<int32_t> var2tot;
    for (int32_t t = 0; t < ntot; t++) if (var[t] > 0) var2tot.push_back(t);

    // ---- xloopn (src/pede.f90:8501-9009) ----
    const int metsol = S.opt.metsol;
    float wolfc2 = (metsol >= 4 && metsol <= 6) ? 0.1f : 0.5f;
    int minf = (metsol >= 4 && metsol <= 6) ? 3 : (metsol == 2 ? 2 : 1);
    const int matrit = std::min(S.matrit, S.mitera);
    int nloopn = 0, iterat = -1, lastit = -1;
    float chicut = S.opt.chicut;
    int64_t nrej_last = 0;
    mpgpu_pass_stats ps_st{};
    std::vector<double> b(nall), x(nall), xls(nvar), corr(nvar);
    std::vector<int64_t> cnt_var(nvar);
    double tp = 0, tsol = 0;
    // times and counts of function passes, matrix passes with their solution, new solutions
    // (the func. value / global matrix / new solution lines of millepede.log)
    double t_cat[3] = {0, 0, 0};
    long long n_cat[3] = {0, 0, 0}, mnrsit = 0;
    std::unique_ptr<FILE, int (*)(FILE *)> fmon_keep(nullptr, std::fclose);   // closed on every exit
    FILE *fmon = nullptr;
    if (S.imonit != 0) {
        fmon = std::fopen((od + "/millepede.mon").c_str(), "w");
        fmon_keep.reset(fmon);
        if (fmon)
            std::fprintf(fmon, "%s\n! LFC     Label   Entries        Median  RMS(MAD)          <error>\n",
                         S.opt.monitor == 2 ? "*** Pulls grouped by first global label (per local fit cycle) ***"
                                            : "*** Normalized residuals grouped by first global label (per local fit cycle) ***");
    }
    auto monres = [&]() {
        if (!fmon) return;
        int32_t ng = 0;
        if (mpgpu_get_monitor(c, &ng, nullptr, nullptr, nullptr) || ng <= 0) return;
        std::vector<int32_t> ml(ng);
        std::vector<float> mr(ng);
        std::vector<int64_t> mh((size_t)ng * S.opt.monitor_bins);
        if (mpgpu_get_monitor(c, &ng, ml.data(), mr.data(), mh.data())) return;
        write_monres(fmon, nloopn, ml, mr, mh, S.opt.monitor_bins);
    };
    auto pass = [&](int icalcm) -> int {
        nloopn++;
        const double tp0 = now_ms();
        struct Acc { double &t; long long &n; double t0; ~Acc() { t += now_ms() - t0; n++; } };
        Acc acc{icalcm == 1 ? t_cat[1] : t_cat[0], icalcm == 1 ? n_cat[1] : n_cat[0], tp0};
        // cut factor of a new iteration (loopn, src/pede.f90:2580-2595)
        if (iterat != lastit && nloopn > 1) {
            if (iterat == 1) chicut = S.opt.chirem;
            else if (iterat > 1) {
                chicut = std::sqrt(chicut);
                if (chicut != 0.0f && chicut < 1.5f) chicut = 1.0f;
                if (chicut != 0.0f && nrej_last == 0) chicut = 1.0f;
            }
        }
        lastit = iterat;
        double t = now_ms();
        int r = mpgpu_loopn(c, icalcm, nloopn, chicut, &ps_st);
        tp += now_ms() - t;
        if (r) return r;
        int64_t nrej = 0;
        for (int k = 0; k < 6; k++) nrej += ps_st.nrej[k];
        nrej_last = nrej;
        if (3 * nrej > ps_st.nrec) { std::fprintf(stderr, "Too many rejects (>33.3%%) - stop\n"); return 26; }
        mpgpu_get_vector(c, b.data(), cnt_var.data());
        if (S.imonit < 0 || (nloopn == 1 && (S.imonit & 1))) monres();   // src/pede.f90:2879
        return 0;
    };
    double fvalue = 0.0;
    LineSearch LS;
    std::vector<double> err_v, gcor_v;
    bool have_err = false;
    Feasibility FS;
    FS.setup(S.cons, labels, var);
    if (sm.ncon > 0) {
        // feasma (src/pede.f90:1494-1507): rank of the product matrix of the constraints; too
        // few independent constraints enforce SUBITO unless 'force'
        both(" Rank of product matrix of constraints is" + ld_i(FS.rank) + "  for" + ld_i(sm.ncon) +
             "  constraint equations");
        if (FS.rank < sm.ncon) {
            both(" Warning: insufficient constraint equations!");
            if (S.iforce == 0) {
                S.isubit = 1;
                both("          --> enforcing SUBITO mode");
            }
        }
    }
    {
        // chi^2 cut table (src/pede.f90:7485-7510), LOOP2 end, outlier treatment (214-222) and
        // the solution algorithm (xloopn, src/pede.f90:8571-8683)
        logl("  ");
        logl("    Cut values of Chi^2/Ndf and Chi2,");
        logl("    corresponding to 2 and 3 standard deviations");
        logl("    Ndf  Chi^2/Ndf(2)  Chi^2(2)     Chi^2/Ndf(3)  Chi^2(3)");
        for (int ndf = 0;;) {
            if (ndf > sm.max_eq_rec) break;
            if (ndf < 10) ndf += 1;
            else if (ndf < 20) ndf += 2;
            else if (ndf < 100) ndf += 5;
            else if (ndf < 200) ndf += 10;
            else break;
            const float c2 = chi2_ndf_cut(2, ndf), c3 = chi2_ndf_cut(3, ndf);
            logl(ld_i(ndf, 6) + "   " + fmt_f(c2, 9, 3) + fmt_f(c2 * (float)ndf, 12, 1) + "      " + fmt_f(c3, 9, 3) +
                 fmt_f(c3 * (float)ndf, 12, 1) + "   ");
        }
        logl(" LOOP2: ending");
        logl("  ");
        if (S.opt.chicut != 0.0f) {
            logl(" Chi square cut equiv 3 st.dev applied ...");
            logl("  in  first iteration with factor" + ld_r4(S.opt.chicut));
            logl("  in second iteration with factor" + ld_r4(S.opt.chirem));
            logl("  (reduced by sqrt in next iterations)");
        }
        if (S.opt.lhuber != 0) {
            logl(" Down-weighting of outliers in" + ld_i(S.opt.lhuber) + "  iterations");
            logl(" Cut on downweight fraction" + ld_r4(S.opt.dwcut));
        }
        const int ms = S.opt.metsol;
        logl("  ");
        logl(" Solution algorithm: ");
        logl(f121("=================================================== "));
        if (ms == MPGPU_INVERSION) logl(f121("solution method:", "matrix inversion"));
        else if (ms == MPGPU_DECOMPOSITION) logl(f121("solution method:", "decomposition"));
        else if (ms == MPGPU_MINRES) logl(f121("solution method:", "minres (Paige/Saunders)"));
        else if (ms == MPGPU_MINRES_QLP) {
            logl(f121("solution method:", "minres-qlp (Choi/Paige/Saunders)"));
            if (S.opt.mrmode == 1) logl(f121(" ", "   using QR factorization"));
            else if (S.opt.mrmode == 2) logl(f121(" ", "   using QLP factorization"));
            else {
                logl(f121(" ", "   using QR and QLP factorization"));
                logl(f123("transition condition", S.opt.mrtcnd));
            }
        } else if (ms == MPGPU_LAPACK_PACKED || ms == MPGPU_LAPACK_UNPACKED) {
            const bool indef = nall_active > sm.nvar;
            logl(f121("solution method:", ms == MPGPU_LAPACK_PACKED
                                              ? (indef ? "LAPACK factorization (DSPTRF)" : "LAPACK factorization (DPPTRF)")
                                              : (indef ? "LAPACK factorization (DSYTRF)" : "LAPACK factorization (DPOTRF)")));
            if (S.opt.lapack_errors) logl(f121(" ", "with error calculation (D??TRI)"));
        } else if (ms == MPGPU_CHOLESKY_FP32_IR) {
            logl(f121("solution method:", "Cholesky, FP32 factor, FP64 refinement"));
        }
        logl(f123("convergence limit at Delta F=", S.dflim));
        logl(f122("maximum number of iterations=", S.mitera));
        if (matrit > 1) logl(f122("matrix recalculation up to ", matrit, ". iteration"));
        if (ms == MPGPU_MINRES || ms == MPGPU_MINRES_QLP) {
            logl(f121("matrix storage:", S.opt.matsto == MPGPU_FULL ? "full" : "sparse"));
            logl(f122("pre-con band-width parameter=", S.opt.mbandw));
            if (S.opt.mbandw == 0) logl(f121("pre-conditioning:", "default"));
            else if (S.opt.mbandw < 0) logl(f121("pre-conditioning:", "none!"));
            else logl(f121("pre-conditioning=", S.opt.lprecm > 0 ? "skyline-matrix (rank preserving)" : "band-matrix"));
        }
        long long npresg = 0;
        for (auto &p : S.params) npresg += p.presigma > 0.0;
        if (S.opt.default_presigma == 0.0 && npresg == 0) {
            logl(f121("using pre-sigmas:", "no"));
        } else {
            logl(" " + fa("pre-sigmas defined for", 40) + "   " +
                 fmt_f((float)(100 * npresg) / (float)sm.nvar, 5, 1) + " % of variable parameters");
            logl(f123("default pre-sigma=", S.opt.default_presigma));
        }
        if (S.opt.nregul == 0) logl(f121("regularization:", "no"));
        else {
            logl(f121("regularization:", "yes"));
            logl(f123("regularization factor=", S.opt.regula));
        }
        if (S.opt.chicut != 0.0f) {
            logl(f121("Chi square cut equiv 3 st.dev applied"));
            logl(f123("... in first iteration with factor", S.opt.chicut));
            logl(f123("... in second iteration with factor", S.opt.chirem));
            logl(f121(" (reduced by sqrt in next iterations)"));
        }
        if (S.scale_given) {
            logl(f121("Scaling of measurement errors applied"));
            logl(f123("... factor for \"global\" measuements", S.opt.scale_errors[0]));
            logl(f123("... factor for \"local\"  measuements", S.opt.scale_errors[1]));
        }
        if (S.opt.lhuber != 0) {
            logl(f122("Down-weighting of outliers in", S.opt.lhuber, " iterations"));
            logl(f123("Cut on downweight fraction", S.opt.dwcut));
        }
    }
    float concut = 1.0e-12f;                     // first check, then 1e-6 (src/pede.f90:8693-8694)
    const float concu2 = 1.0e-6f;
    if (S.nofeas == 0) {
        // feasib (src/pede.f90:8723-8741)
        logl(" Checking feasibility of parameters:");
        if (FS.apply(concut, par, var)) {
            std::printf(" Checking feasibility of parameters: parameters are made feasible\n");
            logl(" Call FEASIB with cut=" + fmt_gd(concut, 10, 3));
            logl("    parameters are made feasible");
            if ((rc = mpgpu_set_parameters(c, par.data()))) return bail(rc);
        } else {
            logl("    parameters are feasible  (i.e. satisfy constraints)");
        }
    }
    concut = concu2;
    int lsinfo = 0, iitera = 0, istopa = 0;
    float angras = 0.0f;
    double rstart = now_ms();

    // The iteration state machine of xloopn (src/pede.f90:8755-9009).  icalcm: 1 = data pass
    // with matrix, then solve; 0 = function pass only (line search); -1 = solve with the
    // matrix of the last matrix pass and the gradient of the last pass.
    int icalcm = 1, jcalcm = -1, litera = 0, iagain = 0, nloopsol = 0;
    bool lsflag = false;
    float delfun = 0.0f;
    iterat = -1;
    // the line of the loop just ended, at the top of the next (src/pede.f90:8764-8794)
    auto loop_line = [&]() {
        if (iterat < 0) return;
        static const char *ccalcm[4] = {" end", "   S", " F  ", " FMS"};
        const float dt = (float)((now_ms() - rstart) / 1000.0);        // sechms, src/pede.f90:5750-5765
        const int nsecd = (int)std::lround(dt);
        const int nhour = nsecd / 3600, minut = nsecd / 60 - 60 * nhour;
        const float secnd = dt - 60.0f * (float)(minut + 60 * nhour);
        char tm[32];
        std::snprintf(tm, sizeof tm, "%3d%02d%02d%s", nhour, minut, (int)std::lround(secnd), ccalcm[jcalcm + 2]);
        const std::string nrej = ld_i(std::min<int64_t>(nrej_last, 9999999), 7);
        auto i3 = [](int v) { return ld_i(v, 3); };
        float slopes[3], steps[3];
        std::string l;
        if (jcalcm + 1 != 0) {
            if (iterat == 0) {
                both("  ");
                both(" it fc   fcn_value dfcn_exp  slpr costh  iit st ls  step cutf rejects hhmmss FMS");
                both(" -- -- ----------- --------  ---- -----  --- -- -- ----- ---- ------- ------ ---");
                l = i3(iterat) + i3(nloopn) + fmt_e(fvalue, 12, 5) + std::string(38, ' ') + fmt_f(chicut, 5, 1) + " " +
                    nrej + tm;
                iterat = 1;
            } else if (iterat != litera) {
                if (lsinfo == 10) {
                    l = i3(iterat) + i3(nloopn) + fmt_e(fvalue, 12, 5) + " " + fmt_e(delfun, 8, 2) + std::string(12, ' ') +
                        ld_i(iitera, 5) + i3(istopa) + std::string(9, ' ') + fmt_f(chicut, 5, 1) + " " + nrej + tm;
                } else {
                    LS.opt(slopes, steps);
                    const float ratae = std::max(-99.9f, std::min(99.9f, slopes[1] / slopes[0]));
                    l = i3(iterat) + i3(nloopn) + fmt_e(fvalue, 12, 5) + " " + fmt_e(delfun, 8, 2) + fmt_f(ratae, 6, 3) +
                        fmt_f(angras, 6, 3) + ld_i(iitera, 5) + i3(istopa) + i3(lsinfo) + fmt_f(steps[1], 6, 3) +
                        fmt_f(chicut, 5, 1) + " " + nrej + tm;
                }
                litera = iterat;
            } else if (lsinfo == 10) {
                l = "   " + i3(nloopn) + fmt_e(fvalue, 12, 5) + std::string(44, ' ') + nrej + tm;
            } else {
                LS.opt(slopes, steps);
                l = "   " + i3(nloopn) + fmt_e(fvalue, 12, 5) + std::string(9, ' ') + fmt_f(std::fabs(slopes[1] / slopes[0]), 6, 3) +
                    std::string(14, ' ') + i3(lsinfo) + fmt_f(steps[1], 6, 3) + std::string(6, ' ') + nrej + tm;
            }
        } else {
            l = std::string(69, ' ') + tm;
        }
        both(l);
        rstart = now_ms();
    };
    for (;;) {
        loop_line();
        jcalcm = icalcm;
        if (icalcm >= 0) {                        // block 1: data pass (b = gradient at par)
            if ((rc = pass(icalcm))) return bail(rc);
            fvalue = ps_st.chi2;
        }
        if (std::abs(icalcm) == 1) {              // block 2: new solution
            for (int32_t v = 0; v < nvar; v++) xls[v] = par[var2tot[v]];
            iterat++;
            double t = now_ms();
            mpgpu_solve_stats sst{};
            if ((rc = mpgpu_solve(c, b.data(), x.data(), &sst))) return bail(rc);
            const double dt_sol = now_ms() - t;
            tsol += dt_sol;
            if (icalcm == 1) t_cat[1] += dt_sol;
            else { t_cat[2] += dt_sol; n_cat[2]++; }
            if (metsol == MPGPU_MINRES || metsol == MPGPU_MINRES_QLP) mnrsit += sst.itn;
            if (metsol == MPGPU_DECOMPOSITION && sst.nrank < nall) {
                // mchdec: a rank defect of the decomposition enforces SUBITO unless 'force'
                // (src/pede.f90:8026-8037)
                char w[160];
                std::snprintf(w, sizeof w, " Warning: the rank defect of the symmetric%s-by-%s matrix is %s (should be zero).",
                              ld_i(nall).c_str(), ld_i(nall).c_str(), ld_i(nall - sst.nrank).c_str());
                both(w);
                if (S.iforce == 0 && S.isubit == 0) {
                    S.isubit = 1;
                    both("          --> enforcing SUBITO mode");
                }
            } else if (icalcm == 1 && metsol != MPGPU_MINRES && metsol != MPGPU_MINRES_QLP) {
                // a factorization or inversion of full rank (src/pede.f90:7923-7928, 8031-8038;
                // src/mplapack.f90:143-148)
                logl(" No rank defect of the symmetric matrix");
                if (metsol == MPGPU_DECOMPOSITION) {
                    logl(" " + std::string("   largest  diagonal element (LDLt)") + ld_r8(sst.evmax));
                    logl(" " + std::string("   smallest diagonal element (LDLt)") + ld_r8(sst.evmin));
                }
            }
            nloopsol = nloopn;
            if (metsol == MPGPU_MINRES || metsol == MPGPU_MINRES_QLP) { iitera = sst.itn; istopa = sst.istop; }
            if (metsol == MPGPU_INVERSION ||
                ((metsol == MPGPU_LAPACK_PACKED || metsol == MPGPU_LAPACK_UNPACKED) && S.opt.lapack_errors)) {
                err_v.assign(nvar, 0.0);
                gcor_v.assign(nvar, 0.0);
                have_err = mpgpu_get_errors(c, err_v.data(), gcor_v.data()) == 0;
            }
            // make the step feasible: apply it, correct the parameters, take the difference
            std::vector<double> trial = par;
            for (int32_t v = 0; v < nvar; v++) trial[var2tot[v]] += x[v];
            FS.apply(concut, trial, var);
            concut = concu2;
            for (int32_t v = 0; v < nvar; v++) corr[v] = trial[var2tot[v]] - par[var2tot[v]];
            double db = 0, db1 = 0, db2 = 0;
            for (int32_t v = 0; v < nvar; v++) { db += corr[v] * b[v]; db1 += corr[v] * corr[v]; db2 += b[v] * b[v]; }
            delfun = (float)db;
            angras = (float)(db / std::sqrt(db1 * db2));
            double dbsig = 16.0 * std::sqrt(std::max(db1, db2)) * 2.220446049250313e-16;
            lsflag = (S.lsearch > 2 || (S.lsearch == 2 && chicut < 2.25f) ||
                      (S.lsearch == 1 && chicut < 2.25f && (delfun <= S.dflim || iterat >= S.mitera)));
            lsflag = lsflag && (db > dbsig);
            if (lsflag) {
                LS.define(wolfc2, 10.0f, minf, 10);
                wolfc2 = (metsol >= 4 && metsol <= 6) ? 0.1f : 0.5f;
                minf = (metsol >= 4 && metsol <= 6) ? 4 : 3;
            }
            if (db <= -dbsig) {                   // function not decreasing
                std::printf(" Function not decreasing: %g\n", db);
                if (db > -1.0e-3 && ++iagain <= 1) { icalcm = 1; continue; }
                if (db <= -1.0e-3) iagain = -1;
                break;
            }
            iagain = 0;
            icalcm = 0;
        }
        // block 3: line search step (or the full step)
        int info = 10;
        if (lsflag) {
            double step = 0, flines = 0.5 * fvalue;
            info = LS.line(nvar, xls.data(), flines, b.data(), corr.data(), step);
        } else if (nloopn == nloopsol) {
            for (int32_t v = 0; v < nvar; v++) xls[v] += corr[v];
        }
        lsinfo = info;
        int nan = 0;
        for (int32_t v = 0; v < nvar; v++) {
            if (!(xls[v] <= 0.0) && !(xls[v] > 0.0)) nan++;
            par[var2tot[v]] = xls[v];
        }
        if (nan > 0) {
            std::fprintf(stderr, "Result vector contains %d NaNs - stop\n", nan);
            mpgpu_destroy(c);
            return finish(25);
        }
        if ((rc = mpgpu_set_parameters(c, par.data()))) return bail(rc);
        if (S.isubit) { std::printf(" Subito!     Exit after first step.\n"); break; }
        if (info == 0 && iagain <= 0) { icalcm = 1; continue; }
        if (info < 0 || nloopn == nloopsol) continue;   // icalcm = 0: function pass next
        // block 4: line search converged
        LS.print(flog, lsinfo);
        int iact = FS.apply(concut, par, var);
        if (iact && (rc = mpgpu_set_parameters(c, par.data()))) return bail(rc);
        if (iact || chicut > 1.0f) { icalcm = iterat < matrit ? 1 : -1; continue; }
        if (delfun <= S.dflim || iterat >= S.mitera) break;
        icalcm = iterat < matrit ? 1 : -1;
    }
    loop_line();
    loc.final_chi2 = ps_st.chi2;
    loc.final_ndf = ps_st.ndf;
    loc.iterations = iterat;
    loc.ms_passes = tp;
    loc.ms_solves = tsol;
    {
        // final chi^2 per degree of freedom (src/pede.f90:9124-9147); pede counts the fit
        // parameters as nvar - ncon with eliminated constraints, nvar + ncon with multipliers
        const double dwmean = ps_st.ndf > 0 ? ps_st.wndf / (double)ps_st.ndf : 1.0;
        const double dratio = fvalue / dwmean / (double)(ps_st.ndf - nfitpar);
        float catio = (float)dratio;
        if (nloopn != 1 && S.opt.lhuber != 0) catio = catio / 0.9326f;
        both("  ");
        both(" Sum(Chi^2)/Sum(Ndf) =" + ld_r8(fvalue));
        both("                     / (" + ld_i(ps_st.ndf, 21) + " -" + ld_i(nfitpar) + " )");
        both("                     =" + ld_r8(dratio));
        both("  ");
        if (nloopn != 1 && S.opt.lhuber != 0) both("             with correction for down-weighting   " + ld_r4(catio));
    }
    if ((metsol == MPGPU_MINRES || metsol == MPGPU_MINRES_QLP) && !S.errlabels.empty()) {
        // errors and global correlations of selected parameters from MINRES (solglo/solgloqlp,
        // src/pede.f90:537-701, called at 9288-9306): the row of the inverse from A x = e_i
        // (at most 200 iterations), error sqrt|x_i| (negative for x_i < 0), gcor^2 = 1 - 1/(x_i A_ii);
        // with eliminated constraints the row is that of the constrained system
        mpgpu_options o2 = S.opt;
        o2.itnlim = 200;
        if ((rc = mpgpu_set_options(c, &o2))) return bail(rc);
        std::vector<double> e(nall, 0.0), xe(nall), ye(nall);
        bool head = false;
        for (int32_t lab : S.errlabels) {
            auto it = std::lower_bound(labels.begin(), labels.end(), lab);
            if (lab == 0 || it == labels.end() || *it != lab) continue;
            const int64_t t = it - labels.begin();
            const int32_t v = var[t];
            if (v <= 0) continue;
            e[v - 1] = 1.0;
            mpgpu_solve_stats s2{};
            rc = mpgpu_solve(c, e.data(), xe.data(), &s2);
            if (!rc) rc = mpgpu_avprod(c, e.data(), ye.data());
            e[v - 1] = 0.0;
            if (rc) return bail(rc);
            const double gmati = xe[v - 1], diag = ye[v - 1];
            float err = std::sqrt(std::fabs((float)gmati));
            if (gmati < 0.0) err = -err;
            const float gcor2 = (float)(1.0 - 1.0 / (gmati * diag));
            if (!head) {
                std::printf(" %s\n %s\n", "    label     parameter    presigma      differ       Error gcor^2   iit",
                            "---------  -------------------------------------------------------  ----");
                head = true;
            }
            std::printf("%10d  %s%s%s%s%7.4f%6d\n", lab, fmt_g((float)par[t], 12, 4).c_str(),
                        fmt_g((float)presig[t], 12, 4).c_str(), fmt_g((float)(par[t] - start[t]), 12, 4).c_str(),
                        fmt_g(err, 12, 4).c_str(), gcor2, s2.itn);
        }
    }
    if (S.imonit > 0 && (S.imonit & 2)) monres();                      // src/pede.f90:9011
    fmon_keep.reset();
    fmon = nullptr;
    rc = write_res(od + "/millepede.res", labels, var, par, start, presig, cnt_var, cnt_tot, have_err ? &err_v : nullptr,
                   (have_err && S.igcorr) ? &gcor_v : nullptr, S.ipcntr, S.opt.mreqena);
    {
        // loop counts, times, peak device memory and rejects (src/pede.f90:455-489, prtrej 4448-4461)
        logl("  ");
        const bool iter3 = S.opt.iteren > S.opt.mreqenf && S.opt.mcount == 0;
        logl(std::string(iter3 ? " In total 3 +" : " In total 2 +") + ld_i(nloopn) + "  loops through the data files");
        if (mnrsit > 0) {
            logl("  ");
            logl(" In total    " + ld_i(mnrsit) + "  internal MINRES iterations");
        }
        auto sec = [](double ms) { return fmt_f(ms / 1000.0, 12, 3); };
        auto cnt = [](long long n) { return rjust(std::to_string(n) + ".", 4); };   // F4.0
        logl(" Times [in sec] for     text processing" + sec(t_text));
        logl("                                  LOOP1" + sec(t_loop1));
        logl("                                  LOOP2" + sec(t_loop2));
        logl("   func. value                         " + sec(t_cat[0]) + " *" + cnt(n_cat[0]));
        logl("   func. value, global matrix, solution" + sec(t_cat[1]) + " *" + cnt(n_cat[1]));
        logl("                           new solution" + sec(t_cat[2]) + " *" + cnt(n_cat[2]));
        logl("");
        const double dt = (now_ms() - t_start) / 1000.0;
        const long long ntsec = std::llround(dt);
        const long long nhour = ntsec / 3600, minut = ntsec / 60 - 60 * nhour;
        const long long nsec = std::llround(dt - 60.0 * (double)(minut + 60 * nhour));
        logl(" Total time =" + ld_i(ntsec) + "  seconds =" + ld_i(nhour) + "  h" + ld_i(minut) + "  m" + ld_i(nsec) +
             "  seconds");
        const std::time_t now = std::time(nullptr);
        char date[32] = "";
        std::strftime(date, sizeof date, "%a %b %d %H:%M:%S %Y", std::localtime(&now));
        logl(std::string(" end                                            ") + date);
        // peak of the device memory taken from the stream-ordered pool during the run
        hipMemPool_t pool = nullptr;
        uint64_t high = 0;
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool &&
            hipMemPoolGetAttribute(pool, hipMemPoolAttrUsedMemHigh, &high) == hipSuccess) {
            logl("  ");
            logl("      Peak dynamic memory allocation: " + fmt_f(1.0e-9 * (double)high, 11, 6) + " GB");
        }
        int64_t nrejt = 0;
        for (int k = 0; k < 6; k++) nrejt += ps_st.nrej[k];
        if (nrejt != 0) {
            static const char *why[6] = {" (local solution contains NaNs)", " (local matrix with rank deficit)",
                                         " (local matrix with ill condition)", " (local fit with Ndf=0)",
                                         " (local fit with huge Chi2(Ndf))", " (local fit with large Chi2(Ndf))"};
            logl("  ");
            logl(" Data records rejected in last iteration:   ");
            for (int k = 0; k < 6; k++)
                if (ps_st.nrej[k] > 0) logl(ld_i(ps_st.nrej[k], 21) + " " + why[k]);
            logl("  ");
        }
    }
    mpgpu_destroy(c);
    if (rc) return finish(rc);
    return finish(0);
}
