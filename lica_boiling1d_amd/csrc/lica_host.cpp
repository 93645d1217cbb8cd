// Host side of the C ABI (include/lica_mi355x.h): deck parsing, per-case
// geometry, the device layout of a batch, the step loop and the getters.
//
// Geometry follows READGRID/GEOMETRIC/GEOMETRIC_LVS/GEOM_COUPLING/LHSINIT
// (1d.f:4746-4805, 4838-4942, 5134-5369, 1963-1978) for the configuration
// lica.in/grid.in describe: one NS cell in x and y (periodic), N cells in z,
// a refined level-set grid r x r x rN, all grids uniform.  grid.dat is the
// ES23.15 print of grid.f's UNIFORM nodes (grid.f:240-242, 326-346); the
// round trip through that text format is part of the reference's geometry
// and is reproduced here.
#include "../../include/lica_mi355x.h"
#include "lica_types.h"
#include "lica_launch.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#ifdef LICA_EMUL
int lica_emul_nthreads = 256;
int lica_emul_depth = 0;
typedef void *hipStream_t;
// the warm-start network runs on MFMA only; the CPU logic-check build has none
int lica_net_guess(lica_net *, const NetCase *, const int *, int, double *, double, void *) { return LICA_UNSUPPORTED_DECK; }
int lica_net_cast(lica_net *, void *) { return LICA_UNSUPPORTED_DECK; }
int lica_net_append(lica_net *, const NetPairSrc *, int, int64_t, int64_t, uint64_t, int64_t *)
{
    return LICA_UNSUPPORTED_DECK;
}
int64_t lica_net_window() { return 256; }
extern "C" {
int lica_net_data_size(lica_net *, int64_t *, int64_t *) { return LICA_UNSUPPORTED_DECK; }
int lica_net_get_data(lica_net *, float *, float *) { return LICA_UNSUPPORTED_DECK; }
int64_t lica_net_param_count(void) { return 0; }
int lica_net_create(uint64_t, lica_net **out) { if (out) *out = nullptr; return LICA_UNSUPPORTED_DECK; }
void lica_net_destroy(lica_net *) {}
int lica_net_get_params(lica_net *, float *) { return LICA_UNSUPPORTED_DECK; }
int lica_net_set_params(lica_net *, const float *) { return LICA_UNSUPPORTED_DECK; }
int lica_net_forward(lica_net *, const float *, const int64_t *, int64_t, float *) { return LICA_UNSUPPORTED_DECK; }
int lica_net_grad(lica_net *, const float *, const float *, const int64_t *, int64_t, float *, double *)
{
    return LICA_UNSUPPORTED_DECK;
}
int lica_net_set_data(lica_net *, const float *, const float *, int64_t, int64_t) { return LICA_UNSUPPORTED_DECK; }
int lica_net_train(lica_net *, int64_t, int64_t, double, uint64_t, double *) { return LICA_UNSUPPORTED_DECK; }
}
#else
#include <hip/hip_runtime.h>
#endif

namespace {

// ---------------------------------------------------------------------
// device memory (the CPU logic-check build keeps "device" buffers in host
// memory; the product build is HIP only)
#ifdef LICA_EMUL
int dmalloc(void **p, size_t n) { *p = calloc(1, n ? n : 1); return *p ? 0 : LICA_HIP_ERROR; }
void dfree(void *p) { free(p); }
int h2d(void *d, const void *h, size_t n, hipStream_t) { memcpy(d, h, n); return 0; }
int d2h(void *h, const void *d, size_t n, hipStream_t) { memcpy(h, d, n); return 0; }
int d2d(void *d, const void *s, size_t n, hipStream_t) { memmove(d, s, n); return 0; }
int dsync(hipStream_t) { return 0; }
int have_device() { return 1; }
#else
// zeroed device memory.  hipMemset runs on the null stream, asynchronously
// to the host, and the batch streams are non-blocking (not ordered after the
// null stream): without the wait a late fill could overwrite what a batch
// stream has uploaded or computed in the buffer since.
int dmalloc(void **p, size_t n)
{
    if (hipMalloc(p, n ? n : 1) != hipSuccess) { *p = nullptr; return LICA_HIP_ERROR; }
    if (hipMemset(*p, 0, n ? n : 1) != hipSuccess) return LICA_HIP_ERROR;
    return hipStreamSynchronize(nullptr) == hipSuccess ? 0 : LICA_HIP_ERROR;
}
void dfree(void *p) { if (p) (void)hipFree(p); }
int h2d(void *d, const void *h, size_t n, hipStream_t s)
{
    if (hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, s) != hipSuccess) return LICA_HIP_ERROR;
    return hipStreamSynchronize(s) == hipSuccess ? 0 : LICA_HIP_ERROR;
}
int d2h(void *h, const void *d, size_t n, hipStream_t s)
{
    if (hipMemcpyAsync(h, d, n, hipMemcpyDeviceToHost, s) != hipSuccess) return LICA_HIP_ERROR;
    return hipStreamSynchronize(s) == hipSuccess ? 0 : LICA_HIP_ERROR;
}
// device-to-device copy, ordered on the stream (no wait)
int d2d(void *d, const void *s, size_t n, hipStream_t st)
{
    return hipMemcpyAsync(d, s, n, hipMemcpyDeviceToDevice, st) == hipSuccess ? 0 : LICA_HIP_ERROR;
}
int dsync(hipStream_t s) { return hipStreamSynchronize(s) == hipSuccess ? 0 : LICA_HIP_ERROR; }
int have_device()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return 0;
    return 1;
}
#endif

// grid.f UNIFORM in REAL*8 printed ES23.15 and read back (grid.f:242)
double es23_15(double v)
{
    char buf[64];
    snprintf(buf, sizeof buf, "%.15E", v);
    return strtod(buf, nullptr);
}

void uniform_nodes(int64_t ncell, double len, std::vector<double> &z)
{
    z.resize((size_t)ncell + 1);
    double dc = (len - 0.0) / (double)((ncell + 1) - 1);
    for (int64_t k = 1; k <= ncell + 1; k++) z[(size_t)k - 1] = es23_15(0.0 + (double)(k - 1) * dc);
}

int64_t round_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

// per-case host image of the device blocks
struct CaseImage {
    InstParam p;
    std::vector<double> ns, rf;
    std::vector<int> nsi, rfi;
    double &A(int f, int k) { return ns[(size_t)f * p.ln + k + 2]; }
    double &R(int f, int k) { return rf[(size_t)f * p.lf + k - p.m3l]; }
    int &AI(int f, int k) { return nsi[(size_t)f * p.lni + k + 2]; }
    int &RI(int f, int k) { return rfi[(size_t)f * p.lfi + k]; }
};

// first k in [lo, hi] where the monotone predicate hit(k) holds (hi + 1 if none)
template <class F>
int first_hit(int lo, int hi, F hit)
{
    while (lo <= hi) {
        const int mid = lo + (hi - lo) / 2;
        if (hit(mid)) hi = mid - 1;
        else lo = mid + 1;
    }
    return lo;
}

// GEOM_COUPLING's search of the NS centre at or below zz (1d.f:5156-5177).
// The reference scans kk = 1, 2, ... and stops at the first kk with
// |zz - ZP(kk)| <= 1e-10 or ZP(kk) > zz (the two end tests can only fire at
// kk = 1); with ZP increasing that predicate stays true once true, so the
// first such kk is found by bisection and the result is the scan's.
int coump_vel_search(CaseImage &c, double zz)
{
    const int n = c.p.n;
    auto zp = [&](int kk) { return c.A(NS_ZP, kk); };
    if (fabs(zz - zp(1)) <= 1.0e-10) return 1;
    if (zp(1) > zz) return 1;
    if (zp(n) < zz) return n;
    auto hit = [&](int kk) { return fabs(zz - zp(kk)) <= 1.0e-10 || zp(kk) > zz; };
    const int kk = first_hit(2, n, hit);
    if (kk > n) return -1;
    return (fabs(zz - zp(kk)) <= 1.0e-10) ? kk : kk - 1;
}

// refined node search of GEOM_COUPLING (1d.f:5190-5240): (kf-1,kf) on a node;
// the scan's first kf with |ZF(kf) - zc| <= 1e-10 or ZF(kf) > zc, by bisection
int node_search(CaseImage &c, double zc, int *k1, int *k2)
{
    const int n3f = c.p.n3f;
    auto hit = [&](int kf) { const double zf = c.R(RF_ZF, kf); return fabs(zf - zc) <= 1.0e-10 || zf > zc; };
    const int kf = first_hit(1, n3f, hit);
    if (kf > n3f) return -1;
    if (fabs(c.R(RF_ZF, kf) - zc) <= 1.0e-10) { *k1 = kf - 1; *k2 = kf; }
    else { *k1 = kf - 1; *k2 = kf - 1; }
    return 0;
}

int build_case(const lica_case_desc &d, CaseImage &c)
{
    if (d.n < 8 || d.r < 1 || d.r > 8 || d.n > (1 << 22) || !(d.zl > 0.0) || !(d.re > 0.0) || !(d.prm > 0.0) ||
        !(d.dt0 > 0.0) || !(d.denr > 0.0) || d.denr == 1.0)
        return LICA_BAD_ARG;
    InstParam &P = c.p;
    memset(&P, 0, sizeof P);
    const int n = (int)d.n, r = (int)d.r;
    P.n = n; P.r = r; P.n3 = n + 1; P.n3fm = r * n; P.n3f = P.n3fm + 1; P.m3l = -2 * r;
    P.ln = (int)round_up(n + 8, 16);
    P.lni = P.ln;
    P.lf = (int)round_up((int64_t)P.n3f + 4 * r + 4, 16);
    P.lfi = (int)round_up((int64_t)P.n3f + 2, 16);
    P.mf_band = (int)(int64_t)(0.1 * (double)((int64_t)r * r * P.n3fm));   // MF_BAND=0.1*MF
    P.idtopt = (int)d.idtopt;
    // INPUT_SIMULATION (1d.f:434-480)
    const double pi = acos(-1.0);
    P.re = d.re; P.prm = d.prm; P.denr = d.denr; P.ja = d.ja; P.tsat = d.tsat;
    P.cflmax = d.cflmax; P.dt0 = d.dt0; P.z0 = d.z0; P.tslope = d.tslope;
    P.rei = 1.0 / d.re;
    P.denm = 1.0; P.denp = d.denr; P.den_diff = P.denp - P.denm;
    P.vism = 1.0 * P.rei; P.visp = d.visr * P.rei;
    P.denscm = P.denm * 1.0;
    double denscp = P.denscm * (d.denr * d.scr);
    P.densc_diff = denscp - P.denscm;
    P.tcm = 1.0; P.tcp = d.tcr; P.tc_diff = P.tcp - P.tcm;
    P.prai = 1.0 / d.prm;
    double surf = (d.surf_j == 0.0) ? 0.0 : 1.0 / d.surf_j;
    double gravity = (d.fr == 0.0) ? 0.0 : 1.0 / (d.fr * d.fr);
    P.time_st = surf * pi / (P.denm + P.denp);
    P.time_gv = P.den_diff / (P.denm + P.denp) * gravity / pi;
    // RK3COEF (1d.f:1314-1319)
    P.gamma[0] = 8.0 / 15.0; P.gamma[1] = 5.0 / 12.0; P.gamma[2] = 3.0 / 4.0;
    P.ro[0] = 0.0; P.ro[1] = -17.0 / 60.0; P.ro[2] = -5.0 / 12.0;

    c.ns.assign((size_t)NS_NF * P.ln, 0.0);
    c.rf.assign((size_t)RF_NF * P.lf, 0.0);
    c.nsi.assign((size_t)NSI_NF * P.lni, 0);
    c.rfi.assign((size_t)RFI_NF * P.lfi, 0);

    std::vector<double> zn, zfn, xfn;
    uniform_nodes(n, d.zl, zn);
    uniform_nodes(P.n3fm, d.zl, zfn);
    P.xl = es23_15(1.0); P.yl = es23_15(1.0);
    uniform_nodes(r, P.xl, xfn);
    P.sdx1 = P.xl - 0.0; P.sdy1 = P.yl - 0.0;
    P.xp1 = 0.0 + 0.5 * P.sdx1; P.yp1 = 0.0 + 0.5 * P.sdy1;
    P.sdxf = xfn[1] - xfn[0]; P.sdyf = P.sdxf;
    std::vector<double> xpf((size_t)r + 1);
    for (int i = 1; i <= r; i++) xpf[(size_t)i] = xfn[(size_t)i - 1] + 0.5 * P.sdxf;

    const int n3 = P.n3;
    for (int k = 1; k <= n3; k++) c.A(NS_Z, k) = zn[(size_t)k - 1];
    c.A(NS_Z, 0) = c.A(NS_Z, 1);
    // GEOMETRIC (1d.f:4838-4880)
    for (int k = 1; k <= n; k++) c.A(NS_SDZ, k) = c.A(NS_Z, k + 1) - c.A(NS_Z, k);
    for (int k = 2; k <= n; k++) c.A(NS_VDZ, k) = 0.5 * (c.A(NS_SDZ, k) + c.A(NS_SDZ, k - 1));
    c.A(NS_VDZ, 1) = 0.5 * c.A(NS_SDZ, 1);
    c.A(NS_VDZ, n3) = 0.5 * c.A(NS_SDZ, n);
    for (int k = 1; k <= n; k++) c.A(NS_SSDZ, k) = 1.0 / c.A(NS_SDZ, k);
    for (int k = 1; k <= n3; k++) c.A(NS_VVDZ, k) = 1.0 / c.A(NS_VDZ, k);
    c.A(NS_SDZ, 0) = 0.0; c.A(NS_SDZ, n3) = 0.0;
    for (int k = 1; k <= n; k++) c.A(NS_ZP, k) = c.A(NS_Z, k) + 0.5 * c.A(NS_SDZ, k);
    c.A(NS_ZP, n3) = c.A(NS_Z, n3);
    c.A(NS_ZP, 0) = c.A(NS_Z, 1);
    // LHSINIT (1d.f:1963-1978)
    for (int k = 2; k <= n; k++) {
        c.A(NS_AKW, k) = -c.A(NS_VVDZ, k) * c.A(NS_SSDZ, k - 1);
        c.A(NS_CKW, k) = -c.A(NS_VVDZ, k) * c.A(NS_SSDZ, k);
    }
    for (int k = 1; k <= n; k++) {
        c.A(NS_AKUV, k) = -c.A(NS_VVDZ, k) * c.A(NS_SSDZ, k);
        c.A(NS_CKUV, k) = -c.A(NS_VVDZ, k + 1) * c.A(NS_SSDZ, k);
    }
    // GEOMETRIC_LVS (1d.f:4885-4942)
    const int n3f = P.n3f, n3fm = P.n3fm;
    for (int k = 1; k <= n3f; k++) c.R(RF_ZF, k) = zfn[(size_t)k - 1];
    P.sdzf = c.R(RF_ZF, 2) - c.R(RF_ZF, 1);
    P.ssdzf = 1.0 / P.sdzf;
    for (int k = 1; k <= n3fm; k++) c.R(RF_ZPF, k) = c.R(RF_ZF, k) + 0.5 * P.sdzf;
    for (int k = 0; k <= 2 * r; k++) {
        c.R(RF_ZF, -k) = c.R(RF_ZF, 1) - P.sdzf * (double)(k + 1);
        c.R(RF_ZF, n3f + 1 + k) = c.R(RF_ZF, n3f) + P.sdzf * (double)(k + 1);
    }
    for (int k = 0; k <= 2 * r; k++) {
        c.R(RF_ZPF, -k) = c.R(RF_ZF, -k) + 0.5 * P.sdzf;
        c.R(RF_ZPF, n3f + k) = c.R(RF_ZF, n3f + k) + 0.5 * P.sdzf;
    }
    // running volume of an all-vapour column prefix (serial sum of DVOL)
    {
        double v = 0.0;
        c.R(RF_VPREFIX, 0) = 0.0;
        for (int j = 1; j <= n3fm; j++) { v = v + (1.0 - 0.0) * P.sdzf; c.R(RF_VPREFIX, j) = v; }
    }
    // GEOM_COUPLING (1d.f:5134-5369)
    for (int k = 1; k <= n3f; k++) {
        int a = coump_vel_search(c, c.R(RF_ZF, k));
        int b = coump_vel_search(c, c.R(RF_ZPF, k));
        if (a < 0 || b < 0) return LICA_GRID_TOO_LARGE;
        c.RI(RFI_KCOU_VEL, k) = a;
        c.RI(RFI_KCOUMP_VEL, k) = b;
    }
    for (int k = 1; k <= n3; k++) {
        int k1, k2;
        if (node_search(c, c.A(NS_Z, k), &k1, &k2)) return LICA_GRID_TOO_LARGE;
        c.AI(NSI_KCOU1, k) = k1; c.AI(NSI_KCOU2, k) = k2;
    }
    for (int f : {NSI_KCOU1, NSI_KCOU2}) {
        c.AI(f, 0) = c.AI(f, 1) - (c.AI(f, n) - c.AI(f, n - 1));
        c.AI(f, -1) = c.AI(f, 0) - (c.AI(f, n - 1) - c.AI(f, n - 2));
        c.AI(f, n3 + 1) = c.AI(f, n3) + (c.AI(f, 2) - c.AI(f, 1));
        c.AI(f, n3 + 2) = c.AI(f, n3 + 1) + (c.AI(f, 3) - c.AI(f, 2));
    }
    for (int k = 1; k <= n; k++) {
        int k1, k2;
        if (node_search(c, c.A(NS_ZP, k), &k1, &k2)) return LICA_GRID_TOO_LARGE;
        c.AI(NSI_KCOUMP1, k) = k1; c.AI(NSI_KCOUMP2, k) = k2;
    }
    for (int f : {NSI_KCOUMP1, NSI_KCOUMP2}) {
        c.AI(f, 0) = c.AI(f, 1) - (c.AI(f, n) - c.AI(f, n - 1));
        c.AI(f, -1) = c.AI(f, 0) - (c.AI(f, n - 1) - c.AI(f, n - 2));
        c.AI(f, n3) = c.AI(f, n) + (c.AI(f, n) - c.AI(f, n - 1));
        c.AI(f, n3 + 1) = c.AI(f, n3) + (c.AI(f, 2) - c.AI(f, 1));
    }
    // PHINSINTP x/y weights (chphase.f:1274-1289): I1=1, I2 = refined node of XP(1)
    {
        int i1 = 1, i2 = -1;
        for (int kf = 1; kf <= r + 1; kf++) {
            if (fabs(xfn[(size_t)kf - 1] - P.xp1) <= 1.0e-10) { i2 = kf; break; }
            else if (xfn[(size_t)kf - 1] > P.xp1) { i2 = kf - 1; break; }
        }
        if (i2 < 1) return LICA_GRID_TOO_LARGE;
        P.phins_xd = (i1 == i2) ? 0.0 : (P.xp1 - xpf[(size_t)i1]) / (xpf[(size_t)i2] - xpf[(size_t)i1]);
        P.phins_yd = P.phins_xd;
    }
    // outer-layer column distance of TEMP_GRAD (chphase.f:198-211): min DX^2+DY^2
    {
        double best = -1.0;
        for (int j = 1; j <= r; j++)
            for (int i = 1; i <= r; i++) {
                double dx = xpf[(size_t)i] - P.xp1, dy = xpf[(size_t)j] - P.yp1;
                double d2 = dx * dx + dy * dy;
                if (best < 0.0 || d2 < best) best = d2;
            }
        P.dxy2_min = best;
    }
    return LICA_OK;
}

// ---------------------------------------------------------------------
// deck reading: list-directed records, one per line (1d.f:347-386, grid.f:29-53)
struct LineReader {
    std::vector<std::string> lines;
    size_t pos = 0;
    bool open(const char *path)
    {
        FILE *f = fopen(path, "r");
        if (!f) return false;
        char buf[4096];
        while (fgets(buf, sizeof buf, f)) lines.emplace_back(buf);
        fclose(f);
        return true;
    }
    bool skip() { if (pos >= lines.size()) return false; pos++; return true; }
    // numeric fields of the next record (Fortran D exponents accepted)
    bool nums(int want, double *out)
    {
        if (pos >= lines.size()) return false;
        std::string s = lines[pos++];
        for (char &ch : s) if (ch == 'd' || ch == 'D') ch = 'e';
        for (char &ch : s) if (ch == ',' || ch == '\t') ch = ' ';
        const char *p = s.c_str();
        for (int i = 0; i < want; i++) {
            char *end;
            while (*p == ' ') p++;
            double v = strtod(p, &end);
            if (end == p) return false;
            out[i] = v;
            p = end;
        }
        return true;
    }
};

// most case sets a batch's step can be split into (one stream each)
constexpr int LICA_MAX_LANES = 8;

// the split over L lanes: lane l holds positions l, l + L, l + 2L, ... of the
// largest-first order (n >= L cases), stored lane after lane
int lane_count(int64_t n, int L, int l) { return (int)((n - l + L - 1) / L); }
int lane_first(int64_t n, int L, int l)
{
    int s = 0;
    for (int j = 0; j < l; j++) s += lane_count(n, L, j);
    return s;
}

// launch shape of the kernels on one set of cases: its largest case, and
// tuned threads per workgroup per kernel class and multigrid LDS (0 and -1:
// the default rule).  The tuned classes (DETERMINE_DT, the energy and
// momentum stages, the POISSON solves, UCALC) do not change a result with
// their shape: every reduction in them is a maximum or in an order set by
// the case, the partitioned solves' chunks are set by the row count, and
// where the multigrid vectors live does not change their arithmetic.  The
// band kernels and the warm-start base keep fixed shapes (one or two waves
// per case; a projection whose sums are ordered by the workgroup size).
static_assert(KL_COUNT == LICA_KERNEL_CLASSES, "kernel classes of the tuning table");
struct LaneCfg {
    int nmax = 0;
    int thr[KL_COUNT] = {};
    long long mg_lds = -1;
};

}  // namespace

// ---------------------------------------------------------------------
struct lica_batch {
    int64_t ninst = 0;
    std::vector<lica_case_desc> cases;
    std::vector<InstParam> par;
    size_t ns_total = 0, rf_total = 0, nsi_total = 0, rfi_total = 0;
    InstParam *d_par = nullptr;
    InstState *d_st = nullptr;
    double *d_ns = nullptr, *d_rf = nullptr;
    int *d_nsi = nullptr, *d_rfi = nullptr, *d_order = nullptr;
    // TEMP_GRAD's lead slots: 4 x LICA_TG_LEAD keys for the launches over
    // every case, then the same for each lane
    unsigned long long *d_lead = nullptr;
    double *d_mg = nullptr;      // multigrid scratch (allocated on first use of a multigrid mode)
    lica_net *net = nullptr;     // warm-start network of the learned mode (not owned)
    double *d_guess = nullptr, *d_scale = nullptr, *d_cap = nullptr, *d_hist = nullptr;
    NetCase *d_netcases = nullptr;
    std::vector<int64_t> lens;
    int capture = 0;             // substep whose POISSON solve is captured (0: none)
    bool cap_valid = false;      // the capture rows hold the last run's last substep
    long long mg_lds_cap = -1;   // dynamic LDS cap of the multigrid kernel (-1: by batch size)
    std::vector<std::vector<double>> img_ns, img_rf;     // initial images (geometry)
    std::vector<std::vector<int>> img_nsi, img_rfi;
    int threads = 256;
    int poisson_mode = LICA_POISSON_THOMAS;
    double rtol = 1.0e-10;
    int max_cycles = 100;
    bool use_graph = false;
    bool timing = false;
    bool initialized = false;
    int last_msub = 1;
    hipStream_t stream = nullptr;
#ifndef LICA_EMUL
    hipGraphExec_t graph = nullptr;
    int graph_threads = -1, graph_mode = -1, graph_lanes = -1;
    int graph_capture = 0;
    const lica_net *graph_net = nullptr;
    long long graph_mg_lds = -1;
    int graph_cfg = -1;
    std::vector<hipEvent_t> ev[LICA_MAX_LANES];     // timing events of the lanes
#endif
    // Lanes: the cases split into L sets (every L-th case of the
    // largest-first order) whose steps run as independent launch sequences
    // on L streams, so that a latency-bound kernel of one set shares the
    // CUs with a bandwidth-bound kernel of another and one set's tail of
    // workgroups overlaps the next set's launch.  In timing mode each lane
    // records its own events (a kernel's time is then its duration while
    // the other lanes run beside it).
    int lanes = 1;
    int cur_lane = -1;                 // set being enqueued (-1: all cases)
    int *d_order_lanes = nullptr;      // per lane count L = 1..LICA_MAX_LANES: that split's lane-major order
    hipStream_t lane_stream[LICA_MAX_LANES] = {};
    // split of the cases over the lanes: 0 every L-th case of the
    // largest-first order, 1 contiguous size bins of that order with equal
    // work (lane l holds positions bin_lo[l] .. bin_lo[l+1]-1), so that each
    // lane's launch shape fits its own largest case
    int lane_split = 0;
    int bin_lo[LICA_MAX_LANES + 1] = {};
    std::vector<int> order_h;          // the largest-first order
    LaneCfg lcfg[LICA_MAX_LANES];      // per lane (lane 0 is every case with one lane)
    LaneCfg all_cfg;                   // launches over every case while there are several lanes
    int cfg_version = 0;               // changes with any launch shape
#ifndef LICA_EMUL
    hipEvent_t fork = nullptr, join[LICA_MAX_LANES] = {};
#endif
    double timings[KL_COUNT] = {0};
    double wall_ms = 0.0;
};

namespace {

size_t lead_bytes() { return sizeof(unsigned long long) * 4 * LICA_TG_LEAD * (LICA_MAX_LANES + 1); }

// launch shape of the set being enqueued
const LaneCfg &cur_cfg(const lica_batch *b)
{
    if (b->cur_lane >= 0) return b->lcfg[b->cur_lane];
    return (b->lanes == 1) ? b->lcfg[0] : b->all_cfg;
}

// dynamic LDS of the multigrid kernel: x, f of every level of the set's
// largest case in the cycles' precision, capped (tuned, else the batch's
// cap, else 60 KB for batches of many cases, so that several workgroups
// share a CU, else the device's per-workgroup limit); levels that do not
// fit stay in the global scratch
long long mg_lds(const lica_batch *b, const LaneCfg &cfg)
{
    if (b->poisson_mode < LICA_POISSON_MG64) return 0;
    const long long need = lica_mg_lds_need(cfg.nmax, b->poisson_mode == LICA_POISSON_MG64 ? 8 : 4);
    long long cap = (cfg.mg_lds >= 0) ? cfg.mg_lds
                  : (b->mg_lds_cap >= 0) ? b->mg_lds_cap : (b->ninst >= 512 ? 65536 - 4096 : lica_mg_lds_max());
    cap = std::min(cap, lica_mg_lds_max());
    return std::min(need, cap) / 16 * 16;
}
long long mg_lds(const lica_batch *b) { return mg_lds(b, cur_cfg(b)); }

BatchDev make_bd(const lica_batch *b, int msub)
{
    BatchDev bd;
    bd.par = b->d_par; bd.st = b->d_st;
    bd.ns = b->d_ns; bd.rf = b->d_rf; bd.nsi = b->d_nsi; bd.rfi = b->d_rfi;
    bd.order = b->d_order;
    bd.ninst = (int)b->ninst;
    // the multigrid modes leave only diagonally dominant rows (energy,
    // momentum) to the tridiagonal solver: no refinement steps there
    bd.tri_mode = (b->poisson_mode == LICA_POISSON_THOMAS) ? 0 : (b->poisson_mode >= LICA_POISSON_MG64) ? 2 : 1;
    bd.msub = msub;
    bd.pmode = b->poisson_mode;
    bd.mg = b->d_mg;
    bd.mg_rtol = b->rtol;
    bd.mg_maxc = b->max_cycles;
    bd.guess = b->d_guess;
    bd.hist = b->d_hist;
    bd.cap = b->capture ? b->d_cap : nullptr;
    // the learned mode's extra pass over B (feature scales, the solve's
    // start verdict); the projected mode writes B alone (the multigrid
    // kernel's own start test is then cheaper than the pass)
    bd.wscale = (b->poisson_mode == LICA_POISSON_LEARNED && b->net) ? b->d_scale : nullptr;
    bd.mg_lds = mg_lds(b);
    bd.lead = nullptr;
    return bd;
}

// threads per workgroup of kernel class `which` on the set being enqueued:
// the tuned count, else the requested one raised (outside the Thomas mode)
// until the set's largest case has a thread per chunk of its partitioned
// solves; a tuned count of a kernel that solves systems is raised likewise
int block_threads(const lica_batch *b, int which)
{
    const LaneCfg &cfg = cur_cfg(b);
    const bool part = b->poisson_mode != LICA_POISSON_THOMAS;
    const int need = part ? lica_tri_chunks(cfg.nmax) : 0;
    const bool solves = which == KL_CONT_ENERGY || which == KL_MOMENTUM || which == KL_POISSON_DIRECT;
    int t = cfg.thr[which];
    if (t > 0) {
        while (solves && t < need && t < 1024) t *= 2;
        return t;
    }
    t = b->threads;
    while (part && t < need && t < 1024) t *= 2;
    return t;
}

hipStream_t cur_stream(const lica_batch *b) { return b->cur_lane < 0 ? b->stream : b->lane_stream[b->cur_lane]; }

// cases of lane l, and the workgroup order of the lane being enqueued
int lane_cases(const lica_batch *b, int l)
{
    return b->lane_split ? b->bin_lo[l + 1] - b->bin_lo[l] : lane_count(b->ninst, b->lanes, l);
}
const int *lane_order(const lica_batch *b)
{
    if (b->lane_split) return b->d_order + b->bin_lo[b->cur_lane];
    return b->d_order_lanes + (size_t)(b->lanes - 1) * (size_t)b->ninst + lane_first(b->ninst, b->lanes, b->cur_lane);
}

// the lanes' case sets and their largest cases for the current lane count
// and split; tuned launch shapes are dropped.  The binned split gives lane l
// the positions of the largest-first order whose cumulative work (n plus a
// fixed 512 rows for the band kernels, whose cost does not grow with n)
// reaches l/L of the total, at least one case each.
void plan_lanes(lica_batch *b)
{
    const int L = b->lanes;
    const int64_t N = b->ninst;
    auto nof = [&](int pos) { return b->par[(size_t)b->order_h[(size_t)pos]].n; };
    if (b->lane_split) {
        double tot = 0.0;
        for (int64_t i = 0; i < N; i++) tot += nof((int)i) + 512.0;
        b->bin_lo[0] = 0;
        double acc = 0.0;
        int64_t pos = 0;
        for (int l = 1; l < L; l++) {
            const double goal = tot * l / L;
            while (pos < N - (L - l) && (pos < b->bin_lo[l - 1] + 1 || acc + (nof((int)pos) + 512.0) * 0.5 < goal)) {
                acc += nof((int)pos) + 512.0;
                pos++;
            }
            b->bin_lo[l] = (int)pos;
        }
        b->bin_lo[L] = (int)N;
    }
    for (int l = 0; l < LICA_MAX_LANES; l++) b->lcfg[l] = LaneCfg();
    for (int l = 0; l < L; l++)
        b->lcfg[l].nmax = (N > 0) ? nof(b->lane_split ? b->bin_lo[l] : l) : 0;
    b->all_cfg = LaneCfg();
    b->all_cfg.nmax = (N > 0) ? nof(0) : 0;
    b->cfg_version++;
}

int launch(lica_batch *b, int which, int msub)
{
    BatchDev bd = make_bd(b, msub);
    if (b->cur_lane >= 0) {
        bd.order = lane_order(b);
        bd.ninst = lane_cases(b, b->cur_lane);
    }
    // the lead slots of the set being enqueued (a key holds a position in 24 bits)
    if (which == KL_TEMP_GRAD && b->d_lead && bd.ninst <= (1 << 24))
        bd.lead = b->d_lead + (size_t)(b->cur_lane + 1) * 4 * LICA_TG_LEAD;
    int rc = lica_launch(which, bd, block_threads(b, which), (void *)cur_stream(b));
    return rc == 0 ? 0 : LICA_HIP_ERROR;
}

// work between the POISSON matrix and its solve: the warm-start base, the
// capture of the starting point, the network's correction
int pre_poisson(lica_batch *b, int msub)
{
    int rc = 0;
    const bool warm = (b->poisson_mode == LICA_POISSON_LEARNED || b->poisson_mode == LICA_POISSON_PROJECTED);
    if (warm) rc = launch(b, KL_WARM_BASE, msub);
    if (!rc && b->capture == msub) rc = launch(b, KL_CAPTURE_PRE, msub);
    if (!rc && b->poisson_mode == LICA_POISSON_LEARNED) {
        if (!b->net) return LICA_BAD_ARG;
        const int *ord = (b->cur_lane >= 0) ? lane_order(b) : nullptr;
        const int cnt = (b->cur_lane >= 0) ? lane_cases(b, b->cur_lane) : (int)b->ninst;
        if (lica_net_guess(b->net, b->d_netcases, ord, cnt, b->d_scale, b->rtol, (void *)cur_stream(b)))
            rc = LICA_HIP_ERROR;
    }
    return rc;
}

int enqueue_step_cases(lica_batch *b);

// streams and join events of lanes 1..L-1, created when first needed
int ensure_lanes(lica_batch *b, int L)
{
#ifndef LICA_EMUL
    for (int l = 1; l < L; l++) {
        if (!b->lane_stream[l] && hipStreamCreateWithFlags(&b->lane_stream[l], hipStreamNonBlocking) != hipSuccess)
            return LICA_HIP_ERROR;
        if (!b->join[l] && hipEventCreateWithFlags(&b->join[l], hipEventDisableTiming) != hipSuccess)
            return LICA_HIP_ERROR;
    }
#else
    (void)b; (void)L;
#endif
    return 0;
}

#ifndef LICA_EMUL
// lanes 1.. start after what the batch's stream holds, and the batch's stream
// goes on after every lane's launches
int fork_lanes(lica_batch *b)
{
    if (hipEventRecord(b->fork, b->stream) != hipSuccess) return 1;
    for (int l = 1; l < b->lanes; l++)
        if (hipStreamWaitEvent(b->lane_stream[l], b->fork, 0) != hipSuccess) return 1;
    return 0;
}

int join_lanes(lica_batch *b)
{
    for (int l = 1; l < b->lanes; l++)
        if (hipEventRecord(b->join[l], b->lane_stream[l]) != hipSuccess ||
            hipStreamWaitEvent(b->stream, b->join[l], 0) != hipSuccess)
            return 1;
    return 0;
}
#endif

// the launches of one time step (1d.f:196-295); with L > 1 lanes the L case
// sets run their own sequences on L streams, forked from and joined back
// into the batch's stream
// the network's BF16 weights for the learned mode (they do not change while
// a batch runs)
int cast_net(lica_batch *b)
{
    if (b->poisson_mode != LICA_POISSON_LEARNED) return 0;
    if (!b->net) return LICA_BAD_ARG;
    return lica_net_cast(b->net, (void *)b->stream) ? LICA_HIP_ERROR : 0;
}

int enqueue_step(lica_batch *b)
{
    if (int rc = cast_net(b)) return rc;
#ifndef LICA_EMUL
    if (b->lanes > 1 && b->cur_lane < 0) {
        if (fork_lanes(b)) return LICA_HIP_ERROR;
        int rc = 0;
        for (int l = 0; l < b->lanes && !rc; l++) {
            b->cur_lane = l;
            rc = enqueue_step_cases(b);
        }
        b->cur_lane = -1;
        if (join_lanes(b)) rc = rc ? rc : LICA_HIP_ERROR;
        return rc;
    }
#endif
    return enqueue_step_cases(b);
}

int enqueue_step_cases(lica_batch *b)
{
    int rc = launch(b, KL_DT, 1);
    const int kp = (b->poisson_mode >= LICA_POISSON_MG64) ? KL_POISSON_MG : KL_POISSON_DIRECT;
    for (int msub = 1; msub <= 3 && !rc; msub++) {
        if (!rc) rc = launch(b, KL_CONT_ENERGY, msub);
        if (!rc) rc = launch(b, KL_TEMP_GRAD, msub);
        if (!rc) rc = launch(b, KL_MOMENTUM, msub);
        if (!rc) rc = pre_poisson(b, msub);
        if (!rc) rc = launch(b, kp, msub);
        if (!rc && b->capture == msub) rc = launch(b, KL_CAPTURE_POST, msub);
        if (!rc) rc = launch(b, KL_UCALC, msub);
        if (!rc) rc = launch(b, KL_TRANSPORT, msub);
    }
    return rc;
}

#ifndef LICA_EMUL
// same launches with an event after each one (per-kernel GPU time)
// (one lane's launches when b->cur_lane >= 0; ev[0] is recorded by the
// caller on the same stream)
int enqueue_step_timed(lica_batch *b, std::vector<hipEvent_t> &ev, std::vector<int> &kinds)
{
    auto rec = [&](int kind) -> int {
        kinds.push_back(kind);
        size_t i = kinds.size();
        while (ev.size() <= i) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return LICA_HIP_ERROR;
            ev.push_back(e);
        }
        return hipEventRecord(ev[i], cur_stream(b)) == hipSuccess ? 0 : LICA_HIP_ERROR;
    };
    int rc = launch(b, KL_DT, 1);
    if (!rc) rc = rec(KL_DT);
    for (int msub = 1; msub <= 3 && !rc; msub++) {
        const int kp = (b->poisson_mode >= LICA_POISSON_MG64) ? KL_POISSON_MG : KL_POISSON_DIRECT;
        const int seq[6] = {KL_CONT_ENERGY, KL_TEMP_GRAD, KL_MOMENTUM, kp, KL_UCALC, KL_TRANSPORT};
        for (int j = 0; j < 6 && !rc; j++) {
            if (seq[j] == kp) {
                rc = pre_poisson(b, msub);
                if (!rc) rc = rec(KL_CAPTURE_PRE);    // time of the guess/capture work
                if (rc) break;
            }
            rc = launch(b, seq[j], msub);
            if (!rc) rc = rec(seq[j]);
            if (!rc && seq[j] == kp && b->capture == msub) rc = launch(b, KL_CAPTURE_POST, msub);
        }
    }
    return rc;
}

int build_graph(lica_batch *b)
{
    if (b->graph) { (void)hipGraphExecDestroy(b->graph); b->graph = nullptr; }
    hipGraph_t g;
    if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return LICA_HIP_ERROR;
    int rc = enqueue_step(b);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc || e != hipSuccess) return LICA_HIP_ERROR;
    e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { b->graph = nullptr; return LICA_HIP_ERROR; }
    b->graph_threads = b->threads;
    b->graph_lanes = b->lanes;
    b->graph_mode = b->poisson_mode;
    b->graph_capture = b->capture;
    b->graph_net = b->net;
    b->graph_mg_lds = b->mg_lds_cap;
    b->graph_cfg = b->cfg_version;
    return 0;
}
#endif

int check_batch(lica_batch *b) { return (b && b->ninst > 0) ? 0 : LICA_BAD_ARG; }

int first_status(lica_batch *b, int64_t *bad)
{
    std::vector<InstState> st((size_t)b->ninst);
    int rc = d2h(st.data(), b->d_st, sizeof(InstState) * st.size(), b->stream);
    if (rc) return rc;
    for (int64_t i = 0; i < b->ninst; i++)
        if (st[(size_t)i].status) { if (bad) *bad = i; return st[(size_t)i].status; }
    if (bad) *bad = -1;
    return 0;
}

int get_state(lica_batch *b, int64_t inst, InstState *s)
{
    return d2h(s, b->d_st + inst, sizeof(InstState), b->stream);
}

}  // namespace

extern "C" {

const char *lica_status_string(int s)
{
    switch (s) {
    case LICA_OK: return "ok";
    case LICA_BAND_OVERFLOW: return "narrow band exceeds MF_BAND";
    case LICA_PSIF_BUG: return "CAL_PSIF_MTP reached its error branch";
    case LICA_GRID_TOO_LARGE: return "grid coupling search failed";
    case LICA_EXTRAP_NOT_CONVERGED: return "temperature extrapolation did not converge";
    case LICA_BAND_CAPACITY: return "band layer holds more planes than the device record";
    case LICA_NONFINITE: return "non-finite W, T or P";
    case LICA_WINDOW: return "band left the device window";
    case LICA_BAD_ARG: return "bad argument";
    case LICA_UNSUPPORTED_DECK: return "deck describes a configuration other than the 1-D film-boiling column";
    case LICA_IO: return "cannot read input file";
    case LICA_HIP_ERROR: return "HIP runtime error";
    case LICA_NO_DEVICE: return "no GPU device";
    case LICA_NOT_CONVERGED: return "Poisson iteration hit its cycle cap";
    case LICA_TOO_LARGE: return "systems exceed the fixed workspace of the allocation-free entry point";
    }
    return "unknown status";
}

int lica_parse_deck(const char *lica_in_path, const char *grid_in_path, lica_case_desc *out)
{
    if (!lica_in_path || !grid_in_path || !out) return LICA_BAD_ARG;
    LineReader L;
    if (!L.open(lica_in_path)) return LICA_IO;
    double v[16];
    double impl, idtopt, dt, cfl, ips, iphs, ibmon, ipx, ipy, ipz, itracking, re, visr, denr, surfj, fr, prm, scr, tcr,
        tsat, ja;
    bool ok = L.skip() && L.nums(5, v) && L.skip() && L.nums(4, v) && L.skip() && L.nums(4, v);
    if (ok) { impl = v[0]; idtopt = v[1]; dt = v[2]; cfl = v[3]; }
    ok = ok && L.skip() && L.nums(9, v) && L.skip() && L.nums(2, v);
    if (ok) { ips = v[0]; iphs = v[1]; }
    ok = ok && L.skip() && L.nums(2, v);
    if (ok) ibmon = v[0];
    ok = ok && L.skip() && L.nums(1, v) && L.skip() && L.nums(4, v);
    if (ok) { ipx = v[1]; ipy = v[2]; ipz = v[3]; }
    ok = ok && L.skip() && L.skip() && L.nums(3, v);
    if (ok) itracking = v[0];
    ok = ok && L.skip() && L.nums(3, v);
    if (ok) { re = v[0]; visr = v[1]; denr = v[2]; }
    ok = ok && L.skip() && L.nums(3, v);
    if (ok) { surfj = v[0]; fr = v[1]; }
    ok = ok && L.skip() && L.nums(3, v);
    if (ok) { prm = v[0]; scr = v[1]; tcr = v[2]; }
    ok = ok && L.skip() && L.nums(2, v);
    if (ok) { tsat = v[0]; ja = v[1]; }
    if (!ok) return LICA_IO;
    (void)ips;
    if (impl != 0.0 || iphs != 1.0 || ibmon != 0.0 || ipx != 1.0 || ipy != 1.0 || ipz != 3.0 || itracking != 2.0 ||
        (idtopt != 0.0 && idtopt != 1.0))
        return LICA_UNSUPPORTED_DECK;

    LineReader G;
    if (!G.open(grid_in_path)) return LICA_IO;
    double g[8], sp[4];
    if (!(G.skip() && G.nums(4, sp) && G.skip() && G.skip() && G.nums(7, g) && G.skip())) return LICA_IO;
    if (sp[0] != 0.0 || sp[1] != 0.0 || sp[2] != 0.0 || sp[3] != 0.0) return LICA_UNSUPPORTED_DECK;
    int nsect1 = (int)g[0];
    double n1 = g[1], n2 = g[2], n3 = g[3], xl = g[4], yl = g[5], zl = g[6];
    auto sections = [&](int ns, int dirbase, double len[3], double np[3]) -> int {
        for (int l = 0; l < ns; l++) {
            double s[8];
            if (!G.nums(8, s)) return LICA_IO;
            int dir = (int)s[0] - dirbase;
            if (dir < 1 || dir > 3) return LICA_UNSUPPORTED_DECK;
            // one uniform section spanning each direction (IOPT=1)
            if (s[1] != 0.0 || s[5] != 1.0 || s[3] != 1.0 || s[2] != len[dir - 1] || s[4] != np[dir - 1])
                return LICA_UNSUPPORTED_DECK;
        }
        return 0;
    };
    if (nsect1 != 3) return LICA_UNSUPPORTED_DECK;
    double len1[3] = {xl, yl, zl}, np1[3] = {n1, n2, n3};
    int rc = sections(nsect1, 0, len1, np1);
    if (rc) return rc;
    double h[8];
    if (!(G.skip() && G.skip() && G.nums(7, h) && G.skip())) return LICA_IO;
    int nsect2 = (int)h[0];
    double len2[3] = {h[4], h[5], h[6]}, np2[3] = {h[1], h[2], h[3]};
    if (nsect2 != 3) return LICA_UNSUPPORTED_DECK;
    rc = sections(nsect2, 3, len2, np2);
    if (rc) return rc;
    // the 1-D column: one NS cell in x/y of unit width, refined r x r x rN
    if (n1 != 2.0 || n2 != 2.0 || xl != 1.0 || yl != 1.0 || h[4] != xl || h[5] != yl || h[6] != zl)
        return LICA_UNSUPPORTED_DECK;
    int64_t ncell = (int64_t)n3 - 1, nf = (int64_t)h[3] - 1, rr = (int64_t)h[1] - 1;
    if (ncell < 8 || rr < 1 || (int64_t)h[2] - 1 != rr || nf != rr * ncell) return LICA_UNSUPPORTED_DECK;

    memset(out, 0, sizeof *out);
    out->n = ncell; out->r = rr; out->zl = zl;
    out->re = re; out->visr = visr; out->denr = denr;
    out->surf_j = surfj; out->fr = fr;
    out->prm = prm; out->scr = scr; out->tcr = tcr;
    out->tsat = tsat; out->ja = ja;
    out->dt0 = dt; out->cflmax = cfl; out->idtopt = (int64_t)idtopt;
    out->z0 = 0.0607;     // ASSIGN_LVS (two_phase.f:108)
    out->tslope = 16.47;  // MAKEFLD vapour temperature T=16.47*ZP (1d.f:909)
    return LICA_OK;
}

int lica_batch_create(int64_t ninst, const lica_case_desc *cases, lica_batch **out)
{
    if (!out || ninst <= 0 || !cases) return LICA_BAD_ARG;
    *out = nullptr;
    if (!have_device()) return LICA_NO_DEVICE;
    lica_batch *b = new lica_batch();
    b->ninst = ninst;
    b->cases.assign(cases, cases + ninst);
    b->par.resize((size_t)ninst);
    b->img_ns.resize((size_t)ninst); b->img_rf.resize((size_t)ninst);
    b->img_nsi.resize((size_t)ninst); b->img_rfi.resize((size_t)ninst);
    long long gs_total = 0;
    // host images of the cases, built on all host threads
    std::vector<CaseImage> imgs((size_t)ninst);
    std::vector<int> brc((size_t)ninst, 0);
    {
        const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(ninst, std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
        std::atomic<int64_t> next(0);
        auto work = [&]() {
            for (int64_t i = next++; i < ninst; i = next++) brc[(size_t)i] = build_case(cases[i], imgs[(size_t)i]);
        };
        std::vector<std::thread> th;
        for (int t = 1; t < nt; t++) th.emplace_back(work);
        work();
        for (auto &t : th) t.join();
    }
    for (int64_t i = 0; i < ninst; i++) {
        CaseImage &c = imgs[(size_t)i];
        int rc = brc[(size_t)i];
        if (rc) { delete b; return rc; }
        c.p.ns_off = (long long)b->ns_total; b->ns_total += c.ns.size();
        c.p.rf_off = (long long)b->rf_total; b->rf_total += c.rf.size();
        c.p.nsi_off = (long long)b->nsi_total; b->nsi_total += c.nsi.size();
        c.p.rfi_off = (long long)b->rfi_total; b->rfi_total += c.rfi.size();
        c.p.gs_off = gs_total; gs_total += c.p.n + 2;
        b->lens.push_back(c.p.n);
        b->par[(size_t)i] = c.p;
        b->img_ns[(size_t)i].swap(c.ns); b->img_rf[(size_t)i].swap(c.rf);
        b->img_nsi[(size_t)i].swap(c.nsi); b->img_rfi[(size_t)i].swap(c.rfi);
    }
#ifndef LICA_EMUL
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) { delete b; return LICA_HIP_ERROR; }
    b->lane_stream[0] = b->stream;
    if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) {
        lica_batch_destroy(b);
        return LICA_HIP_ERROR;
    }
#endif
    int rc = 0;
    rc = rc ? rc : dmalloc((void **)&b->d_par, sizeof(InstParam) * (size_t)ninst);
    rc = rc ? rc : dmalloc((void **)&b->d_st, sizeof(InstState) * (size_t)ninst);
    rc = rc ? rc : dmalloc((void **)&b->d_ns, sizeof(double) * b->ns_total);
    rc = rc ? rc : dmalloc((void **)&b->d_rf, sizeof(double) * b->rf_total);
    rc = rc ? rc : dmalloc((void **)&b->d_nsi, sizeof(int) * b->nsi_total);
    rc = rc ? rc : dmalloc((void **)&b->d_rfi, sizeof(int) * b->rfi_total);
    rc = rc ? rc : dmalloc((void **)&b->d_order, sizeof(int) * (size_t)ninst);
    rc = rc ? rc : dmalloc((void **)&b->d_lead, lead_bytes());
    rc = rc ? rc : dmalloc((void **)&b->d_guess, sizeof(double) * (size_t)gs_total);
    rc = rc ? rc : dmalloc((void **)&b->d_cap, sizeof(double) * 3 * (size_t)gs_total);
    rc = rc ? rc : dmalloc((void **)&b->d_scale, sizeof(double) * LICA_WARM_INFO * (size_t)ninst);
    if (rc) { lica_batch_destroy(b); return rc; }
    // workgroup order: largest cases first (longest-processing-time first)
    std::vector<int> order((size_t)ninst);
    for (int64_t i = 0; i < ninst; i++) order[(size_t)i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return b->par[(size_t)a].n > b->par[(size_t)c].n; });
    rc = h2d(b->d_order, order.data(), sizeof(int) * order.size(), b->stream);
    {
        // the splits over 1..LICA_MAX_LANES lanes, one after the other
        std::vector<int> lo;
        for (int L = 1; L <= LICA_MAX_LANES; L++)
            for (int l = 0; l < L; l++)
                for (size_t j = (size_t)l; j < order.size(); j += (size_t)L) lo.push_back(order[j]);
        b->lanes = (ninst >= 256) ? 2 : 1;
        b->order_h = order;
        plan_lanes(b);
        rc = rc ? rc : ensure_lanes(b, b->lanes);
        rc = rc ? rc : dmalloc((void **)&b->d_order_lanes, sizeof(int) * lo.size());
        rc = rc ? rc : h2d(b->d_order_lanes, lo.data(), sizeof(int) * lo.size(), b->stream);
    }
    rc = rc ? rc : h2d(b->d_par, b->par.data(), sizeof(InstParam) * b->par.size(), b->stream);
    if (rc) { lica_batch_destroy(b); return rc; }
    *out = b;
    return LICA_OK;
}

int lica_batch_init(lica_batch *b)
{
    int rc = check_batch(b);
    if (rc) return rc;
    // fresh geometry and zeroed fields, then MAKEFLD/LVSINIT/GRID_COUPLING(1) on the device
    for (int64_t i = 0; i < b->ninst && !rc; i++) {
        const InstParam &p = b->par[(size_t)i];
        rc = h2d(b->d_ns + p.ns_off, b->img_ns[(size_t)i].data(), sizeof(double) * b->img_ns[(size_t)i].size(), b->stream);
        rc = rc ? rc : h2d(b->d_rf + p.rf_off, b->img_rf[(size_t)i].data(), sizeof(double) * b->img_rf[(size_t)i].size(), b->stream);
        rc = rc ? rc : h2d(b->d_nsi + p.nsi_off, b->img_nsi[(size_t)i].data(), sizeof(int) * b->img_nsi[(size_t)i].size(), b->stream);
        rc = rc ? rc : h2d(b->d_rfi + p.rfi_off, b->img_rfi[(size_t)i].data(), sizeof(int) * b->img_rfi[(size_t)i].size(), b->stream);
    }
    if (rc) return rc;
    std::vector<InstState> st((size_t)b->ninst);
    memset((void *)st.data(), 0, sizeof(InstState) * st.size());
    rc = h2d(b->d_st, st.data(), sizeof(InstState) * st.size(), b->stream);
    if (rc) return rc;
    rc = launch(b, KL_INIT, 1);
    if (!rc) rc = dsync(b->stream);
    if (rc) return rc;
    b->initialized = true;
    int64_t bad;
    return first_status(b, &bad);
}

int lica_batch_run(lica_batch *b, int64_t nsteps)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (!b->initialized || nsteps < 0) return LICA_BAD_ARG;
    for (int i = 0; i < KL_COUNT; i++) b->timings[i] = 0.0;
    auto t0 = std::chrono::steady_clock::now();
#ifdef LICA_EMUL
    for (int64_t s = 0; s < nsteps && !rc; s++) rc = enqueue_step(b);
#else
    if (b->timing || !b->use_graph) {
        // the lanes are forked once for the whole run: each lane's steps
        // follow each other on its stream with no join in between (the cases
        // are independent), so the lanes drift apart and one lane's launches
        // fill the other's idle CUs.  Timing events are read after the run.
        const int nl = b->lanes;
        std::vector<int> kinds[LICA_MAX_LANES];
        rc = cast_net(b);
        if (!rc && nl > 1 && fork_lanes(b)) rc = LICA_HIP_ERROR;
        for (int l = 0; l < nl && b->timing && !rc; l++) {
            b->cur_lane = (nl > 1) ? l : -1;
            if (b->ev[l].empty()) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) { rc = LICA_HIP_ERROR; break; }
                b->ev[l].push_back(e);
            }
            if (hipEventRecord(b->ev[l][0], cur_stream(b)) != hipSuccess) rc = LICA_HIP_ERROR;
        }
        for (int64_t s = 0; s < nsteps && !rc; s++)
            for (int l = 0; l < nl && !rc; l++) {
                b->cur_lane = (nl > 1) ? l : -1;
                rc = b->timing ? enqueue_step_timed(b, b->ev[l], kinds[l]) : enqueue_step_cases(b);
            }
        b->cur_lane = -1;
        if (nl > 1 && join_lanes(b)) rc = rc ? rc : LICA_HIP_ERROR;
        for (int l = 0; l < nl && b->timing && !rc; l++) {
            if (hipEventSynchronize(b->ev[l][kinds[l].size()]) != hipSuccess) return LICA_HIP_ERROR;
            for (size_t j = 0; j < kinds[l].size(); j++) {
                float ms = 0.0f;
                if (hipEventElapsedTime(&ms, b->ev[l][j], b->ev[l][j + 1]) != hipSuccess) return LICA_HIP_ERROR;
                b->timings[kinds[l][j]] += (double)ms;
            }
        }
    } else if (b->use_graph) {
        if (!b->graph || b->graph_threads != b->threads || b->graph_lanes != b->lanes ||
            b->graph_mode != b->poisson_mode ||
            b->graph_capture != b->capture || b->graph_net != b->net || b->graph_mg_lds != b->mg_lds_cap ||
            b->graph_cfg != b->cfg_version)
            rc = build_graph(b);
        for (int64_t s = 0; s < nsteps && !rc; s++)
            if (hipGraphLaunch(b->graph, b->stream) != hipSuccess) rc = LICA_HIP_ERROR;
    }
#endif
    if (!rc) rc = dsync(b->stream);
    b->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rc) return rc;
    if (b->capture && nsteps > 0) b->cap_valid = true;
    int64_t bad;
    return first_status(b, &bad);
}

int lica_batch_set_poisson(lica_batch *b, int64_t mode, double rtol, int64_t max_cycles)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (mode < LICA_POISSON_THOMAS || mode > LICA_POISSON_PROJECTED || !(rtol > 0.0) || max_cycles < 1) return LICA_BAD_ARG;
    if (mode == LICA_POISSON_LEARNED && !b->net) return LICA_BAD_ARG;
    if (mode >= LICA_POISSON_MG64 && !b->d_mg) {
        long long tot = 0;
        for (InstParam &p : b->par) { p.mg_off = tot; tot += lica_mg_scratch_doubles(p.n); }
        rc = dmalloc((void **)&b->d_mg, sizeof(double) * (size_t)tot);
        rc = rc ? rc : h2d(b->d_par, b->par.data(), sizeof(InstParam) * b->par.size(), b->stream);
        if (rc) return rc;
    }
    if ((mode == LICA_POISSON_LEARNED || mode == LICA_POISSON_PROJECTED) && !b->d_hist) {
        long long tot = 0;
        for (InstParam &p : b->par) { p.hist_off = tot; tot += 3LL * MG_HIST * p.n; }
        rc = dmalloc((void **)&b->d_hist, sizeof(double) * (size_t)tot);
        rc = rc ? rc : h2d(b->d_par, b->par.data(), sizeof(InstParam) * b->par.size(), b->stream);
        if (rc) return rc;
    }
    if (b->poisson_mode != (int)mode) {
        b->poisson_mode = (int)mode;
        plan_lanes(b);   // tuned shapes belong to a mode
    }
    b->rtol = rtol;
    b->max_cycles = (int)max_cycles;
    return LICA_OK;
}

int lica_batch_set_mg_lds(lica_batch *b, int64_t bytes)
{
    int rc = check_batch(b);
    if (rc) return rc;
    b->mg_lds_cap = (bytes < 0) ? -1 : (long long)bytes;
    return LICA_OK;
}

int lica_batch_set_lanes(lica_batch *b, int64_t lanes)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (lanes < 1 || lanes > LICA_MAX_LANES) return LICA_BAD_ARG;
    const int L = (int)std::min<int64_t>(lanes, b->ninst);
    if (int e = ensure_lanes(b, L)) return e;
    b->lanes = L;
    plan_lanes(b);
    return LICA_OK;
}

int lica_batch_set_block(lica_batch *b, int64_t threads)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (threads < 64 || threads > 1024 || (threads & (threads - 1))) return LICA_BAD_ARG;
    b->threads = (int)threads;
    return LICA_OK;
}

int lica_batch_set_lane_split(lica_batch *b, int64_t split)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (split != 0 && split != 1) return LICA_BAD_ARG;
    b->lane_split = (int)split;
    plan_lanes(b);
    return LICA_OK;
}

}  // extern "C"

namespace {
long long guess_doubles(const lica_batch *b);
long long hist_doubles(const lica_batch *b);

// a device copy of everything a step reads and writes, to run trial steps
// from and return to the same state
struct StateCopy {
    struct Buf { void *dev; void *copy; size_t bytes; };
    std::vector<Buf> bufs;
    ~StateCopy() { for (Buf &x : bufs) dfree(x.copy); }
    int add(void *dev, size_t bytes)
    {
        if (!dev || !bytes) return 0;
        Buf x{dev, nullptr, bytes};
        if (dmalloc(&x.copy, bytes)) return LICA_HIP_ERROR;
        bufs.push_back(x);
        return 0;
    }
    int save(hipStream_t st)
    {
        for (Buf &x : bufs)
            if (d2d(x.copy, x.dev, x.bytes, st)) return LICA_HIP_ERROR;
        return dsync(st);
    }
    int restore(hipStream_t st)
    {
        for (Buf &x : bufs)
            if (d2d(x.dev, x.copy, x.bytes, st)) return LICA_HIP_ERROR;
        return dsync(st);
    }
};

#ifndef LICA_EMUL
// GPU time of `steps` steps of every lane, the lanes on their streams at
// once as in a run (events on the batch's stream before the fork and after
// the join)
int time_steps(lica_batch *b, int64_t steps, double *ms, double *lane_ms = nullptr)
{
    // lane_ms: each lane's time from its first launch to the end of its last
    // (the host enqueues the lanes' first steps one after the other)
    hipEvent_t e[2], ls[LICA_MAX_LANES] = {}, le[LICA_MAX_LANES] = {};
    const int nl = b->lanes;
    if (hipEventCreate(&e[0]) != hipSuccess) return LICA_HIP_ERROR;
    if (hipEventCreate(&e[1]) != hipSuccess) { (void)hipEventDestroy(e[0]); return LICA_HIP_ERROR; }
    int rc = 0;
    for (int l = 0; l < nl && lane_ms && !rc; l++)
        if (hipEventCreate(&ls[l]) != hipSuccess || hipEventCreate(&le[l]) != hipSuccess) rc = LICA_HIP_ERROR;
    rc = rc ? rc : (hipEventRecord(e[0], b->stream) == hipSuccess) ? 0 : LICA_HIP_ERROR;
    if (!rc && nl > 1 && fork_lanes(b)) rc = LICA_HIP_ERROR;
    for (int64_t s = 0; s < steps && !rc; s++)
        for (int l = 0; l < nl && !rc; l++) {
            b->cur_lane = (nl > 1) ? l : -1;
            if (lane_ms && s == 0 && hipEventRecord(ls[l], cur_stream(b)) != hipSuccess) rc = LICA_HIP_ERROR;
            rc = rc ? rc : enqueue_step_cases(b);
            if (!rc && lane_ms && s == steps - 1 && hipEventRecord(le[l], cur_stream(b)) != hipSuccess)
                rc = LICA_HIP_ERROR;
        }
    b->cur_lane = -1;
    if (nl > 1 && join_lanes(b)) rc = rc ? rc : LICA_HIP_ERROR;
    float t = 0.0f;
    if (!rc && (hipEventRecord(e[1], b->stream) != hipSuccess || hipEventSynchronize(e[1]) != hipSuccess ||
                hipEventElapsedTime(&t, e[0], e[1]) != hipSuccess))
        rc = LICA_HIP_ERROR;
    for (int l = 0; l < nl && lane_ms && !rc; l++) {
        float tl = 0.0f;
        if (hipEventElapsedTime(&tl, ls[l], le[l]) != hipSuccess) rc = LICA_HIP_ERROR;
        lane_ms[l] = (double)tl;
    }
    for (int l = 0; l < nl; l++) {
        if (ls[l]) (void)hipEventDestroy(ls[l]);
        if (le[l]) (void)hipEventDestroy(le[l]);
    }
    (void)hipEventDestroy(e[0]);
    (void)hipEventDestroy(e[1]);
    *ms = (double)t;
    return rc;
}

// New cuts of the binned split from the lanes' own times: a position's
// modelled time is its work in the split's measure (n + 512) times its
// lane's measured time per unit of work, and lane l's cut is where the
// cumulative modelled time reaches l/L of the total (at least one case per
// lane).  The lanes keep their tuned launch shapes; their largest cases are
// updated.
void rebalance_bins(lica_batch *b, const double *lane_ms)
{
    const int L = b->lanes;
    const int64_t N = b->ninst;
    auto wof = [&](int64_t pos) { return b->par[(size_t)b->order_h[(size_t)pos]].n + 512.0; };
    std::vector<double> t((size_t)N, 0.0);
    double tot = 0.0;
    for (int l = 0; l < L; l++) {
        double w = 0.0;
        for (int64_t p = b->bin_lo[l]; p < b->bin_lo[l + 1]; p++) w += wof(p);
        const double rate = (w > 0.0) ? lane_ms[l] / w : 0.0;
        for (int64_t p = b->bin_lo[l]; p < b->bin_lo[l + 1]; p++) { t[(size_t)p] = wof(p) * rate; tot += t[(size_t)p]; }
    }
    int lo[LICA_MAX_LANES + 1];
    lo[0] = 0;
    double acc = 0.0;
    int64_t pos = 0;
    for (int l = 1; l < L; l++) {
        const double goal = tot * l / L;
        while (pos < N - (L - l) && (pos < lo[l - 1] + 1 || acc + t[(size_t)pos] * 0.5 < goal)) {
            acc += t[(size_t)pos];
            pos++;
        }
        lo[l] = (int)pos;
    }
    lo[L] = (int)N;
    for (int l = 0; l <= L; l++) b->bin_lo[l] = lo[l];
    for (int l = 0; l < L; l++) b->lcfg[l].nmax = (N > 0) ? b->par[(size_t)b->order_h[(size_t)lo[l]]].n : 0;
    b->cfg_version++;
}
#endif
}  // namespace

extern "C" {

int lica_batch_autotune(lica_batch *b, int64_t steps)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (!b->initialized || steps < 1 || steps > 16) return LICA_BAD_ARG;
    for (int l = 0; l < b->lanes; l++) {
        for (int k = 0; k < KL_COUNT; k++) b->lcfg[l].thr[k] = 0;
        b->lcfg[l].mg_lds = -1;
    }
    b->cfg_version++;
#ifndef LICA_EMUL
    if (int e = cast_net(b)) return e;
    StateCopy sc;
    rc = sc.add(b->d_st, sizeof(InstState) * (size_t)b->ninst);
    rc = rc ? rc : sc.add(b->d_ns, sizeof(double) * b->ns_total);
    rc = rc ? rc : sc.add(b->d_rf, sizeof(double) * b->rf_total);
    rc = rc ? rc : sc.add(b->d_nsi, sizeof(int) * b->nsi_total);
    rc = rc ? rc : sc.add(b->d_rfi, sizeof(int) * b->rfi_total);
    rc = rc ? rc : sc.add(b->d_guess, sizeof(double) * (size_t)guess_doubles(b));
    rc = rc ? rc : sc.add(b->d_scale, sizeof(double) * LICA_WARM_INFO * (size_t)b->ninst);
    rc = rc ? rc : sc.add(b->d_cap, sizeof(double) * 3 * (size_t)guess_doubles(b));
    if (b->d_hist) rc = rc ? rc : sc.add(b->d_hist, sizeof(double) * (size_t)hist_doubles(b));
    rc = rc ? rc : sc.add(b->d_lead, lead_bytes());
    rc = rc ? rc : sc.save(b->stream);
    if (rc) return rc;
    const bool cap_valid = b->cap_valid;
    // a trial: `steps` steps of every lane from the saved state, back to it after
    auto trial = [&](double *ms, double *lane_ms = nullptr, int64_t nst = 0) -> int {
        b->cfg_version++;
        int e = time_steps(b, nst ? nst : steps, ms, lane_ms);
        int e2 = sc.restore(b->stream);
        return e ? e : e2;
    };
    // Coordinate descent on the whole batch's step time: one kernel class
    // (the longest first) of one lane at a time, each candidate timed with
    // every lane running as in a run.  Trials of `steps` steps spread by 1-2 %
    // from one to the next, so a candidate within 0.5 % of the best or faster
    // is confirmed by trials of four times the steps, the candidate's and the
    // kept shape's in turn, twice (a drift of the clocks falls on both); it
    // is taken when its two long trials together are more than 0.5 % faster.
    // The first trial runs the default shapes once untimed.
    double best = 0.0;
    rc = trial(&best);
    rc = rc ? rc : trial(&best);
    const int64_t long_steps = 4 * steps;
    auto confirm = [&](double t_short, auto &&set_cand, auto &&set_keep, bool *take) -> int {
        *take = false;
        if (!(t_short < best * 1.005)) return 0;
        double tc = 0.0, tk = 0.0;
        int e = 0;
        for (int r = 0; r < 2 && !e; r++) {
            double t = 0.0;
            set_cand();
            e = trial(&t, nullptr, long_steps);
            tc += t;
            set_keep();
            e = e ? e : trial(&t, nullptr, long_steps);
            tk += t;
        }
        set_cand();
        *take = !e && tc < tk * 0.995;
        return e;
    };
    const int classes[5] = {KL_POISSON_MG, KL_MOMENTUM, KL_CONT_ENERGY, KL_UCALC, KL_DT};
    const int cand[5] = {64, 128, 256, 512, 1024};
    for (int ci = 0; ci < 5 && !rc; ci++) {
        const int k = (classes[ci] == KL_POISSON_MG && b->poisson_mode < LICA_POISSON_MG64) ? KL_POISSON_DIRECT
                                                                                          : classes[ci];
        for (int l = 0; l < b->lanes && !rc; l++) {
            LaneCfg &cfg = b->lcfg[l];
            b->cur_lane = (b->lanes > 1) ? l : -1;
            const int cur = block_threads(b, k);
            b->cur_lane = -1;
            int keep = cfg.thr[k];
            for (int j = 0; j < 5 && !rc; j++) {
                if (cand[j] == cur) continue;
                cfg.thr[k] = cand[j];
                b->cur_lane = (b->lanes > 1) ? l : -1;
                const bool raised = block_threads(b, k) != cand[j];   // below a solving kernel's need
                b->cur_lane = -1;
                if (raised) continue;
                double t = 0.0;
                bool take = false;
                rc = trial(&t);
                rc = rc ? rc : confirm(t, [&] { cfg.thr[k] = cand[j]; }, [&] { cfg.thr[k] = keep; }, &take);
                if (take) { best = t; keep = cand[j]; }
            }
            cfg.thr[k] = keep;
        }
    }
    if (b->poisson_mode >= LICA_POISSON_MG64) {
        // LDS of the multigrid workgroup: room for every vector of the lane's
        // largest case, or a cap that lets more workgroups share a CU (the
        // finest levels then use global memory)
        for (int l = 0; l < b->lanes && !rc; l++) {
            LaneCfg &cfg = b->lcfg[l];
            const long long need = std::min(lica_mg_lds_need(cfg.nmax, b->poisson_mode == LICA_POISSON_MG64 ? 8 : 4),
                                            lica_mg_lds_max()) / 16 * 16;
            b->cur_lane = (b->lanes > 1) ? l : -1;
            const long long cur = mg_lds(b);
            b->cur_lane = -1;
            const long long caps[6] = {need, 65536 - 4096, 49152, 32768, 20480, 12288};
            long long keep = cfg.mg_lds;
            for (int j = 0; j < 6 && !rc; j++) {
                const long long c = std::min(caps[j], need);
                if (c == cur || (j > 0 && c == need)) continue;
                cfg.mg_lds = c;
                double t = 0.0;
                bool take = false;
                rc = trial(&t);
                rc = rc ? rc : confirm(t, [&] { cfg.mg_lds = c; }, [&] { cfg.mg_lds = keep; }, &take);
                if (take) { best = t; keep = c; }
            }
            cfg.mg_lds = keep;
        }
    }
    if (b->lane_split && b->lanes > 1) {
        // The binned split from the lanes' measured times: the lanes run
        // independently to the end of a run, which ends with the slowest.
        // The new cuts follow from measurements, so they are kept unless the
        // batch's time grows by more than 0.5 %; trials of four times the
        // steps (the imbalance is a drift), until the lanes' times are
        // within 1 % of each other.
        const int64_t rs = 4 * steps;
        double t0 = 0.0, lt[LICA_MAX_LANES] = {};
        rc = trial(&t0, lt, rs);
        for (int it = 0; it < 3 && !rc; it++) {
            double lmax = 0.0, lmin = 1e300;
            for (int l = 0; l < b->lanes; l++) { lmax = std::max(lmax, lt[l]); lmin = std::min(lmin, lt[l]); }
            if (!(lmax > 0.0) || lmax - lmin < 0.01 * lmax) break;
            int lo[LICA_MAX_LANES + 1], nm[LICA_MAX_LANES];
            for (int l = 0; l <= b->lanes; l++) lo[l] = b->bin_lo[l];
            for (int l = 0; l < b->lanes; l++) nm[l] = b->lcfg[l].nmax;
            rebalance_bins(b, lt);
            double t1 = 0.0, lt1[LICA_MAX_LANES] = {};
            rc = trial(&t1, lt1, rs);
            if (!rc && t1 <= t0 * 1.005) {
                t0 = t1;
                for (int l = 0; l < b->lanes; l++) lt[l] = lt1[l];
                continue;
            }
            for (int l = 0; l <= b->lanes; l++) b->bin_lo[l] = lo[l];
            for (int l = 0; l < b->lanes; l++) b->lcfg[l].nmax = nm[l];
            b->cfg_version++;
            break;
        }
    }
    b->cfg_version++;
    b->cap_valid = cap_valid;
    for (int k = 0; k < KL_COUNT; k++) b->timings[k] = 0.0;
    if (rc) {
        for (int l = 0; l < b->lanes; l++) {
            for (int k = 0; k < KL_COUNT; k++) b->lcfg[l].thr[k] = 0;
            b->lcfg[l].mg_lds = -1;
        }
        b->cfg_version++;
    }
    return rc;
#else
    (void)steps;
    return LICA_OK;
#endif
}

int lica_batch_get_tuning(lica_batch *b, int64_t *out, int64_t cap, int64_t *count)
{
    int rc = check_batch(b);
    if (rc) return rc;
    const int64_t per = 3 + KL_COUNT;
    const int64_t need = per * b->lanes;
    if (count) *count = need;
    if (!out) return LICA_OK;
    if (cap < need) return LICA_BAD_ARG;
    for (int l = 0; l < b->lanes; l++) {
        const int saved = b->cur_lane;
        b->cur_lane = (b->lanes > 1) ? l : -1;
        int64_t *o = out + per * l;
        o[0] = lane_cases(b, l);
        o[1] = b->lcfg[l].nmax;
        o[2] = mg_lds(b);
        for (int k = 0; k < KL_COUNT; k++) o[3 + k] = block_threads(b, k);
        b->cur_lane = saved;
    }
    return LICA_OK;
}

int lica_batch_use_graph(lica_batch *b, int64_t on)
{
    int rc = check_batch(b);
    if (rc) return rc;
    b->use_graph = on != 0;
    return LICA_OK;
}

int lica_batch_set_timing(lica_batch *b, int64_t on)
{
    int rc = check_batch(b);
    if (rc) return rc;
    b->timing = on != 0;
    return LICA_OK;
}

int lica_batch_sync(lica_batch *b)
{
    int rc = check_batch(b);
    return rc ? rc : dsync(b->stream);
}

int64_t lica_batch_size(lica_batch *b) { return b ? b->ninst : 0; }

int lica_batch_status(lica_batch *b, int64_t *first_bad_inst)
{
    int rc = check_batch(b);
    if (rc) return rc;
    rc = dsync(b->stream);
    return rc ? rc : first_status(b, first_bad_inst);
}

int lica_batch_get_progress(lica_batch *b, int64_t *status_out, int64_t *steps_out)
{
    int rc = check_batch(b);
    if (rc) return rc;
    std::vector<InstState> st((size_t)b->ninst);
    rc = d2h(st.data(), b->d_st, sizeof(InstState) * st.size(), b->stream);
    if (rc) return rc;
    for (int64_t i = 0; i < b->ninst; i++) {
        if (status_out) status_out[i] = st[(size_t)i].status;
        if (steps_out) steps_out[i] = st[(size_t)i].m;
    }
    return LICA_OK;
}

int lica_batch_get_scalars(lica_batch *b, int64_t inst, double *out)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (inst < 0 || inst >= b->ninst || !out) return LICA_BAD_ARG;
    InstState s;
    rc = get_state(b, inst, &s);
    if (rc) return rc;
    const InstParam &p = b->par[(size_t)inst];
    double t01[2], vv1;
    rc = d2h(t01, b->d_ns + p.ns_off + (size_t)NS_T * p.ln + 0 + 2, 2 * sizeof(double), b->stream);
    rc = rc ? rc : d2h(&vv1, b->d_ns + p.ns_off + (size_t)NS_VVDZ * p.ln + 1 + 2, sizeof(double), b->stream);
    if (rc) return rc;
    for (int i = 0; i < LICA_NSCAL; i++) out[i] = 0.0;
    out[LICA_S_TIME] = s.time; out[LICA_S_DT] = s.dt; out[LICA_S_STEP] = (double)s.m;
    out[LICA_S_PLACE] = s.place; out[LICA_S_DVMAX] = s.dvmax; out[LICA_S_CFLM] = s.cflm;
    out[LICA_S_VOL_ORI] = s.vol_tot_ori; out[LICA_S_VOL_IPHS] = s.vol_tot_iphs;
    // wall heat flux k_v dT/dz at the wall face
    out[LICA_S_QWALL] = p.tcm * (t01[1] - t01[0]) * vv1;
    out[LICA_S_STATUS] = (double)s.status;
    out[LICA_S_SUBCYCLE_WARN] = s.n_subcycle_warn; out[LICA_S_MASS_WARN] = s.n_mass_warn;
    out[LICA_S_EXTRAP_HALVINGS] = s.n_halvings; out[LICA_S_EXTRAP_ITERS] = s.n_extrap_iters;
    out[LICA_S_REINIT_CALLS] = s.n_reinit; out[LICA_S_BANDGEN_CALLS] = s.n_bandgen;
    out[LICA_S_MASSCOR_ITERS] = s.n_masscor;
    out[LICA_S_POISSON_CYCLES] = s.poisson_cycles; out[LICA_S_POISSON_SOLVES] = s.poisson_solves;
    out[LICA_S_POISSON_RESID] = s.poisson_resid;
    out[LICA_S_DTCONST] = s.dtconst; out[LICA_S_PSHIFT] = s.pshift;
    return LICA_OK;
}

int lica_batch_get_field(lica_batch *b, int64_t field, int64_t inst, double *out, int64_t cap, int64_t *count)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (inst < 0 || inst >= b->ninst || !out || cap < 0) return LICA_BAD_ARG;
    const InstParam &p = b->par[(size_t)inst];
    int64_t m = 0;
    std::vector<double> buf;
    auto ns_field = [&](int f) {
        buf.resize((size_t)p.n3 + 1);
        int r2 = d2h(buf.data(), b->d_ns + p.ns_off + (size_t)f * p.ln + 2, sizeof(double) * buf.size(), b->stream);
        for (size_t k = 0; k < buf.size() && m < cap; k++) out[m++] = buf[k];
        return r2;
    };
    auto rf_field = [&](int f) {
        buf.resize((size_t)p.n3fm);
        int r2 = d2h(buf.data(), b->d_rf + p.rf_off + (size_t)f * p.lf + (1 - p.m3l), sizeof(double) * buf.size(), b->stream);
        for (size_t k = 0; k < buf.size() && m < cap; k++) out[m++] = buf[k];
        return r2;
    };
    switch (field) {
    case LICA_F_W: rc = ns_field(NS_W); break;
    case LICA_F_P: rc = ns_field(NS_P); break;
    case LICA_F_T: rc = ns_field(NS_T); break;
    case LICA_F_PSI_CN: rc = ns_field(NS_PSI_CN); break;
    case LICA_F_PSI_ZN: rc = ns_field(NS_PSI_ZN); break;
    case LICA_F_WO: rc = ns_field(NS_WO); break;
    case LICA_F_PA: case LICA_F_PB: case LICA_F_PC: case LICA_F_PG: {
        const int f = (field == LICA_F_PA) ? NS_PA : (field == LICA_F_PB) ? NS_PB : (field == LICA_F_PC) ? NS_PC : NS_PG;
        buf.resize((size_t)p.n);
        rc = d2h(buf.data(), b->d_ns + p.ns_off + (size_t)f * p.ln + 1 + 2, sizeof(double) * buf.size(), b->stream);
        for (size_t k = 0; k < buf.size() && m < cap; k++) out[m++] = buf[k];
        break;
    }
    case LICA_F_PPREV: case LICA_F_PSOL: case LICA_F_PRHS: {
        buf.resize((size_t)p.n);
        const int which = (field == LICA_F_PPREV) ? 0 : (field == LICA_F_PSOL) ? 1 : 2;
        const long long o = 3 * p.gs_off + (long long)which * (p.n + 2) + 1;
        rc = d2h(buf.data(), b->d_cap + o, sizeof(double) * buf.size(), b->stream);
        for (size_t k = 0; k < buf.size() && m < cap; k++) out[m++] = buf[k];
        break;
    }
    case LICA_F_PSIF: rc = rf_field(RF_PSIF); break;
    case LICA_F_SFGAS: rc = rf_field(RF_SFGAS); break;
    case LICA_F_PHI: case LICA_F_BAND: case LICA_F_ALPHI: case LICA_F_MFLVS: {
        InstState s;
        rc = get_state(b, inst, &s);
        if (rc) break;
        if (field == LICA_F_PHI) {
            // ALPHI_EXT image: zero off the band, band planes from ALPHI
            std::vector<double> phi((size_t)p.n3fm + 1, 0.0);
            for (int nl = 1; nl <= LICA_MLAYER; nl++)
                for (int q = 0; q < s.bn[nl]; q++) {
                    int k = s.bk[nl][q];
                    if (k >= 1 && k <= p.n3fm) phi[(size_t)k] = s.alphi[nl][q];
                }
            for (int k = 1; k <= p.n3fm && m < cap; k++) out[m++] = phi[(size_t)k];
        } else if (field == LICA_F_BAND) {
            for (int nl = 1; nl <= LICA_MLAYER; nl++) {
                if (m < cap) out[m++] = (double)s.bn[nl];
                for (int q = 0; q < s.bn[nl] && m < cap; q++) out[m++] = (double)s.bk[nl][q];
            }
        } else {
            for (int nl = 1; nl <= LICA_MLAYER; nl++)
                for (int q = 0; q < s.bn[nl] && m < cap; q++)
                    out[m++] = (field == LICA_F_ALPHI) ? s.alphi[nl][q] : s.mflvs[nl][q];
        }
        break;
    }
    default: return LICA_BAD_ARG;
    }
    if (count) *count = m;
    return rc;
}

int lica_batch_last_timings(lica_batch *b, double *ms_out, int64_t cap)
{
    if (!b || !ms_out) return 0;
    int64_t m = 0;
    for (int i = 0; i < KL_COUNT && m < cap; i++) ms_out[m++] = b->timings[i];
    if (m < cap) ms_out[m++] = b->wall_ms;
    return (int)m;
}

void lica_batch_destroy(lica_batch *b)
{
    if (!b) return;
#ifndef LICA_EMUL
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (int l = 1; l < LICA_MAX_LANES; l++)
        if (b->lane_stream[l]) (void)hipStreamSynchronize(b->lane_stream[l]);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    for (int l = 0; l < LICA_MAX_LANES; l++)
        for (hipEvent_t e : b->ev[l]) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (int l = 1; l < LICA_MAX_LANES; l++) {
        if (b->join[l]) (void)hipEventDestroy(b->join[l]);
        if (b->lane_stream[l]) (void)hipStreamDestroy(b->lane_stream[l]);
    }
#endif
    dfree(b->d_par); dfree(b->d_st); dfree(b->d_ns); dfree(b->d_rf);
    dfree(b->d_nsi); dfree(b->d_rfi); dfree(b->d_order); dfree(b->d_order_lanes); dfree(b->d_lead); dfree(b->d_mg);
    dfree(b->d_guess); dfree(b->d_scale); dfree(b->d_cap); dfree(b->d_netcases); dfree(b->d_hist);
#ifndef LICA_EMUL
    if (b->stream) (void)hipStreamDestroy(b->stream);
#endif
    delete b;
}

int lica_batch_set_net(lica_batch *b, lica_net *net)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (!net) return LICA_BAD_ARG;
    std::vector<NetCase> nc((size_t)b->ninst);
    for (int64_t i = 0; i < b->ninst; i++) {
        const InstParam &p = b->par[(size_t)i];
        double *base = b->d_ns + p.ns_off + 2;
        NetCase &c = nc[(size_t)i];
        c.a = base + (size_t)NS_PA * p.ln; c.b = base + (size_t)NS_PB * p.ln;
        c.c = base + (size_t)NS_PC * p.ln; c.g = base + (size_t)NS_PG * p.ln;
        c.guess = b->d_guess + p.gs_off;
        c.place = &b->d_st[i].place;
        c.dzi = 1.0 / ((double)p.r * p.sdzf);
        c.n = p.n;
        c.win = (int)std::min<int64_t>(p.n, lica_net_window());
    }
    if (!b->d_netcases) rc = dmalloc((void **)&b->d_netcases, sizeof(NetCase) * nc.size());
    rc = rc ? rc : h2d(b->d_netcases, nc.data(), sizeof(NetCase) * nc.size(), b->stream);
    if (rc) return rc;
    b->net = net;
    return LICA_OK;
}

int lica_batch_set_capture(lica_batch *b, int64_t on)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (on < 0 || on > 3) return LICA_BAD_ARG;
    b->capture = (int)on;
    b->cap_valid = false;
    return LICA_OK;
}

// per case [first row of the network's window, backward error of the
// warm-start base, s_r, s_b, alpha] of the last network correction
int lica_batch_warm_info(lica_batch *b, double *out, int64_t len)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (!out || len < LICA_WARM_INFO * b->ninst) return LICA_BAD_ARG;
    return d2h(out, b->d_scale, sizeof(double) * LICA_WARM_INFO * (size_t)b->ninst, b->stream);
}

// training windows from the captured systems of every healthy case; needs a
// multigrid mode (the solve leaves the matrix rows in place) and a run with
// capture on since capture was switched on
int lica_net_add_pairs(lica_net *net, lica_batch *b, int64_t windows_per_case, int64_t len, uint64_t seed,
                       int64_t *added)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (!net || windows_per_case < 1 || len < 1) return LICA_BAD_ARG;
    if (!b->capture || !b->cap_valid || b->poisson_mode < LICA_POISSON_MG64) return LICA_BAD_ARG;
    std::vector<InstState> st((size_t)b->ninst);
    rc = d2h(st.data(), b->d_st, sizeof(InstState) * st.size(), b->stream);
    if (rc) return rc;
    std::vector<NetPairSrc> src;
    for (int64_t i = 0; i < b->ninst; i++) {
        if (st[(size_t)i].status != 0) continue;
        const InstParam &p = b->par[(size_t)i];
        const double *base = b->d_ns + p.ns_off + 2;
        const double *cap = b->d_cap + 3 * p.gs_off;
        NetPairSrc s;
        s.a = base + (size_t)NS_PA * p.ln; s.b = base + (size_t)NS_PB * p.ln; s.c = base + (size_t)NS_PC * p.ln;
        s.p0 = cap; s.p1 = cap + (p.n + 2); s.g = cap + 2 * (p.n + 2);
        s.n = p.n;
        s.kif = (int)(st[(size_t)i].place / ((double)p.r * p.sdzf)) + 1;
        src.push_back(s);
    }
    return lica_net_append(net, src.data(), (int)src.size(), windows_per_case, len, seed, added);
}

// ---- single-case API -------------------------------------------------
int lica_init(lica_batch **ctx, const char *lica_in_path, const char *grid_in_path)
{
    if (!ctx) return LICA_BAD_ARG;
    *ctx = nullptr;
    lica_case_desc d;
    int rc = lica_parse_deck(lica_in_path, grid_in_path, &d);
    if (rc) return rc;
    rc = lica_batch_create(1, &d, ctx);
    if (rc) return rc;
    return lica_batch_init(*ctx);
}

int lica_step(lica_batch *ctx, int64_t nsteps) { return lica_batch_run(ctx, nsteps); }

int lica_determine_dt(lica_batch *ctx, double *cflm, double *dt)
{
    int rc = check_batch(ctx);
    if (rc) return rc;
    rc = launch(ctx, KL_DT, 1);
    rc = rc ? rc : dsync(ctx->stream);
    InstState s;
    rc = rc ? rc : get_state(ctx, 0, &s);
    if (rc) return rc;
    if (cflm) *cflm = s.cflm;
    if (dt) *dt = s.dt;
    return s.status;
}

static int copy_ns_out(lica_batch *b, int f, double *out)
{
    if (!out) return 0;
    const InstParam &p = b->par[0];
    return d2h(out, b->d_ns + p.ns_off + (size_t)f * p.ln + 2, sizeof(double) * ((size_t)p.n3 + 1), b->stream);
}

int lica_rhsnlhs(lica_batch *ctx, int64_t msub, double *w, double *p, double *psi_zn, double *psi_cn)
{
    int rc = check_batch(ctx);
    if (rc) return rc;
    if (msub < 1 || msub > 3) return LICA_BAD_ARG;
    ctx->last_msub = (int)msub;
    if (ctx->poisson_mode == LICA_POISSON_LEARNED) {
        if (!ctx->net) return LICA_BAD_ARG;
        if (lica_net_cast(ctx->net, (void *)ctx->stream)) return LICA_HIP_ERROR;
    }
    const int kp = (ctx->poisson_mode >= LICA_POISSON_MG64) ? KL_POISSON_MG : KL_POISSON_DIRECT;
    const int seq[5] = {KL_CONT_ENERGY, KL_TEMP_GRAD, KL_MOMENTUM, kp, KL_UCALC};
    for (int j = 0; j < 5 && !rc; j++) {
        if (seq[j] == kp) rc = pre_poisson(ctx, (int)msub);
        if (!rc) rc = launch(ctx, seq[j], (int)msub);
    }
    rc = rc ? rc : dsync(ctx->stream);
    rc = rc ? rc : copy_ns_out(ctx, NS_W, w);
    rc = rc ? rc : copy_ns_out(ctx, NS_P, p);
    rc = rc ? rc : copy_ns_out(ctx, NS_PSI_ZN, psi_zn);
    rc = rc ? rc : copy_ns_out(ctx, NS_PSI_CN, psi_cn);
    if (rc) return rc;
    int64_t bad;
    return first_status(ctx, &bad);
}

int lica_convrge1(lica_batch *ctx, double *dvmax)
{
    int rc = check_batch(ctx);
    if (rc) return rc;
    InstState s;
    rc = get_state(ctx, 0, &s);
    if (rc) return rc;
    if (dvmax) *dvmax = s.dvmax;
    return LICA_OK;
}

int lica_transport(lica_batch *ctx)
{
    int rc = check_batch(ctx);
    if (rc) return rc;
    rc = launch(ctx, KL_TRANSPORT, ctx->last_msub);
    rc = rc ? rc : dsync(ctx->stream);
    if (rc) return rc;
    int64_t bad;
    return first_status(ctx, &bad);
}

int lica_grid_coupling(lica_batch *ctx, int64_t ini_psi, double *psi_zn, double *psi_cn)
{
    int rc = check_batch(ctx);
    if (rc) return rc;
    if (ini_psi != 0) return LICA_BAD_ARG;   // GRID_COUPLING(1) belongs to lica_init
    rc = copy_ns_out(ctx, NS_PSI_ZN, psi_zn);
    return rc ? rc : copy_ns_out(ctx, NS_PSI_CN, psi_cn);
}

int lica_end_step(lica_batch *ctx)
{
    int rc = check_batch(ctx);
    if (rc) return rc;
    int64_t bad;
    rc = dsync(ctx->stream);
    return rc ? rc : first_status(ctx, &bad);
}

// ---------------------------------------------------------------------
// Seams with the reference's own argument lists.  The arrays are the
// reference's REAL*8 (0:M1,0:M2,0:M3) of the 1-D configuration, M1 = M2 = 2
// (1dmodulesX.f:5), in Fortran order: (I,J,K) at I + 3 (J + 3 K).  The flow,
// the energy and the Poisson solve use the column I = J = 1 (1d.f:1554-1555,
// 3529-3530); each call copies the columns it reads into the context at
// entry and writes the ones the subroutine changes back at exit.  The x/y
// arrays (U, V, PSI_XN, PSI_YN, RK3XO, RK3YO) are identically zero in this
// configuration and are accepted and left alone, as are the channel-flow
// scalars QVOL_ORI and VOL1 (ICH = 0) and ANPSO (D7: read as zero).

static size_t fidx(int64_t k) { return (size_t)(1 + 3 * (1 + 3 * k)); }

static int fcol_in(lica_batch *b, int f, const double *a)
{
    if (!a) return 0;
    const InstParam &p = b->par[0];
    std::vector<double> col((size_t)p.n3 + 1);
    for (int64_t k = 0; k <= p.n3; k++) col[(size_t)k] = a[fidx(k)];
    return h2d(b->d_ns + p.ns_off + (size_t)f * p.ln + 2, col.data(), sizeof(double) * col.size(), b->stream);
}

static int fcol_get(lica_batch *b, int f, std::vector<double> &col)
{
    const InstParam &p = b->par[0];
    col.resize((size_t)p.n3 + 2);
    return d2h(col.data(), b->d_ns + p.ns_off + (size_t)f * p.ln + 2, sizeof(double) * col.size(), b->stream);
}

static int fcol_out(lica_batch *b, int f, double *a)
{
    if (!a) return 0;
    std::vector<double> col;
    int rc = fcol_get(b, f, col);
    if (rc) return rc;
    for (int64_t k = 0; k <= b->par[0].n3; k++) a[fidx(k)] = col[(size_t)k];
    return 0;
}

static int put_state(lica_batch *b, const InstState *s) { return h2d(b->d_st, s, sizeof(InstState), b->stream); }

static int f_single(lica_batch *b) { return (check_batch(b) || b->ninst != 1) ? LICA_BAD_ARG : 0; }

int lica_f_determine_dt(lica_batch *ctx, double *cflm, const double *u, const double *v, const double *w,
                        const double *psi_cn, double *dt)
{
    (void)u; (void)v;
    int rc = f_single(ctx);
    if (rc) return rc;
    rc = fcol_in(ctx, NS_W, w);
    rc = rc ? rc : fcol_in(ctx, NS_PSI_CN, psi_cn);
    InstState s;
    rc = rc ? rc : get_state(ctx, 0, &s);
    if (rc) return rc;
    if (s.status) return s.status;
    if (w) {
        // the CFL input of DETERMINE_DT from the caller's W (1d.f:1325-1373),
        // max |W(k)+W(k+1)| SSDZ(k) as the projection records it
        std::vector<double> ssdz;
        rc = fcol_get(ctx, NS_SSDZ, ssdz);
        if (rc) return rc;
        double c3 = 0.0;
        for (int64_t k = 1; k <= ctx->par[0].n; k++)
            c3 = fmax(c3, fabs(w[fidx(k)] + w[fidx(k + 1)]) * ssdz[(size_t)k]);
        s.c3max = c3;
    }
    if (dt) s.dt = *dt;
    rc = put_state(ctx, &s);
    rc = rc ? rc : launch(ctx, KL_DT, 1);
    rc = rc ? rc : dsync(ctx->stream);
    rc = rc ? rc : get_state(ctx, 0, &s);
    if (rc) return rc;
    if (cflm) *cflm = s.cflm;
    if (dt) *dt = s.dt;
    return s.status;
}

int lica_f_rhsnlhs(lica_batch *ctx, const int64_t *msub, const int64_t *itracking, double *u, double *v, double *w,
                   double *p, double *psi_xn, double *psi_yn, double *psi_zn, double *psi_cn, double *rk3xo,
                   double *rk3yo, double *rk3zo, double *anpso, double *qvol_ori, double *vol1)
{
    (void)itracking; (void)u; (void)v; (void)psi_xn; (void)psi_yn; (void)rk3xo; (void)rk3yo; (void)anpso;
    (void)qvol_ori; (void)vol1;
    int rc = f_single(ctx);
    if (rc) return rc;
    if (!msub) return LICA_BAD_ARG;
    rc = fcol_in(ctx, NS_W, w);
    rc = rc ? rc : fcol_in(ctx, NS_P, p);
    rc = rc ? rc : fcol_in(ctx, NS_PSI_ZN, psi_zn);
    rc = rc ? rc : fcol_in(ctx, NS_PSI_CN, psi_cn);
    rc = rc ? rc : fcol_in(ctx, NS_RK3ZO, rk3zo);
    if (rc) return rc;
    const int st = lica_rhsnlhs(ctx, *msub, nullptr, nullptr, nullptr, nullptr);
    if (st && !(st > 0 && st < 20)) return st;
    rc = fcol_out(ctx, NS_W, w);
    rc = rc ? rc : fcol_out(ctx, NS_P, p);
    rc = rc ? rc : fcol_out(ctx, NS_RK3ZO, rk3zo);
    return rc ? rc : st;
}

int lica_f_convrge1(lica_batch *ctx, double *dvmax, const double *u, const double *v, const double *w)
{
    (void)u; (void)v;
    int rc = f_single(ctx);
    if (rc) return rc;
    rc = fcol_in(ctx, NS_W, w);
    rc = rc ? rc : launch(ctx, KL_CONVRGE1, ctx->last_msub);
    rc = rc ? rc : dsync(ctx->stream);
    InstState s;
    rc = rc ? rc : get_state(ctx, 0, &s);
    if (rc) return rc;
    if (dvmax) *dvmax = s.dvmax;
    return s.status;
}

int lica_f_transport(lica_batch *ctx, const double *u, const double *v, const double *w, double *psi_cn,
                     double *vol_tot_ori)
{
    (void)u; (void)v; (void)psi_cn;
    int rc = f_single(ctx);
    if (rc) return rc;
    if (w) {
        // VEL_INTP's time-averaged velocity was prepared by the projection
        // from the W the last RHSNLHS returned; TRANSPORT must get that W
        std::vector<double> cur;
        rc = fcol_get(ctx, NS_W, cur);
        if (rc) return rc;
        for (int64_t k = 0; k <= ctx->par[0].n3; k++)
            if (!(w[fidx(k)] == cur[(size_t)k])) return LICA_BAD_ARG;
    }
    if (vol_tot_ori) {
        InstState s;
        rc = get_state(ctx, 0, &s);
        if (rc) return rc;
        s.vol_tot_ori = *vol_tot_ori;
        rc = put_state(ctx, &s);
        if (rc) return rc;
    }
    const int st = lica_transport(ctx);
    if (st && !(st > 0 && st < 20)) return st;
    if (vol_tot_ori) {
        InstState s;
        rc = get_state(ctx, 0, &s);
        if (rc) return rc;
        *vol_tot_ori = s.vol_tot_ori;
    }
    return st;
}

int lica_f_grid_coupling(lica_batch *ctx, const int64_t *ini_psi, double *psi_xn, double *psi_yn, double *psi_zn,
                         double *psi_cn)
{
    (void)psi_xn; (void)psi_yn;
    int rc = f_single(ctx);
    if (rc) return rc;
    if (!ini_psi || *ini_psi != 0) return LICA_BAD_ARG;   // GRID_COUPLING(1) belongs to lica_init
    rc = fcol_out(ctx, NS_PSI_ZN, psi_zn);
    return rc ? rc : fcol_out(ctx, NS_PSI_CN, psi_cn);
}

int lica_f_poisson(lica_batch *ctx, double *p, const double *divgsum, const double *denf_xi, const double *denf_yi,
                   const double *denf_zi, const double *denf_y, const double *denf_c)
{
    (void)denf_xi; (void)denf_yi; (void)denf_y; (void)denf_c;
    int rc = f_single(ctx);
    if (rc) return rc;
    if (!p || !divgsum || !denf_zi) return LICA_BAD_ARG;
    if (ctx->poisson_mode == LICA_POISSON_LEARNED || ctx->poisson_mode == LICA_POISSON_PROJECTED)
        return LICA_UNSUPPORTED_DECK;   // the warm starts belong to the substep sequence
    const InstParam &ip = ctx->par[0];
    const int64_t n = ip.n;
    std::vector<double> vvdz, ssdz;
    rc = fcol_get(ctx, NS_VVDZ, vvdz);
    rc = rc ? rc : fcol_get(ctx, NS_SSDZ, ssdz);
    if (rc) return rc;
    // the rows of 1d.f:3529-3543: a(1) = 0 (FIXKL(1) = 1), c from DENF_ZI
    // above each face, b = -(a + c), g = DIVGSUM(1,1,K), K = 1..N3M
    std::vector<double> a((size_t)n + 2, 0.0), bb((size_t)n + 2, 0.0), c((size_t)n + 2, 0.0), g((size_t)n + 2, 0.0);
    for (int64_t k = 1; k <= n; k++) {
        const double fixkl = (k == 1) ? 1.0 : 0.0;
        const double ak = vvdz[(size_t)k] * ssdz[(size_t)k] * denf_zi[fidx(k)] * (1.0 - fixkl);
        const double ck = vvdz[(size_t)k + 1] * ssdz[(size_t)k] * denf_zi[fidx(k + 1)];
        a[(size_t)k] = ak; c[(size_t)k] = ck; bb[(size_t)k] = -(ak + ck); g[(size_t)k] = divgsum[k - 1];
    }
    const size_t bytes = sizeof(double) * ((size_t)n + 2);
    double *base = ctx->d_ns + ip.ns_off + 2;
    rc = h2d(base + (size_t)NS_PA * ip.ln, a.data(), bytes, ctx->stream);
    rc = rc ? rc : h2d(base + (size_t)NS_PB * ip.ln, bb.data(), bytes, ctx->stream);
    rc = rc ? rc : h2d(base + (size_t)NS_PC * ip.ln, c.data(), bytes, ctx->stream);
    rc = rc ? rc : h2d(base + (size_t)NS_PG * ip.ln, g.data(), bytes, ctx->stream);
    const int kp = (ctx->poisson_mode >= LICA_POISSON_MG64) ? KL_POISSON_MG : KL_POISSON_DIRECT;
    rc = rc ? rc : launch(ctx, kp, ctx->last_msub);
    rc = rc ? rc : dsync(ctx->stream);
    InstState s;
    rc = rc ? rc : get_state(ctx, 0, &s);
    if (rc) return rc;
    if (s.status) return s.status;
    std::vector<double> col;
    rc = fcol_get(ctx, NS_P, col);
    if (rc) return rc;
    for (int64_t k = 0; k <= ip.n3; k++) p[fidx(k)] = col[(size_t)k];
    return 0;
}

int lica_f_state(lica_batch *ctx, double *w, double *p, double *t, double *psi_zn, double *psi_cn,
                 double *vol_tot_ori, double *time, double *dt)
{
    int rc = f_single(ctx);
    rc = rc ? rc : dsync(ctx->stream);
    rc = rc ? rc : fcol_out(ctx, NS_W, w);
    rc = rc ? rc : fcol_out(ctx, NS_P, p);
    rc = rc ? rc : fcol_out(ctx, NS_T, t);
    rc = rc ? rc : fcol_out(ctx, NS_PSI_ZN, psi_zn);
    rc = rc ? rc : fcol_out(ctx, NS_PSI_CN, psi_cn);
    InstState s;
    rc = rc ? rc : get_state(ctx, 0, &s);
    if (rc) return rc;
    if (vol_tot_ori) *vol_tot_ori = s.vol_tot_ori;
    if (time) *time = s.time;
    if (dt) *dt = s.dt;
    return 0;
}

int lica_f_temp_grad(lica_batch *ctx, const double *psi_cn)
{
    int rc = f_single(ctx);
    if (rc) return rc;
    rc = fcol_in(ctx, NS_PSI_CN, psi_cn);
    rc = rc ? rc : launch(ctx, KL_TEMP_GRAD, ctx->last_msub);
    rc = rc ? rc : dsync(ctx->stream);
    if (rc) return rc;
    int64_t bad;
    return first_status(ctx, &bad);
}

}  // extern "C"

// ---------------------------------------------------------------------
// Line solvers of the 3-D code (1d.f:3721-4034), column-major m1 x m2
// arrays with 1-based ranges as in the reference
namespace {
bool line_shape_ok(int64_t m1, int64_t m2)
{
    return m1 >= 1 && m2 >= 1 && m1 <= INT32_MAX && m2 <= INT32_MAX && (double)m1 * (double)m2 * 8.0 < 1.0e12;
}
bool range_ok(int64_t lo, int64_t hi, int64_t m, int64_t min_len)
{
    return lo >= 1 && hi <= m && hi - lo + 1 >= min_len;
}

// device copies of host arrays (freed with the holder)
struct LineBufs {
    std::vector<double *> p;
    size_t bytes = 0;
    ~LineBufs()
    {
        for (double *x : p) dfree(x);
    }
    double *add(const double *h, hipStream_t st, int *rc)
    {
        double *d = nullptr;
        if (*rc) return nullptr;
        *rc = dmalloc((void **)&d, bytes);
        if (*rc) return nullptr;
        p.push_back(d);
        if (h) *rc = h2d(d, h, bytes, st);
        return d;
    }
};

int trdiag_host(const double *a, const double *b, const double *c, double *r, double *uu, int64_t m1, int64_t m2,
                int64_t l1, int64_t l2, int64_t ll1, int64_t ll2, int overwrite_r)
{
    if (!a || !b || !c || !r || !uu) return LICA_BAD_ARG;
    if (!line_shape_ok(m1, m2) || !range_ok(l1, l2, m2, 1) || !range_ok(ll1, ll2, m1, 1)) return LICA_BAD_ARG;
    if (!have_device()) return LICA_NO_DEVICE;
    hipStream_t st = nullptr;
    LineBufs B;
    B.bytes = sizeof(double) * (size_t)m1 * (size_t)m2;
    int rc = 0;
    double *da = B.add(a, st, &rc), *db = B.add(b, st, &rc), *dc = B.add(c, st, &rc), *dr = B.add(r, st, &rc);
    double *du = B.add(uu, st, &rc), *dg = B.add(nullptr, st, &rc);
    if (!rc && lica_launch_trdiag_lines(da, db, dc, dr, du, dg, m1, (int)l1, (int)l2, (int)ll1, (int)ll2, overwrite_r,
                                        (void *)st))
        rc = LICA_HIP_ERROR;
    rc = rc ? rc : dsync(st);
    rc = rc ? rc : d2h(uu, du, B.bytes, st);
    if (overwrite_r) rc = rc ? rc : d2h(r, dr, B.bytes, st);
    return rc;
}

int trdiag_periodic_host(const double *a, const double *b, const double *c, double *f, int64_t m1, int64_t m2,
                         int64_t j1, int64_t j2, int64_t l1, int64_t l2)
{
    if (!a || !b || !c || !f) return LICA_BAD_ARG;
    if (!line_shape_ok(m1, m2) || !range_ok(j1, j2, m2, 2) || !range_ok(l1, l2, m1, 1)) return LICA_BAD_ARG;
    if (!have_device()) return LICA_NO_DEVICE;
    hipStream_t st = nullptr;
    LineBufs B;
    B.bytes = sizeof(double) * (size_t)m1 * (size_t)m2;
    int rc = 0;
    double *da = B.add(a, st, &rc), *db = B.add(b, st, &rc), *dc = B.add(c, st, &rc), *df = B.add(f, st, &rc);
    double *dq = B.add(nullptr, st, &rc), *ds = B.add(nullptr, st, &rc), *de = B.add(nullptr, st, &rc);
    if (!rc && lica_launch_trdiag_periodic_lines(da, db, dc, df, dq, ds, de, m1, (int)j1, (int)j2, (int)l1, (int)l2,
                                                 (void *)st))
        rc = LICA_HIP_ERROR;
    rc = rc ? rc : dsync(st);
    rc = rc ? rc : d2h(f, df, B.bytes, st);
    return rc;
}
}  // namespace

extern "C" int lica_trdiag1(const double *a, const double *b, const double *c, const double *r, double *uu, int64_t m1,
                            int64_t m2, int64_t l1, int64_t l2, int64_t ll1, int64_t ll2)
{
    return trdiag_host(a, b, c, const_cast<double *>(r), uu, m1, m2, l1, l2, ll1, ll2, 0);
}
extern "C" int lica_trdiag2(const double *a, const double *b, const double *c, const double *r, double *uu, int64_t m1,
                            int64_t m2, int64_t l1, int64_t l2, int64_t ll1, int64_t ll2)
{
    return trdiag_host(a, b, c, const_cast<double *>(r), uu, m1, m2, l1, l2, ll1, ll2, 0);
}
extern "C" int lica_trdiag3(const double *a, const double *b, const double *c, double *r, double *uu, int64_t m1,
                            int64_t m2, int64_t l1, int64_t l2, int64_t ll1, int64_t ll2)
{
    return trdiag_host(a, b, c, r, uu, m1, m2, l1, l2, ll1, ll2, 1);
}
extern "C" int lica_trdiag1p(const double *a, const double *b, const double *c, double *f, int64_t m1, int64_t m2,
                             int64_t j1, int64_t j2, int64_t l1, int64_t l2)
{
    return trdiag_periodic_host(a, b, c, f, m1, m2, j1, j2, l1, l2);
}
extern "C" int lica_trdiag2p(const double *a, const double *b, const double *c, double *f, int64_t m1, int64_t m2,
                             int64_t j1, int64_t j2, int64_t l1, int64_t l2)
{
    return trdiag_periodic_host(a, b, c, f, m1, m2, j1, j2, l1, l2);
}
extern "C" int lica_trdiag3p(const double *a, const double *b, const double *c, double *f, int64_t m1, int64_t m2,
                             int64_t j1, int64_t j2, int64_t l1, int64_t l2)
{
    return trdiag_periodic_host(a, b, c, f, m1, m2, j1, j2, l1, l2);
}
extern "C" int lica_ptdiag1a(const double *a, const double *b, const double *c, const double *d, const double *e,
                             const double *f, double *uu, int64_t m1, int64_t m2, int64_t i1, int64_t in, int64_t j1,
                             int64_t jn, double *o, double *q, double *r)
{
    if (!a || !b || !c || !d || !e || !f || !uu || !o || !q || !r) return LICA_BAD_ARG;
    if (!line_shape_ok(m1, m2) || !range_ok(j1, jn, m2, 2) || !range_ok(i1, in, m1, 1)) return LICA_BAD_ARG;
    if (!have_device()) return LICA_NO_DEVICE;
    hipStream_t st = nullptr;
    LineBufs B;
    B.bytes = sizeof(double) * (size_t)m1 * (size_t)m2;
    int rc = 0;
    double *da = B.add(a, st, &rc), *db = B.add(b, st, &rc), *dc = B.add(c, st, &rc), *dd = B.add(d, st, &rc);
    double *de = B.add(e, st, &rc), *df = B.add(f, st, &rc), *du = B.add(uu, st, &rc), *dO = B.add(o, st, &rc);
    double *dq = B.add(q, st, &rc), *dr = B.add(r, st, &rc);
    if (!rc && lica_launch_ptdiag_lines(da, db, dc, dd, de, df, du, dO, dq, dr, m1, (int)i1, (int)in, (int)j1, (int)jn,
                                        (void *)st))
        rc = LICA_HIP_ERROR;
    rc = rc ? rc : dsync(st);
    rc = rc ? rc : d2h(uu, du, B.bytes, st);
    rc = rc ? rc : d2h(o, dO, B.bytes, st);
    rc = rc ? rc : d2h(q, dq, B.bytes, st);
    rc = rc ? rc : d2h(r, dr, B.bytes, st);
    return rc;
}

extern "C" int lica_trdiag_lines_dev(const double *a, const double *b, const double *c, double *r, double *uu,
                                     double *work, int64_t m1, int64_t m2, int64_t l1, int64_t l2, int64_t ll1,
                                     int64_t ll2, int64_t overwrite_r, void *stream)
{
    if (!a || !b || !c || !r || !uu || !work || (overwrite_r != 0 && overwrite_r != 1)) return LICA_BAD_ARG;
    if (!line_shape_ok(m1, m2) || !range_ok(l1, l2, m2, 1) || !range_ok(ll1, ll2, m1, 1)) return LICA_BAD_ARG;
    if (!have_device()) return LICA_NO_DEVICE;
    return lica_launch_trdiag_lines(a, b, c, r, uu, work, m1, (int)l1, (int)l2, (int)ll1, (int)ll2, (int)overwrite_r,
                                    stream) ? LICA_HIP_ERROR : LICA_OK;
}
extern "C" int lica_trdiag_periodic_lines_dev(const double *a, const double *b, const double *c, double *f,
                                              double *work, int64_t m1, int64_t m2, int64_t j1, int64_t j2,
                                              int64_t l1, int64_t l2, void *stream)
{
    if (!a || !b || !c || !f || !work) return LICA_BAD_ARG;
    if (!line_shape_ok(m1, m2) || !range_ok(j1, j2, m2, 2) || !range_ok(l1, l2, m1, 1)) return LICA_BAD_ARG;
    if (!have_device()) return LICA_NO_DEVICE;
    const size_t sz = (size_t)m1 * (size_t)m2;
    return lica_launch_trdiag_periodic_lines(a, b, c, f, work, work + sz, work + 2 * sz, m1, (int)j1, (int)j2,
                                             (int)l1, (int)l2, stream) ? LICA_HIP_ERROR : LICA_OK;
}
extern "C" int lica_ptdiag_lines_dev(const double *a, const double *b, const double *c, const double *d,
                                     const double *e, const double *f, double *uu, int64_t m1, int64_t m2, int64_t i1,
                                     int64_t in, int64_t j1, int64_t jn, double *o, double *q, double *r,
                                     void *stream)
{
    if (!a || !b || !c || !d || !e || !f || !uu || !o || !q || !r) return LICA_BAD_ARG;
    if (!line_shape_ok(m1, m2) || !range_ok(j1, jn, m2, 2) || !range_ok(i1, in, m1, 1)) return LICA_BAD_ARG;
    if (!have_device()) return LICA_NO_DEVICE;
    return lica_launch_ptdiag_lines(a, b, c, d, e, f, uu, o, q, r, m1, (int)i1, (int)in, (int)j1, (int)jn, stream)
               ? LICA_HIP_ERROR : LICA_OK;
}

extern "C" int lica_poisson(int64_t nsys, int64_t n, const double *a, const double *b, const double *c,
                            const double *g, double *p, int64_t mode, double rtol, int64_t max_cycles,
                            int64_t *cycles_out)
{
    if (nsys <= 0 || n < 1 || !a || !b || !c || !g || !p) return LICA_BAD_ARG;
    if (mode < LICA_POISSON_THOMAS || mode > LICA_POISSON_MIXED) return LICA_UNSUPPORTED_DECK;
    if (mode >= LICA_POISSON_MG64 && (!(rtol > 0.0) || max_cycles < 1)) return LICA_BAD_ARG;
    if (!have_device()) return LICA_NO_DEVICE;
    hipStream_t st = nullptr;
#ifndef LICA_EMUL
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return LICA_HIP_ERROR;
#endif
    const size_t tot = (size_t)nsys * (size_t)n, bytes = tot * sizeof(double);
    double *da = nullptr, *db = nullptr, *dc = nullptr, *dg = nullptr;
    int rc = dmalloc((void **)&da, bytes);
    rc = rc ? rc : dmalloc((void **)&db, bytes);
    rc = rc ? rc : dmalloc((void **)&dc, bytes);
    rc = rc ? rc : dmalloc((void **)&dg, bytes);
    rc = rc ? rc : h2d(da, a, bytes, st);
    rc = rc ? rc : h2d(db, b, bytes, st);
    rc = rc ? rc : h2d(dc, c, bytes, st);
    rc = rc ? rc : h2d(dg, g, bytes, st);
    double *dp = nullptr, *dmg = nullptr, *dres = nullptr;
    int *dcyc = nullptr;
    int64_t maxcyc = 0;
    if (mode >= LICA_POISSON_MG64) {
        const size_t sc = (size_t)lica_mg_scratch_doubles((int)n) * (size_t)nsys;
        rc = rc ? rc : dmalloc((void **)&dp, bytes);
        rc = rc ? rc : dmalloc((void **)&dmg, sizeof(double) * sc);
        rc = rc ? rc : dmalloc((void **)&dres, sizeof(double) * (size_t)nsys);
        rc = rc ? rc : dmalloc((void **)&dcyc, sizeof(int) * (size_t)nsys);
        if (!rc && lica_launch_mg((int)nsys, (int)n, da, db, dc, dg, dp, dmg, (int)mode, rtol, (int)max_cycles, dcyc,
                                  dres, 256, (void *)st))
            rc = LICA_HIP_ERROR;
        rc = rc ? rc : dsync(st);
        rc = rc ? rc : d2h(p, dp, bytes, st);
        std::vector<int> cyc((size_t)nsys);
        rc = rc ? rc : d2h(cyc.data(), dcyc, sizeof(int) * cyc.size(), st);
        for (int cv : cyc) {
            if (cv < 0) { maxcyc = max_cycles; if (!rc) rc = LICA_NOT_CONVERGED; }
            else maxcyc = std::max<int64_t>(maxcyc, cv);
        }
    } else {
        if (!rc && lica_launch_tri((int)nsys, (int)n, da, db, dc, dg, mode == LICA_POISSON_THOMAS ? 0 : 1, 256, (void *)st))
            rc = LICA_HIP_ERROR;
        rc = rc ? rc : dsync(st);
        rc = rc ? rc : d2h(p, dg, bytes, st);
    }
    dfree(da); dfree(db); dfree(dc); dfree(dg); dfree(dp); dfree(dmg); dfree(dres); dfree(dcyc);
#ifndef LICA_EMUL
    (void)hipStreamDestroy(st);
#endif
    if (cycles_out) *cycles_out = maxcyc;
    return rc;
}

// The same solves without any allocation: device arrays in and out, the
// kernels' workspace in the fixed per-GPU array (lica_noalloc_scratch), the
// caller's stream.  Nothing in the workspace needs to start at zero: the
// kernels write every element before they read it.  Calls that overlap on
// one GPU share the workspace and race.
extern "C" int lica_poisson_noalloc(int64_t nsys, int64_t n, const double *a, const double *b, const double *c,
                                    const double *g, double *p, int64_t mode, double rtol, int64_t max_cycles,
                                    int64_t *cycles_out, void *stream)
{
    if (nsys <= 0 || n < 1 || !a || !b || !c || !g || !p) return LICA_BAD_ARG;
    if (mode < LICA_POISSON_THOMAS || mode > LICA_POISSON_MIXED) return LICA_UNSUPPORTED_DECK;
    if (mode >= LICA_POISSON_MG64 && (!(rtol > 0.0) || max_cycles < 1)) return LICA_BAD_ARG;
    if (nsys > INT32_MAX || n > INT32_MAX) return LICA_TOO_LARGE;
    if (!have_device()) return LICA_NO_DEVICE;
    double *scr = nullptr;
    if (lica_noalloc_scratch(&scr)) return LICA_HIP_ERROR;
    const size_t cap = LICA_NOALLOC_SCRATCH_BYTES / sizeof(double);
    const size_t tot = (size_t)nsys * (size_t)n, bytes = tot * sizeof(double);
    hipStream_t st = (hipStream_t)stream;
    int64_t maxcyc = 0;
    int rc = 0;
    if (mode >= LICA_POISSON_MG64) {
        const size_t sc = (size_t)lica_mg_scratch_doubles((int)n);
        if (sc > cap / (size_t)nsys || (size_t)nsys * (sc + 2) > cap) return LICA_TOO_LARGE;
        double *dmg = scr, *dres = scr + (size_t)nsys * sc;
        int *dcyc = (int *)(dres + (size_t)nsys);
        if (lica_launch_mg((int)nsys, (int)n, a, b, c, g, p, dmg, (int)mode, rtol, (int)max_cycles, dcyc, dres, 256,
                           (void *)st))
            return LICA_HIP_ERROR;
        // the cycle counts come back through a fixed host buffer
        int buf[256];
        bool capped = false;
        for (int64_t i0 = 0; i0 < nsys && !rc; i0 += 256) {
            const int64_t m = std::min<int64_t>(256, nsys - i0);
            rc = d2h(buf, dcyc + i0, sizeof(int) * (size_t)m, st);
            for (int64_t i = 0; i < m && !rc; i++) {
                if (buf[i] < 0) { maxcyc = max_cycles; capped = true; }
                else maxcyc = std::max<int64_t>(maxcyc, buf[i]);
            }
        }
        if (!rc && capped) rc = LICA_NOT_CONVERGED;
    } else {
        // the direct solves work in place: a, b, c, g go to the workspace first
        if (tot > cap / 4) return LICA_TOO_LARGE;
        double *w[4] = {scr, scr + tot, scr + 2 * tot, scr + 3 * tot};
        const double *src[4] = {a, b, c, g};
        for (int i = 0; i < 4 && !rc; i++) rc = d2d(w[i], src[i], bytes, st);
        if (!rc && lica_launch_tri((int)nsys, (int)n, w[0], w[1], w[2], w[3], mode == LICA_POISSON_THOMAS ? 0 : 1,
                                   256, (void *)st))
            rc = LICA_HIP_ERROR;
        rc = rc ? rc : d2d(p, w[3], bytes, st);
        rc = rc ? rc : dsync(st);
    }
    if (cycles_out) *cycles_out = maxcyc;
    return rc;
}

// ---------------------------------------------------------------------
// Checkpoint / restart.  The binary checkpoint holds every byte of the
// batch's device state a later step reads (fields, band, scalars, counters,
// the Poisson mode with its tolerance and cycle cap, and in the
// warm-started modes the solution history and the guess rows), so that a
// restart continues bit for bit; the multigrid scratch, the network's
// per-substep scales and the capture rows are rebuilt by every substep and
// are not kept, nor is the network (a batch does not own it: a restart in
// the learned mode needs lica_batch_set_net before it runs).  The field
// file is the text format of WRITEFIELD (1d.f:1121-1232), which post.f's
// READFLD (post.f:238-401) reads.

namespace {
const char CKPT_MAGIC[8] = {'L', 'I', 'C', 'A', 'C', 'K', 'P', '2'};

long long guess_doubles(const lica_batch *b)
{
    long long t = 0;
    for (const InstParam &p : b->par) t += p.n + 2;
    return t;
}
long long hist_doubles(const lica_batch *b)
{
    long long t = 0;
    for (const InstParam &p : b->par) t += 3LL * MG_HIST * p.n;
    return t;
}

template <class T>
bool put(FILE *f, const T *p, size_t n) { return fwrite(p, sizeof(T), n, f) == n; }
template <class T>
bool get(FILE *f, T *p, size_t n) { return fread(p, sizeof(T), n, f) == n; }
}  // namespace

extern "C" int lica_batch_save(lica_batch *b, const char *path)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (!path || !b->initialized) return LICA_BAD_ARG;
    rc = dsync(b->stream);
    if (rc) return rc;
    std::vector<InstState> st((size_t)b->ninst);
    std::vector<double> ns(b->ns_total), rf(b->rf_total);
    std::vector<int> nsi(b->nsi_total), rfi(b->rfi_total);
    rc = d2h(st.data(), b->d_st, sizeof(InstState) * st.size(), b->stream);
    rc = rc ? rc : d2h(ns.data(), b->d_ns, sizeof(double) * ns.size(), b->stream);
    rc = rc ? rc : d2h(rf.data(), b->d_rf, sizeof(double) * rf.size(), b->stream);
    rc = rc ? rc : d2h(nsi.data(), b->d_nsi, sizeof(int) * nsi.size(), b->stream);
    rc = rc ? rc : d2h(rfi.data(), b->d_rfi, sizeof(int) * rfi.size(), b->stream);
    if (rc) return rc;
    FILE *f = fopen(path, "wb");
    if (!f) return LICA_IO;
    const int64_t hdr[6] = {b->ninst, (int64_t)sizeof(InstState), (int64_t)b->ns_total, (int64_t)b->rf_total,
                            (int64_t)b->nsi_total, (int64_t)b->rfi_total};
    bool ok = put(f, CKPT_MAGIC, 8) && put(f, hdr, 6) && put(f, b->cases.data(), b->cases.size()) &&
              put(f, st.data(), st.size()) && put(f, ns.data(), ns.size()) && put(f, rf.data(), rf.size()) &&
              put(f, nsi.data(), nsi.size()) && put(f, rfi.data(), rfi.size());
    // the Poisson solve's setting and, when kept, the warm-start history
    const int64_t pm[3] = {b->poisson_mode, b->max_cycles, b->d_hist ? 1 : 0};
    ok = ok && put(f, pm, 3) && put(f, &b->rtol, 1);
    if (ok && b->d_hist) {
        std::vector<double> hist((size_t)hist_doubles(b)), guess((size_t)guess_doubles(b));
        rc = d2h(hist.data(), b->d_hist, sizeof(double) * hist.size(), b->stream);
        rc = rc ? rc : d2h(guess.data(), b->d_guess, sizeof(double) * guess.size(), b->stream);
        ok = !rc && put(f, hist.data(), hist.size()) && put(f, guess.data(), guess.size());
    }
    ok = (fclose(f) == 0) && ok;
    if (rc) return rc;
    return ok ? LICA_OK : LICA_IO;
}

extern "C" int lica_batch_load(const char *path, lica_batch **out)
{
    if (!path || !out) return LICA_BAD_ARG;
    *out = nullptr;
    FILE *f = fopen(path, "rb");
    if (!f) return LICA_IO;
    char magic[8];
    int64_t hdr[6];
    if (!get(f, magic, 8) || memcmp(magic, CKPT_MAGIC, 8) != 0 || !get(f, hdr, 6) || hdr[0] <= 0 ||
        hdr[1] != (int64_t)sizeof(InstState)) {
        fclose(f);
        return LICA_IO;
    }
    std::vector<lica_case_desc> cases((size_t)hdr[0]);
    if (!get(f, cases.data(), cases.size())) { fclose(f); return LICA_IO; }
    lica_batch *b = nullptr;
    int rc = lica_batch_create(hdr[0], cases.data(), &b);
    if (rc) { fclose(f); return rc; }
    if ((int64_t)b->ns_total != hdr[2] || (int64_t)b->rf_total != hdr[3] || (int64_t)b->nsi_total != hdr[4] ||
        (int64_t)b->rfi_total != hdr[5]) {
        fclose(f);
        lica_batch_destroy(b);
        return LICA_IO;
    }
    std::vector<InstState> st((size_t)hdr[0]);
    std::vector<double> ns(b->ns_total), rf(b->rf_total);
    std::vector<int> nsi(b->nsi_total), rfi(b->rfi_total);
    bool ok = get(f, st.data(), st.size()) && get(f, ns.data(), ns.size()) && get(f, rf.data(), rf.size()) &&
              get(f, nsi.data(), nsi.size()) && get(f, rfi.data(), rfi.size());
    int64_t pm[3] = {0, 0, 0};
    double rtol = 0.0;
    ok = ok && get(f, pm, 3) && get(f, &rtol, 1) && pm[0] >= LICA_POISSON_THOMAS && pm[0] <= LICA_POISSON_PROJECTED &&
         pm[1] >= 1 && rtol > 0.0;
    std::vector<double> hist, guess;
    if (ok && pm[2]) {
        hist.resize((size_t)hist_doubles(b));
        guess.resize((size_t)guess_doubles(b));
        ok = get(f, hist.data(), hist.size()) && get(f, guess.data(), guess.size());
    }
    fclose(f);
    if (!ok) { lica_batch_destroy(b); return LICA_IO; }
    // the mode as lica_batch_set_poisson would set it (its scratch and
    // history allocated); the learned mode's network is the caller's to set
    const int64_t mode = pm[0];
    rc = lica_batch_set_poisson(b, mode == LICA_POISSON_LEARNED ? LICA_POISSON_PROJECTED : mode, rtol, pm[1]);
    if (!rc && mode == LICA_POISSON_LEARNED) b->poisson_mode = LICA_POISSON_LEARNED;
    if (!rc && pm[2]) {
        if (!b->d_hist) rc = LICA_IO;
        rc = rc ? rc : h2d(b->d_hist, hist.data(), sizeof(double) * hist.size(), b->stream);
        rc = rc ? rc : h2d(b->d_guess, guess.data(), sizeof(double) * guess.size(), b->stream);
    }
    if (rc) { lica_batch_destroy(b); return rc; }
    rc = h2d(b->d_st, st.data(), sizeof(InstState) * st.size(), b->stream);
    rc = rc ? rc : h2d(b->d_ns, ns.data(), sizeof(double) * ns.size(), b->stream);
    rc = rc ? rc : h2d(b->d_rf, rf.data(), sizeof(double) * rf.size(), b->stream);
    rc = rc ? rc : h2d(b->d_nsi, nsi.data(), sizeof(int) * nsi.size(), b->stream);
    rc = rc ? rc : h2d(b->d_rfi, rfi.data(), sizeof(int) * rfi.size(), b->stream);
    if (rc) { lica_batch_destroy(b); return rc; }
    b->initialized = true;
    *out = b;
    return LICA_OK;
}

namespace {
// Fortran list of REAL*8 in rows of `per` ESw.12 fields (FORMAT(nES20.12))
void fld_reals(FILE *f, const std::vector<double> &v, int per)
{
    for (size_t i = 0; i < v.size(); i++) {
        fprintf(f, "%20.12E", v[i]);
        if ((i + 1) % (size_t)per == 0 || i + 1 == v.size()) fputc('\n', f);
    }
    if (v.empty()) fputc('\n', f);
}
void fld_ints(FILE *f, const std::vector<long long> &v, int per)
{
    for (size_t i = 0; i < v.size(); i++) {
        fprintf(f, "%8lld", v[i]);
        if ((i + 1) % (size_t)per == 0 || i + 1 == v.size()) fputc('\n', f);
    }
    if (v.empty()) fputc('\n', f);
}
}  // namespace

// WRITEFIELD (1d.f:1121-1232) for case inst: the NS grid has N1=N2=2 nodes
// (one cell) in x and y, every x/y copy of the column holds the same values,
// U=V=0; the band list is written plane by plane with its r x r columns.
extern "C" int lica_batch_write_fld(lica_batch *b, int64_t inst, const char *path, int64_t ihist)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (inst < 0 || inst >= b->ninst || !path) return LICA_BAD_ARG;
    const InstParam &p = b->par[(size_t)inst];
    const lica_case_desc &cd = b->cases[(size_t)inst];
    InstState s;
    rc = get_state(b, inst, &s);
    if (rc) return rc;
    const int n = p.n, n3 = p.n3, r = p.r;
    std::vector<double> W((size_t)n3 + 1), P((size_t)n3 + 1), T((size_t)n3 + 1);
    auto nsget = [&](int fld, std::vector<double> &v) {
        return d2h(v.data(), b->d_ns + p.ns_off + (size_t)fld * p.ln + 2, sizeof(double) * v.size(), b->stream);
    };
    rc = nsget(NS_W, W);
    rc = rc ? rc : nsget(NS_P, P);
    rc = rc ? rc : nsget(NS_T, T);
    if (rc) return rc;
    FILE *f = fopen(path, "w");
    if (!f) return LICA_IO;
    const int N1 = 2, N2 = 2, N3 = n3;
    fprintf(f, "%8d%8d%8d\n", N1, N2, N3);
    fprintf(f, "%8lld%8lld%20.12E%20.12E\n", (long long)ihist, (long long)s.m, s.time, s.dt);
    fprintf(f, "%8d%8d%8d%8d%20.12E\n", 1, 1, 1, 3, cd.prm);
    fld_reals(f, {0.0}, 1);                                   // QVOL_ORI
    std::vector<double> v;
    v.assign((size_t)N1 * (N2 + 1) * (N3 + 1), 0.0);          // U(1:N1,0:N2,0:N3)
    fld_reals(f, v, 5);
    v.assign((size_t)(N1 + 1) * N2 * (N3 + 1), 0.0);          // V(0:N1,1:N2,0:N3)
    fld_reals(f, v, 5);
    v.clear();                                                // W(0:N1,0:N2,1:N3)
    for (int k = 1; k <= N3; k++)
        for (int j = 0; j <= N2; j++)
            for (int i = 0; i <= N1; i++) v.push_back(W[(size_t)k]);
    fld_reals(f, v, 5);
    v.clear();                                                // P(1:N1M,1:N2M,1:N3M)
    for (int k = 1; k <= n; k++) v.push_back(P[(size_t)k]);
    fld_reals(f, v, 5);
    fprintf(f, "%8d%8d%8d\n", r + 1, r + 1, p.n3f);
    fld_reals(f, {cd.denr, cd.visr, cd.fr, cd.surf_j, 0.0, cd.re}, 6);
    fprintf(f, "%8d%8d\n", 1, LICA_MLAYER);
    fld_reals(f, {s.vol_tot_ori}, 8);
    std::vector<long long> numa;
    for (int nl = 1; nl <= LICA_MLAYER; nl++) numa.push_back((long long)s.bn[nl] * r * r);
    fld_ints(f, numa, 8);
    for (int nl = 1; nl <= LICA_MLAYER; nl++)
        for (int q = 0; q < s.bn[nl]; q++)
            for (int j = 1; j <= r; j++)
                for (int i = 1; i <= r; i++) fprintf(f, "%8d%8d%8d%20.12E\n", i, j, s.bk[nl][q], s.alphi[nl][q]);
    fld_ints(f, {0}, 4);                                      // CONTACT
    fld_ints(f, {0}, 4);                                      // CONTACT_OLD
    fld_reals(f, {0.0}, 5);                                   // CONTACT_START
    fld_reals(f, {0.0}, 5);                                   // CONTACT_TIME
    fld_ints(f, {1}, 4);                                      // LVSON
    fld_reals(f, {0.0}, 5);                                   // T_CON
    fld_reals(f, {cd.prm, cd.scr, cd.tcr}, 3);
    v.clear();                                                // T(0:N1,0:N2,0:N3)
    for (int k = 0; k <= N3; k++)
        for (int j = 0; j <= N2; j++)
            for (int i = 0; i <= N1; i++) v.push_back(T[(size_t)k]);
    fld_reals(f, v, 5);
    return (fclose(f) == 0) ? LICA_OK : LICA_IO;
}

// ---------------------------------------------------------------------
// Restart of one case from a field file, as PREFLD reads it
// (1d.f:972-1097) and LVSINIT rebuilds the level set from the band read
// (ITRACKING=1, two_phase.f:146-189), followed by GRID_COUPLING(1) as MAIN
// calls it (1d.f:117-143).  Each READ of PREFLD starts a new line; a list
// longer than its format continues on the next line (format reversion):
// 5ES20.12 lists hold five values a line, 8I8 eight, and so on.

namespace {
struct FldIn {
    FILE *f = nullptr;
    std::string line;
    bool ok = true;
    bool next()
    {
        line.clear();
        int ch;
        bool any = false;
        while ((ch = fgetc(f)) != EOF) {
            any = true;
            if (ch == '\n') break;
            if (ch != '\r') line.push_back((char)ch);
        }
        if (!any) ok = false;
        return ok;
    }
    // the text of field i of width w on the current line (blank beyond its end)
    std::string field(size_t at, int w) const
    {
        if (at >= line.size()) return std::string();
        return line.substr(at, (size_t)w);
    }
    static bool real(const std::string &t0, double *v)
    {
        std::string t;
        for (char c : t0)
            if (c != ' ') t.push_back(c);
        if (t.empty()) { *v = 0.0; return true; }   // a blank field reads as zero
        // Fortran drops the E of a three-digit exponent (1.0-100)
        if (t.find_first_of("EeDd") == std::string::npos) {
            size_t p = t.find_last_of("+-");
            if (p != std::string::npos && p > 0) t.insert(p, "E");
        }
        for (char &c : t)
            if (c == 'D' || c == 'd') c = 'E';
        char *end = nullptr;
        *v = strtod(t.c_str(), &end);
        return end && *end == '\0' && std::isfinite(*v);
    }
    static bool integer(const std::string &t0, long long *v)
    {
        std::string t;
        for (char c : t0)
            if (c != ' ') t.push_back(c);
        if (t.empty()) { *v = 0; return true; }
        char *end = nullptr;
        *v = strtoll(t.c_str(), &end, 10);
        return end && *end == '\0';
    }
    // one READ of a list of n reals, `per` fields of width w a line
    bool reals(std::vector<double> &out, size_t n, int per, int w = 20)
    {
        out.assign(n, 0.0);
        size_t i = 0;
        do {
            if (!next()) return false;
            for (int j = 0; j < per && i < n; j++, i++)
                if (!real(field((size_t)j * w, w), &out[i])) return ok = false;
        } while (i < n);
        return true;
    }
    bool ints(std::vector<long long> &out, size_t n, int per, int w = 8)
    {
        out.assign(n, 0);
        size_t i = 0;
        do {
            if (!next()) return false;
            for (int j = 0; j < per && i < n; j++, i++)
                if (!integer(field((size_t)j * w, w), &out[i])) return ok = false;
        } while (i < n);
        return true;
    }
    // one line of ni I8 fields then nr ES20.12 fields
    bool mixed(int ni, long long *iv, int nr, double *rv)
    {
        if (!next()) return false;
        for (int j = 0; j < ni; j++)
            if (!integer(field((size_t)j * 8, 8), &iv[j])) return ok = false;
        for (int j = 0; j < nr; j++)
            if (!real(field((size_t)ni * 8 + (size_t)j * 20, 20), &rv[j])) return ok = false;
        return true;
    }
};

// a (0:N1,0:N2) plane list of the 1-D column must hold one value per k
bool column_of(const std::vector<double> &v, size_t per_k, size_t nk, std::vector<double> &col)
{
    col.assign(nk, 0.0);
    for (size_t k = 0; k < nk; k++) {
        const double *p = v.data() + k * per_k;
        for (size_t j = 1; j < per_k; j++)
            if (p[j] != p[0]) return false;
        col[k] = p[0];
    }
    return true;
}
}  // namespace

extern "C" int lica_batch_read_fld(lica_batch *b, int64_t inst, const char *path, int64_t flags)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (inst < 0 || inst >= b->ninst || !path || !b->initialized || (flags & ~7LL)) return LICA_BAD_ARG;
    const InstParam &p = b->par[(size_t)inst];
    const int n = p.n, n3 = p.n3, r = p.r;
    FldIn in;
    in.f = fopen(path, "r");
    if (!in.f) return LICA_IO;
    long long iv[5];
    double rv[6];
    std::vector<double> U, V, W, Pf, T, vol, misc;
    std::vector<long long> numa, contact;
    bool ok = in.mixed(3, iv, 0, rv);
    const long long N1 = iv[0], N2 = iv[1], N3 = iv[2];
    // the 1-D column: one interior cell in x and y (N1M = N2M = 1), the case's N3
    ok = ok && N1 == 2 && N2 == 2 && N3 == n3;
    long long ihist_m[2];
    double time_dt[2];
    ok = ok && in.mixed(2, ihist_m, 2, time_dt);
    long long ips[4];
    double pra;
    ok = ok && in.mixed(4, ips, 1, &pra);
    const long long ipss = ips[0];
    ok = ok && (ipss == 0 || ipss == 1);
    ok = ok && in.reals(misc, 1, 1);                                              // QVOL_ORI
    ok = ok && in.reals(U, (size_t)(N1 * (N2 + 1) * (N3 + 1)), 5);
    ok = ok && in.reals(V, (size_t)((N1 + 1) * N2 * (N3 + 1)), 5);
    ok = ok && in.reals(W, (size_t)((N1 + 1) * (N2 + 1) * N3), 5);
    ok = ok && in.reals(Pf, (size_t)((N1 - 1) * (N2 - 1) * (N3 - 1)), 5);
    for (double x : U) ok = ok && x == 0.0;
    for (double x : V) ok = ok && x == 0.0;
    long long nf[3];
    ok = ok && in.mixed(3, nf, 0, rv) && nf[2] == p.n3f;
    ok = ok && in.reals(misc, 6, 6);                                              // DENR .. RE_AIR
    long long nl_nmax[2];
    ok = ok && in.mixed(2, nl_nmax, 0, rv) && nl_nmax[0] == 1 && nl_nmax[1] == LICA_MLAYER;
    ok = ok && in.reals(vol, 1, 8);
    ok = ok && in.ints(numa, LICA_MLAYER, 8);
    // the band: every plane as r x r cells (I, J = 1..r) of one value
    InstState s;
    memset((void *)&s, 0, sizeof(s));
    for (int nl = 1; nl <= LICA_MLAYER && ok; nl++) {
        const long long cells = numa[(size_t)nl - 1];
        if (cells < 0 || cells % ((long long)r * r) != 0 || cells / ((long long)r * r) > LICA_MAXP) { ok = false; break; }
        s.bn[nl] = (int)(cells / ((long long)r * r));
        for (int q = 0; q < s.bn[nl] && ok; q++)
            for (int c = 0; c < r * r && ok; c++) {
                long long e[3];
                double a;
                ok = in.mixed(3, e, 1, &a);
                if (!ok) break;
                if (c == 0) {
                    s.bk[nl][q] = (int)e[2];
                    s.alphi[nl][q] = a;
                    ok = e[2] >= 1 && e[2] <= p.n3fm;
                }
                ok = ok && e[0] == 1 + c % r && e[1] == 1 + c / r && e[2] == s.bk[nl][q] && a == s.alphi[nl][q];
            }
    }
    ok = ok && in.ints(contact, 1, 4) && in.ints(contact, 1, 4);                  // CONTACT, CONTACT_OLD
    ok = ok && in.reals(misc, 1, 5) && in.reals(misc, 1, 5);                      // CONTACT_START, _TIME
    ok = ok && in.ints(contact, 1, 4) && contact[0] == 1;                          // LVSON
    ok = ok && in.reals(misc, 1, 5);                                               // T_CON
    std::vector<double> tcol((size_t)n3 + 1, 0.0);
    if (ok && ipss == 1) {
        ok = in.reals(misc, 3, 3) && in.reals(T, (size_t)((N1 + 1) * (N2 + 1) * (N3 + 1)), 5);
        ok = ok && column_of(T, (size_t)((N1 + 1) * (N2 + 1)), (size_t)N3 + 1, tcol);
    }
    fclose(in.f);
    std::vector<double> wcol, pcol;
    ok = ok && column_of(W, (size_t)((N1 + 1) * (N2 + 1)), (size_t)N3, wcol);     // W(k), k = 1..N3
    ok = ok && column_of(Pf, 1, (size_t)n, pcol);                                 // P(k), k = 1..N3M
    if (!ok) return LICA_IO;
    // fresh geometry, then the fields read (PREFLD), IPZERO / IRESET / IDTOLD
    // as MAIN applies them (1d.f:1070, 183-186, 1078-1083)
    std::vector<double> ns = b->img_ns[(size_t)inst];
    auto at = [&](int fld, int k) -> double & { return ns[(size_t)fld * p.ln + (size_t)(k + 2)]; };
    for (int k = 1; k <= n3; k++) at(NS_W, k) = wcol[(size_t)k - 1];
    for (int k = 1; k <= n; k++) at(NS_P, k) = (flags & 2) ? 0.0 : pcol[(size_t)k - 1];
    for (int k = 0; k <= n3; k++) at(NS_T, k) = tcol[(size_t)k];
    s.dt = (flags & 4) ? time_dt[1] : p.dt0;
    s.time = (flags & 1) ? 0.0 : time_dt[0];
    s.m = 0;
    s.vol_tot_ori = vol[0];
    s.vol_tot_iphs = vol[0];
    s.sf_lo = 1; s.sf_hi = 0; s.im_lo = 1; s.im_hi = 0;
    rc = h2d(b->d_ns + p.ns_off, ns.data(), sizeof(double) * ns.size(), b->stream);
    rc = rc ? rc : h2d(b->d_rf + p.rf_off, b->img_rf[(size_t)inst].data(), sizeof(double) * b->img_rf[(size_t)inst].size(), b->stream);
    rc = rc ? rc : h2d(b->d_nsi + p.nsi_off, b->img_nsi[(size_t)inst].data(), sizeof(int) * b->img_nsi[(size_t)inst].size(), b->stream);
    rc = rc ? rc : h2d(b->d_rfi + p.rfi_off, b->img_rfi[(size_t)inst].data(), sizeof(int) * b->img_rfi[(size_t)inst].size(), b->stream);
    rc = rc ? rc : h2d(b->d_st + inst, &s, sizeof(InstState), b->stream);
    if (rc) return rc;
    // the case's own workgroup: its place in the largest-first order
    size_t pos = 0;
    while (pos < b->order_h.size() && b->order_h[pos] != (int)inst) pos++;
    BatchDev bd = make_bd(b, 1);
    bd.order = b->d_order + pos;
    bd.ninst = 1;
    if (lica_launch_init_fld(bd, (void *)b->stream)) return LICA_HIP_ERROR;
    rc = dsync(b->stream);
    if (rc) return rc;
    rc = get_state(b, inst, &s);
    return rc ? rc : s.status;
}
