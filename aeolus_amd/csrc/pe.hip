// Dry hydrostatic primitive-equation tendency on the sigma grid
// (phy/primitive_equations.py:482-753) and the weak-form vector
// curl/divergence analysis it uses (num/spectral_operators.py:399-502).
//
// One tendency call is six device steps on the caller's stream:
//   1. pe_prep_kernel     every coefficient field the product grid needs, per
//                         level: U, V (the wind numerators), zeta, delta, T,
//                         T_lambda, T_s, eta, eta_lambda, eta_s, plus the two
//                         ln p_s derivatives (one thread per (l, m, level));
//   2. synthesis          of those 10K + 2 fields on the product plan;
//   3. pe_column_kernel   one thread per product point walks its column:
//                         continuity (G, d ln p_s/dt, sigma_dot), omega/p,
//                         centred vertical advection, the thermodynamic and
//                         momentum products and the half-metric vector
//                         components, all in registers;
//   4. analysis           of 4K + 1 product fields on the product plan;
//   5. analysis           of the 2K half-metric components on the degree-(L+1)
//                         plan over the same points;
//   6. pe_assemble_kernel one thread per (l, m, level): weak-form curl/div of the
//                         residual vector, 2/3 truncations, the spectral
//                         hydrostatic geopotential and -lap(Phi + E), the
//                         pinned l = 0 momentum rows.
#include "plan.hpp"
#include "stage.hpp"

#include <cmath>
#include <mutex>
#include <vector>

namespace aeo {

// Per-level fields synthesised on the product grid (index within a level).
enum PeSyn { kSynU = 0, kSynV, kSynZeta, kSynDelta, kSynT, kSynTLam, kSynTS, kSynEta,
             kSynEtaLam, kSynEtaS, kSynPerLevel };

// Sigma constants, [7K + 1] doubles:
//   [0K) Dsigma_k  [1K) alpha_k  [2K) beta_k / Dsigma_k (k >= 1)  [3K) 1 / (2 Dsigma_k)
//   [4K) R_d beta_k (k >= 1)  [5K) R_d alpha_k  [6K, 7K] sigma_{k-1/2}, k = 0..K
constexpr int kSigConsts = 7;

// One wave per workgroup for the column kernel: tens of thousands of columns
// spread over every CU.
constexpr int kColThreads = 64;

__device__ inline double point_coslat(const double* c, int row_len, int i) {
    return fmax(row_len > 0 ? c[i / row_len] : c[i], 1e-8);
}

__device__ inline double2 cdivr(double2 a, double s) { return c2(a.x / s, a.y / s); }

// (S^T_ext c)_{l,m} = C+_{l,m} c_{l+1,m} + C-_{l,m} c_{l-1,m} with the
// unclipped C+ at l = L reading the extended degree; `c` has row stride L + 2.
__device__ inline double2 adj_ext(const double2* c, int l, int m, int L) {
    const int s = L + 2;
    double2 v = cscale(c[(l + 1) * s + m], c_plus(l, m, L + 1));
    if (l >= 1) v = cadd(v, cscale(c[(l - 1) * s + m], c_minus(l, m)));
    return v;
}

// curl and div of the weak form at (l, m) from the extended analyses a
// (east) and b (north): div = (i m/R) a + S^T b / R, curl = (i m/R) b - S^T a / R.
__device__ inline void weak_curl_div(const double2* a, const double2* b, int l, int m, int L,
                                     double R, double2* curl, double2* div) {
    const int s = L + 2;
    *div = cadd(dlam(a[l * s + m], m, R), cdivr(adj_ext(b, l, m, L), R));
    *curl = csub(dlam(b[l * s + m], m, R), cdivr(adj_ext(a, l, m, L), R));
}

__device__ inline double2 trunc23(double2 v, int l, int m, int cut) {
    return (l > cut || m > cut) ? c2(0.0, 0.0) : v;
}

// ---------------------------------------------------------------------------
// 1. Coefficient fields for the product grid
// ---------------------------------------------------------------------------
__global__ void pe_prep_kernel(int L, int K, double R, double f10, const double2* __restrict__ y,
                               double2* __restrict__ prep) {
    const int n = L + 1, fs = n * n;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;            // level, or K for the ln p_s fields
    if (r >= fs) return;
    const int l = r / n, m = r - l * n;
    const int nf = (k < K) ? kSynPerLevel : 2;
    double2* o = prep + (size_t)kSynPerLevel * k * fs + r;
    if (m > l) {
        for (int f = 0; f < nf; ++f) o[(size_t)f * fs] = c2(0.0, 0.0);
        return;
    }
    const bool lo = l >= 1, hi = l <= L - 1;
    const double cpl = lo ? c_plus(l - 1, m, L) : 0.0;
    const double cmh = hi ? c_minus(l + 1, m) : 0.0;
    auto lap = [&](int ll) { return -(double)ll * (ll + 1.0) / (R * R); };
    // psi = lap^-1 a with l = 0 zeroed (phy/primitive_equations.py:266-270)
    auto ilap = [&](double2 a, int ll) {
        return ll == 0 ? c2(0.0, 0.0) : c2(a.x / lap(ll), a.y / lap(ll));
    };
    // (S a)_l = C+_{l-1} a_{l-1} + C-_{l+1} a_{l+1} (num/spectral_operators.py:228-241)
    auto sop = [&](double2 am, double2 ap) {
        double2 v = c2(0.0, 0.0);
        if (lo) v = cadd(v, cscale(am, cpl));
        if (hi) v = cadd(v, cscale(ap, cmh));
        return v;
    };
    const double2 zero = c2(0.0, 0.0);
    auto at = [&](int row, int ll) { return y[(size_t)row * fs + ll * n + m]; };
    if (k == K) {
        const double2 q = at(3 * K, l), qm = lo ? at(3 * K, l - 1) : zero,
                      qp = hi ? at(3 * K, l + 1) : zero;
        o[0] = dlam(q, m, R);
        o[fs] = cdivr(sop(qm, qp), R);
        return;
    }
    const double2 z = at(k, l), zm = lo ? at(k, l - 1) : zero, zp = hi ? at(k, l + 1) : zero;
    const double2 d = at(K + k, l), dm = lo ? at(K + k, l - 1) : zero,
                  dp = hi ? at(K + k, l + 1) : zero;
    const double2 t = at(2 * K + k, l), tm = lo ? at(2 * K + k, l - 1) : zero,
                  tp = hi ? at(2 * K + k, l + 1) : zero;
    const double2 ps = ilap(z, l), psm = lo ? ilap(zm, l - 1) : zero,
                  psp = hi ? ilap(zp, l + 1) : zero;
    const double2 ch = ilap(d, l), chm = lo ? ilap(dm, l - 1) : zero,
                  chp = hi ? ilap(dp, l + 1) : zero;
    // eta = zeta + f, f = 2 Omega sqrt(4 pi / 3) at (1, 0) only
    auto eta = [&](double2 a, int ll) {
        return (ll == 1 && m == 0) ? c2(a.x + f10, a.y + 0.0) : a;
    };
    o[kSynU * fs] = cadd(cdivr(sop(psm, psp), R), dlam(ch, m, R));
    o[kSynV * fs] = csub(dlam(ps, m, R), cdivr(sop(chm, chp), R));
    o[kSynZeta * fs] = z;
    o[kSynDelta * fs] = d;
    o[kSynT * fs] = t;
    o[kSynTLam * fs] = dlam(t, m, R);
    o[kSynTS * fs] = cdivr(sop(tm, tp), R);
    o[kSynEta * fs] = eta(z, l);
    o[kSynEtaLam * fs] = dlam(z, m, R);
    o[kSynEtaS * fs] = cdivr(sop(lo ? eta(zm, l - 1) : zero, hi ? eta(zp, l + 1) : zero), R);
}

// ---------------------------------------------------------------------------
// 3. Column physics at one product point (phy/primitive_equations.py:527-725,
//    phy/sigma_coordinate.py:164-398)
// ---------------------------------------------------------------------------
__global__ void pe_column_kernel(int npts, int row_len, const double* __restrict__ coslat, int K,
                                 const double* __restrict__ sig, double kappa, double r_dry,
                                 const double* __restrict__ syn, double* __restrict__ an,
                                 double* __restrict__ ex) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    const size_t N = npts;
    // one division per column: every 1/cos(lat) factor below is a product
    const double ic = 1.0 / point_coslat(coslat, row_len, i);
    auto F = [&](int f) { return syn[(size_t)f * N + i]; };
    const double* th = sig;
    const double* al = sig + K;
    const double* cb = sig + 2 * K;
    const double* ith2 = sig + 3 * K;   // 1 / (2 Dsigma_k)
    const double* si = sig + 6 * K;
    const double ll = F(kSynPerLevel * K), ls = F(kSynPerLevel * K + 1);

    // column total of Dsigma_k G_k, G_k = delta_k + V_k . grad ln p_s
    double total = 0.0;
#pragma unroll 8
    for (int k = 0; k < K; ++k) {
        const int b = kSynPerLevel * k;
        const double u = F(b + kSynU) * ic, v = F(b + kSynV) * ic;
        const double G = F(b + kSynDelta) + (u * ll - v * ls) * ic;
        total = (k == 0) ? th[0] * G : total + th[k] * G;
    }

    double partial_prev = 0.0, sd_top = 0.0;
    double u0 = 0.0, v0 = 0.0, t0 = 0.0;
    double u1 = F(kSynU) * ic, v1 = F(kSynV) * ic, t1 = F(kSynT);
    // unrolled so that several levels' loads are in flight at once (a column
    // is one thread; there are few waves per CU to hide the latency)
#pragma unroll 4
    for (int k = 0; k < K; ++k) {
        const int b = kSynPerLevel * k;
        const double d = F(b + kSynDelta), z = F(b + kSynZeta);
        const double A = (u1 * ll - v1 * ls) * ic;
        const double G = d + A;
        const double partial = (k == 0) ? th[0] * G : partial_prev + th[k] * G;
        const double sd_bot = (k == K - 1) ? total - total : si[k + 1] * total - partial;
        // next level (clamped so the loads are unconditional; unused at the bottom)
        const int bn = kSynPerLevel * ((k + 1 < K) ? k + 1 : k);
        const double u2 = F(bn + kSynU) * ic, v2 = F(bn + kSynV) * ic, t2 = F(bn + kSynT);
        // centred (sigma_dot dX/dsigma)_k; boundary interfaces never read
        auto vadv = [&](double xm, double x0, double xp) {
            double acc = 0.0;
            bool have = false;
            if (k < K - 1) { acc = sd_bot * (xp - x0); have = true; }
            if (k > 0) {
                const double up = sd_top * (x0 - xm);
                acc = have ? acc + up : up;
                have = true;
            }
            return have ? acc * ith2[k] : 0.0;
        };
        const double vu = vadv(u0, u1, u2), vv = vadv(v0, v1, v2), vt = vadv(t0, t1, t2);
        const double wp = (k == 0) ? A - al[0] * G : (A - cb[k] * partial_prev) - al[k] * G;
        // thermodynamic: -V.grad T - V_adv(T) + kappa T omega/p
        const double advT = (u1 * F(b + kSynTLam) - v1 * F(b + kSynTS)) * ic;
        an[(size_t)k * N + i] = (-advT - vt) + (kappa * t1) * wp;
        // eta V: div = u.grad eta + eta delta, k.curl = eta zeta + (grad eta x V).k
        const double eg = F(b + kSynEta), el = F(b + kSynEtaLam), es = F(b + kSynEtaS);
        const double div_ev = (u1 * el - v1 * es) * ic + eg * d;
        const double curl_ev = eg * z + (v1 * el + u1 * es) * ic;
        an[(size_t)(K + k) * N + i] = -div_ev;
        an[(size_t)(2 * K + k) * N + i] = curl_ev;
        an[(size_t)(3 * K + k) * N + i] = 0.5 * (u1 * u1 + v1 * v1);
        // residual vector Z = sigma_dot dV/dsigma + R_d T grad ln p_s, half-metric
        const double rt = r_dry * t1;
        const double ze = vu + rt * (ll * ic);
        const double zn = vv + rt * (-ls * ic);
        ex[(size_t)k * N + i] = ze * ic;
        ex[(size_t)(K + k) * N + i] = zn * ic;
        u0 = u1; v0 = v1; t0 = t1;
        u1 = u2; v1 = v2; t1 = t2;
        sd_top = sd_bot;
        partial_prev = partial;
    }
    an[(size_t)4 * K * N + i] = -total;
}

// ---------------------------------------------------------------------------
// 6. Spectral assembly (phy/primitive_equations.py:596-725)
// ---------------------------------------------------------------------------
__global__ void pe_assemble_kernel(int L, int K, double R, const double* __restrict__ sig,
                                   const double2* __restrict__ y, const double2* __restrict__ phis,
                                   const double2* __restrict__ sp,
                                   const double2* __restrict__ spx, double2* __restrict__ out) {
    const int n = L + 1, fs = n * n, fx = (L + 2) * (L + 2);
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;            // level, or K for the ln p_s row
    if (r >= fs) return;
    const int l = r / n, m = r - l * n;
    const int cut = (2 * L) / 3;
    if (k == K) {
        out[(size_t)3 * K * fs + r] = trunc23(sp[(size_t)4 * K * fs + r], l, m, cut);
        return;
    }
    const double lap = -(double)l * (l + 1.0) / (R * R);
    const double* rb = sig + 4 * K;
    const double* ra = sig + 5 * K;
    const double2* T = y + (size_t)2 * K * fs + r;
    // spectral Simmons-Burridge geopotential: the interface recursion from the
    // surface up to level k, in the reference's order (phy/sigma_coordinate.py:164-209)
    double2 pb = phis[r];
    for (int j = K - 1; j > k; --j) pb = cadd(pb, cscale(T[(size_t)j * fs], rb[j]));
    const double2 phi = cadd(pb, cscale(T[(size_t)k * fs], ra[k]));
    double2 curl_z, div_z;
    weak_curl_div(spx + (size_t)k * fx, spx + (size_t)(K + k) * fx, l, m, L, R, &curl_z, &div_z);
    const double2 zd = trunc23(csub(sp[(size_t)(K + k) * fs + r], curl_z), l, m, cut);
    const double2 ke = trunc23(sp[(size_t)(3 * K + k) * fs + r], l, m, cut);
    const double2 dd = csub(trunc23(csub(sp[(size_t)(2 * K + k) * fs + r], div_z), l, m, cut),
                            cscale(cadd(ke, phi), lap));
    out[(size_t)k * fs + r] = l == 0 ? c2(0.0, 0.0) : zd;
    out[(size_t)(K + k) * fs + r] = l == 0 ? c2(0.0, 0.0) : dd;
    out[(size_t)(2 * K + k) * fs + r] = trunc23(sp[(size_t)k * fs + r], l, m, cut);
}

// ---------------------------------------------------------------------------
// Stand-alone weak-form vector analysis
// ---------------------------------------------------------------------------
__global__ void halfmetric_kernel(int npts, int row_len, const double* __restrict__ coslat,
                                  int nf, const double* __restrict__ fe,
                                  const double* __restrict__ fn, double* __restrict__ out) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t N = npts;
    if (t >= 2 * N * nf) return;
    const int f = (int)(t / N), i = (int)(t - (size_t)f * N);
    const double v = f < nf ? fe[t] : fn[t - (size_t)nf * N];
    out[t] = v / point_coslat(coslat, row_len, i);
}

__global__ void curl_div_kernel(int L, double R, int nf, int truncate,
                                const double2* __restrict__ spx, double2* __restrict__ curl,
                                double2* __restrict__ div) {
    const int n = L + 1, fs = n * n, fx = (L + 2) * (L + 2);
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)fs * nf) return;
    const int f = (int)(t / fs), r = (int)(t - (size_t)f * fs);
    const int l = r / n, m = r - l * n;
    double2 cv, dv;
    weak_curl_div(spx + (size_t)f * fx, spx + (size_t)(nf + f) * fx, l, m, L, R, &cv, &dv);
    if (truncate) {
        const int cut = (2 * L) / 3;
        cv = trunc23(cv, l, m, cut);
        dv = trunc23(dv, l, m, cut);
    }
    curl[t] = cv;
    div[t] = dv;
}

static int plan_points(const aeo_plan_s* p) {
    return p->kind == kGauss ? p->nlat * p->nlon : p->npts;
}

static const double* plan_coslat(const aeo_plan_s* p, int* row_len) {
    *row_len = p->kind == kGauss ? p->nlon : 0;
    return p->kind == kGauss ? p->d_coslat : p->d_pcoslat;
}

static unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

// Weak-form vector curl/div of nfields tangent fields with the caller's
// workspace `work` (aeo_vector_curl_div_work doubles): the half-metric
// components on the extended plan's points, their degree-(L+1) analyses, then
// curl/div.  Every entry of `work` is written before it is read.
static int vector_curl_div_run(aeo_plan_s* ext, double radius, const double* f_east,
                               const double* f_north, int nfields, int truncate,
                               double* curl_out, double* div_out, double* work, hipStream_t s) {
    AEO_REQUIRE(ext && f_east && f_north && curl_out && div_out && work, AEO_ERR_INVALID,
                "aeo_vector_curl_div: NULL argument");
    AEO_REQUIRE(nfields >= 1 && ext->L >= 1, AEO_ERR_INVALID,
                "aeo_vector_curl_div: needs nfields >= 1 and an extended plan of degree >= 1");
    AEO_REQUIRE(std::isfinite(radius) && radius > 0.0, AEO_ERR_INVALID,
                "aeo_vector_curl_div: radius must be finite and > 0");
    const int L = ext->L - 1;
    const size_t N = plan_points(ext);
    double* grid = work;
    double2* spx = reinterpret_cast<double2*>(work + 2 * N * nfields);
    int row_len = 0;
    const double* cl = plan_coslat(ext, &row_len);
    hipLaunchKernelGGL(halfmetric_kernel, dim3(blocks_for(2 * N * nfields)), dim3(256), 0, s,
                       (int)N, row_len, cl, nfields, f_east, f_north, grid);
    AEO_LAUNCH_CHECK();
    int rc = aeo_analysis(ext, grid, 2 * nfields, reinterpret_cast<double*>(spx), s);
    if (rc) return rc;
    const size_t fs = (size_t)(L + 1) * (L + 1);
    hipLaunchKernelGGL(curl_div_kernel, dim3(blocks_for(fs * nfields)), dim3(256), 0, s, L,
                       radius, nfields, truncate ? 1 : 0, spx,
                       reinterpret_cast<double2*>(curl_out), reinterpret_cast<double2*>(div_out));
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

// Workspace of aeo_vector_curl_div_noalloc: a fixed device array, one copy
// per GPU, resident from the library's load until the process ends (no
// allocation at call time).  In host code `scratch` names a host-side stand-in,
// not device memory: the kernels get the device address that
// hipGetSymbolAddress returns for the current device, looked up once per
// device on the entry point's first call there.
constexpr size_t kNoAllocScratchDoubles = (size_t)2 << 20;   // 16 MiB
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

constexpr int kScratchMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_dev[kScratchMaxDevices];   // device address per device id

static int scratch_on_current_device(double** out) {
    int dev = -1;
    AEO_CHECK_HIP(hipGetDevice(&dev));
    AEO_REQUIRE(dev >= 0 && dev < kScratchMaxDevices, AEO_ERR_UNSUPPORTED,
                "aeo_vector_curl_div_noalloc: device id out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        AEO_REQUIRE(e == hipSuccess && p != nullptr, AEO_ERR_HIP,
                    std::string("aeo_vector_curl_div_noalloc: hipGetSymbolAddress(scratch): ") +
                        hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_dev[dev];
    return AEO_OK;
}

}  // namespace aeo

using namespace aeo;

struct aeo_pe_s {
    aeo_plan_t prod = nullptr, ext = nullptr;
    int K = 0, L = 0, npts = 0;
    double R = 0.0, f10 = 0.0, r_dry = 0.0, kappa = 0.0;
    double* d_sig = nullptr;
    double2* d_phis = nullptr;
    double2* d_prep = nullptr;    // [10K + 2][(L+1)^2]
    double* d_syn = nullptr;      // [10K + 2][npts]
    double* d_an = nullptr;       // [4K + 1][npts]
    double* d_ex = nullptr;       // [2K][npts]
    double2* d_sp = nullptr;      // [4K + 1][(L+1)^2]
    double2* d_spx = nullptr;     // [2K][(L+2)^2]
};

namespace {

template <class T>
int pe_alloc(T** ptr, size_t count) {
    *ptr = nullptr;
    hipError_t e = hipMalloc((void**)ptr, count * sizeof(T));
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                  " bytes failed: " + hipGetErrorString(e));
        *ptr = nullptr;
        return AEO_ERR_HIP;
    }
    return AEO_OK;
}

void pe_free(aeo_pe_s* p) {
    if (!p) return;
    for (void* q : {(void*)p->d_sig, (void*)p->d_phis, (void*)p->d_prep, (void*)p->d_syn,
                    (void*)p->d_an, (void*)p->d_ex, (void*)p->d_sp, (void*)p->d_spx})
        if (q) (void)hipFree(q);
    delete p;
}

int check_ext_pair(aeo_plan_t prod, aeo_plan_t ext, const char* who) {
    AEO_REQUIRE(ext->L == prod->L + 1 && ext->kind == prod->kind &&
                    plan_points(ext) == plan_points(prod) &&
                    (prod->kind != kGauss || (ext->nlat == prod->nlat && ext->nlon == prod->nlon)),
                AEO_ERR_INVALID,
                std::string(who) + ": the extended plan must have degree L+1 on the product "
                "plan's points");
    return AEO_OK;
}

}  // namespace

extern "C" {

int aeo_pe_create(aeo_plan_t prod, aeo_plan_t ext, int nlev, const double* interfaces,
                  double radius, double omega, double r_dry, double cp_dry,
                  const double* phi_surface_lm, aeo_pe_t* out) {
    AEO_REQUIRE(prod && ext && interfaces && out, AEO_ERR_INVALID, "aeo_pe_create: NULL argument");
    *out = nullptr;
    int rc = check_ext_pair(prod, ext, "aeo_pe_create");
    if (rc) return rc;
    AEO_REQUIRE(nlev >= 1, AEO_ERR_INVALID, "aeo_pe_create: nlev must be >= 1");
    AEO_REQUIRE(std::isfinite(radius) && radius > 0.0 && std::isfinite(omega), AEO_ERR_INVALID,
                "aeo_pe_create: radius must be finite and > 0, omega finite");
    AEO_REQUIRE(std::isfinite(r_dry) && r_dry > 0.0, AEO_ERR_INVALID,
                "aeo_pe_create: r_dry must be finite and > 0");
    AEO_REQUIRE(std::isfinite(cp_dry) && cp_dry > r_dry, AEO_ERR_INVALID,
                "aeo_pe_create: cp_dry must be finite and > r_dry");
    const int K = nlev;
    const double* s = interfaces;
    AEO_REQUIRE(s[0] == 0.0 && s[K] == 1.0, AEO_ERR_INVALID,
                "aeo_pe_create: sigma interfaces must start at exactly 0 and end at exactly 1");
    for (int k = 0; k < K; ++k)
        AEO_REQUIRE(std::isfinite(s[k + 1]) && s[k] < s[k + 1], AEO_ERR_INVALID,
                    "aeo_pe_create: sigma interfaces must be finite and strictly increasing");
    // metadata of phy/sigma_coordinate.py:100-139, evaluated as the reference does
    std::vector<double> sig((size_t)kSigConsts * K + 1, 0.0);
    for (int k = 0; k < K; ++k) {
        const double th = s[k + 1] - s[k];
        const double beta = k == 0 ? 0.0 : std::log(s[k + 1] / s[k]);
        const double alpha = k == 0 ? std::log(2.0) : 1.0 - (s[k] / th) * beta;
        sig[k] = th;
        sig[K + k] = alpha;
        sig[2 * K + k] = k == 0 ? 0.0 : beta / th;
        sig[3 * K + k] = 1.0 / (2.0 * th);
        sig[4 * K + k] = k == 0 ? 0.0 : r_dry * beta;
        sig[5 * K + k] = r_dry * alpha;
    }
    for (int k = 0; k <= K; ++k) sig[6 * K + k] = s[k];

    auto* p = new aeo_pe_s();
    p->prod = prod;
    p->ext = ext;
    p->K = K;
    p->L = prod->L;
    p->npts = plan_points(prod);
    p->R = radius;
    p->f10 = 2.0 * omega * std::sqrt(4.0 * M_PI / 3.0);
    p->r_dry = r_dry;
    p->kappa = r_dry / cp_dry;
    const size_t fs = (size_t)(p->L + 1) * (p->L + 1), fx = (size_t)(p->L + 2) * (p->L + 2);
    const size_t N = p->npts, nsyn = (size_t)kSynPerLevel * K + 2;
    do {
        if ((rc = pe_alloc(&p->d_sig, sig.size()))) break;
        if ((rc = pe_alloc(&p->d_phis, fs))) break;
        if ((rc = pe_alloc(&p->d_prep, nsyn * fs))) break;
        if ((rc = pe_alloc(&p->d_syn, nsyn * N))) break;
        if ((rc = pe_alloc(&p->d_an, (size_t)(4 * K + 1) * N))) break;
        if ((rc = pe_alloc(&p->d_ex, (size_t)2 * K * N))) break;
        if ((rc = pe_alloc(&p->d_sp, (size_t)(4 * K + 1) * fs))) break;
        if ((rc = pe_alloc(&p->d_spx, (size_t)2 * K * fx))) break;
        hipError_t e = hipMemcpy(p->d_sig, sig.data(), sig.size() * sizeof(double),
                                 hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = phi_surface_lm ? hipMemcpy(p->d_phis, phi_surface_lm, fs * sizeof(double2),
                                           hipMemcpyDeviceToDevice)
                               : hipMemset(p->d_phis, 0, fs * sizeof(double2));
        if (e != hipSuccess) {
            set_error(std::string("aeo_pe_create: ") + hipGetErrorString(e));
            rc = AEO_ERR_HIP;
        }
    } while (0);
    if (rc) {
        pe_free(p);
        return rc;
    }
    *out = p;
    return AEO_OK;
}

int aeo_pe_destroy(aeo_pe_t pe) {
    pe_free(pe);
    return AEO_OK;
}

int aeo_pe_tendency(aeo_pe_t pe, const double* y, double* out, void* stream) {
    AEO_REQUIRE(pe && y && out, AEO_ERR_INVALID, "aeo_pe_tendency: NULL argument");
    AEO_REQUIRE(y != out, AEO_ERR_INVALID, "aeo_pe_tendency: out must not alias y");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int L = pe->L, K = pe->K, n = L + 1;
    const int nsyn = kSynPerLevel * K + 2;
    hipLaunchKernelGGL(pe_prep_kernel, dim3(blocks_for((size_t)n * n), K + 1), dim3(256), 0, s, L, K,
                       pe->R, pe->f10, reinterpret_cast<const double2*>(y), pe->d_prep);
    AEO_LAUNCH_CHECK();
    int rc = aeo_synthesis(pe->prod, reinterpret_cast<const double*>(pe->d_prep), nsyn,
                           pe->d_syn, stream);
    if (rc) return rc;
    int row_len = 0;
    const double* cl = plan_coslat(pe->prod, &row_len);
    hipLaunchKernelGGL(pe_column_kernel, dim3((pe->npts + kColThreads - 1) / kColThreads),
                       dim3(kColThreads), 0, s, pe->npts,
                       row_len, cl, K, pe->d_sig, pe->kappa, pe->r_dry, pe->d_syn, pe->d_an,
                       pe->d_ex);
    AEO_LAUNCH_CHECK();
    if ((rc = aeo_analysis(pe->prod, pe->d_an, 4 * K + 1, reinterpret_cast<double*>(pe->d_sp),
                           stream)))
        return rc;
    if ((rc = aeo_analysis(pe->ext, pe->d_ex, 2 * K, reinterpret_cast<double*>(pe->d_spx),
                           stream)))
        return rc;
    hipLaunchKernelGGL(pe_assemble_kernel, dim3(blocks_for((size_t)n * n), K + 1), dim3(256), 0, s,
                       L, K,
                       pe->R, pe->d_sig, reinterpret_cast<const double2*>(y), pe->d_phis,
                       pe->d_sp, pe->d_spx, reinterpret_cast<double2*>(out));
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

int aeo_vector_curl_div_work(aeo_plan_t ext, int nfields, long long* n_doubles) {
    AEO_REQUIRE(ext && n_doubles && nfields >= 1, AEO_ERR_INVALID,
                "aeo_vector_curl_div_work: bad argument");
    const long long fx = (long long)(ext->L + 1) * (ext->L + 1);
    *n_doubles = 2LL * nfields * plan_points(ext) + 4LL * nfields * fx;
    return AEO_OK;
}

int aeo_vector_curl_div(aeo_plan_t ext, double radius, const double* f_east, const double* f_north,
                        int nfields, int truncate, double* curl_out, double* div_out,
                        double* work, void* stream) {
    AEO_REQUIRE(work, AEO_ERR_INVALID, "aeo_vector_curl_div: NULL argument");
    return vector_curl_div_run(ext, radius, f_east, f_north, nfields, truncate, curl_out, div_out,
                               work, reinterpret_cast<hipStream_t>(stream));
}

int aeo_vector_curl_div_noalloc(aeo_plan_t ext, double radius, const double* f_east,
                                const double* f_north, int nfields, int truncate,
                                double* curl_out, double* div_out, void* stream) {
    long long need = 0;
    int rc = aeo_vector_curl_div_work(ext, nfields, &need);
    if (rc) return rc;
    AEO_REQUIRE((unsigned long long)need <= kNoAllocScratchDoubles, AEO_ERR_UNSUPPORTED,
                "aeo_vector_curl_div_noalloc: needs " + std::to_string(need) +
                    " doubles of scratch, more than its fixed " +
                    std::to_string(kNoAllocScratchDoubles) +
                    " (16 MiB); use aeo_vector_curl_div with a caller workspace");
    double* work = nullptr;
    if ((rc = scratch_on_current_device(&work))) return rc;
    return vector_curl_div_run(ext, radius, f_east, f_north, nfields, truncate, curl_out, div_out,
                               work, reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
