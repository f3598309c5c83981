// Spectral algebra on the dense (L+1)^2 layout, spectral invariants with
// compensated sums, and the velocity / gradient pointwise stage.
#include "plan.hpp"
#include "stage.hpp"

namespace aeo {

// One thread per (field, l, m).  Formulas: num/spectral_operators.py:213-334
// and phy/barotropic.py:68-93.
__global__ void spec_op_kernel(int op, int L, double R, const double2* __restrict__ in,
                               double2* __restrict__ out, int nf) {
    const int n = L + 1;
    const size_t fs = (size_t)n * n;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= fs * nf) return;
    const int f = (int)(t / fs);
    const int r = (int)(t - (size_t)f * fs);
    const int l = r / n, m = r - l * n;
    const double2* a = in + (size_t)f * fs;
    const double lap = -(double)l * (l + 1.0) / (R * R);
    double2 v;
    switch (op) {
        case AEO_OP_LAP: v = cscale(a[r], lap); break;
        case AEO_OP_INVLAP: {
            const double d = (l == 0) ? 1.0 / (R * R) : lap;
            v = c2(a[r].x / d, a[r].y / d);
            break;
        }
        case AEO_OP_INVLAP_ZERO: {
            v = (l == 0) ? c2(0.0, 0.0) : c2(a[r].x / lap, a[r].y / lap);
            break;
        }
        case AEO_OP_DLAMBDA: v = dlam(a[r], m, R); break;
        case AEO_OP_SINTH_DTH: {
            // g_l = C+_{l-1} a_{l-1} + C-_{l+1} a_{l+1}
            v = c2(0.0, 0.0);
            if (l >= 1) v = cadd(v, cscale(a[r - n], c_plus(l - 1, m, L)));
            if (l <= L - 1) v = cadd(v, cscale(a[r + n], c_minus(l + 1, m)));
            break;
        }
        case AEO_OP_SINTH_DTH_T: {
            // (S^T b)_l = C+_l b_{l+1} + C-_l b_{l-1}
            v = c2(0.0, 0.0);
            if (l <= L - 1) v = cadd(v, cscale(a[r + n], c_plus(l, m, L)));
            if (l >= 1) v = cadd(v, cscale(a[r - n], c_minus(l, m)));
            break;
        }
        default: {  // AEO_OP_TRUNC23
            const int cut = (2 * L) / 3;
            v = (l > cut || m > cut) ? c2(0.0, 0.0) : a[r];
            break;
        }
    }
    out[t] = v;
}

int launch_spec_op(int op, int L, double R, const double2* in, double2* out, int nf,
                   hipStream_t s) {
    const size_t total = (size_t)(L + 1) * (L + 1) * nf;
    if (total == 0) return AEO_OK;
    hipLaunchKernelGGL(spec_op_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, op,
                       L, R, in, out, nf);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

// ---------------------------------------------------------------------------
// Compensated (Neumaier / TwoSum) accumulation.
// ---------------------------------------------------------------------------
struct CSum {
    double s, c;
};

__device__ inline CSum cs_add(CSum a, double x) {
    const double t = a.s + x;
    const double bp = t - a.s;
    const double err = (a.s - (t - bp)) + (x - bp);
    return {t, a.c + err};
}

__device__ inline CSum cs_merge(CSum a, CSum b) {
    CSum r = cs_add(a, b.s);
    r.c += b.c;
    return r;
}

// One block per field: circulation, energy, relative and absolute enstrophy
// (run/bve/diagnostics.py:52-97).
__global__ void spectral_diag_kernel(int L, double R, double omega,
                                     const double2* __restrict__ z, double* __restrict__ out) {
    const int n = L + 1;
    const double2* a = z + (size_t)blockIdx.x * n * n;
    const double f10 = 2.0 * omega * sqrt(4.0 * M_PI / 3.0);
    CSum e = {0, 0}, zr = {0, 0}, za = {0, 0};
    for (int r = threadIdx.x; r < n * n; r += blockDim.x) {
        const int l = r / n, m = r - l * n;
        if (m > l) continue;
        const double2 v = a[r];
        const double p = (m == 0) ? v.x * v.x : 2.0 * (v.x * v.x + v.y * v.y);
        zr = cs_add(zr, p);
        if (l >= 1) e = cs_add(e, p / ((double)l * (l + 1.0)));
        double2 q = v;
        if (l == 1 && m == 0) q.x += f10;
        const double pq = (m == 0) ? q.x * q.x : 2.0 * (q.x * q.x + q.y * q.y);
        za = cs_add(za, pq);
    }
    __shared__ CSum se[256], sz[256], sa[256];
    se[threadIdx.x] = e;
    sz[threadIdx.x] = zr;
    sa[threadIdx.x] = za;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            se[threadIdx.x] = cs_merge(se[threadIdx.x], se[threadIdx.x + w]);
            sz[threadIdx.x] = cs_merge(sz[threadIdx.x], sz[threadIdx.x + w]);
            sa[threadIdx.x] = cs_merge(sa[threadIdx.x], sa[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* o = out + 4 * blockIdx.x;
        o[0] = R * R * sqrt(4.0 * M_PI) * a[0].x;
        o[1] = 0.5 * R * R * R * R * (se[0].s + se[0].c);
        o[2] = 0.5 * R * R * (sz[0].s + sz[0].c);
        o[3] = 0.5 * R * R * (sa[0].s + sa[0].c);
    }
}

int launch_spectral_diag(int L, double R, double omega, const double2* z, int nf, double* out,
                         hipStream_t s) {
    hipLaunchKernelGGL(spectral_diag_kernel, dim3(nf), dim3(256), 0, s, L, R, omega, z, out);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

// Per-step grid statistics of one sampling, one workgroup (fixed reduction
// order, compensated sums; run/bve/diagnostics.py:162-211,
// run/swe/diagnostics.py:113-173, phy/shallow_water.py:456-484).
//   AEO_GRID_STATS_BVE: a = zeta, b = u, c = v
//     out = {max |zeta|, sum w zeta^2, sum w, max sqrt(u^2 + v^2)}
//   AEO_GRID_STATS_SWE: a = u, b = v, c = phi (anomaly), d = phi_s or null;
//     k = {Phi0, mean phi_s, gravity}; Phi = Phi0 + phi
//     out = {max sqrt(u^2 + v^2), min Phi, max Phi,
//            sum w [Phi (u^2 + v^2)/2 + Phi^2/2 + Phi phi_s],
//            min eta, max eta},  eta = (Phi - Phi0 + phi_s - mean phi_s) / g
//   AEO_GRID_STATS_MINMAX: a = field; out = {min a, max a}
constexpr int kStatsThreads = 512;

__global__ __launch_bounds__(kStatsThreads) void grid_stats_kernel(
    int mode, int npts, const double* __restrict__ w, const double* __restrict__ a,
    const double* __restrict__ b, const double* __restrict__ c, const double* __restrict__ d,
    double k0, double k1, double k2, double* __restrict__ out) {
    // per thread: 4 extrema (e0 max, e1 min, e2 max, e3 min/max by mode) and 2 sums
    double x0 = 0.0, x1 = INFINITY, x2 = -INFINITY, x3 = INFINITY, x4 = -INFINITY;
    CSum s0 = {0, 0}, s1 = {0, 0};
    for (int i = threadIdx.x; i < npts; i += kStatsThreads) {
        if (mode == AEO_GRID_STATS_BVE) {
            const double z = a[i], wi = w[i];
            x0 = fmax(x0, fabs(z));
            x4 = fmax(x4, b[i] * b[i] + c[i] * c[i]);
            s0 = cs_add(s0, wi * z * z);
            s1 = cs_add(s1, wi);
        } else if (mode == AEO_GRID_STATS_SWE) {
            const double u = a[i], v = b[i], ph = k0 + c[i], sp = d ? d[i] : 0.0;
            const double q = u * u + v * v;
            x0 = fmax(x0, q);
            x1 = fmin(x1, ph);
            x2 = fmax(x2, ph);
            double e = 0.5 * ph * q + 0.5 * ph * ph;
            if (d) e += ph * sp;
            s0 = cs_add(s0, w[i] * e);
            const double eta = d ? (ph - k0 + (sp - k1)) / k2 : (ph - k0) / k2;
            x3 = fmin(x3, eta);
            x4 = fmax(x4, eta);
        } else {
            x1 = fmin(x1, a[i]);
            x2 = fmax(x2, a[i]);
        }
    }
    __shared__ double r0[kStatsThreads], r1[kStatsThreads], r2[kStatsThreads], r3[kStatsThreads],
        r4[kStatsThreads];
    __shared__ CSum q0[kStatsThreads], q1[kStatsThreads];
    const int t = threadIdx.x;
    r0[t] = x0; r1[t] = x1; r2[t] = x2; r3[t] = x3; r4[t] = x4;
    q0[t] = s0; q1[t] = s1;
    __syncthreads();
    for (int h = kStatsThreads / 2; h > 0; h >>= 1) {
        if (t < h) {
            r0[t] = fmax(r0[t], r0[t + h]);
            r1[t] = fmin(r1[t], r1[t + h]);
            r2[t] = fmax(r2[t], r2[t + h]);
            r3[t] = fmin(r3[t], r3[t + h]);
            r4[t] = fmax(r4[t], r4[t + h]);
            q0[t] = cs_merge(q0[t], q0[t + h]);
            q1[t] = cs_merge(q1[t], q1[t + h]);
        }
        __syncthreads();
    }
    if (t == 0) {
        if (mode == AEO_GRID_STATS_BVE) {
            out[0] = r0[0];
            out[1] = q0[0].s + q0[0].c;
            out[2] = q1[0].s + q1[0].c;
            out[3] = sqrt(r4[0]);
        } else if (mode == AEO_GRID_STATS_SWE) {
            out[0] = sqrt(r0[0]);
            out[1] = r1[0];
            out[2] = r2[0];
            out[3] = q0[0].s + q0[0].c;
            out[4] = r3[0];
            out[5] = r4[0];
        } else {
            out[0] = r1[0];
            out[1] = r2[0];
        }
    }
}

int launch_grid_stats(int mode, int npts, const double* w, const double* a, const double* b,
                      const double* c, const double* d, double k0, double k1, double k2,
                      double* out, hipStream_t s) {
    hipLaunchKernelGGL(grid_stats_kernel, dim3(1), dim3(kStatsThreads), 0, s, mode, npts, w, a,
                       b, c, d, k0, k1, k2, out);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

// mode 0: velocity (num/spectral_operators.py:517-538):
//   u = g_s / (R c), v = g_lam / c, with |c| < 1e-6 -> 0
// mode 1: gradient (num/spectral_operators.py:542-558):
//   dq/dx = g_lam / c, dq/dy = -g_s / (R c)
// coslat holds one value per point (row_len = 0) or per row of row_len points.
__global__ void velocity_kernel(int npts, int row_len, double R,
                                const double* __restrict__ coslat,
                                const double* __restrict__ g_s, const double* __restrict__ g_lam,
                                double* __restrict__ u, double* __restrict__ v, int mode) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    const double c = coslat[row_len > 0 ? i / row_len : i];
    if (mode == 0) {
        if (fabs(c) < 1e-6) {
            u[i] = 0.0;
            v[i] = 0.0;
        } else {
            u[i] = g_s[i] / (R * c);
            v[i] = g_lam[i] / c;
        }
    } else {
        u[i] = g_lam[i] / c;
        v[i] = -g_s[i] / (R * c);
    }
}

int launch_velocity_points(int n, int row_len, double R, const double* coslat,
                           const double* g_s, const double* g_lam, double* u, double* v, int mode,
                           hipStream_t s) {
    hipLaunchKernelGGL(velocity_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, row_len, R,
                       coslat, g_s, g_lam, u, v, mode);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

// out[k][j][i] = in[k][i][j] for nmat square (n x n) complex matrices.
__global__ void transpose_kernel(const double2* __restrict__ in, double2* __restrict__ out,
                                 int n) {
    __shared__ double2 tile[32][33];
    const size_t base = (size_t)blockIdx.z * n * n;
    const int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        const int i = i0 + r, j = j0 + threadIdx.x;
        if (i < n && j < n) tile[r][threadIdx.x] = in[base + (size_t)i * n + j];
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += blockDim.y) {
        const int j = j0 + r, i = i0 + threadIdx.x;
        if (i < n && j < n) out[base + (size_t)j * n + i] = tile[threadIdx.x][r];
    }
}

int launch_transpose(const double2* in, double2* out, int n, int nmat, hipStream_t s) {
    dim3 grid((n + 31) / 32, (n + 31) / 32, nmat);
    hipLaunchKernelGGL(transpose_kernel, grid, dim3(32, 8), 0, s, in, out, n);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

}  // namespace aeo

namespace aeo {

__global__ void scale_kernel(const double2* __restrict__ in, double2* __restrict__ out,
                             double alpha, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = cscale(in[i], alpha);
}

int launch_scale(const double2* in, double2* out, double alpha, size_t n, hipStream_t s) {
    if (n == 0) return AEO_OK;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(scale_kernel, dim3((unsigned)blocks), dim3(256), 0, s, in, out, alpha, n);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

}  // namespace aeo
