// Gauss-Legendre grid back end: FP64 associated-Legendre table, per-m
// Legendre contractions (even/odd folded about the equator) and the LDS FFT
// grid kernels with the fused pointwise products.
//
// A transform on a Gauss grid is the same sum as the reference's dense
// point-set transform (num/fast_geodesic_sh.py:68-107) evaluated in a
// different order:
//   synthesis  f(j,k) = sum_m c_m Re( G_m(j) e^{i m lambda_k} ),
//              G_m(j) = sum_l Pbar_lm(x_j) a_lm     (c_0 = 1 with Re a_l0, c_m = 2)
//   analysis   a_lm   = sum_j w_j Pbar_lm(x_j) F_m(j),
//              F_m(j) = sum_k f(j,k) e^{-i m lambda_k}
// North/south rows share Pbar up to the parity (-1)^(l+m), so the Legendre
// kernels work on the northern half only: synthesis produces the even and
// odd partial sums E, O (G_north = E + O, G_south = E - O); analysis reads
// S = w (F_north + F_south) and D = w (F_north - F_south).  One complex FFT
// of length nlon transforms a north/south row pair at once
// (z = f_north + i f_south).
#include "plan.hpp"
#include "stage.hpp"

namespace aeo {

constexpr int kLegThreads = 256;  // 4 waves; wave w takes l - m = w (mod 4)
constexpr int kRowElems = 4096;   // complex values per LDS pass of the generic grid kernels

// ---------------------------------------------------------------------------
// FP64 normalised associated-Legendre recurrence (one thread per (m, jp)).
//   Pbar_00 = 1/sqrt(4 pi)
//   Pbar_mm = -sqrt((2m+1)/(2m)) s Pbar_{m-1,m-1}      (Condon-Shortley phase)
//   Pbar_{m+1,m} = sqrt(2m+3) x Pbar_mm
//   Pbar_lm = a_lm (x Pbar_{l-1,m} - b_lm Pbar_{l-2,m})
// ---------------------------------------------------------------------------
__global__ void legendre_table_kernel(int L, int nh, const double* __restrict__ x_north,
                                      const long long* __restrict__ offs,
                                      double* __restrict__ Ps, double* __restrict__ Pa) {
    const int m = blockIdx.x;
    const int jp = blockIdx.y * blockDim.x + threadIdx.x;
    if (m > L || jp >= nh) return;
    const double x = x_north[jp];
    const double s = sqrt((1.0 - x) * (1.0 + x));
    double pmm = 0.28209479177387814347;  // 1/sqrt(4 pi)
    for (int k = 1; k <= m; ++k) pmm *= -sqrt((2.0 * k + 1.0) / (2.0 * k)) * s;
    const int K = L - m + 1;
    double* ps = Ps + offs[m];
    double* pa = Pa + offs[m];
    double p2 = 0.0, p1 = pmm;
    ps[jp] = pmm;
    pa[(size_t)jp * K] = pmm;
    if (m + 1 <= L) {
        const double p = sqrt(2.0 * m + 3.0) * x * pmm;
        ps[(size_t)nh + jp] = p;
        pa[(size_t)jp * K + 1] = p;
        p2 = pmm;
        p1 = p;
    }
    for (int l = m + 2; l <= L; ++l) {
        const double ll = (double)l * l, mm = (double)m * m;
        const double a = sqrt((4.0 * ll - 1.0) / (ll - mm));
        const double b = sqrt(((l - 1.0) * (l - 1.0) - mm) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
        const double p = a * (x * p1 - b * p2);
        ps[(size_t)(l - m) * nh + jp] = p;
        pa[(size_t)jp * K + (l - m)] = p;
        p2 = p1;
        p1 = p;
    }
}

// ---------------------------------------------------------------------------
// Legendre synthesis: E/O[f][jp][m] = sum over even/odd (l-m) of Pbar * value.
// grid (L+1, ceil(nh/64), field groups); block 256 = 4 waves x 64 latitudes.
// ---------------------------------------------------------------------------
template <class Loader>
__global__ __launch_bounds__(kLegThreads) void legendre_syn_kernel(GaussDev g, Loader ld,
                                                                  double2* __restrict__ E,
                                                                  double2* __restrict__ O) {
    constexpr int NF = Loader::NF;
    extern __shared__ double2 lds[];
    const int m = blockIdx.x;
    const int L = g.L, nh = g.nh;
    const int K = L - m + 1;
    double2* col = lds;                 // [NF][K]; reused as [2][NF][64] after the sum
    for (int t = threadIdx.x; t < NF * K; t += kLegThreads) {
        const int f = t / K, i = t - f * K;
        col[t] = ld.value(f, m + i, m);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int jp = blockIdx.y * 64 + lane;
    double2 acc[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) acc[f] = c2(0.0, 0.0);
    if (jp < nh) {
        const double* P = g.Ps + g.offs[m] + jp;
        for (int i = wave; i < K; i += 4) {
            const double p = P[(size_t)i * nh];
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const double2 c = col[f * K + i];
                acc[f].x = fma(p, c.x, acc[f].x);
                acc[f].y = fma(p, c.y, acc[f].y);
            }
        }
    }
    __syncthreads();
    double2* red = lds;
    if (wave >= 2) {
#pragma unroll
        for (int f = 0; f < NF; ++f) red[((wave - 2) * NF + f) * 64 + lane] = acc[f];
    }
    __syncthreads();
    // wave 0 -> E = (l-m even) sums, wave 1 -> O = (l-m odd) sums
    if (wave < 2 && jp < nh) {
        double2* dst = (wave == 0) ? E : O;
        const size_t plane = (size_t)nh * (L + 1);
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const double2 b = red[(wave * NF + f) * 64 + lane];
            dst[f * plane + (size_t)jp * (L + 1) + m] = cadd(acc[f], b);
        }
    }
}

// ---------------------------------------------------------------------------
// Legendre analysis: acc_f(l, m) = sum_jp Pa * (S or D)[f][jp][m], then the
// epilogue for every (l, m) of column m.  grid (L+1); block 256.
// ---------------------------------------------------------------------------
template <class Epi>
__global__ __launch_bounds__(kLegThreads) void legendre_ana_kernel(GaussDev g,
                                                                  const double2* __restrict__ S,
                                                                  const double2* __restrict__ D,
                                                                  int lmax_a, int mmax_a,
                                                                  Epi ep) {
    constexpr int NF = Epi::NF;
    extern __shared__ double2 lds[];
    const int m = blockIdx.x;
    const int L = g.L, nh = g.nh;
    const int K = L - m + 1;
    double2* sS = lds;              // [NF][nh]
    double2* sD = lds + NF * nh;    // [NF][nh]
    const bool active = (m <= mmax_a);
    if (active) {
        const size_t plane = (size_t)nh * (L + 1);
        for (int t = threadIdx.x; t < NF * nh; t += kLegThreads) {
            const int f = t / nh, jp = t - f * nh;
            const size_t src = f * plane + (size_t)jp * (L + 1) + m;
            sS[t] = S[src];
            sD[t] = D[src];
        }
    }
    __syncthreads();
    for (int l = threadIdx.x; l <= L; l += kLegThreads) {
        double2 acc[NF];
#pragma unroll
        for (int f = 0; f < NF; ++f) acc[f] = c2(0.0, 0.0);
        if (active && l >= m && l <= lmax_a) {
            const double* P = g.Pa + g.offs[m] + (l - m);
            const double2* X = ((l - m) & 1) ? sD : sS;
            for (int jp = 0; jp < nh; ++jp) {
                const double p = P[(size_t)jp * K];
#pragma unroll
                for (int f = 0; f < NF; ++f) {
                    const double2 x = X[f * nh + jp];
                    acc[f].x = fma(p, x.x, acc[f].x);
                    acc[f].y = fma(p, x.y, acc[f].y);
                }
            }
        }
        ep.apply(l, m, acc);
    }
}

// ---------------------------------------------------------------------------
// Fourier stage helpers.
// ---------------------------------------------------------------------------
enum SpecMode { kSpecPlain = 0, kSpecDlam = 1, kSpecScale = 2 };

// Packed spectrum of a north/south row pair for one field:
//   Z(b) = sum over m = b (mod n) of [H_N(m) + i H_S(m)]
//        + sum over m = -b (mod n), m >= 1, of [conj G_N(m) + i conj G_S(m)]
// with H(0) = Re G(0), H(m) = G(m), G_N = E + O, G_S = E - O and the
// optional (i m / R) or 1/R factor.  Its unnormalised inverse DFT is
// f_N + i f_S.
__device__ inline void pack_spectrum(double2* Z, int n, int L, const double2* Er,
                                     const double2* Or, int mode, double R, int t0,
                                     int tstride) {
    const double invR = 1.0 / R;
    for (int b = t0; b < n; b += tstride) {
        double2 z = c2(0.0, 0.0);
        for (int m = b; m <= L; m += n) {
            double2 e = Er[m], o = Or[m];
            double2 gn = cadd(e, o), gs = csub(e, o);
            if (mode == kSpecDlam) { gn = dlam(gn, m, R); gs = dlam(gs, m, R); }
            else if (mode == kSpecScale) { gn = cscale(gn, invR); gs = cscale(gs, invR); }
            if (m == 0) { gn.y = 0.0; gs.y = 0.0; }
            z = cadd(z, c2(gn.x - gs.y, gn.y + gs.x));
        }
        for (int m = (b == 0 ? n : n - b); m <= L; m += n) {
            double2 e = Er[m], o = Or[m];
            double2 gn = cadd(e, o), gs = csub(e, o);
            if (mode == kSpecDlam) { gn = dlam(gn, m, R); gs = dlam(gs, m, R); }
            else if (mode == kSpecScale) { gn = cscale(gn, invR); gs = cscale(gs, invR); }
            z = cadd(z, c2(gn.x + gs.y, -gn.y + gs.x));
        }
        Z[b] = z;
    }
}

// Split the forward DFT of z = g_N + i g_S into S = w (F_N + F_S) and
// D = w (F_N - F_S) for m = 0..mmax.
__device__ inline void unpack_sd(const double2* Z, int n, int mmax, double w, double2* Sr,
                                 double2* Dr, int t0, int tstride) {
    for (int m = t0; m <= mmax; m += tstride) {
        const int b = m % n;
        const int bn = (n - b) % n;
        const double2 zb = Z[b];
        const double2 zn = cconj(Z[bn]);
        const double2 fn = cscale(cadd(zb, zn), 0.5);
        const double2 dd = csub(zb, zn);
        const double2 fs = c2(0.5 * dd.y, -0.5 * dd.x);  // (zb - zn) / (2i)
        Sr[m] = cscale(cadd(fn, fs), w);
        Dr[m] = cscale(csub(fn, fs), w);
    }
}

// Synthesis of nf fields: E/O -> grid rows.  grid (nh); block 256.
__global__ __launch_bounds__(kGridThreads) void grid_syn_kernel(GaussDev g,
                                                               const double2* __restrict__ E,
                                                               const double2* __restrict__ O,
                                                               int nf, double* __restrict__ f_out) {
    extern __shared__ double2 lds[];
    const int jp = blockIdx.x;
    const int n = g.nlon, L = g.L, nh = g.nh;
    const int jN = jp, jS = g.nlat - 1 - jp;
    const size_t plane = (size_t)nh * (L + 1);
    const int cap = kRowElems / n > 0 ? kRowElems / n : 1;  // rows per LDS pass
    for (int f0 = 0; f0 < nf; f0 += cap) {
        const int nr = (nf - f0 < cap) ? nf - f0 : cap;
        for (int r = 0; r < nr; ++r) {
            const size_t off = (size_t)(f0 + r) * plane + (size_t)jp * (L + 1);
            pack_spectrum(lds + r * n, n, L, E + off, O + off, kSpecPlain, 0.0, threadIdx.x,
                          kGridThreads);
        }
        __syncthreads();
        fft_rows<+1>(lds, nr, g.fft);
        for (int t = threadIdx.x; t < nr * n; t += kGridThreads) {
            const int r = t / n, k = t - r * n;
            const double2 z = lds[t];
            double* fo = f_out + (size_t)(f0 + r) * g.nlat * n;
            fo[(size_t)jN * n + k] = z.x;
            if (jS != jN) fo[(size_t)jS * n + k] = z.y;
        }
        __syncthreads();
    }
}

// Analysis of nf grid fields: rows -> S/D (m <= mmax).  grid (nh); block 256.
__global__ __launch_bounds__(kGridThreads) void grid_ana_kernel(GaussDev g,
                                                               const double* __restrict__ f_in,
                                                               int nf, int mmax,
                                                               double2* __restrict__ S,
                                                               double2* __restrict__ D) {
    extern __shared__ double2 lds[];
    const int jp = blockIdx.x;
    const int n = g.nlon, L = g.L, nh = g.nh;
    const int jN = jp, jS = g.nlat - 1 - jp;
    const size_t plane = (size_t)nh * (L + 1);
    const double w = (jN == jS) ? 0.5 * g.w_rows[jN] : g.w_rows[jN];
    const int cap = kRowElems / n > 0 ? kRowElems / n : 1;
    for (int f0 = 0; f0 < nf; f0 += cap) {
        const int nr = (nf - f0 < cap) ? nf - f0 : cap;
        for (int t = threadIdx.x; t < nr * n; t += kGridThreads) {
            const int r = t / n, k = t - r * n;
            const double* fi = f_in + (size_t)(f0 + r) * g.nlat * n;
            lds[t] = c2(fi[(size_t)jN * n + k], fi[(size_t)jS * n + k]);
        }
        __syncthreads();
        fft_rows<-1>(lds, nr, g.fft);
        for (int r = 0; r < nr; ++r) {
            const size_t off = (size_t)(f0 + r) * plane + (size_t)jp * (L + 1);
            unpack_sd(lds + r * n, n, mmax, w, S + off, D + off, threadIdx.x, kGridThreads);
        }
        __syncthreads();
    }
}

// Jacobian: fields (a, S a, b, S b) -> J = (a_s b_l - a_l b_s)/c^2 -> S/D.
__global__ __launch_bounds__(kGridThreads) void grid_jac_kernel(GaussDev g,
                                                               const double2* __restrict__ E,
                                                               const double2* __restrict__ O,
                                                               double R, int mmax,
                                                               double2* __restrict__ S,
                                                               double2* __restrict__ D) {
    extern __shared__ double2 lds[];
    const int jp = blockIdx.x;
    const int n = g.nlon, L = g.L, nh = g.nh;
    const int jN = jp, jS = g.nlat - 1 - jp;
    const size_t plane = (size_t)nh * (L + 1);
    const size_t roff = (size_t)jp * (L + 1);
    // rows: 0 a_l, 1 a_s, 2 b_l, 3 b_s
    for (int r = 0; r < 4; ++r) {
        const size_t off = (size_t)r * plane + roff;
        pack_spectrum(lds + r * n, n, L, E + off, O + off, (r & 1) ? kSpecScale : kSpecDlam,
                      R, threadIdx.x, kGridThreads);
    }
    __syncthreads();
    fft_rows<+1>(lds, 4, g.fft);
    const double cN = clamp_coslat(g.coslat[jN]);
    const double cS = clamp_coslat(g.coslat[jS]);
    for (int k = threadIdx.x; k < n; k += kGridThreads) {
        const double2 al = lds[k], as = lds[n + k], bl = lds[2 * n + k], bs = lds[3 * n + k];
        const double jn = jac_point(al.x, as.x, bl.x, bs.x, cN);
        const double js = jac_point(al.y, as.y, bl.y, bs.y, cS);
        lds[k] = c2(jn, js);
    }
    __syncthreads();
    fft_rows<-1>(lds, 1, g.fft);
    const double w = (jN == jS) ? 0.5 * g.w_rows[jN] : g.w_rows[jN];
    unpack_sd(lds, n, mmax, w, S + roff, D + roff, threadIdx.x, kGridThreads);
}

// Shallow water: 10 Legendre fields -> 12 grid rows -> 4 products -> S/D.
__global__ __launch_bounds__(kGridThreads) void grid_swe_kernel(GaussDev g,
                                                               const double2* __restrict__ E,
                                                               const double2* __restrict__ O,
                                                               double R, int mmax,
                                                               double2* __restrict__ S,
                                                               double2* __restrict__ D) {
    extern __shared__ double2 lds[];
    const int jp = blockIdx.x;
    const int n = g.nlon, L = g.L, nh = g.nh;
    const int jN = jp, jS = g.nlat - 1 - jp;
    const size_t plane = (size_t)nh * (L + 1);
    const size_t roff = (size_t)jp * (L + 1);
    // grid row r <- (Legendre field, spectral factor)
    constexpr int src[12] = {0, 1, 2, 3, 4, 5, 4, 6, 7, 6, 8, 9};
    constexpr int mode[12] = {kSpecDlam, kSpecScale, kSpecDlam, kSpecScale, kSpecDlam,
                              kSpecScale, kSpecPlain, kSpecDlam, kSpecScale, kSpecPlain,
                              kSpecPlain, kSpecPlain};
    for (int r = 0; r < 12; ++r) {
        const size_t off = (size_t)src[r] * plane + roff;
        pack_spectrum(lds + r * n, n, L, E + off, O + off, mode[r], R, threadIdx.x,
                      kGridThreads);
    }
    __syncthreads();
    fft_rows<+1>(lds, 12, g.fft);
    const double cN = clamp_coslat(g.coslat[jN]);
    const double cS = clamp_coslat(g.coslat[jS]);
    for (int k = threadIdx.x; k < n; k += kGridThreads) {
        double gn[12], gs[12], on[4], os[4];
#pragma unroll
        for (int r = 0; r < 12; ++r) {
            const double2 z = lds[r * n + k];
            gn[r] = z.x;
            gs[r] = z.y;
        }
        swe_point(gn, cN, on);
        swe_point(gs, cS, os);
#pragma unroll
        for (int r = 0; r < 4; ++r) lds[r * n + k] = c2(on[r], os[r]);
    }
    __syncthreads();
    fft_rows<-1>(lds, 4, g.fft);
    const double w = (jN == jS) ? 0.5 * g.w_rows[jN] : g.w_rows[jN];
    for (int r = 0; r < 4; ++r) {
        const size_t off = (size_t)r * plane + roff;
        unpack_sd(lds + r * n, n, mmax, w, S + off, D + off, threadIdx.x, kGridThreads);
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static size_t syn_lds(int L, int NF) {
    const int a = NF * (L + 1), b = 2 * NF * 64;
    return sizeof(double2) * (a > b ? a : b);
}
static size_t ana_lds(int nh, int NF) { return sizeof(double2) * 2 * NF * nh; }

template <class Loader>
static int launch_syn(const aeo_plan_s* p, const Loader& ld, double2* E, double2* O,
                      hipStream_t s) {
    dim3 grid(p->L + 1, (p->nh + 63) / 64);
    hipLaunchKernelGGL(legendre_syn_kernel<Loader>, grid, dim3(kLegThreads),
                       syn_lds(p->L, Loader::NF), s, p->gdev(), ld, E, O);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

template <class Epi>
static int launch_ana(const aeo_plan_s* p, const double2* S, const double2* D, int lmax_a,
                      int mmax_a, const Epi& ep, hipStream_t s) {
    hipLaunchKernelGGL(legendre_ana_kernel<Epi>, dim3(p->L + 1), dim3(kLegThreads),
                       ana_lds(p->nh, Epi::NF), s, p->gdev(), S, D, lmax_a, mmax_a, ep);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

int gauss_build_tables(aeo_plan_s* p, const double* d_x, hipStream_t s) {
    dim3 grid(p->L + 1, (p->nh + 63) / 64);
    hipLaunchKernelGGL(legendre_table_kernel, grid, dim3(64), 0, s, p->L, p->nh, d_x, p->d_offs,
                       p->d_Ps, p->d_Pa);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

int gauss_synthesis(aeo_plan_s* p, const double2* a, int nf, double* f, hipStream_t s) {
    const int L = p->L;
    const size_t fs = (size_t)(L + 1) * (L + 1);
    for (int f0 = 0; f0 < nf; f0 += p->max_fields) {
        const int nb = (nf - f0 < p->max_fields) ? nf - f0 : p->max_fields;
        for (int g0 = 0; g0 < nb; g0 += RawLoader::NF) {
            RawLoader ld{a + (size_t)(f0 + g0) * fs, L,
                         (nb - g0 < RawLoader::NF) ? nb - g0 : RawLoader::NF};
            const size_t plane = (size_t)p->nh * (L + 1);
            int rc = launch_syn(p, ld, p->d_E + g0 * plane, p->d_O + g0 * plane, s);
            if (rc) return rc;
        }
        const size_t lds = sizeof(double2) * (size_t)p->nlon *
                           ((kRowElems / p->nlon > 0) ? kRowElems / p->nlon : 1);
        hipLaunchKernelGGL(grid_syn_kernel, dim3(p->nh), dim3(kGridThreads), lds, s, p->gdev(),
                           p->d_E, p->d_O, nb, f + (size_t)f0 * p->nlat * p->nlon);
        AEO_LAUNCH_CHECK();
    }
    return AEO_OK;
}

int gauss_analysis(aeo_plan_s* p, const double* f, int nf, double2* a, hipStream_t s) {
    const int L = p->L;
    const size_t fs = (size_t)(L + 1) * (L + 1);
    const size_t plane = (size_t)p->nh * (L + 1);
    for (int f0 = 0; f0 < nf; f0 += p->max_fields) {
        const int nb = (nf - f0 < p->max_fields) ? nf - f0 : p->max_fields;
        const size_t lds = sizeof(double2) * (size_t)p->nlon *
                           ((kRowElems / p->nlon > 0) ? kRowElems / p->nlon : 1);
        hipLaunchKernelGGL(grid_ana_kernel, dim3(p->nh), dim3(kGridThreads), lds, s, p->gdev(),
                           f + (size_t)f0 * p->nlat * p->nlon, nb, L, p->d_S, p->d_D);
        AEO_LAUNCH_CHECK();
        for (int g0 = 0; g0 < nb; g0 += EpiStore::NF) {
            EpiStore ep{a + (size_t)(f0 + g0) * fs, L,
                        (nb - g0 < EpiStore::NF) ? nb - g0 : EpiStore::NF};
            int rc = launch_ana(p, p->d_S + g0 * plane, p->d_D + g0 * plane, L, L, ep, s);
            if (rc) return rc;
        }
    }
    return AEO_OK;
}

int gauss_jacobian(aeo_plan_s* p, double R, const double2* a, const double2* b, double2* out,
                   int dealias, hipStream_t s) {
    const int L = p->L;
    SpecK k{L, R, p->d_Cp, p->d_Cm};
    JacLoader<false> ld{a, b, false, nullptr, 0.0, k};
    int rc = launch_syn(p, ld, p->d_E, p->d_O, s);
    if (rc) return rc;
    const int lim = dealias ? (2 * L) / 3 : L;
    hipLaunchKernelGGL(grid_jac_kernel, dim3(p->nh), dim3(kGridThreads),
                       sizeof(double2) * 4 * p->nlon, s, p->gdev(), p->d_E, p->d_O, R,
                       lim, p->d_S, p->d_D);
    AEO_LAUNCH_CHECK();
    EpiStore1 ep{out, L};
    return launch_ana(p, p->d_S, p->d_D, lim, lim, ep, s);
}

int gauss_model_stage(aeo_model_s* md, const double2* ys, const double2* y0,
                      const double2* forcing, double2* kout, int stage, double dt,
                      double2* ynext, hipStream_t s) {
    aeo_plan_s* p = md->plan;
    const int L = p->L;
    const int cut = (2 * L) / 3;
    SpecK k{L, md->R, p->d_Cp, p->d_Cm};
    const double f10 = 2.0 * md->omega * sqrt(4.0 * M_PI / 3.0);
    Rk4 rk{stage, stage == 2 ? dt : 0.5 * dt, dt / 6.0, y0, md->d_ksum, ynext, kout};
    ProfRec& pr = md->prof;
    if (md->kind == kBVE) {
        JacLoader<false> ld{ys, nullptr, true, md->d_lap, f10, k};
        pr.begin(kClsSynthesis, s);
        int rc = launch_syn(p, ld, md->d_E, md->d_O, s);
        pr.end(kClsSynthesis, s);
        if (rc) return rc;
        pr.begin(kClsGrid, s);
        hipLaunchKernelGGL(grid_jac_kernel, dim3(p->nh), dim3(kGridThreads),
                           sizeof(double2) * 4 * p->nlon, s, p->gdev(), md->d_E, md->d_O,
                           md->R, cut, md->d_S, md->d_D);
        pr.end(kClsGrid, s);
        AEO_LAUNCH_CHECK();
        EpiBve ep{L, ys, forcing, md->d_nulap, rk};
        pr.begin(kClsAnalysis, s);
        rc = launch_ana(p, md->d_S, md->d_D, cut, cut, ep, s);
        pr.end(kClsAnalysis, s);
        return rc;
    }
    SweLoader<false> ld{ys, md->d_lap, f10, k};
    pr.begin(kClsSynthesis, s);
    int rc = launch_syn(p, ld, md->d_E, md->d_O, s);
    pr.end(kClsSynthesis, s);
    if (rc) return rc;
    const size_t swe_lds = sizeof(double2) * 12 * p->nlon;
    if (swe_lds > 65536) {
        AEO_CHECK_HIP(hipFuncSetAttribute((const void*)grid_swe_kernel,
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)swe_lds));
    }
    pr.begin(kClsGrid, s);
    hipLaunchKernelGGL(grid_swe_kernel, dim3(p->nh), dim3(kGridThreads), swe_lds, s, p->gdev(), md->d_E, md->d_O,
                       md->R, cut, md->d_S, md->d_D);
    pr.end(kClsGrid, s);
    AEO_LAUNCH_CHECK();
    EpiSwe ep{L, ys, md->d_lap, md->nu4 > 0.0 ? md->d_nu4lap2 : nullptr,
              md->has_topo ? md->d_lap_phi_s : nullptr,
              (md->has_topo && md->nu4 > 0.0) ? md->d_nu4_lap2_phi_s : nullptr, md->phi0, rk};
    pr.begin(kClsAnalysis, s);
    rc = launch_ana(p, md->d_S, md->d_D, cut, cut, ep, s);
    pr.end(kClsAnalysis, s);
    return rc;
}

}  // namespace aeo
