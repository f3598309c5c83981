// Point-set (geodesic) back end: dense FP64 basis built on the device by the
// stable normalised recurrence, synthesis as a dense contraction with fused
// pointwise products, analysis as a chunked reduction over points.
//
// Reference: PointSetSphericalHarmonics (num/fast_geodesic_sh.py:15-107):
//   analysis   a_k = sum_i conj(Y_ik) (w_i f_i)
//   synthesis  f_i = 2 Re(sum_k Y_ik a_k) - Re(sum_{m=0} Y_ik a_k)
//                  = sum_k c_k (Re Y_ik Re a_k - Im Y_ik Im a_k),  c = 1 (m = 0), 2 (m > 0)
#include "plan.hpp"
#include "split.hpp"
#include "stage.hpp"

namespace aeo {

constexpr int kPtsThreads = 256;
constexpr int kPtsSynPts = 32;                          // points per synthesis workgroup
constexpr int kPtsSynSlices = kPtsThreads / kPtsSynPts;  // K (basis) slices per point

__device__ inline void k_to_lm(int k, int& l, int& m) {
    int t = (int)((sqrt(8.0 * k + 1.0) - 1.0) * 0.5);
    while ((t + 1) * (t + 2) / 2 <= k) ++t;
    while (t * (t + 1) / 2 > k) --t;
    l = t;
    m = k - t * (t + 1) / 2;
}

// One thread per point: x = cos(pi/2 - lat) (cu/sh_matrix.cu:25-27), stable
// normalised recurrence, phase (cos m lon, sin m lon) (cu/sh_matrix.cu:62-63).
__global__ void points_basis_kernel(int npts, int L, const double* __restrict__ lat,
                                    const double* __restrict__ lon, double* __restrict__ Yk,
                                    double2* __restrict__ Yp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    const int B = (L + 1) * (L + 2) / 2;
    const double x = cos(M_PI / 2.0 - lat[i]);
    const double s = sqrt((1.0 - x) * (1.0 + x));
    const double lo = lon[i];
    double pmm = 0.28209479177387814347;
    for (int m = 0; m <= L; ++m) {
        if (m > 0) pmm *= -sqrt((2.0 * m + 1.0) / (2.0 * m)) * s;
        double sn, cs;
        sincos(m * lo, &sn, &cs);
        double p2 = 0.0, p1 = 0.0;
        for (int l = m; l <= L; ++l) {
            double p;
            if (l == m) {
                p = pmm;
            } else if (l == m + 1) {
                p = sqrt(2.0 * m + 3.0) * x * pmm;
            } else {
                const double ll = (double)l * l, mm = (double)m * m;
                const double a = sqrt((4.0 * ll - 1.0) / (ll - mm));
                const double b =
                    sqrt(((l - 1.0) * (l - 1.0) - mm) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
                p = a * (x * p1 - b * p2);
            }
            p2 = p1;
            p1 = p;
            const int k = l * (l + 1) / 2 + m;
            const double yr = p * cs, yi = p * sn;
            Yk[(size_t)k * npts + i] = yr;
            Yk[(size_t)(B + k) * npts + i] = yi;
            Yp[(size_t)i * B + k] = c2(yr, yi);
        }
    }
}

// Pointwise stages applied to the synthesised values at one point.
struct PwStore {  // generic synthesis
    double* out;
    int npts;
    int nf;
    template <int NF>
    __device__ void apply(int i, const double* v, double, double) const {
        for (int f = 0; f < nf && f < NF; ++f) out[(size_t)f * npts + i] = v[f];
    }
};

__device__ inline PwStore at_group(const PwStore& x, int g) {
    const int f0 = g * RawLoader::NF;
    return PwStore{x.out + (size_t)f0 * x.npts, x.npts, x.nf - f0};
}

struct PwJac {  // v = (a_l, S a, b_l, S b) -> w J, with a_s = Syn(S a) / R
    double* gw;
    double R;
    template <int NF>
    __device__ void apply(int i, const double* v, double w, double c) const {
        const double as = v[1] / R, bs = v[3] / R;
        gw[i] = w * jac_point(v[0], as, v[2], bs, clamp_coslat(c));
    }
};

struct PwSwe {  // 12 synthesised fields -> w * (zeta_dot, curl, kinetic, phi_dot)
    double* gw;
    int npts;
    double R;
    template <int NF>
    __device__ void apply(int i, const double* v, double w, double c) const {
        double g[12], o[4];
#pragma unroll
        for (int r = 0; r < 12; ++r) g[r] = v[r];
        g[1] /= R;
        g[3] /= R;
        g[5] /= R;
        g[8] /= R;
        swe_point(g, clamp_coslat(c), o);
#pragma unroll
        for (int r = 0; r < 4; ++r) gw[(size_t)r * npts + i] = w * o[r];
    }
};

// Workgroup = 64 points (lanes) x 4 slices of the basis index k (waves):
// each wave sums its quarter of k, the four partials meet in LDS in a fixed
// order, and the first wave applies the pointwise stage.
template <class Loader, class Pw>
// The coefficient values sit in LDS tile by tile, `tile` basis entries at a
// time: one tile holds all B of them when they fit (then every slice sums one
// contiguous k range), several when NF B values would pass the LDS.
__global__ __launch_bounds__(kPtsThreads) void points_syn_kernel(PointsDev p, Loader ld0, Pw pw0,
                                                                 int tile) {
    constexpr int NF = Loader::NF;
    const Loader ld = at_group(ld0, (int)blockIdx.y);
    const Pw pw = at_group(pw0, (int)blockIdx.y);
    extern __shared__ double2 lds[];  // [NF][tile] coefficient values times c_k | partials
    const int B = p.B;
    const int lane = threadIdx.x % kPtsSynPts, sl = threadIdx.x / kPtsSynPts;
    const int i = blockIdx.x * kPtsSynPts + lane;
    const int ic = (i < p.npts) ? i : p.npts - 1;   // clamped: every load stays in range
    double acc[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) acc[f] = 0.0;
    const double* yr = p.Yk + ic;
    const double* yi = p.Yk + (size_t)B * p.npts + ic;
    for (int kt0 = 0; kt0 < B; kt0 += tile) {
        const int nt = (B - kt0 < tile) ? B - kt0 : tile;
        if (kt0 > 0) __syncthreads();               // every slice is done with the last tile
        for (int t = threadIdx.x; t < NF * nt; t += kPtsThreads) {
            const int f = t / nt, kk = t - f * nt;
            int l, m;
            k_to_lm(kt0 + kk, l, m);
            const double2 v = ld.value(f, l, m);
            lds[f * tile + kk] = (m == 0) ? v : cscale(v, 2.0);
        }
        __syncthreads();
        const int kq = (nt + kPtsSynSlices - 1) / kPtsSynSlices;
        const int k0 = kt0 + sl * kq, k1 = (k0 + kq < kt0 + nt) ? k0 + kq : kt0 + nt;
#pragma unroll 8
        for (int k = k0; k < k1; ++k) {
            const double a = yr[(size_t)k * p.npts];
            const double b = yi[(size_t)k * p.npts];
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const double2 c = lds[f * tile + (k - kt0)];
                acc[f] = fma(a, c.x, acc[f]);
                acc[f] = fma(-b, c.y, acc[f]);
            }
        }
    }
    double* part = reinterpret_cast<double*>(lds + (size_t)NF * tile);   // [slices][NF][64]
#pragma unroll
    for (int f = 0; f < NF; ++f) part[(sl * NF + f) * kPtsSynPts + lane] = acc[f];
    __syncthreads();
    if (sl == 0 && i < p.npts) {
        double v[NF];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            double s = part[f * kPtsSynPts + lane];
            for (int q = 1; q < kPtsSynSlices; ++q) s += part[(q * NF + f) * kPtsSynPts + lane];
            v[f] = s;
        }
        pw.template apply<NF>(i, v, p.w[i], p.coslat[i]);
    }
}

// Partial analysis sums: part[chunk][f][k] = sum_{i in chunk} conj(Y_ik) g_i,
// g = values * w (weights != nullptr) or values already weighted.
// grid (ceil(B/64), nchunks, nf); block 256 = 64 k-lanes x 4 point groups.
__global__ __launch_bounds__(kPtsThreads) void points_partial_kernel(
    PointsDev p, const double* __restrict__ vals, const double* __restrict__ weights,
    int nf, double2* __restrict__ part) {
    __shared__ double2 red[4][64];
    __shared__ double gs[kPtsChunk];   // the chunk's weighted values, one vector load each
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + lane;
    const int chunk = blockIdx.y, f = blockIdx.z;
    const int i0 = chunk * kPtsChunk;
    const int i1 = (i0 + kPtsChunk < p.npts) ? i0 + kPtsChunk : p.npts;
    const double* v = vals + (size_t)f * p.npts;
    for (int t = threadIdx.x; t < i1 - i0; t += kPtsThreads)
        gs[t] = weights ? v[i0 + t] * weights[i0 + t] : v[i0 + t];
    __syncthreads();
    double2 acc = c2(0.0, 0.0);
    if (k < p.B) {
#pragma unroll 8
        for (int i = i0 + q; i < i1; i += 4) {
            const double g = gs[i - i0];
            const double2 y = p.Yp[(size_t)i * p.B + k];
            acc.x = fma(y.x, g, acc.x);
            acc.y = fma(-y.y, g, acc.y);
        }
    }
    red[q][lane] = acc;
    __syncthreads();
    if (q == 0 && k < p.B) {
        double2 s = red[0][lane];
        s = cadd(s, red[1][lane]);
        s = cadd(s, red[2][lane]);
        s = cadd(s, red[3][lane]);
        part[((size_t)chunk * nf + f) * p.B + k] = s;
    }
}

// Plain analysis output for few chunks: one thread per (field, l, m) sums the
// chunk partials in chunk order (padding m > l written as 0).  grid
// ceil(nf (L+1)^2 / 256); used when nchunks <= kPtsSeqChunks, where the
// wave-per-entry tree below is bound by its shuffle chain.
constexpr int kPtsSeqChunks = 32;

__global__ __launch_bounds__(kPtsThreads) void points_store_sum_kernel(
    int L, int B, int nchunks, int nf, const double2* __restrict__ part,
    double2* __restrict__ out) {
    const int n = L + 1, fs = n * n;
    const size_t t = (size_t)blockIdx.x * kPtsThreads + threadIdx.x;
    if (t >= (size_t)nf * fs) return;
    const int f = (int)(t / fs), r = (int)(t - (size_t)f * fs);
    const int l = r / n, m = r - l * n;
    double2 acc = c2(0.0, 0.0);
    if (m <= l) {
        const int k = l * (l + 1) / 2 + m;
#pragma unroll 8
        for (int c = 0; c < nchunks; ++c) acc = cadd(acc, part[((size_t)c * nf + f) * B + k]);
    }
    out[t] = acc;
}

// Sum the chunk partials and run the epilogue for every (l, m): one wave per
// entry, lanes over chunks, a fixed shuffle tree (deterministic), lane 0
// applies the epilogue.
template <class Epi>
__global__ __launch_bounds__(kPtsThreads) void points_finish_kernel(
    int L, int B, int nchunks, int nf_part, int nf_valid, const double2* __restrict__ part,
    int lmax_a, int mmax_a, Epi ep0) {
    constexpr int NF = Epi::NF;
    // field group blockIdx.y: its partials start NF * B into each chunk
    const int g = blockIdx.y;
    const Epi ep = at_group(ep0, g);
    part += (size_t)g * NF * B;
    nf_valid = (nf_valid - g * NF < NF) ? nf_valid - g * NF : NF;
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * (kPtsThreads / 64) + (threadIdx.x >> 6);
    const int n = L + 1;
    if (t >= n * n) return;   // whole waves exit together
    const int l = t / n, m = t - l * n;
    double2 acc[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) acc[f] = c2(0.0, 0.0);
    if (m <= l && l <= lmax_a && m <= mmax_a) {
        const int k = l * (l + 1) / 2 + m;
        for (int c = lane; c < nchunks; c += 64) {
#pragma unroll
            for (int f = 0; f < NF; ++f)
                if (f < nf_valid) acc[f] = cadd(acc[f], part[((size_t)c * nf_part + f) * B + k]);
        }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f)
        for (int off = 32; off > 0; off >>= 1) {
            acc[f].x += __shfl_xor(acc[f].x, off);
            acc[f].y += __shfl_xor(acc[f].y, off);
        }
    if (lane == 0) ep.apply(l, m, acc);
}

// ---------------------------------------------------------------------------
// Split precision (INT8 digit MFMA, exact INT32 sums, FP64 recombination; the
// digit scheme is the one of the Gauss path, split.hpp).  The dense complex
// contractions are written as real GEMMs:
//   synthesis  f[i][f] = sum_k' A[i][k'] C[k'][f],  A = (Re Y, -Im Y) over
//              k' < 2B, C = (Re, Im) of the c_k-scaled coefficients;
//   analysis   part[r][f] = sum_i A[r][i] g[i][f],  A rows (Re Y_k, -Im Y_k),
//              r < 2B = (Re, Im) of conj(Y) g summed over one point chunk.
// Tables (one-time build): one thread per row, exponent over the whole row.
// ---------------------------------------------------------------------------
__global__ void points_split_table_kernel(PointsDev p, int ns, signed char* __restrict__ Ysl,
                                          int* __restrict__ eYs, signed char* __restrict__ Yasl,
                                          int* __restrict__ eYa) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int B = p.B, n = p.npts;
    int d[8];
    if (r < p.nprow) {   // synthesis row = point r; Yk[k' * n + r] covers both planes
        double mx = 0.0;
        if (r < n)
            for (int k = 0; k < 2 * B; ++k) mx = fmax(mx, fabs(p.Yk[(size_t)k * n + r]));
        const int e = exp_above(mx);
        eYs[r] = e;
        for (int k = 0; k < p.K2; ++k) {
            const double v = (r < n && k < 2 * B)
                                 ? ((k < B) ? 1.0 : -1.0) * p.Yk[(size_t)k * n + r] : 0.0;
            digits_rt(v, e, ns, d);
            for (int q = 0; q < ns; ++q) Ysl[((size_t)q * p.nprow + r) * p.K2 + k] = (signed char)d[q];
        }
    }
    if (r < p.R2p) {     // analysis row r: Re Y_r (r < B) or -Im Y_(r-B) over the points
        const double sg = (r < B) ? 1.0 : -1.0;
        double mx = 0.0;
        if (r < 2 * B)
            for (int i = 0; i < n; ++i) mx = fmax(mx, fabs(p.Yk[(size_t)r * n + i]));
        const int e = exp_above(mx);
        eYa[r] = e;
        for (int i = 0; i < p.Npad; ++i) {
            const double v = (r < 2 * B && i < n) ? sg * p.Yk[(size_t)r * n + i] : 0.0;
            digits_rt(v, e, ns, d);
            for (int q = 0; q < ns; ++q) Yasl[((size_t)q * p.R2p + r) * p.Npad + i] = (signed char)d[q];
        }
    }
}

constexpr int kPtsSplitKC = 256;   // K values per LDS digit pass

// Max exponent of 16 columns (one wave per column, lanes over K).
template <class F>
__device__ inline void pts_col_exp(int len, int* ex, const F& val) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = wave; c < 16; c += kPtsThreads / 64) {
        double mx = 0.0;
        for (int k = lane; k < len; k += 64) mx = fmax(mx, fabs(val(c, k)));
        for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
        if (lane == 0) ex[c] = exp_above(mx);
    }
}

// Digits of 16 columns x kPtsSplitKC K values (pairs over threads) into
// xs[slice][16][kPtsSplitKC].
template <int NS, class F>
__device__ inline void pts_digits(int k0, const int* ex, unsigned short* xsh, const F& val) {
    constexpr int KP = kPtsSplitKC / 2;
    for (int t = threadIdx.x; t < 16 * KP; t += kPtsThreads) {
        const int c = t / KP, kk = 2 * (t - c * KP);
        int d0[NS], d1[NS];
        digits<NS>(val(c, k0 + kk), ex[c], d0);
        digits<NS>(val(c, k0 + kk + 1), ex[c], d1);
#pragma unroll
        for (int q = 0; q < NS; ++q)
            xsh[(((size_t)q * 16 + c) * kPtsSplitKC + kk) >> 1] = digit_pair(d0[q], d1[q]);
    }
}

// One 16-row tile x one kPtsSplitKC pass of MFMAs: rows of `tab` (row stride
// ld, K offset k0) against xs.
template <int NS>
__device__ inline void pts_mfma_pass(const signed char* tab, size_t slice_stride, size_t ld,
                                     int row0, int k0, const signed char* xs, intx4* acc) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int c4 = 0; c4 < kPtsSplitKC / kSplitK; ++c4) {
        intx4 av[NS];
#pragma unroll
        for (int q = 0; q < NS; ++q)
            av[q] = *reinterpret_cast<const intx4*>(tab + q * slice_stride +
                                                    (size_t)(row0 + (lane & 15)) * ld + k0 +
                                                    c4 * kSplitK + (lane >> 4) * 16);
#pragma unroll
        for (int t2 = 0; t2 < NS; ++t2) {
            const intx4 bv = *reinterpret_cast<const intx4*>(
                xs + ((size_t)t2 * 16 + (lane & 15)) * kPtsSplitKC + c4 * kSplitK + (lane >> 4) * 16);
#pragma unroll
            for (int q = 0; q + t2 < NS; ++q)
                acc[q + t2] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[q], bv, acc[q + t2], 0, 0, 0);
        }
    }
}

template <int NS>
__device__ inline double pts_recombine(const intx4* acc, int r, int e) {
    double v = 0.0;
#pragma unroll
    for (int q = NS - 1; q >= 0; --q) v = fma(v, kInvRadix, (double)acc[q][r]);
    return ldexp(v, e) * kInvTop2;
}

// Synthesis + pointwise stage: workgroup = 4 waves x 16 points; the loader's
// NF <= 16 fields are the 16 MFMA columns.
template <class Loader, class Pw, int NS>
__global__ __launch_bounds__(kPtsThreads) void points_syn_split_kernel(PointsDev p, Loader ld,
                                                                      Pw pw) {
    constexpr int NF = Loader::NF;
    static_assert(NF <= 16, "one 16-column MFMA tile");
    extern __shared__ double2 lds[];
    const int B = p.B, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double2* cv = lds;                                              // [NF][B]
    double* outv = reinterpret_cast<double*>(cv + (size_t)NF * B);  // [4][16][NF]
    int* fexp = reinterpret_cast<int*>(outv + 4 * 16 * NF);          // [16]
    signed char* xs = reinterpret_cast<signed char*>(fexp + 16);     // [NS][16][KC]
    for (int t = threadIdx.x; t < NF * B; t += kPtsThreads) {
        const int f = t / B, k = t - f * B;
        int l, m;
        k_to_lm(k, l, m);
        const double2 v = ld.value(f, l, m);
        cv[t] = (m == 0) ? v : cscale(v, 2.0);
    }
    __syncthreads();
    auto col = [&](int c, int k) -> double {
        if (c >= NF || k >= 2 * B) return 0.0;
        return (k < B) ? cv[c * B + k].x : cv[c * B + k - B].y;
    };
    pts_col_exp(2 * B, fexp, col);
    __syncthreads();
    const int row0 = (blockIdx.x * 4 + wave) * 16;
    intx4 acc[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) acc[q] = intx4{0, 0, 0, 0};
    for (int k0 = 0; k0 < p.K2; k0 += kPtsSplitKC) {
        pts_digits<NS>(k0, fexp, reinterpret_cast<unsigned short*>(xs), col);
        __syncthreads();
        if (row0 < p.nprow)
            pts_mfma_pass<NS>(p.Ysl, (size_t)p.nprow * p.K2, p.K2, row0, k0, xs, acc);
        __syncthreads();
    }
    // recombine: lane = (column f, 4 rows); gather each point's fields in LDS
    double* ow = outv + (size_t)wave * 16 * NF;
    const int f = lane & 15;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int ir = (lane >> 4) * 4 + r;
        const int i = row0 + ir;
        if (f < NF && i < p.nprow) ow[ir * NF + f] = pts_recombine<NS>(acc, r, p.eYs[i] + fexp[f]);
    }
    __syncthreads();
    if (lane < 16) {
        const int i = row0 + lane;
        if (i < p.npts) pw.template apply<NF>(i, ow + lane * NF, p.w[i], p.coslat[i]);
    }
}

// Partial analysis sums of one point chunk: grid (R2p / 64, nchunks); wave =
// 16 rows of the 2B real rows, columns = the nf <= 16 fields; per-chunk
// column exponents (the partials are recombined per chunk in FP64).
template <int NS>
__global__ __launch_bounds__(kPtsThreads) void points_partial_split_kernel(
    PointsDev p, const double* __restrict__ vals, const double* __restrict__ weights, int nf,
    double2* __restrict__ part) {
    __shared__ int gexp[16];
    __shared__ __align__(16) signed char xs[NS * 16 * kPtsSplitKC];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int B = p.B, chunk = blockIdx.y;
    const int i0 = chunk * kPtsChunk;
    auto col = [&](int c, int i) -> double {
        const int ii = i0 + i;
        if (c >= nf || i >= kPtsChunk || ii >= p.npts) return 0.0;
        const double v = vals[(size_t)c * p.npts + ii];
        return weights ? v * weights[ii] : v;
    };
    pts_col_exp(kPtsChunk, gexp, col);
    __syncthreads();
    const int row0 = (blockIdx.x * 4 + wave) * 16;
    intx4 acc[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) acc[q] = intx4{0, 0, 0, 0};
    for (int s0 = 0; s0 < kPtsChunk; s0 += kPtsSplitKC) {
        pts_digits<NS>(s0, gexp, reinterpret_cast<unsigned short*>(xs), col);
        __syncthreads();
        if (row0 < p.R2p)
            pts_mfma_pass<NS>(p.Yasl, (size_t)p.R2p * p.Npad, p.Npad, row0, i0 + s0, xs, acc);
        __syncthreads();
    }
    double* pd = reinterpret_cast<double*>(part);
    const int f = lane & 15;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int rr = row0 + (lane >> 4) * 4 + r;
        if (f < nf && rr < 2 * B) {
            const int k = (rr < B) ? rr : rr - B;
            pd[2 * (((size_t)chunk * nf + f) * B + k) + (rr >= B)] =
                pts_recombine<NS>(acc, r, p.eYa[rr] + gexp[f]);
        }
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <class Loader, class Pw>
static int launch_psyn(const aeo_plan_s* p, const Loader& ld, const Pw& pw, hipStream_t s,
                       int groups = 1) {
    // all B coefficient values of the NF fields in LDS when they fit, else
    // tiles of a multiple of 64 entries
    const size_t fixed = sizeof(double) * kPtsSynSlices * Loader::NF * kPtsSynPts;
    const size_t per = sizeof(double2) * Loader::NF;
    int tile = p->B;
    if (per * p->B + fixed > 163840) {
        AEO_REQUIRE(fixed + 64 * per <= 163840, AEO_ERR_UNSUPPORTED,
                    "point synthesis: LDS demand exceeds 160 KiB for this field batch");
        tile = (int)(((163840 - fixed) / per) & ~(size_t)63);
    }
    const size_t lds = per * tile + fixed;
    dim3 grid((p->npts + kPtsSynPts - 1) / kPtsSynPts, groups);
    if (lds > 65536) {
        AEO_CHECK_HIP(hipFuncSetAttribute((const void*)points_syn_kernel<Loader, Pw>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    hipLaunchKernelGGL((points_syn_kernel<Loader, Pw>), grid, dim3(kPtsThreads), lds, s,
                       p->pdev(), ld, pw, tile);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

static int launch_partial(const aeo_plan_s* p, const double* vals, const double* w, int nf,
                          double2* part, hipStream_t s) {
    dim3 grid((p->B + 63) / 64, p->nchunks, nf);
    hipLaunchKernelGGL(points_partial_kernel, grid, dim3(kPtsThreads), 0, s, p->pdev(), vals, w,
                       nf, part);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

template <class Loader, class Pw, int NS>
static int launch_psyn_split_ns(const aeo_plan_s* p, const Loader& ld, const Pw& pw,
                                hipStream_t s) {
    constexpr int NF = Loader::NF;
    const size_t lds = sizeof(double2) * NF * p->B + sizeof(double) * 4 * 16 * NF +
                       sizeof(int) * 16 + (size_t)NS * 16 * kPtsSplitKC;
    AEO_REQUIRE(lds <= 163840, AEO_ERR_UNSUPPORTED,
                "split point synthesis: LDS demand exceeds 160 KiB at this l_max");
    auto kern = points_syn_split_kernel<Loader, Pw, NS>;
    if (lds > 65536)
        AEO_CHECK_HIP(hipFuncSetAttribute((const void*)kern,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3((p->nprow + 63) / 64), dim3(kPtsThreads), lds, s, p->pdev(),
                       ld, pw);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

template <int NS>
static int launch_partial_split_ns(const aeo_plan_s* p, const double* vals, const double* w,
                                   int nf, double2* part, hipStream_t s) {
    AEO_REQUIRE(nf <= 16, AEO_ERR_UNSUPPORTED, "split point analysis: at most 16 fields");
    hipLaunchKernelGGL(points_partial_split_kernel<NS>, dim3((p->R2p + 63) / 64, p->nchunks),
                       dim3(kPtsThreads), 0, s, p->pdev(), vals, w, nf, part);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

// ns = 0: FP64 contractions; ns in {4, 5, 6, 7}: split digits (plan tables must match).
template <class Loader, class Pw>
static int launch_psyn_ns(const aeo_plan_s* p, int ns, const Loader& ld, const Pw& pw,
                          hipStream_t s) {
    if (ns != 0)
        AEO_REQUIRE(p->nslices == ns, AEO_ERR_STATE,
                    "split point synthesis: the plan's digit tables have a different slice count");
    switch (ns) {
        case 0: return launch_psyn(p, ld, pw, s);
        case 4: return launch_psyn_split_ns<Loader, Pw, 4>(p, ld, pw, s);
        case 5: return launch_psyn_split_ns<Loader, Pw, 5>(p, ld, pw, s);
        case 6: return launch_psyn_split_ns<Loader, Pw, 6>(p, ld, pw, s);
        case 7: return launch_psyn_split_ns<Loader, Pw, 7>(p, ld, pw, s);
        default:
            set_error("split point synthesis: nslices must be 4, 5, 6 or 7");
            return AEO_ERR_UNSUPPORTED;
    }
}

static int launch_partial_ns(const aeo_plan_s* p, int ns, const double* vals, const double* w,
                             int nf, double2* part, hipStream_t s) {
    if (ns != 0)
        AEO_REQUIRE(p->nslices == ns, AEO_ERR_STATE,
                    "split point analysis: the plan's digit tables have a different slice count");
    switch (ns) {
        case 0: return launch_partial(p, vals, w, nf, part, s);
        case 4: return launch_partial_split_ns<4>(p, vals, w, nf, part, s);
        case 5: return launch_partial_split_ns<5>(p, vals, w, nf, part, s);
        case 6: return launch_partial_split_ns<6>(p, vals, w, nf, part, s);
        case 7: return launch_partial_split_ns<7>(p, vals, w, nf, part, s);
        default:
            set_error("split point analysis: nslices must be 4, 5, 6 or 7");
            return AEO_ERR_UNSUPPORTED;
    }
}

template <class Epi>
static int launch_finish(const aeo_plan_s* p, int nf_part, int nf_valid, const double2* part,
                         int lmax_a, int mmax_a, const Epi& ep, hipStream_t s, int groups = 1) {
    const int n = (p->L + 1) * (p->L + 1);
    constexpr int per = kPtsThreads / 64;   // entries (waves) per workgroup
    hipLaunchKernelGGL(points_finish_kernel<Epi>, dim3((n + per - 1) / per, groups),
                       dim3(kPtsThreads), 0,
                       s, p->L, p->B, p->nchunks, nf_part, nf_valid, part, lmax_a, mmax_a, ep);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

int points_build_tables(aeo_plan_s* p, const double* d_lat, const double* d_lon, hipStream_t s) {
    hipLaunchKernelGGL(points_basis_kernel, dim3((p->npts + 255) / 256), dim3(256), 0, s,
                       p->npts, p->L, d_lat, d_lon, p->d_Yk, p->d_Yp);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

int points_build_split_tables(aeo_plan_s* p, int ns, hipStream_t s) {
    const int rows = (p->nprow > p->R2p) ? p->nprow : p->R2p;
    hipLaunchKernelGGL(points_split_table_kernel, dim3((rows + 63) / 64), dim3(64), 0, s,
                       p->pdev(), ns, p->d_Ysl, p->d_eYs, p->d_Yasl, p->d_eYa);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

int points_synthesis(aeo_plan_s* p, const double2* a, int nf, double* f, hipStream_t s) {
    const size_t fs = (size_t)(p->L + 1) * (p->L + 1);
    // one launch: grid y runs over groups of RawLoader::NF fields
    RawLoader ld{a, p->L, nf};
    PwStore pw{f, p->npts, nf};
    (void)fs;
    return launch_psyn(p, ld, pw, s, (nf + RawLoader::NF - 1) / RawLoader::NF);
}

int points_analysis(aeo_plan_s* p, const double* f, int nf, double2* a, hipStream_t s) {
    const size_t fs = (size_t)(p->L + 1) * (p->L + 1);
    for (int f0 = 0; f0 < nf; f0 += p->max_fields) {
        const int nb = (nf - f0 < p->max_fields) ? nf - f0 : p->max_fields;
        int rc = launch_partial(p, f + (size_t)f0 * p->npts, p->d_w, nb, p->d_part, s);
        if (rc) return rc;
        if (p->nchunks <= kPtsSeqChunks) {
            const size_t tot = (size_t)nb * fs;
            hipLaunchKernelGGL(points_store_sum_kernel, dim3((unsigned)((tot + kPtsThreads - 1) /
                                                                        kPtsThreads)),
                               dim3(kPtsThreads), 0, s, p->L, p->B, p->nchunks, nb, p->d_part,
                               a + (size_t)f0 * fs);
            AEO_LAUNCH_CHECK();
            continue;
        }
        // one launch: grid y runs over groups of EpiStore::NF fields
        EpiStore ep{a + (size_t)f0 * fs, p->L, nb};
        rc = launch_finish(p, nb, nb, p->d_part, p->L, p->L, ep, s,
                           (nb + EpiStore::NF - 1) / EpiStore::NF);
        if (rc) return rc;
    }
    return AEO_OK;
}

int points_jacobian(aeo_plan_s* p, double R, const double2* a, const double2* b, double2* out,
                    int dealias, hipStream_t s) {
    SpecK k{p->L, R, p->d_Cp, p->d_Cm};
    JacLoader<true> ld{a, b, false, nullptr, 0.0, k};
    PwJac pw{p->d_gw, R};
    int rc = launch_psyn(p, ld, pw, s);
    if (rc) return rc;
    rc = launch_partial(p, p->d_gw, nullptr, 1, p->d_part, s);
    if (rc) return rc;
    const int lim = dealias ? (2 * p->L) / 3 : p->L;
    EpiStore1 ep{out, p->L};
    return launch_finish(p, 1, 1, p->d_part, lim, lim, ep, s);
}

int points_model_stage(aeo_model_s* md, const double2* ys, const double2* y0,
                       const double2* forcing, double2* kout, int stage, double dt,
                       double2* ynext, hipStream_t s) {
    aeo_plan_s* p = md->plan;
    const int L = p->L;
    const int cut = (2 * L) / 3;
    SpecK k{L, md->R, p->d_Cp, p->d_Cm};
    const double f10 = 2.0 * md->omega * sqrt(4.0 * M_PI / 3.0);
    Rk4 rk{stage, stage == 2 ? dt : 0.5 * dt, dt / 6.0, y0, md->d_ksum, ynext, kout};
    const int ns = md->nslices;
    if (md->kind == kBVE) {
        JacLoader<true> ld{ys, nullptr, true, md->d_lap, f10, k};
        PwJac pw{md->d_gw, md->R};
        int rc = launch_psyn_ns(p, ns, ld, pw, s);
        if (rc) return rc;
        rc = launch_partial_ns(p, ns, md->d_gw, nullptr, 1, md->d_part, s);
        if (rc) return rc;
        EpiBve ep{L, ys, forcing, md->d_nulap, rk};
        return launch_finish(p, 1, 1, md->d_part, cut, cut, ep, s);
    }
    SweLoader<true> ld{ys, md->d_lap, f10, k};
    PwSwe pw{md->d_gw, p->npts, md->R};
    int rc = launch_psyn_ns(p, ns, ld, pw, s);
    if (rc) return rc;
    rc = launch_partial_ns(p, ns, md->d_gw, nullptr, 4, md->d_part, s);
    if (rc) return rc;
    EpiSwe ep{L, ys, md->d_lap, md->nu4 > 0.0 ? md->d_nu4lap2 : nullptr,
              md->has_topo ? md->d_lap_phi_s : nullptr,
              (md->has_topo && md->nu4 > 0.0) ? md->d_nu4_lap2_phi_s : nullptr, md->phi0, rk};
    return launch_finish(p, 4, 4, md->d_part, cut, cut, ep, s);
}

}  // namespace aeo
