// Member-batched split-precision BVE stage in GEMM form.
//
// The fused stage kernel (legendre.hpp) gives a workgroup one zonal
// wavenumber m and a few members: every member group streams the digit
// tables of column m again, an MFMA tile spans at most a few members, and
// each workgroup runs the whole chain (analysis, epilogue, synthesis) at one
// workgroup per CU.  For large member batches the same stage runs here as
// three kernels whose MFMA N dimension spans the batch:
//   gemm_ana_kernel  (m, parity, 32 members): digits of the weighted Fourier
//                    coefficients S (even) or D (odd) with one exponent per
//                    column over the latitudes, INT8 MFMA against the
//                    analysis digit table of m, FP64 recombination
//                    -> T[m][l][member]
//   gemm_epi_kernel  (m, 4 or 8 members): tendency and RK4 stage bookkeeping from
//                    T, then the synthesis fields psi, S psi, eta, S eta of
//                    the new stage state with one exponent per column over
//                    the degrees, written as digits Qd[m][parity][slice][col][k]
//   gemm_syn_kernel  (m, parity, column group): INT8 MFMA of the synthesis
//                    digit table of m against Qd (a wave tile: 16 latitudes x
//                    2 or 4 column tiles of 16), FP64 recombination -> E/O
// Digits, exponents, the exact INT32 sums and the recombination are those of
// the fused split kernel, so both forms give bitwise equal results (the
// order of exact integer sums does not matter, and every FP64 expression is
// written out alike).
#include <string>
#include <vector>

#include "stage_launch.hpp"

namespace aeo {

constexpr int kGemmThreads = 512;
constexpr int kGemmWaves = kGemmThreads / 64;
constexpr int kAnaMembers = 32;      // members per analysis workgroup: 64 columns
constexpr int kSynColAlign = 128;    // synthesis columns per m are padded to this

struct GemmArgs {
    int L, nh, nmem;
    int top, mmax;          // analysed band: l <= top, m <= mmax
    int ncp;                // synthesis columns per (m, parity, slice) row block of Qd
    const double2* S;       // [jp][m][member]
    const double2* D;
    double* T;              // [m][l][member][re, im]
    signed char* Qd;        // block m at 2 NS ncp kpoff[m]: [parity][slice][ncp][kpad_m]
    int* Qe;                // [m][ncp] column exponents
    const int* kpoff;       // sum of kpad over m' < m
    double2* E;             // [jp][m][member][4]
    double2* O;
    const double2* ys;      // stage input states, [member][m][l]
    size_t ys_stride;
    int ana, syn;
    ColK k;
    EpiBveS ep;
};

// ---------------------------------------------------------------------------
// Analysis: T = Pa_digits(m, parity) x X_digits, X = S (even) or D (odd).
// ---------------------------------------------------------------------------
__host__ __device__ inline int gemm_xs_stride(int nh) { return split_nhk(nh) + 16; }

template <int NS>
__host__ __device__ inline size_t gemm_ana_lds(int nh) {
    constexpr int NC = 2 * kAnaMembers;
    const int nhk = split_nhk(nh);
    const size_t red = sizeof(double) * NC * (nhk / 8);
    const size_t xs = (size_t)NS * NC * gemm_xs_stride(nh);
    return 256 + (red > xs ? red : xs);
}

template <int NS>
__global__ __launch_bounds__(kGemmThreads) void gemm_ana_kernel(GaussDev g, GemmArgs a) {
    constexpr int NC = 2 * kAnaMembers;       // columns (member, re/im)
    constexpr int IPT = 2;                    // items (member, 8 latitudes) per thread
    extern __shared__ __align__(16) unsigned char gsm[];
    const int m = blockIdx.x, p = blockIdx.y, b0 = blockIdx.z * kAnaMembers;
    const int L = g.L, nh = g.nh, n = L + 1;
    const int nhk = split_nhk(nh), XS = gemm_xs_stride(nh);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int* ex = reinterpret_cast<int*>(gsm);                           // [NC]
    double* red = reinterpret_cast<double*>(gsm + 256);              // [nhk / 8][NC]
    signed char* xs = reinterpret_cast<signed char*>(gsm + 256);     // [NS][NC][XS] (after red)

    // the wave's first table tile and row exponents, issued with the loads
    const int Mv = (a.top >= m + p) ? (a.top - m - p) / 2 + 1 : 0;
    const int MT = (Mv + 15) / 16;
    const int mpad = rup(split_kp(L, m), 16);
    const signed char* Pt = g.Pasl + g.offs_sa[m];
    const int* er = g.eana + g.offs_ea[m];

    // ---- the S or D coefficients of column m for 32 members, 8 latitudes per item
    const double2* X = p ? a.D : a.S;
    const int nitems = kAnaMembers * (nhk / 8);
    double2 v[IPT][8];
#pragma unroll
    for (int u = 0; u < IPT; ++u) {
        const int it = threadIdx.x + u * kGemmThreads;
        const int bl = it % kAnaMembers, jb = it / kAnaMembers;
        const int b = b0 + bl;
        const bool ok = it < nitems && b < a.nmem;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int j = jb * 8 + i;
            v[u][i] = (ok && j < nh) ? X[((size_t)j * n + m) * a.nmem + b] : c2(0.0, 0.0);
        }
    }
#pragma unroll
    for (int u = 0; u < IPT; ++u) {
        const int it = threadIdx.x + u * kGemmThreads;
        if (it < nitems) {
            const int bl = it % kAnaMembers, jb = it / kAnaMembers;
            double mr = 0.0, mi = 0.0;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                mr = fmax(mr, fabs(v[u][i].x));
                mi = fmax(mi, fabs(v[u][i].y));
            }
            red[jb * NC + 2 * bl] = mr;
            red[jb * NC + 2 * bl + 1] = mi;
        }
    }
    __syncthreads();
    if (threadIdx.x < NC) {
        double mx = 0.0;
        for (int jb = 0; jb < nhk / 8; ++jb) mx = fmax(mx, red[jb * NC + threadIdx.x]);
        ex[threadIdx.x] = exp_above(mx);
    }
    __syncthreads();
    // digits: 8 latitudes of one column per 8-byte LDS store and slice
#pragma unroll
    for (int u = 0; u < IPT; ++u) {
        const int it = threadIdx.x + u * kGemmThreads;
        if (it < nitems) {
            const int bl = it % kAnaMembers, jb = it / kAnaMembers;
#pragma unroll
            for (int ri = 0; ri < 2; ++ri) {
                const int e = ex[2 * bl + ri];
                unsigned int lo[NS], hi[NS];
#pragma unroll
                for (int q = 0; q < NS; ++q) lo[q] = hi[q] = 0u;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    int d[NS];
                    digits<NS>(ri ? v[u][i].y : v[u][i].x, e, d);
#pragma unroll
                    for (int q = 0; q < NS; ++q) {
                        const unsigned int byte = (unsigned int)(d[q] & 0xff) << (8 * (i & 3));
                        if (i < 4) lo[q] |= byte; else hi[q] |= byte;
                    }
                }
#pragma unroll
                for (int q = 0; q < NS; ++q)
                    *reinterpret_cast<uint2*>(xs + ((size_t)q * NC + 2 * bl + ri) * XS + jb * 8) =
                        make_uint2(lo[q], hi[q]);
            }
        }
    }
    __syncthreads();

    // ---- (16-degree tile, pair of 16-column tiles) per wave
    constexpr int NG = NC / 32;
    for (int tt = wave; tt < MT * NG; tt += kGemmWaves) {
        const int mt = tt / NG, ng = tt - mt * NG;
        intx4 acc[NS][2];
#pragma unroll
        for (int q = 0; q < NS; ++q) acc[q][0] = acc[q][1] = intx4{0, 0, 0, 0};
        for (int k0 = 0; k0 < nhk; k0 += kSplitK) {
            intx4 av[NS];
#pragma unroll
            for (int q = 0; q < NS; ++q)
                av[q] = *reinterpret_cast<const intx4*>(
                    Pt + ((size_t)(p * NS + q) * mpad + mt * 16 + (lane & 15)) * nhk + k0 +
                    (lane >> 4) * 16);
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int t2 = 0; t2 < NS; ++t2) {
                    const intx4 bv = *reinterpret_cast<const intx4*>(
                        xs + ((size_t)t2 * NC + (2 * ng + h) * 16 + (lane & 15)) * XS + k0 +
                        (lane >> 4) * 16);
#pragma unroll
                    for (int q = 0; q + t2 < NS; ++q)
                        acc[q + t2][h] =
                            __builtin_amdgcn_mfma_i32_16x16x64_i8(av[q], bv, acc[q + t2][h], 0, 0, 0);
                }
        }
        int erow[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) erow[r] = er[p * mpad + mt * 16 + (lane >> 4) * 4 + r];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int c = (2 * ng + h) * 16 + (lane & 15);
            const int b = b0 + (c >> 1);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int ir = mt * 16 + (lane >> 4) * 4 + r;
                double vv = 0.0;
#pragma unroll
                for (int q = NS - 1; q >= 0; --q) vv = fma(vv, kInvRadix, (double)acc[q][h][r]);
                if (ir < Mv && b < a.nmem) {
                    const int l = m + 2 * ir + p;
                    a.T[(((size_t)m * n + l) * a.nmem + b) * 2 + (c & 1)] =
                        ldexp(vv, erow[r] + ex[c]) * kInvTop2;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Epilogue + synthesis digits for ME members of column m.
// ---------------------------------------------------------------------------
__host__ __device__ inline size_t gemm_epi_lds(int L, int me) {
    const size_t n = L + 1;
    // state, psi [ME][n] complex; C+, C-, lap [n]; exponents [8 ME]; the
    // column-maximum scratch of col_max_exp needs none
    return sizeof(double2) * 2 * me * n + sizeof(double) * (3 * n + (n & 1)) + 256;
}

template <int NS, int ME>
__global__ __launch_bounds__(kGemmThreads) void gemm_epi_kernel(GemmArgs a) {
    constexpr int NF = 4, NCE = ME * NF * 2;
    extern __shared__ __align__(16) unsigned char gsm[];
    const int m = blockIdx.x, b0 = blockIdx.y * ME;
    const int L = a.L, n = L + 1, K = L - m + 1;
    double2* st = reinterpret_cast<double2*>(gsm);            // [ME][n]
    double2* psi = st + (size_t)ME * n;                       // [ME][K]
    double* cpc = reinterpret_cast<double*>(psi + (size_t)ME * n);
    double* cmc = cpc + n;
    double* lapc = cmc + n;
    int* fexp = reinterpret_cast<int*>(lapc + n + (n & 1));   // [NCE]
    ColK kc = a.k;
    kc.cpc = cpc;
    kc.cmc = cmc;
    kc.lap = lapc;

    for (int t = threadIdx.x; t < ME * n; t += kGemmThreads) {
        const int bb = t / n, l = t - bb * n;
        const int b = b0 + bb;
        st[t] = (b < a.nmem) ? a.ys[b * a.ys_stride + (size_t)m * n + l] : c2(0.0, 0.0);
    }
    for (int t = threadIdx.x; t < n; t += kGemmThreads) {
        cpc[t] = a.k.Cp[(size_t)m * n + t];
        cmc[t] = a.k.Cm[(size_t)m * n + t];
        lapc[t] = a.k.lap[t];
    }
    __syncthreads();
    if (a.ana) {
        // every (l, m) of the column, as the fused kernel's epilogue: analysed
        // values in the band, zero elsewhere (members fastest: T rows are
        // member-contiguous)
        const bool active = m <= a.mmax;
        for (int t = threadIdx.x; t < ME * n; t += kGemmThreads) {
            const int bb = t % ME, l = t / ME;
            const int b = b0 + bb;
            if (b >= a.nmem) continue;
            double2 acc = c2(0.0, 0.0);
            if (active && l >= m && l <= a.top) {
                const double* tp = a.T + (((size_t)m * n + l) * a.nmem + b) * 2;
                acc = c2(tp[0], tp[1]);
            }
            a.ep.apply(b, l, m, &acc, st + (size_t)bb * n, true, nullptr);
        }
        __syncthreads();
    }
    if (!a.syn) return;
    for (int t = threadIdx.x; t < ME * K; t += kGemmThreads) {
        const int bb = t / K, i = t - bb * K;
        psi[(size_t)bb * n + i] = KindBve::base(0, st + (size_t)bb * n, m + i, m, kc);
    }
    __syncthreads();
    // column c = (member * 4 + field) * 2 + re/im, entry i = l - m: the fields
    // of KindBve as derive_columns forms them
    auto val = [&](int c, int i) -> double {
        const int bb = c >> 3, f = (c >> 1) & 3, ri = c & 1;
        const double2* sb = st + (size_t)bb * n;
        const double2* pb = psi + (size_t)bb * n;
        auto base = [&](int ii) -> double2 {
            return (f < 2) ? pb[ii] : KindBve::base(2, sb, m + ii, m, kc);
        };
        double2 x;
        if (!(f & 1)) {
            x = base(i);
        } else {
            const int l = m + i;
            x = c2(0.0, 0.0);
            if (i >= 1) x = cfma(base(i - 1), cpc[l - 1], x);
            if (l <= L - 1) x = cfma(base(i + 1), cmc[l + 1], x);
        }
        return cpart(x, ri);
    };
    col_max_exp(NCE, K, fexp, [&](int c, int i) { return val(c, i); });
    __syncthreads();
    const int kpad = split_kpad(L, m);
    const size_t qbase = 2 * (size_t)NS * a.ncp * a.kpoff[m];
    const int cg0 = b0 * 8;
    for (int t = threadIdx.x; t < NCE; t += kGemmThreads)
        if (b0 + (t >> 3) < a.nmem) a.Qe[(size_t)m * a.ncp + cg0 + t] = fexp[t];
    // items (parity, column, 16 degrees of the parity): one 16-byte store per slice
    const int nib = kpad / 16;
    for (int t = threadIdx.x; t < 2 * NCE * nib; t += kGemmThreads) {
        const int ib = t % nib, c = (t / nib) % NCE, p = t / (nib * NCE);
        if (b0 + (c >> 3) >= a.nmem) continue;
        const int e = fexp[c];
        unsigned int w[NS][4];
#pragma unroll
        for (int q = 0; q < NS; ++q) w[q][0] = w[q][1] = w[q][2] = w[q][3] = 0u;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int i = 2 * (ib * 16 + u) + p;
            const double x = (i < K) ? val(c, i) : 0.0;
            int d[NS];
            digits<NS>(x, e, d);
#pragma unroll
            for (int q = 0; q < NS; ++q) w[q][u >> 2] |= (unsigned int)(d[q] & 0xff) << (8 * (u & 3));
        }
#pragma unroll
        for (int q = 0; q < NS; ++q)
            *reinterpret_cast<uint4*>(a.Qd + qbase +
                                      ((size_t)(p * NS + q) * a.ncp + cg0 + c) * kpad + ib * 16) =
                make_uint4(w[q][0], w[q][1], w[q][2], w[q][3]);
    }
}

// ---------------------------------------------------------------------------
// Synthesis: E or O = Ps_digits(m, parity) x Qd, CG columns per workgroup.
// ---------------------------------------------------------------------------
template <int NS, int CG>
__host__ __device__ inline size_t gemm_syn_lds(int L) {
    return sizeof(int) * CG + (size_t)NS * CG * (split_kpad(L, 0) + 16);
}

template <int NS, int CG, int NTW>
__global__ __launch_bounds__(kGemmThreads) void gemm_syn_kernel(GaussDev g, GemmArgs a) {
    constexpr int NT = CG / 16, NG = NT / NTW;   // a wave tile is 16 latitudes x NTW N-tiles
    extern __shared__ __align__(16) unsigned char gsm[];
    const int m = blockIdx.x, p = blockIdx.y, c0 = blockIdx.z * CG;
    const int L = g.L, nh = g.nh, n = L + 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int kpad = split_kpad(L, m), KS = kpad + 16, nhp = rup(nh, 16);
    int* qe = reinterpret_cast<int*>(gsm);                                   // [CG]
    signed char* qs = reinterpret_cast<signed char*>(gsm + sizeof(int) * CG);  // [NS][CG][KS]
    const signed char* src = a.Qd + 2 * (size_t)NS * a.ncp * a.kpoff[m];
    const int nib = kpad / 16;
    for (int t = threadIdx.x; t < NS * CG * nib; t += kGemmThreads) {
        const int ib = t % nib, c = (t / nib) % CG, q = t / (nib * CG);
        *reinterpret_cast<uint4*>(qs + ((size_t)q * CG + c) * KS + ib * 16) =
            *reinterpret_cast<const uint4*>(src + ((size_t)(p * NS + q) * a.ncp + c0 + c) * kpad +
                                            ib * 16);
    }
    for (int t = threadIdx.x; t < CG; t += kGemmThreads) qe[t] = a.Qe[(size_t)m * a.ncp + c0 + t];
    __syncthreads();

    const signed char* Pt = g.Psl + g.offs_ss[m];
    const int* er = g.esyn + (size_t)m * 2 * nhp + p * nhp;
    const int kfull = kpad & ~(kSplitK - 1);
    double2* dst = p ? a.O : a.E;
    const int T = (nhp / 16) * NG;
    for (int tt = wave; tt < T; tt += kGemmWaves) {
        const int jt = tt / NG, ng = tt - jt * NG;
        const int j0 = jt * 16;
        intx4 acc[NS][NTW];
#pragma unroll
        for (int q = 0; q < NS; ++q)
#pragma unroll
            for (int h = 0; h < NTW; ++h) acc[q][h] = intx4{0, 0, 0, 0};
        for (int k0 = 0; k0 < kfull; k0 += kSplitK) {
            intx4 av[NS];
#pragma unroll
            for (int q = 0; q < NS; ++q)
                av[q] = *reinterpret_cast<const intx4*>(
                    Pt + ((size_t)(p * NS + q) * nhp + j0 + (lane & 15)) * kpad + k0 +
                    (lane >> 4) * 16);
#pragma unroll
            for (int h = 0; h < NTW; ++h)
#pragma unroll
                for (int t2 = 0; t2 < NS; ++t2) {
                    const intx4 bv = *reinterpret_cast<const intx4*>(
                        qs + ((size_t)t2 * CG + (NTW * ng + h) * 16 + (lane & 15)) * KS + k0 +
                        (lane >> 4) * 16);
#pragma unroll
                    for (int q = 0; q + t2 < NS; ++q)
                        acc[q + t2][h] =
                            __builtin_amdgcn_mfma_i32_16x16x64_i8(av[q], bv, acc[q + t2][h], 0, 0, 0);
                }
        }
        if (kpad != kfull) {   // the 32-deep tail at kfull = kpad - 32
            long at[NS];
#pragma unroll
            for (int q = 0; q < NS; ++q)
                at[q] = *reinterpret_cast<const long*>(
                    Pt + ((size_t)(p * NS + q) * nhp + j0 + (lane & 15)) * kpad + kfull +
                    (lane >> 4) * 8);
#pragma unroll
            for (int h = 0; h < NTW; ++h)
#pragma unroll
                for (int t2 = 0; t2 < NS; ++t2) {
                    const long bv = *reinterpret_cast<const long*>(
                        qs + ((size_t)t2 * CG + (NTW * ng + h) * 16 + (lane & 15)) * KS + kfull +
                        (lane >> 4) * 8);
#pragma unroll
                    for (int q = 0; q + t2 < NS; ++q)
                        acc[q + t2][h] =
                            __builtin_amdgcn_mfma_i32_16x16x32_i8(at[q], bv, acc[q + t2][h], 0, 0, 0);
                }
        }
        int erow[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) erow[r] = er[j0 + (lane >> 4) * 4 + r];
#pragma unroll
        for (int h = 0; h < NTW; ++h) {
            const int c = (NTW * ng + h) * 16 + (lane & 15);
            const int cg = c0 + c;
            const int b = cg >> 3, f = (cg >> 1) & 3;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = j0 + (lane >> 4) * 4 + r;
                double vv = 0.0;
#pragma unroll
                for (int q = NS - 1; q >= 0; --q) vv = fma(vv, kInvRadix, (double)acc[q][h][r]);
                vv = ldexp(vv, erow[r] + qe[c]) * kInvTop2;
                const double w = __shfl_xor(vv, 1);
                if (!(c & 1) && j < nh && b < a.nmem)
                    dst[(((size_t)j * n + m) * a.nmem + b) * 4 + f] = c2(vv, w);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static int gemm_ncp(int nmem) { return rup(nmem * 8, kSynColAlign); }

bool gemm_stage_supported(const aeo_model_s* md) {
    const aeo_plan_s* p = md->plan;
    return md->kind == kBVE && p->kind == kGauss && split_nhk(p->nh) <= 256 &&
           gemm_epi_lds(p->L, 8) <= 163840 && gemm_syn_lds<7, 128>(p->L) <= 163840 &&
           gemm_ana_lds<7>(p->nh) <= 163840;
}

int gemm_stage_setup(aeo_model_s* md) {
    if (md->d_gT) return AEO_OK;
    AEO_REQUIRE(gemm_stage_supported(md), AEO_ERR_UNSUPPORTED,
                "GEMM-form stage: needs a BVE model on a Gauss plan of at most 256 latitude "
                "pairs");
    const aeo_plan_s* p = md->plan;
    const int L = p->L, n = L + 1;
    std::vector<int> koff(n + 1, 0);
    for (int m = 0; m < n; ++m) koff[m + 1] = koff[m] + split_kpad(L, m);
    const int ncp = gemm_ncp(md->nmem);
    const size_t tbytes = sizeof(double2) * (size_t)n * n * md->nmem;
    const size_t qbytes = 2 * (size_t)7 * ncp * koff[n];   // up to 7 slices
    const size_t ebytes = sizeof(int) * (size_t)n * ncp;
    double* T = nullptr;
    signed char* Q = nullptr;
    int* Qe = nullptr;
    int* ko = nullptr;
    hipError_t e = hipMalloc(&T, tbytes);
    if (e == hipSuccess) e = hipMalloc(&Q, qbytes);
    if (e == hipSuccess) e = hipMalloc(&Qe, ebytes);
    if (e == hipSuccess) e = hipMalloc(&ko, sizeof(int) * (n + 1));
    // digit columns of padding members stay zero
    if (e == hipSuccess) e = hipMemset(Q, 0, qbytes);
    if (e == hipSuccess) e = hipMemset(Qe, 0, ebytes);
    if (e == hipSuccess)
        e = hipMemcpy(ko, koff.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        for (void* q : {(void*)T, (void*)Q, (void*)Qe, (void*)ko})
            if (q) (void)hipFree(q);
        set_error(std::string("GEMM-form stage workspaces: ") + hipGetErrorString(e));
        return AEO_ERR_HIP;
    }
    md->d_gT = T;
    md->d_gQ = Q;
    md->d_gQe = Qe;
    md->d_gkoff = ko;
    md->gemm_ncp = ncp;
    return AEO_OK;
}

template <class K>
static int allow_gemm_lds(K kern, size_t bytes) {
    AEO_REQUIRE(bytes <= 163840, AEO_ERR_UNSUPPORTED, "GEMM-form stage: LDS demand exceeds 160 KiB");
    if (bytes > 65536)
        AEO_CHECK_HIP(hipFuncSetAttribute((const void*)kern,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return AEO_OK;
}

template <int NS, int ME>
static int gemm_epi_launch(const aeo_plan_s* p, const GemmArgs& a, hipStream_t s) {
    const size_t lds = gemm_epi_lds(p->L, ME);
    int rc;
    if ((rc = allow_gemm_lds(gemm_epi_kernel<NS, ME>, lds))) return rc;
    launch_k(gemm_epi_kernel<NS, ME>, dim3(p->L + 1, (a.nmem + ME - 1) / ME), dim3(kGemmThreads),
             lds, s, a);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

template <int NS, int CG, int NTW>
static int gemm_syn_launch(const aeo_plan_s* p, const GemmArgs& a, hipStream_t s) {
    const size_t lds = gemm_syn_lds<NS, CG>(p->L);
    int rc;
    if ((rc = allow_gemm_lds(gemm_syn_kernel<NS, CG, NTW>, lds))) return rc;
    launch_k(gemm_syn_kernel<NS, CG, NTW>, dim3(p->L + 1, 2, a.ncp / CG), dim3(kGemmThreads), lds,
             s, p->gdev(), a);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

template <int NS>
static int gemm_stage_ns(aeo_model_s* md, const GemmArgs& a, hipStream_t s) {
    const aeo_plan_s* p = md->plan;
    const GaussDev g = p->gdev();
    int rc;
    if (a.ana && a.mmax >= 0) {
        const size_t lds = gemm_ana_lds<NS>(p->nh);
        if ((rc = allow_gemm_lds(gemm_ana_kernel<NS>, lds))) return rc;
        launch_k(gemm_ana_kernel<NS>, dim3(a.mmax + 1, 2, (a.nmem + kAnaMembers - 1) / kAnaMembers),
                 dim3(kGemmThreads), lds, s, g, a);
        AEO_LAUNCH_CHECK();
    }
    rc = (md->gemm_me == 4) ? gemm_epi_launch<NS, 4>(p, a, s) : gemm_epi_launch<NS, 8>(p, a, s);
    if (rc || !a.syn) return rc;
    if (md->gemm_ntw == 2)
        return (md->gemm_cg == 128) ? gemm_syn_launch<NS, 128, 2>(p, a, s)
                                    : gemm_syn_launch<NS, 64, 2>(p, a, s);
    return (md->gemm_cg == 128) ? gemm_syn_launch<NS, 128, 4>(p, a, s)
                                : gemm_syn_launch<NS, 64, 4>(p, a, s);
}

int gemm_stage_tiles(aeo_model_s* md, int syn_cols, int epi_members, int syn_ntiles) {
    AEO_REQUIRE(syn_cols == 64 || syn_cols == 128, AEO_ERR_INVALID,
                "aeo_model_set_gemm_tiles: syn_cols must be 64 or 128");
    AEO_REQUIRE(epi_members == 4 || epi_members == 8, AEO_ERR_INVALID,
                "aeo_model_set_gemm_tiles: epi_members must be 4 or 8");
    AEO_REQUIRE(syn_ntiles == 2 || syn_ntiles == 4, AEO_ERR_INVALID,
                "aeo_model_set_gemm_tiles: syn_ntiles must be 2 or 4");
    md->gemm_cg = syn_cols;
    md->gemm_me = epi_members;
    md->gemm_ntw = syn_ntiles;
    return AEO_OK;
}

int gemm_stage_launch(aeo_model_s* md, const double2* ys, const double2* y0,
                      const double2* forcing, double2* kout, int stage, double dt,
                      double2* ynext, bool ana, bool syn, hipStream_t s) {
    const aeo_plan_s* p = md->plan;
    AEO_REQUIRE(md->d_gT != nullptr, AEO_ERR_STATE,
                "GEMM-form stage: workspaces missing (select the tiles with aeo_model_set_tiles)");
    AEO_REQUIRE(md->nslices > 0 && p->nslices == md->nslices, AEO_ERR_STATE,
                "GEMM-form stage: needs split precision with the plan's digit tables");
    const int L = p->L;
    const size_t fs = (size_t)(L + 1) * (L + 1);
    const double f10 = 2.0 * md->omega * sqrt(4.0 * M_PI / 3.0);
    GemmArgs a{};
    a.L = L;
    a.nh = p->nh;
    a.nmem = md->nmem;
    a.top = (2 * L) / 3;
    a.mmax = (2 * L) / 3;
    a.ncp = md->gemm_ncp;
    a.S = md->d_S;
    a.D = md->d_D;
    a.T = md->d_gT;
    a.Qd = md->d_gQ;
    a.Qe = md->d_gQe;
    a.kpoff = md->d_gkoff;
    a.E = md->d_E;
    a.O = md->d_O;
    a.ys = ys;
    a.ys_stride = fs * md->nstate;
    a.ana = ana ? 1 : 0;
    a.syn = syn ? 1 : 0;
    a.k = col_consts(p, md->R, md->d_lap, f10);
    Rk4B rk{stage, stage == 2 ? dt : 0.5 * dt, dt / 6.0, y0, md->d_ksum, ynext, kout};
    a.ep = EpiBveS{L, fs, forcing, md->d_nulap, rk};
    switch (md->nslices) {
        case 4: return gemm_stage_ns<4>(md, a, s);
        case 5: return gemm_stage_ns<5>(md, a, s);
        case 6: return gemm_stage_ns<6>(md, a, s);
        case 7: return gemm_stage_ns<7>(md, a, s);
        default:
            set_error("GEMM-form stage: nslices must be 4, 5, 6 or 7");
            return AEO_ERR_UNSUPPORTED;
    }
}

}  // namespace aeo
