// Single-member FP64 barotropic-vorticity stage on Gauss grids: the
// latency-shaped case (one state, so one workgroup per zonal wavenumber m is
// all the parallelism the Legendre stage has).
//
// It computes what leg_stage_kernel<KindBve, EpiBveS, 1, ANA, SYN, 0>
// computes, to rounding (the per-degree analysis sums associated per
// half-wave and wave, the synthesis sums of a split column per degree
// half), written for a short critical path:
//   * a small argument block and 32-bit indexing (no scalar reloads of a
//     large kernel-argument struct, no 64-bit address arithmetic per load);
//   * every global load that does not depend on this launch's results --
//     the analysis table rows, the state column, S/D, the coupling columns,
//     the recurrence column, the synthesis seeds, the RK4 inputs y0 / ksum /
//     forcing -- is issued at the start, before the first wait;
//   * the heavy columns (small m) get two workgroups that split the degrees
//     of the column at an even s: each analyses its own degrees (plus the
//     neighbour degree the S operator needs), runs the epilogue, writes the
//     RK4 outputs of its degrees and synthesises their even/odd Legendre sums
//     over all latitudes into its own E/O plane (the second half into plane
//     1; the grid stage adds the planes in the fixed order plane 0 + plane 1).
//     No analysis-table row is loaded twice and no workgroup waits on another;
//   * analysis lanes cover degrees l + 32 s (equal parity, so one LDS read
//     of S or D serves all of them) and the two half-waves split the
//     latitudes; in the synthesis each half-wave takes two of the
//     four fields and each lane four latitudes j + 32 q (one LDS read of a
//     coefficient serves four recurrences), and the recurrence pairs come
//     through vector buffer loads: the per-degree LDS traffic, which bounds
//     both loops, is cut from five wave-wide reads to two.
// Per workgroup: analysis of column m (lanes over degree, waves over
// latitude, reduced over waves in LDS), tendency -J + nu lap' zeta (+ forcing)
// with row 0 pinned and the RK4 stage update, then the four synthesis fields
// (psi, S psi, eta, S eta) of the new stage state and their even/odd Legendre
// sums E/O from the Pbar recurrence seeded by two table rows per degree block
// (phy/barotropic.py:120-167, run/bve/runner.py:211-216,
// num/spectral_operators.py:228-241).
#include "legendre.hpp"

namespace aeo {

constexpr int kB1Threads = 512;
constexpr int kB1Waves = kB1Threads / 64;
constexpr int kB1AG = 2;                  // analysis latitude groups per wave
constexpr int kB1AW = 64 / kB1AG;         // degrees per analysis slot (lanes per group)
constexpr int kB1Slots = 128 / kB1AW;     // analysis degree slots (128 degrees at most)
constexpr int kB1Pre = 16 / kB1AG;        // analysis-table rows prefetched per lane and slot
constexpr int kB1PalQ = kB1Pre / 2;       // 16-byte lane-order entries per lane and slot
constexpr int kB1PassRows = kB1Waves * kB1AG * kB1Pre;   // latitudes per analysis pass
static_assert(kB1AW == 32 && kB1PalQ == 4 && kB1PassRows == 128, "analysis lane layout");
constexpr int kB1SynWaves = 4;   // synthesis degree blocks, one wave each
constexpr int kB1SplitQ8 = 93;   // split point of a heavy column: s ~ K q8 / 256


struct Bve1Args {
    int L, nh, top;           // top: analysed degrees l <= top (2L/3)
    int msplit;               // columns m < msplit run as two degree halves
    int split_q8;             // their split point: i = l - m < s = (K q8 / 256) & ~1
    int ana, syn;             // analysis + epilogue; synthesis of the (new) state
    const double2* Pal;       // analysis-table blocks in lane order (bve1_pal_kernel)
    int npass;                // analysis passes of kB1PassRows latitudes
    const double* Ps;         // synthesis table, block m: [l - m][jp]
    const double* xn;         // nh northern nodes
    const double2* rec;       // [m][l] recurrence pairs
    const double* CpT;        // [m][l]
    const double* CmT;        // [m][l]
    const double* lap;        // -l(l+1)/R^2, l = 0 entry 0
    const double* nulap;      // nu lap'_l
    double f10;               // f at (1, 0)
    const double2* S;         // [jp][m]; hemi: FN [m][jp], the north row's unweighted spectrum
    const double2* D;         // hemi: FS [m][jp]
    const double* wrow;       // hemi: nlat row weights (S = w (FN + FS), D = w (FN - FS))
                              // hemi also writes the synthesis as row sums (E + O, E - O)
    int hemi, nlat;
    double2* E;               // plane 0: [jp][m][4]; hemi: the north row sums E + O
    double2* O;               // hemi: the south row sums E - O
    double2* E1;              // plane 1 (degree half 1 of the split columns): [jp][m][4]
    double2* O1;
    const double2* ys;        // stage input state, internal [m][l]
    const double2* y0;        // step input state
    double2* ksum;
    double2* ynext;
    double2* kout;            // stage < 0: tendency output
    const double2* forcing;   // or nullptr
    int stage;
    double cst, dt6;
};

// Degree split point of a heavy column with K degrees: even, 2 <= s <= K - 2,
// about K q8 / 256 but at most 62, so the first half analyses degrees 0..s
// (the neighbour included) in two 32-degree slots (first chosen for one
// 64-lane slot: a second then doubled its table loads, 0.0803 -> 0.0821 ms
// per L = 170 step; re-swept with 32-degree slots and the lane-order table,
// 93 / 62 stayed the fastest).
__host__ __device__ inline int bve1_split_point(int K, int q8) {
    int s = ((K * q8) >> 8) & ~1;
    if (s > 62) s = 62;
    const int hi = (K - 2) & ~1;
    return s < 2 ? 2 : (s > hi ? hi : s);
}

// Workgroup b of a launch: column m and degree half h (-1: whole column;
// the split columns first, heaviest first); its degrees i = l - m in [d0, d1)
// (RK4 outputs and synthesis), the state needed on [e0, e1) (the S
// operator's neighbours) and the analysed degrees [e0, a1].
struct Bve1Wg {
    int m, h, K, d0, d1, e0, e1, a1;
};
__host__ __device__ inline Bve1Wg bve1_wg(int b, int L, int msplit, int q8, int top) {
    Bve1Wg w;
    if (b < 2 * msplit) {
        w.m = b >> 1;
        w.h = b & 1;
    } else {
        w.m = b - msplit;
        w.h = -1;
    }
    w.K = L - w.m + 1;
    w.d0 = 0;
    w.d1 = w.K;
    if (w.h >= 0) {
        const int s = bve1_split_point(w.K, q8);
        if (w.h == 0) w.d1 = s; else w.d0 = s;
    }
    w.e0 = (w.d0 > 0) ? w.d0 - 1 : 0;
    w.e1 = (w.d1 < w.K) ? w.d1 + 1 : w.K;
    w.a1 = (w.e1 - 1 < top - w.m) ? w.e1 - 1 : top - w.m;
    return w;
}

// The analysis-table rows of each workgroup in the order its lanes read
// them: entry ((p kB1Slots + s) kB1PalQ + q) kB1Threads + t of workgroup b
// holds, for lane t % 64 of wave t / 64 (degree slot ad = lane % 32,
// latitude group ag = lane / 32), degree i = e0 + ad + 32 s and the rows
// jp = 128 p + wave + 8 (2 u + ag) of u = 2 q (x) and 2 q + 1 (y); zero past
// a1 or nh.  Every load of the stage kernel's analysis is then one
// contiguous kilobyte per wave (16 bytes a lane) where the [jp][i] blocks
// gave two 256-byte runs of 8-byte lanes: a 32 MiB read of the two
// patterns took 2.8 and 1.2 us.
__global__ __launch_bounds__(kB1Threads) void bve1_pal_kernel(const double* __restrict__ Pa,
                                                               double2* __restrict__ out, int L,
                                                               int nh, int msplit, int q8, int top,
                                                               int npass) {
    const int n = L + 1, b = blockIdx.x, t = threadIdx.x;
    const Bve1Wg w = bve1_wg(b, L, msplit, q8, top);
    const long long boff = (long long)nh * ((long long)w.m * n - (long long)w.m * (w.m - 1) / 2);
    const int lane = t & 63, wave = t >> 6, ad = lane % kB1AW, ag = lane / kB1AW;
    double2* o = out + (size_t)b * npass * kB1Slots * kB1PalQ * kB1Threads + t;
    for (int p = 0; p < npass; ++p)
        for (int s = 0; s < kB1Slots; ++s) {
            const int i = w.e0 + ad + kB1AW * s;
            for (int q = 0; q < kB1PalQ; ++q) {
                double v[2];
                for (int e = 0; e < 2; ++e) {
                    const int u = 2 * q + e, jp = kB1PassRows * p + wave + 8 * (kB1AG * u + ag);
                    v[e] = (i <= w.a1 && jp < nh) ? Pa[boff + (long long)jp * w.K + i] : 0.0;
                }
                o[((p * kB1Slots + s) * kB1PalQ + q) * kB1Threads] = c2(v[0], v[1]);
            }
        }
}

// kR: latitudes per synthesis lane (g + 32 q, q < kR, g = lane % 32): 4 when
// nh > 64.  recg: the [m][l] recurrence pairs (a.rec) as a restrict pointer,
// so the synthesis reads them with wave-uniform scalar loads.
template <int kR>
__global__ __launch_bounds__(kB1Threads) void bve1_stage_kernel(Bve1Args a,
                                                                const double2* __restrict__ recg) {
    extern __shared__ double2 lds[];
    // the arguments into scalar registers at entry, in one batch: loaded
    // where first used, they came in four rounds, each waited for before the
    // next (the analysis-table loads waited for three, the column loads for
    // four)
    asm volatile("" ::"s"(a.L), "s"(a.nh), "s"(a.top), "s"(a.msplit), "s"(a.split_q8),
                 "s"(a.ana), "s"(a.syn), "s"(a.Pal), "s"(a.npass), "s"(a.Ps), "s"(a.xn),
                 "s"(a.CpT), "s"(a.CmT), "s"(a.lap), "s"(a.nulap));
    asm volatile("" ::"s"(a.S), "s"(a.D), "s"(a.wrow), "s"(a.hemi), "s"(a.nlat), "s"(a.ys),
                 "s"(a.y0), "s"(a.ksum), "s"(a.forcing), "s"(a.stage), "s"(recg));
    const int L = a.L, n = L + 1, nh = a.nh;
    // wave index as a wave-uniform (scalar) value: block bounds and loop
    // counters derived from it stay in scalar registers
    const int tid = threadIdx.x, lane = tid & 63,
              wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    // workgroup -> (m, degree half) and its degree ranges (bve1_wg)
    const Bve1Wg wg = bve1_wg(blockIdx.x, L, a.msplit, a.split_q8, a.top);
    const int m = wg.m, h = wg.h, K = wg.K, d0 = wg.d0, d1 = wg.d1, e0 = wg.e0, e1 = wg.e1,
              a1 = wg.a1;
    const bool active = a.ana && e0 <= a1;
    const int nsl = active ? (a1 - e0) / kB1AW + 1 : 0;   // degree slots in use (uniform)
    const long long boff = (long long)nh * ((long long)m * n - (long long)m * (m - 1) / 2);
    const int col = m * n;                       // internal [m][l] column offset

    // LDS carve-up: persistent columns, then one scratch region used by the
    // analysis (S/D, wave partials) and later by the synthesis block sums
    double2* st = lds;                           // [n] state column (stage input, then next state)
    double2* cols = st + n;                      // [4][K] synthesis fields
    double* cpc = reinterpret_cast<double*>(cols + 4 * n);   // [n] C+ of m
    double* cmc = cpc + n;                       // [n] C- of m
    double* lapc = cmc + n;                      // [n] lap
    double2* scratch = reinterpret_cast<double2*>(lapc + n + (n & 1));
    // S then D, planes nh + 1 apart: the S and D reads of one latitude (lanes
    // of both parities in one instruction) fall in different LDS banks
    const int nhx = nh + 1;
    double2* X = scratch;                        // [2][nh + 1]
    double2* red = X + 2 * nhx;                  // [8][slots][AW] analysis wave partials
    double2* sred = scratch;                     // [4 blocks][2 parity][2][kR][64] block sums

    // ---- prologue: all independent global loads first
    // analysis lanes: degree slot d = lane % AW (degrees e0 + d + AW s, s <
    // nsl, all of one parity), latitude group ag = lane / AW (rows
    // wave + 8 (AG u + ag) of each pass of 128): one LDS read of S or D
    // serves up to 128 / AW degrees of a lane
    const int ad = lane % kB1AW, ag = lane / kB1AW;
    double pa[kB1Slots][kB1Pre];
    const int par = (e0 + ad) & 1;               // parity of every degree of this lane
    // this workgroup's table rows in lane order (bve1_pal_kernel): load
    // (s, q) of a lane is one 16-byte entry, rows u = 2 q and 2 q + 1
    const double2* PL =
        a.Pal + (size_t)blockIdx.x * a.npass * kB1Slots * kB1PalQ * kB1Threads + tid;
    if (active) {
#pragma unroll
        for (int s = 0; s < kB1Slots; ++s) {
#pragma unroll
            for (int q = 0; q < kB1PalQ; ++q) {
                double2 v = c2(0.0, 0.0);
                if (s < nsl && 32 * q < nh) v = PL[(s * kB1PalQ + q) * kB1Threads];
                pa[s][2 * q] = v.x;
                pa[s][2 * q + 1] = v.y;
            }
        }
    }
    // synthesis seeds of this lane's first latitude pass
    const int kb = wave;
    const int Kd = d1 - d0;
    const int Kb = ((Kd + kB1SynWaves - 1) / kB1SynWaves + 1) & ~1;
    const int i0 = d0 + kb * Kb, i1 = (i0 + Kb < d1) ? i0 + Kb : d1;
    const bool syn_wave = a.syn && wave < kB1SynWaves;
    const int g = lane & 31;                     // synthesis latitude group of the lane
    double sx[kR], sp0[kR], sp1[kR];
#pragma unroll
    for (int q = 0; q < kR; ++q) {
        sx[q] = sp0[q] = sp1[q] = 0.0;
        const int jp = g + 32 * q;
        if (syn_wave && jp < nh && i0 < i1) {
            const double* P = a.Ps + boff + jp;
            sx[q] = a.xn[jp];
            sp0[q] = P[i0 * nh];
            if (i0 + 1 < i1) sp1[q] = P[(i0 + 1) * nh];
        }
    }
    // column data (thread t <-> degree l = t) and S/D (thread t <-> latitude t)
    double2 vs = c2(0.0, 0.0), vy0 = c2(0.0, 0.0), vks = c2(0.0, 0.0);
    double2 vf = c2(0.0, 0.0), vS = c2(0.0, 0.0), vD = c2(0.0, 0.0);
    double vcp = 0.0, vcm = 0.0, vlap = 0.0, vnl = 0.0;
    const int tl = (tid < n) ? tid : n - 1;
    vs = a.ys[col + tl];
    vcp = a.CpT[col + tl];
    vcm = a.CmT[col + tl];
    vlap = a.lap[tl];
    if (a.ana) {
        vnl = a.nulap[tl];
        if (a.stage >= 0) vy0 = a.y0[col + tl];
        if (a.stage >= 1) vks = a.ksum[col + tl];
        if (a.forcing) vf = a.forcing[col + tl];
    }
    double vw = 0.0;
    if (active) {
        const int tj = (tid < nh) ? tid : nh - 1;
        const int sdo = a.hemi ? m * nh + tj : tj * n + m;
        vS = a.S[sdo];
        vD = a.D[sdo];
        if (a.hemi) vw = (2 * tj + 1 == a.nlat) ? 0.5 * a.wrow[tj] : a.wrow[tj];
    }
    // stores
    if (tid < n) {
        st[tid] = vs;
        cpc[tid] = vcp;
        cmc[tid] = vcm;
        lapc[tid] = vlap;
    }
    if (active && tid < nh) {
        if (a.hemi) {
            const double2 fn = vS, fs = vD;
            vS = cscale(cadd(fn, fs), vw);
            vD = cscale(csub(fn, fs), vw);
        }
        X[tid] = vS;
        X[nhx + tid] = vD;
    }
    for (int t = tid + kB1Threads; t < n; t += kB1Threads) {    // n > 512 (L > 511)
        st[t] = a.ys[col + t];
        cpc[t] = a.CpT[col + t];
        cmc[t] = a.CmT[col + t];
        lapc[t] = a.lap[t];
    }
    if (active)
        for (int t = tid + kB1Threads; t < nh; t += kB1Threads) {
            const int sdo = a.hemi ? m * nh + t : t * n + m;
            double2 s = a.S[sdo], d = a.D[sdo];
            if (a.hemi) {
                const double w = (2 * t + 1 == a.nlat) ? 0.5 * a.wrow[t] : a.wrow[t];
                const double2 fn = s, fs = d;
                s = cscale(cadd(fn, fs), w);
                d = cscale(csub(fn, fs), w);
            }
            X[t] = s;
            X[nhx + t] = d;
        }
    __syncthreads();

    // ---- analysis: lanes over degrees e0 + d + AW s, lane groups and waves
    // over latitude; the groups' sums meet by lane swaps, then the waves'
    // sums in wave order
    if (active) {
        double2 acc[kB1Slots];
#pragma unroll
        for (int s = 0; s < kB1Slots; ++s) acc[s] = c2(0.0, 0.0);
        for (int j0 = 0; j0 < nh; j0 += kB1PassRows) {
            double pv[kB1Slots][kB1Pre];
            if (j0 == 0) {
#pragma unroll
                for (int s = 0; s < kB1Slots; ++s)
#pragma unroll
                    for (int u = 0; u < kB1Pre; ++u) pv[s][u] = pa[s][u];
            } else {                                      // nh > 128 (L > 170)
                const double2* PP = PL + (j0 / kB1PassRows) * kB1Slots * kB1PalQ * kB1Threads;
#pragma unroll
                for (int s = 0; s < kB1Slots; ++s) {
#pragma unroll
                    for (int q = 0; q < kB1PalQ; ++q) {
                        double2 v = c2(0.0, 0.0);
                        if (s < nsl && j0 + 32 * q < nh) v = PP[(s * kB1PalQ + q) * kB1Threads];
                        pv[s][2 * q] = v.x;
                        pv[s][2 * q + 1] = v.y;
                    }
                }
            }
            // rows past nh have pv = 0 and read row nh - 1 (branch-free);
            // four LDS reads in flight per group
#pragma unroll
            for (int u0 = 0; u0 < kB1Pre; u0 += 4) {
                double2 x[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int jq = j0 + wave + 8 * (kB1AG * (u0 + u) + ag);
                    x[u] = X[par * nhx + ((jq < nh) ? jq : nh - 1)];
                }
#pragma unroll
                for (int s = 0; s < kB1Slots; ++s) {
                    if (s < nsl) {
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            acc[s].x = fma(pv[s][u0 + u], x[u].x, acc[s].x);
                            acc[s].y = fma(pv[s][u0 + u], x[u].y, acc[s].y);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int s = 0; s < kB1Slots; ++s) {
            acc[s].x = sum_xor32(acc[s].x);      // the two latitude groups (kB1AW = 32)
            acc[s].y = sum_xor32(acc[s].y);
            if (ag == 0) red[(wave * kB1Slots + s) * kB1AW + ad] = acc[s];
        }
        __syncthreads();
    }
    // the analysed value of degree l: the waves' partial sums in wave order,
    // summed by the epilogue thread of l (0 for degrees not analysed here)
    auto analysed = [&](int l) {
        const int i = l - m - e0;
        double2 sv = c2(0.0, 0.0);
        if (active && i >= 0 && l - m <= a1) {
            const int s = i / kB1AW, d = i % kB1AW;
            sv = red[s * kB1AW + d];
#pragma unroll
            for (int w = 1; w < kB1Waves; ++w) sv = cadd(sv, red[(w * kB1Slots + s) * kB1AW + d]);
        }
        return sv;
    };

    // synthesis fields psi = zeta / lap (0 at l = 0) and eta = zeta (+ f at
    // (1, 0)) of degree l on [e0, e1) ([f][i], i = l - m; S psi and S eta
    // follow once every degree is in)
    auto syn_fields = [&](int l, double2 z) {
        const int i = l - m;
        if (i >= e0 && i < e1) {
            double2 v = c2(0.0, 0.0);
            if (l > 0) {
                const double d = lapc[l];
                v = c2(z.x / d, z.y / d);
            }
            cols[i] = v;
            if (l == 1 && m == 0) z.x += a.f10;
            cols[2 * K + i] = z;
        }
    };

    // ---- epilogue: tendency and RK4 stage bookkeeping for degree l = tid;
    // the global RK4 outputs of the degrees this workgroup owns (the padding
    // rows l < m belong to the first half); stored right away (stores held
    // back to the end of the kernel delay its completion: measured slower)
    auto store_out = [&](int l, double2 kv, double2 yv) {
        const int idx = col + l;
        if (a.stage < 0) {
            a.kout[idx] = kv;
        } else {
            if (a.stage < 3) a.ksum[idx] = kv;
            a.ynext[idx] = yv;
        }
    };
    if (a.ana) {
        for (int l = tid; l < n; l += kB1Threads) {
            double2 y0v = vy0, ksv = vks, fv = vf;
            double nl = vnl;
            if (l >= kB1Threads) {                   // beyond the prologue's one degree per thread
                nl = a.nulap[l];
                y0v = (a.stage >= 0) ? a.y0[col + l] : c2(0.0, 0.0);
                ksv = (a.stage >= 1) ? a.ksum[col + l] : c2(0.0, 0.0);
                fv = a.forcing ? a.forcing[col + l] : c2(0.0, 0.0);
            }
            const int i = l - m;
            const bool writer = (h < 0) || (h == 0 ? i < d1 : i >= d0);
            double2 k = c2(0.0, 0.0);
            if (l > 0) {
                const double2 acc = analysed(l);
                k = cfma(st[l], nl, c2(-acc.x, -acc.y));
                if (a.forcing) k = cadd(k, fv);
            }
            double2 nx, kv = k;                  // kv: the kout or ksum value
            if (a.stage < 0) {
                nx = k;
            } else if (a.stage == 0) {
                nx = cfma(k, a.cst, y0v);
            } else if (a.stage < 3) {
                kv = cfma(k, 2.0, ksv);
                nx = cfma(k, a.cst, y0v);
            } else {
                const double2 s = cadd(ksv, k);
                nx = cfma(s, a.dt6, y0v);
            }
            if (writer) store_out(l, kv, nx);
            // psi and eta of the new state, straight into the synthesis columns
            if (a.syn) syn_fields(l, nx);
        }
    } else if (a.syn) {
        for (int l = m + e0 + tid; l < m + e1; l += kB1Threads) syn_fields(l, st[l]);
    }
    if (!a.syn) return;
    __syncthreads();
    for (int t = tid; t < 2 * Kd; t += kB1Threads) {
        const int src = (t < Kd) ? 0 : 2, i = d0 + ((t < Kd) ? t : t - Kd), l = m + i;
        const double2* s = cols + src * K;
        double2 v = c2(0.0, 0.0);
        if (i >= 1) v = cfma(s[i - 1], cpc[l - 1], v);
        if (l <= L - 1) v = cfma(s[i + 1], cmc[l + 1], v);
        cols[(src + 1) * K + i] = v;
    }
    __syncthreads();

    // ---- Pbar recurrence synthesis: wave = degree block kb of [d0, d1);
    // half-wave hf = lane / 32 takes fields 2 hf, 2 hf + 1 ((psi, S psi) or
    // (eta, S eta)), lane latitudes jbase + g + 32 q (q < kR)
    auto slot = [&](int g2, int par2, int f, int q) -> double2& {
        return sred[(((g2 * 2 + par2) * 2 + f) * kR + q) * 64 + lane];
    };
    double2* Ep = (h == 1) ? a.E1 : a.E;
    double2* Op = (h == 1) ? a.O1 : a.O;
    const int hf = lane >> 5;
    const double2* cfh = cols + 2 * hf * K;      // [2][K] this half-wave's fields
    const double2* rg = recg + col + m;          // recurrence pair of degree i at rg[i]
    for (int jbase = 0; jbase < nh; jbase += 32 * kR) {
        double2 accE[kR][2], accO[kR][2];
#pragma unroll
        for (int q = 0; q < kR; ++q)
#pragma unroll
            for (int f = 0; f < 2; ++f) accE[q][f] = accO[q][f] = c2(0.0, 0.0);
        if (syn_wave && i0 < i1) {
            const bool first = (jbase == 0);
            double x[kR], p1[kR], p2[kR];
#pragma unroll
            for (int q = 0; q < kR; ++q) {
                const int jp = jbase + g + 32 * q;
                if (first) {
                    x[q] = sx[q];
                    p2[q] = sp0[q];
                    p1[q] = sp1[q];
                } else {
                    const bool ok = jp < nh;
                    const double* P = a.Ps + boff + (ok ? jp : 0);
                    x[q] = ok ? a.xn[jp] : 0.0;
                    p2[q] = ok ? P[i0 * nh] : 0.0;
                    p1[q] = (ok && i0 + 1 < i1) ? P[(i0 + 1) * nh] : 0.0;
                }
            }
#pragma unroll
            for (int f = 0; f < 2; ++f) {
                const double2* cf = cfh + f * K;
                const double2 ce = cf[i0];
                const double2 co = cf[(i0 + 1 < i1) ? i0 + 1 : i0];
#pragma unroll
                for (int q = 0; q < kR; ++q) {
                    accE[q][f] = c2(pe_mul(p2[q], ce.x), pe_mul(p2[q], ce.y));
                    accO[q][f] = c2(pe_mul(p1[q], co.x), pe_mul(p1[q], co.y));
                }
            }
            // two degrees per step, double-buffered: the next pair's
            // coefficients are read while this pair's products run (the LDS
            // reads, not the FMAs, bound this loop); the recurrence pairs
            // come through the idle vector-memory path (buffer loads of a
            // wave-uniform address, counted apart from the LDS reads)
            const __amdgpu_buffer_rsrc_t rr = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<double2*>(rg), (short)0, K * (int)sizeof(double2), 0x00020000);
            struct Pair {
                double2 ab[2], cv[2][2];
            };
            auto load2 = [&](Pair& b, int ii) {
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    b.ab[u] = __builtin_bit_cast(
                        double2, __builtin_amdgcn_raw_buffer_load_b128(rr, (ii + u) * 16, 0, 0));
#pragma unroll
                    for (int f = 0; f < 2; ++f) b.cv[u][f] = cfh[f * K + ii + u];
                }
            };
            auto step2 = [&](const Pair& b) {
#pragma unroll
                for (int u = 0; u < 2; ++u) {
#pragma unroll
                    for (int q = 0; q < kR; ++q) {
                        const double p = rec_step(b.ab[u], x[q], p1[q], p2[q]);
                        p2[q] = p1[q];
                        p1[q] = p;
#pragma unroll
                        for (int f = 0; f < 2; ++f) {
                            double2& acc = (u & 1) ? accO[q][f] : accE[q][f];
                            acc.x = fma(p, b.cv[u][f].x, acc.x);
                            acc.y = fma(p, b.cv[u][f].y, acc.y);
                        }
                    }
                }
            };
            // (the reads run at most two degrees past the block: in-bounds
            // LDS, buffer reads past the column return 0; the values go
            // unused); the scheduling barriers keep each buffer's reads after
            // its last use, so the two buffers stay two register sets
            int i = i0 + 2;
            Pair pA, pB;
            load2(pA, i);
            for (; i + 4 <= i1; i += 4) {
                load2(pB, i + 2);
                step2(pA);
                __builtin_amdgcn_sched_barrier(0);
                load2(pA, i + 4);
                step2(pB);
                __builtin_amdgcn_sched_barrier(0);
            }
            if (i + 2 <= i1) {
                step2(pA);
                i += 2;
            }
            if (i < i1) {               // the block's last degree (odd count)
                const double2 ab = rg[i];
                double2 c[2];
#pragma unroll
                for (int f = 0; f < 2; ++f) c[f] = cfh[f * K + i];
#pragma unroll
                for (int q = 0; q < kR; ++q) {
                    const double p = rec_step(ab, x[q], p1[q], p2[q]);
#pragma unroll
                    for (int f = 0; f < 2; ++f) {
                        accE[q][f].x = fma(p, c[f].x, accE[q][f].x);
                        accE[q][f].y = fma(p, c[f].y, accE[q][f].y);
                    }
                }
            }
        }
        // the four blocks' sums meet in one round: each wave writes the
        // slots of the (latitude slot q, field f) pairs other waves finish,
        // then wave w finishes its own pairs (q = w for kR = 4; q = w % 2,
        // f = w / 2 for kR = 2) as (b0 + b2) + (b1 + b3) and stores them
        static_assert(kB1SynWaves == 4, "four synthesis blocks");
        auto owner = [](int q, int f) { return kR == 4 ? q : (q | (f << 1)); };
        if (syn_wave) {
#pragma unroll
            for (int q = 0; q < kR; ++q)
#pragma unroll
                for (int f = 0; f < 2; ++f)
                    if (owner(q, f) != kb) {
                        slot(kb, 0, f, q) = accE[q][f];
                        slot(kb, 1, f, q) = accO[q][f];
                    }
        }
        __syncthreads();
        if (syn_wave) {
#pragma unroll
            for (int q = 0; q < kR; ++q) {
                const int jp = jbase + g + 32 * q;
#pragma unroll
                for (int f = 0; f < 2; ++f) {
                    if (owner(q, f) != kb || jp >= nh) continue;
                    double2 e4[4], o4[4];
#pragma unroll
                    for (int bb = 0; bb < 4; ++bb) {
                        e4[bb] = (bb == kb) ? accE[q][f] : slot(bb, 0, f, q);
                        o4[bb] = (bb == kb) ? accO[q][f] : slot(bb, 1, f, q);
                    }
                    const double2 ev = cadd(cadd(e4[0], e4[2]), cadd(e4[1], e4[3]));
                    const double2 od = cadd(cadd(o4[0], o4[2]), cadd(o4[1], o4[3]));
                    const int o = (jp * n + m) * 4 + 2 * hf + f;
                    if (a.hemi) {                // the row values' spectra N = E + O, S = E - O
                        Ep[o] = cadd(ev, od);
                        Op[o] = csub(ev, od);
                    } else {
                        Ep[o] = ev;
                        Op[o] = od;
                    }
                }
            }
        }
        __syncthreads();
    }
}

static size_t bve1_lds(int L, int nh) {
    const size_t n = L + 1, R = (nh > 64) ? 4 : 2;
    const size_t ana = 2 * ((size_t)nh + 1) + kB1Waves * kB1Slots * kB1AW,
                 syn = kB1SynWaves * 2 * 2 * R * 64;
    return sizeof(double2) * (5 * n + (ana > syn ? ana : syn)) + sizeof(double) * (3 * n + (n & 1));
}

// The most degrees any workgroup of a launch analyses: the kernel's four
// 32-degree analysis slots hold 128 (degrees past them would be left out).
static int bve1_max_analysed(const aeo_plan_s* p) {
    const int n = p->L + 1, top = (2 * p->L) / 3, ms = bve1_msplit(p);
    int mx = 0;
    for (int m = 0; m <= top && m < n; ++m) {
        const int K = n - m;
        int cnt = top - m + 1;                       // unsplit: degrees 0 .. top - m
        if (m < ms) {
            const int s = bve1_split_point(K, kB1SplitQ8);
            const int c0 = ((s < top - m) ? s : top - m) + 1;   // first half: 0 .. min(s, top - m)
            const int c1 = (top - m) - (s - 1) + 1;             // second: s - 1 .. top - m
            cnt = (c0 > c1) ? c0 : c1;
        }
        if (cnt > mx) mx = cnt;
    }
    return mx;
}

static bool bve1_fits(const aeo_plan_s* p) {
    return p->kind == kGauss && bve1_lds(p->L, p->nh) <= 163840 && bve1_max_analysed(p) <= 128;
}

// Whether the single-member path applies to this model and plan (its
// lane-order table is built with the model: bve1_prepare).
bool bve1_supported(const aeo_model_s* md) {
    const aeo_plan_s* p = md->plan;
    return md->kind == kBVE && md->nmem == 1 && md->nslices == 0 && bve1_fits(p) &&
           p->d_b1pa != nullptr;
}

static int bve1_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        hipDeviceProp_t prop;
        cus = (hipGetDevice(&dev) == hipSuccess &&
               hipGetDeviceProperties(&prop, dev) == hipSuccess)
                  ? prop.multiProcessorCount
                  : 256;
    }
    return cus;
}

// Split columns of a synthesising launch: as many of the heaviest columns
// as the CUs left over by one workgroup per column (each split column keeps
// at least 8 degrees per half).
int bve1_msplit(const aeo_plan_s* p) {
    const int n = p->L + 1;
    int ms = bve1_cus() - n;
    if (ms > n - 16) ms = n - 16;
    return ms < 0 ? 0 : ms;
}

static int bve1_npass(const aeo_plan_s* p) { return (p->nh + kB1PassRows - 1) / kB1PassRows; }

int bve1_prepare(aeo_plan_s* p, hipStream_t s) {
    if (p->d_b1pa != nullptr || !bve1_fits(p)) return AEO_OK;
    const int L = p->L, ms = bve1_msplit(p), np = bve1_npass(p), nwg = L + 1 + ms;
    const size_t count = (size_t)nwg * np * kB1Slots * kB1PalQ * kB1Threads;
    double2* d = nullptr;
    AEO_CHECK_HIP(hipMalloc((void**)&d, count * sizeof(double2)));
    launch_k(bve1_pal_kernel, dim3(nwg), dim3(kB1Threads), 0, s, (const double*)p->d_Pa, d, L,
             p->nh, ms, kB1SplitQ8, (2 * L) / 3, np);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        (void)hipFree(d);
        set_error(std::string("bve1_prepare: ") + hipGetErrorString(e));
        return AEO_ERR_HIP;
    }
    p->d_b1pa = d;
    return AEO_OK;
}

int bve1_launch(aeo_model_s* md, const double2* ys, const double2* y0, const double2* forcing,
                double2* kout, int stage, double dt, double2* ynext, bool ana, bool syn,
                bool hemi, hipStream_t s) {
    const aeo_plan_s* p = md->plan;
    const int L = p->L, n = L + 1;
    Bve1Args a{};
    a.L = L;
    a.nh = p->nh;
    a.top = (2 * L) / 3;
    // the heavy columns run as two degree halves in every launch (in one
    // without synthesis too: a whole heavy column can analyse more degrees
    // than the four slots hold)
    const int ms = bve1_msplit(p);
    a.msplit = ms;
    a.split_q8 = kB1SplitQ8;
    a.ana = ana ? 1 : 0;
    a.syn = syn ? 1 : 0;
    a.Pal = p->d_b1pa;
    a.npass = bve1_npass(p);
    a.Ps = p->d_Ps;
    a.xn = p->d_xn;
    a.rec = p->d_rec;
    a.CpT = p->d_CpT;
    a.CmT = p->d_CmT;
    a.lap = md->d_lap;
    a.nulap = md->d_nulap;
    a.f10 = 2.0 * md->omega * sqrt(4.0 * M_PI / 3.0);
    a.S = md->d_S;
    a.D = md->d_D;
    a.wrow = p->d_w_rows;
    a.hemi = hemi ? 1 : 0;
    a.nlat = p->nlat;
    a.E = md->d_E;
    a.O = md->d_O;
    a.E1 = md->d_E1;
    a.O1 = md->d_O1;
    a.ys = ys;
    a.y0 = y0;
    a.ksum = md->d_ksum;
    a.ynext = ynext;
    a.kout = kout;
    a.forcing = forcing;
    a.stage = stage;
    a.cst = (stage == 2) ? dt : 0.5 * dt;
    a.dt6 = dt / 6.0;
    const size_t lds = bve1_lds(L, p->nh);
    const void* kern = (p->nh > 64) ? (const void*)bve1_stage_kernel<4>
                                    : (const void*)bve1_stage_kernel<2>;
    if (lds > 65536)
        AEO_CHECK_HIP(
            hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const double2* recg = p->d_rec;
    if (p->nh > 64)
        launch_k(bve1_stage_kernel<4>, dim3(n + ms), dim3(kB1Threads), lds, s, a, recg);
    else
        launch_k(bve1_stage_kernel<2>, dim3(n + ms), dim3(kB1Threads), lds, s, a, recg);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

}  // namespace aeo
