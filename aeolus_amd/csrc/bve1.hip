// Single-member FP64 barotropic-vorticity stage on Gauss grids: the
// latency-shaped case (one state, so one workgroup per zonal wavenumber m is
// all the parallelism the Legendre stage has).
//
// It computes exactly what leg_stage_kernel<KindBve, EpiBveS, 1, ANA, SYN, 0>
// computes, in the same arithmetic order (so results are bitwise equal), but
// is written for a short critical path:
//   * a small argument block and 32-bit indexing (no scalar reloads of a
//     large kernel-argument struct, no 64-bit address arithmetic per load);
//   * every global load that does not depend on this launch's results --
//     the state column, S/D, the analysis table rows, the coupling columns,
//     the recurrence column, the synthesis seeds, the RK4 inputs y0 / ksum /
//     forcing -- is issued at the start, before the first wait;
//   * the heavy columns (small m) get two workgroups, each synthesising half
//     of the latitudes; both run the (cheap, table-bound) analysis and the
//     epilogue, only the first writes the global RK4 outputs.
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
constexpr int kB1Pre = 16;    // analysis-table rows prefetched per lane and chunk
constexpr int kB1Chunks = 2;  // prefetched 64-degree chunks

// Raw buffer over `count` doubles (the gfx9 raw-buffer descriptor word 3).
__device__ inline __amdgpu_buffer_rsrc_t table_rsrc(const double* p, size_t count) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(p), (short)0,
                                             (int)(count * sizeof(double)), 0x00020000);
}
constexpr unsigned kOutOfRange = 0x80000000u;   // byte offset past every table block
__device__ inline double load_f64(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

struct Bve1Args {
    int L, nh, top;           // top: analysed degrees l <= top (2L/3)
    int msplit;               // columns m < msplit run as two latitude halves
    int ana, syn;             // analysis + epilogue; synthesis of the (new) state
    const double* Pa;         // analysis table, block m: [jp][l - m]
    const double* Ps;         // synthesis table, block m: [l - m][jp]
    const double* xn;         // nh northern nodes
    const double2* rec;       // [m][l] recurrence pairs
    const double* CpT;        // [m][l]
    const double* CmT;        // [m][l]
    const double* lap;        // -l(l+1)/R^2, l = 0 entry 0
    const double* nulap;      // nu lap'_l
    double f10;               // f at (1, 0)
    const double2* S;         // [jp][m]
    const double2* D;
    double2* E;               // [jp][m][4]
    double2* O;
    const double2* ys;        // stage input state, internal [m][l]
    const double2* y0;        // step input state
    double2* ksum;
    double2* ynext;
    double2* kout;            // stage < 0: tendency output
    const double2* forcing;   // or nullptr
    int stage;
    double cst, dt6;
};

__global__ __launch_bounds__(kB1Threads) void bve1_stage_kernel(Bve1Args a) {
    extern __shared__ double2 lds[];
    const int L = a.L, n = L + 1, nh = a.nh;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // workgroup -> (m, latitude half): the split columns first (heaviest first)
    int m, half;
    if ((int)blockIdx.x < 2 * a.msplit) {
        m = blockIdx.x >> 1;
        half = blockIdx.x & 1;
    } else {
        m = blockIdx.x - a.msplit;
        half = -1;
    }
    const int K = L - m + 1;
    const int top = a.top;
    const bool active = a.ana && m <= top;
    const bool writer = half <= 0;
    const long long boff = (long long)nh * ((long long)m * n - (long long)m * (m - 1) / 2);
    const int col = m * n;                       // internal [m][l] column offset

    // LDS carve-up: persistent columns, then one scratch region used by the
    // analysis (S/D, analysed column, wave partials) and later by the
    // synthesis tree
    double2* st = lds;                           // [n] state column (stage input, then next state)
    double2* rcc = st + n;                       // [n] recurrence pairs of m
    double2* cols = rcc + n;                     // [4][K] synthesis fields
    double* cpc = reinterpret_cast<double*>(cols + 4 * n);   // [n] C+ of m
    double* cmc = cpc + n;                       // [n] C- of m
    double* lapc = cmc + n;                      // [n] lap
    double2* scratch = reinterpret_cast<double2*>(lapc + n + (n & 1));
    double2* X = scratch;                        // [2][nh] S then D
    double2* accl = X + 2 * nh;                  // [n] analysed column
    double2* red = accl + n;                     // [8][64] analysis wave partials
    double2* sred = scratch;                     // [2 lg][2][2 parity][4][64] synthesis tree

    // ---- prologue: all independent global loads first
    double pa[kB1Chunks][kB1Pre];
    if (active) {
        const double* P = a.Pa + boff;
        // buffer loads over block m: lanes past the band get an offset past the
        // buffer's range, which reads 0 without a memory request (no branch per
        // load; clamping them onto one address instead serialises the wave)
        const __amdgpu_buffer_rsrc_t rs = table_rsrc(P, (size_t)K * nh);
#pragma unroll
        for (int c = 0; c < kB1Chunks; ++c)
#pragma unroll
            for (int u = 0; u < kB1Pre; ++u) {
                const int l = m + 64 * c + lane, jp = wave + kB1Waves * u;
                const unsigned off = (l <= top && jp < nh) ? (unsigned)(jp * K + (l - m)) * 8u
                                                           : kOutOfRange;
                pa[c][u] = load_f64(rs, off);
            }
    }
    // synthesis seeds of this lane's first latitude pass
    const int kb = wave & 3, lg = wave >> 2;
    const int Kb = ((K + 3) / 4 + 1) & ~1;
    const int i0 = kb * Kb, i1 = (i0 + Kb < K) ? i0 + Kb : K;
    const int lat_lo = (half < 0) ? 0 : 64 * half;             // latitude range of this workgroup
    const int lat_hi = (half < 0) ? nh : ((64 * half + 64 < nh) ? 64 * half + 64 : nh);
    const bool syn_wave = a.syn && (half < 0 || lg == 0);      // split halves use waves 0-3
    const int lat_first = (half < 0) ? lg * 64 + lane : lat_lo + lane;
    double sx = 0.0, sp0 = 0.0, sp1 = 0.0;
    if (syn_wave && lat_first < lat_hi && i0 < K) {
        const double* P = a.Ps + boff + lat_first;
        sx = a.xn[lat_first];
        sp0 = P[i0 * nh];
        if (i0 + 1 < K) sp1 = P[(i0 + 1) * nh];
    }
    // column data (thread t <-> degree l = t) and S/D (thread t <-> latitude t)
    double2 vs = c2(0.0, 0.0), vr = c2(0.0, 0.0), vy0 = c2(0.0, 0.0), vks = c2(0.0, 0.0);
    double2 vf = c2(0.0, 0.0), vS = c2(0.0, 0.0), vD = c2(0.0, 0.0);
    double vcp = 0.0, vcm = 0.0, vlap = 0.0, vnl = 0.0;
    const int tl = (tid < n) ? tid : n - 1;
    vs = a.ys[col + tl];
    vcp = a.CpT[col + tl];
    vcm = a.CmT[col + tl];
    vlap = a.lap[tl];
    if (a.syn) vr = a.rec[col + tl];
    if (a.ana) {
        vnl = a.nulap[tl];
        if (a.stage >= 0) vy0 = a.y0[col + tl];
        if (a.stage >= 1) vks = a.ksum[col + tl];
        if (a.forcing) vf = a.forcing[col + tl];
    }
    if (active) {
        const int tj = (tid < nh) ? tid : nh - 1;
        vS = a.S[tj * n + m];
        vD = a.D[tj * n + m];
    }
    // stores
    if (tid < n) {
        st[tid] = vs;
        rcc[tid] = vr;
        cpc[tid] = vcp;
        cmc[tid] = vcm;
        lapc[tid] = vlap;
        accl[tid] = c2(0.0, 0.0);
    }
    if (active && tid < nh) {
        X[tid] = vS;
        X[nh + tid] = vD;
    }
    for (int t = tid + kB1Threads; t < n; t += kB1Threads) {    // n > 512 (L > 511)
        st[t] = a.ys[col + t];
        rcc[t] = a.syn ? a.rec[col + t] : c2(0.0, 0.0);
        cpc[t] = a.CpT[col + t];
        cmc[t] = a.CmT[col + t];
        lapc[t] = a.lap[t];
        accl[t] = c2(0.0, 0.0);
    }
    if (active)
        for (int t = tid + kB1Threads; t < nh; t += kB1Threads) {
            X[t] = a.S[t * n + m];
            X[nh + t] = a.D[t * n + m];
        }
    __syncthreads();

    // ---- analysis: lanes over degree (64 per chunk), waves over latitude
    if (active) {
        const double* P = a.Pa + boff;
        for (int l0 = m, c = 0; l0 <= top; l0 += 64, ++c) {
            double2 acc = c2(0.0, 0.0);
            if (l0 + lane <= top) {
                const int par = (l0 + lane - m) & 1;
                for (int j0 = wave; j0 < nh; j0 += kB1Waves * kB1Pre) {
                    double pv[kB1Pre];
                    if (j0 == wave && c < kB1Chunks) {
#pragma unroll
                        for (int u = 0; u < kB1Pre; ++u) pv[u] = (c == 0) ? pa[0][u] : pa[kB1Chunks - 1][u];
                    } else {
                        const int l = l0 + lane;
#pragma unroll
                        for (int u = 0; u < kB1Pre; ++u) {
                            const int jp = j0 + kB1Waves * u;
                            pv[u] = (jp < nh && l <= top) ? P[jp * K + (l - m)] : 0.0;
                        }
                    }
                    // rows past nh have pv = 0 and read row nh - 1 (branch-free);
                    // four LDS reads in flight per group
#pragma unroll
                    for (int u0 = 0; u0 < kB1Pre; u0 += 4) {
                        double2 x[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int jq = j0 + kB1Waves * (u0 + u);
                            x[u] = X[par * nh + ((jq < nh) ? jq : nh - 1)];
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            acc.x = fma(pv[u0 + u], x[u].x, acc.x);
                            acc.y = fma(pv[u0 + u], x[u].y, acc.y);
                        }
                    }
                }
            }
            red[wave * 64 + lane] = acc;
            __syncthreads();
            if (tid < 64 && l0 + tid <= top) {
                double2 sv = red[tid];
#pragma unroll
                for (int w = 1; w < kB1Waves; ++w) sv = cadd(sv, red[w * 64 + tid]);
                accl[l0 + tid] = sv;
            }
            __syncthreads();
        }
    }

    // ---- epilogue: tendency and RK4 stage bookkeeping for degree l = tid
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
            const int idx = col + l;
            double2 k = c2(0.0, 0.0);
            if (l > 0) {
                const double2 acc = accl[l];
                k = cfma(st[l], nl, c2(-acc.x, -acc.y));
                if (a.forcing) k = cadd(k, fv);
            }
            double2 nx;
            if (a.stage < 0) {
                if (writer) a.kout[idx] = k;
                nx = k;
            } else {
                if (a.stage == 0) {
                    if (writer) a.ksum[idx] = k;
                    nx = cfma(k, a.cst, y0v);
                } else if (a.stage < 3) {
                    if (writer) a.ksum[idx] = cfma(k, 2.0, ksv);
                    nx = cfma(k, a.cst, y0v);
                } else {
                    const double2 s = cadd(ksv, k);
                    nx = cfma(s, a.dt6, y0v);
                }
                if (writer) a.ynext[idx] = nx;
            }
            st[l] = nx;
        }
        __syncthreads();
    }
    if (!a.syn) return;

    // ---- synthesis fields of the column: psi, S psi, eta, S eta ([f][i], i = l - m)
    for (int t = tid; t < 2 * K; t += kB1Threads) {
        const int f = (t < K) ? 0 : 2, i = (t < K) ? t : t - K, l = m + i;
        double2 v;
        if (f == 0) {
            if (l == 0) {
                v = c2(0.0, 0.0);
            } else {
                const double2 z = st[l];
                const double d = lapc[l];
                v = c2(z.x / d, z.y / d);
            }
        } else {
            v = st[l];
            if (l == 1 && m == 0) v.x += a.f10;
        }
        cols[f * K + i] = v;
    }
    __syncthreads();
    for (int t = tid; t < 2 * K; t += kB1Threads) {
        const int src = (t < K) ? 0 : 2, i = (t < K) ? t : t - K, l = m + i;
        const double2* s = cols + src * K;
        double2 v = c2(0.0, 0.0);
        if (i >= 1) v = cfma(s[i - 1], cpc[l - 1], v);
        if (l <= L - 1) v = cfma(s[i + 1], cmc[l + 1], v);
        cols[(src + 1) * K + i] = v;
    }
    __syncthreads();

    // ---- Pbar recurrence synthesis: wave = (degree block kb, latitude group)
    auto slot = [&](int g2, int s2, int par, int f) -> double2& {
        return sred[(((g2 * 2 + s2) * 2 + par) * 4 + f) * 64 + lane];
    };
    const int glat = (half < 0) ? lg : 0;         // LDS tree group of this wave
    for (int jbase = (half < 0 ? 0 : lat_lo); jbase < lat_hi; jbase += (half < 0 ? 128 : 64)) {
        const int jp = (half < 0) ? jbase + lg * 64 + lane : jbase + lane;
        const bool ok = syn_wave && jp < lat_hi;
        double2 accE[4], accO[4];
#pragma unroll
        for (int f = 0; f < 4; ++f) accE[f] = accO[f] = c2(0.0, 0.0);
        if (ok && i0 < i1) {
            const bool first = (jbase == (half < 0 ? 0 : lat_lo));
            const double* P = a.Ps + boff + jp;
            const double x = first ? sx : a.xn[jp];
            double p2 = first ? sp0 : P[i0 * nh];
            double p1 = first ? sp1 : ((i0 + 1 < i1) ? P[(i0 + 1) * nh] : 0.0);
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                const double2* cf = cols + f * K;
                const double2 ce = cf[i0];
                const double2 co = cf[(i0 + 1 < i1) ? i0 + 1 : i0];
                accE[f] = c2(pe_mul(p2, ce.x), pe_mul(p2, ce.y));
                accO[f] = c2(pe_mul(p1, co.x), pe_mul(p1, co.y));
            }
            int i = i0 + 2;
            for (; i + 4 <= i1; i += 4) {
                double2 ab[4], cv[4][4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    ab[u] = rcc[m + i + u];
#pragma unroll
                    for (int f = 0; f < 4; ++f) cv[u][f] = cols[f * K + i + u];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const double p = rec_step(ab[u], x, p1, p2);
                    p2 = p1;
                    p1 = p;
#pragma unroll
                    for (int f = 0; f < 4; ++f) {
                        double2& acc = (u & 1) ? accO[f] : accE[f];
                        acc.x = fma(p, cv[u][f].x, acc.x);
                        acc.y = fma(p, cv[u][f].y, acc.y);
                    }
                }
            }
            for (; i < i1; ++i) {       // the block's last degrees (fewer than a group)
                const double p = rec_step(rcc[m + i], x, p1, p2);
                p2 = p1;
                p1 = p;
                const bool odd = (i - i0) & 1;
#pragma unroll
                for (int f = 0; f < 4; ++f) {
                    const double2 c = cols[f * K + i];
                    double2& acc = odd ? accO[f] : accE[f];
                    acc.x = fma(p, c.x, acc.x);
                    acc.y = fma(p, c.y, acc.y);
                }
            }
        }
        // blocks meet as (b0 + b2) + (b1 + b3), the order of the generic kernel
        if (syn_wave && kb >= 2) {
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                slot(glat, kb - 2, 0, f) = accE[f];
                slot(glat, kb - 2, 1, f) = accO[f];
            }
        }
        __syncthreads();
        if (syn_wave && kb < 2) {
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                accE[f] = cadd(accE[f], slot(glat, kb, 0, f));
                accO[f] = cadd(accO[f], slot(glat, kb, 1, f));
            }
        }
        if (syn_wave && kb == 1) {
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                slot(glat, 1, 0, f) = accE[f];
                slot(glat, 1, 1, f) = accO[f];
            }
        }
        __syncthreads();
        if (syn_wave && kb == 0 && ok) {
            const int o = (jp * n + m) * 4;
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                a.E[o + f] = cadd(accE[f], slot(glat, 1, 0, f));
                a.O[o + f] = cadd(accO[f], slot(glat, 1, 1, f));
            }
        }
        __syncthreads();
    }
}

static size_t bve1_lds(int L, int nh) {
    const size_t n = L + 1;
    const size_t ana = 2 * (size_t)nh + n + kB1Waves * 64, syn = 2 * 2 * 2 * 4 * 64;
    return sizeof(double2) * (6 * n + (ana > syn ? ana : syn)) + sizeof(double) * (3 * n + (n & 1));
}

// Whether the single-member path applies to this model and plan.
bool bve1_supported(const aeo_model_s* md) {
    const aeo_plan_s* p = md->plan;
    return md->kind == kBVE && md->nmem == 1 && md->nslices == 0 && p->kind == kGauss &&
           bve1_lds(p->L, p->nh) <= 163840;
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

int bve1_launch(aeo_model_s* md, const double2* ys, const double2* y0, const double2* forcing,
                double2* kout, int stage, double dt, double2* ynext, bool ana, bool syn,
                hipStream_t s) {
    const aeo_plan_s* p = md->plan;
    const int L = p->L, n = L + 1;
    Bve1Args a{};
    a.L = L;
    a.nh = p->nh;
    a.top = (2 * L) / 3;
    // latitude halves only when one pass of 128 latitudes covers the hemisphere
    int ms = 0;
    if (syn && p->nh <= 128 && p->nh > 64) {
        ms = bve1_cus() - n;
        ms = ms < 0 ? 0 : (ms > n ? n : ms);
    }
    a.msplit = ms;
    a.ana = ana ? 1 : 0;
    a.syn = syn ? 1 : 0;
    a.Pa = p->d_Pa;
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
    a.E = md->d_E;
    a.O = md->d_O;
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
    if (lds > 65536)
        AEO_CHECK_HIP(hipFuncSetAttribute((const void*)bve1_stage_kernel,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    launch_k(bve1_stage_kernel, dim3(n + ms), dim3(kB1Threads), lds, s, a);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

}  // namespace aeo
