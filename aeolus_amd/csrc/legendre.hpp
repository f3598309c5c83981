// Per-m Legendre stage kernel for Gauss grids (FP64), batched over members.
//
// One workgroup owns one zonal wavenumber m (grid.x), one 64-latitude tile
// (grid.y) and a group of MB members (grid.z).  Depending on ANA / SYN it
//   ANA: contracts the analysis table block Pa[m] with the weighted Fourier
//        coefficients S/D of column m (lanes over degree l, the four waves
//        over latitude quarters, reduced in LDS), then runs the epilogue for
//        every (l, m) of the column: store, tendency, RK4 stage update.  The
//        epilogue leaves the new stage state of column m in LDS.
//   SYN: derives the synthesis fields of column m from the state column in
//        LDS (psi, eta, the S recurrence, ...), and contracts them with the
//        synthesis table block Ps[m] (lanes over latitude, the four waves over
//        l - m mod 4, so waves 0/2 hold the even and 1/3 the odd parity), giving
//        E/O = even/odd partial sums per latitude.
// With both set, the analysis of RK4 stage s and the synthesis of stage s+1
// run in one launch: column m of the new stage state never leaves the CU.
// The analysis part is recomputed by every latitude tile of a column (only
// tile 0 writes the epilogue's global outputs).
#pragma once

#include "plan.hpp"
#include "split.hpp"
#include "stage.hpp"

namespace aeo {

constexpr int kStageThreads = 512;
constexpr int kStageWaves = kStageThreads / 64;
// FP64 analysis-table loads issued ahead per wave (per 64 degrees, for the
// first PC chunks of 64), fewer for wide member/field tiles
template <int MB, int NA, int NF>
struct Prefetch {
    static constexpr int PA = (MB == 1 && NA <= 2) ? 16 : 8;
    static constexpr int PC = (MB == 1 && NA <= 2) ? 2 : 1;
};

// One step of the Pbar recurrence Pbar_l = a (x Pbar_{l-1} - b Pbar_{l-2}),
// written as fma(a x, p1, -(a b) p2) with ab = (a, a b): the only operation
// on the dependent path from p1 is one fma (the FP64 latency chain of a
// synthesis is one fma per degree).  The table kernel and the synthesis
// kernels share this exact expression.
__device__ inline double rec_step(double2 ab, double x, double p1, double p2) {
    return fma(ab.x * x, p1, -(ab.y * p2));
}

// p * c as the first term of a multiply-add chain (fma(p, c, 0) rounds alike)
__device__ inline double pe_mul(double p, double c) { return fma(p, c, 0.0); }

struct ColK {
    int L;
    double R;
    const double* Cp;   // (L+1)^2, column-major [m][l] (a column is contiguous)
    const double* Cm;   // (L+1)^2, column-major [m][l]
    const double* lap;  // -l(l+1)/R^2 (l = 0 entry 0), or nullptr
    double f10;         // f at (1, 0)
    // set inside the stage kernel: LDS copies of column m of Cp / Cm ([l])
    const double* cpc = nullptr;
    const double* cmc = nullptr;
};

// ---------------------------------------------------------------------------
// Synthesis kinds: which fields a column contributes.  SF[f] >= 0 marks field
// f as S applied to field SF[f]; other fields come from base().
// st points at the member's state columns, NSTATE x (L+1) complex.
// ---------------------------------------------------------------------------
struct KindNone {
    static constexpr int NSTATE = 1, NF = 1;
    static constexpr int SF[1] = {-1};
    __device__ static double2 base(int, const double2*, int, int, const ColK&) {
        return c2(0.0, 0.0);
    }
};

// Raw field (generic synthesis: one member per field).
struct KindRaw {
    static constexpr int NSTATE = 1, NF = 1;
    static constexpr int SF[1] = {-1};
    __device__ static double2 base(int, const double2* st, int l, int, const ColK&) {
        return st[l];
    }
};

// Jacobian of two given fields: a, S a, b, S b.
struct KindJac {
    static constexpr int NSTATE = 2, NF = 4;
    static constexpr int SF[4] = {-1, 0, -1, 2};
    __device__ static double2 base(int f, const double2* st, int l, int, const ColK& k) {
        return st[(f >> 1) * (k.L + 1) + l];
    }
};

// BVE: psi = zeta / lap (l >= 1), S psi, eta = zeta + f, S eta.
struct KindBve {
    static constexpr int NSTATE = 1, NF = 4;
    static constexpr int SF[4] = {-1, 0, -1, 2};
    __device__ static double2 base(int f, const double2* st, int l, int m, const ColK& k) {
        double2 z = st[l];
        if (f == 0) {
            if (l == 0) return c2(0.0, 0.0);
            const double d = k.lap[l];
            return c2(z.x / d, z.y / d);
        }
        if (l == 1 && m == 0) z.x += k.f10;
        return z;
    }
};

// SWE: psi, S psi, chi, S chi, eta, S eta, phi, S phi, zeta, delta.
struct KindSwe {
    static constexpr int NSTATE = 3, NF = 10;
    static constexpr int SF[10] = {-1, 0, -1, 2, -1, 4, -1, 6, -1, -1};
    __device__ static double2 base(int f, const double2* st, int l, int m, const ColK& k) {
        const int n = k.L + 1;
        switch (f) {
            case 0:
            case 2: {
                if (l == 0) return c2(0.0, 0.0);
                const double2 z = st[(f == 0 ? 0 : n) + l];
                const double d = k.lap[l];
                return c2(z.x / d, z.y / d);
            }
            case 4: {
                double2 z = st[l];
                if (l == 1 && m == 0) z.x += k.f10;
                return z;
            }
            case 6: return st[2 * n + l];
            case 8: return st[l];
            default: return st[n + l];
        }
    }
};

// BVE adjoint: the forward fields of the stage state (column 0; psi, S psi,
// eta, S eta as in KindBve) and, from the tendency cotangent kb (column 1),
// the synthesis input of the Jacobian cotangent
//     jb_lm = -kb_lm / c_m   for 1 <= l <= cut, m <= cut,   0 elsewhere
// (c_0 = 1, c_m = 2, cut = 2L/3).  Synthesising jb (which applies c_m) gives
// the grid field jt with Ana^T(-trunc kb) = w jt, w the quadrature weight.
struct KindBveAdj {
    static constexpr int NSTATE = 2, NF = 5;
    static constexpr int SF[5] = {-1, 0, -1, 2, -1};
    __device__ static double2 base(int f, const double2* st, int l, int m, const ColK& k) {
        if (f < 4) return KindBve::base(f, st, l, m, k);
        const int cut = (2 * k.L) / 3;
        if (l == 0 || l > cut || m > cut) return c2(0.0, 0.0);
        const double s = (m == 0) ? -1.0 : -0.5;
        const double2 v = st[(k.L + 1) + l];
        return c2(s * v.x, s * v.y);
    }
};

// Fill cols[f][i] (i = l - m) for the member's state column st.
template <class KS>
__device__ inline void derive_columns(const double2* st, int m, const ColK& k, double2* cols) {
    const int K = k.L - m + 1;
    for (int t = threadIdx.x; t < KS::NF * K; t += kStageThreads) {
        const int f = t / K, i = t - f * K;
        if (KS::SF[f] < 0) cols[t] = KS::base(f, st, m + i, m, k);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < KS::NF * K; t += kStageThreads) {
        const int f = t / K, i = t - f * K;
        const int src = KS::SF[f];
        if (src >= 0) {
            const int l = m + i;
            const double2* s = cols + src * K;
            double2 v = c2(0.0, 0.0);
            if (i >= 1) v = cfma(s[i - 1], k.cpc[l - 1], v);
            if (l <= k.L - 1) v = cfma(s[i + 1], k.cmc[l + 1], v);
            cols[t] = v;
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Stage epilogues.  apply(b, l, m, acc, st, writer): b = member, acc = analysed
// values (zero outside the band), st = the member's state columns in LDS
// (read the stage input, overwrite with the next stage state), writer =
// whether this workgroup writes the global outputs.
// ---------------------------------------------------------------------------
// Global inputs of an epilogue entry (y0, ksum, forcing, topography, ...) are
// read by prefetch(b, l, m, pf) into NPF double2 slots; the single-member
// stage kernel issues these loads with its prologue (they do not depend on
// the launch's own results) and hands them to apply(); other launches pass
// nullptr and apply() reads them itself.
struct EpiNoneS {
    static constexpr int NA = 1, NSTATE = 1, NPF = 0;
    __device__ void prefetch(int, int, int, double2*) const {}
    __device__ void apply(int, int, int, const double2*, double2*, bool, const double2*) const {}
};

// Store the analysed field (generic analysis: one member per field; Jacobian).
struct EpiStore1S {
    static constexpr int NA = 1, NSTATE = 1, NPF = 0;
    double2* out;
    int L;
    size_t member_stride;
    __device__ void prefetch(int, int, int, double2*) const {}
    __device__ void apply(int b, int l, int m, const double2* acc, double2*, bool writer,
                          const double2*) const {
        if (writer) out[b * member_stride + l * (L + 1) + m] = acc[0];
    }
};

// RK4 stage bookkeeping on one entry (see Rk4 in stage.hpp); returns the
// next stage state (or k for the tendency-only mode).  Inputs y0[idx] and
// ksum[idx] come from pre() (2 slots).
struct Rk4B {
    int stage;      // -1: tendency only
    double cst;
    double dt6;
    const double2* y0;
    double2* ksum;
    double2* ynext;
    double2* kout;
    __device__ void pre(size_t idx, double2* pf) const {
        pf[0] = (stage >= 0) ? y0[idx] : c2(0.0, 0.0);
        pf[1] = (stage >= 1) ? ksum[idx] : c2(0.0, 0.0);
    }
    __device__ double2 apply(size_t idx, double2 k, bool writer, const double2* pf) const {
        if (stage < 0) {
            if (writer) kout[idx] = k;
            return k;
        }
        const double2 y = pf[0];
        double2 nx;
        if (stage == 0) {
            if (writer) ksum[idx] = k;
            nx = cfma(k, cst, y);
        } else if (stage < 3) {
            if (writer) ksum[idx] = cfma(k, 2.0, pf[1]);
            nx = cfma(k, cst, y);
        } else {
            const double2 s = cadd(pf[1], k);
            nx = cfma(s, dt6, y);
        }
        if (writer) ynext[idx] = nx;
        return nx;
    }
};

// BVE: k = -J + nu lap' zeta + forcing, row 0 pinned.  Slots: y0, ksum,
// forcing, (nu lap'_l, 0).
struct EpiBveS {
    static constexpr int NA = 1, NSTATE = 1, NPF = 4;
    int L;
    size_t member_stride;
    const double2* forcing;  // per member, or nullptr
    const double* nulap;
    Rk4B rk;
    __device__ size_t index(int b, int l, int m) const {
        return b * member_stride + (size_t)m * (L + 1) + l;   // internal [m][l]
    }
    __device__ void prefetch(int b, int l, int m, double2* pf) const {
        const size_t idx = index(b, l, m);
        rk.pre(idx, pf);
        pf[2] = forcing ? forcing[idx] : c2(0.0, 0.0);
        pf[3] = c2(nulap[l], 0.0);
    }
    __device__ void apply(int b, int l, int m, const double2* acc, double2* st, bool writer,
                          const double2* pin) const {
        double2 pf[NPF];
        if (pin) {
#pragma unroll
            for (int i = 0; i < NPF; ++i) pf[i] = pin[i];
        } else {
            prefetch(b, l, m, pf);
        }
        const size_t idx = index(b, l, m);
        double2 k = c2(0.0, 0.0);
        if (l > 0) {
            k = cfma(st[l], pf[3].x, c2(-acc[0].x, -acc[0].y));
            if (forcing) k = cadd(k, pf[2]);
        }
        st[l] = rk.apply(idx, k, writer, pf);
    }
};

// SWE assembly (phy/shallow_water.py:389-418).  Slots: (y0, ksum) per field,
// lap phi_s, nu4 lap^2 phi_s, (lap_l, nu4 lap_l^2).
struct EpiSweS {
    static constexpr int NA = 4, NSTATE = 3, NPF = 9;
    int L;
    size_t member_stride;
    const double* lap;
    const double* nu4lap2;
    const double2* lap_phi_s;
    const double2* nu4_lap2_phi_s;
    double phi0;
    Rk4B rk;
    __device__ void prefetch(int b, int l, int m, double2* pf) const {
        const int n = L + 1;
        const size_t fs = (size_t)n * n;
        const size_t e = (size_t)m * n + l;   // internal [m][l]
        const size_t base = b * member_stride;
        rk.pre(base + e, pf);
        rk.pre(base + fs + e, pf + 2);
        rk.pre(base + 2 * fs + e, pf + 4);
        const size_t eu = (size_t)l * n + m;  // topography terms keep the boundary layout
        pf[6] = lap_phi_s ? lap_phi_s[eu] : c2(0.0, 0.0);
        pf[7] = nu4_lap2_phi_s ? nu4_lap2_phi_s[eu] : c2(0.0, 0.0);
        pf[8] = c2(lap[l], nu4lap2 ? nu4lap2[l] : 0.0);
    }
    __device__ void apply(int b, int l, int m, const double2* acc, double2* st, bool writer,
                          const double2* pin) const {
        double2 pf[NPF];
        if (pin) {
#pragma unroll
            for (int i = 0; i < NPF; ++i) pf[i] = pin[i];
        } else {
            prefetch(b, l, m, pf);
        }
        const int n = L + 1;
        const size_t fs = (size_t)n * n;
        const size_t e = (size_t)m * n + l;   // internal [m][l]
        const size_t base = b * member_stride;
        double2 kz = c2(0.0, 0.0), kd = c2(0.0, 0.0), kp = c2(0.0, 0.0);
        if (l > 0) {
            const double2 zeta = st[l], delta = st[n + l], phi = st[2 * n + l];
            kz = acc[0];
            kp = csub(acc[3], cscale(delta, phi0));
            kd = csub(acc[1], cscale(cadd(acc[2], phi), pf[8].x));
            if (lap_phi_s) kd = csub(kd, pf[6]);
            if (nu4lap2) {
                const double q = pf[8].y;
                kz = csub(kz, cscale(zeta, q));
                kd = csub(kd, cscale(delta, q));
                kp = csub(kp, cscale(phi, q));
                if (nu4_lap2_phi_s) kp = csub(kp, pf[7]);
            }
        }
        st[l] = rk.apply(base + e, kz, writer, pf);
        st[n + l] = rk.apply(base + fs + e, kd, writer, pf + 2);
        st[2 * n + l] = rk.apply(base + 2 * fs + e, kp, writer, pf + 4);
    }
};

// Reverse-mode RK4 bookkeeping on one entry.  With a = cotangent of the step
// output and yb_i = T'(y_i)^T kb_i the stage results (adjoint stage r runs
// forward stage 3 - r):
//   kb_4 = dt/6 a                       (set before stage 0)
//   r=0: asum  = a + yb_4,  kb_3 = dt/3 a + dt yb_4
//   r=1: asum += yb_3,      kb_2 = dt/3 a + dt/2 yb_3
//   r=2: asum += yb_2,      kb_1 = dt/6 a + dt/2 yb_2
//   r=3: anext = asum + yb_1 (cotangent of the step input), kb_4' = dt/6 anext
// Inputs a[idx] and asum[idx] come from pre() (2 slots).  Returns the next
// stage's kb.
struct Rk4Adj {
    int stage;          // 0..3; -1: tendency adjoint only (kout = yb)
    double dt;
    const double2* a;
    double2* asum;
    double2* anext;
    double2* kout;
    double2* knext;     // next kb (global copy for the next launch's epilogue)
    __device__ void pre(size_t idx, double2* pf) const {
        pf[0] = (stage >= 0) ? a[idx] : c2(0.0, 0.0);
        pf[1] = (stage >= 1) ? asum[idx] : c2(0.0, 0.0);
    }
    __device__ double2 apply(size_t idx, double2 yb, bool writer, const double2* pf) const {
        if (stage < 0) {
            if (writer) kout[idx] = yb;
            return c2(0.0, 0.0);
        }
        const double2 av = pf[0];
        double2 nx;
        if (stage == 0) {
            if (writer) asum[idx] = cadd(av, yb);
            nx = cadd(cscale(av, dt / 3.0), cscale(yb, dt));
        } else if (stage == 1) {
            if (writer) asum[idx] = cadd(pf[1], yb);
            nx = cadd(cscale(av, dt / 3.0), cscale(yb, 0.5 * dt));
        } else if (stage == 2) {
            if (writer) asum[idx] = cadd(pf[1], yb);
            nx = cadd(cscale(av, dt / 6.0), cscale(yb, 0.5 * dt));
        } else {
            const double2 tot = cadd(pf[1], yb);
            if (writer) anext[idx] = tot;
            nx = cscale(tot, dt / 6.0);
        }
        if (writer && knext) knext[idx] = nx;
        return nx;
    }
};

// Transpose of one BVE stage k = -trunc Ana(J(psi, eta)) + nu lap' zeta,
// row 0 pinned, psi = zeta / lap (l >= 1), eta = zeta + f.  The four analysed
// columns are weighted analyses of the cotangents of the Jacobian's grid
// inputs (acc 0: a_l, 1: a_s, 2: b_l, 3: b_s); then
//   psib = c_m ( -(i m / R) A_l + S^T A_s / R ),
//   etab = c_m ( -(i m / R) B_l + S^T B_s / R ),
//   zb   = nu lap' kb' + psib / lap (l >= 1) + etab,     kb' = kb with row 0 = 0,
// with (S^T x)_l = C+_l x_{l+1} + C-_l x_{l-1} (l-1 >= m).  State column 0
// receives the forward state of the next adjoint stage, column 1 its kb.
// Slots: a, asum, the next forward state, (nu lap'_l, lap_l).
template <class T, class = void>
struct epi_column { static constexpr bool value = false; };
template <class T>
struct epi_column<T, decltype((void)T::COLUMN)> { static constexpr bool value = T::COLUMN; };

struct EpiBveAdjS {
    static constexpr int NA = 4, NSTATE = 2, NPF = 4;
    static constexpr bool COLUMN = true;
    int L;
    double R;
    size_t member_stride;
    const double* nulap;
    const double* lap;
    const double* Cp;          // column-major [m][l]
    const double* Cm;          // column-major [m][l]
    const double2* fwd_next;   // forward stage state of the next adjoint stage, or nullptr
    Rk4Adj rk;
    __device__ size_t index(int b, int l, int m) const {
        return b * member_stride + (size_t)m * (L + 1) + l;   // internal [m][l]
    }
    __device__ void prefetch(int b, int l, int m, double2* pf) const {
        const size_t idx = index(b, l, m);
        rk.pre(idx, pf);
        pf[2] = fwd_next ? fwd_next[idx] : c2(0.0, 0.0);
        pf[3] = c2(nulap[l], lap[l]);
    }
    // (S^T x)_l with the column-m coefficients cpc / cmc ([l], LDS or global)
    __device__ double2 st_term(const double2* x, int l, int m, const double* cpc,
                               const double* cmc) const {
        double2 v = c2(0.0, 0.0);
        if (l >= m) {
            if (l + 1 <= L) v = cadd(v, cscale(x[l + 1], cpc[l]));
            if (l - 1 >= m) v = cadd(v, cscale(x[l - 1], cmc[l]));
        }
        return v;
    }
    // acc: the member's analysed columns [NA][n]; st: its state columns [2][n];
    // cpc / cmc: C+ / C- of column m
    __device__ void apply_col(int b, int l, int m, const double2* acc, double2* st,
                              bool writer, const double2* pin, const double* cpc,
                              const double* cmc) const {
        double2 pf[NPF];
        if (pin) {
#pragma unroll
            for (int i = 0; i < NPF; ++i) pf[i] = pin[i];
        } else {
            prefetch(b, l, m, pf);
        }
        const int n = L + 1;
        const size_t idx = index(b, l, m);
        const double q = (double)m / R, invR = 1.0 / R;
        const double cm = (m == 0) ? 1.0 : 2.0;
        const double2 al = acc[l], bl = acc[2 * n + l];
        const double2 sa = st_term(acc + n, l, m, cpc, cmc), sb = st_term(acc + 3 * n, l, m, cpc, cmc);
        // -(i q) x = (q x.y, -q x.x)
        const double2 psib = cscale(c2(q * al.y + invR * sa.x, -q * al.x + invR * sa.y), cm);
        const double2 etab = cscale(c2(q * bl.y + invR * sb.x, -q * bl.x + invR * sb.y), cm);
        double2 zb = etab;
        if (l > 0) {
            const double d = pf[3].y;
            zb = cadd(zb, c2(psib.x / d, psib.y / d));
            zb = cadd(zb, cscale(st[n + l], pf[3].x));
        }
        const double2 kn = rk.apply(idx, zb, writer, pf);
        st[l] = fwd_next ? pf[2] : c2(0.0, 0.0);
        st[n + l] = kn;
    }
};

// ---------------------------------------------------------------------------
// The stage kernel.
// ---------------------------------------------------------------------------
struct StageArgs {
    const double2* ys;      // stage input states [member][NSTATE][n][n] (nullptr: zeros)
    size_t ys_stride;       // complex entries per member
    const double2* ys_aux;  // if set: state fields 1.. come from here instead
    size_t aux_stride;      // complex entries per member of ys_aux
    int ys_mmajor;          // 1: ys is [..][m][l] (internal layout); 0: [..][l][m]
    int nmem;
    int nstate_valid;       // state fields present in ys (<= NSTATE)
    const double2* S;       // [jp][m][member][NA]
    const double2* D;
    int lmax_a, mmax_a;
    double2* E;             // [jp][m][member][NF]
    double2* O;
    ColK k;
};

// double2 slots of the per-column constants in LDS: C+, C-, lap (FP64, 3n
// doubles, rounded up) and the recurrence pairs (a_lm, b_lm) (n double2)
__host__ __device__ inline int coef_slots(int n) { return (3 * n + 1) / 2 + n; }

// double2 slots per degree of the epilogue prefetch: single-member launches
// with an analysis load the epilogue's global inputs with the prologue.
template <class Epi, int MB>
__host__ __device__ constexpr int epi_pre_slots(bool ana) {
    return (MB == 1 && ana) ? Epi::NPF : 0;
}

template <class KS, class Epi, int MB, int NS>
struct StageLds {
    static constexpr int NST = (KS::NSTATE > Epi::NSTATE) ? KS::NSTATE : Epi::NSTATE;
    static constexpr int NCA = MB * Epi::NA * 2, NTA = (NCA + 15) / 16;
    static constexpr int NCS = MB * KS::NF * 2, NTS = (NCS + 15) / 16;
    static size_t bytes(int L, int nh, bool ana, bool syn) {
        const int n = L + 1;
        // state + C+/C-/lap + the single-member epilogue prefetch
        size_t s = sizeof(double2) * ((size_t)MB * NST * n + coef_slots(n) +
                                      (size_t)epi_pre_slots<Epi, MB>(ana) * n);
        size_t a = 0, c = 0;
        if (ana) {
            a = sizeof(double2) * ((size_t)MB * Epi::NA * (2 * (nh + 1) + n) +
                                   (size_t)kStageWaves * MB * Epi::NA * 64);
            if (NS > 0) a = sizeof(double2) * ((size_t)MB * Epi::NA * (2 * (nh + 1) + n)) +
                            sizeof(int) * 2 * NTA * 16 + 2 * NS * NTA * 16 * split_nhk(nh) + 64;
        }
        if (syn) {
            c = sizeof(double2) * ((size_t)MB * KS::NF * n +
                                   (MB > 1 ? 0 : 2 * 2 * 2 * (size_t)MB * KS::NF * 64));
            if (NS > 0) c = sizeof(double2) * ((size_t)MB * KS::NF * n) + sizeof(int) * NTS * 16 +
                            2 * NS * NTS * 16 * split_kpad(L, 0) + 64;
        }
        return s + (a > c ? a : c);
    }
};

__device__ inline double cpart(double2 v, int ri) { return ri ? v.y : v.x; }

// ex[c] = exp_above(max_i |val(c, i)|) for c < ncol, i < len: every thread of
// the workgroup cooperates, tpc (a power of two <= 64) threads per column,
// reduced with lane shuffles.
template <class F>
__device__ inline void col_max_exp(int ncol, int len, int* ex, const F& val) {
    int tpc = 64;
    while (tpc > 1 && tpc * ncol > kStageThreads) tpc >>= 1;
    const int cpb = kStageThreads / tpc;
    const int sub = threadIdx.x & (tpc - 1);
    for (int c0 = 0; c0 < ncol; c0 += cpb) {
        const int c = c0 + threadIdx.x / tpc;
        double mx = 0.0;
        if (c < ncol)
            for (int i = sub; i < len; i += tpc) mx = fmax(mx, fabs(val(c, i)));
        for (int off = tpc >> 1; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
        if (sub == 0 && c < ncol) ex[c] = exp_above(mx);
    }
}


// Column m of stage workgroup (bx, bz) for launches that read a row-major
// [l][m] batch (grid.x = stage_grid_x(L + 1)): workgroups are dealt to the
// 8 XCDs round-robin by id (XCD = bx mod 8, grid.x a multiple of 8), so
// groups of 8 consecutive m -- which share the 128-byte lines of the batch --
// go to one XCD (one L2) instead of eight, and the groups are dealt
// round-robin, rotated by bz, so every XCD gets light and heavy columns.
// Ids past the last column return at once.  Launches on the m-major internal
// layout keep m = bx (heaviest columns first).
__host__ __device__ inline int stage_grid_x(int n) { return (n + 63) / 64 * 64; }
__device__ inline int stage_column(int bx, int bz) {
    const int x = (bx + bz) & 7, s = bx >> 3;
    return (((s >> 3) << 3) + x) * 8 + (s & 7);
}

template <class KS, class Epi, int MB, bool ANA, bool SYN, int NS>
__global__ __launch_bounds__(kStageThreads) void leg_stage_kernel(GaussDev g, StageArgs a,
                                                                  Epi ep) {
    constexpr int NST = StageLds<KS, Epi, MB, NS>::NST;
    constexpr int NA = Epi::NA;
    constexpr int NF = KS::NF;
    constexpr int kPre = Prefetch<MB, NA, NF>::PA;
    constexpr int kPC = Prefetch<MB, NA, NF>::PC;
    extern __shared__ double2 lds[];
    // the plan and stage arguments the prologue reads into scalar registers
    // at entry, in one batch (loaded where first used, they came in rounds
    // waited for between the prologue's global loads)
    asm volatile("" ::"s"(g.L), "s"(g.nh), "s"(g.Pa), "s"(g.Ps), "s"(g.xn), "s"(g.rec),
                 "s"(a.ys), "s"(a.ys_stride), "s"(a.ys_aux), "s"(a.ys_mmajor), "s"(a.nmem),
                 "s"(a.nstate_valid), "s"(a.S), "s"(a.D), "s"(a.lmax_a), "s"(a.mmax_a));
    asm volatile("" ::"s"(a.E), "s"(a.O), "s"(a.k.Cp), "s"(a.k.Cm), "s"(a.k.lap));
    const int m = a.ys_mmajor ? (int)blockIdx.x : stage_column(blockIdx.x, blockIdx.z);
    const int b0 = blockIdx.z * MB;
    const int L = g.L, nh = g.nh, n = L + 1;
    if (m >= n) return;
    const int K = L - m + 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int top = (a.lmax_a < L) ? a.lmax_a : L;      // analysed degrees l <= top
    const bool active = ANA && (m <= a.mmax_a);
    double2* st = lds;                                  // [MB][NST][n]
    double* cpc = reinterpret_cast<double*>(lds + (size_t)MB * NST * n);   // C+ column m
    double* cmc = cpc + n;                                                 // C- column m
    double* lapc = cmc + n;                                                // lap
    double2* rcc = lds + (size_t)MB * NST * n + (3 * n + 1) / 2;           // (a, b) of m
    constexpr int kNPF = epi_pre_slots<Epi, MB>(ANA);
    constexpr bool kEpiPre = kNPF > 0;
    double2* epf = lds + (size_t)MB * NST * n + coef_slots(n);              // [n][kNPF]
    double2* work = epf + (size_t)kNPF * n;
    ColK kc = a.k;
    kc.cpc = cpc;
    kc.cmc = cmc;
    if (a.k.lap) kc.lap = lapc;

    // -- prologue: every load that does not depend on another workgroup's
    //    results of this launch is issued before the first wait.  FP64
    //    analysis-table rows go to registers (pa: the first 64 kPC analysed
    //    degrees); state, S/D and coefficient columns go through registers
    //    into LDS.
    double pa[kPC][kPre];
    if constexpr (ANA && NS == 0) {
        const double* Pa = g.Pa + g.block_offset(m);
#pragma unroll
        for (int c = 0; c < kPC; ++c)
#pragma unroll
            for (int u = 0; u < kPre; ++u) {
                const int l = m + 64 * c + lane, jp = wave + kStageWaves * u;
                pa[c][u] = (active && l <= top && jp < nh) ? Pa[(size_t)jp * K + (l - m)] : 0.0;
            }
    }

    // Split precision: per-m table offsets, and each wave's first analysis and
    // synthesis tile: its digit rows (first K-chunks) and row exponents.  The
    // loads are unconditional (waves without a tile read tile 0) so the
    // compiler counts them exactly and no earlier wait has to drain them.
    constexpr int NSX = (NS > 0) ? NS : 1;
    constexpr int kNTA = StageLds<KS, Epi, MB, NS>::NTA, kNTS = StageLds<KS, Epi, MB, NS>::NTS;
    // K-chunks (kSplitK deep) loaded per group, and prefetched
    constexpr int kKCA = (NS * kNTA <= 7) ? 2 : 1;
    constexpr int kKCS = (NS * kNTS <= 7) ? 2 : 1;
    constexpr int kPFA = kKCA, kPFS = kKCS;
    intx4 avA[kPFA][NSX];
    int erA[4] = {0, 0, 0, 0};
    intx4 avS[kPFS][NSX];
    long avSt[NSX];                      // synthesis K tail (32 deep)
    int erS[4] = {0, 0, 0, 0};
    long long oSA = 0, oEA = 0, oSS = 0;
    if constexpr (NS > 0) {
        if constexpr (ANA) {
            oSA = g.offs_sa[m];
            oEA = g.offs_ea[m];
        }
        if constexpr (SYN) oSS = g.offs_ss[m];
    }
    // analysis: issued after the prologue's own loads, before their LDS stores
    auto prefetch_ana = [&]() {
        if constexpr (ANA && NS > 0) {
            const int nhk = split_nhk(nh);
            const int kp = split_kp(L, m), mpad = rup(kp, 16);
            const int Mv0 = (top >= m) ? (top - m) / 2 + 1 : 0;
            const int Mv1 = (top >= m + 1) ? (top - m - 1) / 2 + 1 : 0;
            const int MT0 = (Mv0 + 15) / 16, MT1 = (Mv1 + 15) / 16;
            const bool has = wave < MT0 + MT1;
            const int p = (has && wave >= MT0) ? 1 : 0;
            const int mt = has ? wave - p * MT0 : 0;
            const signed char* Pt = g.Pasl + oSA;
#pragma unroll
            for (int c4 = 0; c4 < kPFA; ++c4) {
                const int k0 = (kSplitK * c4 < nhk) ? kSplitK * c4 : 0;
#pragma unroll
                for (int q = 0; q < NSX; ++q)
                    avA[c4][q] = *reinterpret_cast<const intx4*>(
                        Pt + ((size_t)(p * NS + q) * mpad + mt * 16 + (lane & 15)) * nhk + k0 +
                        (lane >> 4) * 16);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r)
                erA[r] = g.eana[oEA + p * mpad + mt * 16 + (lane >> 4) * 4 + r];
        }
    };
    // synthesis: issued with the prologue loads for synthesis-only launches,
    // else after the epilogue (whose global loads it would otherwise delay)
    auto prefetch_syn = [&]() {
        if constexpr (SYN && NS > 0) {
            const int kpad = split_kpad(L, m), nhp = rup(nh, 16);
            const int kfull = kpad & ~(kSplitK - 1);
            // rows of 32 bytes (no full chunk): the lane offsets stay in the row
            const int lo = (kfull > 0) ? (lane >> 4) * 16 : ((lane >> 4) & 1) * 16;
            const int tt0 = (wave < 2 * (nhp / 16)) ? wave : 0;
            const int p = tt0 & 1, j0 = (tt0 >> 1) * 16;
            const signed char* Pt = g.Psl + oSS;
#pragma unroll
            for (int c4 = 0; c4 < kPFS; ++c4) {
                const int k0 = (kSplitK * c4 + kSplitK <= kfull) ? kSplitK * c4 : 0;
#pragma unroll
                for (int q = 0; q < NSX; ++q)
                    avS[c4][q] = *reinterpret_cast<const intx4*>(
                        Pt + ((size_t)(p * NS + q) * nhp + j0 + (lane & 15)) * kpad + k0 + lo);
            }
#pragma unroll
            for (int q = 0; q < NSX; ++q)
                avSt[q] = *reinterpret_cast<const long*>(
                    Pt + ((size_t)(p * NS + q) * nhp + j0 + (lane & 15)) * kpad + kpad - 32 +
                    (lane >> 4) * 8);
#pragma unroll
            for (int r = 0; r < 4; ++r)
                erS[r] = g.esyn[(size_t)m * 2 * nhp + p * nhp + j0 + (lane >> 4) * 4 + r];
        }
    };

    // FP64 synthesis seeds of this lane for the first latitude pass: node x
    // and the table rows at the start of the lane's degree block
    double sx = 0.0, sp0 = 0.0, sp1 = 0.0;
    if constexpr (SYN && NS == 0) {
        const int kb = (MB > 1) ? (lane >> 4) : (wave & 3);
        const int jp = (MB > 1) ? (wave * 16 + (lane & 15)) : ((wave >> 2) * 64 + lane);
        const int Kb = ((K + 3) / 4 + 1) & ~1;
        const int i0 = kb * Kb;
        if (jp < nh && i0 < K) {
            const double* P = g.Ps + g.block_offset(m) + jp;
            sx = g.xn[jp];
            sp0 = P[(size_t)i0 * nh];
            if (i0 + 1 < K) sp1 = P[(size_t)(i0 + 1) * nh];
        }
    }
    // [MB][NA][2][nh + 1] (S then D): planes nh + 1 apart, so the S and D
    // reads of one latitude (lanes of both parities in one instruction) fall
    // in different LDS banks
    const int nhx = nh + 1;
    double2* X = work;
    double2* accl = X + (size_t)MB * NA * 2 * nhx;      // [MB][NA][n]
    {
        const int tot_s = MB * NST * n;
        const int tot_x = active ? MB * NA * nh : 0;
        const int tot = (tot_s > tot_x ? tot_s : tot_x) > n ? (tot_s > tot_x ? tot_s : tot_x) : n;
        // loads are unconditional where possible (indices clamped in range) so
        // the values stay in registers; only the LDS stores are predicated.
        // The first pass also issues the split tile prefetch between its
        // loads and its stores (peeled, so the prefetch stays behind them).
        for (int t0 = 0; t0 < tot; t0 += 2 * kStageThreads) {
            double2 vs[2], vx[2][2], vr[2];
            double vc[2][3];
            double2 vp[2][kNPF > 0 ? kNPF : 1];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int t = t0 + threadIdx.x + u * kStageThreads;
                vs[u] = c2(0.0, 0.0);
                {
                    const int ts = (t < tot_s) ? t : tot_s - 1;
                    const int bb = ts / (NST * n);
                    const int r = ts - bb * NST * n;
                    const int sf = r / n, l = r - sf * n;
                    const int b = b0 + bb;
                    if (b < a.nmem && sf < a.nstate_valid) {
                        const size_t e = a.ys_mmajor ? (size_t)m * n + l : (size_t)l * n + m;
                        if (sf >= 1 && a.ys_aux)
                            vs[u] = a.ys_aux[b * a.aux_stride + (size_t)(sf - 1) * n * n + e];
                        else if (a.ys)
                            vs[u] = a.ys[b * a.ys_stride + (size_t)sf * n * n + e];
                    }
                }
                vx[u][0] = vx[u][1] = c2(0.0, 0.0);
                if (tot_x > 0) {
                    const int tx = (t < tot_x) ? t : tot_x - 1;
                    const int f = tx % NA;
                    const int bb = (tx / NA) % MB;
                    const int jp = tx / (NA * MB);
                    const int b = b0 + bb;
                    if (b < a.nmem) {
                        const size_t src = (((size_t)jp * n + m) * a.nmem + b) * NA + f;
                        vx[u][0] = a.S[src];
                        vx[u][1] = a.D[src];
                    }
                }
                const int tc = (t < n) ? t : n - 1;
                vc[u][0] = a.k.Cp[(size_t)m * n + tc];
                vc[u][1] = a.k.Cm[(size_t)m * n + tc];
                vc[u][2] = a.k.lap ? a.k.lap[tc] : 0.0;
                vr[u] = c2(0.0, 0.0);
                if constexpr (SYN && NS == 0) vr[u] = g.rec[(size_t)m * n + tc];   // FP64 recurrence
                if constexpr (kEpiPre) ep.prefetch(b0, tc, m, vp[u]);                // epilogue inputs
            }
            if (t0 == 0) {
                if constexpr (ANA && NS > 0) prefetch_ana();
                if constexpr (SYN && NS > 0 && !ANA) prefetch_syn();
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int t = t0 + threadIdx.x + u * kStageThreads;
                if (t < tot_s) st[t] = vs[u];
                if (t < tot_x) {
                    const int f = t % NA;
                    const int bb = (t / NA) % MB;
                    const int jp = t / (NA * MB);
                    X[((bb * NA + f) * 2 + 0) * nhx + jp] = vx[u][0];
                    X[((bb * NA + f) * 2 + 1) * nhx + jp] = vx[u][1];
                }
                // u = 0: unconditional at the clamped index (threads past n
                // rewrite entry n - 1 with its own value), so the loads are
                // not sunk behind later ones
                if (u == 0 || t < n) {
                    const int tc = (t < n) ? t : n - 1;
                    cpc[tc] = vc[u][0];
                    cmc[tc] = vc[u][1];
                    lapc[tc] = vc[u][2];
                    if constexpr (SYN && NS == 0) rcc[tc] = vr[u];
                    if constexpr (kEpiPre) {
#pragma unroll
                        for (int i = 0; i < kNPF; ++i) epf[(size_t)tc * kNPF + i] = vp[u][i];
                    }
                }
            }
        }
    }
    if constexpr (ANA)
        for (int t = threadIdx.x; t < MB * NA * n; t += kStageThreads) accl[t] = c2(0.0, 0.0);
    __syncthreads();

    if constexpr (ANA) {
        double2* red = accl + (size_t)MB * NA * n;          // [waves][MB][NA][64]
        if constexpr (NS > 0) if (active) {
            // ---- split-precision analysis: INT8 digit MFMA, exact INT32, FP64 recombination
            constexpr int NCA = StageLds<KS, Epi, MB, NS>::NCA;
            constexpr int NTA = kNTA;
            constexpr int NCP = NTA * 16;
            const int nhk = split_nhk(nh);
            const int kp = split_kp(L, m), mpad = rup(kp, 16);
            int* gexp = reinterpret_cast<int*>(accl + (size_t)MB * NA * n);   // [2][NCP]
            // 2 NCP ints after a 16-byte boundary keep xs 16-byte aligned; a plain
            // cast (no integer round trip) keeps the LDS address space
            signed char* xs = reinterpret_cast<signed char*>(gexp + 2 * NCP);  // [2][NS][NCP][nhk]
            unsigned short* xsh = reinterpret_cast<unsigned short*>(xs);
            col_max_exp(2 * NCP, nh, gexp, [&](int t, int j) {
                const int p = t / NCP, c = t - p * NCP;
                if (c >= NCA) return 0.0;
                const int bb = c / (NA * 2), f = (c >> 1) % NA, ri = c & 1;
                return cpart(X[((bb * NA + f) * 2 + p) * nhx + j], ri);
            });
            __syncthreads();
            // rows (p, c) over waves, latitude pairs over lanes (one 16-bit
            // store per slice): no runtime divisions
            // (padding columns c >= NCA: zero rows, no digit work; the branch is
            // uniform over the wave)
            for (int rr = wave; rr < 2 * NCP; rr += kStageWaves) {
                const int p = rr / NCP, c = rr - p * NCP;
                unsigned short* row = xsh + (((size_t)p * NS * NCP + c) * nhk >> 1);
                const size_t qstride = (size_t)NCP * nhk >> 1;
                if (c >= NCA) {
                    for (int j = 2 * lane; j < nhk; j += 128)
#pragma unroll
                        for (int q = 0; q < NS; ++q) row[q * qstride + (j >> 1)] = 0;
                    continue;
                }
                const int bb = c / (NA * 2), f = (c >> 1) % NA, ri = c & 1;
                const double2* xr = X + ((bb * NA + f) * 2 + p) * nhx;
                const int e = gexp[p * NCP + c];
                for (int j = 2 * lane; j < nhk; j += 128) {
                    const double v0 = (j < nh) ? cpart(xr[j], ri) : 0.0;
                    const double v1 = (j + 1 < nh) ? cpart(xr[j + 1], ri) : 0.0;
                    int d0[NS], d1[NS];
                    digits<NS>(v0, e, d0);
                    digits<NS>(v1, e, d1);
#pragma unroll
                    for (int q = 0; q < NS; ++q) row[q * qstride + (j >> 1)] = digit_pair(d0[q], d1[q]);
                }
            }
            __syncthreads();
            const signed char* Pt = g.Pasl + oSA;
            const int* er = g.eana + oEA;
            double* accd = reinterpret_cast<double*>(accl);
            // 16-degree tiles of both parities spread over all waves
            const int Mv0 = (top >= m) ? (top - m) / 2 + 1 : 0;
            const int Mv1 = (top >= m + 1) ? (top - m - 1) / 2 + 1 : 0;
            const int MT0 = (Mv0 + 15) / 16, MT1 = (Mv1 + 15) / 16;
            // Tiles are software-pipelined: the next tile's first K-group and
            // row exponents are loaded while this tile's MFMAs drain and it is
            // recombined; the wave's first tile came from the prologue.
            constexpr int kKC = kKCA;
            intx4 cur[kKC][NS], nxt[kKC][NS];
            int ecur[4], enxt[4];
#pragma unroll
            for (int c4 = 0; c4 < kKC; ++c4)
#pragma unroll
                for (int q = 0; q < NS; ++q) cur[c4][q] = avA[c4][q];
#pragma unroll
            for (int r = 0; r < 4; ++r) ecur[r] = erA[r];
            for (int tt = wave; tt < MT0 + MT1; tt += kStageWaves) {
                const int p = (tt < MT0) ? 0 : 1;
                const int mt = (tt < MT0) ? tt : tt - MT0;
                const int Mv = p ? Mv1 : Mv0;
                intx4 acc[NS][NTA];
#pragma unroll
                for (int q = 0; q < NS; ++q)
#pragma unroll
                    for (int nt = 0; nt < NTA; ++nt) acc[q][nt] = intx4{0, 0, 0, 0};
                for (int kg = 0; kg < nhk; kg += kSplitK * kKC) {
                    intx4 av[kKC][NS];
                    if (kg == 0) {
#pragma unroll
                        for (int c4 = 0; c4 < kKC; ++c4)
#pragma unroll
                            for (int q = 0; q < NS; ++q) av[c4][q] = cur[c4][q];
                    } else {
#pragma unroll
                        for (int c4 = 0; c4 < kKC; ++c4) {
                            const int k0 = (kg + kSplitK * c4 < nhk) ? kg + kSplitK * c4 : kg;
#pragma unroll
                            for (int q = 0; q < NS; ++q)
                                av[c4][q] = *reinterpret_cast<const intx4*>(
                                    Pt + ((size_t)(p * NS + q) * mpad + mt * 16 + (lane & 15)) * nhk +
                                    k0 + (lane >> 4) * 16);
                        }
                    }
#pragma unroll
                    for (int c4 = 0; c4 < kKC; ++c4) {
                        const int k0 = kg + kSplitK * c4;
                        if (k0 < nhk) {
#pragma unroll
                            for (int nt = 0; nt < NTA; ++nt)
#pragma unroll
                                for (int t2 = 0; t2 < NS; ++t2) {
                                    const intx4 bv = *reinterpret_cast<const intx4*>(
                                        xs + ((size_t)(p * NS + t2) * NCP + nt * 16 + (lane & 15)) * nhk +
                                        k0 + (lane >> 4) * 16);
#pragma unroll
                                    for (int q = 0; q + t2 < NS; ++q)
                                        acc[q + t2][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                                            av[c4][q], bv, acc[q + t2][nt], 0, 0, 0);
                                }
                        }
                    }
                }
                // next tile's first K-group and exponents: issued after this
                // tile's MFMAs, read at the next iteration (no reload on the last)
                const int tn = tt + kStageWaves;
                if (tn < MT0 + MT1) {
                    const int pn = (tn < MT0) ? 0 : 1;
                    const int mtn = (tn < MT0) ? tn : tn - MT0;
#pragma unroll
                    for (int c4 = 0; c4 < kKC; ++c4) {
                        const int k0 = (kSplitK * c4 < nhk) ? kSplitK * c4 : 0;
#pragma unroll
                        for (int q = 0; q < NS; ++q)
                            nxt[c4][q] = *reinterpret_cast<const intx4*>(
                                Pt + ((size_t)(pn * NS + q) * mpad + mtn * 16 + (lane & 15)) * nhk +
                                k0 + (lane >> 4) * 16);
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r) enxt[r] = er[pn * mpad + mtn * 16 + (lane >> 4) * 4 + r];
                }
#pragma unroll
                for (int nt = 0; nt < NTA; ++nt) {
                    const int c = nt * 16 + (lane & 15);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int ir = mt * 16 + (lane >> 4) * 4 + r;
                        // explicit fma: the same rounding in every member tiling
                        double v = 0.0;
#pragma unroll
                        for (int q = NS - 1; q >= 0; --q) v = fma(v, kInvRadix, (double)acc[q][nt][r]);
                        if (c < NCA && ir < Mv) {
                            const int bb = c / (NA * 2), f = (c >> 1) % NA, ri = c & 1;
                            const int l = m + 2 * ir + p;
                            accd[2 * ((size_t)(bb * NA + f) * n + l) + ri] =
                                ldexp(v, ecur[r] + gexp[p * NCP + c]) * kInvTop2;
                        }
                    }
                }
#pragma unroll
                for (int c4 = 0; c4 < kKC; ++c4)
#pragma unroll
                    for (int q = 0; q < NS; ++q) cur[c4][q] = nxt[c4][q];
#pragma unroll
                for (int r = 0; r < 4; ++r) ecur[r] = enxt[r];
            }
            __syncthreads();
        }
        if constexpr (NS == 0) if (active) {
            // one chunk of 64 degrees per pass (lanes = degrees); the first kPC
            // chunks use the rows prefetched in the prologue
            constexpr int kCH = 1;
            const double* Pa = g.Pa + g.block_offset(m);
            for (int l0 = m, c = 0; l0 <= top; l0 += 64 * kCH, ++c) {
                double2 acc[kCH][MB][NA];
#pragma unroll
                for (int h = 0; h < kCH; ++h)
#pragma unroll
                    for (int bb = 0; bb < MB; ++bb)
#pragma unroll
                        for (int f = 0; f < NA; ++f) acc[h][bb][f] = c2(0.0, 0.0);
                if (l0 + lane <= top) {
                    const int par = (l0 + lane - m) & 1;
                    for (int j0 = wave; j0 < nh; j0 += kStageWaves * kPre) {
                        double pv[kCH][kPre];
                        if (j0 == wave && c < kPC) {
#pragma unroll
                            for (int u = 0; u < kPre; ++u)
                                pv[0][u] = (c == 0) ? pa[0][u] : pa[kPC - 1][u];
                        } else {
#pragma unroll
                            for (int h = 0; h < kCH; ++h) {
                                const int l = l0 + 64 * h + lane;
#pragma unroll
                                for (int u = 0; u < kPre; ++u) {
                                    const int jp = j0 + kStageWaves * u;
                                    pv[h][u] = (jp < nh && l <= top) ? Pa[(size_t)jp * K + (l - m)]
                                                                     : 0.0;
                                }
                            }
                        }
                        // rows past nh have pv = 0 and read row nh - 1 (branch-free,
                        // so every LDS read of the pass is issued before the first fma)
#pragma unroll
                        for (int u = 0; u < kPre; ++u) {
                            const int jq = j0 + kStageWaves * u;
                            const int jp = (jq < nh) ? jq : nh - 1;
#pragma unroll
                            for (int bb = 0; bb < MB; ++bb)
#pragma unroll
                                for (int f = 0; f < NA; ++f) {
                                    const double2 x = X[((bb * NA + f) * 2 + par) * nhx + jp];
#pragma unroll
                                    for (int h = 0; h < kCH; ++h) {
                                        acc[h][bb][f].x = fma(pv[h][u], x.x, acc[h][bb][f].x);
                                        acc[h][bb][f].y = fma(pv[h][u], x.y, acc[h][bb][f].y);
                                    }
                                }
                        }
                    }
                }
#pragma unroll
                for (int h = 0; h < kCH; ++h)
#pragma unroll
                    for (int bb = 0; bb < MB; ++bb)
#pragma unroll
                        for (int f = 0; f < NA; ++f)
                            red[((wave * MB + bb) * NA + f) * (64 * kCH) + 64 * h + lane] =
                                acc[h][bb][f];
                __syncthreads();
                for (int t = threadIdx.x; t < MB * NA * 64 * kCH; t += kStageThreads) {
                    const int r = t % (64 * kCH);
                    const int ll = l0 + r;
                    if (ll <= top) {
                        const int q = t / (64 * kCH);  // (bb * NA + f)
                        double2 sv = red[q * (64 * kCH) + r];
#pragma unroll
                        for (int w = 1; w < kStageWaves; ++w)
                            sv = cadd(sv, red[(w * MB * NA + q) * (64 * kCH) + r]);
                        accl[q * n + ll] = sv;
                    }
                }
                __syncthreads();
            }
        }
        // epilogue over every (l, m) of the column and every member
        const bool writer = true;   // one workgroup per (m, member group)
        for (int t = threadIdx.x; t < MB * n; t += kStageThreads) {
            const int bb = t / n, l = t - bb * n;
            const int b = b0 + bb;
            if (b >= a.nmem) continue;
            const double2* pin = kEpiPre ? epf + (size_t)t * kNPF : nullptr;
            if constexpr (epi_column<Epi>::value) {
                ep.apply_col(b, l, m, accl + (size_t)bb * NA * n, st + (size_t)bb * NST * n,
                             writer, pin, cpc, cmc);
            } else {
                double2 acc[NA];
#pragma unroll
                for (int f = 0; f < NA; ++f) acc[f] = accl[(bb * NA + f) * n + l];
                ep.apply(b, l, m, acc, st + (size_t)bb * NST * n, writer, pin);
            }
        }
        __syncthreads();
    }

    if constexpr (SYN && NS > 0) {
        // ---- split-precision synthesis: INT8 digit MFMA, exact INT32, FP64 recombination
        if constexpr (ANA) prefetch_syn();
        constexpr int NCS = StageLds<KS, Epi, MB, NS>::NCS;
        constexpr int NTS = kNTS;
        constexpr int NCP = NTS * 16;
        double2* cols = work;                               // [MB][NF][K]
        for (int bb = 0; bb < MB; ++bb)
            derive_columns<KS>(st + (size_t)bb * NST * n, m, kc, cols + (size_t)bb * NF * K);
        const int kpad = split_kpad(L, m), nhp = rup(nh, 16);
        int* fexp = reinterpret_cast<int*>(cols + (size_t)MB * NF * n);       // [NCP]
        signed char* qs = reinterpret_cast<signed char*>(fexp + NCP);        // [2][NS][NCP][kpad]
        unsigned short* qsh = reinterpret_cast<unsigned short*>(qs);
        col_max_exp(NCP, K, fexp, [&](int c, int i) {
            return (c < NCS) ? cpart(cols[(size_t)(c >> 1) * K + i], c & 1) : 0.0;
        });
        __syncthreads();
        // rows (p, c) over waves, degree pairs (ip, ip + 1) over lanes
        // (padding columns c >= NCS: zero rows, no digit work)
        for (int rr = wave; rr < 2 * NCP; rr += kStageWaves) {
            const int p = rr / NCP, c = rr - p * NCP;
            unsigned short* row = qsh + (((size_t)p * NS * NCP + c) * kpad >> 1);
            const size_t qstride = (size_t)NCP * kpad >> 1;
            if (c >= NCS) {
                for (int ip = 2 * lane; ip < kpad; ip += 128)
#pragma unroll
                    for (int q = 0; q < NS; ++q) row[q * qstride + (ip >> 1)] = 0;
                continue;
            }
            const double2* cr = cols + (size_t)(c >> 1) * K;
            const int e = fexp[c];
            for (int ip = 2 * lane; ip < kpad; ip += 128) {
                const int i = 2 * ip + p;
                const double v0 = (i < K) ? cpart(cr[i], c & 1) : 0.0;
                const double v1 = (i + 2 < K) ? cpart(cr[i + 2], c & 1) : 0.0;
                int d0[NS], d1[NS];
                digits<NS>(v0, e, d0);
                digits<NS>(v1, e, d1);
#pragma unroll
                for (int q = 0; q < NS; ++q) row[q * qstride + (ip >> 1)] = digit_pair(d0[q], d1[q]);
            }
        }
        __syncthreads();
        // (16-latitude tile, parity) pairs spread over the waves
        const signed char* Pt = g.Psl + oSS;
        const int* er = g.esyn + (size_t)m * 2 * nhp;
        // software-pipelined like the analysis tiles; the wave's first tile
        // was issued before derive_columns
        constexpr int kKC = kKCS;
        const int T = 2 * (nhp / 16);
        const int kfull = kpad & ~(kSplitK - 1);     // 64-deep chunks; a 32-deep tail follows
        const int lo = (kfull > 0) ? (lane >> 4) * 16 : ((lane >> 4) & 1) * 16;
        intx4 cur[kKC][NS], nxt[kKC][NS];
        long curt[NS], nxtt[NS];
        int ecur[4], enxt[4];
#pragma unroll
        for (int c4 = 0; c4 < kKC; ++c4)
#pragma unroll
            for (int q = 0; q < NS; ++q) cur[c4][q] = avS[c4][q];
#pragma unroll
        for (int q = 0; q < NS; ++q) curt[q] = avSt[q];
#pragma unroll
        for (int r = 0; r < 4; ++r) ecur[r] = erS[r];
        for (int tt = wave; tt < T; tt += kStageWaves) {
            const int p = tt & 1;
            const int j0 = (tt >> 1) * 16;
            intx4 acc[NS][NTS];
#pragma unroll
            for (int q = 0; q < NS; ++q)
#pragma unroll
                for (int nt = 0; nt < NTS; ++nt) acc[q][nt] = intx4{0, 0, 0, 0};
            for (int kg = 0; kg < kfull; kg += kSplitK * kKC) {
                intx4 av[kKC][NS];
                if (kg == 0) {
#pragma unroll
                    for (int c4 = 0; c4 < kKC; ++c4)
#pragma unroll
                        for (int q = 0; q < NS; ++q) av[c4][q] = cur[c4][q];
                } else {
#pragma unroll
                    for (int c4 = 0; c4 < kKC; ++c4) {
                        const int k0 = (kg + kSplitK * c4 < kfull) ? kg + kSplitK * c4 : kg;
#pragma unroll
                        for (int q = 0; q < NS; ++q)
                            av[c4][q] = *reinterpret_cast<const intx4*>(
                                Pt + ((size_t)(p * NS + q) * nhp + j0 + (lane & 15)) * kpad + k0 +
                                (lane >> 4) * 16);
                    }
                }
#pragma unroll
                for (int c4 = 0; c4 < kKC; ++c4) {
                    const int k0 = kg + kSplitK * c4;
                    if (k0 < kfull) {
#pragma unroll
                        for (int nt = 0; nt < NTS; ++nt)
#pragma unroll
                            for (int t2 = 0; t2 < NS; ++t2) {
                                const intx4 bv = *reinterpret_cast<const intx4*>(
                                    qs + ((size_t)(p * NS + t2) * NCP + nt * 16 + (lane & 15)) * kpad +
                                    k0 + (lane >> 4) * 16);
#pragma unroll
                                for (int q = 0; q + t2 < NS; ++q)
                                    acc[q + t2][nt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                                        av[c4][q], bv, acc[q + t2][nt], 0, 0, 0);
                            }
                    }
                }
            }
            if (kpad != kfull) {   // the 32-deep tail at kfull = kpad - 32
#pragma unroll
                for (int nt = 0; nt < NTS; ++nt)
#pragma unroll
                    for (int t2 = 0; t2 < NS; ++t2) {
                        const long bv = *reinterpret_cast<const long*>(
                            qs + ((size_t)(p * NS + t2) * NCP + nt * 16 + (lane & 15)) * kpad + kfull +
                            (lane >> 4) * 8);
#pragma unroll
                        for (int q = 0; q + t2 < NS; ++q)
                            acc[q + t2][nt] = __builtin_amdgcn_mfma_i32_16x16x32_i8(
                                curt[q], bv, acc[q + t2][nt], 0, 0, 0);
                    }
            }
            // next tile's first K-group and exponents (after this tile's MFMAs)
            const int tn = tt + kStageWaves;
            if (tn < T) {
                const int pn = tn & 1, jn = (tn >> 1) * 16;
#pragma unroll
                for (int c4 = 0; c4 < kKC; ++c4) {
                    const int k0 = (kSplitK * c4 + kSplitK <= kfull) ? kSplitK * c4 : 0;
#pragma unroll
                    for (int q = 0; q < NS; ++q)
                        nxt[c4][q] = *reinterpret_cast<const intx4*>(
                            Pt + ((size_t)(pn * NS + q) * nhp + jn + (lane & 15)) * kpad + k0 + lo);
                }
#pragma unroll
                for (int q = 0; q < NS; ++q)
                    nxtt[q] = *reinterpret_cast<const long*>(
                        Pt + ((size_t)(pn * NS + q) * nhp + jn + (lane & 15)) * kpad + kpad - 32 +
                        (lane >> 4) * 8);
#pragma unroll
                for (int r = 0; r < 4; ++r) enxt[r] = er[pn * nhp + jn + (lane >> 4) * 4 + r];
            }
            double2* dst = (p == 0) ? a.E : a.O;
#pragma unroll
            for (int nt = 0; nt < NTS; ++nt) {
                const int c = nt * 16 + (lane & 15);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = j0 + (lane >> 4) * 4 + r;
                    // explicit fma: the same rounding in every member tiling
                    double v = 0.0;
#pragma unroll
                    for (int q = NS - 1; q >= 0; --q) v = fma(v, kInvRadix, (double)acc[q][nt][r]);
                    v = ldexp(v, ecur[r] + fexp[c]) * kInvTop2;
                    const double w = __shfl_xor(v, 1);
                    if (!(c & 1) && c < NCS && j < nh) {
                        const int bb = c / (NF * 2), f = (c >> 1) % NF;
                        const int b = b0 + bb;
                        if (b < a.nmem) dst[(((size_t)j * n + m) * a.nmem + b) * NF + f] = c2(v, w);
                    }
                }
            }
#pragma unroll
            for (int c4 = 0; c4 < kKC; ++c4)
#pragma unroll
                for (int q = 0; q < NS; ++q) cur[c4][q] = nxt[c4][q];
#pragma unroll
            for (int q = 0; q < NS; ++q) curt[q] = nxtt[q];
#pragma unroll
            for (int r = 0; r < 4; ++r) ecur[r] = enxt[r];
        }
    }
    if constexpr (SYN && NS == 0) {
        // FP64 synthesis, every latitude of column m in this workgroup, 128
        // latitudes per pass.  Each latitude's degrees are split into 4
        // contiguous blocks of even length; Pbar along a block comes from the
        // recurrence (rec_step) seeded by two table rows, so the table is read
        // twice per block instead of once per degree.
        //   MB == 1: wave w takes block w % 4 of latitude group w / 4 (64
        //            lanes = 64 latitudes; all lanes read the same column entry)
        //            and the four block sums meet in a two-round LDS tree.
        //   MB > 1:  lane = (16 latitudes, 4 blocks), waves = 16-latitude
        //            groups; the block sums meet by lane shuffles, so no LDS
        //            beyond the columns (more workgroups per CU for batches).
        static_assert(kStageWaves == 8, "synthesis maps 8 waves to 128 latitudes");
        constexpr bool kLaneBlocks = (MB > 1);
        double2* cols = work;                               // [MB][NF][K]
        double2* red = cols + (size_t)MB * NF * K;          // [2 lg][2][2 parity][MB][NF][64]
        for (int bb = 0; bb < MB; ++bb)
            derive_columns<KS>(st + (size_t)bb * NST * n, m, kc, cols + (size_t)bb * NF * K);
        constexpr int kRecG = (MB * NF <= 4) ? 4 : 2;      // degrees per recurrence group
        const int kb = kLaneBlocks ? (lane >> 4) : (wave & 3);
        const int lg = wave >> 2;
        const int lat0 = kLaneBlocks ? (wave * 16 + (lane & 15)) : (lg * 64 + lane);
        const int Kb = ((K + 3) / 4 + 1) & ~1;              // even block length
        const int i0 = kb * Kb, i1 = (i0 + Kb < K) ? i0 + Kb : K;
        auto slot = [&](int s2, int par, int bb, int f) -> double2& {
            return red[((((lg * 2 + s2) * 2 + par) * MB + bb) * NF + f) * 64 + lane];
        };
        for (int jbase = 0; jbase < nh; jbase += 128) {
            const int jp = jbase + lat0;
            const bool ok = jp < nh;
            double2 accE[MB][NF], accO[MB][NF];
#pragma unroll
            for (int bb = 0; bb < MB; ++bb)
#pragma unroll
                for (int f = 0; f < NF; ++f) accE[bb][f] = accO[bb][f] = c2(0.0, 0.0);
            if (ok && i0 < i1) {
                const double* P = g.Ps + g.block_offset(m) + jp;
                const bool first = (jbase == 0);        // seeds issued in the prologue
                const double x = first ? sx : g.xn[jp];
                double p2 = first ? sp0 : P[(size_t)i0 * nh];
                double p1 = first ? sp1 : ((i0 + 1 < i1) ? P[(size_t)(i0 + 1) * nh] : 0.0);
#pragma unroll
                for (int bb = 0; bb < MB; ++bb)
#pragma unroll
                    for (int f = 0; f < NF; ++f) {
                        const double2* cf = cols + (bb * NF + f) * K;
                        const double2 ce = cf[i0];
                        const double2 co = cf[(i0 + 1 < i1) ? i0 + 1 : i0];
                        accE[bb][f] = c2(pe_mul(p2, ce.x), pe_mul(p2, ce.y));
                        accO[bb][f] = c2(pe_mul(p1, co.x), pe_mul(p1, co.y));
                    }
                // groups of kRecG degrees (i0 + 2 is even, so degree parity =
                // parity of u): every LDS read of a group is issued before its
                // recurrence and multiply-adds; out-of-block degrees get p = 0
                int i = i0 + 2;
                for (; i + kRecG <= i1; i += kRecG) {
                    double2 ab[kRecG];
                    double2 cv[kRecG][MB][NF];
#pragma unroll
                    for (int u = 0; u < kRecG; ++u) {
                        ab[u] = rcc[m + i + u];
#pragma unroll
                        for (int bb = 0; bb < MB; ++bb)
#pragma unroll
                            for (int f = 0; f < NF; ++f) cv[u][bb][f] = cols[(bb * NF + f) * K + i + u];
                    }
#pragma unroll
                    for (int u = 0; u < kRecG; ++u) {
                        const double p = rec_step(ab[u], x, p1, p2);
                        p2 = p1;
                        p1 = p;
#pragma unroll
                        for (int bb = 0; bb < MB; ++bb)
#pragma unroll
                            for (int f = 0; f < NF; ++f) {
                                double2& acc = (u & 1) ? accO[bb][f] : accE[bb][f];
                                acc.x = fma(p, cv[u][bb][f].x, acc.x);
                                acc.y = fma(p, cv[u][bb][f].y, acc.y);
                            }
                    }
                }
                for (; i < i1; ++i) {   // the block's last degrees (fewer than a group)
                    const double p = rec_step(rcc[m + i], x, p1, p2);
                    p2 = p1;
                    p1 = p;
                    const bool odd = (i - i0) & 1;
#pragma unroll
                    for (int bb = 0; bb < MB; ++bb)
#pragma unroll
                        for (int f = 0; f < NF; ++f) {
                            const double2 c = cols[(bb * NF + f) * K + i];
                            double2& acc = odd ? accO[bb][f] : accE[bb][f];
                            acc.x = fma(p, c.x, acc.x);
                            acc.y = fma(p, c.y, acc.y);
                        }
                }
            }
            if constexpr (kLaneBlocks) {
#pragma unroll
                for (int bb = 0; bb < MB; ++bb)
#pragma unroll
                    for (int f = 0; f < NF; ++f) {
                        // (b0 + b2) + (b1 + b3): the association of the LDS tree,
                        // so batched and single-member results agree bitwise
                        double2 e = accE[bb][f], o = accO[bb][f];
                        e.x = sum_xor32(e.x); e.y = sum_xor32(e.y);
                        o.x = sum_xor32(o.x); o.y = sum_xor32(o.y);
                        e.x = sum_xor16(e.x); e.y = sum_xor16(e.y);
                        o.x = sum_xor16(o.x); o.y = sum_xor16(o.y);
                        accE[bb][f] = e;
                        accO[bb][f] = o;
                    }
                if (kb == 0 && ok) {
#pragma unroll
                    for (int bb = 0; bb < MB; ++bb) {
                        const int b = b0 + bb;
                        if (b >= a.nmem) continue;
#pragma unroll
                        for (int f = 0; f < NF; ++f) {
                            const size_t o = (((size_t)jp * n + m) * a.nmem + b) * NF + f;
                            a.E[o] = accE[bb][f];
                            a.O[o] = accO[bb][f];
                        }
                    }
                }
                continue;
            }
            // tree step 1: blocks 2, 3 -> blocks 0, 1
            if (kb >= 2) {
#pragma unroll
                for (int bb = 0; bb < MB; ++bb)
#pragma unroll
                    for (int f = 0; f < NF; ++f) {
                        slot(kb - 2, 0, bb, f) = accE[bb][f];
                        slot(kb - 2, 1, bb, f) = accO[bb][f];
                    }
            }
            __syncthreads();
            if (kb < 2) {
#pragma unroll
                for (int bb = 0; bb < MB; ++bb)
#pragma unroll
                    for (int f = 0; f < NF; ++f) {
                        accE[bb][f] = cadd(accE[bb][f], slot(kb, 0, bb, f));
                        accO[bb][f] = cadd(accO[bb][f], slot(kb, 1, bb, f));
                    }
            }
            // tree step 2: block 1 -> block 0 (block 1 reuses its own slot)
            if (kb == 1) {
#pragma unroll
                for (int bb = 0; bb < MB; ++bb)
#pragma unroll
                    for (int f = 0; f < NF; ++f) {
                        slot(1, 0, bb, f) = accE[bb][f];
                        slot(1, 1, bb, f) = accO[bb][f];
                    }
            }
            __syncthreads();
            if (kb == 0 && ok) {
#pragma unroll
                for (int bb = 0; bb < MB; ++bb) {
                    const int b = b0 + bb;
                    if (b >= a.nmem) continue;
#pragma unroll
                    for (int f = 0; f < NF; ++f) {
                        const size_t o = (((size_t)jp * n + m) * a.nmem + b) * NF + f;
                        a.E[o] = cadd(accE[bb][f], slot(1, 0, bb, f));
                        a.O[o] = cadd(accO[bb][f], slot(1, 1, bb, f));
                    }
                }
            }
            __syncthreads();
        }
    }
}

}  // namespace aeo
