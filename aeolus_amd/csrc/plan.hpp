// Host-side plan and model objects plus the device views passed to kernels.
#pragma once

#include "common.hpp"
#include "fft.hpp"

namespace aeo {

enum PlanKind { kGauss = 0, kPoints = 1 };

// Device view of a Gauss plan (passed by value to kernels).
struct GaussDev {
    int L, nlat, nlon, nh;
    const double* Ps;         // synthesis table: block m, [l-m][jp]   (jp fastest)
    const double* Pa;         // analysis table:  block m, [jp][l-m]   (l fastest)
    const long long* offs;    // element offset of block m (L+2 entries)
    // the same offset in closed form, nh sum_{m' < m} (L - m' + 1): kernels
    // use it instead of a dependent load of offs[m]
    __host__ __device__ long long block_offset(int m) const {
        return (long long)nh * ((long long)m * (L + 1) - (long long)m * (m - 1) / 2);
    }
    const double* w_rows;     // nlat solid-angle weights per point of a row
    const double* coslat;     // nlat cos(latitude), unclamped
    const double* xn;         // nh northern nodes x_j = cos(colatitude)
    const double2* rec;       // [m][l] (a_lm, b_lm) of the Pbar recurrence (l >= m + 2)
    FftDev fft;
    // Split-precision tables (nslices > 0): INT8 radix-254 digits of Pbar
    // with one power-of-two exponent per table row (see split.hpp).
    int nslices;
    const signed char* Psl;   // synthesis digits, block m: [parity][slice][nhp][kpad_m]
    const long long* offs_ss; // element offset of block m
    const int* esyn;          // row exponents, block m: [parity][nhp] at m * 2 * nhp
    const signed char* Pasl;  // analysis digits, block m: [parity][slice][mpad_m][nhk]
    const long long* offs_sa;
    const int* eana;          // row exponents, block m: [parity][mpad_m] at offs_ea[m]
    const long long* offs_ea;
};

// Device view of a point-set plan.
struct PointsDev {
    int L, B, npts;
    const double* Yk;         // [2][B][npts]: real plane then imaginary plane
    const double2* Yp;        // [npts][B]
    const double* w;          // npts quadrature weights
    const double* coslat;     // npts cos(latitude), unclamped
    // Split-precision tables (nslices > 0): INT8 digits of the real GEMM
    // operands with one exponent per row (see split.hpp).
    //   synthesis  A[i][k'] = (Re Y_ik', -Im Y_i(k'-B)),  Ysl [slice][nprow][K2]
    //   analysis   A[r][i]  = (Re Y_ir, -Im Y_i(r-B)),   Yasl [slice][R2p][Npad]
    int nslices, nprow, K2, R2p, Npad;
    const signed char* Ysl;
    const int* eYs;           // [nprow]
    const signed char* Yasl;
    const int* eYa;           // [R2p]
};

}  // namespace aeo

struct aeo_plan_s {
    int kind = aeo::kGauss;
    int L = 0, B = 0, max_fields = 0;
    // Gauss
    int nlat = 0, nlon = 0, nh = 0;
    double* d_Ps = nullptr;
    double* d_Pa = nullptr;
    double2* d_b1pa = nullptr;     // single-member BVE stage: its analysis-table rows in lane order
    long long* d_offs = nullptr;
    double* d_w_rows = nullptr;
    double* d_coslat = nullptr;
    double* d_xn = nullptr;
    double2* d_rec = nullptr;
    double2* d_tw = nullptr;
    aeo::FftDev fft{};
    int nslices = 0;
    // bumped whenever the split tables are rebuilt (their device addresses
    // change): captured graphs that read them compare it before a replay
    long long tables_gen = 0;
    signed char* d_Psl = nullptr;
    long long* d_offs_ss = nullptr;
    int* d_esyn = nullptr;
    signed char* d_Pasl = nullptr;
    long long* d_offs_sa = nullptr;
    int* d_eana = nullptr;
    long long* d_offs_ea = nullptr;
    // Gauss workspaces: [fields][nh][L+1] complex
    double2* d_E = nullptr;
    double2* d_O = nullptr;
    double2* d_S = nullptr;
    double2* d_D = nullptr;
    // Points
    int npts = 0;
    double* d_Yk = nullptr;
    double2* d_Yp = nullptr;
    double* d_w = nullptr;
    double* d_pcoslat = nullptr;
    int nchunks = 0;
    double* d_gw = nullptr;        // [fields][npts]
    double2* d_part = nullptr;     // [nchunks][fields][B]
    // Spectral constants ((L+1)^2 real each; lap has L+1 entries)
    double* d_Cp = nullptr;
    double* d_Cm = nullptr;
    double* d_CpT = nullptr;       // the same, column-major [m][l] (stage kernels)
    double* d_CmT = nullptr;
    // generic scratch for composed operations: [fields][(L+1)^2] complex
    double2* d_spec = nullptr;
    // Points split-precision tables (nslices > 0; layouts in PointsDev)
    int nprow = 0, K2 = 0, R2p = 0, Npad = 0;
    signed char* d_Ysl = nullptr;
    int* d_eYs = nullptr;
    signed char* d_Yasl = nullptr;
    int* d_eYa = nullptr;

    aeo::GaussDev gdev() const {
        aeo::GaussDev g;
        g.L = L; g.nlat = nlat; g.nlon = nlon; g.nh = nh;
        g.Ps = d_Ps; g.Pa = d_Pa; g.offs = d_offs;
        g.w_rows = d_w_rows; g.coslat = d_coslat; g.xn = d_xn; g.rec = d_rec; g.fft = fft;
        g.nslices = nslices;
        g.Psl = d_Psl; g.offs_ss = d_offs_ss; g.esyn = d_esyn;
        g.Pasl = d_Pasl; g.offs_sa = d_offs_sa; g.eana = d_eana; g.offs_ea = d_offs_ea;
        return g;
    }
    aeo::PointsDev pdev() const {
        aeo::PointsDev p;
        p.L = L; p.B = B; p.npts = npts;
        p.Yk = d_Yk; p.Yp = d_Yp; p.w = d_w; p.coslat = d_pcoslat;
        p.nslices = nslices; p.nprow = nprow; p.K2 = K2; p.R2p = R2p; p.Npad = Npad;
        p.Ysl = d_Ysl; p.eYs = d_eYs; p.Yasl = d_Yasl; p.eYa = d_eYa;
        return p;
    }
};

enum ModelKind { kBVE = 0, kSWE = 1 };

// Kernel classes of one model stage (for aeo_model_profile).
enum KernelClass { kClsSynthesis = 0, kClsGrid = 1, kClsAnalysis = 2 };

// Start/stop events for the next kernel launch of this host thread: armed
// by ProfRec::arm, consumed by launch_k, which then launches with
// hipExtLaunchKernelGGL so the events carry the dispatch's own begin and end
// timestamps (no gap between an event record and the kernel).
struct ArmedEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};
inline ArmedEvents& armed_events() {
    static thread_local ArmedEvents a;
    return a;
}

template <class K, class... Args>
inline void launch_k(K kern, dim3 grid, dim3 block, size_t lds, hipStream_t s, Args... args) {
    ArmedEvents& a = armed_events();
    if (a.start) {
        const hipEvent_t e0 = a.start, e1 = a.stop;
        a = ArmedEvents{};
        hipExtLaunchKernelGGL(kern, grid, block, (std::uint32_t)lds, s, e0, e1, 0u, args...);
    } else {
        hipLaunchKernelGGL(kern, grid, block, lds, s, args...);
    }
}

// Optional timing of every launch of one kernel class (see launch_k).
struct ProfRec {
    int cls = -1;
    int cap = 0;
    int n = 0;
    hipEvent_t* ev = nullptr;  // 2 * cap events
    // the next launch_k of this thread is timed if it belongs to class c
    void arm(int c) {
        armed_events() = ArmedEvents{};
        if (c == cls && n < cap) {
            armed_events() = ArmedEvents{ev[2 * n], ev[2 * n + 1]};
            ++n;
        }
    }
};

struct aeo_model_s {
    int kind = kBVE;
    aeo_plan_s* plan = nullptr;
    int L = 0;
    int nstate = 1;                 // fields in the state (1 BVE, 3 SWE)
    double R = 1.0, omega = 0.0;
    double nu = 0.0;                // BVE viscosity
    double gravity = 0.0, mean_depth = 0.0, phi0 = 0.0, nu4 = 0.0;
    bool has_topo = false;
    // constants
    double* d_nulap = nullptr;      // BVE: nu * laplacian_eig (l = 0 entry nu / R^2)
    double* d_lap = nullptr;        // -l(l+1)/R^2 (l = 0 entry 0)
    double* d_nu4lap2 = nullptr;    // SWE: nu4 * lap^2
    double2* d_lap_phi_s = nullptr; // SWE topography: lap * phi_s
    double2* d_nu4_lap2_phi_s = nullptr;
    // workspaces
    double2* d_E = nullptr;
    double2* d_O = nullptr;
    double2* d_E1 = nullptr;       // single member: plane 1 of the split columns' E/O
    double2* d_O1 = nullptr;
    double2* d_S = nullptr;
    double2* d_D = nullptr;
    double* d_gw = nullptr;
    double2* d_part = nullptr;
    double2* d_ystage = nullptr;   // 2 slots (ping-pong between stages)
    double2* d_ybuf = nullptr;     // 2 slots (ping-pong between steps)
    double2* d_ksum = nullptr;
    double2* d_fint = nullptr;     // forcing in the internal [m][l] layout
    // adjoint workspaces (allocated on the first adjoint call)
    double2* d_aE = nullptr;       // 5 synthesis columns per member
    double2* d_aO = nullptr;
    double2* d_aS = nullptr;       // 4 analysis columns per member
    double2* d_aD = nullptr;
    double2* d_asum = nullptr;
    double2* d_abuf = nullptr;     // 2 slots: step-end / step-start cotangents
    double2* d_kbuf = nullptr;     // 2 slots: stage tendency cotangents
    int nmem = 1;                  // members (independent states) in the group
    int mb = 1;                    // members per workgroup in the Legendre stage kernel;
                                   // 0: the member-batched GEMM form (split_gemm.hip)
    int nslices = 0;               // 0: FP64 contractions; > 0: split INT8 MFMA digits
    // GEMM-form stage workspaces (allocated when those tiles are selected)
    double* d_gT = nullptr;        // analysed tendency terms [m][l][member] complex
    signed char* d_gQ = nullptr;   // synthesis digit columns
    int* d_gQe = nullptr;          // their exponents
    int* d_gkoff = nullptr;        // per-m offsets of the digit blocks
    int gemm_ncp = 0;              // synthesis columns per m (padded)
    int gemm_cg = 64;              // synthesis columns per workgroup (64 or 128)
    int gemm_me = 8;               // members per epilogue workgroup (4 or 8)
    int gemm_ntw = 4;              // 16-column tiles per synthesis wave (2 or 4)
    ProfRec prof;
};

namespace aeo {

// Launchers implemented in gauss.hip / points.hip / spectral.hip.
int gauss_build_tables(aeo_plan_s* p, hipStream_t s);
int gauss_build_split_tables(aeo_plan_s* p, int nslices, hipStream_t s);
int gauss_analysis(aeo_plan_s* p, const double* f, int nf, double2* a, hipStream_t s);
int gauss_synthesis(aeo_plan_s* p, const double2* a, int nf, double* f, hipStream_t s);
int gauss_jacobian(aeo_plan_s* p, double R, const double2* a, const double2* b, double2* out,
                   int dealias, hipStream_t s);
// Synthesis of the model's Legendre fields from states y (all members) into
// the model's E/O workspace (first stage of a run, or a tendency call).
int gauss_model_prime(aeo_model_s* m, const double2* y, hipStream_t s);
// Fourier stage + analysis/epilogue of one RK4 stage (stage -1: tendency into
// kout); with syn_next the same launch also synthesises the next stage.
int gauss_model_stage(aeo_model_s* m, const double2* ys, const double2* y0,
                      const double2* forcing, double2* kout, int stage, double dt,
                      double2* ynext, bool syn_next, hipStream_t s);

// Single-member FP64 BVE stage kernel (bve1.hip): the results of the generic
// stage kernel (to rounding) for models with one member and FP64 contractions.
// ana: analysis + tendency + RK4 epilogue of stage `stage` (-1: tendency into
// kout); syn: synthesis of the resulting (ana) or given (!ana) state into E/O.
bool bve1_supported(const aeo_model_s* md);
// Heavy columns a synthesising launch splits into two degree halves (the
// second half's E/O in d_E1/d_O1; the Jacobian grid stage adds the planes).
int bve1_msplit(const aeo_plan_s* p);
// The single-member stage's lane-order copy of the analysis table
// (p->d_b1pa), built once per plan where the single-member path fits it.
int bve1_prepare(aeo_plan_s* p, hipStream_t s);
// hemi: S/D hold the unweighted row spectra FN/FS ([m][jp]) of the per-row
// Jacobian grid kernel (the stage forms S = w (FN + FS), D = w (FN - FS)).
int bve1_launch(aeo_model_s* md, const double2* ys, const double2* y0, const double2* forcing,
                double2* kout, int stage, double dt, double2* ynext, bool ana, bool syn,
                bool hemi, hipStream_t s);

// Member-batched split-precision BVE stage in GEMM form (split_gemm.hip):
// bitwise the results of the fused stage kernel at the same precision.
// setup allocates its workspaces (once); launch runs the analysis (ana),
// the epilogue and, with syn, the synthesis of the resulting (ana) or given
// (!ana) state.
bool gemm_stage_supported(const aeo_model_s* md);
int gemm_stage_setup(aeo_model_s* md);
int gemm_stage_tiles(aeo_model_s* md, int syn_cols, int epi_members, int syn_ntiles);
int gemm_stage_launch(aeo_model_s* md, const double2* ys, const double2* y0,
                      const double2* forcing, double2* kout, int stage, double dt,
                      double2* ynext, bool ana, bool syn, hipStream_t s);

// Adjoint of the BVE stage (Gauss plans).  prime: synthesis of the forward
// fields of yfwd and of the tendency cotangent kb into the adjoint E/O
// workspace.  stage: Fourier adjoint + analysis and the reverse RK4
// bookkeeping of adjoint stage r (0..3; -1: yb of one tendency into kout);
// with fwd_next set the same launch synthesises the next adjoint stage.
int gauss_adjoint_prime(aeo_model_s* m, const double2* yfwd, const double2* kb, hipStream_t s);
int gauss_adjoint_stage(aeo_model_s* m, const double2* kb, int stage, double dt,
                        const double2* a, double2* anext, double2* kout, double2* knext,
                        const double2* fwd_next, hipStream_t s);

int points_build_tables(aeo_plan_s* p, const double* d_lat, const double* d_lon, hipStream_t s);
int points_build_split_tables(aeo_plan_s* p, int ns, hipStream_t s);
int points_analysis(aeo_plan_s* p, const double* f, int nf, double2* a, hipStream_t s);
int points_synthesis(aeo_plan_s* p, const double2* a, int nf, double* f, hipStream_t s);
int points_jacobian(aeo_plan_s* p, double R, const double2* a, const double2* b, double2* out,
                    int dealias, hipStream_t s);
int points_model_stage(aeo_model_s* m, const double2* ys, const double2* y0,
                       const double2* forcing, double2* kout, int stage, double dt,
                       double2* ynext, hipStream_t s);

int launch_spec_op(int op, int L, double R, const double2* in, double2* out, int nf,
                   hipStream_t s);
// out = alpha * in, n complex entries.
int launch_scale(const double2* in, double2* out, double alpha, size_t n, hipStream_t s);
// out[k] = transpose(in[k]) for nmat (n x n) complex matrices (in != out).
int launch_transpose(const double2* in, double2* out, int n, int nmat, hipStream_t s);
int launch_spectral_diag(int L, double R, double omega, const double2* z, int nf, double* out,
                         hipStream_t s);
int launch_grid_stats(int mode, int npts, const double* w, const double* a, const double* b,
                      const double* c, const double* d, double k0, double k1, double k2,
                      double* out, hipStream_t s);
int launch_velocity_points(int n, int row_len, double R, const double* coslat,
                           const double* g_s, const double* g_lam, double* u, double* v, int mode,
                           hipStream_t s);

}  // namespace aeo
