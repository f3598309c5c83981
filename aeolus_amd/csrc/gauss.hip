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
#include "split.hpp"
#include "stage_launch.hpp"

namespace aeo {

constexpr int kRowElems = 4096;   // complex values per LDS pass of the generic grid kernels

// Rows per LDS pass of the generic grid kernels (the twiddle table takes one row).
__host__ __device__ inline int row_cap(int n) {
    const int c = (kRowElems - n) / lds_row_stride(n);
    return c > 0 ? c : 1;
}

// ---------------------------------------------------------------------------
// FP64 normalised associated-Legendre recurrence (one thread per (m, jp)).
//   Pbar_00 = 1/sqrt(4 pi)
//   Pbar_mm = -sqrt((2m+1)/(2m)) s Pbar_{m-1,m-1}      (Condon-Shortley phase)
//   Pbar_{m+1,m} = sqrt(2m+3) x Pbar_mm
//   Pbar_lm = a_lm (x Pbar_{l-1,m} - b_lm Pbar_{l-2,m})
// with (a_lm, a_lm b_lm) from the plan's recurrence table; rec_step is the exact
// expression the synthesis kernels use to continue the recurrence from two
// table rows, so both produce the same values.
// ---------------------------------------------------------------------------
__global__ void legendre_table_kernel(GaussDev g, double* __restrict__ Ps,
                                      double* __restrict__ Pa) {
    const int m = blockIdx.x;
    const int jp = blockIdx.y * blockDim.x + threadIdx.x;
    const int L = g.L, nh = g.nh;
    if (m > L || jp >= nh) return;
    const double x = g.xn[jp];
    const double s = sqrt((1.0 - x) * (1.0 + x));
    double pmm = 0.28209479177387814347;  // 1/sqrt(4 pi)
    for (int k = 1; k <= m; ++k) pmm *= -sqrt((2.0 * k + 1.0) / (2.0 * k)) * s;
    const int K = L - m + 1;
    double* ps = Ps + g.offs[m];
    double* pa = Pa + g.offs[m];
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
    const double2* rc = g.rec + (size_t)m * (L + 1);
    for (int l = m + 2; l <= L; ++l) {
        const double p = rec_step(rc[l], x, p1, p2);
        ps[(size_t)(l - m) * nh + jp] = p;
        pa[(size_t)jp * K + (l - m)] = p;
        p2 = p1;
        p1 = p;
    }
}

// ---------------------------------------------------------------------------
// Split-precision digit tables (see split.hpp).  grid (L+1, 2 parities,
// row tiles of 64); one thread per table row.
// ---------------------------------------------------------------------------
__global__ void split_table_kernel(GaussDev g, int ns, signed char* __restrict__ Psl,
                                   int* __restrict__ esyn, signed char* __restrict__ Pasl,
                                   int* __restrict__ eana, int nhp, int nhk) {
    const int m = blockIdx.x, p = blockIdx.y;
    const int r = blockIdx.z * blockDim.x + threadIdx.x;
    const int L = g.L, nh = g.nh;
    const int K = L - m + 1;
    const int kp = split_kp(L, m), kpad = split_kpad(L, m), mpad = rup(kp, 16);
    int d[8];
    if (r < nhp) {
        const double* P = g.Ps + g.offs[m] + r;
        double mx = 0.0;
        if (r < nh)
            for (int ip = 0; ip < kp; ++ip) {
                const int i = 2 * ip + p;
                if (i < K) mx = fmax(mx, fabs(P[(size_t)i * nh]));
            }
        const int e = exp_above(mx);
        esyn[((size_t)m * 2 + p) * nhp + r] = e;
        signed char* dst = Psl + g.offs_ss[m];
        for (int ip = 0; ip < kpad; ++ip) {
            const int i = 2 * ip + p;
            const double v = (r < nh && ip < kp && i < K) ? P[(size_t)i * nh] : 0.0;
            digits_rt(v, e, ns, d);
            for (int s = 0; s < ns; ++s)
                dst[((size_t)(p * ns + s) * nhp + r) * kpad + ip] = (signed char)d[s];
        }
    }
    if (r < mpad) {
        const int l = m + 2 * r + p;
        const bool valid = (r < kp) && (l <= L);
        const double* P = g.Pa + g.offs[m] + (l - m);
        double mx = 0.0;
        if (valid)
            for (int j = 0; j < nh; ++j) mx = fmax(mx, fabs(P[(size_t)j * K]));
        const int e = exp_above(mx);
        eana[g.offs_ea[m] + (size_t)p * mpad + r] = e;
        signed char* dst = Pasl + g.offs_sa[m];
        for (int j = 0; j < nhk; ++j) {
            const double v = (valid && j < nh) ? P[(size_t)j * K] : 0.0;
            digits_rt(v, e, ns, d);
            for (int s = 0; s < ns; ++s)
                dst[((size_t)(p * ns + s) * mpad + r) * nhk + j] = (signed char)d[s];
        }
    }
}

// ---------------------------------------------------------------------------
// Fourier stage helpers.  Grid rows live in LDS with the padded layout of
// fft.hpp: row r, longitude k at r * lds_row_stride(n) + lp(k).
// ---------------------------------------------------------------------------
enum SpecMode { kSpecPlain = 0, kSpecDlam = 1, kSpecScale = 2 };

// Packed spectrum of a north/south row pair for one field:
//   Z(b) = sum over m = b (mod n) of [H_N(m) + i H_S(m)]
//        + sum over m = -b (mod n), m >= 1, of [conj G_N(m) + i conj G_S(m)]
// with H(0) = Re G(0), H(m) = G(m), G_N = E + O, G_S = E - O and the
// optional (i m / R) or 1/R factor.  Its unnormalised inverse DFT is
// f_N + i f_S.
__device__ inline void pack_spectrum(double2* Z, int n, int L, const double2* Er,
                                     const double2* Or, size_t ms, int mode, double invR, int t0,
                                     int tstride) {
    // Er[m * ms], Or[m * ms]: even/odd partial sums of one field at wavenumber m;
    // invR = 1/R (the (i m / R) factor is formed as i (m * invR))
    if (n > 2 * L) {
        // no aliasing: bin b holds m = b (b <= L) and m = n - b (1 <= n - b <= L);
        // both loads are clamped into range and issued unconditionally, the
        // masks drop what does not belong to the bin
        for (int b = t0; b < n; b += tstride) {
            const int m1 = (b <= L) ? b : L;
            const int m2 = (b > 0 && n - b <= L) ? n - b : L;
            const double2 e1 = Er[m1 * ms], o1 = Or[m1 * ms];
            const double2 e2 = Er[m2 * ms], o2 = Or[m2 * ms];
            double2 gn = cadd(e1, o1), gs = csub(e1, o1);
            double2 hn = cadd(e2, o2), hs = csub(e2, o2);
            if (mode == kSpecDlam) {
                gn = dlam_q(gn, m1 * invR); gs = dlam_q(gs, m1 * invR);
                hn = dlam_q(hn, m2 * invR); hs = dlam_q(hs, m2 * invR);
            } else if (mode == kSpecScale) {
                gn = cscale(gn, invR); gs = cscale(gs, invR);
                hn = cscale(hn, invR); hs = cscale(hs, invR);
            }
            if (b == 0) { gn.y = 0.0; gs.y = 0.0; }
            double2 z = c2(0.0, 0.0);
            if (b <= L) z = c2(gn.x - gs.y, gn.y + gs.x);
            if (b > 0 && n - b <= L) z = cadd(z, c2(hn.x + hs.y, -hn.y + hs.x));
            Z[lp(b)] = z;
        }
        return;
    }
    for (int b = t0; b < n; b += tstride) {
        double2 z = c2(0.0, 0.0);
        for (int m = b; m <= L; m += n) {
            double2 e = Er[m * ms], o = Or[m * ms];
            double2 gn = cadd(e, o), gs = csub(e, o);
            if (mode == kSpecDlam) { gn = dlam_q(gn, m * invR); gs = dlam_q(gs, m * invR); }
            else if (mode == kSpecScale) { gn = cscale(gn, invR); gs = cscale(gs, invR); }
            if (m == 0) { gn.y = 0.0; gs.y = 0.0; }
            z = cadd(z, c2(gn.x - gs.y, gn.y + gs.x));
        }
        for (int m = (b == 0 ? n : n - b); m <= L; m += n) {
            double2 e = Er[m * ms], o = Or[m * ms];
            double2 gn = cadd(e, o), gs = csub(e, o);
            if (mode == kSpecDlam) { gn = dlam_q(gn, m * invR); gs = dlam_q(gs, m * invR); }
            else if (mode == kSpecScale) { gn = cscale(gn, invR); gs = cscale(gs, invR); }
            z = cadd(z, c2(gn.x + gs.y, -gn.y + gs.x));
        }
        Z[lp(b)] = z;
    }
}

// Split the forward DFT of z = g_N + i g_S into S = w (F_N + F_S) and
// D = w (F_N - F_S) for m = 0..mmax, stored at Sr[m * ms], Dr[m * ms].
__device__ inline void unpack_sd(const double2* Z, int n, int mmax, double w, double2* Sr,
                                 double2* Dr, size_t ms, int t0, int tstride) {
    for (int m = t0; m <= mmax; m += tstride) {
        const int b = m % n;
        const int bn = (n - b) % n;
        const double2 zb = Z[lp(b)];
        const double2 zn = cconj(Z[lp(bn)]);
        const double2 fn = cscale(cadd(zb, zn), 0.5);
        const double2 dd = csub(zb, zn);
        const double2 fs = c2(0.5 * dd.y, -0.5 * dd.x);  // (zb - zn) / (2i)
        Sr[m * ms] = cscale(cadd(fn, fs), w);
        Dr[m * ms] = cscale(csub(fn, fs), w);
    }
}

// Intermediate layout of every grid kernel: E/O/S/D[jp][m][column], one
// column per (member, field) pair, `ncol` columns in all.

// (row pair, member) of this workgroup from a 1-D grid of nh x nmem.  Member
// batches: workgroups are dealt to the 8 XCDs round-robin by id, so the
// decode gives each XCD runs of consecutive members of one row pair, whose
// E/O and S/D columns are adjacent in memory ([jp][m][member][field]) and
// share that XCD's L2; otherwise id = jp + nh * b.
__device__ inline void grid_wg_index(int nh, int nmem, int& jp, int& b) {
    const int id = blockIdx.x;
    if (nmem > 1 && (nh & 7) == 0) {
        const int x = id & 7, k = id >> 3;
        b = k % nmem;
        jp = (k / nmem) * 8 + x;
    } else {
        jp = id % nh;
        b = id / nh;
    }
}

// Row constants of a north/south row pair (clamped cos(latitude) of both
// rows, the row weight; halved for an equator row) through the vector-memory
// path: as scalar loads through a pointer argument they came in a round of
// their own, waited for (with the argument loads) before the kernel's first
// global load.
__device__ inline double vload_f64(const double* p, int i) {
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<double*>(p), (short)0, 0x7fffffff, 0x00020000);
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, i * 8, 0, 0));
}
__device__ inline void row_pair_consts(const GaussDev& g, int jN, int jS, double& cN, double& cS,
                                       double& w) {
    cN = clamp_coslat(vload_f64(g.coslat, jN));
    cS = clamp_coslat(vload_f64(g.coslat, jS));
    const double wr = vload_f64(g.w_rows, jN);
    w = (jN == jS) ? 0.5 * wr : wr;
}

// Prologue of the fused grid kernels: the twiddle table (n entries at twl)
// and the NF even/odd columns of one member of a row pair into LDS, field
// major (sE/sO[f * (L+1) + m], m <= L), so the per-row spectrum packing reads
// consecutive wavenumbers.  Each thread issues all of its loads before its first
// LDS store, so the gather costs one memory latency.
// Columns m < msplit of a single-member synthesis arrive as two planes (the
// degree halves of the split stage workgroups): E = E + E1, O = O + O1.
template <int NF>
__device__ inline void grid_prologue(const GaussDev& g, const double2* __restrict__ E,
                                       const double2* __restrict__ O, size_t eo, size_t ms,
                                       double2* twl, double2* sE, double2* sO, int ntw = -1,
                                       const double2* __restrict__ E1 = nullptr,
                                       const double2* __restrict__ O1 = nullptr,
                                       int msplit = 0) {
    // loads are unconditional (indices clamped in range) so the values stay
    // in registers; only the LDS stores are predicated.  ntw: twiddle entries
    // staged (the first ntw of the table; default the whole table)
    constexpr int PER = 4;
    const int n = (ntw < 0) ? g.nlon : ntw;
    const int tot = (g.L + 1) * NF;
    const int tot1 = msplit * NF;                // plane-1 entries (m < msplit)
    for (int t0 = 0; t0 < tot || t0 < n; t0 += PER * kGridThreads) {
        double2 ve[PER], vo[PER], vt[PER], ve1[PER], vo1[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int t = t0 + threadIdx.x + u * kGridThreads;
            const int te = (t < tot) ? t : tot - 1;
            const int m = te / NF, f = te - m * NF;
            ve[u] = E[eo + m * ms + f];
            vo[u] = O[eo + m * ms + f];
            vt[u] = g.fft.tw[(t < n) ? t : n - 1];
            if (tot1 > 0) {
                const int t1 = (t < tot1) ? t : tot1 - 1;
                const int m1 = t1 / NF, f1 = t1 - m1 * NF;
                ve1[u] = E1[eo + m1 * ms + f1];
                vo1[u] = O1[eo + m1 * ms + f1];
            }
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int t = t0 + threadIdx.x + u * kGridThreads;
            if (t < tot) {
                const int m = t / NF, f = t - m * NF;
                const bool two = t < tot1;
                sE[f * (g.L + 1) + m] = two ? cadd(ve[u], ve1[u]) : ve[u];
                sO[f * (g.L + 1) + m] = two ? cadd(vo[u], vo1[u]) : vo[u];
            }
            if (t < n) twl[t] = vt[u];
        }
    }
}

// Prologue of the single-member power-of-two Jacobian kernel: like
// grid_prologue<4>, but it stores the packed half-spectra of each field's
// north/south rows instead of E/O (field f: d/dlambda for even f, 1/R for odd
// f): Z1[f][m] = the bin-m term and Z2[f][m] = the mirrored bin (n - m) term
// of packed_bin_combine, so a packed bin is Z1[b] + Z2[n - b] (two reads and
// one add instead of four reads and the spectrum arithmetic per bin).
__device__ inline void grid_prologue_z(const GaussDev& g, const double2* __restrict__ E,
                                       const double2* __restrict__ O, size_t eo, size_t ms,
                                       double2* twl, double2* sZ1, double2* sZ2,
                                       const double2* __restrict__ E1,
                                       const double2* __restrict__ O1, int msplit,
                                       double invR) {
    constexpr int NF = 4, PER = 4;
    const int n = g.nlon;
    const int tot = (g.L + 1) * NF;
    const int tot1 = msplit * NF;
    for (int t0 = 0; t0 < tot || t0 < n; t0 += PER * kGridThreads) {
        double2 ve[PER], vo[PER], vt[PER], ve1[PER], vo1[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int t = t0 + threadIdx.x + u * kGridThreads;
            const int te = (t < tot) ? t : tot - 1;
            const int m = te / NF, f = te - m * NF;
            ve[u] = E[eo + m * ms + f];
            vo[u] = O[eo + m * ms + f];
            vt[u] = g.fft.tw[(t < n) ? t : n - 1];
            if (tot1 > 0) {
                const int t1 = (t < tot1) ? t : tot1 - 1;
                const int m1 = t1 / NF, f1 = t1 - m1 * NF;
                ve1[u] = E1[eo + m1 * ms + f1];
                vo1[u] = O1[eo + m1 * ms + f1];
            }
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int t = t0 + threadIdx.x + u * kGridThreads;
            if (t < tot) {
                const int m = t / NF, f = t - m * NF;
                const bool two = t < tot1;
                const double2 e = two ? cadd(ve[u], ve1[u]) : ve[u];
                const double2 o = two ? cadd(vo[u], vo1[u]) : vo[u];
                double2 gn = cadd(e, o), gs = csub(e, o);
                if (f & 1) {
                    gn = cscale(gn, invR);
                    gs = cscale(gs, invR);
                } else {
                    gn = dlam_q(gn, m * invR);
                    gs = dlam_q(gs, m * invR);
                }
                sZ2[f * (g.L + 1) + m] = c2(gn.x + gs.y, -gn.y + gs.x);
                if (m == 0) { gn.y = 0.0; gs.y = 0.0; }
                sZ1[f * (g.L + 1) + m] = c2(gn.x - gs.y, gn.y + gs.x);
            }
            if (t < n) twl[t] = vt[u];
        }
    }
}

// Synthesis of nf fields (columns): E/O -> grid rows.  grid (nh, field
// batches of `cap` rows, cap <= row_cap); block 256.
template <bool P2>
__global__ __launch_bounds__(kGridThreads) void grid_syn_kernel(GaussDev g,
                                                               const double2* __restrict__ E,
                                                               const double2* __restrict__ O,
                                                               int nf, int cap,
                                                               double* __restrict__ f_out) {
    extern __shared__ double2 lds[];
    const int jp = blockIdx.x;
    const int n = g.nlon, L = g.L, rs = lds_row_stride(n);
    const int jN = jp, jS = g.nlat - 1 - jp;
    const size_t row = (size_t)jp * (L + 1) * nf;
    const double2* twl = lds + (size_t)cap * rs;
    fft_twiddles_to_lds(g.fft, lds + (size_t)cap * rs);
    for (int f0 = blockIdx.y * cap; f0 < nf; f0 += cap * gridDim.y) {
        const int nr = (nf - f0 < cap) ? nf - f0 : cap;
        for (int r = 0; r < nr; ++r) {
            pack_spectrum(lds + r * rs, n, L, E + row + f0 + r, O + row + f0 + r, nf, kSpecPlain,
                          1.0, threadIdx.x, kGridThreads);
        }
        __syncthreads();
        fft_rows<+1, P2>(lds, nr, g.fft, twl);
        for (int t = threadIdx.x; t < nr * n; t += kGridThreads) {
            const int r = t / n, k = t - r * n;
            const double2 z = lds[r * rs + lp(k)];
            double* fo = f_out + (size_t)(f0 + r) * g.nlat * n;
            fo[(size_t)jN * n + k] = z.x;
            if (jS != jN) fo[(size_t)jS * n + k] = z.y;
        }
        __syncthreads();
    }
}

// Analysis of nf grid fields: rows -> S/D (m <= mmax).  grid (nh, field
// batches of `cap` rows, cap <= row_cap); block 256.
template <bool P2>
__global__ __launch_bounds__(kGridThreads) void grid_ana_kernel(GaussDev g,
                                                               const double* __restrict__ f_in,
                                                               int nf, int cap, int mmax,
                                                               double2* __restrict__ S,
                                                               double2* __restrict__ D) {
    extern __shared__ double2 lds[];
    const int jp = blockIdx.x;
    const int n = g.nlon, L = g.L, rs = lds_row_stride(n);
    const int jN = jp, jS = g.nlat - 1 - jp;
    const size_t row = (size_t)jp * (L + 1) * nf;
    const double w = (jN == jS) ? 0.5 * g.w_rows[jN] : g.w_rows[jN];
    const double2* twl = lds + (size_t)cap * rs;
    fft_twiddles_to_lds(g.fft, lds + (size_t)cap * rs);
    for (int f0 = blockIdx.y * cap; f0 < nf; f0 += cap * gridDim.y) {
        const int nr = (nf - f0 < cap) ? nf - f0 : cap;
        for (int t = threadIdx.x; t < nr * n; t += kGridThreads) {
            const int r = t / n, k = t - r * n;
            const double* fi = f_in + (size_t)(f0 + r) * g.nlat * n;
            lds[r * rs + lp(k)] = c2(fi[(size_t)jN * n + k], fi[(size_t)jS * n + k]);
        }
        __syncthreads();
        fft_rows<-1, P2>(lds, nr, g.fft, twl);
        for (int r = 0; r < nr; ++r)
            unpack_sd(lds + r * rs, n, mmax, w, S + row + f0 + r, D + row + f0 + r, nf,
                      threadIdx.x, kGridThreads);
        __syncthreads();
    }
}

// Jacobian: fields (a, S a, b, S b) -> J = (a_s b_l - a_l b_s)/c^2 -> S/D.
// grid (nh, members); E/O columns (member, 4 fields), S/D columns (member).
template <bool P2>
__global__ __launch_bounds__(kGridThreads) void grid_jac_kernel(GaussDev g, const double2* E,
                                                               const double2* O, double R,
                                                               int mmax, int nmem, double2* S,
                                                               double2* D, const double2* E1,
                                                               const double2* O1, int msplit) {
    extern __shared__ double2 lds[];
    int jp, b;
    grid_wg_index(g.nh, nmem, jp, b);
    const int n = g.nlon, L = g.L, rs = lds_row_stride(n);
    const int jN = jp, jS = g.nlat - 1 - jp;
    const double invR = 1.0 / R;
    const size_t eo = ((size_t)jp * (L + 1) * nmem + b) * 4;
    const size_t ms = (size_t)nmem * 4;
    // row constants are loaded first: no later phase waits on global memory
    const double cN = clamp_coslat(g.coslat[jN]);
    const double cS = clamp_coslat(g.coslat[jS]);
    const double w = (jN == jS) ? 0.5 * g.w_rows[jN] : g.w_rows[jN];
    // LDS: rows [4][rs] | twiddles [n] | E, O columns [4][L+1] each
    double2* sE = lds + 4 * rs + n;
    double2* sO = sE + (L + 1) * 4;
    const double2* twl = lds + 4 * rs;
    grid_prologue<4>(g, E, O, eo, ms, lds + 4 * rs, sE, sO, -1, E1, O1, msplit);
    __syncthreads();
    // rows: 0 a_l, 1 a_s, 2 b_l, 3 b_s
    for (int r = 0; r < 4; ++r) {
        pack_spectrum(lds + r * rs, n, L, sE + r * (L + 1), sO + r * (L + 1), 1,
                      (r & 1) ? kSpecScale : kSpecDlam, invR, threadIdx.x, kGridThreads);
    }
    __syncthreads();
    fft_rows<+1, P2>(lds, 4, g.fft, twl);
    const double qN = 1.0 / (cN * cN), qS = 1.0 / (cS * cS);
    for (int k = threadIdx.x; k < n; k += kGridThreads) {
        const int e = lp(k);
        const double2 al = lds[e], as = lds[rs + e], bl = lds[2 * rs + e], bs = lds[3 * rs + e];
        const double jn = jac_point_q(al.x, as.x, bl.x, bs.x, qN);
        const double js = jac_point_q(al.y, as.y, bl.y, bs.y, qS);
        lds[e] = c2(jn, js);
    }
    __syncthreads();
    fft_rows<-1, P2>(lds, 1, g.fft, twl);
    const size_t so = (size_t)jp * (L + 1) * nmem + b;
    unpack_sd(lds, n, mmax, w, S + so, D + so, (size_t)nmem, threadIdx.x, kGridThreads);
}

// LDS rows of the split grid kernels: the larger SWE half transforms 9
// grid rows, each adjoint half 3.
constexpr int kSweLdsRows = 9;
constexpr int kAdjLdsRows = 3;

// Shallow water: 10 Legendre fields -> 12 grid rows -> 4 products -> S/D.
// grid (nh, members); E/O columns (member, 10 fields), S/D columns (member, 4).
template <bool P2>
__global__ __launch_bounds__(kGridThreads) void grid_swe_kernel(GaussDev g, const double2* E,
                                                               const double2* O, double R,
                                                               int mmax, int nmem, double2* S,
                                                               double2* D) {
    extern __shared__ double2 lds[];
    // the arguments into scalar registers at entry, in one batch
    asm volatile("" ::"s"(g.nh), "s"(g.nlat), "s"(g.L), "s"(g.nlon), "s"(g.coslat),
                 "s"(g.w_rows), "s"(g.fft.tw), "s"(E), "s"(O), "s"(R), "s"(mmax), "s"(nmem),
                 "s"(S), "s"(D));
    int jp, b;
    grid_wg_index(g.nh, nmem, jp, b);
    // two workgroups per (row pair, member): half 0 forms the products
    // 0, 1 (eta flux divergence and curl) from 9 grid rows, half 1 the
    // products 2, 3 (kinetic energy, phi flux divergence) from 8; both
    // transform the wind rows 0-3 themselves
    const int half = blockIdx.y;
    const int n = g.nlon, L = g.L, rs = lds_row_stride(n);
    const int jN = jp, jS = g.nlat - 1 - jp;
    const double invR = 1.0 / R;
    const size_t eo = ((size_t)jp * (L + 1) * nmem + b) * 10;
    const size_t ms = (size_t)nmem * 10;
    double cN, cS, w;
    row_pair_consts(g, jN, jS, cN, cS, w);
    // grid row r <- (Legendre field, spectral factor)
    constexpr int src[12] = {0, 1, 2, 3, 4, 5, 4, 6, 7, 6, 8, 9};
    constexpr int mode[12] = {kSpecDlam, kSpecScale, kSpecDlam, kSpecScale, kSpecDlam,
                              kSpecScale, kSpecPlain, kSpecDlam, kSpecScale, kSpecPlain,
                              kSpecPlain, kSpecPlain};
    // grid rows of each half, in LDS row order
    constexpr int rows0[9] = {0, 1, 2, 3, 4, 5, 6, 10, 11};
    constexpr int rows1[9] = {0, 1, 2, 3, 7, 8, 9, 11, 11};
    const int nr = half == 0 ? 9 : 8;
    // LDS: rows [9][rs] (the larger half's grid rows) | twiddles [n] | E, O
    // columns [10][L+1] each
    double2* sE = lds + kSweLdsRows * rs + n;
    double2* sO = sE + (L + 1) * 10;
    const double2* twl = lds + kSweLdsRows * rs;
    grid_prologue<10>(g, E, O, eo, ms, lds + kSweLdsRows * rs, sE, sO);
    __syncthreads();
    for (int i = 0; i < nr; ++i) {
        // LDS row i <- grid row rows0[i] (half 0) or rows1[i] (half 1)
        const int r = half == 0 ? (i < 7 ? i : i + 3) : (i < 4 ? i : (i < 7 ? i + 3 : 11));
        pack_spectrum(lds + i * rs, n, L, sE + src[r] * (L + 1), sO + src[r] * (L + 1), 1, mode[r],
                      invR, threadIdx.x, kGridThreads);
    }
    __syncthreads();
    fft_rows<+1, P2>(lds, nr, g.fft, twl);
    const double icN = 1.0 / cN, icS = 1.0 / cS;
    for (int k = threadIdx.x; k < n; k += kGridThreads) {
        const int e = lp(k);
        double gn[12], gs[12], on[4], os[4];
#pragma unroll
        for (int r = 0; r < 12; ++r) gn[r] = gs[r] = 0.0;
        // constant row indices in each branch keep gn / gs in registers
        if (half == 0) {
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                const double2 z = lds[i * rs + e];
                gn[rows0[i]] = z.x;
                gs[rows0[i]] = z.y;
            }
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const double2 z = lds[i * rs + e];
                gn[rows1[i]] = z.x;
                gs[rows1[i]] = z.y;
            }
        }
        swe_point_ic(gn, icN, on);
        swe_point_ic(gs, icS, os);
        const double2 p0 = half == 0 ? c2(on[0], os[0]) : c2(on[2], os[2]);
        const double2 p1 = half == 0 ? c2(on[1], os[1]) : c2(on[3], os[3]);
        lds[e] = p0;
        lds[rs + e] = p1;
    }
    __syncthreads();
    fft_rows<-1, P2>(lds, 2, g.fft, twl);
    const size_t so = ((size_t)jp * (L + 1) * nmem + b) * 4;
    for (int r = 0; r < 2; ++r)
        unpack_sd(lds + r * rs, n, mmax, w, S + so + 2 * half + r, D + so + 2 * half + r,
                  (size_t)nmem * 4, threadIdx.x, kGridThreads);
}

// Adjoint of the BVE Jacobian on one north/south row pair.  E/O columns
// (member, 5 fields): psi, S psi, eta, S eta of the forward stage and the
// Jacobian cotangent input jb; grid rows after the inverse FFT
//   0 a_l, 1 a_s, 2 b_l, 3 b_s (as in grid_jac_kernel), 4 jt.
// The cotangents of the four rows are jt times the partial derivatives of
// J = (a_s b_l - a_l b_s)/c^2 (the quadrature weight of Ana^T cancels the
// unweighted Syn^T, leaving a weighted analysis):
//   0: -jt b_s/c^2,  1: jt b_l/c^2,  2: jt a_s/c^2,  3: -jt a_l/c^2,
// analysed for m <= L into S/D columns (member, 4).  grid (nh, members).
template <bool P2>
__global__ __launch_bounds__(kGridThreads) void grid_adj_kernel(GaussDev g, const double2* E,
                                                               const double2* O, double R,
                                                               int nmem, double2* S,
                                                               double2* D) {
    extern __shared__ double2 lds[];
    int jp, b;
    grid_wg_index(g.nh, nmem, jp, b);
    const int n = g.nlon, L = g.L, rs = lds_row_stride(n);
    const int jN = jp, jS = g.nlat - 1 - jp;
    const double invR = 1.0 / R;
    const size_t eo = ((size_t)jp * (L + 1) * nmem + b) * 5;
    const size_t ms = (size_t)nmem * 5;
    const double cN = clamp_coslat(g.coslat[jN]);
    const double cS = clamp_coslat(g.coslat[jS]);
    const double w = (jN == jS) ? 0.5 * g.w_rows[jN] : g.w_rows[jN];
    // LDS: rows [3][rs] (each half transforms 3 grid rows) | twiddles [n] |
    // E, O columns [5][L+1] each
    double2* sE = lds + kAdjLdsRows * rs + n;
    double2* sO = sE + (L + 1) * 5;
    const double2* twl = lds + kAdjLdsRows * rs;
    grid_prologue<5>(g, E, O, eo, ms, lds + kAdjLdsRows * rs, sE, sO);
    __syncthreads();
    // two workgroups per (row pair, member): half 0 forms cotangents 0, 1
    // from rows b_l, b_s, jt; half 1 cotangents 2, 3 from a_l, a_s, jt
    const int half = blockIdx.y;
    for (int i = 0; i < 3; ++i) {
        const int r = half == 0 ? i + 2 : (i < 2 ? i : 4);
        const int mode = (r == 4) ? kSpecPlain : ((r & 1) ? kSpecScale : kSpecDlam);
        pack_spectrum(lds + i * rs, n, L, sE + r * (L + 1), sO + r * (L + 1), 1, mode, invR,
                      threadIdx.x, kGridThreads);
    }
    __syncthreads();
    fft_rows<+1, P2>(lds, 3, g.fft, twl);
    const double qN = 1.0 / (cN * cN), qS = 1.0 / (cS * cS);
    for (int k = threadIdx.x; k < n; k += kGridThreads) {
        const int e = lp(k);
        const double2 xl = lds[e], xs = lds[rs + e];      // (b_l, b_s) or (a_l, a_s)
        const double2 jt = lds[2 * rs + e];
        const double tN = jt.x * qN, tS = jt.y * qS;
        if (half == 0) {
            lds[e] = c2(-tN * xs.x, -tS * xs.y);
            lds[rs + e] = c2(tN * xl.x, tS * xl.y);
        } else {
            lds[e] = c2(tN * xs.x, tS * xs.y);
            lds[rs + e] = c2(-tN * xl.x, -tS * xl.y);
        }
    }
    __syncthreads();
    fft_rows<-1, P2>(lds, 2, g.fft, twl);
    const size_t so = ((size_t)jp * (L + 1) * nmem + b) * 4;
    for (int r = 0; r < 2; ++r)
        unpack_sd(lds + r * rs, n, L, w, S + so + 2 * half + r, D + so + 2 * half + r,
                  (size_t)nmem * 4, threadIdx.x, kGridThreads);
}

// ---------------------------------------------------------------------------
// Jacobian grid stage for power-of-two rows n = 2^LGN <= 512 with every index
// a compile-time constant: the same arithmetic as grid_jac_kernel (the same
// packed spectra, twiddles, butterflies, products and unpacking), with the
// spectrum packing of the four inverse rows fused into the first radix-8
// stage (each thread packs the eight bins it transforms straight from the
// staged E/O columns, so the packed rows never pass through LDS).
// ---------------------------------------------------------------------------
// Twiddle W^t = e^{DIR 2 pi i t / N}, t in [0, N), from the quarter table
// tw4[r] = e^{-2 pi i r / N}, r < N/4: W^(r + q N/4) is tw4[r] rotated by q
// quarter turns (exact: a swap for odd q, a sign flip for q >= 2),
// conjugated for DIR = +1.  (Written with a four-way conditional choosing
// among double2 values instead, this gave wrong twiddles inside the grid
// kernel on gfx950 while the same function alone was right.)
template <int DIR, int LGN>
__device__ inline double2 rotate_q(double2 w, int t) {
    const int q = t >> (LGN - 2);
    double x = (q & 1) ? w.y : w.x;
    double y = (q & 1) ? -w.x : w.y;
    if (q & 2) {
        x = -x;
        y = -y;
    }
    if (DIR > 0) y = -y;
    return c2(x, y);
}
template <int DIR, int LGN>
__device__ inline double2 twiddle_q(const double2* tw4, int t) {
    return rotate_q<DIR, LGN>(tw4[t & ((1 << (LGN - 2)) - 1)], t);
}

// QT: tw is the quarter table of twiddle_q, else the whole table.
template <int DIR, int R, int LGR, int LGN, int LGNS, int NROWS, bool QT>
__device__ inline void fft_stage_c(double2* buf, const double2* tw) {
    constexpr int N = 1 << LGN, LGNB = LGN - LGR, NB = 1 << LGNB, TOT = NROWS * NB;
    constexpr int RS = N + (N >> 3), LGSPAN = LGN - LGNS - LGR;
    constexpr int MAXB = (TOT + kGridThreads - 1) / kGridThreads;
    // every data and table read of the stage is issued before the first
    // product (one LDS latency per stage instead of one per radix input)
    double2 v[MAXB][R], w[MAXB][R];
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int t = threadIdx.x + b * kGridThreads;
        if (t < TOT) {
            const int row = t >> LGNB, j = t & (NB - 1), k = j & ((1 << LGNS) - 1);
            const double2* rb = buf + row * RS;
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = rb[lp(j + r * NB)];
            // no k > 0 test: W^0 = (1, 0) leaves x unchanged
            if (LGNS > 0) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const int e = (r * k) << LGSPAN;
                    w[b][r] = QT ? tw[e & ((N >> 2) - 1)] : tw[e];
                }
            }
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    if (LGNS > 0) {
#pragma unroll
        for (int b = 0; b < MAXB; ++b) {
            const int t = threadIdx.x + b * kGridThreads;
            if (t < TOT) {
                const int k = (t & (NB - 1)) & ((1 << LGNS) - 1);
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    // explicit fma: both table forms give the same twiddle
                    // bits, and the product must round alike in both instances
                    const int e = (r * k) << LGSPAN;
                    double2 wr = w[b][r];
                    if (QT) {
                        wr = rotate_q<DIR, LGN>(wr, e);
                    } else if (DIR > 0) {
                        wr.y = -wr.y;
                    }
                    const double2 x = v[b][r];
                    v[b][r] = c2(fma(x.x, wr.x, -(x.y * wr.y)), fma(x.x, wr.y, x.y * wr.x));
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int t = threadIdx.x + b * kGridThreads;
        if (t < TOT) {
            const int row = t >> LGNB, j = t & (NB - 1), k = j & ((1 << LGNS) - 1);
            butterfly<DIR, R>(v[b]);
            double2* rb = buf + row * RS;
            const int base = ((j >> LGNS) << (LGNS + LGR)) + k;
#pragma unroll
            for (int q = 0; q < R; ++q) rb[lp(base + (q << LGNS))] = v[b][q];
        }
    }
    __syncthreads();
}

// Stages after the first: radix 8 while at least 3 bits remain, then 4 or 2.
template <int DIR, int LGN, int LGNS, int NROWS, bool QT>
__device__ inline void fft_rest_c(double2* buf, const double2* tw) {
    if constexpr (LGN - LGNS >= 3) {
        fft_stage_c<DIR, 8, 3, LGN, LGNS, NROWS, QT>(buf, tw);
        fft_rest_c<DIR, LGN, LGNS + 3, NROWS, QT>(buf, tw);
    } else if constexpr (LGN - LGNS == 2) {
        fft_stage_c<DIR, 4, 2, LGN, LGNS, NROWS, QT>(buf, tw);
    } else if constexpr (LGN - LGNS == 1) {
        fft_stage_c<DIR, 2, 1, LGN, LGNS, NROWS, QT>(buf, tw);
    }
}

// One radix-R Stockham stage from rows `src` to rows `dst` (two LDS
// buffers): no barrier between the reads and the writes, one after the
// writes (the in-place fft_stage_c needs both).  `tw` is the per-row
// kernel's stage-major table (hemi_tw).
template <int DIR, int R, int LGR, int LGN, int LGNS, int NROWS>
__device__ inline void fft_stage_pp(const double2* src, double2* dst, const double2* tw) {
    constexpr int N = 1 << LGN, LGNB = LGN - LGR, NB = 1 << LGNB, TOT = NROWS * NB;
    constexpr int RS = N + (N >> 3);
    constexpr int MAXB = (TOT + kGridThreads - 1) / kGridThreads;
    constexpr int TWO = LGNS > 0 ? hemi_tw_off(LGN, LGR, LGNS) : 0;
    static_assert(TWO >= 0, "a stage with a twiddle block");
    double2 v[MAXB][R], w[MAXB][R];
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int t = threadIdx.x + b * kGridThreads;
        if (t < TOT) {
            const int row = t >> LGNB, j = t & (NB - 1), k = j & ((1 << LGNS) - 1);
            const double2* rb = src + row * RS;
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = rb[lp(j + r * NB)];
            if (LGNS > 0) {
#pragma unroll
                for (int r = 1; r < R; ++r) w[b][r] = tw[TWO + ((r - 1) << LGNS) + k];
            }
        }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int t = threadIdx.x + b * kGridThreads;
        if (t < TOT) {
            const int row = t >> LGNB, j = t & (NB - 1), k = j & ((1 << LGNS) - 1);
            if (LGNS > 0) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    double2 wr = w[b][r];
                    if (DIR > 0) wr.y = -wr.y;
                    const double2 x = v[b][r];
                    v[b][r] = c2(fma(x.x, wr.x, -(x.y * wr.y)), fma(x.x, wr.y, x.y * wr.x));
                }
            }
            butterfly<DIR, R>(v[b]);
            double2* rb = dst + row * RS;
            const int base = ((j >> LGNS) << (LGNS + LGR)) + k;
#pragma unroll
            for (int q = 0; q < R; ++q) rb[lp(base + (q << LGNS))] = v[b][q];
        }
    }
    __syncthreads();
}

// The last Stockham stage of one row (NROWS 1), its outputs m <= mmax
// stored straight to global memory, out[m * ostride], instead of LDS;
// `tw` is the stage-major table (hemi_tw).
template <int DIR, int R, int LGR, int LGN, int LGNS>
__device__ inline void fft_stage_out(const double2* src, const double2* tw, double2* out,
                                     int ostride, int mmax) {
    static_assert(LGNS + LGR == LGN, "the last stage");
    constexpr int N = 1 << LGN, NB = 1 << (LGN - LGR);
    constexpr int TWO = LGNS > 0 ? hemi_tw_off(LGN, LGR, LGNS) : 0;
    static_assert(TWO >= 0, "a stage with a twiddle block");
    static_assert(NB <= kGridThreads, "one butterfly per thread");
    const int j = threadIdx.x;
    if (j < NB) {
        double2 v[R], w[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = src[lp(j + r * NB)];
        if (LGNS > 0) {
#pragma unroll
            for (int r = 1; r < R; ++r) w[r] = tw[TWO + ((r - 1) << LGNS) + j];
        }
        __builtin_amdgcn_sched_barrier(0);
        if (LGNS > 0) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                double2 wr = w[r];
                if (DIR > 0) wr.y = -wr.y;
                const double2 x = v[r];
                v[r] = c2(fma(x.x, wr.x, -(x.y * wr.y)), fma(x.x, wr.y, x.y * wr.x));
            }
        }
        butterfly<DIR, R>(v);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int m = j + (q << LGNS);           // j < 2^LGNS: base = j
            if (m <= mmax) out[(size_t)m * ostride] = v[q];
        }
    }
    (void)N;
}

// fft_rest_pp of one row whose last stage stores outputs m <= mmax to
// global memory (fft_stage_out).
template <int DIR, int LGN, int LGNS>
__device__ inline void fft_rest_pp_out(double2* a, double2* b, const double2* tw, double2* out,
                                       int ostride, int mmax) {
    if constexpr (LGN - LGNS > 3) {
        fft_stage_pp<DIR, 8, 3, LGN, LGNS, 1>(a, b, tw);
        fft_rest_pp_out<DIR, LGN, LGNS + 3>(b, a, tw, out, ostride, mmax);
    } else if constexpr (LGN - LGNS == 3) {
        fft_stage_out<DIR, 8, 3, LGN, LGNS>(a, tw, out, ostride, mmax);
    } else if constexpr (LGN - LGNS == 2) {
        fft_stage_out<DIR, 4, 2, LGN, LGNS>(a, tw, out, ostride, mmax);
    } else if constexpr (LGN - LGNS == 1) {
        fft_stage_out<DIR, 2, 1, LGN, LGNS>(a, tw, out, ostride, mmax);
    }
}

// Packed spectrum bin b of one row (pack_spectrum, case n > 2L), in two
// parts so a caller can issue the LDS reads of many bins before combining:
// packed_bin_load gathers E/O at m1 = b and m2 = n - b, packed_bin_combine
// forms the bin.
struct PackedBinIn {
    double2 e1, o1, e2, o2;
};
__device__ inline PackedBinIn packed_bin_load(int b, int n, int L, const double2* Er,
                                              const double2* Or) {
    const int m1 = (b <= L) ? b : L;
    const int m2 = (b > 0 && n - b <= L) ? n - b : L;
    return PackedBinIn{Er[m1], Or[m1], Er[m2], Or[m2]};
}
__device__ inline double2 packed_bin_combine(const PackedBinIn& in, int b, int n, int L, int mode,
                                             double invR) {
    const int m1 = (b <= L) ? b : L;
    const int m2 = (b > 0 && n - b <= L) ? n - b : L;
    double2 gn = cadd(in.e1, in.o1), gs = csub(in.e1, in.o1);
    double2 hn = cadd(in.e2, in.o2), hs = csub(in.e2, in.o2);
    if (mode == kSpecDlam) {
        gn = dlam_q(gn, m1 * invR); gs = dlam_q(gs, m1 * invR);
        hn = dlam_q(hn, m2 * invR); hs = dlam_q(hs, m2 * invR);
    } else if (mode == kSpecScale) {
        gn = cscale(gn, invR); gs = cscale(gs, invR);
        hn = cscale(hn, invR); hs = cscale(hs, invR);
    }
    if (b == 0) { gn.y = 0.0; gs.y = 0.0; }
    double2 z = c2(0.0, 0.0);
    if (b <= L) z = c2(gn.x - gs.y, gn.y + gs.x);
    if (b > 0 && n - b <= L) z = cadd(z, c2(hn.x + hs.y, -hn.y + hs.x));
    return z;
}

template <int LGN, bool COMPACT>
__global__ __launch_bounds__(kGridThreads) void grid_jac_p2_kernel(GaussDev g, const double2* E,
                                                                  const double2* O, double R,
                                                                  int mmax, int nmem, double2* S,
                                                                  double2* D, const double2* E1,
                                                                  const double2* O1, int msplit) {
    constexpr int N = 1 << LGN, RS = N + (N >> 3), NB = N >> 3;
    extern __shared__ double2 lds[];
    // the arguments into scalar registers at entry, in one batch
    asm volatile("" ::"s"(g.nh), "s"(g.nlat), "s"(g.L), "s"(g.coslat), "s"(g.w_rows),
                 "s"(g.fft.tw), "s"(E), "s"(O), "s"(R), "s"(mmax), "s"(nmem), "s"(S), "s"(D),
                 "s"(E1), "s"(O1), "s"(msplit));
    int jp, b;
    grid_wg_index(g.nh, nmem, jp, b);
    const int L = g.L;
    const int jN = jp, jS = g.nlat - 1 - jp;
    const double invR = 1.0 / R;
    const size_t eo = ((size_t)jp * (L + 1) * nmem + b) * 4;
    const size_t ms = (size_t)nmem * 4;
    double cN, cS, w;
    row_pair_consts(g, jN, jS, cN, cS, w);
    // LDS: rows [4][RS] | twiddle table [N] | E, O columns [4][L+1] each.
    // COMPACT (member batches: more workgroups per CU): a quarter twiddle
    // table [N/4], and the E/O columns (8 (L+1) <= 4 RS as N > 2L) alias the
    // rows, which the first stage writes only after reading all of them
    double2* twl = lds + 4 * RS;
    double2* sE = COMPACT ? lds : twl + N;      // !COMPACT: Z1 of grid_prologue_z
    double2* sO = sE + (L + 1) * 4;             // !COMPACT: Z2
    if constexpr (COMPACT)
        grid_prologue<4>(g, E, O, eo, ms, twl, sE, sO, N / 4, E1, O1, msplit);
    else
        grid_prologue_z(g, E, O, eo, ms, twl, sE, sO, E1, O1, msplit, invR);
    __syncthreads();
    // inverse stage 1 (radix 8, no twiddles) fed by the packed bins:
    // rows 0 a_l, 1 a_s, 2 b_l, 3 b_s
    constexpr int NBT = (4 * NB + kGridThreads - 1) / kGridThreads;
    static_assert(NBT <= 2, "two first-stage butterflies per thread at most");
    double2 v[NBT][8];
    if constexpr (COMPACT) {
        // member batches: bin by bin (registers kept for occupancy)
#pragma unroll
        for (int bt = 0; bt < NBT; ++bt) {
            const int t = threadIdx.x + bt * kGridThreads;
            if (t < 4 * NB) {
                const int row = t / NB, j = t % NB;
                const int mode = (row & 1) ? kSpecScale : kSpecDlam;
#pragma unroll
                for (int r = 0; r < 8; ++r)
                    v[bt][r] = packed_bin_combine(
                        packed_bin_load(j + r * NB, N, L, sE + row * (L + 1), sO + row * (L + 1)),
                        j + r * NB, N, L, mode, invR);
                butterfly<+1, 8>(v[bt]);
            }
        }
    } else {
        // single member (one workgroup per CU): the packed half-spectra of
        // the prologue, every read issued before the first add
        double2 z1[NBT][8], z2[NBT][8];
#pragma unroll
        for (int bt = 0; bt < NBT; ++bt) {
            const int t = threadIdx.x + bt * kGridThreads;
            if (t < 4 * NB) {
                const int row = t / NB, j = t % NB;
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    const int b = j + r * NB;
                    const int m1 = (b <= L) ? b : L;
                    const int m2 = (b > 0 && N - b <= L) ? N - b : L;
                    z1[bt][r] = sE[row * (L + 1) + m1];
                    z2[bt][r] = sO[row * (L + 1) + m2];
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int bt = 0; bt < NBT; ++bt) {
            const int t = threadIdx.x + bt * kGridThreads;
            if (t < 4 * NB) {
                const int j = t % NB;
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    const int b = j + r * NB;
                    double2 z = c2(0.0, 0.0);
                    if (b <= L) z = z1[bt][r];
                    if (b > 0 && N - b <= L) z = cadd(z, z2[bt][r]);
                    v[bt][r] = z;
                }
                butterfly<+1, 8>(v[bt]);
            }
        }
    }
    if constexpr (COMPACT) __syncthreads();   // every E/O read before the rows are written
#pragma unroll
    for (int bt = 0; bt < NBT; ++bt) {
        const int t = threadIdx.x + bt * kGridThreads;
        if (t < 4 * NB) {
            const int row = t / NB, j = t % NB;
            double2* rb = lds + row * RS;
#pragma unroll
            for (int q = 0; q < 8; ++q) rb[lp(8 * j + q)] = v[bt][q];
        }
    }
    __syncthreads();
    fft_rest_c<+1, LGN, 3, 4, COMPACT>(lds, twl);
    const double qN = 1.0 / (cN * cN), qS = 1.0 / (cS * cS);
#pragma unroll
    for (int k = threadIdx.x; k < N; k += kGridThreads) {
        const int e = lp(k);
        const double2 al = lds[e], as = lds[RS + e], bl = lds[2 * RS + e], bs = lds[3 * RS + e];
        const double jn = jac_point_q(al.x, as.x, bl.x, bs.x, qN);
        const double js = jac_point_q(al.y, as.y, bl.y, bs.y, qS);
        lds[e] = c2(jn, js);
    }
    __syncthreads();
    fft_stage_c<-1, 8, 3, LGN, 0, 1, COMPACT>(lds, twl);
    fft_rest_c<-1, LGN, 3, 1, COMPACT>(lds, twl);
    const size_t so = (size_t)jp * (L + 1) * nmem + b;
    unpack_sd(lds, N, mmax, w, S + so, D + so, (size_t)nmem, threadIdx.x, kGridThreads);
}

// Single-member Jacobian grid stage, one workgroup per latitude row (both
// rows of a pair are separate workgroups: 2 nh workgroups fill the chip where
// the row-pair kernel leaves half of it idle).  Its input is the row sums the
// single-member stage kernel writes for it: E + O (north rows) in E, E - O
// (south rows) in O, each workgroup reading only its own.  The four fields of the row are
// paired into two complex rows, (a_l + i a_s) and (b_l + i b_s); one inverse
// FFT each gives the four grid rows, the Jacobian is formed per point and one
// forward FFT of the real product row gives its Fourier coefficients F(m),
// m <= mmax, stored unweighted: the north row's into FN[m][jp], the south
// row's into FS[m][jp].  The single-member stage kernel forms S = w (FN + FS),
// D = w (FN - FS) while staging them (phy/barotropic.py:120-167).
template <int LGN>
__global__ __launch_bounds__(kGridThreads) void grid_jac_hemi_kernel(
    GaussDev g, const double2* __restrict__ E, const double2* __restrict__ O, double R, int mmax,
    double2* __restrict__ FN, double2* __restrict__ FS, const double2* __restrict__ E1,
    const double2* __restrict__ O1, int msplit) {
    constexpr int N = 1 << LGN, RS = N + (N >> 3), NB = N >> 3;
    extern __shared__ double2 lds[];
    // the arguments into scalar registers at entry, in one batch (loaded
    // where first used, msplit came in a second round waited for between
    // the coefficient loads)
    asm volatile("" ::"s"(g.nh), "s"(g.L), "s"(g.nlat), "s"(g.coslat), "s"(g.fft.tw), "s"(E),
                 "s"(O), "s"(R), "s"(mmax), "s"(FN), "s"(FS), "s"(E1), "s"(O1), "s"(msplit));
    // workgroup -> (row pair jp, hemisphere hs): ids are dealt to the 8 XCDs
    // round-robin; XCD x takes the nh/8 consecutive row pairs from x nh/8
    // (both rows of each), so the 8 rows whose spectra share a line of
    // FN/FS[m][jp] write it through one L2 (full-line write-backs)
    int jp, hs;
    const int id = blockIdx.x;
    if ((g.nh & 7) == 0) {
        const int x = id & 7, k = id >> 3;
        hs = k & 1;
        jp = x * (g.nh >> 3) + (k >> 1);
    } else {
        jp = id >> 1;
        hs = id & 1;
    }
    const int L = g.L;
    const int j = hs ? g.nlat - 1 - jp : jp;
    const double invR = 1.0 / R;
    const double c = clamp_coslat(g.coslat[j]);
    const size_t eo = (size_t)jp * (L + 1) * 4;
    // LDS: rows A [2][RS] | rows B [2][RS] | stage-major twiddle table
    // (hemi_tw); the FFT stages alternate between A and B (one barrier per
    // stage)
    double2* rowA = lds;
    double2* rowB = lds + 2 * RS;
    double2* twl = lds + 4 * RS;
    {
        // prologue and inverse stage 1 (radix 8, no twiddles) in registers:
        // lane (row, jj) reads, for each of its bins b = jj + r N/8, the two
        // fields 2 row, 2 row + 1 of the one m the bin takes (N > 2L: the
        // bin-m term m = b when r < 4, the mirrored bin's m = N - b when
        // r >= 4) and forms the packed bin: Z1 = g_2r(m) + i g_2r+1(m) (real
        // at m = 0) or Z2 = conj(g_2r(m)) + i conj(g_2r+1(m)), g_f = this
        // row's sum E + O (north, in E) or E - O (south, in O), the split
        // columns' second plane added, with d/dlambda (even f) or 1/R (odd
        // f); every thread also loads its share of the stage-major twiddle
        // table
        static_assert(2 * N + 16 <= 2 * RS, "the staged coefficients fit in rows B");
        const double2* G = hs ? O : E;
        const double2* G1 = hs ? O1 : E1;
        constexpr int TWN = hemi_tw(LGN).total;
        static_assert(TWN <= 2 * kGridThreads, "two table entries per thread");
        // two named values, not an array: the array was kept in scratch
        // memory, whose store waited for the table loads before the
        // coefficient loads below were issued
        const double2* twh = g.fft.tw + N;           // the plan's copy, after the full table
        const int t0 = threadIdx.x, t1 = threadIdx.x + kGridThreads;
        const double2 vt0 = twh[t0 < TWN ? t0 : TWN - 1];
        const double2 vt1 = twh[t1 < TWN ? t1 : TWN - 1];
        // the row's coefficients [m][4 fields] in coalesced loads (a lane
        // per 16 bytes: the per-bin gathers took 16 bytes of each 64 and
        // one cache line per lane), the split columns' second plane added,
        // staged field-major in rows B, cs[f RC + m]: RC = 4 mod 16 keeps
        // the 16 lanes of a write pass (4 m x 4 f) and the bin reads
        // (consecutive m) on distinct banks
        const int NE = 4 * (L + 1), RC = ((L + 16) & ~15) + 4;
        constexpr int EPT = (4 * (N / 2) + kGridThreads - 1) / kGridThreads;  // 4 (L + 1) <= 2 N
        double2* cs = rowB;
        double2 ce[EPT], ce1[EPT];
#pragma unroll
        for (int u = 0; u < EPT; ++u) {
            const int e = threadIdx.x + u * kGridThreads;
            ce[u] = ce1[u] = c2(0.0, 0.0);
            if (e < NE) ce[u] = G[eo + e];
            if (e < NE && (e >> 2) < msplit) ce1[u] = G1[eo + e];
        }
        __builtin_amdgcn_sched_barrier(0);
        if (t0 < TWN) twl[t0] = vt0;
        if (t1 < TWN) twl[t1] = vt1;
#pragma unroll
        for (int u = 0; u < EPT; ++u) {
            const int e = threadIdx.x + u * kGridThreads, m = e >> 2;
            if (e < NE) cs[(e & 3) * RC + m] = (m < msplit) ? cadd(ce[u], ce1[u]) : ce[u];
        }
        __syncthreads();
        const int t = threadIdx.x;
        const bool first = t < 2 * NB;
        const int row = first ? t / NB : 0, jj = t % NB;
        if (first) {
            double2 v[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const int b = jj + r * NB, mm = (r < 4) ? b : N - b;
                const int mc = (mm <= L) ? mm : L;
                double2 ge = c2(0.0, 0.0), go = c2(0.0, 0.0);
                if (mm <= L) {                       // bins past the spectrum read nothing
                    ge = cs[2 * row * RC + mc];
                    go = cs[(2 * row + 1) * RC + mc];
                }
                ge = dlam_q(ge, mc * invR);
                go = cscale(go, invR);
                double2 z = c2(0.0, 0.0);
                if (r < 4) {
                    if (mc == 0) { ge.y = 0.0; go.y = 0.0; }
                    if (mm <= L) z = c2(ge.x - go.y, ge.y + go.x);
                } else if (mm <= L) {
                    z = cadd(z, c2(ge.x + go.y, -ge.y + go.x));
                }
                v[r] = z;
            }
            butterfly<+1, 8>(v);
            double2* rb = rowA + row * RS;
#pragma unroll
            for (int q = 0; q < 8; ++q) rb[lp(8 * jj + q)] = v[q];
        }
    }
    __syncthreads();
    const double q = 1.0 / (c * c);
    {
        // inverse stages radix 8, then 2^(LGN - 6) (none for 64 points, 2, 4
        // or 8 for 128, 256, 512), then a last radix-8 stage: that stage
        // leaves lane j both rows' values at the points j + N/8 q, exactly
        // the inputs of the first (radix-8) forward stage's lane j, so the
        // last inverse stage, the Jacobian at those points and the first
        // forward stage run in registers (no LDS round trip)
        static_assert(LGN >= 6 && LGN <= 9, "rows of 64 to 512 points");
        if constexpr (LGN == 7) fft_stage_pp<+1, 2, 1, LGN, 3, 2>(rowA, rowB, twl);
        if constexpr (LGN == 8) fft_stage_pp<+1, 4, 2, LGN, 3, 2>(rowA, rowB, twl);
        if constexpr (LGN == 9) fft_stage_pp<+1, 8, 3, LGN, 3, 2>(rowA, rowB, twl);
        double2* src = (LGN == 6) ? rowA : rowB;     // the last inverse stage's input rows
        double2* dst = (LGN == 6) ? rowB : rowA;
        constexpr int LS = LGN - 3;                  // the last stage's LGNS (N/8 = 2^LS lanes)
        constexpr int TWF = hemi_tw_off(LGN, 3, LS);   // its twiddles W^(r jl)
        const int jl = threadIdx.x;
        double2 v[8];
        if (jl < NB) {
            double2 xa[8], xb[8], wt[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                xa[r] = src[lp(jl + r * NB)];
                xb[r] = src[RS + lp(jl + r * NB)];
                if (r > 0) wt[r] = twl[TWF + (r - 1) * NB + jl];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int r = 1; r < 8; ++r) {
                const double2 w = c2(wt[r].x, -wt[r].y);   // inverse: conjugate
                xa[r] = c2(fma(xa[r].x, w.x, -(xa[r].y * w.y)), fma(xa[r].x, w.y, xa[r].y * w.x));
                xb[r] = c2(fma(xb[r].x, w.x, -(xb[r].y * w.y)), fma(xb[r].x, w.y, xb[r].y * w.x));
            }
            butterfly<+1, 8>(xa);
            butterfly<+1, 8>(xb);
#pragma unroll
            for (int r = 0; r < 8; ++r)
                v[r] = c2(jac_point_q(xa[r].x, xa[r].y, xb[r].x, xb[r].y, q), 0.0);
            butterfly<-1, 8>(v);
        }
        static_assert(LS >= 3, "fused stages need N >= 64");
        if (jl < NB) {
#pragma unroll
            for (int r = 0; r < 8; ++r) dst[lp(8 * jl + r)] = v[r];
        }
        __syncthreads();
        // column-major [m][jp]: the stage kernel of column m reads a
        // contiguous column (the row-major layout cost it one 128-byte line
        // per 16 bytes)
        fft_rest_pp_out<-1, LGN, 3>(dst, src, twl, (hs ? FS : FN) + jp, g.nh, mmax);
    }
}

// Kernel instance for the plan's FFT length (power-of-two stages or general radices).
#define AEO_GRID_PICK(name)                                                        \
    static decltype(&name<true>) name##_for(const aeo_plan_s* p) {                 \
        return fft_p2(p->fft) ? &name<true> : &name<false>;                        \
    }
AEO_GRID_PICK(grid_syn_kernel)
AEO_GRID_PICK(grid_ana_kernel)
AEO_GRID_PICK(grid_jac_kernel)
AEO_GRID_PICK(grid_swe_kernel)
AEO_GRID_PICK(grid_adj_kernel)
#undef AEO_GRID_PICK

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
int gauss_build_tables(aeo_plan_s* p, hipStream_t s) {
    dim3 grid(p->L + 1, (p->nh + 63) / 64);
    hipLaunchKernelGGL(legendre_table_kernel, grid, dim3(64), 0, s, p->gdev(), p->d_Ps, p->d_Pa);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

int gauss_build_split_tables(aeo_plan_s* p, int ns, hipStream_t s) {
    const int nhp = rup(p->nh, 16), nhk = split_nhk(p->nh);
    const int rows = (nhp > rup(split_kp(p->L, 0), 16)) ? nhp : rup(split_kp(p->L, 0), 16);
    aeo::GaussDev g = p->gdev();
    dim3 grid(p->L + 1, 2, (rows + 63) / 64);
    hipLaunchKernelGGL(split_table_kernel, grid, dim3(64), 0, s, g, ns, p->d_Psl, p->d_esyn,
                       p->d_Pasl, p->d_eana, nhp, nhk);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

// Dynamic LDS beyond 64 KiB needs the kernel attribute.
static int allow_lds(const void* kern, size_t bytes) {
    AEO_REQUIRE(bytes <= 163840, AEO_ERR_UNSUPPORTED, "grid kernel: LDS demand exceeds 160 KiB");
    if (bytes > 65536)
        AEO_CHECK_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)bytes));
    return AEO_OK;
}

// LDS of a fused grid kernel: `rows` padded rows, the twiddle table and the
// staged E/O columns (nf fields each).
static size_t grid_lds(const aeo_plan_s* p, int rows, int nf) {
    return sizeof(double2) * ((size_t)rows * lds_row_stride(p->nlon) + p->nlon +
                              2 * (size_t)nf * (p->L + 1));
}

static bool jac_p2(const aeo_plan_s* p) {
    return fft_p2(p->fft) && p->nlon > 2 * p->L && p->fft.lgn >= 6 && p->fft.lgn <= 9;
}

// LDS of the Jacobian grid kernel jac_kernel_for(p, nmem) picks: for member
// batches on the power-of-two instance four padded rows and a quarter
// twiddle table (E/O staged in the rows), else grid_lds.
static size_t jac_lds(const aeo_plan_s* p, int nmem) {
    if (jac_p2(p) && nmem > 1)
        return sizeof(double2) * (4 * (size_t)lds_row_stride(p->nlon) + p->nlon / 4);
    return grid_lds(p, 4, 4);
}

// The Jacobian grid kernel for this plan: the compile-time power-of-two
// instance for rows of 64 to 512 points without aliasing (n > 2L), else the
// general one.
using JacKernel = void (*)(GaussDev, const double2*, const double2*, double, int, int, double2*,
                           double2*, const double2*, const double2*, int);
static JacKernel jac_kernel_for(const aeo_plan_s* p, int nmem) {
    if (jac_p2(p)) {
        const bool c = nmem > 1;
        switch (p->fft.lgn) {
            case 6: return c ? &grid_jac_p2_kernel<6, true> : &grid_jac_p2_kernel<6, false>;
            case 7: return c ? &grid_jac_p2_kernel<7, true> : &grid_jac_p2_kernel<7, false>;
            case 8: return c ? &grid_jac_p2_kernel<8, true> : &grid_jac_p2_kernel<8, false>;
            case 9: return c ? &grid_jac_p2_kernel<9, true> : &grid_jac_p2_kernel<9, false>;
            default: break;
        }
    }
    return grid_jac_kernel_for(p);
}

// The single-member BVE stage pairs with the per-row Jacobian grid kernel
// (grid_jac_hemi_kernel) where the power-of-two instance applies; its S/D
// arrays then hold the unweighted row spectra FN/FS.
static bool bve1_hemi(const aeo_model_s* md) {
    return md->kind == kBVE && bve1_supported(md) && jac_p2(md->plan);
}
using HemiKernel = void (*)(GaussDev, const double2*, const double2*, double, int, double2*,
                            double2*, const double2*, const double2*, int);
static HemiKernel hemi_kernel_for(const aeo_plan_s* p) {
    switch (p->fft.lgn) {
        case 6: return &grid_jac_hemi_kernel<6>;
        case 7: return &grid_jac_hemi_kernel<7>;
        case 8: return &grid_jac_hemi_kernel<8>;
        default: return &grid_jac_hemi_kernel<9>;
    }
}
static size_t hemi_lds(const aeo_plan_s* p) {
    const int lgn = p->fft.lgn < 6 ? 6 : (p->fft.lgn > 9 ? 9 : p->fft.lgn);
    return sizeof(double2) * (4 * (size_t)lds_row_stride(p->nlon) + hemi_tw(lgn).total);
}

// Rows per workgroup of the plain transforms' grid kernels for a batch of nf
// fields: enough workgroups (nh x batches >= 512) to fill the chip, at most
// row_cap rows.
static int rows_per_wg(const aeo_plan_s* p, int nf) {
    const int want = (p->nh * nf + 511) / 512;
    const int cap = row_cap(p->nlon);
    return want < 1 ? 1 : (want > cap ? cap : want);
}

static size_t row_lds(const aeo_plan_s* p, int rows) {
    // padded rows of one LDS pass + the twiddle table
    return sizeof(double2) * ((size_t)rows * lds_row_stride(p->nlon) + p->nlon);
}

// Fields per Legendre workgroup of a plain transform batch: each workgroup
// streams its m column of the table once for all of its fields; a tile whose
// LDS demand is too large is halved by the caller.
static int batch_tile(int nf) { return nf >= 4 ? 4 : (nf >= 2 ? 2 : 1); }

// Workgroups along the field axis of the plain transforms' grid kernels.
static unsigned field_batches(int nf, int rows) { return (unsigned)((nf + rows - 1) / rows); }

int gauss_synthesis(aeo_plan_s* p, const double2* a, int nf, double* f, hipStream_t s) {
    const int L = p->L;
    const size_t fs = (size_t)(L + 1) * (L + 1);
    for (int f0 = 0; f0 < nf; f0 += p->max_fields) {
        const int nb = (nf - f0 < p->max_fields) ? nf - f0 : p->max_fields;
        StageArgs sa{};
        sa.ys = a + (size_t)f0 * fs;
        sa.ys_stride = fs;
        sa.nmem = nb;
        sa.nstate_valid = 1;
        sa.E = p->d_E;
        sa.O = p->d_O;
        sa.k = col_consts(p, 1.0, nullptr, 0.0);
        int rc = AEO_ERR_UNSUPPORTED;
        for (int mb = batch_tile(nb); rc == AEO_ERR_UNSUPPORTED && mb >= 1; mb /= 2)
            rc = launch_stage<KindRaw, EpiNoneS, false, true>(p, mb, sa, EpiNoneS{}, s);
        if (rc) return rc;
        const int rows = rows_per_wg(p, nb);
        int rc2 = allow_lds((const void*)grid_syn_kernel_for(p), row_lds(p, rows));
        if (rc2) return rc2;
        hipLaunchKernelGGL(grid_syn_kernel_for(p), dim3(p->nh, field_batches(nb, rows)),
                           dim3(kGridThreads), row_lds(p, rows), s, p->gdev(), p->d_E, p->d_O,
                           nb, rows, f + (size_t)f0 * p->nlat * p->nlon);
        AEO_LAUNCH_CHECK();
    }
    return AEO_OK;
}

int gauss_analysis(aeo_plan_s* p, const double* f, int nf, double2* a, hipStream_t s) {
    const int L = p->L;
    const size_t fs = (size_t)(L + 1) * (L + 1);
    for (int f0 = 0; f0 < nf; f0 += p->max_fields) {
        const int nb = (nf - f0 < p->max_fields) ? nf - f0 : p->max_fields;
        const int rows = rows_per_wg(p, nb);
        int rc2 = allow_lds((const void*)grid_ana_kernel_for(p), row_lds(p, rows));
        if (rc2) return rc2;
        hipLaunchKernelGGL(grid_ana_kernel_for(p), dim3(p->nh, field_batches(nb, rows)),
                           dim3(kGridThreads), row_lds(p, rows), s, p->gdev(),
                           f + (size_t)f0 * p->nlat * p->nlon, nb, rows, L, p->d_S, p->d_D);
        AEO_LAUNCH_CHECK();
        StageArgs sa{};
        sa.nmem = nb;
        sa.nstate_valid = 0;
        sa.S = p->d_S;
        sa.D = p->d_D;
        sa.lmax_a = L;
        sa.mmax_a = L;
        sa.k = col_consts(p, 1.0, nullptr, 0.0);
        EpiStore1S ep{a + (size_t)f0 * fs, L, fs};
        int rc = AEO_ERR_UNSUPPORTED;
        for (int mb = batch_tile(nb); rc == AEO_ERR_UNSUPPORTED && mb >= 1; mb /= 2)
            rc = launch_stage<KindNone, EpiStore1S, true, false>(p, mb, sa, ep, s);
        if (rc) return rc;
    }
    return AEO_OK;
}

int gauss_jacobian(aeo_plan_s* p, double R, const double2* a, const double2* b, double2* out,
                   int dealias, hipStream_t s) {
    const int L = p->L;
    const size_t fs = (size_t)(L + 1) * (L + 1);
    AEO_CHECK_HIP(hipMemcpyAsync(p->d_spec, a, fs * sizeof(double2), hipMemcpyDeviceToDevice, s));
    AEO_CHECK_HIP(hipMemcpyAsync(p->d_spec + fs, b, fs * sizeof(double2), hipMemcpyDeviceToDevice, s));
    StageArgs sa{};
    sa.ys = p->d_spec;
    sa.ys_stride = 2 * fs;
    sa.nmem = 1;
    sa.nstate_valid = 2;
    sa.E = p->d_E;
    sa.O = p->d_O;
    sa.k = col_consts(p, R, nullptr, 0.0);
    int rc = launch_stage<KindJac, EpiNoneS, false, true>(p, 1, sa, EpiNoneS{}, s);
    if (rc) return rc;
    const int lim = dealias ? (2 * L) / 3 : L;
    if ((rc = allow_lds((const void*)jac_kernel_for(p, 1), jac_lds(p, 1)))) return rc;
    hipLaunchKernelGGL(jac_kernel_for(p, 1), dim3(p->nh), dim3(kGridThreads), jac_lds(p, 1), s,
                       p->gdev(), p->d_E, p->d_O, R, lim, 1, p->d_S, p->d_D,
                       (const double2*)nullptr, (const double2*)nullptr, 0);
    AEO_LAUNCH_CHECK();
    StageArgs sb{};
    sb.nmem = 1;
    sb.nstate_valid = 0;
    sb.S = p->d_S;
    sb.D = p->d_D;
    sb.lmax_a = lim;
    sb.mmax_a = lim;
    sb.k = sa.k;
    EpiStore1S ep{out, L, fs};
    return launch_stage<KindNone, EpiStore1S, true, false>(p, 1, sb, ep, s);
}

int gauss_model_prime(aeo_model_s* md, const double2* y, hipStream_t s) {
    aeo_plan_s* p = md->plan;
    const size_t fs = (size_t)(p->L + 1) * (p->L + 1);
    const double f10 = 2.0 * md->omega * sqrt(4.0 * M_PI / 3.0);
    StageArgs sa{};
    sa.ys = y;
    sa.ys_mmajor = 1;
    sa.ys_stride = fs * md->nstate;
    sa.nmem = md->nmem;
    sa.nstate_valid = md->nstate;
    sa.E = md->d_E;
    sa.O = md->d_O;
    sa.k = col_consts(p, md->R, md->d_lap, f10);
    md->prof.arm(kClsSynthesis);
    if (bve1_supported(md))
        return bve1_launch(md, y, nullptr, nullptr, nullptr, -1, 0.0, nullptr, false, true,
                           bve1_hemi(md), s);
    if (md->mb == 0)
        return gemm_stage_launch(md, y, nullptr, nullptr, nullptr, -1, 0.0, nullptr, false, true,
                                 s);
    int rc = (md->kind == kBVE)
                 ? launch_stage<KindBve, EpiNoneS, false, true>(p, md->mb, sa, EpiNoneS{}, s,
                                                                 md->nslices)
                 : launch_stage<KindSwe, EpiNoneS, false, true>(p, md->mb, sa, EpiNoneS{}, s,
                                                                 md->nslices);
    return rc;
}

template <class KS, class Epi>
static int stage_ana(aeo_model_s* md, const StageArgs& sa, const Epi& ep, bool syn_next,
                     hipStream_t s) {
    return syn_next
               ? launch_stage<KS, Epi, true, true>(md->plan, md->mb, sa, ep, s, md->nslices)
               : launch_stage<KS, Epi, true, false>(md->plan, md->mb, sa, ep, s, md->nslices);
}

int gauss_model_stage(aeo_model_s* md, const double2* ys, const double2* y0,
                      const double2* forcing, double2* kout, int stage, double dt,
                      double2* ynext, bool syn_next, hipStream_t s) {
    aeo_plan_s* p = md->plan;
    const int L = p->L;
    const int cut = (2 * L) / 3;
    const size_t fs = (size_t)(L + 1) * (L + 1);
    const double f10 = 2.0 * md->omega * sqrt(4.0 * M_PI / 3.0);
    ProfRec& pr = md->prof;
    if (bve1_hemi(md)) {
        // single member: one workgroup per latitude row; the split columns
        // of the last synthesis arrive in two planes
        int rc0 = allow_lds((const void*)hemi_kernel_for(p), hemi_lds(p));
        if (rc0) return rc0;
        pr.arm(kClsGrid);
        launch_k(hemi_kernel_for(p), dim3(2 * p->nh), dim3(kGridThreads), hemi_lds(p), s,
                 p->gdev(), (const double2*)md->d_E, (const double2*)md->d_O, md->R, cut,
                 md->d_S, md->d_D, (const double2*)md->d_E1, (const double2*)md->d_O1,
                 bve1_msplit(p));
    } else if (md->kind == kBVE) {
        int rc0 = allow_lds((const void*)jac_kernel_for(p, md->nmem), jac_lds(p, md->nmem));
        if (rc0) return rc0;
        // a single-member synthesis leaves its split columns in two planes
        const int ms1 = bve1_supported(md) ? bve1_msplit(p) : 0;
        pr.arm(kClsGrid);
        launch_k(jac_kernel_for(p, md->nmem), dim3(p->nh * md->nmem), dim3(kGridThreads),
                 jac_lds(p, md->nmem), s,
                 p->gdev(), (const double2*)md->d_E, (const double2*)md->d_O, md->R, cut,
                 md->nmem, md->d_S, md->d_D, (const double2*)md->d_E1,
                 (const double2*)md->d_O1, ms1);
    } else {
        const size_t swe_lds = grid_lds(p, kSweLdsRows, 10);
        int rc0 = allow_lds((const void*)grid_swe_kernel_for(p), swe_lds);
        if (rc0) return rc0;
        pr.arm(kClsGrid);
        launch_k(grid_swe_kernel_for(p), dim3(p->nh * md->nmem, 2), dim3(kGridThreads), swe_lds, s,
                 p->gdev(), (const double2*)md->d_E, (const double2*)md->d_O, md->R, cut,
                 md->nmem, md->d_S, md->d_D);
    }
    AEO_LAUNCH_CHECK();
    StageArgs sa{};
    sa.ys = ys;
    sa.ys_mmajor = 1;
    sa.ys_stride = fs * md->nstate;
    sa.nmem = md->nmem;
    sa.nstate_valid = md->nstate;
    sa.S = md->d_S;
    sa.D = md->d_D;
    sa.lmax_a = cut;
    sa.mmax_a = cut;
    sa.E = md->d_E;
    sa.O = md->d_O;
    sa.k = col_consts(p, md->R, md->d_lap, f10);
    Rk4B rk{stage, stage == 2 ? dt : 0.5 * dt, dt / 6.0, y0, md->d_ksum, ynext, kout};
    int rc;
    pr.arm(kClsAnalysis);
    if (md->kind == kBVE && bve1_supported(md)) {
        rc = bve1_launch(md, ys, y0, forcing, kout, stage, dt, ynext, true, syn_next,
                         bve1_hemi(md), s);
    } else if (md->mb == 0) {
        rc = gemm_stage_launch(md, ys, y0, forcing, kout, stage, dt, ynext, true, syn_next, s);
    } else if (md->kind == kBVE) {
        EpiBveS ep{L, fs, forcing, md->d_nulap, rk};
        rc = stage_ana<KindBve>(md, sa, ep, syn_next, s);
    } else {
        EpiSweS ep{L, 3 * fs, md->d_lap, md->nu4 > 0.0 ? md->d_nu4lap2 : nullptr,
                   md->has_topo ? md->d_lap_phi_s : nullptr,
                   (md->has_topo && md->nu4 > 0.0) ? md->d_nu4_lap2_phi_s : nullptr, md->phi0,
                   rk};
        rc = stage_ana<KindSwe>(md, sa, ep, syn_next, s);
    }
    return rc;
}

// ---------------------------------------------------------------------------
// BVE adjoint launchers
// ---------------------------------------------------------------------------
// Members per workgroup of the adjoint stages: the forward's fused tile, or
// 4 when the forward runs in GEMM form (which has no adjoint counterpart).
static int adjoint_mb(const aeo_model_s* md) { return md->mb ? md->mb : 4; }

int gauss_adjoint_prime(aeo_model_s* md, const double2* yfwd, const double2* kb, hipStream_t s) {
    aeo_plan_s* p = md->plan;
    const size_t fs = (size_t)(p->L + 1) * (p->L + 1);
    const double f10 = 2.0 * md->omega * sqrt(4.0 * M_PI / 3.0);
    StageArgs sa{};
    sa.ys = yfwd;
    sa.ys_mmajor = 1;
    sa.ys_stride = fs;
    sa.ys_aux = kb;
    sa.aux_stride = fs;
    sa.nmem = md->nmem;
    sa.nstate_valid = 2;
    sa.E = md->d_aE;
    sa.O = md->d_aO;
    sa.k = col_consts(p, md->R, md->d_lap, f10);
    md->prof.arm(kClsSynthesis);
    return launch_stage<KindBveAdj, EpiNoneS, false, true>(p, adjoint_mb(md), sa, EpiNoneS{}, s,
                                                           md->nslices);
}

int gauss_adjoint_stage(aeo_model_s* md, const double2* kb, int stage, double dt,
                        const double2* a, double2* anext, double2* kout, double2* knext,
                        const double2* fwd_next, hipStream_t s) {
    aeo_plan_s* p = md->plan;
    const int L = p->L;
    const size_t fs = (size_t)(L + 1) * (L + 1);
    const double f10 = 2.0 * md->omega * sqrt(4.0 * M_PI / 3.0);
    ProfRec& pr = md->prof;
    const size_t lds = grid_lds(p, kAdjLdsRows, 5);
    int rc0 = allow_lds((const void*)grid_adj_kernel_for(p), lds);
    if (rc0) return rc0;
    pr.arm(kClsGrid);
    launch_k(grid_adj_kernel_for(p), dim3(p->nh * md->nmem, 2), dim3(kGridThreads), lds, s,
             p->gdev(), (const double2*)md->d_aE, (const double2*)md->d_aO, md->R, md->nmem,
             md->d_aS, md->d_aD);
    AEO_LAUNCH_CHECK();
    StageArgs sa{};
    sa.ys = nullptr;              // column 0 is filled by the epilogue (fwd_next)
    sa.ys_mmajor = 1;
    sa.ys_stride = fs;
    sa.ys_aux = kb;
    sa.aux_stride = fs;
    sa.nmem = md->nmem;
    sa.nstate_valid = 2;
    sa.S = md->d_aS;
    sa.D = md->d_aD;
    sa.lmax_a = L;
    sa.mmax_a = L;
    sa.E = md->d_aE;
    sa.O = md->d_aO;
    sa.k = col_consts(p, md->R, md->d_lap, f10);
    Rk4Adj rk{stage, dt, a, md->d_asum, anext, kout, knext};
    EpiBveAdjS ep{L, md->R, fs, md->d_nulap, md->d_lap, p->d_CpT, p->d_CmT, fwd_next, rk};
    pr.arm(kClsAnalysis);
    if (fwd_next)
        return launch_stage<KindBveAdj, EpiBveAdjS, true, true>(p, adjoint_mb(md), sa, ep, s,
                                                                md->nslices);
    return launch_stage<KindBveAdj, EpiBveAdjS, true, false>(p, adjoint_mb(md), sa, ep, s,
                                                             md->nslices);
}

}  // namespace aeo
