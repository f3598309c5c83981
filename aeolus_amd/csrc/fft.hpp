// Batched in-place Stockham FFT over rows held in LDS (gfx950).
//
// A workgroup transforms `nrows` contiguous rows of length n (complex FP64)
// in LDS.  Each stage of radix R reads the R inputs of every butterfly into
// registers, synchronises, and writes the R outputs to their Stockham
// positions, so no second LDS buffer is needed:
//
//   j in [0, n/R),  k = j mod Ns,
//   v[r] = x[j + r n/R] * W^(r k n/(Ns R)),        (W = e^{dir 2 pi i / n})
//   V    = R-point DFT of v,
//   x[(j / Ns) Ns R + k + q Ns] = V[q],            Ns *= R after the stage.
//
// Radices 2, 3, 4 and 5 use hand-written butterflies; any other prime factor
// uses a direct R-point DFT computed one output per thread.  Rows are
// processed in chunks of at most kChunkElems complex values so each thread
// keeps at most 8 complex values in registers per stage.
#pragma once

#include "common.hpp"

namespace aeo {

constexpr int kChunkElems = 2048;  // complex values per chunk (256 threads x 8)

// LDS rows are padded by one element per 8 (element i of a row at i + i/8,
// row stride n + n/8): Stockham writes at stride 8 then spread over all banks.
__host__ __device__ inline int lds_row_stride(int n) { return n + (n >> 3); }
__device__ inline int lp(int i) { return i + (i >> 3); }

struct FftDev {
    int n;
    int nstages;
    int radix[kMaxRadices];
    const double2* tw;  // tw[t] = exp(-2 pi i t / n), t in [0, n)
    int lgn;            // log2(n) when n is a power of two (shift/mask stages), else -1
};

// Twiddle table of the per-row Jacobian kernel (grid_jac_hemi_kernel),
// stage-major so that the lanes of a stage read consecutive entries: read
// from the full table W^t, t < N, the span-8 radix-8 stage took W^(8 r k) at
// a stride of 8 r entries (eight lanes on one bank) and the span-1 stages
// W^(r k) at a stride of r.  Block (LGR, LGNS), the radix-2^LGR stage whose
// spans are 2^LGNS, holds W^((r k) 2^(LGN - LGNS - LGR)) at (r - 1) 2^LGNS + k,
// 0 < r < 2^LGR, k < 2^LGNS.  The blocks are those of the stages the kernel
// runs after the untwiddled first one, each once: the inverse middle stage
// (radix 2^(LGN - 6), spans 8), the last inverse stage (radix 8, spans
// N/8), the forward stage of spans 8 (radix 8) and the forward last stage
// (radix 2^(LGN - 6), spans 64); N = 512 and N = 64 share theirs between
// the two directions.  A plan of 64 to 512 (power-of-two) longitudes keeps
// it after its full table, at tw + N.
struct HemiTw {
    int n, total;
    int lgr[4], lgns[4], off[4];
};
__host__ __device__ constexpr HemiTw hemi_tw(int lgn) {
    HemiTw t{};
    const int cr[4] = {lgn - 6, 3, 3, lgn - 6}, cs[4] = {3, lgn - 3, 3, 6};
    for (int i = 0; i < 4; ++i) {
        bool have = cr[i] <= 0;
        for (int k = 0; k < t.n; ++k) have = have || (t.lgr[k] == cr[i] && t.lgns[k] == cs[i]);
        if (have) continue;
        t.lgr[t.n] = cr[i];
        t.lgns[t.n] = cs[i];
        t.off[t.n] = t.total;
        t.total += ((1 << cr[i]) - 1) << cs[i];
        ++t.n;
    }
    return t;
}
__host__ __device__ constexpr int hemi_tw_off(int lgn, int lgr, int lgns) {
    const HemiTw t = hemi_tw(lgn);
    for (int k = 0; k < t.n; ++k)
        if (t.lgr[k] == lgr && t.lgns[k] == lgns) return t.off[k];
    return -1;
}

template <int DIR>
__device__ inline double2 twiddle(const double2* tw, int t) {
    double2 w = tw[t];
    if (DIR > 0) w.y = -w.y;
    return w;
}

// Multiply by DIR * i.
template <int DIR>
__device__ inline double2 mul_i(double2 a) {
    return DIR > 0 ? c2(-a.y, a.x) : c2(a.y, -a.x);
}

template <int DIR, int R>
__device__ inline void butterfly(double2* v) {
    if constexpr (R == 2) {
        double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        double2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        double2 s13 = cadd(v[1], v[3]), d13 = mul_i<DIR>(csub(v[1], v[3]));
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13);
        v[3] = csub(d02, d13);
    } else if constexpr (R == 3) {
        const double h = 0.86602540378443864676;  // sin(2 pi / 3)
        double2 s = cadd(v[1], v[2]);
        double2 d = csub(v[1], v[2]);
        double2 m = c2(v[0].x - 0.5 * s.x, v[0].y - 0.5 * s.y);
        double2 t = mul_i<DIR>(cscale(d, h));
        v[0] = cadd(v[0], s);
        v[1] = cadd(m, t);
        v[2] = csub(m, t);
    } else if constexpr (R == 8) {
        // two 4-point DFTs (even / odd inputs) joined with w^q, w = e^{DIR 2 pi i / 8}
        const double h = 0.70710678118654752440;  // 1/sqrt(2)
        double2 e[4] = {v[0], v[2], v[4], v[6]};
        double2 o[4] = {v[1], v[3], v[5], v[7]};
        butterfly<DIR, 4>(e);
        butterfly<DIR, 4>(o);
        // o1 (1 + DIR i)/sqrt2, o2 DIR i, o3 (-1 + DIR i)/sqrt2
        const double2 t1 = c2(h * (o[1].x - DIR * o[1].y), h * (o[1].y + DIR * o[1].x));
        const double2 t2 = mul_i<DIR>(o[2]);
        const double2 t3 = c2(h * (-o[3].x - DIR * o[3].y), h * (-o[3].y + DIR * o[3].x));
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], t1);
        v[5] = csub(e[1], t1);
        v[2] = cadd(e[2], t2);
        v[6] = csub(e[2], t2);
        v[3] = cadd(e[3], t3);
        v[7] = csub(e[3], t3);
    } else if constexpr (R == 5) {
        const double c1 = 0.30901699437494742410;   // cos(2 pi / 5)
        const double c2_ = -0.80901699437494742410; // cos(4 pi / 5)
        const double s1 = 0.95105651629515357212;   // sin(2 pi / 5)
        const double s2 = 0.58778525229247312917;   // sin(4 pi / 5)
        double2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        double2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        double2 x0 = v[0];
        double2 r1 = c2(x0.x + c1 * a1.x + c2_ * a2.x, x0.y + c1 * a1.y + c2_ * a2.y);
        double2 r2 = c2(x0.x + c2_ * a1.x + c1 * a2.x, x0.y + c2_ * a1.y + c1 * a2.y);
        double2 i1 = mul_i<DIR>(c2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
        double2 i2 = mul_i<DIR>(c2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
        v[0] = cadd(x0, cadd(a1, a2));
        v[1] = cadd(r1, i1);
        v[4] = csub(r1, i1);
        v[2] = cadd(r2, i2);
        v[3] = csub(r2, i2);
    }
}

// One radix-R stage over `nrows` rows of length n starting at buf.
template <int DIR, int R>
__device__ inline void fft_stage(double2* buf, int nrows, int n, int Ns, const double2* tw) {
    constexpr int MAXB = (kChunkElems / R + kGridThreads - 1) / kGridThreads;
    const int nb = n / R;
    const int total = nrows * nb;
    const int span = n / (Ns * R);
    double2 v[MAXB][R];
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int t = threadIdx.x + b * kGridThreads;
        if (t < total) {
            const int row = t / nb;
            const int j = t - row * nb;
            const int k = j % Ns;
            const double2* rb = buf + row * lds_row_stride(n);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                double2 x = rb[lp(j + r * nb)];
                if (r > 0) x = cmul(x, twiddle<DIR>(tw, r * k * span));
                v[b][r] = x;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int t = threadIdx.x + b * kGridThreads;
        if (t < total) {
            const int row = t / nb;
            const int j = t - row * nb;
            const int k = j % Ns;
            butterfly<DIR, R>(v[b]);
            double2* rb = buf + row * lds_row_stride(n);
            const int base = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) rb[lp(base + q * Ns)] = v[b][q];
        }
    }
    __syncthreads();
}

// Direct-DFT stage for a prime radix R > 5: one output per thread slot.
template <int DIR>
__device__ inline void fft_stage_generic(double2* buf, int nrows, int n, int Ns, int R,
                                         const double2* tw) {
    constexpr int MAXO = kChunkElems / kGridThreads;
    const int total = nrows * n;
    const int nb = n / R;
    const int span = n / (Ns * R);
    double2 out[MAXO];
#pragma unroll
    for (int b = 0; b < MAXO; ++b) {
        const int t = threadIdx.x + b * kGridThreads;
        double2 acc = c2(0.0, 0.0);
        if (t < total) {
            const int row = t / n;
            const int p = t - row * n;           // output position
            const int k = p % Ns;
            const int q = (p / Ns) % R;
            const int j = (p / (Ns * R)) * Ns + k;
            const double2* rb = buf + row * lds_row_stride(n);
            for (int r = 0; r < R; ++r) {
                double2 x = rb[lp(j + r * nb)];
                if (r > 0) x = cmul(x, twiddle<DIR>(tw, r * k * span));
                const int e = ((r * q) % R) * nb;  // W_R^{rq} = W_n^{(rq mod R) n/R}
                acc = cadd(acc, e ? cmul(x, twiddle<DIR>(tw, e)) : x);
            }
        }
        out[b] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MAXO; ++b) {
        const int t = threadIdx.x + b * kGridThreads;
        if (t < total) {
            const int row = t / n;
            buf[row * lds_row_stride(n) + lp(t - row * n)] = out[b];
        }
    }
    __syncthreads();
}

// Power-of-two stage: n = 2^lgn, Ns = 2^lgNs, all index arithmetic by shifts
// and masks; rows * n / R butterflies spread over the workgroup.
template <int DIR, int R, int LGR>
__device__ inline void fft_stage_p2(double2* buf, int nrows, int lgn, int lgNs, const double2* tw) {
    constexpr int MAXB = (kChunkElems / R + kGridThreads - 1) / kGridThreads;
    const int lgnb = lgn - LGR;
    const int total = nrows << lgnb;
    const int nmask = (1 << lgnb) - 1, kmask = (1 << lgNs) - 1;
    const int lgspan = lgn - lgNs - LGR;
    double2 v[MAXB][R];
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int t = threadIdx.x + b * kGridThreads;
        if (t < total) {
            const int row = t >> lgnb;
            const int j = t & nmask;
            const int k = j & kmask;
            const double2* rb = buf + row * lds_row_stride(1 << lgn);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                double2 x = rb[lp(j + (r << lgnb))];
                if (r > 0) x = cmul(x, twiddle<DIR>(tw, (r * k) << lgspan));
                v[b][r] = x;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int t = threadIdx.x + b * kGridThreads;
        if (t < total) {
            const int row = t >> lgnb;
            const int j = t & nmask;
            const int k = j & kmask;
            butterfly<DIR, R>(v[b]);
            double2* rb = buf + row * lds_row_stride(1 << lgn);
            const int base = ((j >> lgNs) << (lgNs + LGR)) + k;
#pragma unroll
            for (int q = 0; q < R; ++q) rb[lp(base + (q << lgNs))] = v[b][q];
        }
    }
    __syncthreads();
}

// Copy the twiddle table into LDS (n entries at tw_lds), so no stage waits
// on a global load; pass tw_lds to fft_rows.  The caller synchronises the
// workgroup before the first stage.
__device__ inline void fft_twiddles_to_lds(const FftDev& fp, double2* tw_lds) {
    for (int t = threadIdx.x; t < fp.n; t += kGridThreads) tw_lds[t] = fp.tw[t];
}

// Transform `nrows` rows of length fp.n in LDS (in place), stored with the
// padded layout above (row r, element i at r * lds_row_stride(n) + lp(i));
// tw: the twiddle table (global or an LDS copy).
// DIR = -1: X_m = sum_k x_k e^{-2 pi i m k / n};  DIR = +1: unnormalised inverse.
// Must be called by all threads of the (kGridThreads-wide) workgroup.
// P2: n is a power of two <= kChunkElems (fp.lgn >= 0), compiled without the
// general-radix stages.
template <int DIR, bool P2>
__device__ void fft_rows(double2* buf, int nrows, const FftDev& fp, const double2* tw) {
    const int n = fp.n;
    const int cb = (n <= kChunkElems) ? (kChunkElems / n) : 1;
    if constexpr (P2) {
        // radix-8 stages, then one radix-4 or radix-2 stage for the remainder
        for (int r0 = 0; r0 < nrows; r0 += cb) {
            const int nr = (nrows - r0 < cb) ? (nrows - r0) : cb;
            double2* b = buf + r0 * lds_row_stride(n);
            int lgNs = 0;
            while (fp.lgn - lgNs >= 3) {
                fft_stage_p2<DIR, 8, 3>(b, nr, fp.lgn, lgNs, tw);
                lgNs += 3;
            }
            if (fp.lgn - lgNs == 2) fft_stage_p2<DIR, 4, 2>(b, nr, fp.lgn, lgNs, tw);
            if (fp.lgn - lgNs == 1) fft_stage_p2<DIR, 2, 1>(b, nr, fp.lgn, lgNs, tw);
        }
        return;
    } else {
    for (int r0 = 0; r0 < nrows; r0 += cb) {
        const int nr = (nrows - r0 < cb) ? (nrows - r0) : cb;
        double2* b = buf + r0 * lds_row_stride(n);
        int Ns = 1;
        for (int s = 0; s < fp.nstages; ++s) {
            const int R = fp.radix[s];
            switch (R) {
                case 2: fft_stage<DIR, 2>(b, nr, n, Ns, tw); break;
                case 3: fft_stage<DIR, 3>(b, nr, n, Ns, tw); break;
                case 4: fft_stage<DIR, 4>(b, nr, n, Ns, tw); break;
                case 5: fft_stage<DIR, 5>(b, nr, n, Ns, tw); break;
                default: fft_stage_generic<DIR>(b, nr, n, Ns, R, tw); break;
            }
            Ns *= R;
        }
    }
    }
}

__host__ __device__ inline bool fft_p2(const FftDev& fp) {
    return fp.lgn >= 0 && fp.n <= kChunkElems;
}

}  // namespace aeo
