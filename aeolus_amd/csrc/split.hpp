// Split-precision digits for the Legendre contractions on INT8 MFMA.
//
// A value v of a row (or column) whose largest magnitude is below 2^e is
// written with balanced radix-254 digits, one signed byte each:
//     v = (2^e / 127) * sum_s d_s 254^(-s),   d_s integer in [-127, 127],
// r_0 = 127 v 2^-e (|r_0| < 127), d_s = rint(r_s), r_(s+1) = 254 (r_s - d_s).
// |r_s - d_s| <= 1/2 keeps every later digit in [-127, 127]; r_s - d_s is
// exact, the products by 127 and 254 carry one FP64 rounding each (relative
// 2^-53, far below the truncation 254^-NS).  A digit product is at most
// 127^2 and the INT32 MFMA accumulator of each shift q = s + t collects
// (q+1) K of them, exact while NS K 127^2 < 2^31; the result is recombined
// in FP64:
//     sum_i P A = 2^(eP + eA) / 127^2 * sum_q 254^(-q) acc_q.
// Shifts q >= nslices are dropped (relative truncation ~ 254^-nslices).
//
// Table layouts (m-blocks; kp(m) = ceil((L-m+1)/2) degrees per parity; K
// padded with zero digits to the MFMA depths: 64, or 32 for the synthesis
// tail, which takes one v_mfma_i32_16x16x32_i8):
//   synthesis  [parity][slice][nhp = roundup(nh,16)][kpad = roundup(kp,32)]
//              rows = latitude, K = degree index within the parity class
//   analysis   [parity][slice][mpad = roundup(kp,16)][nhk = roundup(nh,64)]
//              rows = degree within the parity class, K = latitude
#pragma once

#include "plan.hpp"

namespace aeo {

typedef int intx4 __attribute__((ext_vector_type(4)));   // 16 INT8 digits or 4 INT32 sums

constexpr int kSplitK = 64;              // K depth of one v_mfma_i32_16x16x64_i8
constexpr double kDigitTop = 127.0;      // first digit scale
constexpr double kDigitRadix = 254.0;
constexpr double kInvRadix = 1.0 / 254.0;
constexpr double kInvTop2 = 1.0 / (127.0 * 127.0);

__host__ __device__ inline int rup(int x, int a) { return (x + a - 1) / a * a; }
__host__ __device__ inline int split_kp(int L, int m) { return (L - m + 2) / 2; }
__host__ __device__ inline int split_kpad(int L, int m) { return rup(split_kp(L, m), 32); }
__host__ __device__ inline int split_nhk(int nh) { return rup(nh, kSplitK); }

// Exponent e with max < 2^e (frexp convention; 0 for an all-zero row).
__device__ inline int exp_above(double mx) {
    if (!(mx > 0.0)) return 0;
    int e;
    (void)frexp(mx, &e);
    return e;
}

// The scalings are written as fma(x, c, 0): a rounded product that the
// compiler cannot fuse with the following subtraction, so every build and
// every member tiling yields the same digits.
template <int NS>
__device__ inline void digits(double v, int e, int* d) {
    double r = fma(ldexp(v, -e), kDigitTop, 0.0);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const double q = rint(r);
        d[s] = (int)q;
        r = fma(r - q, kDigitRadix, 0.0);
    }
}

__device__ inline void digits_rt(double v, int e, int ns, int* d) {
    double r = fma(ldexp(v, -e), kDigitTop, 0.0);
    for (int s = 0; s < ns; ++s) {
        const double q = rint(r);
        d[s] = (int)q;
        r = fma(r - q, kDigitRadix, 0.0);
    }
}

// Two digits as one 16-bit word (lower address = first digit).
__device__ inline unsigned short digit_pair(int d0, int d1) {
    return (unsigned short)((d0 & 0xff) | ((d1 & 0xff) << 8));
}

// INT32 accumulation stays exact: nslices * K * 127^2 < 2^31.
inline bool split_exact(int nslices, int kmax) {
    return (long long)nslices * kmax * 16129LL < (1LL << 31);
}

}  // namespace aeo
