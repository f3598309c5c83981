// Split-precision digits for the Legendre contractions on BF16 MFMA.
//
// A value v of a row (or column) whose largest magnitude is below 2^e is
// written as
//     v = 2^(e-7) * sum_s d_s 2^(-8 s),   d_s integer in [-128, 128],
// with round-to-nearest digits (r = v 2^(7-e); d_s = rint(r); r = (r - d_s) 256).
// Every digit is an exact BF16 value.  A product of two digits is below 2^14
// and a K-term sum of products of equal shift s + t = q stays below 2^24 when
// (q+1) K 2^14 < 2^24, so the FP32 MFMA accumulator of each shift q is exact;
// the result is recombined in FP64:
//     sum_i P A = 2^(eP + eA - 14) sum_q 2^(-8 q) acc_q.
// Shifts q >= nslices are dropped (relative truncation ~ nslices 2^(-8 nslices)).
//
// Table layouts (m-blocks; kp(m) = ceil((L-m+1)/2) degrees per parity):
//   synthesis  [parity][slice][nhp = roundup(nh,16)][kpad = roundup(kp,32)]
//              rows = latitude, K = degree index within the parity class
//   analysis   [parity][slice][mpad = roundup(kp,16)][nhk = roundup(nh,32)]
//              rows = degree within the parity class, K = latitude
#pragma once

#include "plan.hpp"

namespace aeo {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline int rup(int x, int a) { return (x + a - 1) / a * a; }
__host__ __device__ inline int split_kp(int L, int m) { return (L - m + 2) / 2; }

// Exponent e with max < 2^e (frexp convention; 0 for an all-zero row).
__device__ inline int exp_above(double mx) {
    if (!(mx > 0.0)) return 0;
    int e;
    (void)frexp(mx, &e);
    return e;
}

template <int NS>
__device__ inline void digits(double v, int e, float* d) {
    double r = ldexp(v, 7 - e);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const double q = rint(r);
        d[s] = (float)q;
        r = (r - q) * 256.0;
    }
}

__device__ inline void digits_rt(double v, int e, int ns, float* d) {
    double r = ldexp(v, 7 - e);
    for (int s = 0; s < ns; ++s) {
        const double q = rint(r);
        d[s] = (float)q;
        r = (r - q) * 256.0;
    }
}

// BF16 bit pattern of a digit: an integer of magnitude <= 256 is exact in
// FP32 and in BF16, whose bits are the upper half of the FP32 bits.
__device__ inline unsigned short bf16_bits(float d) {
    return (unsigned short)(__float_as_uint(d) >> 16);
}

// Largest number of FP32-exact accumulations: (nslices) * K * 2^14 < 2^24.
inline bool split_exact(int nslices, int kmax) {
    return (long long)nslices * kmax * 16384LL < (1LL << 24);
}

}  // namespace aeo
