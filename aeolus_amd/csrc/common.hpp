// Shared definitions for the aeolus_hip library (gfx950 / CDNA4).
//
// Complex numbers are stored as interleaved FP64 pairs (double2: x = real,
// y = imaginary), which is the memory layout of a complex128 tensor.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/aeolus_hip.h"

namespace aeo {

// ---------------------------------------------------------------------------
// Error handling: every C entry point returns an int status and leaves a
// thread-local message readable through aeo_last_error().
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
const char* get_error();

struct Status {
    int code;
    std::string msg;
};

#define AEO_CHECK_HIP(expr)                                                    \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) {                                                \
            ::aeo::set_error(std::string("HIP error at " __FILE__ ":") +       \
                             std::to_string(__LINE__) + ": " +                 \
                             hipGetErrorString(e_) + " (" #expr ")");          \
            return AEO_ERR_HIP;                                                \
        }                                                                      \
    } while (0)

#define AEO_REQUIRE(cond, code, msg)                                           \
    do {                                                                       \
        if (!(cond)) {                                                         \
            ::aeo::set_error(msg);                                             \
            return (code);                                                     \
        }                                                                      \
    } while (0)

#define AEO_LAUNCH_CHECK()                                                     \
    do {                                                                       \
        hipError_t e_ = hipGetLastError();                                     \
        if (e_ != hipSuccess) {                                                \
            ::aeo::set_error(std::string("kernel launch failed at " __FILE__  \
                                         ":") +                                \
                             std::to_string(__LINE__) + ": " +                 \
                             hipGetErrorString(e_));                           \
            return AEO_ERR_HIP;                                                \
        }                                                                      \
    } while (0)

// ---------------------------------------------------------------------------
// Device complex helpers.
// ---------------------------------------------------------------------------
__host__ __device__ inline double2 c2(double r, double i) { return make_double2(r, i); }
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return c2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return c2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cmul(double2 a, double2 b) {
    return c2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cscale(double2 a, double s) { return c2(a.x * s, a.y * s); }
__host__ __device__ inline double2 cconj(double2 a) { return c2(a.x, -a.y); }
// a * s + c with one rounding per component (written out, so every kernel
// that forms it rounds alike whatever the compiler would contract)
__host__ __device__ inline double2 cfma(double2 a, double s, double2 c) {
    return c2(fma(a.x, s, c.x), fma(a.y, s, c.y));
}

// x(l) + x(l ^ 32) in lane l of a wave (in that order for l < 32):
// v_permlane32_swap exchanges the half-waves in a VALU op, where
// __shfl_xor(x, 32) takes two trips through the LDS crossbar (ds_bpermute)
// per double.  [0] keeps x(l) in lanes 0-31 and takes x(l - 32) in lanes
// 32-63; [1] takes x(l + 32) in lanes 0-31 and keeps x(l) in lanes 32-63.
__device__ inline double sum_xor32(double x) {
    const unsigned lo = (unsigned)__double2loint(x), hi = (unsigned)__double2hiint(x);
    const auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    return __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
}
// x(l) + x(l ^ 16) (in that order where l & 16 is 0): v_permlane16_swap
// exchanges the odd 16-lane rows of its first operand with the even rows of
// its second.
__device__ inline double sum_xor16(double x) {
    const unsigned lo = (unsigned)__double2loint(x), hi = (unsigned)__double2hiint(x);
    const auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    return __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
}

inline int n_basis(int L) { return (L + 1) * (L + 2) / 2; }

// Block size used by the grid (FFT) kernels and the maximum number of
// complex values one thread keeps in registers across an in-place FFT stage.
constexpr int kGridThreads = 256;
constexpr int kPtsChunk = 256;    // points per analysis partial sum (point-set plans)
constexpr int kMaxPerThread = 24;
constexpr int kMaxRadices = 16;

// Mixed-radix factorisation for the LDS FFT (radix 4 first, then 2, 3, 5,
// then any remaining prime handled by a direct-DFT stage).
int fft_factor(int n, int* radices);

}  // namespace aeo
