// Coefficient loaders (what a synthesis pass contracts) and analysis
// epilogues (what happens to analysed coefficients: store, truncate, form
// the tendency, advance an RK4 stage).  Shared by the Gauss (FFT + Legendre)
// and point-set (dense basis) back ends.
//
// Reference semantics reproduced here:
//   psi = zeta / lap for l >= 1, psi_0 = 0           phy/barotropic.py:68-93
//   eta = zeta + f, f only at (1,0)                    phy/barotropic.py:60-61,147
//   (S a)_l = C+_{l-1} a_{l-1} + C-_{l+1} a_{l+1}      num/spectral_operators.py:228-241
//   d/dlambda = (i m / R) a                            num/spectral_operators.py:220-226
//   tendency = -J + nu lap' zeta (+ forcing), row 0 = 0   phy/barotropic.py:153-165
//   SWE tendency assembly                              phy/shallow_water.py:389-418
//   RK4 stage arithmetic                               run/bve/runner.py:211-216
#pragma once

#include "common.hpp"

namespace aeo {

// ---------------------------------------------------------------------------
// Shared spectral helpers
// ---------------------------------------------------------------------------
struct SpecK {
    int L;
    double R;
    const double* Cp;   // (L+1)^2
    const double* Cm;   // (L+1)^2
};

// (i q) * a for q = m / R given as a product m * (1/R)
__device__ inline double2 dlam_q(double2 a, double q) { return c2(-q * a.y, q * a.x); }

__device__ inline double2 dlam(double2 a, int m, double R) {
    // (i m / R) * a, as a complex product with (0, m/R)
    const double q = (double)m / R;
    return c2(0.0 * a.x - q * a.y, 0.0 * a.y + q * a.x);
}

// C+ and C- of num/spectral_operators.py:135-176 (clamped numerators and
// denominators, zero for m > l, C+[L,:] = 0, C-[0,:] = 0).
__device__ inline double c_plus(int l, int m, int L) {
    if (m > l || l == L) return 0.0;
    const double Lb = l, Mb = m;
    const double nu = fmax((Lb + 1.0) * (Lb + 1.0) - Mb * Mb, 0.0);
    const double de = fmax((2.0 * Lb + 1.0) * (2.0 * Lb + 3.0), 1.0);
    return Lb * sqrt(nu / de);
}

__device__ inline double c_minus(int l, int m) {
    if (m > l || l == 0) return 0.0;
    const double Lb = l, Mb = m;
    const double nu = fmax(Lb * Lb - Mb * Mb, 0.0);
    const double de = fmax((2.0 * Lb - 1.0) * (2.0 * Lb + 1.0), 1.0);
    return -(Lb + 1.0) * sqrt(nu / de);
}

// (S a)_{l,m} for a field given by a functor base(l, m).
template <class F>
__device__ inline double2 s_apply(const F& base, int l, int m, const SpecK& k) {
    const int n = k.L + 1;
    double2 t = c2(0.0, 0.0);
    if (l >= 1 && m <= l - 1) t = cadd(t, cscale(base(l - 1, m), k.Cp[(l - 1) * n + m]));
    if (l <= k.L - 1) t = cadd(t, cscale(base(l + 1, m), k.Cm[(l + 1) * n + m]));
    return t;
}

// ---------------------------------------------------------------------------
// Loaders.  value(f, l, m) returns field f's coefficient at (l, m), m <= l.
// POINTS = true produces the zonal-derivative fields in coefficient space
// (the dense back end synthesises them directly); POINTS = false leaves the
// (i m / R) factor to the Fourier stage of the Gauss back end.
// ---------------------------------------------------------------------------

// Raw fields a[f][l][m] (generic synthesis).
struct RawLoader {
    static constexpr int NF = 4;
    const double2* a;
    int L;
    int nf;  // valid fields in this group
    __device__ double2 value(int f, int l, int m) const {
        if (f >= nf) return c2(0.0, 0.0);
        const int n = L + 1;
        return a[(size_t)f * n * n + l * n + m];
    }
};

// Jacobian inputs.  BVE mode derives (psi, eta) from zeta; JAC mode takes
// (a, b) as given.  Fields: Gauss  {a, S a, b, S b};
//                           points {(im/R) a, S a, (im/R) b, S b}.
template <bool POINTS>
struct JacLoader {
    static constexpr int NF = 4;
    const double2* za;
    const double2* zb;
    bool bve;
    const double* lap;  // -l(l+1)/R^2
    double f10;
    SpecK k;
    __device__ double2 base(int which, int l, int m) const {
        const int n = k.L + 1;
        if (!bve) return (which == 0 ? za : zb)[l * n + m];
        double2 z = za[l * n + m];
        if (which == 0) {
            if (l == 0) return c2(0.0, 0.0);
            const double d = lap[l];
            return c2(z.x / d, z.y / d);
        }
        if (l == 1 && m == 0) z.x += f10;
        return z;
    }
    __device__ double2 value(int f, int l, int m) const {
        const int which = f >> 1;
        if (f & 1) {
            auto b = [&](int ll, int mm) { return base(which, ll, mm); };
            return s_apply(b, l, m, k);
        }
        double2 v = base(which, l, m);
        return POINTS ? dlam(v, m, k.R) : v;
    }
};

// Shallow-water inputs from the (zeta, delta, phi) stack.
//   Gauss  fields: psi, S psi, chi, S chi, eta, S eta, phi, S phi, zeta, delta
//   points fields: psi_l, S psi, chi_l, S chi, eta_l, S eta, eta, phi_l, S phi,
//                  phi, zeta, delta      (x_l = (i m / R) x)
template <bool POINTS>
struct SweLoader {
    static constexpr int NF = POINTS ? 12 : 10;
    const double2* y;   // 3 stacked fields
    const double* lap;
    double f10;
    SpecK k;
    // base fields: 0 psi, 1 chi, 2 eta, 3 phi, 4 zeta, 5 delta
    __device__ double2 base(int which, int l, int m) const {
        const int n = k.L + 1;
        const size_t fs = (size_t)n * n;
        const int idx = l * n + m;
        switch (which) {
            case 0:
            case 1: {
                if (l == 0) return c2(0.0, 0.0);
                const double2 z = y[(which == 0 ? 0 : fs) + idx];
                const double d = lap[l];
                return c2(z.x / d, z.y / d);
            }
            case 2: {
                double2 z = y[idx];
                if (l == 1 && m == 0) z.x += f10;
                return z;
            }
            case 3: return y[2 * fs + idx];
            case 4: return y[idx];
            default: return y[fs + idx];
        }
    }
    __device__ double2 sv(int which, int l, int m) const {
        auto b = [&](int ll, int mm) { return base(which, ll, mm); };
        return s_apply(b, l, m, k);
    }
    __device__ double2 value(int f, int l, int m) const {
        if constexpr (POINTS) {
            switch (f) {
                case 0: return dlam(base(0, l, m), m, k.R);
                case 1: return sv(0, l, m);
                case 2: return dlam(base(1, l, m), m, k.R);
                case 3: return sv(1, l, m);
                case 4: return dlam(base(2, l, m), m, k.R);
                case 5: return sv(2, l, m);
                case 6: return base(2, l, m);
                case 7: return dlam(base(3, l, m), m, k.R);
                case 8: return sv(3, l, m);
                case 9: return base(3, l, m);
                case 10: return base(4, l, m);
                default: return base(5, l, m);
            }
        } else {
            switch (f) {
                case 0: return base(0, l, m);
                case 1: return sv(0, l, m);
                case 2: return base(1, l, m);
                case 3: return sv(1, l, m);
                case 4: return base(2, l, m);
                case 5: return sv(2, l, m);
                case 6: return base(3, l, m);
                case 7: return sv(3, l, m);
                case 8: return base(4, l, m);
                default: return base(5, l, m);
            }
        }
    }
};

// ---------------------------------------------------------------------------
// Pointwise products on the product grid (one point, real values).
// ---------------------------------------------------------------------------
__device__ inline double clamp_coslat(double c) { return c > 1e-8 ? c : 1e-8; }

// J = (a_s b_l - a_l b_s) / c^2      num/spectral_operators.py:638
__device__ inline double jac_point(double al, double as, double bl, double bs, double c) {
    return (as * bl - al * bs) / (c * c);
}

// The same with the reciprocal q = 1/c^2 of the row (one division per row).
__device__ inline double jac_point_q(double al, double as, double bl, double bs, double q) {
    return (as * bl - al * bs) * q;
}

// phy/shallow_water.py:377-386.  In: psi_l psi_s chi_l chi_s eta_l eta_s eta
// phi_l phi_s phi zeta delta.  Out: zeta_dot, curl, kinetic, phi_dot.
__device__ inline void swe_point(const double* g, double c, double* o) {
    const double psi_l = g[0], psi_s = g[1], chi_l = g[2], chi_s = g[3];
    const double eta_l = g[4], eta_s = g[5], eta = g[6];
    const double phi_l = g[7], phi_s = g[8], phi = g[9];
    const double zeta = g[10], delta = g[11];
    const double u = (psi_s + chi_l) / c;
    const double v = (psi_l - chi_s) / c;
    const double adv_eta = (u * eta_l - v * eta_s) / c;
    o[0] = -(adv_eta + eta * delta);
    o[1] = eta * zeta + (v * eta_l + u * eta_s) / c;
    o[2] = 0.5 * (u * u + v * v);
    const double adv_phi = (u * phi_l - v * phi_s) / c;
    o[3] = -(adv_phi + phi * delta);
}

// swe_point with the row reciprocal ic = 1/c (multiplications only).
__device__ inline void swe_point_ic(const double* g, double ic, double* o) {
    const double psi_l = g[0], psi_s = g[1], chi_l = g[2], chi_s = g[3];
    const double eta_l = g[4], eta_s = g[5], eta = g[6];
    const double phi_l = g[7], phi_s = g[8], phi = g[9];
    const double zeta = g[10], delta = g[11];
    const double u = (psi_s + chi_l) * ic;
    const double v = (psi_l - chi_s) * ic;
    const double adv_eta = (u * eta_l - v * eta_s) * ic;
    o[0] = -(adv_eta + eta * delta);
    o[1] = eta * zeta + (v * eta_l + u * eta_s) * ic;
    o[2] = 0.5 * (u * u + v * v);
    const double adv_phi = (u * phi_l - v * phi_s) * ic;
    o[3] = -(adv_phi + phi * delta);
}

// ---------------------------------------------------------------------------
// Epilogues: apply(l, m, acc) is called exactly once for every (l, m) of the
// dense (L+1)^2 layout, with acc = analysed values (zero outside the
// analysed band and in the m > l padding).
// ---------------------------------------------------------------------------
struct EpiStore {
    static constexpr int NF = 4;
    double2* out;
    int L;
    int nf;
    __device__ void apply(int l, int m, const double2* acc) const {
        const int n = L + 1;
        for (int f = 0; f < nf; ++f) out[(size_t)f * n * n + l * n + m] = acc[f];
    }
};

// Field group g of a launch whose grid y index runs over groups of NF fields:
// a loader or epilogue covering all of the launch's fields is narrowed to its
// group; any other loader or epilogue covers one group and is used as given.
template <class T>
__device__ inline T at_group(const T& x, int) { return x; }

__device__ inline RawLoader at_group(const RawLoader& x, int g) {
    const int f0 = g * RawLoader::NF;
    const int nf = (x.nf - f0 < RawLoader::NF) ? x.nf - f0 : RawLoader::NF;
    return RawLoader{x.a + (size_t)f0 * (x.L + 1) * (x.L + 1), x.L, nf};
}

__device__ inline EpiStore at_group(const EpiStore& x, int g) {
    const int f0 = g * EpiStore::NF;
    const int nf = (x.nf - f0 < EpiStore::NF) ? x.nf - f0 : EpiStore::NF;
    return EpiStore{x.out + (size_t)f0 * (x.L + 1) * (x.L + 1), x.L, nf};
}

struct EpiStore1 {
    static constexpr int NF = 1;
    double2* out;
    int L;
    __device__ void apply(int l, int m, const double2* acc) const {
        out[l * (L + 1) + m] = acc[0];
    }
};

// RK4 stage bookkeeping (run/bve/runner.py:211-216; run/engine.py:427-431):
//   stage 0: ksum = k1;          next = y + (dt/2) k1
//   stage 1: ksum = ksum + 2 k2; next = y + (dt/2) k2
//   stage 2: ksum = ksum + 2 k3; next = y + dt k3
//   stage 3: ksum = ksum + k4;   next = y + (dt/6) ksum
struct Rk4 {
    int stage;      // -1: tendency only (write kout)
    double cst;     // dt/2, dt/2, dt for stages 0..2
    double dt6;     // dt / 6
    const double2* y0;
    double2* ksum;
    double2* ynext;
    double2* kout;
    __device__ void apply(size_t idx, double2 k) const {
        if (stage < 0) {
            kout[idx] = k;
            return;
        }
        const double2 y = y0[idx];
        if (stage == 0) {
            ksum[idx] = k;
            ynext[idx] = cadd(y, cscale(k, cst));
        } else if (stage < 3) {
            const double2 s = cadd(ksum[idx], cscale(k, 2.0));
            ksum[idx] = s;
            ynext[idx] = cadd(y, cscale(k, cst));
        } else {
            const double2 s = cadd(ksum[idx], k);
            ynext[idx] = cadd(y, cscale(s, dt6));
        }
    }
};

// BVE: k = -J + nu lap' zeta_s (+ forcing); row l = 0 pinned to zero.
struct EpiBve {
    static constexpr int NF = 1;
    int L;
    const double2* ys;
    const double2* forcing;
    const double* nulap;
    Rk4 rk;
    __device__ void apply(int l, int m, const double2* acc) const {
        const size_t idx = (size_t)l * (L + 1) + m;
        double2 k = c2(0.0, 0.0);
        if (l > 0) {
            const double2 adv = c2(-acc[0].x, -acc[0].y);
            k = cadd(adv, cscale(ys[idx], nulap[l]));
            if (forcing) k = cadd(k, forcing[idx]);
        }
        rk.apply(idx, k);
    }
};

// SWE (phy/shallow_water.py:389-418).  acc = truncated analyses of
// zeta_dot, curl, kinetic, phi_dot.
struct EpiSwe {
    static constexpr int NF = 4;
    int L;
    const double2* ys;          // stage state (3 stacked fields)
    const double* lap;          // -l(l+1)/R^2
    const double* nu4lap2;      // nu4 * lap^2 (nullptr when nu4 == 0)
    const double2* lap_phi_s;   // nullptr when flat
    const double2* nu4_lap2_phi_s;
    double phi0;
    Rk4 rk;
    __device__ void apply(int l, int m, const double2* acc) const {
        const size_t n = L + 1;
        const size_t fs = n * n;
        const size_t idx = (size_t)l * n + m;
        double2 kz = c2(0.0, 0.0), kd = c2(0.0, 0.0), kp = c2(0.0, 0.0);
        if (l > 0) {
            const double2 delta = ys[fs + idx];
            const double2 phi = ys[2 * fs + idx];
            kz = acc[0];
            kp = csub(acc[3], cscale(delta, phi0));
            kd = csub(acc[1], cscale(cadd(acc[2], phi), lap[l]));
            if (lap_phi_s) kd = csub(kd, lap_phi_s[idx]);
            if (nu4lap2) {
                const double q = nu4lap2[l];
                kz = csub(kz, cscale(ys[idx], q));
                kd = csub(kd, cscale(delta, q));
                kp = csub(kp, cscale(phi, q));
                if (nu4_lap2_phi_s) kp = csub(kp, nu4_lap2_phi_s[idx]);
            }
        }
        rk.apply(idx, kz);
        rk.apply(fs + idx, kd);
        rk.apply(2 * fs + idx, kp);
    }
};

}  // namespace aeo
