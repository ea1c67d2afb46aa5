// Device-side building blocks of the spherical-Chebyshev descriptor (FP32,
// and FP64 where a template parameter says so):
// radial Chebyshev basis, smooth cutoff, real spherical harmonics and their
// theta-derivatives.  The functions define the same quantities as the
// reference's nnap_util.hpp:226-271 and basis_SphericalUtil.hpp:12-182.  The
// theta-derivative rests on the same ladder identity as the reference's
// calYPtheta_ (basis_SphericalUtil.hpp:112-147) but is only ever needed
// contracted with dE/dY, so it is evaluated in the phi-free frame, one row of
// l at a time, and the dY/dtheta array is never formed.
#pragma once
#include <hip/hip_runtime.h>

namespace jnn {

// ---- compile-time coefficient tables (closed forms, Newton sqrt) ----
constexpr double csqrt(double x) {
    if (!(x > 0.0)) return 0.0;
    double c = x > 1.0 ? x : 1.0;
    for (int it = 0; it < 200; ++it) c = 0.5 * (c + x / c);
    return c;
}

constexpr int kMaxL = 12;

template <typename T>
struct ShTables {
    // P_lm = A[l][m] (cos t P_{l-1,m} + B[l][m] P_{l-2,m}) for m <= l-2
    T A[kMaxL + 1][kMaxL + 1];
    T B[kMaxL + 1][kMaxL + 1];
    T D1[kMaxL + 1];  // P_{l,l-1} = D1[l] cos t P_{l-1,l-1},  D1 = sqrt(2l+1)
    T D2[kMaxL + 1];  // P_{l,l}   = D2[l] sin t P_{l-1,l-1},  D2 = -sqrt((2l+1)/(2l))  (Condon-Shortley)
    // ladder identity dP_lm/dt = LB[l][m] P_{l,m+1} - LA[l][m] P_{l,m-1}:
    // LA = sqrt((l+m)(l-m+1))/2, LB = sqrt((l-m)(l+m+1))/2; L0 = sqrt(l(l+1)/2)
    T LA[kMaxL + 1][kMaxL + 1];
    T LB[kMaxL + 1][kMaxL + 1];
    T L0[kMaxL + 1];
    constexpr ShTables() : A(), B(), D1(), D2(), LA(), LB(), L0() {
        for (int l = 0; l <= kMaxL; ++l) {
            D1[l] = (T)csqrt(2.0 * l + 1.0);
            D2[l] = l > 0 ? (T)(-csqrt((2.0 * l + 1.0) / (2.0 * l))) : (T)0;
            L0[l] = (T)csqrt(0.5 * l * (l + 1.0));
            for (int m = 0; m <= kMaxL; ++m) {
                LA[l][m] = m <= l ? (T)(0.5 * csqrt((double)(l + m) * (l - m + 1))) : (T)0;
                LB[l][m] = m <= l ? (T)(0.5 * csqrt((double)(l - m) * (l + m + 1))) : (T)0;
                if (l >= 2 && m <= l - 2) {
                    A[l][m] = (T)csqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
                    B[l][m] = (T)(-csqrt((((double)(l - 1) * (l - 1)) - (double)m * m) /
                                         (4.0 * (l - 1) * (l - 1) - 1.0)));
                } else {
                    A[l][m] = (T)0;
                    B[l][m] = (T)0;
                }
            }
        }
    }
};
static constexpr ShTables<float> kSh{};
static constexpr ShTables<double> kShD{};
template <typename T>
__device__ __forceinline__ constexpr const ShTables<T> &sh_tables() {
    if constexpr (sizeof(T) == 8) return kShD;
    else return kSh;
}

constexpr float kInvSqrt4Pi = 0.28209479177387814f;
constexpr float kSqrt2 = 1.4142135623730951f;
constexpr float kPi4 = 12.566370614359172f;
constexpr double kInvSqrt4PiD = 0.28209479177387814347;
constexpr double kSqrt2D = 1.41421356237309504880;

// Normalised associated Legendre P_lm(cos t) for m >= 0, stored at l*l+l+m
// (T = float or double: the arithmetic precision).
template <int LMAX, typename T>
__device__ __forceinline__ void legendre(T ct, T st, T *y) {
    const ShTables<T> &sh = sh_tables<T>();
    T pll = (T)kInvSqrt4PiD;
    y[0] = pll;
#pragma unroll
    for (int l = 1; l <= LMAX; ++l) {
        const int idx = l * l + l;
#pragma unroll
        for (int m = 0; m <= l - 2; ++m)
            y[idx + m] = sh.A[l][m] * (ct * y[(l - 1) * (l - 1) + (l - 1) + m] + sh.B[l][m] * y[(l - 2) * (l - 2) + (l - 2) + m]);
        y[idx + l - 1] = sh.D1[l] * ct * pll;
        pll *= sh.D2[l] * st;
        y[idx + l] = pll;
    }
}

// Real orthonormal harmonics at index l*l+l+m: m>0 -> sqrt2 P cos(m phi), m<0 -> sqrt2 P sin(|m| phi).
template <int LMAX, typename T>
__device__ __forceinline__ void real_sph(T ct, T st, T cp, T sp, T *y) {
    legendre<LMAX>(ct, st, y);
    T cm = (T)1, sm = (T)0;
#pragma unroll
    for (int m = 1; m <= LMAX; ++m) {
        const T c2 = cm * cp - sm * sp;
        const T s2 = sm * cp + cm * sp;
        cm = c2;
        sm = s2;
#pragma unroll
        for (int l = m; l <= LMAX; ++l) {
            const int idx = l * l + l;
            const T p = (T)kSqrt2D * y[idx + m];
            y[idx + m] = p * cm;
            y[idx - m] = p * sm;
        }
    }
}

// The same harmonics evaluated in T and stored as O (FP64 arithmetic, FP32
// storage: no FP64 copy of the whole set is held).  The recurrence keeps only
// the rows l-1 and l-2 of P.
template <int LMAX, typename T, typename O>
__device__ __forceinline__ void real_sph_rows(T ct, T st, T cp, T sp, O *y) {
    const ShTables<T> &sh = sh_tables<T>();
    T cmv[LMAX + 1], smv[LMAX + 1];
    cmv[0] = (T)1;
    smv[0] = (T)0;
#pragma unroll
    for (int m = 1; m <= LMAX; ++m) {
        cmv[m] = cmv[m - 1] * cp - smv[m - 1] * sp;
        smv[m] = smv[m - 1] * cp + cmv[m - 1] * sp;
    }
    T p1[LMAX + 1] = {}, p2[LMAX + 1] = {};
    T pll = (T)kInvSqrt4PiD;
    p1[0] = pll;
    y[0] = (O)pll;
#pragma unroll
    for (int l = 1; l <= LMAX; ++l) {
        T p[LMAX + 1] = {};
#pragma unroll
        for (int m = 0; m <= l - 2; ++m) p[m] = sh.A[l][m] * (ct * p1[m] + sh.B[l][m] * p2[m]);
        p[l - 1] = sh.D1[l] * ct * pll;
        pll *= sh.D2[l] * st;
        p[l] = pll;
        const int idx = l * l + l;
        y[idx] = (O)p[0];
#pragma unroll
        for (int m = 1; m <= l; ++m) {
            const T q = (T)kSqrt2D * p[m];
            y[idx + m] = (O)(q * cmv[m]);
            y[idx - m] = (O)(q * smv[m]);
        }
#pragma unroll
        for (int m = 0; m <= l; ++m) {
            p2[m] = p1[m];
            p1[m] = p[m];
        }
    }
}

// sum_{lm} g[lm] dY_lm/dtheta from the harmonics themselves, in T.  Both y
// and g are turned to the phi-free frame row by row: p_m = Y_{l,m} cos(m phi)
// + Y_{l,-m} sin(m phi) = sqrt2 P_lm (m > 0), p_0 = sqrt2 Y_l0, and likewise
// gm_m = g_{l,m} cos(m phi) + g_{l,-m} sin(m phi); then the ladder identity
// dP_lm/dtheta = LB P_{l,m+1} - LA P_{l,m-1} (P_{l,-1} = -P_{l,1}, P_{l,l+1} = 0)
// gives the row's term g_l0 L0 p_1 + sum_{m>0} gm_m (LB p_{m+1} - LA p_{m-1}).
// No Legendre recurrence is rerun and dY/dtheta is never formed.
template <int LMAX, typename T, typename Y, typename G>
__device__ __forceinline__ T dtheta_ladder(T cp, T sp, const Y *y, const G *g) {
    const ShTables<T> &sh = sh_tables<T>();
    T cmv[LMAX + 1], smv[LMAX + 1];
    cmv[0] = (T)1;
    smv[0] = (T)0;
#pragma unroll
    for (int m = 1; m <= LMAX; ++m) {
        cmv[m] = cmv[m - 1] * cp - smv[m - 1] * sp;
        smv[m] = smv[m - 1] * cp + cmv[m - 1] * sp;
    }
    T acc = (T)0;
#pragma unroll
    for (int l = 1; l <= LMAX; ++l) {
        const int idx = l * l + l;
        T p[LMAX + 1];
        p[0] = (T)kSqrt2D * (T)y[idx];
#pragma unroll
        for (int m = 1; m <= l; ++m) p[m] = (T)y[idx + m] * cmv[m] + (T)y[idx - m] * smv[m];
        T al = (T)g[idx] * sh.L0[l] * p[1];  // per-l partial sum: shorter accumulation chains
#pragma unroll
        for (int m = 1; m <= l; ++m) {
            const T gm = (T)g[idx + m] * cmv[m] + (T)g[idx - m] * smv[m];
            T dp = -sh.LA[l][m] * p[m - 1];
            if (m < l) dp += sh.LB[l][m] * p[m + 1];
            al += gm * dp;
        }
        acc += al;
    }
    return acc;
}

// Pair geometry from hardware reciprocal square roots (about 1 ulp each): r,
// 1/r, cos/sin theta, cos/sin phi, dxy = |(dx, dy)| and 1/dxy.  dx and dy are
// rescaled by 2^64 when dx^2 + dy^2 would leave the normal range; a pair with
// dxy below 8 FLT_MIN is on the pole (cos phi = 1, sin phi = 0), the
// reference's dxy == 0 convention (basis_SphericalUtil.hpp:165-181) with a
// guard against denormal dxy.
// dY_lm/dtheta of every real harmonic from the harmonics themselves (index
// l^2+l+m), by the ladder identity of the associated Legendre functions
// (dP_lm/dt = LB P_{l,m+1} - LA P_{l,m-1}) written back in the real basis.
template <int LMAX, typename T>
__device__ __forceinline__ void sph_dtheta(const T *y, T cp, T sp, T *yt) {
    const ShTables<T> &sh = sh_tables<T>();
    yt[0] = (T)0;
#pragma unroll
    for (int l = 1; l <= LMAX; ++l) {
        const int idx = l * l + l;
        yt[idx] = sh.L0[l] * (cp * y[idx + 1] + sp * y[idx - 1]);
        yt[idx + 1] = -sh.L0[l] * cp * y[idx];
        yt[idx - 1] = -sh.L0[l] * sp * y[idx];
#pragma unroll
        for (int m = 1; m <= l; ++m) {
            if (m > 1) {
                yt[idx + m] = -sh.LA[l][m] * (cp * y[idx + m - 1] - sp * y[idx - m + 1]);
                yt[idx - m] = -sh.LA[l][m] * (cp * y[idx - m + 1] + sp * y[idx + m - 1]);
            }
            if (m < l) {
                yt[idx + m] += sh.LB[l][m] * (cp * y[idx + m + 1] + sp * y[idx - m - 1]);
                yt[idx - m] += sh.LB[l][m] * (cp * y[idx - m - 1] - sp * y[idx + m + 1]);
            }
        }
    }
}

struct PairGeomR {
    float r, ir, ct, st, cp, sp, dxy, idxy;
    bool pole;
};
__device__ __forceinline__ PairGeomR pair_geom_fast(float dx, float dy, float dz) {
    PairGeomR g;
    const float r2 = dx * dx + dy * dy + dz * dz;
    g.ir = __builtin_amdgcn_rsqf(r2);
    g.r = r2 * g.ir;
    float sx = dx, sy = dy, unscale = 1.f, rescale = 1.f;
    float q = dx * dx + dy * dy;
    if (q < 0x1p-100f) {
        sx = dx * 0x1p+64f;
        sy = dy * 0x1p+64f;
        unscale = 0x1p-64f;
        rescale = 0x1p+64f;
        q = sx * sx + sy * sy;
    }
    const float iq = q > 0.f ? __builtin_amdgcn_rsqf(q) : 0.f;
    g.dxy = (q * iq) * unscale;
    g.pole = !(g.dxy >= 8.f * 1.17549435e-38f);
    g.st = g.dxy * g.ir;
    g.ct = dz * g.ir;
    g.cp = g.pole ? 1.f : sx * iq;
    g.sp = g.pole ? 0.f : sy * iq;
    g.idxy = g.pole ? 0.f : iq * rescale;
    return g;
}

// 1/sqrt(x) in FP64: the hardware estimate and one Newton step (the estimate
// alone is not to FP64 accuracy).
__device__ __forceinline__ double rsq64(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x * y;
    const double e = __builtin_fma(-h, y, 0.5);
    return __builtin_fma(y, e, y);
}
// The pair geometry of pair_geom_fast in FP64, from the displacement held as
// the sum of two floats (hi + lo, exact in FP64).  A pair with dxy below
// 8 DBL_MIN is on the pole; no displacement reaches the FP64 underflow range.
struct PairGeomD {
    double dx, dy, dz, r, ir, ct, st, cp, sp, idxy;
    bool pole;
};
__device__ __forceinline__ PairGeomD pair_geom64(float4 hi, float4 lo) {
    PairGeomD g;
    g.dx = (double)hi.x + (double)lo.x;
    g.dy = (double)hi.y + (double)lo.y;
    g.dz = (double)hi.z + (double)lo.z;
    const double q = g.dx * g.dx + g.dy * g.dy;
    const double r2 = q + g.dz * g.dz;
    g.ir = rsq64(r2);
    g.r = r2 * g.ir;
    const double iq = q > 0.0 ? rsq64(q) : 0.0;
    const double dxy = q * iq;
    g.pole = !(dxy >= 8.0 * 2.2250738585072014e-308);
    g.st = dxy * g.ir;
    g.ct = g.dz * g.ir;
    g.cp = g.pole ? 1.0 : g.dx * iq;
    g.sp = g.pole ? 0.0 : g.dy * iq;
    g.idxy = g.pole ? 0.0 : iq;
    return g;
}

// Chebyshev radial basis Rn = T_n(1 - 2r/rc), n = 0..K-1
template <int K, typename T>
__device__ __forceinline__ void cheb_rn(T x, T *rn) {
    rn[0] = (T)1;
    if (K > 1) rn[1] = x;
#pragma unroll
    for (int q = 2; q < K; ++q) rn[q] = (T)2 * x * rn[q - 1] - rn[q - 2];
}
// dRn/dr = -(2/rc) n U_{n-1}(x)
template <int K, typename T>
__device__ __forceinline__ void cheb_drn(T x, T inv_rc, T *drn) {
    T u[K > 1 ? K - 1 : 1];
    if (K > 1) u[0] = (T)1;
    if (K > 2) u[1] = (T)2 * x;
#pragma unroll
    for (int q = 2; q < K - 1; ++q) u[q] = (T)2 * x * u[q - 1] - u[q - 2];
    drn[0] = (T)0;
#pragma unroll
    for (int q = 1; q < K; ++q) drn[q] = (T)(-2) * inv_rc * (T)q * u[q - 1];
}

}  // namespace jnn
