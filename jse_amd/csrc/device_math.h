// Device-side building blocks of the spherical-Chebyshev descriptor (FP32):
// radial Chebyshev basis, smooth cutoff, real spherical harmonics and their
// theta-derivatives.  The functions define the same quantities as the
// reference's nnap_util.hpp:226-271 and basis_SphericalUtil.hpp:12-182; the
// theta-derivative here is obtained by differentiating the Legendre recurrence
// (forward-mode), not from the reference's closed form.
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

struct ShTables {
    // P_lm = A[l][m] (cos t P_{l-1,m} + B[l][m] P_{l-2,m}) for m <= l-2
    float A[kMaxL + 1][kMaxL + 1];
    float B[kMaxL + 1][kMaxL + 1];
    float D1[kMaxL + 1];  // P_{l,l-1} = D1[l] cos t P_{l-1,l-1},  D1 = sqrt(2l+1)
    float D2[kMaxL + 1];  // P_{l,l}   = D2[l] sin t P_{l-1,l-1},  D2 = -sqrt((2l+1)/(2l))  (Condon-Shortley)
    constexpr ShTables() : A(), B(), D1(), D2() {
        for (int l = 0; l <= kMaxL; ++l) {
            D1[l] = (float)csqrt(2.0 * l + 1.0);
            D2[l] = l > 0 ? (float)(-csqrt((2.0 * l + 1.0) / (2.0 * l))) : 0.f;
            for (int m = 0; m <= kMaxL; ++m) {
                if (l >= 2 && m <= l - 2) {
                    A[l][m] = (float)csqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
                    B[l][m] = (float)(-csqrt((((double)(l - 1) * (l - 1)) - (double)m * m) /
                                             (4.0 * (l - 1) * (l - 1) - 1.0)));
                } else {
                    A[l][m] = 0.f;
                    B[l][m] = 0.f;
                }
            }
        }
    }
};
static constexpr ShTables kSh{};

constexpr float kInvSqrt4Pi = 0.28209479177387814f;
constexpr float kSqrt2 = 1.4142135623730951f;
constexpr float kPi4 = 12.566370614359172f;

// Normalised associated Legendre P_lm(cos t) for m >= 0, stored at l*l+l+m.
template <int LMAX>
__device__ __forceinline__ void legendre(float ct, float st, float *y) {
    float pll = kInvSqrt4Pi;
    y[0] = pll;
#pragma unroll
    for (int l = 1; l <= LMAX; ++l) {
        const int idx = l * l + l;
#pragma unroll
        for (int m = 0; m <= l - 2; ++m)
            y[idx + m] = kSh.A[l][m] * (ct * y[(l - 1) * (l - 1) + (l - 1) + m] + kSh.B[l][m] * y[(l - 2) * (l - 2) + (l - 2) + m]);
        y[idx + l - 1] = kSh.D1[l] * ct * pll;
        pll *= kSh.D2[l] * st;
        y[idx + l] = pll;
    }
}

// Real orthonormal harmonics at index l*l+l+m: m>0 -> sqrt2 P cos(m phi), m<0 -> sqrt2 P sin(|m| phi).
template <int LMAX>
__device__ __forceinline__ void real_sph(float ct, float st, float cp, float sp, float *y) {
    legendre<LMAX>(ct, st, y);
    float cm = 1.f, sm = 0.f;
#pragma unroll
    for (int m = 1; m <= LMAX; ++m) {
        const float c2 = cm * cp - sm * sp;
        const float s2 = sm * cp + cm * sp;
        cm = c2;
        sm = s2;
#pragma unroll
        for (int l = m; l <= LMAX; ++l) {
            const int idx = l * l + l;
            const float p = kSqrt2 * y[idx + m];
            y[idx + m] = p * cm;
            y[idx - m] = p * sm;
        }
    }
}

// sum_{lm} g[lm] * dY_lm/dtheta, with dP/dtheta from the differentiated recurrence.
template <int LMAX>
__device__ __forceinline__ float dtheta_contract(float ct, float st, float cp, float sp, const float *g) {
    // rows of P and dP for l-1 and l-2 (m = 0..l)
    float p1[LMAX + 1] = {}, d1[LMAX + 1] = {}, p2[LMAX + 1] = {}, d2[LMAX + 1] = {};
    float cmv[LMAX + 1], smv[LMAX + 1];
    cmv[0] = 1.f;
    smv[0] = 0.f;
#pragma unroll
    for (int m = 1; m <= LMAX; ++m) {
        cmv[m] = cmv[m - 1] * cp - smv[m - 1] * sp;
        smv[m] = smv[m - 1] * cp + cmv[m - 1] * sp;
    }
    float acc = 0.f;
    float pll = kInvSqrt4Pi, dpll = 0.f;
    p1[0] = pll;
    d1[0] = 0.f;
#pragma unroll
    for (int l = 1; l <= LMAX; ++l) {
        float p[LMAX + 1] = {}, d[LMAX + 1] = {};
#pragma unroll
        for (int m = 0; m <= l - 2; ++m) {
            p[m] = kSh.A[l][m] * (ct * p1[m] + kSh.B[l][m] * p2[m]);
            d[m] = kSh.A[l][m] * (ct * d1[m] - st * p1[m] + kSh.B[l][m] * d2[m]);
        }
        p[l - 1] = kSh.D1[l] * ct * pll;
        d[l - 1] = kSh.D1[l] * (ct * dpll - st * pll);
        const float npll = kSh.D2[l] * st * pll;
        dpll = kSh.D2[l] * (st * dpll + ct * pll);
        pll = npll;
        p[l] = pll;
        d[l] = dpll;
        const int idx = l * l + l;
        float al = g[idx] * d[0];  // per-l partial sum: shorter FP32 accumulation chains
#pragma unroll
        for (int m = 1; m <= l; ++m) al += kSqrt2 * d[m] * (g[idx + m] * cmv[m] + g[idx - m] * smv[m]);
        acc += al;
#pragma unroll
        for (int m = 0; m <= l; ++m) {
            p2[m] = p1[m];
            d2[m] = d1[m];
            p1[m] = p[m];
            d1[m] = d[m];
        }
    }
    return acc;
}

// Pair geometry from hardware reciprocal square roots (about 1 ulp each): r,
// 1/r, cos/sin theta, cos/sin phi, dxy = |(dx, dy)| and 1/dxy.  dx and dy are
// rescaled by 2^64 when dx^2 + dy^2 would leave the normal range; a pair with
// dxy below 8 FLT_MIN is on the pole (cos phi = 1, sin phi = 0), the
// reference's dxy == 0 convention (basis_SphericalUtil.hpp:165-181) with a
// guard against denormal dxy.
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

// Chebyshev radial basis Rn = T_n(1 - 2r/rc), n = 0..K-1
template <int K>
__device__ __forceinline__ void cheb_rn(float x, float *rn) {
    rn[0] = 1.f;
    if (K > 1) rn[1] = x;
#pragma unroll
    for (int q = 2; q < K; ++q) rn[q] = 2.f * x * rn[q - 1] - rn[q - 2];
}
// dRn/dr = -(2/rc) n U_{n-1}(x)
template <int K>
__device__ __forceinline__ void cheb_drn(float x, float inv_rc, float *drn) {
    float u[K > 1 ? K - 1 : 1];
    if (K > 1) u[0] = 1.f;
    if (K > 2) u[1] = 2.f * x;
#pragma unroll
    for (int q = 2; q < K - 1; ++q) u[q] = 2.f * x * u[q - 1] - u[q - 2];
    drn[0] = 0.f;
#pragma unroll
    for (int q = 1; q < K; ++q) drn[q] = -2.f * inv_rc * (float)q * u[q - 1];
}

}  // namespace jnn
