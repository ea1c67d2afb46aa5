// Shared pieces of the untiled descriptor kernels (desc_gen.hip) and their
// training derivatives (desc_train.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace jnn {

constexpr int kGenKMax = 21;        // nmax <= 20
constexpr int kGenPMax = 128;       // mixing rows handled here
constexpr size_t kGenLds = 160 * 1024;
__host__ __device__ constexpr int pad4(int v) { return (v + 3) & ~3; }

// The descriptor shapes (lmax, K, P, mixing) with register-tiled kernels
// (desc.hip forward / backward, desc_train.hip tangent); other shapes run the
// untiled kernels.
#define JNN_SHAPES(X)              \
    X(6, 6, 6, MIX_NONE)           \
    X(6, 6, 12, MIX_DEFAULT)       \
    X(6, 6, 12, MIX_DENSE)         \
    X(6, 9, 9, MIX_NONE)           \
    X(6, 9, 18, MIX_DEFAULT)       \
    X(6, 9, 18, MIX_DENSE)         \
    X(4, 6, 6, MIX_NONE)           \
    X(4, 6, 12, MIX_DEFAULT)

// c[p] = fc * sum_n W[t][p][n] Rn[n] for the three mixing specialisations
template <int K, int P, int MIX, typename T>
__device__ __forceinline__ void mix_rnp(const T *rn, int tj, const T *__restrict__ W, T *rnp) {
    if (MIX == MIX_NONE) {
#pragma unroll
        for (int p = 0; p < P; ++p) rnp[p] = rn[p];
    } else if (MIX == MIX_DEFAULT) {
        const T wt = (tj & 1) ? (T)tj : -(T)tj;
#pragma unroll
        for (int q = 0; q < K; ++q) {
            rnp[q] = rn[q];
            rnp[K + q] = wt * rn[q];
        }
    } else {
        const T *w = W + (size_t)(tj - 1) * P * K;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            T s = (T)0;
#pragma unroll
            for (int q = 0; q < K; ++q) s += w[p * K + q] * rn[q];
            rnp[p] = s;
        }
    }
}

// dE/dA of one atom into LDS (P rows of MP floats, zero past M): Gc[p][l]
// A[p][m] on the L2 columns (G carries 2 4pi/(2l+1) dE/dfp there) plus the L3 /
// L4 invariants' terms (G = dE/dfp on their columns); a 64-lane workgroup.
__device__ inline void fill_dEdA(const SegmentArgs &a, int row, int P, int M, int MP, float *sGA) {
    const int PM = P * M;
    for (int e = threadIdx.x; e < P * MP; e += 64) {
        const int p = e / MP, m = e - p * MP;
        int l = 0;
        while ((l + 1) * (l + 1) <= m) ++l;
        float v = 0.f;
        if (m < M) {
            const float *Gp = a.G + (size_t)row * a.xs + p * a.sl;
            const float *Ap = a.A + (size_t)row * PM + p * M;
            if (l < a.l2n) v = Gp[l] * Ap[m];
            if (a.n34) {
                double t = 0.0;
                for (int q = a.cplg_off[m]; q < a.cplg_off[m + 1]; ++q) {
                    const int4 ix = a.cplg_idx[q];
                    double u = a.cplg_coef[q] * (double)Gp[a.l2n + ix.w] * (double)Ap[ix.x] * (double)Ap[ix.y];
                    if (ix.z >= 0) u *= (double)Ap[ix.z];
                    t += u;
                }
                v += (float)t;
            }
        }
        sGA[e] = v;
    }
}

}  // namespace jnn
