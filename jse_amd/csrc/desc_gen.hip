// Descriptor kernels for every basis shape the reference can load that has no
// tiled kernel of its own: spherical-Chebyshev with any lmax in [0, 12], nmax
// in [0, 20] (SphericalChebyshev.java:32, WTypeBasis.java:56) and any radial
// mixing (all wtypes go through the dense W[t][P][K] of
// WTypeBasis.updateParameters_, WTypeBasis.java:202-304), and the radial-only
// chebyshev basis with any nmax and mixing.  The quantities are those of the
// tiled kernels (basis_SphericalChebyshev.hpp:264-465, basis_Chebyshev.hpp:
// 9-80); only the work layout differs:
//
// Forward, four wavefronts per atom: chunks of `ch` neighbours, one lane each,
// put c_j[p] = fc W[t_j] Rn and Y_j into LDS (rows padded with zeros to a
// multiple of 4); A (P x M) is cut into 4 x 4 tiles, which the 256 lanes take
// round-robin: per neighbour a tile costs two 16-byte LDS reads and 16 FP64
// multiply-adds (every c * Y product of two floats is exact in FP64), summed
// over the chunk in list order in registers and then added into the FP64
// copy of A in LDS.  fp and x are formed in FP64.
//
// Backward, one lane per neighbour: dE/dA[p][m] = Gc[p][l(m)] A[p][m] goes to
// LDS once per atom and is folded through every species' mixing, Gt[t][n][m]
// = sum_p W[t][p][n] dE/dA[p][m] (c_j is linear in Rn), as in the tiled
// backward; a lane then forms gY[m] = fc sum_n Rn[n] Gt[n][m] and gct[n] =
// sum_m Gt[n][m] Y[m] (2KM multiply-adds per pair for any P, rows read 16
// bytes at a time by every lane at once), dE/dRn = fc gct, dE/dfc = sum_n
// Rn[n] gct[n], and the chain rule of the tiled backward gives g_j =
// -dE_i/d(r_j - r_i).
#include <hip/hip_runtime.h>

#include "common.h"
#include "device_math.h"
#include "kernels.h"

namespace jnn {
namespace {

constexpr int kGenKMax = 21;        // nmax <= 20
constexpr int kGenPMax = 128;       // mixing rows handled here
constexpr size_t kGenLds = 160 * 1024;

__device__ __forceinline__ int seg_of(const SegmentSet &set, int gr) {
    int sg = 0;
    while (sg + 1 < set.nseg && gr >= set.row0[sg + 1]) ++sg;
    return sg;
}

// Rnp[p] = sum_{n < K} w[p][n] rn[n] for one mixing row (w = W[t] + p K)
__device__ __forceinline__ float mix_row(const float *__restrict__ w, int K, const float *rn) {
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < kGenKMax; ++q)
        if (q < K) s += w[q] * rn[q];
    return s;
}

constexpr int kTile = 4;  // A tile edge of the forward
__host__ __device__ constexpr int pad4(int v) { return (v + 3) & ~3; }

// The centre species' mixing W[ntypes][P][K] is copied into LDS whenever the
// workgroup's rows move to another segment.
__device__ __forceinline__ void stage_w(const SegmentArgs &a, int wsz, float *sW) {
    __syncthreads();  // every lane is done with the previous segment's copy
    for (int e = threadIdx.x; e < wsz; e += blockDim.x) sW[e] = a.W[e];
}

// LDS: A accumulators (FP64, P*M), then Y (ch x MS floats), c (ch x PS
// floats), MS and PS the row lengths M and P padded to a multiple of 4, and W.
constexpr int kFwdThreads = 256;  // four wavefronts per atom in the forward
template <int LMAX>
__global__ __launch_bounds__(kFwdThreads) void k_gen_fwd(SegmentSet set, NeighborList nl, int K, int P, int ch) {
    constexpr int L1 = LMAX + 1, M = L1 * L1, MS = pad4(M);
    constexpr int NMB = MS / kTile;
    extern __shared__ __align__(16) double gen_smem[];
    double *sA = gen_smem;
    const int PS = pad4(P);
    float *sY = reinterpret_cast<float *>(sA + pad4(P * M));
    float *sC = sY + ch * MS;
    float *sW = sC + ch * PS;
    const int lane = threadIdx.x;
    const int PM = P * M, S = P * L1;
    const int ntile = (PS / kTile) * NMB;
    const int wsz = set.a[0].ntypes * P * K;
    int cur = -1;
    for (int gr = blockIdx.x; gr < set.row0[set.nseg]; gr += gridDim.x) {
        const int sg = seg_of(set, gr);
        const SegmentArgs &a = set.a[sg];
        const int row = gr - set.row0[sg];
        if (sg != cur) {
            stage_w(a, wsz, sW);
            cur = sg;
        }
        const float rc = a.rc, inv_rc = 1.f / a.rc;
        const int i = a.nl_of ? a.nl_of[a.rows_atom[row]] : a.rows_atom[row];
        const int n = min(nl.cnt[i], nl.cap);
        const size_t base = (size_t)i * nl.cap;
        __syncthreads();  // the previous atom's epilogue has read sA
        for (int e = lane; e < PM; e += kFwdThreads) sA[e] = 0.0;
        for (int c0 = 0; c0 < n; c0 += ch) {
            const int nc = min(ch, n - c0);
            __syncthreads();  // the previous chunk's tiles have read sY / sC (and sA is cleared)
            if (lane < nc) {
                const float4 dd = nl.d[base + c0 + lane];
                const int tj = role_species(tag_species(dd.w), a.self_t, a.mirror);
                const PairGeomR g = pair_geom_fast(dd.x, dd.y, dd.z);
                float y[M];
                real_sph<LMAX>(g.ct, g.st, g.cp, g.sp, y);
#pragma unroll
                for (int m = 0; m < MS; ++m) sY[lane * MS + m] = m < M ? y[m] : 0.f;
                float rn[kGenKMax];
                cheb_rn<kGenKMax>(1.f - 2.f * g.r * inv_rc, rn);
                const float t = 1.f - (g.r * inv_rc) * (g.r * inv_rc);
                const float t2 = t * t;
                const float fc = g.r < rc ? t2 * t2 : 0.f;  // pairs beyond this species' cutoff add nothing
                const float *w = sW + (tj - 1) * P * K;
#pragma unroll 4
                for (int p = 0; p < PS; ++p) sC[lane * PS + p] = p < P ? fc * mix_row(w + p * K, K, rn) : 0.f;
            }
            __syncthreads();
            for (int tl = lane; tl < ntile; tl += kFwdThreads) {
                const int p0 = (tl / NMB) * kTile, m0 = (tl % NMB) * kTile;
                double acc[kTile][kTile];
#pragma unroll
                for (int u = 0; u < kTile; ++u)
#pragma unroll
                    for (int v = 0; v < kTile; ++v) acc[u][v] = 0.0;
#pragma unroll 2
                for (int j = 0; j < nc; ++j) {
                    const float4 c4 = *reinterpret_cast<const float4 *>(sC + j * PS + p0);
                    const float4 y4 = *reinterpret_cast<const float4 *>(sY + j * MS + m0);
                    const double cv[kTile] = {c4.x, c4.y, c4.z, c4.w};
                    const double yv[kTile] = {y4.x, y4.y, y4.z, y4.w};
#pragma unroll
                    for (int u = 0; u < kTile; ++u)
#pragma unroll
                        for (int v = 0; v < kTile; ++v) acc[u][v] += cv[u] * yv[v];
                }
#pragma unroll
                for (int u = 0; u < kTile; ++u)
#pragma unroll
                    for (int v = 0; v < kTile; ++v)
                        if (p0 + u < P && m0 + v < M) sA[(p0 + u) * M + m0 + v] += acc[u][v];
            }
        }
        __syncthreads();
        for (int e = lane; e < PM; e += kFwdThreads) a.A[(size_t)row * PM + e] = (float)sA[e];
        for (int e = lane; e < S; e += kFwdThreads) {
            const int p = e / L1, l = e - p * L1;
            double s = 0.0;
            for (int m = l * l; m < (l + 1) * (l + 1); ++m) {
                const double v = sA[p * M + m];
                s += v * v;
            }
            const double w4 = 4.0 * 3.14159265358979323846 / (double)(2 * l + 1);
            a.x[(size_t)row * S + e] = (float)((w4 * s - a.mu64[e]) * (1.0 / a.sig64[e]));
        }
    }
}

// LDS: dE/dA of the current atom (P rows of MP floats, M padded to a multiple
// of 4), its fold through each species' mixing, Gt[t][n][m] = sum_p W[t][p][n]
// dE/dA[p][m] (ntypes x K rows of MP floats), then W.
template <int LMAX>
__global__ __launch_bounds__(64) void k_gen_bwd(SegmentSet set, NeighborList nl, int K, int P) {
    constexpr int L1 = LMAX + 1, M = L1 * L1, MP = pad4(M);
    extern __shared__ __align__(16) double gen_smem[];
    const int nt = set.a[0].ntypes;
    float *sGA = reinterpret_cast<float *>(gen_smem);
    float *sGt = sGA + P * MP;
    float *sW = sGt + nt * K * MP;
    const int lane = threadIdx.x;
    const int PM = P * M, S = P * L1;
    const int wsz = nt * P * K;
    int cur = -1;
    for (int gr = blockIdx.x; gr < set.row0[set.nseg]; gr += gridDim.x) {
        const int sg = seg_of(set, gr);
        const SegmentArgs &a = set.a[sg];
        const int row = gr - set.row0[sg];
        if (sg != cur) {
            stage_w(a, wsz, sW);
            cur = sg;
        }
        const float rc = a.rc, inv_rc = 1.f / a.rc;
        const int i = a.nl_of ? a.nl_of[a.rows_atom[row]] : a.rows_atom[row];
        const int n = min(nl.cnt[i], nl.cap);
        const size_t base = (size_t)i * nl.cap;
        __syncthreads();  // the previous atom's pairs have read sGt
        for (int e = lane; e < P * MP; e += 64) {
            const int p = e / MP, m = e - p * MP;
            int l = 0;
            while ((l + 1) * (l + 1) <= m) ++l;
            sGA[e] = m < M ? a.G[(size_t)row * S + p * L1 + l] * a.A[(size_t)row * PM + p * M + m] : 0.f;
        }
        __syncthreads();
        for (int e = lane; e < nt * K * MP; e += 64) {
            const int tq = e / MP, m = e - tq * MP;  // tq = t K + n
            const int t = tq / K, q = tq - t * K;
            const float *w = sW + t * P * K + q;
            float acc = 0.f;
#pragma unroll 8
            for (int p = 0; p < P; ++p) acc += w[p * K] * sGA[p * MP + m];
            sGt[e] = acc;
        }
        __syncthreads();
        for (int j = lane; j < n; j += 64) {
            const float4 dd = nl.d[base + j];
            const float dx = dd.x, dy = dd.y, dz = dd.z;
            const int tj = role_species(tag_species(dd.w), a.self_t, a.mirror);
            const PairGeomR g = pair_geom_fast(dx, dy, dz);
            float ox = 0.f, oy = 0.f, oz = 0.f;
            if (g.r < rc) {
                float y[M];
                real_sph<LMAX>(g.ct, g.st, g.cp, g.sp, y);
                const float xr = 1.f - 2.f * g.r * inv_rc;
                float rn[kGenKMax];
                cheb_rn<kGenKMax>(xr, rn);
                const float t = 1.f - (g.r * inv_rc) * (g.r * inv_rc);
                const float t2 = t * t;
                const float fc = t2 * t2;
                const float *gt = sGt + (tj - 1) * K * MP;
                // gY[m] = fc sum_n Rn[n] Gt[n][m],  gct[n] = sum_m Gt[n][m] Y[m]
                // (then dE/dRn = fc gct, dE/dfc = sum_n Rn[n] gct[n])
                float gy[M];
#pragma unroll
                for (int m = 0; m < M; ++m) gy[m] = 0.f;
                float grn[kGenKMax];
                float gfc = 0.f;
#pragma unroll
                for (int q = 0; q < kGenKMax; ++q) {
                    grn[q] = 0.f;
                    if (q < K) {
                        const float4 *g4 = reinterpret_cast<const float4 *>(gt + q * MP);
                        const float cq = fc * rn[q];
                        float s4[4] = {0.f, 0.f, 0.f, 0.f};  // shorter FP32 accumulation chains
#pragma unroll
                        for (int mb = 0; mb < MP / 4; ++mb) {
                            const float4 v4 = g4[mb];
                            const float ga[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
                            for (int c = 0; c < 4; ++c) {
                                const int m = 4 * mb + c;
                                if (m < M) {
                                    gy[m] += ga[c] * cq;
                                    s4[c] += ga[c] * y[m];
                                }
                            }
                        }
                        const float sq = (s4[0] + s4[1]) + (s4[2] + s4[3]);
                        gfc += sq * rn[q];
                        grn[q] = sq * fc;
                    }
                }
                float drn[kGenKMax];
                cheb_drn<kGenKMax>(xr, inv_rc, drn);
                float dEdr = 0.f;
#pragma unroll
                for (int q = 0; q < kGenKMax; ++q)
                    if (q < K) dEdr += grn[q] * drn[q];
                dEdr += gfc * (-8.f * g.r * inv_rc * inv_rc * t2 * t);
                float dEdph = 0.f;
#pragma unroll
                for (int l = 1; l < L1; ++l) {
                    float pl = 0.f;
#pragma unroll
                    for (int m = 1; m <= l; ++m) {
                        const int idx = l * l + l;
                        // dY_{l,m}/dphi = -m Y_{l,-m},  dY_{l,-m}/dphi = m Y_{l,m}
                        pl += (float)m * (gy[idx - m] * y[idx + m] - gy[idx + m] * y[idx - m]);
                    }
                    dEdph += pl;
                }
                const float dEdth = dtheta_contract<LMAX>(g.ct, g.st, g.cp, g.sp, gy);
                const float ir = g.ir;
                // g = -(dE/dr d/r + dE/dtheta grad(theta) + dE/dphi grad(phi))
                ox = -(dEdr * dx * ir + dEdth * g.ct * g.cp * ir);
                oy = -(dEdr * dy * ir + dEdth * g.ct * g.sp * ir);
                oz = -(dEdr * dz * ir - dEdth * g.st * ir);
                if (!g.pole) {
                    ox -= dEdph * (-g.sp * g.idxy);
                    oy -= dEdph * (g.cp * g.idxy);
                }
            }
            nl.g[base + j] = make_float4(ox, oy, oz, 0.f);
        }
    }
}

// Radial-only basis, forward: chunks of 64 neighbours put c_j[p] = fc Rnp[p]
// into LDS; lane p (and p + 64) sums them in FP64, neighbours in list order.
// LDS: c (64 x P floats), then W.
__global__ __launch_bounds__(64) void k_gen_rad_fwd(SegmentSet set, NeighborList nl, int K, int P) {
    extern __shared__ __align__(16) double gen_smem[];
    float *sC = reinterpret_cast<float *>(gen_smem);
    float *sW = sC + 64 * P;
    const int lane = threadIdx.x;
    const int wsz = set.a[0].ntypes * P * K;
    int cur = -1;
    for (int gr = blockIdx.x; gr < set.row0[set.nseg]; gr += gridDim.x) {
        const int sg = seg_of(set, gr);
        const SegmentArgs &a = set.a[sg];
        const int row = gr - set.row0[sg];
        if (sg != cur) {
            stage_w(a, wsz, sW);
            cur = sg;
        }
        const float rc = a.rc, inv_rc = 1.f / a.rc;
        const int i = a.nl_of ? a.nl_of[a.rows_atom[row]] : a.rows_atom[row];
        const int n = min(nl.cnt[i], nl.cap);
        const size_t base = (size_t)i * nl.cap;
        double acc0 = 0.0, acc1 = 0.0;
        for (int c0 = 0; c0 < n; c0 += 64) {
            const int nc = min(64, n - c0);
            __syncthreads();
            if (lane < nc) {
                const float4 dd = nl.d[base + c0 + lane];
                const int tj = role_species(tag_species(dd.w), a.self_t, a.mirror);
                const PairGeomR g = pair_geom_fast(dd.x, dd.y, dd.z);
                float rn[kGenKMax];
                cheb_rn<kGenKMax>(1.f - 2.f * g.r * inv_rc, rn);
                const float t = 1.f - (g.r * inv_rc) * (g.r * inv_rc);
                const float t2 = t * t;
                const float fc = g.r < rc ? t2 * t2 : 0.f;
                const float *w = sW + (tj - 1) * P * K;
                for (int p = 0; p < P; ++p) sC[lane * P + p] = fc * mix_row(w + p * K, K, rn);
            }
            __syncthreads();
            for (int j = 0; j < nc; ++j) {
                if (lane < P) acc0 += (double)sC[j * P + lane];
                if (lane + 64 < P) acc1 += (double)sC[j * P + lane + 64];
            }
        }
        if (lane < P) a.x[(size_t)row * P + lane] = (float)((acc0 - a.mu64[lane]) * (1.0 / a.sig64[lane]));
        if (lane + 64 < P)
            a.x[(size_t)row * P + lane + 64] = (float)((acc1 - a.mu64[lane + 64]) * (1.0 / a.sig64[lane + 64]));
    }
}

// Radial-only basis, backward: lane per pair, dE/dfc = sum_p G[p] Rnp[p] and
// dE/dRn[n] = fc sum_p G[p] W[t][p][n] with G = dE/dfp of the atom.
// LDS: G (P floats), then W.
__global__ __launch_bounds__(64) void k_gen_rad_bwd(SegmentSet set, NeighborList nl, int K, int P) {
    extern __shared__ __align__(16) double gen_smem[];
    float *sG = reinterpret_cast<float *>(gen_smem);
    float *sW = sG + P;
    const int lane = threadIdx.x;
    const int wsz = set.a[0].ntypes * P * K;
    int cur = -1;
    for (int gr = blockIdx.x; gr < set.row0[set.nseg]; gr += gridDim.x) {
        const int sg = seg_of(set, gr);
        const SegmentArgs &a = set.a[sg];
        const int row = gr - set.row0[sg];
        if (sg != cur) {
            stage_w(a, wsz, sW);
            cur = sg;
        }
        const float rc = a.rc, inv_rc = 1.f / a.rc;
        const int i = a.nl_of ? a.nl_of[a.rows_atom[row]] : a.rows_atom[row];
        const int n = min(nl.cnt[i], nl.cap);
        const size_t base = (size_t)i * nl.cap;
        __syncthreads();
        for (int p = lane; p < P; p += 64) sG[p] = a.G[(size_t)row * P + p];
        __syncthreads();
        for (int j = lane; j < n; j += 64) {
            const float4 dd = nl.d[base + j];
            const int tj = role_species(tag_species(dd.w), a.self_t, a.mirror);
            const PairGeomR g = pair_geom_fast(dd.x, dd.y, dd.z);
            float ox = 0.f, oy = 0.f, oz = 0.f;
            if (g.r < rc) {
                const float xr = 1.f - 2.f * g.r * inv_rc;
                float rn[kGenKMax], drn[kGenKMax];
                cheb_rn<kGenKMax>(xr, rn);
                cheb_drn<kGenKMax>(xr, inv_rc, drn);
                const float t = 1.f - (g.r * inv_rc) * (g.r * inv_rc);
                const float t2 = t * t;
                const float fc = t2 * t2;
                const float *w = sW + (tj - 1) * P * K;
                float gw[kGenKMax];
#pragma unroll
                for (int q = 0; q < kGenKMax; ++q) gw[q] = 0.f;
                float gfc = 0.f;
                for (int p = 0; p < P; ++p) {
                    const float *wp = w + p * K;
                    const float gp = sG[p];
                    gfc += gp * mix_row(wp, K, rn);
#pragma unroll
                    for (int q = 0; q < kGenKMax; ++q)
                        if (q < K) gw[q] += gp * wp[q];
                }
                float dEdr = 0.f;
#pragma unroll
                for (int q = 0; q < kGenKMax; ++q)
                    if (q < K) dEdr += gw[q] * drn[q];
                dEdr *= fc;
                dEdr += gfc * (-8.f * g.r * inv_rc * inv_rc * t2 * t);
                ox = -dEdr * dd.x * g.ir;
                oy = -dEdr * dd.y * g.ir;
                oz = -dEdr * dd.z * g.ir;
            }
            nl.g[base + j] = make_float4(ox, oy, oz, 0.f);
        }
    }
}

int gen_grid(const SegmentSet &set) {
    const int rows = set.row0[set.nseg];
    return rows < 262144 ? rows : 262144;
}

// LDS of the forward for a chunk of ch neighbours, wsz floats of W
size_t gen_fwd_lds(int M, int P, int ch, int wsz) {
    return sizeof(double) * (size_t)pad4(P * M) + sizeof(float) * ((size_t)ch * (pad4(M) + pad4(P)) + wsz);
}

template <int LMAX>
bool run_gen(hipStream_t s, int K, int P, const SegmentSet &set, const NeighborList &nl, bool backward) {
    const int wsz = set.a[0].ntypes * P * K;
    constexpr int M = (LMAX + 1) * (LMAX + 1);
    // once per process (thread-safe static initialisation; retried if it throws)
    static const bool attr_set = [] {
        JNN_HIP(hipFuncSetAttribute((const void *)k_gen_fwd<LMAX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)kGenLds));
        JNN_HIP(hipFuncSetAttribute((const void *)k_gen_bwd<LMAX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)kGenLds));
        return true;
    }();
    (void)attr_set;
    const int grid = gen_grid(set);
    if (backward) {
        const size_t lds = sizeof(float) * ((size_t)(P + set.a[0].ntypes * K) * pad4(M) + wsz);
        if (lds > kGenLds) return false;
        hipLaunchKernelGGL((k_gen_bwd<LMAX>), dim3(grid), dim3(64), lds, s, set, nl, K, P);
        return true;
    }
    int ch = 64;  // the largest chunk that fits
    while (ch > 8 && gen_fwd_lds(M, P, ch, wsz) > kGenLds) ch /= 2;
    if (gen_fwd_lds(M, P, ch, wsz) > kGenLds) return false;
    hipLaunchKernelGGL((k_gen_fwd<LMAX>), dim3(grid), dim3(kFwdThreads), gen_fwd_lds(M, P, ch, wsz), s, set, nl, K, P,
                       ch);
    return true;
}

}  // namespace

bool launch_desc_gen_set(hipStream_t s, int lmax, int K, int P, const SegmentSet &set, const NeighborList &nl,
                         bool backward) {
    if (set.nseg <= 0 || set.row0[set.nseg] <= 0) return true;
    if (K < 1 || K > kGenKMax || P < 1 || P > kGenPMax) return false;
    if (lmax == kRadialLmax) {
        static const bool attr_set = [] {
            JNN_HIP(hipFuncSetAttribute((const void *)k_gen_rad_fwd, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)kGenLds));
            JNN_HIP(hipFuncSetAttribute((const void *)k_gen_rad_bwd, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)kGenLds));
            return true;
        }();
        (void)attr_set;
        const size_t wsz = (size_t)set.a[0].ntypes * P * K;
        const size_t lds = sizeof(float) * ((backward ? (size_t)P : 64 * (size_t)P) + wsz);
        if (lds > kGenLds) return false;
        const int grid = gen_grid(set);
        if (backward) hipLaunchKernelGGL(k_gen_rad_bwd, dim3(grid), dim3(64), lds, s, set, nl, K, P);
        else hipLaunchKernelGGL(k_gen_rad_fwd, dim3(grid), dim3(64), lds, s, set, nl, K, P);
        return true;
    }
    switch (lmax) {
        case 0: return run_gen<0>(s, K, P, set, nl, backward);
        case 1: return run_gen<1>(s, K, P, set, nl, backward);
        case 2: return run_gen<2>(s, K, P, set, nl, backward);
        case 3: return run_gen<3>(s, K, P, set, nl, backward);
        case 4: return run_gen<4>(s, K, P, set, nl, backward);
        case 5: return run_gen<5>(s, K, P, set, nl, backward);
        case 6: return run_gen<6>(s, K, P, set, nl, backward);
        case 7: return run_gen<7>(s, K, P, set, nl, backward);
        case 8: return run_gen<8>(s, K, P, set, nl, backward);
        case 9: return run_gen<9>(s, K, P, set, nl, backward);
        case 10: return run_gen<10>(s, K, P, set, nl, backward);
        case 11: return run_gen<11>(s, K, P, set, nl, backward);
        case 12: return run_gen<12>(s, K, P, set, nl, backward);
        default: return false;
    }
}

}  // namespace jnn
