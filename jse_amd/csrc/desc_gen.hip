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
// copy of A in LDS.  fp and x are formed in FP64: per radial channel the L2
// power spectrum, then (l3max / l4max > 0) the L3 and L4 invariants of the
// channel's A row from their coupling terms (basis_SphericalUtil.hpp:395-467;
// A and Y then run to max(lmax, l3max, l4max), the kernels' LMAX).
//
// Backward, one lane per neighbour: dE/dA[p][m] = Gc[p][l(m)] A[p][m] (plus
// the L3 / L4 invariants' terms) goes to LDS once per atom and is folded
// through every species' mixing, Gt[t][n][m]
// = sum_p W[t][p][n] dE/dA[p][m] (c_j is linear in Rn), as in the tiled
// backward; a lane then forms gY[m] = fc sum_n Rn[n] Gt[n][m] and gct[n] =
// sum_m Gt[n][m] Y[m] (2KM multiply-adds per pair for any P, rows read 16
// bytes at a time by every lane at once), dE/dRn = fc gct, dE/dfc = sum_n
// Rn[n] gct[n], and the chain rule of the tiled backward gives g_j =
// -dE_i/d(r_j - r_i).
#include <hip/hip_runtime.h>

#include "common.h"
#include "desc_gen.h"
#include "device_math.h"
#include "kernels.h"

namespace jnn {
namespace {


__device__ __forceinline__ int seg_of(const SegmentSet &set, int gr) {
    int sg = 0;
    while (sg + 1 < set.nseg && gr >= set.row0[sg + 1]) ++sg;
    return sg;
}

// Rnp[p] = sum_{n < K} w[p][n] rn[n] for one mixing row (w = W[t] + p K)
template <typename T>
__device__ __forceinline__ T mix_row(const T *__restrict__ w, int K, const T *rn) {
    T s = (T)0;
#pragma unroll
    for (int q = 0; q < kGenKMax; ++q)
        if (q < K) s += w[q] * rn[q];
    return s;
}

constexpr int kTile = 4;  // A tile edge of the forward

// The centre species' mixing W[ntypes][P][K] is copied into LDS whenever the
// workgroup's rows move to another segment.
__device__ __forceinline__ void stage_w(const SegmentArgs &a, int wsz, float *sW) {
    __syncthreads();  // every lane is done with the previous segment's copy
    for (int e = threadIdx.x; e < wsz; e += blockDim.x) sW[e] = a.W[e];
}
__device__ __forceinline__ void stage_w(const SegmentArgs &a, int wsz, double *sW) {
    __syncthreads();
    for (int e = threadIdx.x; e < wsz; e += blockDim.x) sW[e] = a.W64[e];
}

// LDS: A accumulators (FP64, P*M), then Y (ch x MS values), c (ch x PS
// values), MS and PS the row lengths M and P padded to a multiple of 4, and
// W.  ACC (the plain mode): the displacement d + dlo, the pair geometry, Y,
// Rn, fc, c and W in FP64 (Y written to LDS by the row recurrence, not held
// in registers); otherwise FP32 values whose products are exact in FP64.
constexpr int kFwdThreads = 256;  // four wavefronts per atom in the forward
template <int LMAX, bool ACC>
__global__ __launch_bounds__(kFwdThreads) void k_gen_fwd(SegmentSet set, NeighborList nl, int K, int P, int ch) {
    using ST = typename std::conditional<ACC, double, float>::type;  // staged values
    constexpr int L1 = LMAX + 1, M = L1 * L1, MS = pad4(M);
    constexpr int NMB = MS / kTile;
    extern __shared__ __align__(16) double gen_smem[];
    double *sA = gen_smem;
    const int PS = pad4(P);
    ST *sY = reinterpret_cast<ST *>(sA + pad4(P * M));
    ST *sC = sY + ch * MS;
    ST *sW = sC + ch * PS;
    const int lane = threadIdx.x;
    const int PM = P * M;
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
        const ST rc = ACC ? (ST)a.rc64 : (ST)a.rc, inv_rc = (ST)1 / rc;
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
                ST r;
                if constexpr (ACC) {
                    const PairGeomD g = pair_geom64(dd, pair_lo(nl, base + c0 + lane));
                    r = g.r;
                    real_sph_rows<LMAX, double, double>(g.ct, g.st, g.cp, g.sp, sY + lane * MS);
#pragma unroll
                    for (int m = M; m < MS; ++m) sY[lane * MS + m] = 0.0;
                } else {
                    const PairGeomR g = pair_geom_fast(dd.x, dd.y, dd.z);
                    r = g.r;
                    float y[M];
                    real_sph<LMAX>(g.ct, g.st, g.cp, g.sp, y);
#pragma unroll
                    for (int m = 0; m < MS; ++m) sY[lane * MS + m] = m < M ? y[m] : 0.f;
                }
                ST rn[kGenKMax];
                cheb_rn<kGenKMax>((ST)1 - (ST)2 * r * inv_rc, rn);
                const ST t = (ST)1 - (r * inv_rc) * (r * inv_rc);
                const ST t2 = t * t;
                const ST fc = r < rc ? t2 * t2 : (ST)0;  // pairs beyond this species' cutoff add nothing
                const ST *w = sW + (tj - 1) * P * K;
#pragma unroll 4
                for (int p = 0; p < PS; ++p) sC[lane * PS + p] = p < P ? fc * mix_row(w + p * K, K, rn) : (ST)0;
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
                    double cv[kTile], yv[kTile];
                    if constexpr (ACC) {
                        const double2 c0v = *reinterpret_cast<const double2 *>(sC + j * PS + p0);
                        const double2 c1v = *reinterpret_cast<const double2 *>(sC + j * PS + p0 + 2);
                        const double2 y0v = *reinterpret_cast<const double2 *>(sY + j * MS + m0);
                        const double2 y1v = *reinterpret_cast<const double2 *>(sY + j * MS + m0 + 2);
                        cv[0] = c0v.x; cv[1] = c0v.y; cv[2] = c1v.x; cv[3] = c1v.y;
                        yv[0] = y0v.x; yv[1] = y0v.y; yv[2] = y1v.x; yv[3] = y1v.y;
                    } else {
                        const float4 c4 = *reinterpret_cast<const float4 *>(sC + j * PS + p0);
                        const float4 y4 = *reinterpret_cast<const float4 *>(sY + j * MS + m0);
                        cv[0] = c4.x; cv[1] = c4.y; cv[2] = c4.z; cv[3] = c4.w;
                        yv[0] = y4.x; yv[1] = y4.y; yv[2] = y4.z; yv[3] = y4.w;
                    }
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
        const int sl = a.sl, l2n = a.l2n;
        for (int e = lane; e < P * sl; e += kFwdThreads) {
            const int p = e / sl, l = e - p * sl;
            const double *Ap = sA + p * M;
            double f;
            if (l < l2n) {
                double s = 0.0;
                for (int m = l * l; m < (l + 1) * (l + 1); ++m) s += Ap[m] * Ap[m];
                f = 4.0 * 3.14159265358979323846 / (double)(2 * l + 1) * s;
            } else {  // an L3 / L4 invariant
                const int k = l - l2n;
                f = 0.0;
                for (int q = a.cpl_off[k]; q < a.cpl_off[k + 1]; ++q) {
                    const int4 ix = a.cpl_idx[q];
                    double v = a.cpl_coef[q] * Ap[ix.x] * Ap[ix.y] * Ap[ix.z];
                    if (ix.w >= 0) v *= Ap[ix.w];
                    f += v;
                }
            }
            store_x(a, (size_t)row * a.xs + e, ((f - a.mu64[e]) * (1.0 / a.sig64[e])));
        }
    }
}

// LDS: dE/dA of the current atom (P rows of MP floats, M padded to a multiple
// of 4), its fold through each species' mixing, Gt[t][n][m] = sum_p W[t][p][n]
// dE/dA[p][m] (ntypes x K rows of MP floats), then W.
// ACC (the plain mode): the pair geometry, Y, Rn, fc and the derivatives in
// FP64 from d + dlo; Y, Rn and fc rounded to FP32 for the contraction.
// Two waves per SIMD are asked for up to lmax 6, where they fit in the
// register file; the higher lmax need more registers than two waves hold.
template <int LMAX, bool ACC>
__global__ __launch_bounds__(64, (LMAX <= 6 ? 2 : 1)) void k_gen_bwd(SegmentSet set, NeighborList nl, int K, int P) {
    using GT = typename std::conditional<ACC, double, float>::type;  // geometry precision
    constexpr int L1 = LMAX + 1, M = L1 * L1, MP = pad4(M);
    extern __shared__ __align__(16) double gen_smem[];
    const int nt = set.a[0].ntypes;
    float *sGA = reinterpret_cast<float *>(gen_smem);
    float *sGt = sGA + P * MP;
    float *sW = sGt + nt * K * MP;
    const int lane = threadIdx.x;
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
        const GT rc = ACC ? (GT)a.rc64 : (GT)a.rc, inv_rc = (GT)1 / rc;
        const int i = a.nl_of ? a.nl_of[a.rows_atom[row]] : a.rows_atom[row];
        const int n = min(nl.cnt[i], nl.cap);
        const size_t base = (size_t)i * nl.cap;
        __syncthreads();  // the previous atom's pairs have read sGt
        fill_dEdA(a, row, P, M, MP, sGA);
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
            const int tj = role_species(tag_species(dd.w), a.self_t, a.mirror);
            GT gr_, gir, gct, gst, gcp, gsp, gidxy;
            bool gpole;
            if constexpr (ACC) {
                const PairGeomD g = pair_geom64(dd, pair_lo(nl, base + j));
                gr_ = g.r; gir = g.ir; gct = g.ct; gst = g.st; gcp = g.cp; gsp = g.sp; gidxy = g.idxy;
                gpole = g.pole;
            } else {
                const PairGeomR g = pair_geom_fast(dd.x, dd.y, dd.z);
                gr_ = g.r; gir = g.ir; gct = g.ct; gst = g.st; gcp = g.cp; gsp = g.sp; gidxy = g.idxy;
                gpole = g.pole;
            }
            float ox = 0.f, oy = 0.f, oz = 0.f;
            if (gr_ < rc) {
                float y[M];
                if constexpr (ACC) real_sph_rows<LMAX, double, float>(gct, gst, gcp, gsp, y);
                else real_sph<LMAX>(gct, gst, gcp, gsp, y);
                const GT xr = (GT)1 - (GT)2 * gr_ * inv_rc;
                GT rnt[kGenKMax];
                cheb_rn<kGenKMax>(xr, rnt);
                float rn[kGenKMax];
#pragma unroll
                for (int q = 0; q < kGenKMax; ++q) rn[q] = (float)rnt[q];
                const GT t = (GT)1 - (gr_ * inv_rc) * (gr_ * inv_rc);
                const GT t2 = t * t;
                const float fc = (float)(t2 * t2);
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
                GT drn[kGenKMax];
                cheb_drn<kGenKMax>(xr, inv_rc, drn);
                GT dEdr = (GT)0;
#pragma unroll
                for (int q = 0; q < kGenKMax; ++q)
                    if (q < K) dEdr += (GT)grn[q] * drn[q];
                dEdr += (GT)gfc * ((GT)(-8) * gr_ * inv_rc * inv_rc * t2 * t);
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
                const GT dEdth = dtheta_ladder<LMAX>(gcp, gsp, y, gy);
                const GT ir = gir;
                // g = -(dE/dr d/r + dE/dtheta grad(theta) + dE/dphi grad(phi)), with
                // d/r = (st cp, st sp, ct) and grad(theta) = (ct cp, ct sp, -st) / r
                const GT hr = dEdr * gst + dEdth * gct * ir;
                GT tx = -(hr * gcp);
                GT ty = -(hr * gsp);
                const GT tz = -(dEdr * gct - dEdth * gst * ir);
                if (!gpole) {
                    tx -= (GT)dEdph * (-gsp * gidxy);
                    ty -= (GT)dEdph * (gcp * gidxy);
                }
                ox = (float)tx;
                oy = (float)ty;
                oz = (float)tz;
            }
            put_g(nl, base + j, ox, oy, oz, a.accum_g);
        }
    }
}

// Radial-only basis, forward: chunks of 64 neighbours put c_j[p] = fc Rnp[p]
// into LDS; lane p (and p + 64) sums them in FP64, neighbours in list order.
// LDS: c (64 x P floats), then W.
template <bool ACC>
__global__ __launch_bounds__(64) void k_gen_rad_fwd(SegmentSet set, NeighborList nl, int K, int P) {
    using ST = typename std::conditional<ACC, double, float>::type;  // pair arithmetic and staged values
    extern __shared__ __align__(16) double gen_smem[];
    ST *sC = reinterpret_cast<ST *>(gen_smem);
    ST *sW = sC + 64 * P;
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
        const ST rc = ACC ? (ST)a.rc64 : (ST)a.rc, inv_rc = (ST)1 / rc;
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
                ST r;
                if constexpr (ACC) r = pair_geom64(dd, pair_lo(nl, base + c0 + lane)).r;
                else r = pair_geom_fast(dd.x, dd.y, dd.z).r;
                ST rn[kGenKMax];
                cheb_rn<kGenKMax>((ST)1 - (ST)2 * r * inv_rc, rn);
                const ST t = (ST)1 - (r * inv_rc) * (r * inv_rc);
                const ST t2 = t * t;
                const ST fc = r < rc ? t2 * t2 : (ST)0;
                const ST *w = sW + (tj - 1) * P * K;
                for (int p = 0; p < P; ++p) sC[lane * P + p] = fc * mix_row(w + p * K, K, rn);
            }
            __syncthreads();
            for (int j = 0; j < nc; ++j) {
                if (lane < P) acc0 += (double)sC[j * P + lane];
                if (lane + 64 < P) acc1 += (double)sC[j * P + lane + 64];
            }
        }
        if (lane < P) store_x(a, (size_t)row * a.xs + lane, ((acc0 - a.mu64[lane]) * (1.0 / a.sig64[lane])));
        if (lane + 64 < P)
            store_x(a, (size_t)row * a.xs + lane + 64, ((acc1 - a.mu64[lane + 64]) * (1.0 / a.sig64[lane + 64])));
    }
}

// Radial-only basis, backward: lane per pair, dE/dfc = sum_p G[p] Rnp[p] and
// dE/dRn[n] = fc sum_p G[p] W[t][p][n] with G = dE/dfp of the atom.
// LDS: G (P floats), then W.
template <bool ACC>
__global__ __launch_bounds__(64) void k_gen_rad_bwd(SegmentSet set, NeighborList nl, int K, int P) {
    using ST = typename std::conditional<ACC, double, float>::type;  // pair arithmetic and staged values
    extern __shared__ __align__(16) double gen_smem[];
    ST *sW = reinterpret_cast<ST *>(gen_smem);
    float *sG = reinterpret_cast<float *>(sW + set.a[0].ntypes * P * K);
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
        const ST rc = ACC ? (ST)a.rc64 : (ST)a.rc, inv_rc = (ST)1 / rc;
        const int i = a.nl_of ? a.nl_of[a.rows_atom[row]] : a.rows_atom[row];
        const int n = min(nl.cnt[i], nl.cap);
        const size_t base = (size_t)i * nl.cap;
        __syncthreads();
        for (int p = lane; p < P; p += 64) sG[p] = a.G[(size_t)row * a.xs + p];
        __syncthreads();
        for (int j = lane; j < n; j += 64) {
            const float4 dd = nl.d[base + j];
            const int tj = role_species(tag_species(dd.w), a.self_t, a.mirror);
            ST r, ir, dx, dy, dz;
            if constexpr (ACC) {
                const PairGeomD g = pair_geom64(dd, pair_lo(nl, base + j));
                r = g.r; ir = g.ir; dx = g.dx; dy = g.dy; dz = g.dz;
            } else {
                const PairGeomR g = pair_geom_fast(dd.x, dd.y, dd.z);
                r = g.r; ir = g.ir; dx = dd.x; dy = dd.y; dz = dd.z;
            }
            float ox = 0.f, oy = 0.f, oz = 0.f;
            if (r < rc) {
                const ST xr = (ST)1 - (ST)2 * r * inv_rc;
                ST rn[kGenKMax], drn[kGenKMax];
                cheb_rn<kGenKMax>(xr, rn);
                cheb_drn<kGenKMax>(xr, inv_rc, drn);
                const ST t = (ST)1 - (r * inv_rc) * (r * inv_rc);
                const ST t2 = t * t;
                const ST fc = t2 * t2;
                const ST *w = sW + (tj - 1) * P * K;
                ST gw[kGenKMax];
#pragma unroll
                for (int q = 0; q < kGenKMax; ++q) gw[q] = (ST)0;
                ST gfc = (ST)0;
                for (int p = 0; p < P; ++p) {
                    const ST *wp = w + p * K;
                    const ST gp = (ST)sG[p];
                    gfc += gp * mix_row(wp, K, rn);
#pragma unroll
                    for (int q = 0; q < kGenKMax; ++q)
                        if (q < K) gw[q] += gp * wp[q];
                }
                ST dEdr = (ST)0;
#pragma unroll
                for (int q = 0; q < kGenKMax; ++q)
                    if (q < K) dEdr += gw[q] * drn[q];
                dEdr *= fc;
                dEdr += gfc * ((ST)(-8) * r * inv_rc * inv_rc * t2 * t);
                ox = (float)(-dEdr * dx * ir);
                oy = (float)(-dEdr * dy * ir);
                oz = (float)(-dEdr * dz * ir);
            }
            put_g(nl, base + j, ox, oy, oz, a.accum_g);
        }
    }
}

int gen_grid(const SegmentSet &set) {
    const int rows = set.row0[set.nseg];
    return rows < 262144 ? rows : 262144;
}

// LDS of the forward for a chunk of ch neighbours and wsz mixing weights; v
// is the size of a staged value (4, or 8 in the plain mode)
size_t gen_fwd_lds(int M, int P, int ch, int wsz, size_t v) {
    return sizeof(double) * (size_t)pad4(P * M) + v * ((size_t)ch * (pad4(M) + pad4(P)) + wsz);
}
size_t gen_bwd_lds(int M, int P, int K, int nt) {
    return sizeof(float) * ((size_t)(P + nt * K) * pad4(M) + (size_t)nt * P * K);
}
// the largest forward chunk (64, halved down to 8) that fits, 0 if none
int gen_fwd_chunk(int M, int P, int wsz, size_t v) {
    int ch = 64;
    while (ch > 8 && gen_fwd_lds(M, P, ch, wsz, v) > kGenLds) ch /= 2;
    return gen_fwd_lds(M, P, ch, wsz, v) > kGenLds ? 0 : ch;
}

template <int LMAX, bool ACC>
bool run_gen(hipStream_t s, int K, int P, const SegmentSet &set, const NeighborList &nl, bool backward) {
    const int nt = set.a[0].ntypes;
    const int wsz = nt * P * K;
    constexpr int M = (LMAX + 1) * (LMAX + 1);
    // once per process (thread-safe static initialisation; retried if it throws)
    static const bool attr_set = [] {
        JNN_HIP(hipFuncSetAttribute((const void *)k_gen_fwd<LMAX, ACC>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)kGenLds));
        JNN_HIP(hipFuncSetAttribute((const void *)k_gen_bwd<LMAX, ACC>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)kGenLds));
        return true;
    }();
    (void)attr_set;
    const int grid = gen_grid(set);
    if (backward) {
        const size_t lds = gen_bwd_lds(M, P, K, nt);
        if (lds > kGenLds) return false;
        hipLaunchKernelGGL((k_gen_bwd<LMAX, ACC>), dim3(grid), dim3(64), lds, s, set, nl, K, P);
        return true;
    }
    const size_t v = ACC ? sizeof(double) : sizeof(float);
    const int ch = gen_fwd_chunk(M, P, wsz, v);
    if (!ch) return false;
    hipLaunchKernelGGL((k_gen_fwd<LMAX, ACC>), dim3(grid), dim3(kFwdThreads), gen_fwd_lds(M, P, ch, wsz, v), s, set,
                       nl, K, P, ch);
    return true;
}
template <bool ACC>
bool run_gen_rad(hipStream_t s, int K, int P, const SegmentSet &set, const NeighborList &nl, bool backward) {
    static const bool attr_set = [] {
        JNN_HIP(hipFuncSetAttribute((const void *)k_gen_rad_fwd<ACC>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)kGenLds));
        JNN_HIP(hipFuncSetAttribute((const void *)k_gen_rad_bwd<ACC>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)kGenLds));
        return true;
    }();
    (void)attr_set;
    const size_t v = ACC ? sizeof(double) : sizeof(float);
    const size_t wsz = (size_t)set.a[0].ntypes * P * K;
    const size_t lds = backward ? v * wsz + sizeof(float) * P : v * (64 * (size_t)P + wsz);
    if (lds > kGenLds) return false;
    const int grid = gen_grid(set);
    if (backward) hipLaunchKernelGGL((k_gen_rad_bwd<ACC>), dim3(grid), dim3(64), lds, s, set, nl, K, P);
    else hipLaunchKernelGGL((k_gen_rad_fwd<ACC>), dim3(grid), dim3(64), lds, s, set, nl, K, P);
    return true;
}
template <bool ACC>
bool launch_gen(hipStream_t s, int lmax, int K, int P, const SegmentSet &set, const NeighborList &nl, bool backward) {
    switch (lmax) {
        case kRadialLmax: return run_gen_rad<ACC>(s, K, P, set, nl, backward);
        case 0: return run_gen<0, ACC>(s, K, P, set, nl, backward);
        case 1: return run_gen<1, ACC>(s, K, P, set, nl, backward);
        case 2: return run_gen<2, ACC>(s, K, P, set, nl, backward);
        case 3: return run_gen<3, ACC>(s, K, P, set, nl, backward);
        case 4: return run_gen<4, ACC>(s, K, P, set, nl, backward);
        case 5: return run_gen<5, ACC>(s, K, P, set, nl, backward);
        case 6: return run_gen<6, ACC>(s, K, P, set, nl, backward);
        case 7: return run_gen<7, ACC>(s, K, P, set, nl, backward);
        case 8: return run_gen<8, ACC>(s, K, P, set, nl, backward);
        case 9: return run_gen<9, ACC>(s, K, P, set, nl, backward);
        case 10: return run_gen<10, ACC>(s, K, P, set, nl, backward);
        case 11: return run_gen<11, ACC>(s, K, P, set, nl, backward);
        case 12: return run_gen<12, ACC>(s, K, P, set, nl, backward);
        default: return false;
    }
}

}  // namespace

bool launch_desc_gen_set(hipStream_t s, int lmax, int K, int P, const SegmentSet &set, const NeighborList &nl,
                         bool backward) {
    if (set.nseg <= 0 || set.row0[set.nseg] <= 0) return true;
    if (K < 1 || K > kGenKMax || P < 1 || P > kGenPMax) return false;
    return set.a[0].acc ? launch_gen<true>(s, lmax, K, P, set, nl, backward)
                        : launch_gen<false>(s, lmax, K, P, set, nl, backward);
}

std::string desc_gen_limit(int lmax, int K, int P, int ntypes, bool acc) {
    if (K < 1 || K > kGenKMax || P < 1 || P > kGenPMax)
        return "nmax + 1 <= " + std::to_string(kGenKMax) + " and size_np <= " + std::to_string(kGenPMax);
    const size_t v = acc ? sizeof(double) : sizeof(float);
    const size_t wsz = (size_t)ntypes * P * K;
    const size_t cap = kGenLds;
    if (lmax == kRadialLmax) {
        if (v * (64 * (size_t)P + wsz) > cap || v * wsz + sizeof(float) * P > cap)
            return "the radial descriptor kernels need (64 size_np + ntypes size_np (nmax + 1)) values within 160 KB of LDS";
        return "";
    }
    const int M = (lmax + 1) * (lmax + 1);
    if (!gen_fwd_chunk(M, P, (int)wsz, v))
        return "the descriptor forward needs 8 P M + 8 (M + size_np) + ntypes size_np (nmax + 1) values within 160 KB of LDS";
    if (gen_bwd_lds(M, P, K, ntypes) > cap)
        return "the descriptor backward needs (size_np + ntypes (nmax + 1)) (lmax + 1)^2 + ntypes size_np (nmax + 1) floats within 160 KB of LDS";
    return "";
}

}  // namespace jnn
