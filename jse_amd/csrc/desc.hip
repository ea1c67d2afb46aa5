// Spherical-Chebyshev descriptor, forward and backward, FP32, one wavefront
// per atom.
//
// Forward (reference: basis_SphericalChebyshev.hpp:264-351, the L2 power
// spectrum of basis_SphericalUtil.hpp:378-394):
//   per neighbour j (r < rc):  c_j[p] = fc(r) * sum_n W[type_j][p][n] T_n(1-2r/rc),
//                              Y_j[lm] real orthonormal harmonics
//   A[p][lm] = sum_j c_j[p] Y_j[lm]          (per-atom rank-n contraction)
//   fp[p*(lmax+1)+l] = 4pi/(2l+1) sum_m A[p][l^2+l+m]^2
//   x = (fp - mu) / sigma                    (nnap_main.hpp:274-279)
// Lanes first compute one neighbour each (c_j, Y_j -> LDS); then the P x M
// output is cut into TP x TM register tiles, NT tiles per neighbour group and
// G = 64/NT groups that take every G-th neighbour, and the G partial tiles
// are summed in a fixed order through LDS.
//
// Backward (reference: basis_SphericalChebyshev.hpp:353-465): with
// Gc[p][l] = 2*4pi/(2l+1) * dE/dfp[p][l] (from the MLP pass) the gradient of
// E with respect to A is Gc[p][l(m)] * A[p][m].  Because c[p] = fc sum_n
// W[t][p][n] Rn[n] is linear in Rn, the per-atom dE/dA is first taken back
// through each species' mixing, Gt[n][m] = sum_p W[t][p][n] Gc[p][l] A[p][m]
// (LDS, K x M per species), and one lane per neighbour of species t forms
//   gY[m] = fc sum_n Rn[n] Gt[n][m],  gct[n] = sum_m Gt[n][m] Y[m],
// dE/dRn[n] = fc gct[n], dE/dfc = sum_n Rn[n] gct[n]: 2KM multiply-adds per
// pair instead of 2PM; then the chain rule through fc, Rn and
// (r, theta, phi) gives g_j = -dE_i/d(r_j - r_i).  The pole
// convention of the reference (dxy == 0 -> cos phi = 1, sin phi = 0, the phi
// term dropped; basis_SphericalUtil.hpp:165-181) is kept.
#include "common.h"
#include "desc_gen.h"
#include "device_math.h"
#include "kernels.h"

namespace jnn {

template <int LMAX, int K, int P, int MIX, int TP0 = 6>
struct DescShape {
    static constexpr int L1 = LMAX + 1;
    static constexpr int M = L1 * L1;
    static constexpr int S = P * L1;
    static constexpr int TP = P >= TP0 ? TP0 : P;
    static constexpr int NPB = (P + TP - 1) / TP;
    static constexpr int TM = 7;
    static constexpr int NMB = (M + TM - 1) / TM;
    static constexpr int NT = NPB * NMB;
    static constexpr int G = NT >= 64 ? 1 : 64 / NT;
    static_assert(NT <= 64, "descriptor tile layout needs NT <= 64");
    static constexpr int RED = G * NPB * TP * NMB * TM;  // reduction buffer floats
    static constexpr int YBUF = (64 * M > RED) ? 64 * M : RED;
};

// The plain mode's forward on the FP64 matrix cores (v_mfma_f64_16x16x4_f64).
// Each lane evaluates one neighbour from its FP64 displacement (d + dlo): the
// pair geometry, Y, Rn, fc and c in FP64.  c goes to LDS as C^T[p][j]; Y stays
// in registers and is staged one 16-row m-tile at a time as Y^T[m][j]; each
// m-tile runs ceil(n/4) k-steps of A^T += Y^T(16 m x 4 neighbours) C(4
// neighbours x 16 p) against the p-tiles of C, all in FP64.  The epilogue
// writes A[p][m] (FP64) to LDS, stores A as FP32 for the backward and forms fp
// and x in FP64.  Lane maps of the FP64 instruction: A[i][k] from lane
// i + 16 k, B[k][j] from lane j + 16 k (as in the FP32 one), and D[i][j] in
// lane j + 16 (i % 4), element i / 4 (the FP32 one has lane j + 16 (i / 4),
// element i % 4).  Rows of 65 doubles put the 16 lanes of each
// quarter-wave of an operand read on distinct bank pairs.  With the default
// mixing the last m-tile carries the rows w_t Y[m] as in the FP32 kernel.
template <int LMAX, int K, int P, int MIX>
__global__ __launch_bounds__(64) void k_desc_fwd_acc(SegmentSet set, NeighborList nl) {
    constexpr int L1 = LMAX + 1, M = L1 * L1, S = P * L1;
    constexpr int NMT = (M + 15) / 16, NPT = (P + 15) / 16;
    constexpr int MLAST = M - 16 * (NMT - 1);  // data rows of the last m-tile
    constexpr bool WROWS = MIX == MIX_DEFAULT && P == 2 * K && K <= 16 && 2 * MLAST <= 16;
    constexpr int YLAST = WROWS ? 2 * MLAST : MLAST;  // operand rows of the last m-tile
    constexpr int NS = 65;
    constexpr int FC = (P + 1) * NS, FY = 16 * NS, FA = P * M;  // C^T rows, zero row P
    constexpr int F = FC + (FY > FA ? FY : FA);
    constexpr int NE = (S + 63) / 64;
    __shared__ double sbuf[F];
    double *sC = sbuf, *sY = sbuf + FC, *sA = sbuf + FC;
    using d4 = __attribute__((__vector_size__(4 * sizeof(double)))) double;
    constexpr d4 kZero = {0.0, 0.0, 0.0, 0.0};
    const int lane = threadIdx.x;
    const int li = lane & 15, lk = lane >> 4;
    sC[P * NS + lane] = 0.0;  // the zero row
    if (lane < NS - 64) sC[P * NS + 64 + lane] = 0.0;
    int boff[NPT];
#pragma unroll
    for (int pt = 0; pt < NPT; ++pt) boff[pt] = (pt * 16 + li < P ? pt * 16 + li : P) * NS;
    const int ylast = li < YLAST ? li * NS : (P * NS - FC);
    double ew[NE], emu[NE], eis[NE];
    double rc = 0.0, inv_rc = 0.0;
    int cur = -1;

    for (int gr = blockIdx.x; gr < set.row0[set.nseg]; gr += gridDim.x) {
        int sg = 0;  // the row's segment (species)
        while (sg + 1 < set.nseg && gr >= set.row0[sg + 1]) ++sg;
        const SegmentArgs &a = set.a[sg];
        const int row = gr - set.row0[sg];
        if (sg != cur) {
            cur = sg;
            rc = a.rc64;
            inv_rc = 1.0 / a.rc64;
#pragma unroll
            for (int q = 0; q < NE; ++q) {
                const int e = lane + 64 * q;
                ew[q] = 4.0 * 3.14159265358979323846 / (double)(2 * (e % L1) + 1);
                emu[q] = e < S ? a.mu64[e] : 0.0;
                eis[q] = e < S ? 1.0 / a.sig64[e] : 0.0;
            }
        }
        const int i = a.nl_of ? a.nl_of[a.rows_atom[row]] : a.rows_atom[row];
        const int n = min(nl.cnt[i], nl.cap);
        const size_t base = (size_t)i * nl.cap;
        d4 acc[NMT][NPT];
#pragma unroll
        for (int mt = 0; mt < NMT; ++mt)
#pragma unroll
            for (int pt = 0; pt < NPT; ++pt) acc[mt][pt] = kZero;

        for (int c0 = 0; c0 < n; c0 += 64) {
            // lanes past the last neighbour evaluate a valid pair with c = 0
            const int j = c0 + lane;
            const int jr = j < n ? j : n - 1;
            const float4 dd = nl.d[base + jr];
            const int tj = role_species(tag_species(dd.w), a.self_t, a.mirror);
            const PairGeomD g = pair_geom64(dd, pair_lo(nl, base + jr));
            double y[M];
            real_sph<LMAX>(g.ct, g.st, g.cp, g.sp, y);
            double rn[K];
            cheb_rn<K>(1.0 - 2.0 * g.r * inv_rc, rn);
            const double t = 1.0 - (g.r * inv_rc) * (g.r * inv_rc);
            const double t2 = t * t;
            const double fc = (j < n && g.r < rc) ? t2 * t2 : 0.0;
            double rnp[P];
            mix_rnp<K, P, MIX>(rn, tj, a.W64, rnp);
            const int nks = (min(64, n - c0) + 3) >> 2;
            const double wt = (tj & 1) ? (double)tj : -(double)tj;  // the default mixing's weight
            __syncthreads();  // the previous chunk's operand reads (and the previous atom's A) are done
#pragma unroll
            for (int p = 0; p < P; ++p) sC[p * NS + lane] = fc * rnp[p];
#pragma unroll
            for (int mt = 0; mt < NMT; ++mt) {
                if (mt > 0) __syncthreads();  // the previous m-tile's operand reads are done
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    if (mt * 16 + r < M) sY[r * NS + lane] = y[mt * 16 + r];
                    else if (WROWS && mt == NMT - 1 && r < YLAST) sY[r * NS + lane] = wt * y[mt * 16 + r - MLAST];
                }
                const int npt = (WROWS && mt == NMT - 1) ? 1 : NPT;
                __syncthreads();
                const int aoff = mt < NMT - 1 ? li * NS : ylast;
                int ks = 0;
                for (; ks + 4 <= nks; ks += 4) {
                    double av[4], bv[4][NPT];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int col = 4 * (ks + u) + lk;
                        av[u] = sY[aoff + col];
#pragma unroll
                        for (int pt = 0; pt < npt; ++pt) bv[u][pt] = sC[boff[pt] + col];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int pt = 0; pt < npt; ++pt)
                            acc[mt][pt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bv[u][pt], acc[mt][pt], 0, 0, 0);
                }
                for (; ks < nks; ++ks) {  // the last (shorter) batch
                    const int col = 4 * ks + lk;
                    const double av = sY[aoff + col];
#pragma unroll
                    for (int pt = 0; pt < npt; ++pt)
                        acc[mt][pt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, sC[boff[pt] + col], acc[mt][pt], 0, 0, 0);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int mt = 0; mt < NMT; ++mt)
#pragma unroll
            for (int pt = 0; pt < NPT; ++pt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int ti = 4 * r + lk;  // row of the tile
                    int m = mt * 16 + ti, p = pt * 16 + li;
                    bool keep = (mt < NMT - 1 || ti < MLAST) && (pt < NPT - 1 || p < P);
                    if (WROWS && mt == NMT - 1) {
                        // data rows give A[n][m] for n < K, the w_t Y rows A[K + n][m]
                        keep = pt == 0 && li < K && ti < 2 * MLAST;
                        if (ti >= MLAST) {
                            m -= MLAST;
                            p += K;
                        }
                    }
                    if (keep) sA[p * M + m] = acc[mt][pt][r];
                }
        __syncthreads();
        for (int e = lane; e < P * M; e += 64) a.A[(size_t)row * (P * M) + e] = (float)sA[e];
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            const int e = lane + 64 * q;
            if (e < S) {
                const int p = e / L1, l = e % L1;
                double s = 0.0;
                for (int m = l * l; m < (l + 1) * (l + 1); ++m) {
                    const double v = sA[p * M + m];
                    s += v * v;
                }
                store_x(a, (size_t)row * a.xs + e, ((ew[q] * s - emu[q]) * eis[q]));
            }
        }
    }
}

// Forward of the FP32-geometry modes on the matrix cores
// (v_mfma_f32_16x16x4_f32), A accumulated in FP32.  Per atom, A^T (M x P) =
// sum over k-steps of
// Y^T (16 m x 4 neighbours) * C (4 neighbours x 16 p).  Lanes first compute one
// neighbour each: c_j goes to LDS as C^T[p][j]; Y_j stays in registers and is
// staged into LDS one 16-row m-tile at a time as Y^T[m][j].  Each m-tile then
// runs ceil(n/4) k-steps against the p-tiles of C, four k-steps per batch of
// operand reads.  Lanes of padding rows (m >= M, p >= P) read one zero row of
// LDS; lanes past the last neighbour evaluate a valid pair with c = 0.  The
// epilogue writes the tiles to LDS as A[p][m] and forms fp and x in FP64.
// Lane maps of the 16x16x4 instruction: A[i][k] from lane i + 16 k,
// B[k][j] from lane j + 16 k, D[i][j] in lane j + 16 (i / 4), register i % 4.
// A row stride of 68 floats puts the 64 lanes of every operand read on 64
// distinct banks.
// With the default mixing (c[K + n] = w_t c[n]) and at most 8 data rows in
// the last m-tile, that tile also carries the rows w_t Y[m] below its data
// rows: against the columns c[n] of p-tile 0 they give A[K + n][m], so the
// last m-tile needs no other p-tile.
template <int LMAX, int K, int P, int MIX>
__global__ __launch_bounds__(64) void k_desc_fwd_mf(SegmentSet set, NeighborList nl) {
    using AT = float;
    constexpr int L1 = LMAX + 1, M = L1 * L1, S = P * L1;
    constexpr int NMT = (M + 15) / 16, NPT = (P + 15) / 16;
    constexpr int MLAST = M - 16 * (NMT - 1);  // data rows of the last m-tile
    constexpr bool WROWS = MIX == MIX_DEFAULT && P == 2 * K && K <= 16 && 2 * MLAST <= 16;
    constexpr int YLAST = WROWS ? 2 * MLAST : MLAST;  // operand rows of the last m-tile
    constexpr int NS = 68;
    constexpr int FC = (P + 1) * NS, FY = 16 * NS, FA = P * M;  // C^T rows, zero row P
    constexpr int F = FC + (FY > FA ? FY : FA);
    constexpr int NE = (S + 63) / 64;
    __shared__ __align__(16) float sbuf[F];
    float *sC = sbuf, *sY = sbuf + FC;
    AT *sA = reinterpret_cast<AT *>(sbuf + FC);
    using f4 = __attribute__((__vector_size__(4 * sizeof(float)))) float;
    constexpr f4 kZero = {0.f, 0.f, 0.f, 0.f};
    const int lane = threadIdx.x;
    const int li = lane & 15, lk = lane >> 4;
    sC[P * NS + lane] = 0.f;  // the zero row
    if (lane < NS - 64) sC[P * NS + 64 + lane] = 0.f;
    // operand rows of this lane: C^T row pt * 16 + li (or the zero row) and, in the
    // last m-tile, Y^T row li (or the zero row, addressed from sY)
    int boff[NPT];
#pragma unroll
    for (int pt = 0; pt < NPT; ++pt) boff[pt] = (pt * 16 + li < P ? pt * 16 + li : P) * NS;
    const int ylast = li < YLAST ? li * NS : (P * NS - FC);
    // per-lane descriptor slots e = lane + 64 q: 4 pi / (2l + 1), mu, 1 / sigma of the current segment
    double ew[NE], emu[NE], eis[NE];
    float rc = 0.f, inv_rc = 0.f;
    int cur = -1;

    for (int gr = blockIdx.x; gr < set.row0[set.nseg]; gr += gridDim.x) {
        int sg = 0;  // the row's segment (species)
        while (sg + 1 < set.nseg && gr >= set.row0[sg + 1]) ++sg;
        const SegmentArgs &a = set.a[sg];
        const int row = gr - set.row0[sg];
        if (sg != cur) {
            cur = sg;
            rc = a.rc;
            inv_rc = 1.f / a.rc;
#pragma unroll
            for (int q = 0; q < NE; ++q) {
                const int e = lane + 64 * q;
                ew[q] = 4.0 * 3.14159265358979323846 / (double)(2 * (e % L1) + 1);
                emu[q] = e < S ? a.mu64[e] : 0.0;
                eis[q] = e < S ? 1.0 / a.sig64[e] : 0.0;
            }
        }
        const int i = a.nl_of ? a.nl_of[a.rows_atom[row]] : a.rows_atom[row];
        const int n = min(nl.cnt[i], nl.cap);
        const size_t base = (size_t)i * nl.cap;
        f4 acc[NMT][NPT];
#pragma unroll
        for (int mt = 0; mt < NMT; ++mt)
#pragma unroll
            for (int pt = 0; pt < NPT; ++pt) acc[mt][pt] = kZero;

        for (int c0 = 0; c0 < n; c0 += 64) {
            const int j = c0 + lane;
            const int jr = j < n ? j : n - 1;
            const float4 dd = nl.d[base + jr];
            const int tj = role_species(tag_species(dd.w), a.self_t, a.mirror);
            const PairGeomR g = pair_geom_fast(dd.x, dd.y, dd.z);
            float y[M];
            real_sph<LMAX>(g.ct, g.st, g.cp, g.sp, y);
            float rn[K];
            cheb_rn<K>(1.f - 2.f * g.r * inv_rc, rn);
            const float t = 1.f - (g.r * inv_rc) * (g.r * inv_rc);
            const float t2 = t * t;
            const float fc = (j < n && g.r < rc) ? t2 * t2 : 0.f;
            float rnp[P];
            mix_rnp<K, P, MIX>(rn, tj, a.W, rnp);
            const int nks = (min(64, n - c0) + 3) >> 2;
            const float wt = (tj & 1) ? (float)tj : -(float)tj;  // the default mixing's weight
            __syncthreads();  // the previous chunk's operand reads (and the previous atom's A) are done
#pragma unroll
            for (int p = 0; p < P; ++p) sC[p * NS + lane] = fc * rnp[p];
#pragma unroll
            for (int mt = 0; mt < NMT; ++mt) {
                if (mt > 0) __syncthreads();  // the previous m-tile's operand reads are done
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    if (mt * 16 + r < M) sY[r * NS + lane] = y[mt * 16 + r];
                    else if (WROWS && mt == NMT - 1 && r < YLAST) sY[r * NS + lane] = wt * y[mt * 16 + r - MLAST];
                }
                // p-tiles this m-tile runs (a constant once the m-tile loop is unrolled)
                const int npt = (WROWS && mt == NMT - 1) ? 1 : NPT;
                __syncthreads();
                const int aoff = mt < NMT - 1 ? li * NS : ylast;
                int ks = 0;
                for (; ks + 4 <= nks; ks += 4) {
                    float av[4], bv[4][NPT];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int col = 4 * (ks + u) + lk;
                        av[u] = sY[aoff + col];
#pragma unroll
                        for (int pt = 0; pt < npt; ++pt) bv[u][pt] = sC[boff[pt] + col];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int pt = 0; pt < npt; ++pt)
                            acc[mt][pt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u], bv[u][pt], acc[mt][pt], 0, 0, 0);
                }
                for (; ks < nks; ++ks) {  // the last (shorter) batch
                    const int col = 4 * ks + lk;
                    const float av = sY[aoff + col];
#pragma unroll
                    for (int pt = 0; pt < npt; ++pt)
                        acc[mt][pt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, sC[boff[pt] + col], acc[mt][pt], 0, 0, 0);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int mt = 0; mt < NMT; ++mt)
#pragma unroll
            for (int pt = 0; pt < NPT; ++pt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = 4 * lk + r;  // row of the tile
                    int m = mt * 16 + i, p = pt * 16 + li;
                    bool keep = (mt < NMT - 1 || i < MLAST) && (pt < NPT - 1 || p < P);
                    if (WROWS && mt == NMT - 1) {
                        // data rows give A[n][m] for n < K, the w_t Y rows A[K + n][m]
                        keep = pt == 0 && li < K && i < 2 * MLAST;
                        if (i >= MLAST) {
                            m -= MLAST;
                            p += K;
                        }
                    }
                    if (keep) sA[p * M + m] = acc[mt][pt][r];
                }
        __syncthreads();
        for (int e = lane; e < P * M; e += 64) a.A[(size_t)row * (P * M) + e] = (float)sA[e];
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            const int e = lane + 64 * q;
            if (e < S) {
                const int p = e / L1, l = e % L1;
                double s = 0.0;
                for (int m = l * l; m < (l + 1) * (l + 1); ++m) {
                    const double v = sA[p * M + m];
                    s += v * v;
                }
                store_x(a, (size_t)row * a.xs + e, ((ew[q] * s - emu[q]) * eis[q]));
            }
        }
    }
}

// Number of per-species dE/dA blocks the backward keeps in LDS: mixing
// without weights (P = K, one block) or one block per species.
template <int MIX>
__host__ __device__ constexpr int bwd_blocks(int ntypes) {
    return MIX == MIX_NONE ? 1 : ntypes;
}
// LDS row length of a block (K rows of M values): M rounded up to 4, so each
// lane reads its block 16 bytes at a time with immediate offsets
__host__ __device__ constexpr int bwd_row(int m) { return (m + 3) & ~3; }

// BAR: a scheduling barrier after every BAR rows of the per-pair contraction
// (bounds how many rows of LDS operands are in flight; 0: none)
// ACC (the plain mode): each pair's geometry, Y, Rn, fc and their derivatives
// are evaluated in FP64 from the FP64 displacement (d + dlo); Y, Rn and fc are
// rounded to FP32 for the dE/dA contraction, which stays FP32; the radial and
// theta derivatives and the final chain rule are FP64.
template <int LMAX, int K, int P, int MIX, int BAR, bool ACC>
__global__ __launch_bounds__(64, 2) void k_desc_bwd(SegmentSet set, NeighborList nl) {
    using GT = typename std::conditional<ACC, double, float>::type;  // geometry precision
    using SH = DescShape<LMAX, K, P, MIX>;
    constexpr int BSTEP = BAR > 0 ? BAR : 1;
    constexpr int M = SH::M, S = SH::S, L1 = SH::L1;
    constexpr int MP = bwd_row(M), MB = MP / 4;
    // sGt[t][n][m] = sum_p W[t][p][n] Gc[p][l(m)] A[p][m]: dE/dA taken back
    // through species t's radial mixing, so a neighbour of species t needs
    // K (not P) rows; rows of MP floats (m >= M unused)
    extern __shared__ float4 sGt4[];
    float *sGt = reinterpret_cast<float *>(sGt4);
    const int lane = threadIdx.x;
    const int nblk = bwd_blocks<MIX>(set.a[0].ntypes);
    const int total = set.row0[set.nseg];
    const int accum = set.a[0].accum_g;  // add into g (a merge basis' later parts; one segment then)
    // The next row's Gc and A inputs of the dE/dA fold (mixing without weights or the
    // default mixing) are loaded while this row's pairs are worked on: pg[it][h], pa[it][h]
    // for slot e = lane + 64 it and half h (p = q or K + q).
    constexpr bool PF = MIX != MIX_DENSE;
    constexpr int NIT = (K * M + 63) / 64;
    constexpr int NH = MIX == MIX_DEFAULT ? 2 : 1;
    float pg[PF ? NIT : 1][NH], pa[PF ? NIT : 1][NH];
    auto prefetch = [&](int gr2) {
        if (!PF || gr2 >= total) return;
        int s2 = 0;
        while (s2 + 1 < set.nseg && gr2 >= set.row0[s2 + 1]) ++s2;
        const SegmentArgs &b = set.a[s2];
        const int row2 = gr2 - set.row0[s2];
        const float *__restrict__ A2 = b.A + (size_t)row2 * (P * M);
        const float *__restrict__ G2 = b.G + (size_t)row2 * b.xs;
#pragma unroll
        for (int it = 0; it < (PF ? NIT : 1); ++it) {
            const int e = lane + 64 * it;
            const int q = e / M, m = e % M;
            const int l = (int)__builtin_amdgcn_sqrtf((float)m + 0.5f);  // l(m) = floor(sqrt(m))
#pragma unroll
            for (int h = 0; h < NH; ++h) {
                pg[it][h] = e < K * M ? G2[(h * K + q) * L1 + l] : 0.f;
                pa[it][h] = e < K * M ? A2[(h * K + q) * M + m] : 0.f;
            }
        }
    };
    // Each wave works on one contiguous range of rows as a single stream of
    // pairs: a 64-lane chunk takes the rest of row A and, when that is short of
    // 64, the first pairs of the next row B, whose fold sits in the other of two
    // LDS slots.  The fold inputs of the next row are loaded after the
    // displacements of the following chunk (or right before its fold).
    const int slot_f4 = nblk * K * MB;  // float4 per slot
    const int per = (total + (int)gridDim.x - 1) / (int)gridDim.x;
    const int rbeg = min(total, (int)blockIdx.x * per), rend = min(total, rbeg + per);
    struct RowRef {
        int gr, nlrow, n;
        GT rc, inv_rc;
        int self_t, mirror;
    };
    auto row_ref = [&](int gr) {
        RowRef r{gr, 0, 0, (GT)1, (GT)1, 0, 0};
        if (gr >= rend) return r;
        int sg = 0;  // the row's segment (species)
        while (sg + 1 < set.nseg && gr >= set.row0[sg + 1]) ++sg;
        const SegmentArgs &a = set.a[sg];
        const int row = gr - set.row0[sg];
        r.nlrow = a.nl_of ? a.nl_of[a.rows_atom[row]] : a.rows_atom[row];
        r.n = min(nl.cnt[r.nlrow], nl.cap);
        if constexpr (ACC) {
            r.rc = a.rc64;
            r.inv_rc = 1.0 / a.rc64;
        } else {
            r.rc = a.rc;
            r.inv_rc = 1.f / a.rc;
        }
        r.self_t = a.self_t;
        r.mirror = a.mirror;
        return r;
    };
    auto next_row = [&](int gr) {  // the first row at or after gr with pairs (gr == rend: none)
        RowRef r = row_ref(gr);
        while (r.gr < rend && r.n == 0) r = row_ref(r.gr + 1);
        return r;
    };
    // sGt of row r into LDS slot `slot` (with PF from the prefetched inputs of r)
    auto fold = [&](const RowRef &r, int slot) {
        float *sG = sGt + (size_t)slot * slot_f4 * 4;
        if constexpr (PF) {
#pragma unroll
            for (int it = 0; it < NIT; ++it) {
                const int e = lane + 64 * it;
                const int w = e + (MP - M) * (e / M);  // (q, m) -> q MP + m
                if (e < K * M) {
                    if (MIX == MIX_NONE) {
                        sG[w] = pg[it][0] * pa[it][0];
                    } else {
                        const float lo = pg[it][0] * pa[it][0];
                        const float hi = pg[it][NH - 1] * pa[it][NH - 1];
                        for (int t = 1; t <= nblk; ++t) {
                            const float wt = (t & 1) ? (float)t : -(float)t;
                            sG[(t - 1) * K * MP + w] = lo + wt * hi;
                        }
                    }
                }
            }
        } else {
            int sg = 0;
            while (sg + 1 < set.nseg && r.gr >= set.row0[sg + 1]) ++sg;
            const SegmentArgs &a = set.a[sg];
            const int row = r.gr - set.row0[sg];
            const float *__restrict__ Arow = a.A + (size_t)row * (P * M);
            const float *__restrict__ Grow = a.G + (size_t)row * a.xs;
            for (int e = lane; e < K * M; e += 64) {
                const int q = e / M, m = e % M;
                const int l = (int)__builtin_amdgcn_sqrtf((float)m + 0.5f);  // l(m) = floor(sqrt(m))
                for (int t = 1; t <= nblk; ++t) {
                    const float *w = a.W + (size_t)(t - 1) * P * K;
                    float sacc = 0.f;
                    for (int p = 0; p < P; ++p) sacc += w[p * K + q] * (Grow[p * L1 + l] * Arow[p * M + m]);
                    sG[(t - 1) * K * MP + q * MP + m] = sacc;
                }
            }
        }
    };
    RowRef A = next_row(rbeg);
    if (A.gr >= rend) return;
    prefetch(A.gr);
    fold(A, 0);
    RowRef N = next_row(A.gr + 1);  // the next row to fold
    int pf = N.gr;                   // row whose inputs are still to be loaded (rend: none)
    RowRef B = A;
    int kA = 0, slotA = 0, slotB = 1;
    bool haveB = false;
    __syncthreads();
    while (true) {
        const int remA = A.n - kA;
        if (remA < 64 && !haveB && N.gr < rend) {  // row B joins this chunk
            if (pf == N.gr) {
                prefetch(N.gr);
                pf = rend;
            }
            B = N;
            slotB = slotA ^ 1;
            __syncthreads();  // the last reads of that slot are done
            fold(B, slotB);
            __syncthreads();
            haveB = true;
            N = next_row(B.gr + 1);
            pf = N.gr;
        }
        const int cntA = min(remA, 64);
        const int takeB = haveB ? min(64 - cntA, B.n) : 0;
        {
            const bool inA = lane < cntA;
            const bool on = inA || lane < cntA + takeB;
            const int j = inA ? kA + lane : lane - cntA;
            const int slot = inA ? slotA : slotB;
            const GT rc = inA ? A.rc : B.rc, inv_rc = inA ? A.inv_rc : B.inv_rc;
            const int self_t = inA ? A.self_t : B.self_t, mirror = inA ? A.mirror : B.mirror;
            const size_t base = (size_t)(inA ? A.nlrow : B.nlrow) * nl.cap;
            const float4 dd = on ? nl.d[base + j] : make_float4(1.f, 0.f, 0.f, 0.f);
            if (pf < rend) {  // issued after this chunk's displacements
                prefetch(pf);
                pf = rend;
            }
            if (on) {
                const int tj = role_species(tag_species(dd.w), self_t, mirror);
                // pair geometry in GT: r, 1/r, cos/sin theta, cos/sin phi, 1/dxy, d
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
                    // Rn and fc rounded to FP32 for the contraction; x, t and the
                    // derivatives are formed again after it (fewer registers held across it)
                    float rn[K], fc;
                    {
                        GT rnt[K];
                        cheb_rn<K>((GT)1 - (GT)2 * gr_ * inv_rc, rnt);
    #pragma unroll
                        for (int q = 0; q < K; ++q) rn[q] = (float)rnt[q];
                        const GT t = (GT)1 - (gr_ * inv_rc) * (gr_ * inv_rc);
                        fc = (float)((t * t) * (t * t));
                    }
                    int blk = slot * slot_f4 + (MIX == MIX_NONE ? 0 : (tj - 1) * K * MB);
                    // opaque to the compiler, so every read below is this base plus an immediate
                    // offset (folding the -1 into the reads would make the offsets negative)
                    __asm__("" : "+v"(blk));
                    const float4 *sG = sGt4 + blk;
                    // gY[m] = fc sum_n Rn[n] Gt[n][m],  gct[n] = sum_m Gt[n][m] Y[m]
                    // (then dE/dRn = fc gct, dE/dfc = sum_n Rn[n] gct[n])
                    float gy[M];
    #pragma unroll
                    for (int m = 0; m < M; ++m) gy[m] = 0.f;
                    float gfc = 0.f;
                    float grn[K];
    #pragma unroll
                    for (int q = 0; q < K; ++q) {
                        const float cq = fc * rn[q];
                        // four interleaved partial sums: shorter FP32 accumulation chains
                        float s4[4] = {0.f, 0.f, 0.f, 0.f};
    #pragma unroll
                        for (int mb = 0; mb < MB; ++mb) {
                            const float4 g4 = sG[q * MB + mb];
                            const float ga[4] = {g4.x, g4.y, g4.z, g4.w};
    #pragma unroll
                            for (int c = 0; c < 4; ++c) {
                                const int m = 4 * mb + c;
                                if (m < M) {
                                    gy[m] += ga[c] * cq;
                                    s4[c] += ga[c] * y[m];
                                }
                            }
                        }
                        const float s = (s4[0] + s4[1]) + (s4[2] + s4[3]);
                        gfc += s * rn[q];
                        grn[q] = s * fc;
                        if (BAR > 0 && q % BSTEP == BSTEP - 1) __builtin_amdgcn_sched_barrier(0);
                    }
                    const GT u = gr_ * inv_rc, t = (GT)1 - u * u;
                    GT drn[K];
                    cheb_drn<K>((GT)1 - (GT)2 * u, inv_rc, drn);
                    GT dEdr = (GT)0;
    #pragma unroll
                    for (int q = 0; q < K; ++q) dEdr += (GT)grn[q] * drn[q];
                    const GT dfc = (GT)(-8) * u * inv_rc * (t * t) * t;
                    dEdr += (GT)gfc * dfc;
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
                put_g(nl, base + j, ox, oy, oz, accum);
            }
        }
        kA += cntA;
        if (kA < A.n) continue;
        if (haveB) {  // A is done: B carries on
            A = B;
            slotA = slotB;
            kA = takeB;
            haveB = false;
            if (kA < A.n) continue;
        }
        // the next row starts a chunk of its own
        if (N.gr >= rend) break;
        if (pf == N.gr) {
            prefetch(N.gr);
            pf = rend;
        }
        A = N;
        slotA ^= 1;
        kA = 0;
        __syncthreads();
        fold(A, slotA);
        __syncthreads();
        N = next_row(A.gr + 1);
        pf = N.gr;
    }
}

// ---- radial-only chebyshev basis (Chebyshev.java; basis_Chebyshev.hpp:9-80) ----
// fp[p] = sum_j fc(r) Rnp[p], Rnp = W[type_j] Rn: no angular part, so the
// descriptor rows have S = P and no A.  One wavefront per atom: each lane sums
// its neighbours' fc Rnp[p] in FP64 and the lanes are combined in a fixed
// order (DPP row rotations, then readlane across the four rows).

__device__ __forceinline__ double dpp_row_ror(double v, int n) {
    const long long b = __double_as_longlong(v);
    int lo = (int)(b & 0xffffffffll), hi = (int)(b >> 32);
    switch (n) {  // row_ror:n is DPP control 0x120 + n
        case 8: lo = __builtin_amdgcn_update_dpp(0, lo, 0x128, 0xf, 0xf, false);
                hi = __builtin_amdgcn_update_dpp(0, hi, 0x128, 0xf, 0xf, false); break;
        case 4: lo = __builtin_amdgcn_update_dpp(0, lo, 0x124, 0xf, 0xf, false);
                hi = __builtin_amdgcn_update_dpp(0, hi, 0x124, 0xf, 0xf, false); break;
        case 2: lo = __builtin_amdgcn_update_dpp(0, lo, 0x122, 0xf, 0xf, false);
                hi = __builtin_amdgcn_update_dpp(0, hi, 0x122, 0xf, 0xf, false); break;
        default: lo = __builtin_amdgcn_update_dpp(0, lo, 0x121, 0xf, 0xf, false);
                 hi = __builtin_amdgcn_update_dpp(0, hi, 0x121, 0xf, 0xf, false); break;
    }
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double readlane_f64(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// the wave's sum, the same in every lane
__device__ __forceinline__ double wave_sum_f64(double v) {
    v += dpp_row_ror(v, 8);
    v += dpp_row_ror(v, 4);
    v += dpp_row_ror(v, 2);
    v += dpp_row_ror(v, 1);
    return ((readlane_f64(v, 0) + readlane_f64(v, 16)) + readlane_f64(v, 32)) + readlane_f64(v, 48);
}

// ACC (the plain mode): the displacement d + dlo and every pair term in FP64.
template <int K, int P, int MIX, bool ACC>
__global__ __launch_bounds__(64) void k_rad_fwd(SegmentSet set, NeighborList nl) {
    using GT = typename std::conditional<ACC, double, float>::type;  // pair arithmetic
    const int lane = threadIdx.x;
    for (int gr = blockIdx.x; gr < set.row0[set.nseg]; gr += gridDim.x) {
        int sg = 0;  // the row's segment (species)
        while (sg + 1 < set.nseg && gr >= set.row0[sg + 1]) ++sg;
        const SegmentArgs &a = set.a[sg];
        const int row = gr - set.row0[sg];
        const GT rc = ACC ? (GT)a.rc64 : (GT)a.rc, inv_rc = (GT)1 / rc;
        const int i = a.nl_of ? a.nl_of[a.rows_atom[row]] : a.rows_atom[row];
        const int n = min(nl.cnt[i], nl.cap);
        const size_t base = (size_t)i * nl.cap;
        const GT *W = nullptr;
        if constexpr (ACC) W = a.W64;
        else W = a.W;
        double acc[P];
#pragma unroll
        for (int p = 0; p < P; ++p) acc[p] = 0.0;
        for (int j = lane; j < n; j += 64) {
            const float4 dd = nl.d[base + j];
            const int tj = role_species(tag_species(dd.w), a.self_t, a.mirror);
            GT r;
            if constexpr (ACC) r = pair_geom64(dd, pair_lo(nl, base + j)).r;
            else r = pair_geom_fast(dd.x, dd.y, dd.z).r;
            if (r < rc) {
                GT rn[K];
                cheb_rn<K>((GT)1 - (GT)2 * r * inv_rc, rn);
                const GT t = (GT)1 - (r * inv_rc) * (r * inv_rc);
                const GT t2 = t * t;
                const GT fc = t2 * t2;
                GT rnp[P];
                mix_rnp<K, P, MIX>(rn, tj, W, rnp);
#pragma unroll
                for (int p = 0; p < P; ++p) acc[p] += (double)(fc * rnp[p]);
            }
        }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const double fp = wave_sum_f64(acc[p]);
            if (lane == p) store_x(a, (size_t)row * a.xs + p, ((fp - a.mu64[p]) * (1.0 / a.sig64[p])));
        }
    }
}

// Lane per pair: with G[p] = dE/dfp[p] of the atom (the MLP's gradient),
// dE/dfc = sum_p G[p] Rnp[p] and dE/dRn[n] = fc sum_p G[p] W[t][p][n], so
// g_j = -(sum_n dE/dRn[n] dRn/dr + dE/dfc dfc/dr) d / r.
template <int K, int P, int MIX, bool ACC>
__global__ __launch_bounds__(64) void k_rad_bwd(SegmentSet set, NeighborList nl) {
    using GT = typename std::conditional<ACC, double, float>::type;  // pair arithmetic
    const int lane = threadIdx.x;
    for (int gr = blockIdx.x; gr < set.row0[set.nseg]; gr += gridDim.x) {
        int sg = 0;
        while (sg + 1 < set.nseg && gr >= set.row0[sg + 1]) ++sg;
        const SegmentArgs &a = set.a[sg];
        const int row = gr - set.row0[sg];
        const GT rc = ACC ? (GT)a.rc64 : (GT)a.rc, inv_rc = (GT)1 / rc;
        const int i = a.nl_of ? a.nl_of[a.rows_atom[row]] : a.rows_atom[row];
        const int n = min(nl.cnt[i], nl.cap);
        const size_t base = (size_t)i * nl.cap;
        const GT *W = nullptr;
        if constexpr (ACC) W = a.W64;
        else W = a.W;
        GT gp[P];
#pragma unroll
        for (int p = 0; p < P; ++p) gp[p] = (GT)a.G[(size_t)row * a.xs + p];
        for (int j = lane; j < n; j += 64) {
            const float4 dd = nl.d[base + j];
            const int tj = role_species(tag_species(dd.w), a.self_t, a.mirror);
            GT r, ir, dx, dy, dz;
            if constexpr (ACC) {
                const PairGeomD g = pair_geom64(dd, pair_lo(nl, base + j));
                r = g.r; ir = g.ir; dx = g.dx; dy = g.dy; dz = g.dz;
            } else {
                const PairGeomR g = pair_geom_fast(dd.x, dd.y, dd.z);
                r = g.r; ir = g.ir; dx = dd.x; dy = dd.y; dz = dd.z;
            }
            float ox = 0.f, oy = 0.f, oz = 0.f;
            if (r < rc) {
                const GT xr = (GT)1 - (GT)2 * r * inv_rc;
                GT rn[K], drn[K];
                cheb_rn<K>(xr, rn);
                cheb_drn<K>(xr, inv_rc, drn);
                const GT t = (GT)1 - (r * inv_rc) * (r * inv_rc);
                const GT t2 = t * t;
                const GT fc = t2 * t2;
                GT rnp[P];
                mix_rnp<K, P, MIX>(rn, tj, W, rnp);
                GT gfc = (GT)0;
#pragma unroll
                for (int p = 0; p < P; ++p) gfc += gp[p] * rnp[p];
                // gw[n] = sum_p G[p] W[t][p][n]
                GT gw[K];
                if (MIX == MIX_NONE) {
#pragma unroll
                    for (int q = 0; q < K; ++q) gw[q] = gp[q];
                } else if (MIX == MIX_DEFAULT) {
                    const GT wt = (tj & 1) ? (GT)tj : -(GT)tj;
#pragma unroll
                    for (int q = 0; q < K; ++q) gw[q] = gp[q] + wt * gp[K + q];
                } else {
                    const GT *w = W + (size_t)(tj - 1) * P * K;
#pragma unroll
                    for (int q = 0; q < K; ++q) {
                        GT sacc = (GT)0;
#pragma unroll
                        for (int p = 0; p < P; ++p) sacc += gp[p] * w[p * K + q];
                        gw[q] = sacc;
                    }
                }
                GT dEdr = (GT)0;
#pragma unroll
                for (int q = 0; q < K; ++q) dEdr += gw[q] * drn[q];
                dEdr *= fc;
                dEdr += gfc * ((GT)(-8) * r * inv_rc * inv_rc * t2 * t);
                ox = (float)(-dEdr * dx * ir);
                oy = (float)(-dEdr * dy * ir);
                oz = (float)(-dEdr * dz * ir);
            }
            put_g(nl, base + j, ox, oy, oz, a.accum_g);
        }
    }
}

template <int K, int P, int MIX>
static void run_rad(hipStream_t s, const SegmentSet &set, const NeighborList &nl, bool backward) {
    const SegmentArgs &a = set.a[0];
    const int rows = set.row0[set.nseg];
    int grid = rows < 262144 ? rows : 262144;
    if (a.wg_per_cu > 0 && grid > a.wg_per_cu * device_cu_count()) grid = a.wg_per_cu * device_cu_count();
    if (a.acc) {
        if (backward) hipLaunchKernelGGL((k_rad_bwd<K, P, MIX, true>), dim3(grid), dim3(64), 0, s, set, nl);
        else hipLaunchKernelGGL((k_rad_fwd<K, P, MIX, true>), dim3(grid), dim3(64), 0, s, set, nl);
    } else {
        if (backward) hipLaunchKernelGGL((k_rad_bwd<K, P, MIX, false>), dim3(grid), dim3(64), 0, s, set, nl);
        else hipLaunchKernelGGL((k_rad_fwd<K, P, MIX, false>), dim3(grid), dim3(64), 0, s, set, nl);
    }
}
#define JNN_RAD_SHAPES(X) \
    X(6, 6, MIX_NONE)      \
    X(6, 12, MIX_DEFAULT)  \
    X(6, 12, MIX_DENSE)    \
    X(9, 9, MIX_NONE)      \
    X(9, 18, MIX_DEFAULT)  \
    X(9, 18, MIX_DENSE)
static bool launch_rad_set(hipStream_t s, int K, int P, int mix, const SegmentSet &set, const NeighborList &nl,
                           bool backward) {
    if (set.nseg <= 0 || set.row0[set.nseg] <= 0) return true;
#define X(K_, P_, M_)                                  \
    if (K == K_ && P == P_ && mix == M_) {            \
        run_rad<K_, P_, M_>(s, set, nl, backward);    \
        return true;                                   \
    }
    JNN_RAD_SHAPES(X)
#undef X
    return false;
}

// ---- dispatch table over the tiled shapes (other shapes: the kernels of desc_gen.hip) ----
template <int LMAX, int K, int P, int MIX>
static void run_fwd(hipStream_t s, const SegmentSet &set, const NeighborList &nl) {
    const SegmentArgs &a = set.a[0];  // launch shape, variant and precision are the set's
    const int rows = set.row0[set.nseg];
    int grid = rows < 262144 ? rows : 262144;
    if (a.wg_per_cu > 0 && grid > a.wg_per_cu * device_cu_count()) grid = a.wg_per_cu * device_cu_count();
    // plain mode: FP64 geometry and A on the FP64 matrix cores; the FP32-geometry
    // modes: FP32 matrix cores with FP32 accumulation
    if (a.acc)
        hipLaunchKernelGGL((k_desc_fwd_acc<LMAX, K, P, MIX>), dim3(grid), dim3(64), 0, s, set, nl);
    else
        hipLaunchKernelGGL((k_desc_fwd_mf<LMAX, K, P, MIX>), dim3(grid), dim3(64), 0, s, set, nl);
}
template <int LMAX, int K, int P, int MIX>
static void run_bwd(hipStream_t s, const SegmentSet &set, const NeighborList &nl) {
    const SegmentArgs &a = set.a[0];
    const int rows = set.row0[set.nseg];
    int grid = rows < 262144 ? rows : 262144;
    if (a.wg_per_cu > 0 && grid > a.wg_per_cu * device_cu_count()) grid = a.wg_per_cu * device_cu_count();
    // two slots: the rows a chunk of pairs can straddle
    const size_t lds = 2 * sizeof(float) * (size_t)bwd_blocks<MIX>(a.ntypes) * K * bwd_row((LMAX + 1) * (LMAX + 1));
    if (a.acc) {
        if (a.variant == 1)
            hipLaunchKernelGGL((k_desc_bwd<LMAX, K, P, MIX, 2, true>), dim3(grid), dim3(64), lds, s, set, nl);
        else if (a.variant == 2)
            hipLaunchKernelGGL((k_desc_bwd<LMAX, K, P, MIX, 0, true>), dim3(grid), dim3(64), lds, s, set, nl);
        else
            hipLaunchKernelGGL((k_desc_bwd<LMAX, K, P, MIX, 1, true>), dim3(grid), dim3(64), lds, s, set, nl);
    } else {
        if (a.variant == 1)
            hipLaunchKernelGGL((k_desc_bwd<LMAX, K, P, MIX, 2, false>), dim3(grid), dim3(64), lds, s, set, nl);
        else if (a.variant == 2)
            hipLaunchKernelGGL((k_desc_bwd<LMAX, K, P, MIX, 0, false>), dim3(grid), dim3(64), lds, s, set, nl);
        else
            hipLaunchKernelGGL((k_desc_bwd<LMAX, K, P, MIX, 1, false>), dim3(grid), dim3(64), lds, s, set, nl);
    }
}

bool launch_desc_fwd_set(hipStream_t s, int lmax, int K, int P, int mix, const SegmentSet &set,
                         const NeighborList &nl) {
    if (set.nseg <= 0 || set.row0[set.nseg] <= 0) return true;
    if (set.a[0].n34 > 0) return launch_desc_gen_set(s, lmax, K, P, set, nl, false);  // L3 / L4 invariants
    if (lmax == kRadialLmax)
        return launch_rad_set(s, K, P, mix, set, nl, false) || launch_desc_gen_set(s, lmax, K, P, set, nl, false);
#define X(L_, K_, P_, M_)                                              \
    if (lmax == L_ && K == K_ && P == P_ && mix == M_) {              \
        run_fwd<L_, K_, P_, M_>(s, set, nl);                          \
        return true;                                                   \
    }
    JNN_SHAPES(X)
#undef X
    return launch_desc_gen_set(s, lmax, K, P, set, nl, false);
}
bool launch_desc_bwd_set(hipStream_t s, int lmax, int K, int P, int mix, const SegmentSet &set,
                         const NeighborList &nl) {
    if (set.nseg <= 0 || set.row0[set.nseg] <= 0) return true;
    if (set.a[0].n34 > 0) return launch_desc_gen_set(s, lmax, K, P, set, nl, true);
    if (lmax == kRadialLmax)
        return launch_rad_set(s, K, P, mix, set, nl, true) || launch_desc_gen_set(s, lmax, K, P, set, nl, true);
#define X(L_, K_, P_, M_)                                              \
    if (lmax == L_ && K == K_ && P == P_ && mix == M_) {              \
        run_bwd<L_, K_, P_, M_>(s, set, nl);                          \
        return true;                                                   \
    }
    JNN_SHAPES(X)
#undef X
    return launch_desc_gen_set(s, lmax, K, P, set, nl, true);
}
bool desc_tiled(int lmax, int K, int P, int mix) {
#define X(L_, K_, P_, M_) \
    if (lmax == L_ && K == K_ && P == P_ && mix == M_) return true;
    JNN_SHAPES(X)
#undef X
    if (lmax == kRadialLmax) {
#define X(K_, P_, M_) \
    if (K == K_ && P == P_ && mix == M_) return true;
        JNN_RAD_SHAPES(X)
#undef X
    }
    return false;
}
static SegmentSet single(const SegmentArgs &a) {
    SegmentSet set{};
    set.a[0] = a;
    set.row0[0] = 0;
    set.row0[1] = a.nrows;
    set.nseg = 1;
    return set;
}
bool launch_desc_fwd(hipStream_t s, int lmax, int K, int P, int mix, const SegmentArgs &a, const NeighborList &nl) {
    return launch_desc_fwd_set(s, lmax, K, P, mix, single(a), nl);
}
bool launch_desc_bwd(hipStream_t s, int lmax, int K, int P, int mix, const SegmentArgs &a, const NeighborList &nl) {
    return launch_desc_bwd_set(s, lmax, K, P, mix, single(a), nl);
}

}  // namespace jnn
