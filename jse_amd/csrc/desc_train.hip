// Training derivatives of the descriptor's trainable mixing weights (fuse,
// rfuse, exfuse W[ntypes][P][K]) for every basis shape the untiled kernels
// run (spherical lmax 0-12 with L3 / L4 invariants, radial), FP64 sums in a
// fixed order.
#include <hip/hip_runtime.h>

#include "common.h"
#include "desc_gen.h"
#include "device_math.h"
#include "kernels.h"

namespace jnn {
namespace {

// Training derivative of a trainable mixing W[ntypes][P][K] (fuse, rfuse,
// exfuse: the reference's rGradFpParam, basis_SphericalChebyshev.hpp:353-440
// and basis_Chebyshev.hpp:135-220 with GRAD_PARAM; backwardRnp,
// nnap_util.hpp:282-296), summed over the segment's atoms with weights w:
//   dE/dW[t][p][n] += w_i sum_{j of species t} fc_j Rn_j[n] q_j[p],
//   q_j[p] = sum_m dE_i/dA[p][m] Y_j[m]   (RADIAL: q_j[p] = dE_i/dfp[p]).
// One 64-lane workgroup per atom at a time (persistent).  Per chunk of ch
// neighbours a lane per neighbour puts Y_j (FP64, from the FP64 displacement),
// fc_j Rn_j and the species into LDS; then the lanes form q for every
// (neighbour, p), and lane o of the P K outputs adds, species by species,
// w_i sum_j q_j[p] fc_j Rn_j[n] over the chunk's neighbours in list order to
// its entries of partial[blockIdx.x][ntypes][P][K] (the workgroup's own row,
// zeroed by it first).  Deterministic; reduced by k_pgrad_reduce.
// LDS: Y [ch][MP] and q [ch][P] and fc Rn [ch][K] (FP64), dE/dA [P][MP]
// (FP32), species [ch].
template <int LMAX, bool RADIAL>
__global__ __launch_bounds__(64) void k_gen_wgrad(SegmentArgs a, NeighborList nl, int K, int P, int ch,
                                                  const double *__restrict__ w, double *__restrict__ partial) {
    constexpr int L1 = LMAX + 1, M = L1 * L1, MP = pad4(M);
    extern __shared__ __align__(16) double gen_smem[];
    double *sY = gen_smem;
    double *sQ = sY + (size_t)ch * MP;
    double *sR = sQ + (size_t)ch * P;
    float *sGA = reinterpret_cast<float *>(sR + (size_t)ch * K);
    int *sT = reinterpret_cast<int *>(sGA + (size_t)P * MP);
    const int lane = threadIdx.x;
    const int nt = a.ntypes, PK = P * K;
    double *part = partial + (size_t)blockIdx.x * nt * PK;
    for (int e = lane; e < nt * PK; e += 64) part[e] = 0.0;
    const double rc = a.rc64, inv_rc = 1.0 / rc;
    for (int row = blockIdx.x; row < a.nrows; row += gridDim.x) {
        const int atom = a.rows_atom[row];
        const int i = a.nl_of ? a.nl_of[atom] : atom;
        const int n = min(nl.cnt[i], nl.cap);
        const size_t base = (size_t)i * nl.cap;
        const double wi = w[atom];
        __syncthreads();  // the previous atom's reads of sGA are done
        if constexpr (RADIAL) {
            for (int p = lane; p < P; p += 64) sGA[p] = a.G[(size_t)row * a.xs + p];
        } else {
            fill_dEdA(a, row, P, M, MP, sGA);
        }
        for (int c0 = 0; c0 < n; c0 += ch) {
            const int nc = min(ch, n - c0);
            __syncthreads();  // dE/dA is written; the previous chunk's reads are done
            if (lane < nc) {
                const float4 dd = nl.d[base + c0 + lane];
                sT[lane] = role_species(tag_species(dd.w), a.self_t, a.mirror);
                const PairGeomD g = pair_geom64(dd, pair_lo(nl, base + c0 + lane));
                if constexpr (!RADIAL) real_sph_rows<LMAX, double, double>(g.ct, g.st, g.cp, g.sp, sY + lane * MP);
                double rn[kGenKMax];
                cheb_rn<kGenKMax>(1.0 - 2.0 * g.r * inv_rc, rn);
                const double t = 1.0 - (g.r * inv_rc) * (g.r * inv_rc);
                const double fc = g.r < rc ? t * t * t * t : 0.0;  // pairs beyond this basis' cutoff add nothing
#pragma unroll
                for (int q = 0; q < kGenKMax; ++q)  // unrolled: rn stays in registers
                    if (q < K) sR[lane * K + q] = fc * rn[q];
            }
            __syncthreads();
            for (int e = lane; e < nc * P; e += 64) {
                const int j = e / P, p = e - j * P;
                double s;
                if constexpr (RADIAL) {
                    s = (double)sGA[p];
                } else {
                    const float *gp = sGA + p * MP;
                    const double *y = sY + j * MP;
                    s = 0.0;
                    for (int m = 0; m < M; ++m) s += (double)gp[m] * y[m];
                }
                sQ[e] = s;
            }
            __syncthreads();
            for (int o = lane; o < PK; o += 64) {
                const int p = o / K, q = o - p * K;
                for (int t = 1; t <= nt; ++t) {
                    double s = 0.0;
                    bool any = false;
                    for (int j = 0; j < nc; ++j)
                        if (sT[j] == t) {
                            s += sQ[j * P + p] * sR[j * K + q];
                            any = true;
                        }
                    if (any) part[(size_t)(t - 1) * PK + o] += wi * s;
                }
            }
        }
    }
}

// ---- the force / virial loss ----
// One pair's values for the tangent and its adjoint: the pair's upstream
// direction u = bf_j - bf_i + sum_a bv_ab d_a (bv of atom i's replica in a batch), Y and D_u Y (the angular
// gradient along u: (u.theta_hat)/r dY/dtheta + (u.phi_hat)/(r sin theta)
// dY/dphi, the phi term dropped at the pole as in the backward), R = fc Rn and
// D_u R = (u.d/r) dR/dr, and the neighbour's species; beyond the cutoff every
// value is zero.  FP64 from the FP64 displacement.
template <int LMAX, bool RADIAL>
__device__ __forceinline__ void pair_tangent(const SegmentArgs &a, const NeighborList &nl, const TangentArgs &ta, size_t k,
                             int iatom, int K, double *y, double *yd, double *R, double *Rd, int &tj) {
    constexpr int L1 = LMAX + 1, M = L1 * L1;
    const float4 dd = nl.d[k];
    tj = role_species(tag_species(dd.w), a.self_t, a.mirror);
    const PairGeomD g = pair_geom64(dd, pair_lo(nl, k));
    const double rc = a.rc64, inv_rc = 1.0 / rc;
    if (!(g.r < rc)) {
        if constexpr (!RADIAL)
            for (int m = 0; m < M; ++m) y[m] = yd[m] = 0.0;
#pragma unroll
        for (int q = 0; q < kGenKMax; ++q)
            if (q < K) R[q] = Rd[q] = 0.0;
        return;
    }
    double u[3] = {0.0, 0.0, 0.0};
    if (ta.bf) {
        const int jatom = nl.key[k] >> 5;
        for (int c = 0; c < 3; ++c) u[c] = ta.bf[3 * jatom + c] - ta.bf[3 * iatom + c];
    }
    const double d[3] = {g.dx, g.dy, g.dz};
    const int ca[9] = {0, 1, 2, 0, 0, 1, 1, 2, 2}, cb[9] = {0, 1, 2, 1, 2, 2, 0, 0, 1};
    const double *bv = ta.bv_rep ? ta.bv_rep + 9 * (size_t)ta.rep_of[iatom] : ta.bv;
    for (int q = 0; q < 9; ++q) u[cb[q]] += bv[q] * d[ca[q]];
    const double ud = (u[0] * g.dx + u[1] * g.dy + u[2] * g.dz) * g.ir;
    const double xr = 1.0 - 2.0 * g.r * inv_rc;
    double rn[kGenKMax], drn[kGenKMax];
    cheb_rn<kGenKMax>(xr, rn);
    cheb_drn<kGenKMax>(xr, inv_rc, drn);
    const double t = 1.0 - (g.r * inv_rc) * (g.r * inv_rc);
    const double t2 = t * t;
    const double fc = t2 * t2, dfc = -8.0 * g.r * inv_rc * inv_rc * t2 * t;
#pragma unroll
    for (int q = 0; q < kGenKMax; ++q)  // unrolled: rn and drn stay in registers
        if (q < K) {
            R[q] = fc * rn[q];
            Rd[q] = ud * (dfc * rn[q] + fc * drn[q]);
        }
    if constexpr (!RADIAL) {
        real_sph_rows<LMAX, double, double>(g.ct, g.st, g.cp, g.sp, y);
        double yt[M];
        sph_dtheta<LMAX, double>(y, g.cp, g.sp, yt);
        const double uth = (g.ct * g.cp * u[0] + g.ct * g.sp * u[1] - g.st * u[2]) * g.ir;
        const double uph = g.pole ? 0.0 : (-g.sp * u[0] + g.cp * u[1]) * g.idxy;
#pragma unroll
        for (int l = 0; l < L1; ++l)
#pragma unroll
            for (int m = -l; m <= l; ++m) {
                const int idx = l * l + l + m;
                yd[idx] = uth * yt[idx] + uph * (-(double)m * y[l * l + l - m]);  // dY_lm/dphi = -m Y_{l,-m}
            }
    }
}

// D_u A[p][m] = sum_j (D_u c_j[p] Y_j[m] + c_j[p] D_u Y_j[m]) (c = W R, D_u c = W D_u R)
// and D_u x = D_u fp / sigma: 2 4pi/(2l+1) sum_m A D_u A on the L2 columns, the
// invariants' directional derivatives on the L3 / L4 columns, sum_j D_u c_j
// for the radial basis.  One 64-lane workgroup per atom (persistent), FP64;
// D_u A in 4 x 4 register tiles (eight 16-byte LDS reads and 32 multiply-adds
// per neighbour and tile), summed over the neighbours in list order.
// LDS: Y, D_u Y [ch][MP], c, D_u c [ch][PS] (rows padded with zeros), R, D_u R [ch][K],
// species [ch].
template <int LMAX, bool RADIAL>
__global__ __launch_bounds__(64) void k_gen_jvp(SegmentArgs a, NeighborList nl, int K, int P, int ch,
                                                TangentArgs ta) {
    constexpr int L1 = LMAX + 1, M = L1 * L1, MP = pad4(M);
    extern __shared__ __align__(16) double gen_smem[];
    const int PS = pad4(P);
    double *sY = gen_smem, *sYd = sY + (size_t)ch * MP;
    double *sC = sYd + (size_t)ch * MP, *sCd = sC + (size_t)ch * PS;
    double *sR = sCd + (size_t)ch * PS, *sRd = sR + (size_t)ch * K;
    int *sT = reinterpret_cast<int *>(sRd + (size_t)ch * K);
    const int lane = threadIdx.x;
    const int PM = P * M;
    const int ntile = (PS / 4) * (MP / 4);  // 4 x 4 tiles of D_u A, taken by the lanes round-robin
    for (int row = blockIdx.x; row < a.nrows; row += gridDim.x) {
        const int atom = a.rows_atom[row];
        const int i = a.nl_of ? a.nl_of[atom] : atom;
        const int n = min(nl.cnt[i], nl.cap);
        const size_t base = (size_t)i * nl.cap;
        double *adot = RADIAL ? nullptr : ta.adot + (size_t)row * PM;
        if constexpr (!RADIAL)
            for (int e = lane; e < PM; e += 64) adot[e] = 0.0;
        double racc[2] = {0.0, 0.0};  // radial: D_u fp[p] for p = lane, lane + 64
        for (int c0 = 0; c0 < n; c0 += ch) {
            const int nc = min(ch, n - c0);
            __syncthreads();  // the previous chunk's reads are done
            if (lane < nc) {
                int tj;
                pair_tangent<LMAX, RADIAL>(a, nl, ta, base + c0 + lane, atom, K, sY + lane * MP, sYd + lane * MP,
                                           sR + lane * K, sRd + lane * K, tj);
                sT[lane] = tj;
                if constexpr (!RADIAL)
                    for (int m = M; m < MP; ++m) sY[lane * MP + m] = sYd[lane * MP + m] = 0.0;
            }
            __syncthreads();
            for (int e = lane; e < nc * PS; e += 64) {  // c, D_u c: every lane, the mixing read from L1 / L2
                const int j = e / PS, p = e - j * PS;
                double c = 0.0, cd = 0.0;
                if (p < P) {
                    const double *w = a.W64 + ((size_t)(sT[j] - 1) * P + p) * K;
                    const double *R = sR + j * K, *Rd = sRd + j * K;
                    for (int q = 0; q < K; ++q) {
                        c += w[q] * R[q];
                        cd += w[q] * Rd[q];
                    }
                }
                sC[e] = c;
                sCd[e] = cd;
            }
            __syncthreads();
            if constexpr (RADIAL) {
                for (int j = 0; j < nc; ++j) {
                    if (lane < P) racc[0] += sCd[j * PS + lane];
                    if (lane + 64 < P) racc[1] += sCd[j * PS + lane + 64];
                }
            } else {
                for (int tl = lane; tl < ntile; tl += 64) {
                    const int p0 = (tl / (MP / 4)) * 4, m0 = (tl % (MP / 4)) * 4;
                    double acc[4][4] = {};
                    for (int j = 0; j < nc; ++j) {
                        const double2 c01 = *reinterpret_cast<const double2 *>(sC + j * PS + p0);
                        const double2 c23 = *reinterpret_cast<const double2 *>(sC + j * PS + p0 + 2);
                        const double2 d01 = *reinterpret_cast<const double2 *>(sCd + j * PS + p0);
                        const double2 d23 = *reinterpret_cast<const double2 *>(sCd + j * PS + p0 + 2);
                        const double2 y01 = *reinterpret_cast<const double2 *>(sY + j * MP + m0);
                        const double2 y23 = *reinterpret_cast<const double2 *>(sY + j * MP + m0 + 2);
                        const double2 e01 = *reinterpret_cast<const double2 *>(sYd + j * MP + m0);
                        const double2 e23 = *reinterpret_cast<const double2 *>(sYd + j * MP + m0 + 2);
                        const double c[4] = {c01.x, c01.y, c23.x, c23.y}, cd[4] = {d01.x, d01.y, d23.x, d23.y};
                        const double y[4] = {y01.x, y01.y, y23.x, y23.y}, yd[4] = {e01.x, e01.y, e23.x, e23.y};
#pragma unroll
                        for (int u = 0; u < 4; ++u)
#pragma unroll
                            for (int v = 0; v < 4; ++v) acc[u][v] += cd[u] * y[v] + c[u] * yd[v];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            if (p0 + u < P && m0 + v < M) adot[(p0 + u) * M + m0 + v] += acc[u][v];
                }
            }
        }
        if constexpr (RADIAL) {
            if (lane < P) ta.xdot[(size_t)row * a.xs + lane] = racc[0] / a.sig64[lane];
            if (lane + 64 < P) ta.xdot[(size_t)row * a.xs + lane + 64] = racc[1] / a.sig64[lane + 64];
        } else {
            __syncthreads();  // every lane's D_u A is stored
            const float *A = a.A + (size_t)row * PM;
            const int sl = a.sl, l2n = a.l2n;
            for (int e = lane; e < P * sl; e += 64) {
                const int p = e / sl, c = e - p * sl;
                const float *Ap = A + p * M;
                const double *Adp = adot + p * M;
                double f = 0.0;
                if (c < l2n) {
                    for (int m = c * c; m < (c + 1) * (c + 1); ++m) f += (double)Ap[m] * Adp[m];
                    f *= 2.0 * 4.0 * 3.14159265358979323846 / (double)(2 * c + 1);
                } else {  // the invariant's terms, each differentiated factor by factor
                    const int k = c - l2n;
                    for (int q = a.cpl_off[k]; q < a.cpl_off[k + 1]; ++q) {
                        const int4 ix = a.cpl_idx[q];
                        const double x0 = Ap[ix.x], x1 = Ap[ix.y], x2 = Ap[ix.z], x3 = ix.w >= 0 ? (double)Ap[ix.w] : 1.0;
                        double dv = Adp[ix.x] * x1 * x2 * x3 + x0 * Adp[ix.y] * x2 * x3 + x0 * x1 * Adp[ix.z] * x3;
                        if (ix.w >= 0) dv += x0 * x1 * x2 * Adp[ix.w];
                        f += a.cpl_coef[q] * dv;
                    }
                }
                ta.xdot[(size_t)row * a.xs + e] = f / a.sig64[e];
            }
        }
    }
}

// The tangent on the FP64 matrix cores for the register-tiled shapes (no L3 /
// L4 columns): one wavefront per atom, a lane per neighbour of each chunk of
// 64 forms the pair's Y, D_u Y (registers), R, D_u R and then c = W R, D_u c =
// W D_u R (LDS as C^T[p][j] rows, a zero row P); each 16-row m-tile of Y^T and
// D_u Y^T is staged in LDS, and its k-steps add
//   D_u A^T += Y^T(16 m x 4 neighbours) D_u C(4 x 16 p) + D_u Y^T C
// with v_mfma_f64_16x16x4_f64 (lane maps as in the plain forward: A[i][k] from
// lane i + 16 k, B[k][j] from lane j + 16 k, D[i][j] in lane j + 16 (i % 4),
// element i / 4).  Lanes past the last neighbour take c = D_u c = 0.  The
// epilogue puts D_u A (FP64) in LDS, stores it, and forms D_u x = 2 4pi/(2l+1)
// sum_m A D_u A / sigma.  Rows of 65 doubles: distinct bank pairs per
// quarter-wave.
template <int LMAX, int K, int P, int MIX>
__global__ __launch_bounds__(64) void k_desc_jvp_acc(SegmentArgs a, NeighborList nl, TangentArgs ta) {
    constexpr int L1 = LMAX + 1, M = L1 * L1, S = P * L1;
    constexpr int NMT = (M + 15) / 16, NPT = (P + 15) / 16;
    constexpr int NS = 65;
    constexpr int FC = (P + 1) * NS, FY = 16 * NS;
    constexpr int F = 2 * FC + 2 * FY;
    static_assert(P * M <= F, "D_u A must fit the operand buffer");
    __shared__ double sbuf[F];
    double *sC = sbuf, *sCd = sbuf + FC, *sY = sbuf + 2 * FC, *sYd = sY + FY, *sA = sbuf;
    using d4 = __attribute__((__vector_size__(4 * sizeof(double)))) double;
    constexpr d4 kZero = {0.0, 0.0, 0.0, 0.0};
    const int lane = threadIdx.x;
    const int li = lane & 15, lk = lane >> 4;
    int boff[NPT];
#pragma unroll
    for (int pt = 0; pt < NPT; ++pt) boff[pt] = (pt * 16 + li < P ? pt * 16 + li : P) * NS;
    for (int row = blockIdx.x; row < a.nrows; row += gridDim.x) {
        const int atom = a.rows_atom[row];
        const int i = a.nl_of ? a.nl_of[atom] : atom;
        const int n = min(nl.cnt[i], nl.cap);
        const size_t base = (size_t)i * nl.cap;
        d4 acc[NMT][NPT];
#pragma unroll
        for (int mt = 0; mt < NMT; ++mt)
#pragma unroll
            for (int pt = 0; pt < NPT; ++pt) acc[mt][pt] = kZero;
        for (int c0 = 0; c0 < n; c0 += 64) {
            const int j = c0 + lane;
            const int jr = j < n ? j : n - 1;
            const bool live = j < n;
            const int nks = (min(64, n - c0) + 3) >> 2;
            __syncthreads();  // the previous chunk's (or atom's) LDS reads are done
            double y[M], yd[M], R[K], Rd[K];
            int tj;
            pair_tangent<LMAX, false>(a, nl, ta, base + jr, atom, K, y, yd, R, Rd, tj);
            if constexpr (MIX == MIX_DENSE) {  // c, D_u c row by row into LDS (few registers live)
                const double *w = a.W64 + (size_t)(tj - 1) * P * K;
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    double cp = 0.0, cdp = 0.0;
#pragma unroll
                    for (int q = 0; q < K; ++q) {
                        cp += w[p * K + q] * R[q];
                        cdp += w[p * K + q] * Rd[q];
                    }
                    sC[p * NS + lane] = live ? cp : 0.0;
                    sCd[p * NS + lane] = live ? cdp : 0.0;
                }
            } else {
                double c[P], cd[P];
                mix_rnp<K, P, MIX>(R, tj, a.W64, c);
                mix_rnp<K, P, MIX>(Rd, tj, a.W64, cd);
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    sC[p * NS + lane] = live ? c[p] : 0.0;
                    sCd[p * NS + lane] = live ? cd[p] : 0.0;
                }
            }
            sC[P * NS + lane] = 0.0;  // the zero row (padding p of the last p-tile)
            sCd[P * NS + lane] = 0.0;
#pragma unroll
            for (int mt = 0; mt < NMT; ++mt) {
                if (mt > 0) __syncthreads();  // the previous m-tile's operand reads are done
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int m = mt * 16 + r;
                    sY[r * NS + lane] = m < M ? y[m] : 0.0;
                    sYd[r * NS + lane] = m < M ? yd[m] : 0.0;
                }
                __syncthreads();
                for (int ks = 0; ks < nks; ++ks) {
                    const int col = 4 * ks + lk;
                    const double av = sY[li * NS + col], avd = sYd[li * NS + col];
#pragma unroll
                    for (int pt = 0; pt < NPT; ++pt) {
                        acc[mt][pt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, sCd[boff[pt] + col], acc[mt][pt], 0, 0, 0);
                        acc[mt][pt] = __builtin_amdgcn_mfma_f64_16x16x4f64(avd, sC[boff[pt] + col], acc[mt][pt], 0, 0, 0);
                    }
                }
            }
        }
        __syncthreads();  // operand reads are done: the buffer takes D_u A
#pragma unroll
        for (int mt = 0; mt < NMT; ++mt)
#pragma unroll
            for (int pt = 0; pt < NPT; ++pt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = mt * 16 + 4 * r + lk, p = pt * 16 + li;
                    if (m < M && p < P) sA[p * M + m] = acc[mt][pt][r];
                }
        __syncthreads();
        double *adot = ta.adot + (size_t)row * (P * M);
        for (int e = lane; e < P * M; e += 64) adot[e] = sA[e];
        const float *A = a.A + (size_t)row * (P * M);
        for (int e = lane; e < S; e += 64) {
            const int p = e / L1, l = e - p * L1;
            double f = 0.0;
            for (int m = l * l; m < (l + 1) * (l + 1); ++m) f += (double)A[p * M + m] * sA[p * M + m];
            f *= 2.0 * 4.0 * 3.14159265358979323846 / (double)(2 * l + 1);
            ta.xdot[(size_t)row * a.xs + e] = f / a.sig64[e];
        }
        __syncthreads();  // the epilogue's reads of sA are done before the next atom's operands
    }
}

// dL/dW of a trainable mixing for the force / virial loss, from dL/dfp =
// xbar / sigma and dL/dD_u fp = xdbar / sigma:
//   Abar[p][m] = dL/dA  = J^T fbar + (d(J D_u A)/dA)^T fdbar,  Adbar = dL/dD_u A = J^T fdbar
// (J = dfp/dA: 2 4pi/(2l+1) A on the L2 columns, the invariants' gradients on
// the L3 / L4 columns), then with A = sum_j c_j Y_j^T, D_u A = sum_j (D_u c_j
// Y_j^T + c_j D_u Y_j^T), c = W R, D_u c = W D_u R:
//   dL/dW[t][p][n] += sum_{j of species t} R_j[n] q1_j[p] + D_u R_j[n] q2_j[p],
//   q1 = sum_m Abar[p][m] Y_j[m] + Adbar[p][m] D_u Y_j[m],  q2 = sum_m Adbar[p][m] Y_j[m]
// (radial basis: q1 = fbar[p], q2 = fdbar[p]).  Partial rows per workgroup as
// in k_gen_wgrad.  LDS: Abar, Adbar [P][MP]; per chunk Y, D_u Y [ch][MP],
// R, D_u R [ch][K], q1, q2 [ch][P] (FP64), species [ch].
template <int LMAX, bool RADIAL>
__global__ __launch_bounds__(64) void k_gen_wgrad2(SegmentArgs a, NeighborList nl, int K, int P, int ch,
                                                   TangentArgs ta, double *__restrict__ partial) {
    constexpr int L1 = LMAX + 1, M = L1 * L1, MP = pad4(M);
    extern __shared__ __align__(16) double gen_smem[];
    double *sAb = gen_smem, *sAd = sAb + (size_t)P * MP;
    double *sY = sAd + (size_t)P * MP, *sYd = sY + (size_t)ch * MP;
    double *sR = sYd + (size_t)ch * MP, *sRd = sR + (size_t)ch * K;
    double *sQ1 = sRd + (size_t)ch * K, *sQ2 = sQ1 + (size_t)ch * P;
    int *sT = reinterpret_cast<int *>(sQ2 + (size_t)ch * P);
    const int lane = threadIdx.x;
    const int nt = a.ntypes, PK = P * K, PM = P * M;
    double *part = partial + (size_t)blockIdx.x * nt * PK;
    for (int e = lane; e < nt * PK; e += 64) part[e] = 0.0;
    for (int row = blockIdx.x; row < a.nrows; row += gridDim.x) {
        const int atom = a.rows_atom[row];
        const int i = a.nl_of ? a.nl_of[atom] : atom;
        const int n = min(nl.cnt[i], nl.cap);
        const size_t base = (size_t)i * nl.cap;
        const double *xb = ta.xbar + (size_t)row * a.xs, *xdb = ta.xdbar + (size_t)row * a.xs;
        __syncthreads();  // the previous atom's reads of Abar / Adbar are done
        if constexpr (RADIAL) {
            for (int p = lane; p < P; p += 64) {
                sAb[p * MP] = xb[p] / a.sig64[p];
                sAd[p * MP] = xdb[p] / a.sig64[p];
            }
        } else {
            const float *A = a.A + (size_t)row * PM;
            const double *Ad = ta.adot + (size_t)row * PM;
            const int sl = a.sl, l2n = a.l2n;
            for (int e = lane; e < P * MP; e += 64) {
                const int p = e / MP, m = e - p * MP;
                double ab = 0.0, ad = 0.0;
                if (m < M) {
                    int l = 0;
                    while ((l + 1) * (l + 1) <= m) ++l;
                    const float *Ap = A + p * M;
                    const double *Adp = Ad + p * M;
                    const double *fb = xb + p * sl, *fdb = xdb + p * sl, *sg = a.sig64 + p * sl;
                    if (l < l2n) {
                        const double w4 = 2.0 * 4.0 * 3.14159265358979323846 / (double)(2 * l + 1);
                        const double f1 = fb[l] / sg[l], f2 = fdb[l] / sg[l];
                        ab = w4 * (f1 * (double)Ap[m] + f2 * Adp[m]);
                        ad = w4 * f2 * (double)Ap[m];
                    }
                    if (a.n34) {
                        for (int q = a.cplg_off[m]; q < a.cplg_off[m + 1]; ++q) {
                            const int4 ix = a.cplg_idx[q];
                            const double f1 = fb[l2n + ix.w] / sg[l2n + ix.w], f2 = fdb[l2n + ix.w] / sg[l2n + ix.w];
                            const double x0 = Ap[ix.x], x1 = Ap[ix.y], x2 = ix.z >= 0 ? (double)Ap[ix.z] : 1.0;
                            double dv = Adp[ix.x] * x1 * x2 + x0 * Adp[ix.y] * x2;
                            if (ix.z >= 0) dv += x0 * x1 * Adp[ix.z];
                            const double cf = a.cplg_coef[q];
                            ab += cf * (f1 * x0 * x1 * x2 + f2 * dv);
                            ad += cf * f2 * x0 * x1 * x2;
                        }
                    }
                }
                sAb[e] = ab;
                sAd[e] = ad;
            }
        }
        for (int c0 = 0; c0 < n; c0 += ch) {
            const int nc = min(ch, n - c0);
            __syncthreads();  // Abar / Adbar are written; the previous chunk's reads are done
            if (lane < nc) {
                int tj;
                pair_tangent<LMAX, RADIAL>(a, nl, ta, base + c0 + lane, atom, K, sY + lane * MP, sYd + lane * MP,
                                           sR + lane * K, sRd + lane * K, tj);
                sT[lane] = tj;
            }
            __syncthreads();
            for (int e = lane; e < nc * P; e += 64) {
                const int j = e / P, p = e - j * P;
                double q1, q2;
                if constexpr (RADIAL) {
                    q1 = sAb[p * MP];
                    q2 = sAd[p * MP];
                } else {
                    const double *ab = sAb + p * MP, *ad = sAd + p * MP, *y = sY + j * MP, *yd = sYd + j * MP;
                    q1 = 0.0;
                    q2 = 0.0;
                    for (int m = 0; m < M; ++m) {
                        q1 += ab[m] * y[m] + ad[m] * yd[m];
                        q2 += ad[m] * y[m];
                    }
                }
                sQ1[e] = q1;
                sQ2[e] = q2;
            }
            __syncthreads();
            for (int o = lane; o < PK; o += 64) {
                const int p = o / K, q = o - p * K;
                for (int t = 1; t <= nt; ++t) {
                    double s = 0.0;
                    bool any = false;
                    for (int j = 0; j < nc; ++j)
                        if (sT[j] == t) {
                            s += sR[j * K + q] * sQ1[j * P + p] + sRd[j * K + q] * sQ2[j * P + p];
                            any = true;
                        }
                    if (any) part[(size_t)(t - 1) * PK + o] += s;
                }
            }
        }
    }
}

}  // namespace

// Neighbour chunks are cut so that a workgroup's LDS stays near 24 KB (six
// one-wave workgroups per CU) where the shape allows; up to 160 KB otherwise.
constexpr size_t kTrainLds = 24 * 1024;
static size_t wgrad_lds(int M, int P, int K, int ch) {
    return sizeof(double) * (size_t)ch * (pad4(M) + P + K) + sizeof(float) * (size_t)P * pad4(M) +
           sizeof(int) * (size_t)ch;
}
template <int LMAX, bool RADIAL>
static bool run_wgrad(hipStream_t s, const SegmentArgs &a, const NeighborList &nl, int K, int P, const double *w,
               double *partial, int nwg) {
    constexpr int M = (LMAX + 1) * (LMAX + 1);
    static const bool attr_set = [] {
        JNN_HIP(hipFuncSetAttribute((const void *)k_gen_wgrad<LMAX, RADIAL>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)kGenLds));
        return true;
    }();
    (void)attr_set;
    int ch = 64;
    while (ch > 8 && wgrad_lds(M, P, K, ch) > kTrainLds) ch /= 2;
    if (wgrad_lds(M, P, K, ch) > kGenLds) return false;
    hipLaunchKernelGGL((k_gen_wgrad<LMAX, RADIAL>), dim3(nwg), dim3(64), wgrad_lds(M, P, K, ch), s, a, nl, K, P, ch,
                       w, partial);
    return true;
}

int mix_pgrad_groups(int nrows) { return std::max(1, std::min(nrows, 8 * device_cu_count())); }

bool launch_mix_param_grad(hipStream_t s, int lmax, int K, int P, const SegmentArgs &a, const NeighborList &nl,
                           const double *w, double *partial, double *out) {
    if (K < 1 || K > kGenKMax || P < 1 || P > kGenPMax) return false;
    const int nwg = mix_pgrad_groups(a.nrows);
    bool ok;
    switch (lmax) {
        case kRadialLmax: ok = run_wgrad<0, true>(s, a, nl, K, P, w, partial, nwg); break;
        case 0: ok = run_wgrad<0, false>(s, a, nl, K, P, w, partial, nwg); break;
        case 1: ok = run_wgrad<1, false>(s, a, nl, K, P, w, partial, nwg); break;
        case 2: ok = run_wgrad<2, false>(s, a, nl, K, P, w, partial, nwg); break;
        case 3: ok = run_wgrad<3, false>(s, a, nl, K, P, w, partial, nwg); break;
        case 4: ok = run_wgrad<4, false>(s, a, nl, K, P, w, partial, nwg); break;
        case 5: ok = run_wgrad<5, false>(s, a, nl, K, P, w, partial, nwg); break;
        case 6: ok = run_wgrad<6, false>(s, a, nl, K, P, w, partial, nwg); break;
        case 7: ok = run_wgrad<7, false>(s, a, nl, K, P, w, partial, nwg); break;
        case 8: ok = run_wgrad<8, false>(s, a, nl, K, P, w, partial, nwg); break;
        case 9: ok = run_wgrad<9, false>(s, a, nl, K, P, w, partial, nwg); break;
        case 10: ok = run_wgrad<10, false>(s, a, nl, K, P, w, partial, nwg); break;
        case 11: ok = run_wgrad<11, false>(s, a, nl, K, P, w, partial, nwg); break;
        case 12: ok = run_wgrad<12, false>(s, a, nl, K, P, w, partial, nwg); break;
        default: return false;
    }
    if (!ok) return false;
    const int np = a.ntypes * P * K;
    launch_pgrad_reduce(s, nwg, np, partial, out);
    return true;
}


static size_t jvp_lds(int M, int P, int K, int ch) {
    return sizeof(double) * (size_t)ch * 2 * (pad4(M) + pad4(P) + K) + sizeof(int) * (size_t)ch;
}
static size_t wgrad2_lds(int M, int P, int K, int ch) {
    return sizeof(double) * ((size_t)2 * P * pad4(M) + (size_t)ch * (2 * pad4(M) + 2 * K + 2 * P)) +
           sizeof(int) * (size_t)ch;
}
template <int LMAX, bool RADIAL>
static bool run_jvp(hipStream_t s, const SegmentArgs &a, const NeighborList &nl, int K, int P,
                    const TangentArgs &ta) {
    constexpr int M = (LMAX + 1) * (LMAX + 1);  // the kernel's LDS layout (the radial basis: LMAX 0)
    static const bool attr_set = [] {
        JNN_HIP(hipFuncSetAttribute((const void *)k_gen_jvp<LMAX, RADIAL>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)kGenLds));
        return true;
    }();
    (void)attr_set;
    int ch = 64;
    while (ch > 8 && jvp_lds(M, P, K, ch) > kTrainLds) ch /= 2;
    if (jvp_lds(M, P, K, ch) > kGenLds) return false;
    const int grid = std::max(1, std::min(a.nrows, 8 * device_cu_count()));
    hipLaunchKernelGGL((k_gen_jvp<LMAX, RADIAL>), dim3(grid), dim3(64), jvp_lds(M, P, K, ch), s, a, nl, K, P, ch, ta);
    return true;
}
template <int LMAX, bool RADIAL>
static bool run_wgrad2(hipStream_t s, const SegmentArgs &a, const NeighborList &nl, int K, int P,
                       const TangentArgs &ta, double *partial, int nwg) {
    constexpr int M = (LMAX + 1) * (LMAX + 1);  // the kernel's LDS layout (the radial basis: LMAX 0)
    static const bool attr_set = [] {
        JNN_HIP(hipFuncSetAttribute((const void *)k_gen_wgrad2<LMAX, RADIAL>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)kGenLds));
        return true;
    }();
    (void)attr_set;
    int ch = 64;
    while (ch > 8 && wgrad2_lds(M, P, K, ch) > kTrainLds) ch /= 2;
    if (wgrad2_lds(M, P, K, ch) > kGenLds) return false;
    hipLaunchKernelGGL((k_gen_wgrad2<LMAX, RADIAL>), dim3(nwg), dim3(64), wgrad2_lds(M, P, K, ch), s, a, nl, K, P,
                       ch, ta, partial);
    return true;
}
#define JNN_LMAX_SWITCH(CALL_SPH, CALL_RAD)                            \
    switch (lmax) {                                                    \
        case kRadialLmax: ok = CALL_RAD; break;                        \
        case 0: ok = CALL_SPH(0); break;                               \
        case 1: ok = CALL_SPH(1); break;                               \
        case 2: ok = CALL_SPH(2); break;                               \
        case 3: ok = CALL_SPH(3); break;                               \
        case 4: ok = CALL_SPH(4); break;                               \
        case 5: ok = CALL_SPH(5); break;                               \
        case 6: ok = CALL_SPH(6); break;                               \
        case 7: ok = CALL_SPH(7); break;                               \
        case 8: ok = CALL_SPH(8); break;                               \
        case 9: ok = CALL_SPH(9); break;                               \
        case 10: ok = CALL_SPH(10); break;                             \
        case 11: ok = CALL_SPH(11); break;                             \
        case 12: ok = CALL_SPH(12); break;                             \
        default: return false;                                         \
    }
bool launch_desc_jvp(hipStream_t s, int lmax, int K, int P, int mix, const SegmentArgs &a, const NeighborList &nl,
                     const TangentArgs &ta) {
    if (a.nrows <= 0) return true;
    if (a.n34 == 0 && lmax != kRadialLmax) {  // a register-tiled shape: the matrix-core kernel
        const int grid = std::max(1, std::min(a.nrows, 4 * device_cu_count()));
#define X(L_, K_, P_, M_)                                                                                   \
        if (lmax == L_ && K == K_ && P == P_ && mix == M_) {                                                 \
            hipLaunchKernelGGL((k_desc_jvp_acc<L_, K_, P_, M_>), dim3(grid), dim3(64), 0, s, a, nl, ta);      \
            return true;                                                                                     \
        }
        JNN_SHAPES(X)
#undef X
    }
    if (K < 1 || K > kGenKMax || P < 1 || P > kGenPMax) return false;
    bool ok;
#define SPH(L) run_jvp<L, false>(s, a, nl, K, P, ta)
    JNN_LMAX_SWITCH(SPH, (run_jvp<0, true>(s, a, nl, K, P, ta)))
#undef SPH
    return ok;
}
bool launch_mix_param_grad2(hipStream_t s, int lmax, int K, int P, const SegmentArgs &a, const NeighborList &nl,
                            const TangentArgs &ta, double *partial, double *out) {
    if (K < 1 || K > kGenKMax || P < 1 || P > kGenPMax) return false;
    const int nwg = mix_pgrad_groups(a.nrows);
    bool ok;
#define SPH(L) run_wgrad2<L, false>(s, a, nl, K, P, ta, partial, nwg)
    JNN_LMAX_SWITCH(SPH, (run_wgrad2<0, true>(s, a, nl, K, P, ta, partial, nwg)))
#undef SPH
    if (!ok) return false;
    launch_pgrad_reduce(s, nwg, a.ntypes * P * K, partial, out);
    return true;
}
#undef JNN_LMAX_SWITCH

}  // namespace jnn
