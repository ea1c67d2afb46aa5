// Host-callable launchers of the HIP kernels in libjsenn_hip.  Every launcher
// enqueues on the given stream and never synchronises.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace jnn {

// Compute units of the current device (cached after the first query).
int device_cu_count();

// MI355X deals consecutive workgroups round-robin to its 8 XCDs, each with
// its own L2.  With a grid of 8 * ceil(nblocks / 8) workgroups, this maps
// workgroup b to logical block (b % 8) * ceil(nblocks / 8) + b / 8, so each
// XCD takes one contiguous range of logical blocks (>= nblocks: no work).
constexpr int kXcds = 8;
__host__ __device__ __forceinline__ int xcd_grid(int nblocks) { return ((nblocks + kXcds - 1) / kXcds) * kXcds; }
__device__ __forceinline__ int xcd_block(int b, int grid) {
    return (b % kXcds) * (grid / kXcds) + b / kXcds;
}

// Launch shapes of one system's kernels, chosen by the autotuner
// (jse_nnap_system_autotune); 0 keeps a kernel's built-in default.
struct LaunchTune {
    int desc_fwd_wg_per_cu = 0;  // descriptor forward: persistent workgroups per CU (0: one per atom)
    int desc_bwd_wg_per_cu = 0;  // descriptor backward: likewise
    int mlp_wg_per_cu = 0;       // MLP: persistent workgroups per CU (0: 2)
    int nl_wg_per_cu = 0;        // neighbour build: persistent workgroups per CU (0: one per atom)
    int gather_waves = 0;        // force gather: atoms (waves) per workgroup (0: 4)
    int desc_fwd_variant = 0;    // descriptor forward: reserved (one compiled schedule per precision mode)
    int desc_bwd_variant = 0;    // descriptor backward: scheduling barrier every 1 / 2 / no rows
};
constexpr int kTuneKnobs = 7;

// Periodic cell grid of one replica's orthogonal box (reference:
// NeighborListGetter.java:952-1001).  A system holds one or more replicas;
// their atoms and cells are numbered consecutively.
struct ReplicaGrid {
    double box[3];
    double cb[3];   // cell edge per axis = L_d / size_d
    int size[3];
    int ncell;
    int cell_off;   // first global cell of the replica
    int atom_off;   // first atom of the replica
    int natoms;
    int pad_;
};

// Neighbour-list storage of one system (device pointers), rows of `cap` slots per atom,
// each row sorted by key = j*32 + shift_code.
struct NeighborList {
    int cap = 0;
    int *cnt = nullptr;     // [N] true neighbour count (may exceed cap -> overflow)
    int *key = nullptr;     // [N*cap]
    float4 *d = nullptr;    // [N*cap] (dx, dy, dz, tag): FP32 of the FP64 displacement r_j + shift - r_i and
                            // the pair tag of j (pair_tag: its model species and its own row), so the
                            // descriptor kernels skip key -> type and the force gather key -> row
    const int *types = nullptr;  // [atoms] model species (1-based) of the atoms the keys name; read by
                                 // the builders that write d
    int *rev = nullptr;     // [N*cap] slot of the reverse pair in j's row, -1 if absent (statistics only)
    float4 *g = nullptr;    // [N*cap] (gx, gy, gz, 0): g_j = -dE_i/d(r_j - r_i)
    float4 *dlo = nullptr;  // [N*cap] (dx, dy, dz, 0) of the FP64 displacement minus d's FP32 rounding, so
                            // d + dlo holds the FP64 displacement to 48 bits (nullptr: FP32 only, the
                            // accurate kernels then take lo = 0)
};
// The FP64 displacement of list slot k (d + dlo, or d alone without dlo).
__device__ __forceinline__ float4 pair_lo(const NeighborList &nl, size_t k) {
    return nl.dlo ? nl.dlo[k] : make_float4(0.f, 0.f, 0.f, 0.f);
}

// Pair tag, carried as the int bits of d.w: bits 0-3 the model species of j
// (1-based), bits 4-30 the row of the list that holds j's own neighbours
// (0 where j has no row, e.g. a ghost atom).  Rows are limited to kTagRows.
constexpr int kTagRows = 1 << 27;
__device__ __forceinline__ float pair_tag(int jrow, int tj) { return __int_as_float((jrow << 4) | tj); }
__device__ __forceinline__ int tag_species(float w) { return __float_as_int(w) & 15; }
__device__ __forceinline__ int tag_row(float w) { return __float_as_int(w) >> 4; }

// Orders one wavefront's LDS writes before its own later LDS reads of other
// lanes' data (no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Slot of the reverse pair (i, -shift) of atom i's pair `key` = j*32 + code in
// j's sorted row, -1 if j's row does not hold it.  Keys use atom numbers; row
// j of the list is row_of[j] (nullptr: j).  *jrow receives that row.
__device__ __forceinline__ int find_reverse(const NeighborList &nl, int i, int key, const int *row_of, int *jrow) {
    const int j = key >> 5, code = key & 31;
    const int want = i * 32 + (26 - code);
    const int rj = row_of ? row_of[j] : j;
    *jrow = rj;
    const int cj = min(nl.cnt[rj], nl.cap);
    const int *kj = nl.key + (size_t)rj * nl.cap;
    int lo = 0, hi = cj - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int v = kj[mid];
        if (v == want) return mid;
        if (v < want) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

// ---- neighbour build ----
void launch_fill_int(hipStream_t s, int n, int *p, int v);  // kernel fill (graph-capture safe)
// Copies of a box's atoms for its supercell of cp[0] x cp[1] x cp[2] boxes:
// out row c * nuser + i = xyz[i] + (cx, cy, cz) * box, c = cx + cp[0] (cy + cp[1] cz).
void launch_expand_copies(hipStream_t s, int nuser, const int *cp, const double *box, const double *xyz,
                          double *out);
// `atoms` lists the atoms to process (nullptr = atoms 0..n-1); `rep_list`
// the replicas whose cells are scanned (nullptr = replicas 0..nrep-1);
// `need` (nullptr = all) skips the atoms and cells of replicas whose flag is 0.
// Each replica owns ncell + 1 entries of cell_start from its cell_off.
// With nl.d == nullptr the build stores keys only (a Verlet candidate list).
void launch_wrap_bin(hipStream_t s, int n, const int *atoms, const double *xyz_in, const int *rep_of,
                     const ReplicaGrid *reps, const int *need, double *xw, int *cell, int *cell_cnt, int *rank);
// With `clear` the scan zeroes the counts it consumed (the next binning starts
// from zero without a fill) and *clear_flag when non-null.
void launch_cell_scan(hipStream_t s, int nrep, const int *rep_list, const ReplicaGrid *reps, const int *need,
                      int *cell_cnt, int *cell_start, bool clear = false, int *clear_flag = nullptr);
// Also writes slot_of[atom] when non-null.
void launch_cell_scatter(hipStream_t s, int n, const int *atoms, const int *rep_of, const int *need, const int *cell,
                         const int *rank, const int *cell_start, const double *xw, int *sidx, double *sxw,
                         int *slot_of = nullptr);
void launch_nl_build(hipStream_t s, int n, const int *atoms, const int *rep_of, const ReplicaGrid *reps,
                     const int *need, double rc2, const double *xw, const int *cell, const int *cell_start,
                     const int *sidx, const double *sxw, NeighborList nl, int *overflow, int wg_per_cu = 0);
// The rows of every atom of a system, placed in cell-sorted slot order (row s
// is atom sidx[s]; keys keep atom numbers j*32 + code, so each row's content
// and order do not depend on the slot order), one workgroup per cell (27-cell
// candidates staged in LDS once per cell); cell_rep[c] = replica of global
// cell c, -1 for a replica's end entry.
void launch_nl_build_cells(hipStream_t s, int ncell_total, const int *cell_rep, const ReplicaGrid *reps, double rc2,
                           const int *cell_start, const int *sidx, const double *sxw, NeighborList nl, int *overflow,
                           int wg_per_cu = 0, bool large_cells = false);
void launch_nl_reverse(hipStream_t s, int n, const int *atoms, NeighborList nl, const int *out_of = nullptr,
                       const int *row_of = nullptr);

// ---- descriptor / MLP / backward for one species segment ----
struct SegmentArgs {
    int nrows;              // atoms of this species
    const int *rows_atom;   // atom index of each row (species-sorted order)
    const int *nl_of;       // neighbour-list row of an atom (nullptr: the atom index)
    const int *types;       // [N] 1-based types by atom
    float rc;
    double rc64;            // the same cut-off in FP64 (the plain mode's kernels)
    int ntypes;             // species of the model (neighbour types are 1..ntypes)
    const float *W;         // [ntypes][P][K] FP32 mixing (used by the dense kernel)
    const float *mu, *sig;  // normalisation [S]
    const double *mu64, *sig64;  // the same in FP64 (descriptor normalisation)
    float *x;               // [nrows][S] normalised descriptors (forward out, MLP in)
    double *x64;            // the same in FP64 (the plain mode writes these instead)
    float *A;               // [nrows][P*M]
    float *G;               // [nrows][S]  2*(4pi/(2l+1)) * dE/dfp
    int wg_per_cu;          // persistent workgroups per CU (0: one workgroup per row)
    int variant;            // compiled schedule variant (autotuned; 0: default)
    int fp32_acc;           // 1: the matrix-core forward accumulates in FP32 (FP32 geometry modes)
    int acc;                // 1: pair geometry, harmonics, radial basis and A in FP64 (the plain mode)
    const double *W64;      // [ntypes][P][K] FP64 mixing (the plain mode's forward)
    int self_t;             // this segment's species (1-based)
    int mirror;             // mirror basis: neighbours of species `mirror` and of self_t swap roles (0: none)
    int xs;                 // row stride of x and G: the species' descriptor size (a merge basis' part
                            // writes its slice of the species' row; x, G, mu and sigma then point at it)
    int accum_g;            // backward: add into the pairs' g (the parts of a merge basis after the first)
    // descriptor columns: per radial channel l2n L2 values (lmax + 1), then n34 L3 / L4
    // invariants, sl = l2n + n34 columns in all; A and the harmonics run to
    // max(lmax, l3max, l4max).  Invariant k of a channel's A row is the sum over
    // terms off[k] .. off[k+1]-1 of coef A[i.x] A[i.y] A[i.z] (times A[i.w] when
    // i.w >= 0); dE/dA[m] gains, over terms goff[m] .. goff[m+1]-1, gcoef G[k]
    // A[i.x] A[i.y] (times A[i.z] when i.z >= 0) with k = i.w.  Only the untiled
    // kernels evaluate n34 > 0.
    int l2n, sl, n34;
    const int4 *cpl_idx, *cplg_idx;
    const double *cpl_coef, *cplg_coef;
    const int *cpl_off, *cplg_off;
};
// g of a pair: stored, or added to what an earlier part of a merge basis stored
__device__ __forceinline__ void put_g(const NeighborList &nl, size_t k, float gx, float gy, float gz, int accum) {
    if (accum) {
        const float4 o = nl.g[k];
        gx += o.x;
        gy += o.y;
        gz += o.z;
    }
    nl.g[k] = make_float4(gx, gy, gz, 0.f);
}
// Stores a normalised descriptor value: FP64 rows in the plain mode, else FP32.
__device__ __forceinline__ void store_x(const SegmentArgs &a, size_t idx, double v) {
    if (a.x64) a.x64[idx] = v;
    else a.x[idx] = (float)v;
}
// A neighbour's species as the centre's basis sees it (MirrorBasis: the mirrored
// species and the centre's own swap roles; nnap_main.hpp:40-47)
__device__ __forceinline__ int role_species(int tj, int self_t, int mirror) {
    return mirror == 0 ? tj : (tj == mirror ? self_t : (tj == self_t ? mirror : tj));
}
// returns false when no kernel is compiled for this (lmax, K, P, mix); lmax =
// kRadialLmax selects the radial-only chebyshev basis (fp[p] = sum_j fc Rnp[p],
// rows of S = P and no A)
constexpr int kRadialLmax = -1;
bool launch_desc_fwd(hipStream_t s, int lmax, int K, int P, int mix, const SegmentArgs &a, const NeighborList &nl);
bool launch_desc_bwd(hipStream_t s, int lmax, int K, int P, int mix, const SegmentArgs &a, const NeighborList &nl);
// Several species segments of one descriptor shape in one launch (no tail per
// species): rows row0[k] .. row0[k+1]-1 of the launch are row (r - row0[k]) of
// segment k.  Launch shape, variant and precision come from a[0].
constexpr int kSegMax = 8;
struct SegmentSet {
    SegmentArgs a[kSegMax];
    int row0[kSegMax + 1];
    int nseg;
};
bool launch_desc_fwd_set(hipStream_t s, int lmax, int K, int P, int mix, const SegmentSet &set,
                         const NeighborList &nl);
bool launch_desc_bwd_set(hipStream_t s, int lmax, int K, int P, int mix, const SegmentSet &set,
                         const NeighborList &nl);
// The untiled kernels for the shapes above without a tiled kernel: lmax in
// [0, 12] or kRadialLmax, K <= 21, P <= 128, any mixing (through the dense W);
// false when the shape is outside these limits or its A does not fit in LDS.
bool launch_desc_gen_set(hipStream_t s, int lmax, int K, int P, const SegmentSet &set, const NeighborList &nl,
                         bool backward);
// true when (lmax, K, P, mix) has a tiled kernel (lmax = kRadialLmax: radial)
bool desc_tiled(int lmax, int K, int P, int mix);
// Why the untiled kernels cannot run a shape (its LDS needs exceed 160 KB;
// `acc`: the plain mode's FP64 staging), or "" when they can.
std::string desc_gen_limit(int lmax, int K, int P, int ntypes, bool acc);

struct MlpArgs {
    int nrows, S, L1, nlayers;
    int hidden[8];
    const int *rows_atom;
    const float *x;          // [nrows][S]
    const float *w_in_out;   // per layer [in][out] (forward layout) then output W
    const float *w_out_in;   // per layer [out][in] (backward layout)
    const float *bias;       // hidden biases, then output bias
    const float *sig;        // [S]
    double sig_e, mu_e;
    // the FP64 MLP of the plain mode reads these instead of x, the weights and sigma
    const double *x64;       // [nrows][S]
    const double *w_in_out64, *w_out_in64, *bias64, *sig64;
    double *E;               // [N] per-atom energy (by atom index)
    float *G;                // [nrows][S]
    int want_grad;
    int wg_per_cu;           // persistent grid size (0: 2 per CU)
    int radial;              // radial-only basis: G = dE/dfp / sigma (no L2 power-spectrum factor)
    const double *l2f;       // [S] per-column factor of G (a merge basis: 2 4pi/(2l+1) per spherical
                             // part's column, 1 per radial part's); nullptr: from L1 and radial
};
// Any layer count and widths; false when a 32-row tile of the input and the
// hidden activations (and their SiLU derivatives) exceeds 160 KB of LDS:
// S + 2 * sum(hidden) > 1280 (FP32) or 640 (FP64).
bool launch_mlp_fp32(hipStream_t s, const MlpArgs &a);
constexpr const char *kMlpFp32Limit = "the FP32 MLP needs input_dim + 2 * sum(hidden_dims) <= 1280";
// MFMA kernels for 2 hidden layers of supported widths; false when the shape has no compiled kernel
bool launch_mlp_mfma_fp32(hipStream_t s, const MlpArgs &a);
bool launch_mlp_mfma_bf16(hipStream_t s, const MlpArgs &a);
bool launch_mlp_mfma_fp64(hipStream_t s, const MlpArgs &a);
// FP64 vector MLP of the plain mode: any depth with S + 2 * sum(hidden) <= 640
bool launch_mlp_fp64(hipStream_t s, const MlpArgs &a);
constexpr const char *kMlpFp64Limit = "the FP64 MLP needs input_dim + 2 * sum(hidden_dims) <= 640";
enum { kMlpFp32 = 0, kMlpBf16 = 1, kMlpFp64 = 2 };  // element types of the MLP kernels
// Training derivative: out[p] = sum_rows w[atom] dE/dtheta_p over the packed
// network parameters (FP64 weights; x64 or x as input); partial holds
// mlp_pgrad_groups(a) rows of mlp_param_count(a) doubles.  False when a
// 16-row tile does not fit in LDS.
int mlp_param_count(const MlpArgs &a);
int mlp_pgrad_groups(const MlpArgs &a);
bool launch_mlp_param_grad(hipStream_t s, const MlpArgs &a, const double *w, double *partial, double *out);
// out[e] = sum over the nwg rows of partial[g][e], in row order (np entries)
void launch_pgrad_reduce(hipStream_t s, int nwg, int np, const double *partial, double *out);
// Training derivative of a segment's trainable mixing: out[ntypes][P][K] = sum
// over its atoms of w[atom] dE/dW (a's G from an MLP pass with gradients; lmax
// = kRadialLmax: the radial basis).  partial holds mix_pgrad_groups(nrows)
// rows of ntypes P K doubles.  False outside the untiled kernels' limits.
int mix_pgrad_groups(int nrows);
bool launch_mix_param_grad(hipStream_t s, int lmax, int K, int P, const SegmentArgs &a, const NeighborList &nl,
                           const double *w, double *partial, double *out);

// ---- training derivative of the force / virial loss ----
// The loss is taken linear in the outputs: L = sum_i w_i E_i + sum_k bf_k . F_k
// + sum_ab bv_ab V_ab (w, bf, bv: its gradients by the energies, forces and
// total virial).  With F_k = sum over pairs (i,j) of (delta_kj - delta_ki) g_ij
// and V_ab = sum over pairs of d_a g_b, the pairs' g = -dE_i/dd_ij enter as
// sum_ij u_ij . g_ij = -D_u E_i, u_ij = bf_j - bf_i + sum_a bv_ab d_a: a
// directional derivative of E_i with every displacement moved along u_ij.
// The tangents and adjoints below are per segment row (a part's slices, as
// part_segment narrows x and A).
struct TangentArgs {
    const double *bf;       // [N][3] dL/dF by atom (nullptr: none)
    double bv[9];           // dL/dV, components xx yy zz xy xz yz yx zx zy
    const double *bv_rep;   // [nrep][9] dL/dV of each replica of a batch (nullptr: bv for every pair)
    const int *rep_of;      // [N] replica of each atom (with bv_rep)
    double *xdot;           // [nrows][xs] D_u x (the x slice of this part)
    double *adot;           // [nrows][P M] D_u A of this part
    const double *xbar;     // [nrows][xs] dL/dx
    const double *xdbar;    // [nrows][xs] dL/d(D_u x)
};
// D_u A and D_u x of a segment (lmax = kRadialLmax: the radial basis, no A).
// False outside the untiled kernels' limits.
bool launch_desc_jvp(hipStream_t s, int lmax, int K, int P, int mix, const SegmentArgs &a, const NeighborList &nl,
                     const TangentArgs &ta);
// out[ntypes][P][K] = dL/dW of a trainable mixing from dL/dx and dL/d(D_u x)
// (tangent args), summed over the segment's atoms; partial as
// launch_mix_param_grad.
bool launch_mix_param_grad2(hipStream_t s, int lmax, int K, int P, const SegmentArgs &a, const NeighborList &nl,
                            const TangentArgs &ta, double *partial, double *out);
// The networks' parameter gradient of sum_i (w_i E_i - D_u E_i) from x and
// xdot = D_u x, with dL/dx and dL/dxdot of every row into xbar / xdbar
// (nullptr: not written); partial holds mlp_dual_groups rows.  False when an
// 8-row tile of the layers does not fit in LDS.
int mlp_dual_groups(const MlpArgs &a);
bool launch_mlp_dual_grad(hipStream_t s, const MlpArgs &a, const double *w, const double *xdot, double *xbar,
                          double *xdbar, double *partial, double *out);
// All species in one launch: the persistent workgroups are split among the
// species in proportion to their row tiles (each stages its species' weights
// once).  False when the species' MLP shapes differ, there are more than
// kMlpMaxSpecies of them or the shape has no compiled kernel.
constexpr int kMlpMaxSpecies = 8;
struct MlpMulti {
    MlpArgs a[kMlpMaxSpecies];
    int wg0[kMlpMaxSpecies + 1];  // first workgroup of each species, then the total
    int nsp;
};
bool launch_mlp_mfma_multi(hipStream_t s, const MlpArgs *a, int nsp, int mode);

// ---- forces / virials / totals ----
// Row t of the list belongs to atom out_of[t] (nullptr: t); atom j's row is
// the row in the pair tag of d.w.  F and V are written by atom.  The rows'
// capacity must be a multiple of 4 (keys are read 16 bytes at a time).
// `unmatched` (optional) counts pairs whose reverse pair j's row lacks.
void launch_gather(hipStream_t s, int n, const int *atoms, const NeighborList &nl, double *F, double *V,
                   int want_virial, int waves = 0, const int *out_of = nullptr, int *unmatched = nullptr);
// Largest chunk the totals kernel sums in one workgroup.
constexpr int kTotalsChunkMax = 1024;
// Per-replica totals [nrep][10] = (E, V[9]): fixed-size chunks of each
// replica's atoms are summed (Neumaier, then a fixed tree) and the chunk sums
// of a replica are added in order, so a replica's totals do not depend on
// which other replicas share the system.  chunks[c] = (replica, first, end).
// One launch: the last chunk of each replica to finish adds the chunk sums.
// arrivals[nrep] must be zero before the first call (each call leaves it
// zero); out rows of replicas without atoms are not written.
void launch_replica_totals(hipStream_t s, int nchunk, const int4 *chunks, int nrep, const int *rep_chunk_off,
                           const double *E, const double *V, int want_virial, double *partial, double *out,
                           int *arrivals);

// ---- served path (serve.hip) ----
// One replica of a request batch: arena slots [atom_off, atom_off + natoms),
// positions at [t_off, t_off + natoms) of the batch's position buffer.
struct SrvRep {
    int rid, atom_off, natoms, t_off, force, pad_[3];
};
// A replica admitted into the arena: its grid (with cell_off / atom_off /
// natoms set) and where its types sit in the batch's type buffer.
struct SrvAdmit {
    ReplicaGrid grid;
    int rid, types_off, pad_[2];
};
void launch_srv_admit(hipStream_t s, int nadm, const SrvAdmit *adm, const int *types_in, ReplicaGrid *reps,
                      int *rep_of, int *types);
void launch_srv_check(hipStream_t s, int nrep, const SrvRep *br, const double *pos, double *x, const double *xb,
                      double half_skin2, int *need, int *need_out);
void launch_srv_zero_cells(hipStream_t s, int nrep, const SrvRep *br, const ReplicaGrid *reps, const int *need,
                           int *cell_cnt);
void launch_srv_update(hipStream_t s, int n, const int *atoms, const int *rep_of, const int *need, const double *x,
                       double *xb, double *xwb, double *xw);
void launch_srv_refresh(hipStream_t s, int n, const int *atoms, const int *rep_of, const ReplicaGrid *reps,
                        double rc2, const double *xw, const NeighborList &vl, const NeighborList &nl, int *overflow);
void launch_srv_pack(hipStream_t s, int n, const int *atoms, const double *F, double *out);

}  // namespace jnn
