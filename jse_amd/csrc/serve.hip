// Kernels of the served path: replicas resident in a device arena keep a
// Verlet candidate list at rc + skin between requests.  Per request batch:
//   k_srv_check   new positions into the arena, per-replica skin test
//   (rebuild of the flagged replicas' candidate lists: nl.hip kernels)
//   k_srv_update  current continuous positions of every batch atom
//   k_srv_refresh candidate rows -> rows within rc with their displacements
// and, after the evaluation, k_srv_pack gathers the batch's forces.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace jnn {

namespace {

__device__ __forceinline__ double dist2_no_fma(double dx, double dy, double dz) {
#pragma clang fp contract(off)
    const double a = dx * dx;
    const double b = dy * dy;
    const double c = dz * dz;
    return (a + b) + c;
}

// One workgroup per batch replica: copies its new positions into the arena
// and sets need[rid] when it is forced or some atom moved more than half the
// skin since the replica's last candidate-list build.
__global__ __launch_bounds__(256) void k_srv_check(const SrvRep *__restrict__ br, const double *__restrict__ pos,
                                                   double *__restrict__ x, const double *__restrict__ xb,
                                                   double half_skin2, int *__restrict__ need,
                                                   int *__restrict__ need_out) {
    __shared__ double red[256];
    const SrvRep r = br[blockIdx.x];
    double m = 0.0;
    for (int k = threadIdx.x; k < r.natoms; k += 256) {
        const size_t i = (size_t)(r.atom_off + k), t = (size_t)(r.t_off + k);
        const double px = pos[3 * t], py = pos[3 * t + 1], pz = pos[3 * t + 2];
        x[3 * i] = px;
        x[3 * i + 1] = py;
        x[3 * i + 2] = pz;
        if (!r.force) {
            const double dx = px - xb[3 * i], dy = py - xb[3 * i + 1], dz = pz - xb[3 * i + 2];
            m = fmax(m, dist2_no_fma(dx, dy, dz));
        }
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int nd = (r.force || !(red[0] <= half_skin2)) ? 1 : 0;
        need[r.rid] = nd;
        need_out[blockIdx.x] = nd;
    }
}

// One workgroup per batch replica: zero the cell counts of replicas being rebuilt.
__global__ __launch_bounds__(256) void k_srv_zero_cells(const SrvRep *__restrict__ br, const ReplicaGrid *__restrict__ reps,
                                                        const int *__restrict__ need, int *__restrict__ cell_cnt) {
    const SrvRep r = br[blockIdx.x];
    if (!need[r.rid]) return;
    const ReplicaGrid &g = reps[r.rid];
    for (int c = threadIdx.x; c <= g.ncell; c += 256) cell_cnt[g.cell_off + c] = 0;
}

// Continuous positions: a rebuilt replica takes its freshly wrapped positions
// as the new reference; the others move their reference wrap by the
// displacement since the build, so image codes stay valid between builds.
__global__ void k_srv_update(int n, const int *__restrict__ atoms, const int *__restrict__ rep_of,
                             const int *__restrict__ need, const double *__restrict__ x, double *__restrict__ xb,
                             double *__restrict__ xwb, double *__restrict__ xw) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const size_t i = (size_t)atoms[t];
    if (need[rep_of[i]]) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            xb[3 * i + d] = x[3 * i + d];
            xwb[3 * i + d] = xw[3 * i + d];
        }
    } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) xw[3 * i + d] = xwb[3 * i + d] + (x[3 * i + d] - xb[3 * i + d]);
    }
}

// Wave per atom: keeps the candidates within rc, in the candidate row's
// (sorted) order, with the displacement computed exactly as the cell-list
// build computes it: (x_j + shift) - x_i in FP64, no FMA, rounded once.
__global__ __launch_bounds__(256) void k_srv_refresh(int n, const int *__restrict__ atoms, const int *__restrict__ rep_of,
                                                     const ReplicaGrid *__restrict__ reps, double rc2,
                                                     const double *__restrict__ xw, NeighborList vl, NeighborList nl,
                                                     int *__restrict__ overflow) {
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= n) return;
    const int i = atoms[t];
    const ReplicaGrid &g = reps[rep_of[i]];
    const double xi = xw[3 * (size_t)i], yi = xw[3 * (size_t)i + 1], zi = xw[3 * (size_t)i + 2];
    const int cv = min(vl.cnt[i], vl.cap);
    const int *vrow = vl.key + (size_t)i * vl.cap;
    const size_t row = (size_t)i * nl.cap;
    int hits = 0;
    for (int k0 = 0; k0 < cv; k0 += 64) {
        const int k = k0 + lane;
        bool hit = false;
        int key = 0;
        float fdx = 0.f, fdy = 0.f, fdz = 0.f;
        float4 lo = make_float4(0.f, 0.f, 0.f, 0.f);
        if (k < cv) {
            key = vrow[k];
            const int j = key >> 5, code = key & 31;
            const int sx = code % 3 - 1, sy = (code / 3) % 3 - 1, sz = code / 9 - 1;
            double xj = xw[3 * (size_t)j], yj = xw[3 * (size_t)j + 1], zj = xw[3 * (size_t)j + 2];
            if (sx) xj = xj + (sx > 0 ? g.box[0] : -g.box[0]);
            if (sy) yj = yj + (sy > 0 ? g.box[1] : -g.box[1]);
            if (sz) zj = zj + (sz > 0 ? g.box[2] : -g.box[2]);
            const double dx = xj - xi, dy = yj - yi, dz = zj - zi;
            hit = dist2_no_fma(dx, dy, dz) < rc2;
            fdx = (float)dx;
            fdy = (float)dy;
            fdz = (float)dz;
            lo = make_float4((float)(dx - (double)fdx), (float)(dy - (double)fdy), (float)(dz - (double)fdz), 0.f);
        }
        const unsigned long long mask = __ballot(hit);
        const unsigned long long below = lane ? (mask & ((1ull << lane) - 1ull)) : 0ull;
        const int pos = hits + __popcll(below);
        if (hit && pos < nl.cap) {
            nl.key[row + pos] = key;
            nl.d[row + pos] = make_float4(fdx, fdy, fdz, pair_tag(key >> 5, nl.types[key >> 5]));  // row j: atom j
            if (nl.dlo) nl.dlo[row + pos] = lo;
        }
        hits += __popcll(mask);
    }
    if (lane == 0) {
        nl.cnt[i] = hits;
        if (hits > nl.cap) atomicMax(overflow, hits);
    }
}

__global__ void k_srv_pack(int n, const int *__restrict__ atoms, const double *__restrict__ F,
                           double *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const size_t i = (size_t)atoms[t];
    out[3 * (size_t)t] = F[3 * i];
    out[3 * (size_t)t + 1] = F[3 * i + 1];
    out[3 * (size_t)t + 2] = F[3 * i + 2];
}

// Per-replica arena metadata of newly admitted replicas: grid, replica id and type of every slot.
__global__ __launch_bounds__(256) void k_srv_admit(const SrvAdmit *__restrict__ adm, const int *__restrict__ types_in,
                                                   ReplicaGrid *__restrict__ reps, int *__restrict__ rep_of,
                                                   int *__restrict__ types) {
    const SrvAdmit a = adm[blockIdx.x];
    if (threadIdx.x == 0) reps[a.rid] = a.grid;
    for (int k = threadIdx.x; k < a.grid.natoms; k += 256) {
        rep_of[a.grid.atom_off + k] = a.rid;
        types[a.grid.atom_off + k] = types_in[a.types_off + k];
    }
}

}  // namespace

void launch_srv_admit(hipStream_t s, int nadm, const SrvAdmit *adm, const int *types_in, ReplicaGrid *reps,
                      int *rep_of, int *types) {
    if (nadm <= 0) return;
    hipLaunchKernelGGL(k_srv_admit, dim3(nadm), dim3(256), 0, s, adm, types_in, reps, rep_of, types);
}
void launch_srv_check(hipStream_t s, int nrep, const SrvRep *br, const double *pos, double *x, const double *xb,
                      double half_skin2, int *need, int *need_out) {
    if (nrep <= 0) return;
    hipLaunchKernelGGL(k_srv_check, dim3(nrep), dim3(256), 0, s, br, pos, x, xb, half_skin2, need, need_out);
}
void launch_srv_zero_cells(hipStream_t s, int nrep, const SrvRep *br, const ReplicaGrid *reps, const int *need,
                           int *cell_cnt) {
    if (nrep <= 0) return;
    hipLaunchKernelGGL(k_srv_zero_cells, dim3(nrep), dim3(256), 0, s, br, reps, need, cell_cnt);
}
void launch_srv_update(hipStream_t s, int n, const int *atoms, const int *rep_of, const int *need, const double *x,
                       double *xb, double *xwb, double *xw) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_srv_update, dim3((n + 255) / 256), dim3(256), 0, s, n, atoms, rep_of, need, x, xb, xwb, xw);
}
void launch_srv_refresh(hipStream_t s, int n, const int *atoms, const int *rep_of, const ReplicaGrid *reps,
                        double rc2, const double *xw, const NeighborList &vl, const NeighborList &nl, int *overflow) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_srv_refresh, dim3((n + 3) / 4), dim3(256), 0, s, n, atoms, rep_of, reps, rc2, xw, vl, nl,
                       overflow);
}
void launch_srv_pack(hipStream_t s, int n, const int *atoms, const double *F, double *out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_srv_pack, dim3((n + 255) / 256), dim3(256), 0, s, n, atoms, F, out);
}

}  // namespace jnn
