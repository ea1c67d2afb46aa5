// Device side of the GPUMD hook (reference: initGpumdNeiKernel and
// collectGpumdResultsKernel, nnap_interface_cuda.cu:172-298).  GPUMD hands
// over a full neighbour list on the device, column-major [jj * N + i], with
// FP32 displacements r_j - r_i and 0-based types.
#include <hip/hip_runtime.h>

#include "common.h"
#include "kernels.h"

namespace jnn {
namespace {

// model type (1-based) of every particle
__global__ void k_gpumd_types(int n, const int *__restrict__ g_type, int *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = g_type[t] + 1;
}

// Thread per centre i = N1 + r (adjacent threads read adjacent entries of
// GPUMD's column-major list): list row r keeps the pairs with
// dx^2 + dy^2 + dz^2 < rc^2 in FP32 (the reference's test), in GPUMD's order;
// key j*32 + 13 (no periodic shift: GPUMD's displacement already carries it).
// The list is read in batches of 8 entries whose loads are all in flight.
// With tag_rows (every particle a centre) the pair tag carries j's row, j.
constexpr int kRowBatch = 8;
__global__ void k_gpumd_rows(int N, int N1, int N2, const int *__restrict__ num, const int *__restrict__ list,
                             const float *__restrict__ nl_dx, const float *__restrict__ nl_dy,
                             const float *__restrict__ nl_dz, float rc2, int tag_rows, NeighborList nl,
                             int *__restrict__ overflow) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = N1 + r;
    if (i >= N2) return;
    const int jnum = num[i];
    const size_t row = (size_t)r * nl.cap;
    int c = 0;
    for (int j0 = 0; j0 < jnum; j0 += kRowBatch) {
        float dx[kRowBatch], dy[kRowBatch], dz[kRowBatch];
        int jv[kRowBatch];
#pragma unroll
        for (int u = 0; u < kRowBatch; ++u) {
            const bool on = j0 + u < jnum;
            const size_t e = (size_t)(on ? j0 + u : 0) * N + i;
            dx[u] = on ? nl_dx[e] : 1e30f;
            dy[u] = on ? nl_dy[e] : 0.f;
            dz[u] = on ? nl_dz[e] : 0.f;
            jv[u] = on ? list[e] : 0;
        }
#pragma unroll
        for (int u = 0; u < kRowBatch; ++u) {
            if (dx[u] * dx[u] + dy[u] * dy[u] + dz[u] * dz[u] < rc2) {
                const int j = jv[u];
                if (c < nl.cap) {
                    nl.key[row + c] = j * 32 + 13;
                    nl.d[row + c] = make_float4(dx[u], dy[u], dz[u], pair_tag(tag_rows ? j : 0, nl.types[j]));
                }
                ++c;
            }
        }
    }
    nl.cnt[r] = c;
    if (c > nl.cap) atomicMax(overflow, c);
}

// Wavefront per row: the row's entries go to key2 / d2 sorted by key (a
// stable rank: smaller keys, then equal keys earlier in the row; up to 64
// entries the keys are read lane by lane, longer rows from memory).  Equal
// keys (one neighbour listed twice) raise *flag: the force gather's reverse
// search needs unique keys.
__global__ __launch_bounds__(256) void k_gpumd_sort(int nrow, NeighborList nl, int *__restrict__ key2,
                                                    float4 *__restrict__ d2, int *__restrict__ flag) {
    const int lane = threadIdx.x & 63;
    const int r = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (r >= nrow) return;  // wave-uniform
    const int n = min(nl.cnt[r], nl.cap);
    const size_t row = (size_t)r * nl.cap;
    for (int k0 = 0; k0 < n; k0 += 64) {
        const int k = k0 + lane;
        const bool on = k < n;
        const int kt = on ? nl.key[row + k] : 0x7fffffff;
        const float4 dt = on ? nl.d[row + k] : make_float4(0.f, 0.f, 0.f, 0.f);
        int rank = 0;
        bool dup = false;
        if (n <= 64) {
            for (int u = 0; u < n; ++u) {
                const int ku = __builtin_amdgcn_readlane(kt, u);
                rank += (ku < kt || (ku == kt && u < lane)) ? 1 : 0;
                dup = dup || (ku == kt && u != lane);
            }
        } else {
            for (int u = 0; u < n; ++u) {
                const int ku = nl.key[row + u];
                rank += (ku < kt || (ku == kt && u < k)) ? 1 : 0;
                dup = dup || (ku == kt && u != k);
            }
        }
        if (on) {
            key2[row + rank] = kt;
            d2[row + rank] = dt;
            if (dup) atomicAdd(flag, 1);
        }
    }
}

// GPUMD's arrays += the gathered per-atom results: forces [3][N] from F[N][3],
// virials [9][N] from V[N][9], potential from the row energies.
__global__ void k_gpumd_add(int N, const double *__restrict__ F, const double *__restrict__ V,
                            const double *__restrict__ E, double *__restrict__ fx, double *__restrict__ fy,
                            double *__restrict__ fz, double *__restrict__ vir, double *__restrict__ pot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    fx[i] += F[3 * (size_t)i];
    fy[i] += F[3 * (size_t)i + 1];
    fz[i] += F[3 * (size_t)i + 2];
#pragma unroll
    for (int c = 0; c < 9; ++c) vir[(size_t)c * N + i] += V[9 * (size_t)i + c];
    pot[i] += E[i];
}

// Sum over each row of 16 lanes (DPP row rotations); every lane of the row gets it.
__device__ __forceinline__ double row16_sum(double v) {
#pragma unroll
    for (int sh = 8; sh >= 1; sh >>= 1) {
        const long long b = __double_as_longlong(v);
        int lo = (int)(b & 0xffffffffll), hi = (int)(b >> 32);
        switch (sh) {
            case 8: lo = __builtin_amdgcn_update_dpp(0, lo, 0x128, 0xf, 0xf, false);
                    hi = __builtin_amdgcn_update_dpp(0, hi, 0x128, 0xf, 0xf, false); break;
            case 4: lo = __builtin_amdgcn_update_dpp(0, lo, 0x124, 0xf, 0xf, false);
                    hi = __builtin_amdgcn_update_dpp(0, hi, 0x124, 0xf, 0xf, false); break;
            case 2: lo = __builtin_amdgcn_update_dpp(0, lo, 0x122, 0xf, 0xf, false);
                    hi = __builtin_amdgcn_update_dpp(0, hi, 0x122, 0xf, 0xf, false); break;
            default: lo = __builtin_amdgcn_update_dpp(0, lo, 0x121, 0xf, 0xf, false);
                     hi = __builtin_amdgcn_update_dpp(0, hi, 0x121, 0xf, 0xf, false); break;
        }
        v += __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    }
    return v;
}

// Sixteen lanes per centre, four centres per wavefront: lane l takes the
// centre's pairs l, l + 16, ...: F[j] += g and the per-neighbour virial
// V[c][j] += (dx gx, dy gy, dz gz, dx gy, dx gz, dy gz, dy gx, dz gx, dz gy)
// with FP64 atomics (as the reference), then F[i] -= sum g over the row of
// lanes and potential[i] += E_i.
__global__ __launch_bounds__(256) void k_gpumd_scatter(int N, int N1, int nrow, NeighborList nl,
                                                       const double *__restrict__ E, double *__restrict__ fx,
                                                       double *__restrict__ fy, double *__restrict__ fz,
                                                       double *__restrict__ vir, double *__restrict__ pot) {
    const int lid = threadIdx.x & 15;
    const int r = (blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const bool act = r < nrow;  // no early return: every lane takes part in the row sums
    const int n = act ? min(nl.cnt[r], nl.cap) : 0;
    const size_t row = (size_t)r * nl.cap;
    double f0x = 0.0, f0y = 0.0, f0z = 0.0;
    for (int k = lid; k < n; k += 16) {
        const int j = nl.key[row + k] >> 5;
        const float4 g = nl.g[row + k];
        const float4 d = nl.d[row + k];
        const double gx = g.x, gy = g.y, gz = g.z;
        const double dx = d.x, dy = d.y, dz = d.z;
        f0x -= gx;
        f0y -= gy;
        f0z -= gz;
        atomicAdd(fx + j, gx);
        atomicAdd(fy + j, gy);
        atomicAdd(fz + j, gz);
        atomicAdd(vir + 0 * (size_t)N + j, dx * gx);
        atomicAdd(vir + 1 * (size_t)N + j, dy * gy);
        atomicAdd(vir + 2 * (size_t)N + j, dz * gz);
        atomicAdd(vir + 3 * (size_t)N + j, dx * gy);
        atomicAdd(vir + 4 * (size_t)N + j, dx * gz);
        atomicAdd(vir + 5 * (size_t)N + j, dy * gz);
        atomicAdd(vir + 6 * (size_t)N + j, dy * gx);
        atomicAdd(vir + 7 * (size_t)N + j, dz * gx);
        atomicAdd(vir + 8 * (size_t)N + j, dz * gy);
    }
    f0x = row16_sum(f0x);
    f0y = row16_sum(f0y);
    f0z = row16_sum(f0z);
    if (act && lid == 0) {
        const int i = N1 + r;
        atomicAdd(fx + i, f0x);
        atomicAdd(fy + i, f0y);
        atomicAdd(fz + i, f0z);
        pot[i] += E[r];
    }
}

}  // namespace

// Species-grouped row order on the device, stable within a species: rows of
// centres i = N1 + r grouped by model type t = g_type[i] + 1 in [1, nt];
// seg[t] = first row of species t (seg[nt] = nrow).  One wavefront walks the
// centres 64 at a time, ranking each lane among the lanes of its species with
// a ballot.  A type outside [1, nt] sets *bad and the rows are not used.
__global__ __launch_bounds__(64) void k_gpumd_group(int nrow, const int *__restrict__ g_type, int nt,
                                                    int *__restrict__ rows, int *__restrict__ seg,
                                                    int *__restrict__ bad) {
    const int lane = threadIdx.x;
    __shared__ int cnt[kMaxSpecies + 1];
    if (lane <= kMaxSpecies) cnt[lane] = 0;
    __syncthreads();
    for (int r0 = 0; r0 < nrow; r0 += 64) {
        const int r = r0 + lane;
        const int t = r < nrow ? g_type[r] + 1 : 0;
        if (r < nrow && (t < 1 || t > nt)) atomicOr(bad, 1);
        if (t >= 1 && t <= nt) atomicAdd(&cnt[t], 1);
    }
    __syncthreads();
    if (lane == 0) {
        int run = 0;
        for (int t = 1; t <= nt; ++t) {
            const int c = cnt[t];
            cnt[t] = run;
            seg[t - 1] = run;
            run += c;
        }
        seg[nt] = run;
    }
    __syncthreads();
    for (int r0 = 0; r0 < nrow; r0 += 64) {
        const int r = r0 + lane;
        const int t = r < nrow ? g_type[r] + 1 : 0;
        for (int u = 1; u <= nt; ++u) {
            const unsigned long long m = __ballot(t == u);
            if (t == u) rows[cnt[u] + __popcll(m & ((1ull << lane) - 1ull))] = r;
            if (lane == 0) cnt[u] += __popcll(m);
            __syncthreads();
        }
    }
}

void launch_gpumd_group(hipStream_t s, int nrow, const int *g_type, int nt, int *rows, int *seg, int *bad) {
    hipLaunchKernelGGL(k_gpumd_group, dim3(1), dim3(64), 0, s, nrow, g_type, nt, rows, seg, bad);
}
void launch_gpumd_types(hipStream_t s, int n, const int *g_type, int *out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gpumd_types, dim3((n + 255) / 256), dim3(256), 0, s, n, g_type, out);
}
void launch_gpumd_rows(hipStream_t s, int N, int N1, int N2, const int *num, const int *list, const float *nl_dx,
                       const float *nl_dy, const float *nl_dz, float rc2, int tag_rows, const NeighborList &nl,
                       int *overflow) {
    if (N2 <= N1) return;
    hipLaunchKernelGGL(k_gpumd_rows, dim3((N2 - N1 + 255) / 256), dim3(256), 0, s, N, N1, N2, num, list, nl_dx, nl_dy,
                       nl_dz, rc2, tag_rows, nl, overflow);
}
void launch_gpumd_sort(hipStream_t s, int nrow, const NeighborList &nl, int *key2, float4 *d2, int *flag) {
    if (nrow <= 0) return;
    hipLaunchKernelGGL(k_gpumd_sort, dim3((nrow + 3) / 4), dim3(256), 0, s, nrow, nl, key2, d2, flag);  // 4 rows each
}
void launch_gpumd_add(hipStream_t s, int N, const double *F, const double *V, const double *E, double *fx, double *fy,
                      double *fz, double *vir, double *pot) {
    if (N <= 0) return;
    hipLaunchKernelGGL(k_gpumd_add, dim3((N + 255) / 256), dim3(256), 0, s, N, F, V, E, fx, fy, fz, vir, pot);
}
void launch_gpumd_scatter(hipStream_t s, int N, int N1, int nrow, const NeighborList &nl, const double *E, double *fx,
                          double *fy, double *fz, double *vir, double *pot) {
    if (nrow <= 0) return;
    hipLaunchKernelGGL(k_gpumd_scatter, dim3((nrow + 15) / 16), dim3(256), 0, s, N, N1, nrow, nl, E, fx, fy, fz,
                       vir, pot);  // 16 centres per workgroup
}

}  // namespace jnn
