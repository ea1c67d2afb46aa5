// Device side of the GPUMD hook (reference: initGpumdNeiKernel and
// collectGpumdResultsKernel, nnap_interface_cuda.cu:172-298).  GPUMD hands
// over a full neighbour list on the device, column-major [jj * N + i], with
// FP32 displacements r_j - r_i and 0-based types.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace jnn {
namespace {

// model type (1-based) of every particle
__global__ void k_gpumd_types(int n, const int *__restrict__ g_type, int *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = g_type[t] + 1;
}

// Thread per centre i = N1 + r: list row r keeps the pairs with
// dx^2 + dy^2 + dz^2 < rc^2 in FP32 (the reference's test), in GPUMD's order;
// key j*32 + 13 (no periodic shift: GPUMD's displacement already carries it).
__global__ void k_gpumd_rows(int N, int N1, int N2, const int *__restrict__ num, const int *__restrict__ list,
                             const float *__restrict__ nl_dx, const float *__restrict__ nl_dy,
                             const float *__restrict__ nl_dz, float rc2, NeighborList nl, int *__restrict__ overflow) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = N1 + r;
    if (i >= N2) return;
    const int jnum = num[i];
    const size_t row = (size_t)r * nl.cap;
    int c = 0;
    for (int jj = 0; jj < jnum; ++jj) {
        const size_t e = (size_t)jj * N + i;
        const float dx = nl_dx[e], dy = nl_dy[e], dz = nl_dz[e];
        if (dx * dx + dy * dy + dz * dz < rc2) {
            const int j = list[e];
            if (c < nl.cap) {
                nl.key[row + c] = j * 32 + 13;
                nl.d[row + c] = make_float4(dx, dy, dz, pair_tag(0, nl.types[j]));  // species only
            }
            ++c;
        }
    }
    nl.cnt[r] = c;
    if (c > nl.cap) atomicMax(overflow, c);
}

// Thread per centre: F[i] -= g, F[j] += g, the per-neighbour virial
// V[c][j] += (dx gx, dy gy, dz gz, dx gy, dx gz, dy gz, dy gx, dz gx, dz gy)
// (FP64 atomics, as the reference) and potential[i] += E_i.
__global__ void k_gpumd_scatter(int N, int N1, int nrow, NeighborList nl, const double *__restrict__ E,
                                double *__restrict__ fx, double *__restrict__ fy, double *__restrict__ fz,
                                double *__restrict__ vir, double *__restrict__ pot) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    const int i = N1 + r;
    const int n = min(nl.cnt[r], nl.cap);
    const size_t row = (size_t)r * nl.cap;
    double f0x = 0.0, f0y = 0.0, f0z = 0.0;
    for (int k = 0; k < n; ++k) {
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
    atomicAdd(fx + i, f0x);
    atomicAdd(fy + i, f0y);
    atomicAdd(fz + i, f0z);
    pot[i] += E[r];
}

}  // namespace

void launch_gpumd_types(hipStream_t s, int n, const int *g_type, int *out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gpumd_types, dim3((n + 255) / 256), dim3(256), 0, s, n, g_type, out);
}
void launch_gpumd_rows(hipStream_t s, int N, int N1, int N2, const int *num, const int *list, const float *nl_dx,
                       const float *nl_dy, const float *nl_dz, float rc2, const NeighborList &nl, int *overflow) {
    if (N2 <= N1) return;
    hipLaunchKernelGGL(k_gpumd_rows, dim3((N2 - N1 + 255) / 256), dim3(256), 0, s, N, N1, N2, num, list, nl_dx, nl_dy,
                       nl_dz, rc2, nl, overflow);
}
void launch_gpumd_scatter(hipStream_t s, int N, int N1, int nrow, const NeighborList &nl, const double *E, double *fx,
                          double *fy, double *fz, double *vir, double *pot) {
    if (nrow <= 0) return;
    hipLaunchKernelGGL(k_gpumd_scatter, dim3((nrow + 255) / 256), dim3(256), 0, s, N, N1, nrow, nl, E, fx, fy, fz,
                       vir, pot);
}

}  // namespace jnn
