// Kernels of the LAMMPS pair hook (the GPU triple of the reference,
// nnap_interface_cuda.cu:12-169, 319-483): neighbour rows built from LAMMPS'
// full lists (local atoms and explicit ghosts, no periodic images), and the
// scatter of pair gradients to forces and virials of local and ghost atoms.
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

// Thread per listed atom ii (i = ilist[ii]): the entries of LAMMPS' list
// (firstneigh transposed as [jj * inum + ii], the layout the reference copies
// to the device) within rc, in list order.  The displacement x_j - x_i is
// formed in FP64 from the double positions and rounded once (the reference
// forms it from FP32 positions).  Keys are j*32 + 13 (no image shift).
__global__ void k_lmp_rows(int inum, const int *__restrict__ ilist, const int *__restrict__ numneigh,
                           const int *__restrict__ firstneigh, const double *__restrict__ x, double rc2,
                           NeighborList nl, int *__restrict__ overflow) {
    const int ii = blockIdx.x * blockDim.x + threadIdx.x;
    if (ii >= inum) return;
    const int i = ilist[ii];
    const double xi = x[3 * (size_t)i], yi = x[3 * (size_t)i + 1], zi = x[3 * (size_t)i + 2];
    const int jnum = numneigh[ii];
    const size_t row = (size_t)ii * nl.cap;
    int c = 0;
    for (int jj = 0; jj < jnum; ++jj) {
        const int j = firstneigh[(size_t)jj * inum + ii];
        const double dx = x[3 * (size_t)j] - xi, dy = x[3 * (size_t)j + 1] - yi, dz = x[3 * (size_t)j + 2] - zi;
        if (dist2_no_fma(dx, dy, dz) < rc2) {
            if (c < nl.cap) {
                nl.key[row + c] = j * 32 + 13;
                nl.d[row + c] = make_float4((float)dx, (float)dy, (float)dz, pair_tag(0, nl.types[j]));  // species only
            }
            ++c;
        }
    }
    nl.cnt[ii] = c;
    if (c > nl.cap) atomicMax(overflow, c);
}

// model type of every local and ghost atom through LAMMPS' type map
__global__ void k_lmp_types(int n, const int *__restrict__ type, const int *__restrict__ map, int *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = map[type[t]];
}

// Thread per listed atom: F[i] -= g, F[j] += g (FP64 atomics: j may be a
// ghost), the per-centre virial vatom0[ii] and the per-neighbour virial
// vatom1[j] (6 or 9 components), as collectLammpsResultsKernel accumulates
// them (nnap_interface_cuda.cu:114-169), in FP64.  Layouts are SoA:
// F[3][nall], V0[6][inum], V1[9][nall].
__global__ void k_lmp_scatter(int inum, int nall, const int *__restrict__ ilist, NeighborList nl, int vflag,
                              int vflag_atom, int cvflag_atom, double *__restrict__ F, double *__restrict__ V0,
                              double *__restrict__ V1) {
    const int ii = blockIdx.x * blockDim.x + threadIdx.x;
    if (ii >= inum) return;
    const int i = ilist[ii];
    const int n = min(nl.cnt[ii], nl.cap);
    const size_t row = (size_t)ii * nl.cap;
    double fx0 = 0.0, fy0 = 0.0, fz0 = 0.0;
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < n; ++k) {
        const int j = nl.key[row + k] >> 5;
        const float4 g = nl.g[row + k];
        const double fx = g.x, fy = g.y, fz = g.z;
        fx0 -= fx;
        fy0 -= fy;
        fz0 -= fz;
        atomicAdd(F + j, fx);
        atomicAdd(F + nall + j, fy);
        atomicAdd(F + 2 * (size_t)nall + j, fz);
        if (vflag || vflag_atom || cvflag_atom) {
            const float4 d = nl.d[row + k];
            const double dx = d.x, dy = d.y, dz = d.z;
            const double c6[6] = {dx * fx, dy * fy, dz * fz, dx * fy, dx * fz, dy * fz};
#pragma unroll
            for (int q = 0; q < 6; ++q) v[q] += c6[q];
            if (cvflag_atom) {
#pragma unroll
                for (int q = 0; q < 6; ++q) atomicAdd(V1 + (size_t)q * nall + j, c6[q]);
                atomicAdd(V1 + 6 * (size_t)nall + j, dy * fx);
                atomicAdd(V1 + 7 * (size_t)nall + j, dz * fx);
                atomicAdd(V1 + 8 * (size_t)nall + j, dz * fy);
            } else if (vflag_atom) {
#pragma unroll
                for (int q = 0; q < 6; ++q) atomicAdd(V1 + (size_t)q * nall + j, c6[q]);
            }
        }
    }
    atomicAdd(F + i, fx0);
    atomicAdd(F + nall + i, fy0);
    atomicAdd(F + 2 * (size_t)nall + i, fz0);
    if (vflag) {
#pragma unroll
        for (int q = 0; q < 6; ++q) V0[(size_t)q * inum + ii] = v[q];
    }
}

__global__ void k_zero_double(size_t n, double *__restrict__ p) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0.0;
}

// FP64 results -> the FP32 buffers the reference's caller owns
__global__ void k_lmp_out(int inum, int nall, int eflag, int vflag, int nv1, const double *__restrict__ F,
                          const double *__restrict__ E, const double *__restrict__ V0, const double *__restrict__ V1,
                          float *__restrict__ f, float *__restrict__ eatom0, float *__restrict__ vatom0,
                          float *__restrict__ vatom1) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nall) {
#pragma unroll
        for (int d = 0; d < 3; ++d) f[(size_t)d * nall + t] = (float)F[(size_t)d * nall + t];
        for (int q = 0; q < nv1; ++q) vatom1[(size_t)q * nall + t] = (float)V1[(size_t)q * nall + t];
    }
    if (t < inum) {
        if (eflag) eatom0[t] = (float)E[t];
        if (vflag)
#pragma unroll
            for (int q = 0; q < 6; ++q) vatom0[(size_t)q * inum + t] = (float)V0[(size_t)q * inum + t];
    }
}

}  // namespace

void launch_lmp_rows(hipStream_t s, int inum, const int *ilist, const int *numneigh, const int *firstneigh,
                     const double *x, double rc2, const NeighborList &nl, int *overflow) {
    if (inum <= 0) return;
    hipLaunchKernelGGL(k_lmp_rows, dim3((inum + 255) / 256), dim3(256), 0, s, inum, ilist, numneigh, firstneigh, x,
                       rc2, nl, overflow);
}
void launch_lmp_types(hipStream_t s, int n, const int *type, const int *map, int *out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_lmp_types, dim3((n + 255) / 256), dim3(256), 0, s, n, type, map, out);
}
void launch_zero_double(hipStream_t s, size_t n, double *p) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_zero_double, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, p);
}
void launch_lmp_scatter(hipStream_t s, int inum, int nall, const int *ilist, const NeighborList &nl, int vflag,
                        int vflag_atom, int cvflag_atom, double *F, double *V0, double *V1) {
    if (inum <= 0) return;
    hipLaunchKernelGGL(k_lmp_scatter, dim3((inum + 255) / 256), dim3(256), 0, s, inum, nall, ilist, nl, vflag,
                       vflag_atom, cvflag_atom, F, V0, V1);
}
void launch_lmp_out(hipStream_t s, int inum, int nall, int eflag, int vflag, int nv1, const double *F, const double *E,
                    const double *V0, const double *V1, float *f, float *eatom0, float *vatom0, float *vatom1) {
    const int n = nall > inum ? nall : inum;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_lmp_out, dim3((n + 255) / 256), dim3(256), 0, s, inum, nall, eflag, vflag, nv1, F, E, V0, V1,
                       f, eatom0, vatom0, vatom1);
}

}  // namespace jnn
