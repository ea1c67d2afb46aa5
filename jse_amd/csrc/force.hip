// Forces, per-atom virials and totals in FP64.
//
// The reference accumulates, for every pair (i -> j) with g = -dE_i/d(r_j - r_i):
//   F[i] -= g, F[j] += g, V[j] += (dx gx, dy gy, dz gz, dx gy, dx gz, dy gz, dy gx, dz gx, dz gy)
// (src/nnap/java/jsex/nnap/NNAP.java:575-592, src/main/java/jse/atom/
// IPairPotential.java:938-1006).  Here every atom gathers instead of being
// scattered into: the incoming g_{j->i} is read through the reverse-pair slot
// of the neighbour list, so the sums are deterministic, need no atomics and
// are carried in FP64 with Neumaier compensation.  The displacement of the
// incoming pair is recomputed in FP64 exactly as the neighbour build formed
// it: dx = (x_i + s L) - x_j with s the reverse shift.
#include "kernels.h"

namespace jnn {

struct Neum {
    double s = 0.0, c = 0.0;
    __device__ __forceinline__ void add(double v) {
        const double t = s + v;
        if (fabs(s) >= fabs(v)) c += (s - t) + v;
        else c += (v - t) + s;
        s = t;
    }
    __device__ __forceinline__ double get() const { return s + c; }
};

__global__ __launch_bounds__(256) void k_gather(int n, const double *__restrict__ xw, double bx, double by, double bz,
                                                NeighborList nl, double *__restrict__ F, double *__restrict__ V,
                                                int want_virial) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int ci = min(nl.cnt[i], nl.cap);
    const size_t base = (size_t)i * nl.cap;
    Neum fx, fy, fz;
    Neum v[9];
    const double xi = xw[3 * i], yi = xw[3 * i + 1], zi = xw[3 * i + 2];
    for (int k = 0; k < ci; ++k) {
        fx.add(-(double)nl.gx[base + k]);
        fy.add(-(double)nl.gy[base + k]);
        fz.add(-(double)nl.gz[base + k]);
        const int r = nl.rev[base + k];
        if (r < 0) continue;
        const int key = nl.key[base + k];
        const int j = key >> 5;
        const size_t sj = (size_t)j * nl.cap + r;
        const double gx = (double)nl.gx[sj], gy = (double)nl.gy[sj], gz = (double)nl.gz[sj];
        fx.add(gx);
        fy.add(gy);
        fz.add(gz);
        if (want_virial) {
            const int code = 26 - (key & 31);  // shift of the pair j -> i
            const int sx = code % 3 - 1, sy = (code / 3) % 3 - 1, sz = code / 9 - 1;
            double ax = xi, ay = yi, az = zi;
            if (sx) ax = ax + (sx > 0 ? bx : -bx);
            if (sy) ay = ay + (sy > 0 ? by : -by);
            if (sz) az = az + (sz > 0 ? bz : -bz);
            const double dx = ax - xw[3 * j], dy = ay - xw[3 * j + 1], dz = az - xw[3 * j + 2];
            v[0].add(dx * gx); v[1].add(dy * gy); v[2].add(dz * gz);
            v[3].add(dx * gy); v[4].add(dx * gz); v[5].add(dy * gz);
            v[6].add(dy * gx); v[7].add(dz * gx); v[8].add(dz * gy);
        }
    }
    F[3 * i] = fx.get();
    F[3 * i + 1] = fy.get();
    F[3 * i + 2] = fz.get();
    if (want_virial) {
#pragma unroll
        for (int c = 0; c < 9; ++c) V[9 * (size_t)i + c] = v[c].get();
    }
}

// stage 1: each block sums a contiguous slice (fixed tree order), stage 2: one block sums the partials
constexpr int kRedBlock = 256;
__global__ __launch_bounds__(kRedBlock) void k_reduce_partial(int n, int stride, int ncomp, const double *__restrict__ in,
                                                             double *__restrict__ partial) {
    __shared__ double sh[kRedBlock];
    const int per = (n + gridDim.x - 1) / gridDim.x;
    const int b = blockIdx.x * per, e = min(b + per, n);
    for (int c = 0; c < ncomp; ++c) {
        Neum acc;
        for (int i = b + threadIdx.x; i < e; i += kRedBlock) acc.add(in[(size_t)i * stride + c]);
        sh[threadIdx.x] = acc.get();
        __syncthreads();
        for (int off = kRedBlock / 2; off > 0; off >>= 1) {
            if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[(size_t)blockIdx.x * ncomp + c] = sh[0];
        __syncthreads();
    }
}
__global__ __launch_bounds__(kRedBlock) void k_reduce_final(int np, int ncomp, const double *__restrict__ partial,
                                                           double *__restrict__ out) {
    if (threadIdx.x >= ncomp) return;
    Neum acc;
    for (int b = 0; b < np; ++b) acc.add(partial[(size_t)b * ncomp + threadIdx.x]);
    out[threadIdx.x] = acc.get();
}

void launch_gather(hipStream_t s, int n, const double *xw, const double *box, const NeighborList &nl, double *F,
                   double *V, int want_virial) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, s, n, xw, box[0], box[1], box[2], nl, F, V,
                       want_virial);
}
void launch_reduce(hipStream_t s, int n, int stride, int ncomp, const double *in, double *partial, int npartial,
                   double *out) {
    if (n <= 0) {  // empty system: the final pass over zero partials writes zeros
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kRedBlock), 0, s, 0, ncomp, partial, out);
        return;
    }
    hipLaunchKernelGGL(k_reduce_partial, dim3(npartial), dim3(kRedBlock), 0, s, n, stride, ncomp, in, partial);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kRedBlock), 0, s, npartial, ncomp, partial, out);
}

}  // namespace jnn
