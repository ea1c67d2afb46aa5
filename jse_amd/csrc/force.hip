// Forces, per-atom virials and totals in FP64.
//
// The reference accumulates, for every pair (i -> j) with g = -dE_i/d(r_j - r_i):
//   F[i] -= g, F[j] += g, V[j] += (dx gx, dy gy, dz gz, dx gy, dx gz, dy gz, dy gx, dz gx, dz gy)
// (src/nnap/java/jsex/nnap/NNAP.java:575-592, src/main/java/jse/atom/
// IPairPotential.java:938-1006).  Here every atom gathers instead of being
// scattered into: the incoming g_{j->i} is read through the reverse-pair slot
// of the neighbour list (found by a binary search of j's sorted row), so the
// sums are deterministic and need no atomics;
// each lane sums its slots in FP64 and the wave combines lanes in a fixed
// order on the vector ALU, the totals use Neumaier-compensated FP64 partial
// sums.  The
// virial of an incoming pair j -> i takes its displacement r_i - r_j as the
// negated displacement of the pair i -> j in i's own row (read coalesced;
// j's row holds the same vector formed as (x_i + s L) - x_j, which can differ
// from it in the last FP32 bit).
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

// One wavefront per atom: lane k takes slot k of the row (and k+64, ...), and
// the 12 sums are reduced across the wave in a fixed order on the vector ALU
// (no LDS traffic): DPP rotations within each row of 16 lanes (every lane then
// holds its row's sum), then lane 0 adds rows 1-3 read with readlane.  The
// result is valid in lane 0.
__device__ __forceinline__ double dpp_ror(double v, int ctrl_ror) {
    const long long b = __double_as_longlong(v);
    int lo = (int)(b & 0xffffffffll), hi = (int)(b >> 32);
    switch (ctrl_ror) {  // row_ror:n is DPP control 0x120 + n
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
__device__ __forceinline__ double lane_of(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_ror(v, 8);
    v += dpp_ror(v, 4);
    v += dpp_ror(v, 2);
    v += dpp_ror(v, 1);
    return ((v + lane_of(v, 16)) + lane_of(v, 32)) + lane_of(v, 48);
}

__global__ __launch_bounds__(512) void k_gather(int n, const int *__restrict__ atoms, NeighborList nl,
                                                double *__restrict__ F, double *__restrict__ V, int want_virial,
                                                const int *__restrict__ out_of, const int *__restrict__ row_of) {
    const int lane = threadIdx.x & 63;
    // rows are in cell order: each XCD gathers one contiguous run of rows, so
    // the neighbours' rows it reads sit in its own L2
    const int t = xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (t >= n) return;
    const int i = atoms ? atoms[t] : t;       // row of the list
    const int ia = out_of ? out_of[i] : i;    // its atom
    const int ci = min(nl.cnt[i], nl.cap);
    const size_t base = (size_t)i * nl.cap;
    double f[3] = {0.0, 0.0, 0.0};
    double v[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = lane; k < ci; k += 64) {
        const float4 go = nl.g[base + k];
        f[0] -= (double)go.x;
        f[1] -= (double)go.y;
        f[2] -= (double)go.z;
        const int key = nl.key[base + k];
        const float4 di = want_virial ? nl.d[base + k] : make_float4(0.f, 0.f, 0.f, 0.f);
        int jrow;
        const int r = find_reverse(nl, ia, key, row_of, &jrow);  // binary search in j's sorted row
        if (r < 0) continue;
        const size_t sj = (size_t)jrow * nl.cap + r;
        const float4 gi = nl.g[sj];
        const double gx = (double)gi.x, gy = (double)gi.y, gz = (double)gi.z;
        f[0] += gx;
        f[1] += gy;
        f[2] += gz;
        if (want_virial) {
            // displacement of the pair j -> i: r_i - r_j = -(r_j - r_i)
            const double dx = -(double)di.x, dy = -(double)di.y, dz = -(double)di.z;
            v[0] += dx * gx; v[1] += dy * gy; v[2] += dz * gz;
            v[3] += dx * gy; v[4] += dx * gz; v[5] += dy * gz;
            v[6] += dy * gx; v[7] += dz * gx; v[8] += dz * gy;
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) f[c] = wave_sum(f[c]);
    if (want_virial) {
#pragma unroll
        for (int c = 0; c < 9; ++c) v[c] = wave_sum(v[c]);
    }
    if (lane == 0) {
        const size_t o = (size_t)ia;
        F[3 * o] = f[0];
        F[3 * o + 1] = f[1];
        F[3 * o + 2] = f[2];
        if (want_virial) {
#pragma unroll
            for (int c = 0; c < 9; ++c) V[9 * o + c] = v[c];
        }
    }
}

constexpr int kRedBlock = 256;
constexpr int kComp = 10;      // E, V[9]
constexpr int kPerThread = 4;  // atoms per thread: one chunk is at most kRedBlock * kPerThread atoms
static_assert(kRedBlock * kPerThread == kTotalsChunkMax, "totals chunk size");
// One workgroup per chunk (replica, first, end): every thread loads its (at
// most 4) atoms' E and V at once, sums them with Neumaier compensation in
// atom order, and the workgroup combines the 10 sums in a fixed tree into the
// chunk's partial.  The last chunk of a replica to finish (an arrival count
// per replica) then sums that replica's partials in chunk order, Neumaier
// again, and resets the count for the next evaluation; the order of every
// sum is fixed, so the totals do not depend on which chunk finishes last.
// Replicas without atoms have no chunk: their totals keep the zeros they
// were allocated with.
__global__ __launch_bounds__(kRedBlock) void k_rep_totals(const int4 *__restrict__ chunks,
                                                         const int *__restrict__ rep_chunk_off,
                                                         const double *__restrict__ E, const double *__restrict__ V,
                                                         int want_virial, double *__restrict__ partial,
                                                         double *__restrict__ out, int *__restrict__ arrivals) {
    __shared__ double sh[kComp][kRedBlock];
    __shared__ int s_last;
    const int4 ch = chunks[blockIdx.x];  // (replica, first atom, end atom, -)
    const int tid = threadIdx.x;
    double v[kPerThread][kComp];
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
        const int i = ch.y + tid + kRedBlock * u;
        const bool on = i < ch.z;
        v[u][0] = on ? E[i] : 0.0;
#pragma unroll
        for (int c = 1; c < kComp; ++c) v[u][c] = (on && want_virial) ? V[9 * (size_t)i + c - 1] : 0.0;
    }
#pragma unroll
    for (int c = 0; c < kComp; ++c) {
        Neum acc;
#pragma unroll
        for (int u = 0; u < kPerThread; ++u) acc.add(v[u][c]);
        sh[c][tid] = acc.get();
    }
    __syncthreads();
    for (int off = kRedBlock / 2; off > 0; off >>= 1) {
        if (tid < off) {
#pragma unroll
            for (int c = 0; c < kComp; ++c) sh[c][tid] += sh[c][tid + off];
        }
        __syncthreads();
    }
    const int r = ch.x;
    const int q0 = rep_chunk_off[r], q1 = rep_chunk_off[r + 1];
    if (tid < kComp) partial[(size_t)blockIdx.x * kComp + tid] = sh[tid][0];
    __threadfence();  // the partial is visible device-wide before this chunk's arrival counts
    __syncthreads();
    if (tid == 0) s_last = atomicAdd(arrivals + r, 1) == q1 - q0 - 1;
    __syncthreads();
    if (!s_last) return;
    __threadfence();  // every earlier chunk's partial is read after its arrival
    if (tid < kComp && (tid == 0 || want_virial)) {
        Neum acc;
        for (int q = q0; q < q1; ++q) acc.add(partial[(size_t)q * kComp + tid]);
        out[(size_t)r * kComp + tid] = acc.get();
    }
    if (tid == 0) arrivals[r] = 0;
}

void launch_gather(hipStream_t s, int n, const int *atoms, const NeighborList &nl, double *F, double *V,
                   int want_virial, int waves, const int *out_of, const int *row_of) {
    if (n <= 0) return;
    const int w = (waves >= 1 && waves <= 8) ? waves : 4;
    hipLaunchKernelGGL(k_gather, dim3(xcd_grid((n + w - 1) / w)), dim3(64 * w), 0, s, n, atoms, nl, F, V, want_virial,
                       out_of, row_of);
}
void launch_replica_totals(hipStream_t s, int nchunk, const int4 *chunks, int nrep, const int *rep_chunk_off,
                           const double *E, const double *V, int want_virial, double *partial, double *out,
                           int *arrivals) {
    if (nrep <= 0 || nchunk <= 0) return;
    hipLaunchKernelGGL(k_rep_totals, dim3(nchunk), dim3(kRedBlock), 0, s, chunks, rep_chunk_off, E, V, want_virial,
                       partial, out, arrivals);
}

}  // namespace jnn
