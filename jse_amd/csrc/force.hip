// Forces, per-atom virials and totals in FP64.
//
// The reference accumulates, for every pair (i -> j) with g = -dE_i/d(r_j - r_i):
//   F[i] -= g, F[j] += g, V[j] += (dx gx, dy gy, dz gz, dx gy, dx gz, dy gz, dy gx, dz gx, dz gy)
// (src/nnap/java/jsex/nnap/NNAP.java:575-592, src/main/java/jse/atom/
// IPairPotential.java:938-1006).  Here every atom gathers instead of being
// scattered into, so the sums are deterministic and need no atomics: the
// incoming g_{j->i} sits at the reverse pair's slot of j's sorted row, found
// by a binary search (j's row comes with the pair, in the pair tag of d.w).
// Each lane sums in FP64 and the lanes of an atom are combined in a fixed
// order on the vector ALU; the totals use
// Neumaier-compensated FP64 partial sums.  The virial of an incoming pair
// j -> i takes its displacement r_i - r_j as the negated displacement of the
// pair i -> j in i's own row (j's row holds the same vector formed as
// (x_i + s L) - x_j, which can differ from it in the last FP32 bit).
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

// DPP row rotation of a double (row_ror:n is DPP control 0x120 + n): lane l of
// each row of 16 lanes receives lane (l + n) mod 16 of the same row.
__device__ __forceinline__ double dpp_ror(double v, int ctrl_ror) {
    const long long b = __double_as_longlong(v);
    int lo = (int)(b & 0xffffffffll), hi = (int)(b >> 32);
    switch (ctrl_ror) {
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
// Sum over each row of 16 lanes in a fixed order (DPP row rotations); every
// lane of the row gets it.
__device__ __forceinline__ double row16_sum(double v) {
    v += dpp_ror(v, 8);
    v += dpp_ror(v, 4);
    v += dpp_ror(v, 2);
    v += dpp_ror(v, 1);
    return v;
}

constexpr int kGatherMaxWaves = 8;
constexpr int kGatherSlots = 4;  // pairs per lane and pass: 16 lanes x 4 = 64 pairs of an atom

// Sixteen lanes per atom (a DPP row), four atoms per wavefront.  In each pass
// over 64 pairs of the atom's row, lane l takes pairs l, l + 16, l + 32 and
// l + 48: their own g (F_i -= g) and four binary searches of the neighbours'
// sorted rows for the reverse pairs, run in lockstep so their loads overlap;
// j's row comes with the pair (the tag in d.w).  The 12 sums of the row of
// lanes are then combined in a fixed order with DPP row rotations.
__global__ __launch_bounds__(64 * kGatherMaxWaves) void k_gather(int n, const int *__restrict__ atoms,
                                                                 NeighborList nl, double *__restrict__ F,
                                                                 double *__restrict__ V, int want_virial,
                                                                 const int *__restrict__ out_of,
                                                                 int *__restrict__ unmatched) {
    constexpr int U = kGatherSlots;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int grp = lane >> 4, lid = lane & 15;
    // rows are in cell order: each XCD gathers one contiguous run of rows, so
    // the neighbours' rows it reads sit in its own L2
    const int t = (xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + wv) * 4 + grp;
    const bool act = t < n;  // no early return: every lane takes part in the row sums
    const int i = act ? (atoms ? atoms[t] : t) : 0;  // row of the list
    const int ia = act && out_of ? out_of[i] : i;    // its atom
    const int cap = nl.cap;
    const int ci = act ? min(nl.cnt[i], cap) : 0;
    const size_t base = (size_t)i * cap;
    double f[3] = {0.0, 0.0, 0.0};
    double v[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < ci; k0 += 16 * U) {
        bool on[U], live[U];
        int want[U], jr[U], lo[U], hi[U], res[U];
        double dx[U], dy[U], dz[U];  // d, plus dlo where the list keeps FP64 displacements
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + lid + 16 * u;
            on[u] = k < ci;
            jr[u] = 0;
            want[u] = 0;
            dx[u] = dy[u] = dz[u] = 0.0;
            if (on[u]) {
                const float4 go = nl.g[base + k];
                f[0] -= (double)go.x;
                f[1] -= (double)go.y;
                f[2] -= (double)go.z;
                const float4 dk = nl.d[base + k];
                const float4 dl = (want_virial && nl.dlo) ? nl.dlo[base + k] : make_float4(0.f, 0.f, 0.f, 0.f);
                dx[u] = (double)dk.x + (double)dl.x;
                dy[u] = (double)dk.y + (double)dl.y;
                dz[u] = (double)dk.z + (double)dl.z;
                jr[u] = tag_row(dk.w);
                want[u] = ia * 32 + (26 - (nl.key[base + k] & 31));
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            lo[u] = 0;
            hi[u] = on[u] ? min(nl.cnt[jr[u]], cap) - 1 : -1;
            res[u] = -1;
            live[u] = hi[u] >= 0;
        }
        // the four searches advance together until every one has ended
        while (true) {
            bool any = false;
#pragma unroll
            for (int u = 0; u < U; ++u) any = any || live[u];
            if (!__any(any)) break;
            int mid[U], key[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                mid[u] = (lo[u] + hi[u]) >> 1;
                key[u] = live[u] ? nl.key[(size_t)jr[u] * cap + mid[u]] : 0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (!live[u]) continue;
                if (key[u] == want[u]) {
                    res[u] = mid[u];
                    live[u] = false;
                } else {
                    if (key[u] < want[u]) lo[u] = mid[u] + 1;
                    else hi[u] = mid[u] - 1;
                    live[u] = lo[u] <= hi[u];
                }
            }
        }
        float4 gi[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            gi[u] = res[u] >= 0 ? nl.g[(size_t)jr[u] * cap + res[u]] : make_float4(0.f, 0.f, 0.f, 0.f);
            if (unmatched && on[u] && res[u] < 0) atomicAdd(unmatched, 1);  // j's row lacks the pair
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (res[u] < 0) continue;
            const double gx = (double)gi[u].x, gy = (double)gi[u].y, gz = (double)gi[u].z;
            f[0] += gx;
            f[1] += gy;
            f[2] += gz;
            if (want_virial) {
                // displacement of the pair j -> i: r_i - r_j = -(r_j - r_i)
                const double ex = -dx[u], ey = -dy[u], ez = -dz[u];
                v[0] += ex * gx; v[1] += ey * gy; v[2] += ez * gz;
                v[3] += ex * gy; v[4] += ex * gz; v[5] += ey * gz;
                v[6] += ey * gx; v[7] += ez * gx; v[8] += ez * gy;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) f[c] = row16_sum(f[c]);
    if (want_virial) {
#pragma unroll
        for (int c = 0; c < 9; ++c) v[c] = row16_sum(v[c]);
    }
    if (act && lid == 0) {
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
    // the partials (written from other XCDs) are loaded in parallel, up to one per thread
    // at a time, and summed from LDS in chunk order
    Neum acc;
    for (int q = q0; q < q1; q += kRedBlock) {
        const int nq = min(kRedBlock, q1 - q);
        __syncthreads();
        if (tid < nq) {
#pragma unroll
            for (int c = 0; c < kComp; ++c) sh[c][tid] = partial[(size_t)(q + tid) * kComp + c];
        }
        __syncthreads();
        if (tid < kComp)
            for (int u = 0; u < nq; ++u) acc.add(sh[tid][u]);
    }
    if (tid < kComp && (tid == 0 || want_virial)) out[(size_t)r * kComp + tid] = acc.get();
    if (tid == 0) arrivals[r] = 0;
}

void launch_gather(hipStream_t s, int n, const int *atoms, const NeighborList &nl, double *F, double *V,
                   int want_virial, int waves, const int *out_of, int *unmatched) {
    if (n <= 0) return;
    const int w = (waves >= 1 && waves <= kGatherMaxWaves) ? waves : 4;
    const int atoms_per_wg = 4 * w;  // four atoms per wavefront
    hipLaunchKernelGGL(k_gather, dim3(xcd_grid((n + atoms_per_wg - 1) / atoms_per_wg)), dim3(64 * w), 0, s, n, atoms,
                       nl, F, V, want_virial, out_of, unmatched);
}
void launch_replica_totals(hipStream_t s, int nchunk, const int4 *chunks, int nrep, const int *rep_chunk_off,
                           const double *E, const double *V, int want_virial, double *partial, double *out,
                           int *arrivals) {
    if (nrep <= 0 || nchunk <= 0) return;
    hipLaunchKernelGGL(k_rep_totals, dim3(nchunk), dim3(kRedBlock), 0, s, chunks, rep_chunk_off, E, V, want_virial,
                       partial, out, arrivals);
}

}  // namespace jnn
