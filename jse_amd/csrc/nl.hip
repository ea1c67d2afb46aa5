// Periodic neighbour-list build on the GPU.
//
// Semantics follow the reference's cell list (paths relative to the reference
// repository root): positions wrapped as IBox.wrapPBC (src/main/java/jse/atom/
// IBox.java:145-151); cell grid of NeighborListGetter.getProperLinkedCell
// (NeighborListGetter.java:952-1001); a pair (j, image) is kept when the FP64
// displacement dx = (x_j + shift) - x_i satisfies dx*dx+dy*dy+dz*dz < rc*rc,
// evaluated left to right without fused multiply-add
// (NeighborListGetter.java:1041-1048); self is excluded only at zero shift
// (:129, :153-171).  Each atom's row is sorted by key = j*32 + shift_code
// (shift_code = (sx+1) + 3(sy+1) + 9(sz+1)), so the set can be compared
// bit-exactly with the reference's (which visits cells in a fixed order).
//
// Layout: positions wrapped into SoA-free AoS FP64 [N][3]; atoms counting-
// sorted by cell so a cell's members are contiguous; one wavefront per atom
// scans the 27 neighbouring cells as one flattened candidate range, compacts
// hits with a wave ballot into LDS and writes each hit straight to its slot of
// the sorted row: keys are unique, so the slot is the number of smaller keys
// among the atom's hits (counted against the LDS hit buffer).
#include <algorithm>

#include "common.h"
#include "kernels.h"

namespace jnn {

constexpr int kNlBlock = 64;
constexpr int kNlLds = 512;  // candidate hits staged per atom (rows longer than this overflow)

__global__ void k_wrap_bin(int n, const int *__restrict__ atoms, const double *__restrict__ xyz_in,
                           const int *__restrict__ rep_of, const ReplicaGrid *__restrict__ reps,
                           const int *__restrict__ need, double *__restrict__ xw, int *__restrict__ cell,
                           int *__restrict__ cell_cnt, int *__restrict__ rank) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = atoms ? atoms[t] : t;
    if (need && !need[rep_of[i]]) return;
    const ReplicaGrid &g = reps[rep_of[i]];
    int c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double s = xyz_in[3 * i + d] / g.box[d];
        if (s < 0.0 || s >= 1.0) s -= floor(s);
        const double x = s * g.box[d];
        xw[3 * i + d] = x;
        int ci = (int)floor(x / g.cb[d]);
        ci = ci < 0 ? 0 : (ci >= g.size[d] ? g.size[d] - 1 : ci);
        c[d] = ci;
    }
    const int cc = g.cell_off + c[0] + g.size[0] * (c[1] + g.size[1] * c[2]);
    cell[i] = cc;
    rank[i] = atomicAdd(&cell_cnt[cc], 1);
}

// Exclusive scan of each replica's cell counts, one workgroup per replica:
// start[cell_off + c] = atom_off + (atoms in the replica's cells before c), so
// every atom gets a unique slot in its replica's slot range.
// With `clear`, the counts are zeroed once consumed (ready for the next
// binning) and *clear_flag (nullable) is zeroed by workgroup 0.
__global__ __launch_bounds__(1024) void k_cell_scan(const int *__restrict__ rep_list, const ReplicaGrid *__restrict__ reps,
                                                   const int *__restrict__ need, int *__restrict__ cnt,
                                                   int *__restrict__ start, int clear, int *__restrict__ clear_flag) {
    __shared__ int part[1024];
    const int r = rep_list ? rep_list[blockIdx.x] : blockIdx.x;
    if (need && !need[r]) return;
    const ReplicaGrid &g = reps[r];
    const int t = threadIdx.x;
    const int ncell = g.ncell;
    if (clear && clear_flag && blockIdx.x == 0 && threadIdx.x == 0) *clear_flag = 0;
    int *c_in = cnt + g.cell_off;
    int *c_out = start + g.cell_off;
    const int per = (ncell + 1023) / 1024;
    const int b = t * per, e = min(b + per, ncell);
    int s = 0;
    for (int c = b; c < e; ++c) s += c_in[c];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = (t >= off) ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = g.atom_off + part[t] - s;
    for (int c = b; c < e; ++c) {
        c_out[c] = run;
        run += c_in[c];
        if (clear) c_in[c] = 0;
    }
    if (t == 1023) c_out[ncell] = g.atom_off + part[1023];
}

__global__ void k_cell_scatter(int n, const int *__restrict__ atoms, const int *__restrict__ rep_of,
                               const int *__restrict__ need, const int *__restrict__ cell,
                               const int *__restrict__ rank, const int *__restrict__ start,
                               const double *__restrict__ xw, int *__restrict__ sidx, double *__restrict__ sxw,
                               int *__restrict__ slot_of) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = atoms ? atoms[t] : t;
    if (need && !need[rep_of[i]]) return;
    const int slot = start[cell[i]] + rank[i];
    sidx[slot] = i;
    if (slot_of) slot_of[i] = slot;
    sxw[3 * slot] = xw[3 * i];
    sxw[3 * slot + 1] = xw[3 * i + 1];
    sxw[3 * slot + 2] = xw[3 * i + 2];
}

// neighbour cell w (0..26) holding flattened candidate q: the last w with s_pref[w] <= q
__device__ __forceinline__ int cand_cell(const int *s_pref, int q) {
    int lo = 0, hi = 26;
#pragma unroll
    for (int it = 0; it < 5; ++it) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_pref[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ double dist2_no_fma(double dx, double dy, double dz) {
#pragma clang fp contract(off)
    const double a = dx * dx;
    const double b = dy * dy;
    const double c = dz * dz;
    return (a + b) + c;
}

__global__ __launch_bounds__(kNlBlock) void k_nl_build(int n, const int *__restrict__ atoms, const int *__restrict__ rep_of,
                                                       const ReplicaGrid *__restrict__ reps,
                                                       const int *__restrict__ need, double rc2,
                                                       const double *__restrict__ xw,
                                                       const int *__restrict__ cell, const int *__restrict__ cstart,
                                                       const int *__restrict__ sidx, const double *__restrict__ sxw,
                                                       NeighborList nl, int *__restrict__ overflow) {
    __shared__ int s_start[27], s_cnt[27], s_code[27], s_pref[28];
    __shared__ int s_key[kNlLds];
    __shared__ float s_d[3][kNlLds];
    const int lane = threadIdx.x;
    for (int t_ = blockIdx.x; t_ < n; t_ += gridDim.x) {
        const int i = atoms ? atoms[t_] : t_;
        const int ri = rep_of[i];
        if (need && !need[ri]) continue;   // uniform over the wave: no barrier is skipped by part of it
        const double xi = xw[3 * i], yi = xw[3 * i + 1], zi = xw[3 * i + 2];
        const ReplicaGrid &g = reps[ri];
        const int c = cell[i] - g.cell_off;
        const int cx = c % g.size[0], cy = (c / g.size[0]) % g.size[1], cz = c / (g.size[0] * g.size[1]);
        if (lane < 27) {
            const int w = lane;
            int ix = cx + (w % 3) - 1, iy = cy + (w / 3) % 3 - 1, iz = cz + w / 9 - 1;
            int sx = 0, sy = 0, sz = 0;
            if (ix >= g.size[0]) { ix -= g.size[0]; sx = 1; } else if (ix < 0) { ix += g.size[0]; sx = -1; }
            if (iy >= g.size[1]) { iy -= g.size[1]; sy = 1; } else if (iy < 0) { iy += g.size[1]; sy = -1; }
            if (iz >= g.size[2]) { iz -= g.size[2]; sz = 1; } else if (iz < 0) { iz += g.size[2]; sz = -1; }
            const int cc = g.cell_off + ix + g.size[0] * (iy + g.size[1] * iz);
            s_start[w] = cstart[cc];
            s_cnt[w] = cstart[cc + 1] - cstart[cc];
            s_code[w] = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
        }
        __syncthreads();
        if (lane == 0) {
            int run = 0;
            for (int w = 0; w < 27; ++w) {
                s_pref[w] = run;
                run += s_cnt[w];
            }
            s_pref[27] = run;
        }
        __syncthreads();
        const int total = s_pref[27];
        int hits = 0;
        for (int q0 = 0; q0 < total; q0 += kNlBlock) {
            const int q = q0 + lane;
            bool hit = false;
            int key = 0;
            float fdx = 0.f, fdy = 0.f, fdz = 0.f;
            if (q < total) {
                const int w = cand_cell(s_pref, q);
                const int slot = s_start[w] + (q - s_pref[w]);
                const int j = sidx[slot];
                const int code = s_code[w];
                const int sx = code % 3 - 1, sy = (code / 3) % 3 - 1, sz = code / 9 - 1;
                double xj = sxw[3 * slot], yj = sxw[3 * slot + 1], zj = sxw[3 * slot + 2];
                if (sx) xj = xj + (sx > 0 ? g.box[0] : -g.box[0]);
                if (sy) yj = yj + (sy > 0 ? g.box[1] : -g.box[1]);
                if (sz) zj = zj + (sz > 0 ? g.box[2] : -g.box[2]);
                const double dx = xj - xi, dy = yj - yi, dz = zj - zi;
                const double d2 = dist2_no_fma(dx, dy, dz);
                hit = (d2 < rc2) && !(j == i && code == 13);
                key = j * 32 + code;
                fdx = (float)dx;
                fdy = (float)dy;
                fdz = (float)dz;
            }
            const unsigned long long mask = __ballot(hit);
            const unsigned long long below = lane ? (mask & ((1ull << lane) - 1ull)) : 0ull;
            const int pos = hits + __popcll(below);
            if (hit && pos < kNlLds) {
                s_key[pos] = key;
                s_d[0][pos] = fdx;
                s_d[1][pos] = fdy;
                s_d[2][pos] = fdz;
            }
            hits += __popcll(mask);
        }
        const int nst = hits < kNlLds ? hits : kNlLds;
        __syncthreads();
        // keys are unique: a hit's slot in the sorted row is the number of smaller keys
        const size_t row = (size_t)i * nl.cap;
        for (int t = lane; t < nst; t += kNlBlock) {
            const int kt = s_key[t];
            int rank = 0;
            for (int u = 0; u < nst; ++u) rank += s_key[u] < kt ? 1 : 0;
            if (rank < nl.cap) {
                nl.key[row + rank] = kt;
                if (nl.d)  // rows by atom: j's row is j
                    nl.d[row + rank] = make_float4(s_d[0][t], s_d[1][t], s_d[2][t], pair_tag(kt >> 5, nl.types[kt >> 5]));
            }
        }
        if (lane == 0) {
            nl.cnt[i] = hits;
            if (hits > nl.cap) atomicMax(overflow, hits);
        }
        __syncthreads();
    }
}

// ---- cell-tile build: one workgroup per cell ----------------------------
// Rows are placed in cell-sorted slot order (row s is atom sidx[s]), so the
// rows the force gather visits for an atom's neighbours sit close together in
// memory; keys keep atom numbers.  The atoms of one cell share its 27-cell neighbourhood: the workgroup stages
// those candidates once in LDS (FP64 positions with the periodic shift
// already added, exactly as the per-atom build forms x_j + shift, and the key
// j*32 + code, and slot), then each wave tests one member atom at a time
// against the staged block, compacts its hits (key, candidate index) with a
// ballot and writes each hit to its rank in the row, its displacement formed
// again from the staged candidate by the same FP64 arithmetic.  A neighbourhood larger than the staging buffer is read
// from global memory instead (same arithmetic).
constexpr int kCellWaves = 4;
constexpr int kStage = 512;


// BLOCKS: the variant for replicas of large cells (a neighbour block is skipped
// whole when its cell lies beyond the cutoff from the atom); without it the
// code is the plain walk.
template <bool BLOCKS>
__global__ __launch_bounds__(64 * kCellWaves) void k_nl_build_cells(
    int ncell_total, const int *__restrict__ cell_rep, const ReplicaGrid *__restrict__ reps, double rc2,
    const int *__restrict__ cstart, const int *__restrict__ sidx, const double *__restrict__ sxw, NeighborList nl,
    int *__restrict__ overflow) {
    __shared__ int s_start[27], s_cnt[27], s_code[27], s_pref[28];
    __shared__ double s_lo[27][3], s_hi[27][3];  // each neighbour cell's extent, image shift applied
    __shared__ double s_cx[kStage], s_cy[kStage], s_cz[kStage];
    __shared__ int s_ckey[kStage], s_cslot[kStage], s_ctype[kStage];
    __shared__ int s_hkey[kCellWaves][kNlLds], s_hq[kCellWaves][kNlLds];  // hits: key, candidate index
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    // each XCD takes a contiguous run of cells, whose 27-cell blocks overlap in its L2
    for (int c = xcd_block(blockIdx.x, gridDim.x); c < ncell_total; c += gridDim.x) {
        const int r = cell_rep[c];
        if (r < 0) continue;  // a replica's end entry, not a cell
        const int cs = cstart[c], ce = cstart[c + 1];
        if (cs == ce) continue;
        const ReplicaGrid &g = reps[r];
        const int lc = c - g.cell_off;
        const int cx = lc % g.size[0], cy = (lc / g.size[0]) % g.size[1], cz = lc / (g.size[0] * g.size[1]);
        if (tid < 64) {  // wave 0: the 27 neighbour cells and their prefix (a shuffle scan)
            int cnt = 0;
            if (tid < 27) {
                const int w = tid;
                int ix = cx + (w % 3) - 1, iy = cy + (w / 3) % 3 - 1, iz = cz + w / 9 - 1;
                int sx = 0, sy = 0, sz = 0;
                if (ix >= g.size[0]) { ix -= g.size[0]; sx = 1; } else if (ix < 0) { ix += g.size[0]; sx = -1; }
                if (iy >= g.size[1]) { iy -= g.size[1]; sy = 1; } else if (iy < 0) { iy += g.size[1]; sy = -1; }
                if (iz >= g.size[2]) { iz -= g.size[2]; sz = 1; } else if (iz < 0) { iz += g.size[2]; sz = -1; }
                const int cc = g.cell_off + ix + g.size[0] * (iy + g.size[1] * iz);
                const int st = cstart[cc];
                cnt = cstart[cc + 1] - st;
                s_start[w] = st;
                s_cnt[w] = cnt;
                s_code[w] = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
                if constexpr (BLOCKS) {  // widened by 1e-6 of a cell edge: a binned atom never lies outside
                    const int ic[3] = {ix, iy, iz}, sc[3] = {sx, sy, sz};
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const double m = 1e-6 * g.cb[d];
                        s_lo[w][d] = ic[d] * g.cb[d] + sc[d] * g.box[d] - m;
                        s_hi[w][d] = (ic[d] + 1) * g.cb[d] + sc[d] * g.box[d] + m;
                    }
                }
            }
            int incl = cnt;
#pragma unroll
            for (int off = 1; off < 32; off <<= 1) {
                const int u = __shfl_up(incl, off);
                if (tid >= off) incl += u;
            }
            if (tid < 27) s_pref[tid] = incl - cnt;
            if (tid == 26) s_pref[27] = incl;
        }
        __syncthreads();
        const int total = s_pref[27];
        const bool staged = total <= kStage;
        // candidate q: its key, slot (= its own row) and FP64 shifted position
        auto fetch = [&](int q, int &key, int &slot, double &xj, double &yj, double &zj) {
            const int w = cand_cell(s_pref, q);
            slot = s_start[w] + (q - s_pref[w]);
            const int code = s_code[w];
            const int sx = code % 3 - 1, sy = (code / 3) % 3 - 1, sz = code / 9 - 1;
            xj = sxw[3 * (size_t)slot];
            yj = sxw[3 * (size_t)slot + 1];
            zj = sxw[3 * (size_t)slot + 2];
            if (sx) xj = xj + (sx > 0 ? g.box[0] : -g.box[0]);
            if (sy) yj = yj + (sy > 0 ? g.box[1] : -g.box[1]);
            if (sz) zj = zj + (sz > 0 ? g.box[2] : -g.box[2]);
            key = sidx[slot] * 32 + code;
        };
        if (staged) {
            for (int q = tid; q < total; q += 64 * kCellWaves) {
                int key, slot;
                double xj, yj, zj;
                fetch(q, key, slot, xj, yj, zj);
                s_cx[q] = xj;
                s_cy[q] = yj;
                s_cz[q] = zj;
                s_ckey[q] = key;
                s_cslot[q] = slot;
                s_ctype[q] = nl.types[key >> 5];
            }
        }
        __syncthreads();
        int *hkey = s_hkey[wv], *hq = s_hq[wv];
        for (int sl = cs + wv; sl < ce; sl += kCellWaves) {
            // the atom itself: from the staged centre cell (no shift), else from memory
            int i;  // the row goes to slot sl; keys use atom numbers
            double xi, yi, zi;
            if (staged) {
                const int q = s_pref[13] + (sl - cs);
                i = s_ckey[q] >> 5;
                xi = s_cx[q];
                yi = s_cy[q];
                zi = s_cz[q];
            } else {
                i = sidx[sl];
                xi = sxw[3 * (size_t)sl];
                yi = sxw[3 * (size_t)sl + 1];
                zi = sxw[3 * (size_t)sl + 2];
            }
            // candidate q: key, slot and displacement from atom i (the same FP64 arithmetic wherever it is formed)
            auto cand = [&](int q, int &key, int &slot, double &dx, double &dy, double &dz) {
                double xj, yj, zj;
                if (staged) {
                    key = s_ckey[q];
                    slot = s_cslot[q];
                    xj = s_cx[q];
                    yj = s_cy[q];
                    zj = s_cz[q];
                } else {
                    fetch(q, key, slot, xj, yj, zj);
                }
                dx = xj - xi;
                dy = yj - yi;
                dz = zj - zi;
            };
            int hits = 0;
            auto test_range = [&](int qa, int qb) {  // candidates qa .. qb-1, 64 at a time
                for (int q0 = qa; q0 < qb; q0 += 64) {
                    const int q = q0 + lane;
                    bool hit = false;
                    int key = 0;
                    if (q < qb) {
                        int slot;
                        double dx, dy, dz;
                        cand(q, key, slot, dx, dy, dz);
                        hit = (dist2_no_fma(dx, dy, dz) < rc2) && !((key >> 5) == i && (key & 31) == 13);
                    }
                    const unsigned long long mask = __ballot(hit);
                    const unsigned long long below = lane ? (mask & ((1ull << lane) - 1ull)) : 0ull;
                    const int pos = hits + __popcll(below);
                    if (hit && pos < kNlLds) {
                        hkey[pos] = key;
                        hq[pos] = q;
                    }
                    hits += __popcll(mask);
                }
            };
            if (BLOCKS && total >= 27 * 32) {
                // large cells (a replica of one or two cells per axis): a neighbour cell
                // whose shifted extent lies rc or farther from atom i along the three axes
                // together holds no neighbour, and its block is skipped (rows are ranked
                // by key below, so the order the blocks are visited in does not matter)
                const double xa[3] = {xi, yi, zi};
                for (int w = 0; w < 27; ++w) {
                    const int qa = s_pref[w], qb = s_pref[w + 1];
                    if (qa == qb) continue;
                    double d2 = 0.0;
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const double t = s_lo[w][d] > xa[d] ? s_lo[w][d] - xa[d] : (xa[d] > s_hi[w][d] ? xa[d] - s_hi[w][d] : 0.0);
                        d2 += t * t;
                    }
                    if (d2 < rc2) test_range(qa, qb);
                }
            } else {
                test_range(0, total);
            }
            wave_lds_sync();
            // keys are unique: a hit's slot in the sorted row is the number of smaller
            // keys.  Up to 64 hits each lane holds its hit's key and counts against the
            // others read lane by lane (no LDS round trips); longer rows count against
            // the whole hit buffer (broadcast reads).
            const int nst = hits < kNlLds ? hits : kNlLds;
            const size_t row = (size_t)sl * nl.cap;
            for (int t = lane; t < nst; t += 64) {
                const int kt = hkey[t];
                int rank = 0;
                if (nst <= 64) {
                    for (int u = 0; u < nst; ++u) rank += __builtin_amdgcn_readlane(kt, u) < kt ? 1 : 0;
                } else {
                    for (int u = 0; u < nst; ++u) rank += hkey[u] < kt ? 1 : 0;
                }
                if (rank < nl.cap) {
                    nl.key[row + rank] = kt;
                    if (nl.d) {
                        const int q = hq[t];
                        int key, slot;
                        double dx, dy, dz;
                        cand(q, key, slot, dx, dy, dz);
                        const int tj = staged ? s_ctype[q] : nl.types[kt >> 5];
                        // j's own row is its slot
                        const float fx = (float)dx, fy = (float)dy, fz = (float)dz;
                        nl.d[row + rank] = make_float4(fx, fy, fz, pair_tag(slot, tj));
                        if (nl.dlo)
                            nl.dlo[row + rank] = make_float4((float)(dx - (double)fx), (float)(dy - (double)fy),
                                                             (float)(dz - (double)fz), 0.f);
                    }
                }
            }
            if (lane == 0) {
                nl.cnt[sl] = hits;
                if (hits > nl.cap) atomicMax(overflow, hits);
            }
            wave_lds_sync();  // this wave's hit buffer is reused for its next atom
        }
        __syncthreads();  // the staged block is reused for the next cell
    }
}

// rev[i][k] = slot of (i, -shift) in j's row (binary search on the sorted keys)
__global__ void k_nl_reverse(int n, const int *__restrict__ atoms, NeighborList nl, const int *__restrict__ out_of,
                             const int *__restrict__ row_of) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ti = (int)(t / nl.cap), k = (int)(t % nl.cap);
    if (ti >= n) return;
    const int r = atoms ? atoms[ti] : ti;
    const int ci = min(nl.cnt[r], nl.cap);
    if (k >= ci) return;
    int jrow;
    nl.rev[(size_t)r * nl.cap + k] = find_reverse(nl, out_of ? out_of[r] : r, nl.key[(size_t)r * nl.cap + k], row_of, &jrow);
}

__global__ void k_fill_int(int n, int *__restrict__ p, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
void launch_fill_int(hipStream_t s, int n, int *p, int v) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, s, n, p, v);
}

// Copy c of atom i at xyz[i] + (cx L_x, cy L_y, cz L_z), c = cx + cp0 (cy + cp1 cz),
// written to row c * nuser + i of out.
__global__ void k_expand_copies(int nuser, int cp0, int cp1, int ncopy, double lx, double ly, double lz,
                                const double *__restrict__ xyz, double *__restrict__ out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nuser * ncopy) return;
    const int c = (int)(t / nuser), i = (int)(t % nuser);
    const int cx = c % cp0, cy = (c / cp0) % cp1, cz = c / (cp0 * cp1);
    out[3 * t] = xyz[3 * i] + lx * cx;
    out[3 * t + 1] = xyz[3 * i + 1] + ly * cy;
    out[3 * t + 2] = xyz[3 * i + 2] + lz * cz;
}
void launch_expand_copies(hipStream_t s, int nuser, const int *cp, const double *box, const double *xyz,
                          double *out) {
    const long long tot = (long long)nuser * cp[0] * cp[1] * cp[2];
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_expand_copies, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, nuser, cp[0], cp[1],
                       cp[0] * cp[1] * cp[2], box[0], box[1], box[2], xyz, out);
}

void launch_wrap_bin(hipStream_t s, int n, const int *atoms, const double *xyz_in, const int *rep_of,
                     const ReplicaGrid *reps, const int *need, double *xw, int *cell, int *cell_cnt, int *rank) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_wrap_bin, dim3((n + 255) / 256), dim3(256), 0, s, n, atoms, xyz_in, rep_of, reps, need, xw,
                       cell, cell_cnt, rank);
}
void launch_cell_scan(hipStream_t s, int nrep, const int *rep_list, const ReplicaGrid *reps, const int *need,
                      int *cell_cnt, int *cell_start, bool clear, int *clear_flag) {
    if (nrep <= 0) return;
    hipLaunchKernelGGL(k_cell_scan, dim3(nrep), dim3(1024), 0, s, rep_list, reps, need, cell_cnt, cell_start,
                       clear ? 1 : 0, clear_flag);
}
void launch_cell_scatter(hipStream_t s, int n, const int *atoms, const int *rep_of, const int *need, const int *cell,
                         const int *rank, const int *cell_start, const double *xw, int *sidx, double *sxw,
                         int *slot_of) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cell_scatter, dim3((n + 255) / 256), dim3(256), 0, s, n, atoms, rep_of, need, cell, rank,
                       cell_start, xw, sidx, sxw, slot_of);
}
int device_cu_count() {
    static int ncu = 0;
    if (!ncu) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            ncu = v;
        else
            ncu = 256;
    }
    return ncu;
}

void launch_nl_build(hipStream_t s, int n, const int *atoms, const int *rep_of, const ReplicaGrid *reps,
                     const int *need, double rc2, const double *xw, const int *cell, const int *cell_start,
                     const int *sidx, const double *sxw, NeighborList nl, int *overflow, int wg_per_cu) {
    if (n <= 0) return;
    // the per-atom build keeps FP32 displacements only (its callers build key-only candidate lists)
    JNN_CHECK(!nl.dlo, ERR_ARG, "the per-atom neighbour build does not write FP64 displacements");
    int grid = n < 65536 * 4 ? n : 65536 * 4;
    if (wg_per_cu > 0) grid = std::min(grid, wg_per_cu * device_cu_count());
    hipLaunchKernelGGL(k_nl_build, dim3(grid), dim3(kNlBlock), 0, s, n, atoms, rep_of, reps, need, rc2, xw, cell,
                       cell_start, sidx, sxw, nl, overflow);
}
void launch_nl_build_cells(hipStream_t s, int ncell_total, const int *cell_rep, const ReplicaGrid *reps, double rc2,
                           const int *cell_start, const int *sidx, const double *sxw, NeighborList nl, int *overflow,
                           int wg_per_cu, bool large_cells) {
    if (ncell_total <= 0) return;
    int grid = ncell_total;
    if (wg_per_cu > 0) grid = std::min(grid, wg_per_cu * device_cu_count());
    grid = xcd_grid(grid);
    if (large_cells)
        hipLaunchKernelGGL(k_nl_build_cells<true>, dim3(grid), dim3(64 * kCellWaves), 0, s, ncell_total, cell_rep,
                           reps, rc2, cell_start, sidx, sxw, nl, overflow);
    else
        hipLaunchKernelGGL(k_nl_build_cells<false>, dim3(grid), dim3(64 * kCellWaves), 0, s, ncell_total, cell_rep,
                           reps, rc2, cell_start, sidx, sxw, nl, overflow);
}
void launch_nl_reverse(hipStream_t s, int n, const int *atoms, NeighborList nl, const int *out_of, const int *row_of) {
    if (n <= 0) return;
    const long long tot = (long long)n * nl.cap;
    hipLaunchKernelGGL(k_nl_reverse, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, atoms, nl, out_of,
                       row_of);
}

}  // namespace jnn
