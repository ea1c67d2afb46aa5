// Per-species 2-hidden-layer MLP, forward and input-gradient backward, on
// MFMA matrix cores (reference semantics: nn_FeedForward.hpp:8-21,131-261;
// normalisation nnap_main.hpp:267-323).
//
// A workgroup of 4 wavefronts owns a tile of 64 atoms of one species (rows);
// wave w owns rows 16w..16w+15 and every output column, so the four GEMMs of
// one tile
//     Z1 = X W1^T + b1, H1 = silu(Z1)       [64 x S]  x [S x H1]
//     Z2 = H1 W2^T + b2, y = wo . silu(Z2)  [64 x H1] x [H1 x H2]
//     gZ1 = (gZ2 W2) * silu'(Z1)            [64 x H2] x [H2 x H1]
//     gX  = gZ1 W1                          [64 x H1] x [H1 x S]
// keep each wave's SiLU derivatives in its own registers between forward and
// backward (the same lane -> (row, column) map serves both), so the only
// cross-wave traffic is the staged weights.  X goes from memory straight into
// the registers of the first product's A operand (the next tile's X is loaded
// while the current tile's other three products run); hidden activations and
// weights live in LDS with row strides chosen against bank conflicts.  The
// grid is persistent (weights staged once per workgroup).
//
// Three element types share the code:
//   double : v_mfma_f64_16x16x4_f64   (the plain mode: FP64 inputs, weights,
//            activations and SiLU; its error stays far below the force bar)
//   float  : v_mfma_f32_16x16x4_f32   (the FP32-geometry mode)
//   __bf16 : v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32 accumulation); the
//            two forward GEMMs carry each operand as a hi + lo BF16 pair
//            (three MFMAs: hi*hi + lo*hi + hi*lo), which keeps the energy at
//            FP32-like accuracy, while the two backward GEMMs use single BF16.
// Lane maps (16x16 output tile, 64 lanes): A[i][k] from lane i + 16*(k/E),
// B[k][j] from lane j + 16*(k/E), with E = 1 (f64 and f32, 4 k per
// instruction) or 8 consecutive k per lane (bf16, 32 k); D[i][j] in lane
// j + 16*(i/4), element i%4 (f32, bf16) or lane j + 16*(i%4), element i/4
// (f64).  SiLU, biases, the energy and the output scaling are FP64 in the
// FP64 kernel and FP32 in the other two.
#include <type_traits>

#include "common.h"
#include "kernels.h"

namespace jnn {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr float kPi4m = 12.566370614359172f;

__device__ __forceinline__ __bf16 to_bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return __builtin_bit_cast(__bf16, (uint16_t)(u >> 16));
}

template <typename T>
struct Mma;

template <>
struct Mma<double> {
    static constexpr int KSTEP = 4;
    static constexpr int KPL = 1;
    static constexpr int PAD = 4;  // row stride == 8 (mod 64) dwords
    static constexpr bool kTransposedCopies = false;
    static constexpr bool kSplit = false;
    using AV = double;
    using CT = double;  // epilogue arithmetic
    using ACC = f64x4;
    __device__ static __forceinline__ double cvt(double v) { return v; }
    __device__ static __forceinline__ double lo(double v) { return 0.0; }
    __device__ static __forceinline__ int lane_k(int lane) { return lane >> 4; }
    // tile row of accumulator element r of this lane
    __device__ static __forceinline__ int row(int lane, int r) { return 4 * r + (lane >> 4); }
    __device__ static __forceinline__ f64x4 mma(const double *a, const double *b, f64x4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], b[0], c, 0, 0, 0);
    }
    __device__ static __forceinline__ AV a_hi(const double *x) { return x[0]; }
    __device__ static __forceinline__ AV a_lo(const double *x) { return 0.0; }
    __device__ static __forceinline__ f64x4 mma_r(AV a, const double *b, f64x4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[0], c, 0, 0, 0);
    }
};

template <>
struct Mma<float> {
    static constexpr int KSTEP = 4;
    static constexpr int KPL = 1;  // k values per lane and instruction
    static constexpr int PAD = 2;  // row stride == 2 (mod 32) dwords
    static constexpr bool kTransposedCopies = false;
    static constexpr bool kSplit = false;
    using AV = float;  // an A operand held in registers
    using CT = float;
    using ACC = f32x4;
    __device__ static __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
    __device__ static __forceinline__ float cvt(float v) { return v; }
    __device__ static __forceinline__ float lo(float v) { return 0.f; }
    __device__ static __forceinline__ int lane_k(int lane) { return lane >> 4; }
    // A and B both contiguous in k (row-major [row][k])
    __device__ static __forceinline__ f32x4 mma(const float *a, const float *b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], b[0], c, 0, 0, 0);
    }
    __device__ static __forceinline__ AV a_hi(const float *x) { return x[0]; }
    __device__ static __forceinline__ AV a_lo(const float *x) { return 0.f; }
    __device__ static __forceinline__ f32x4 mma_r(AV a, const float *b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b[0], c, 0, 0, 0);
    }
};

template <>
struct Mma<__bf16> {
    static constexpr int KSTEP = 32;
    static constexpr int KPL = 8;
    static constexpr int PAD = 8;  // row stride == 4 (mod 32) dwords: conflict-free 16-byte reads
    using AV = bf16x8;
    using CT = float;
    using ACC = f32x4;
    __device__ static __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
    static constexpr bool kTransposedCopies = true;
    // forward operands as hi + lo BF16 pairs: x w ~ xh wh + xl wh + xh wl (FP32 accumulation)
    static constexpr bool kSplit = true;
    __device__ static __forceinline__ __bf16 cvt(float v) { return to_bf16_rne(v); }
    __device__ static __forceinline__ __bf16 lo(float v) { return to_bf16_rne(v - (float)to_bf16_rne(v)); }
    __device__ static __forceinline__ int lane_k(int lane) { return 8 * (lane >> 4); }
    __device__ static __forceinline__ f32x4 mma(const __bf16 *a, const __bf16 *b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const bf16x8 *>(a),
                                                       *reinterpret_cast<const bf16x8 *>(b), c, 0, 0, 0);
    }
    __device__ static __forceinline__ AV a_hi(const float *x) {
        AV v;
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = cvt(x[e]);
        return v;
    }
    __device__ static __forceinline__ AV a_lo(const float *x) {
        AV v;
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = lo(x[e]);
        return v;
    }
    __device__ static __forceinline__ f32x4 mma_r(AV a, const __bf16 *b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, *reinterpret_cast<const bf16x8 *>(b), c, 0, 0, 0);
    }
};

__device__ __forceinline__ float silu_grad(float z, float *g) {
    const float s = 1.f / (1.f + expf(-z));
    *g = s * (1.f + z * (1.f - s));
    return z * s;
}
__device__ __forceinline__ double silu_grad(double z, double *g) {
    const double s = 1.0 / (1.0 + exp(-z));
    *g = s * (1.0 + z * (1.0 - s));
    return z * s;
}

// Stages NTOT elements with the workgroup's 256 threads: every thread first
// issues the loads of all its ceil(NTOT/256) elements (all in flight at once),
// then converts and stores them, so staging costs one memory latency, not one
// per element.
template <int NTOT, typename V, typename Ld, typename St>
__device__ __forceinline__ void stage(int tid, Ld ld, St st) {
    constexpr int PT = (NTOT + 255) / 256;
    V v[PT];
#pragma unroll
    for (int u = 0; u < PT; ++u) {
        const int e = tid + 256 * u;
        v[u] = e < NTOT ? ld(e) : (V)0;
    }
#pragma unroll
    for (int u = 0; u < PT; ++u) {
        const int e = tid + 256 * u;
        if (e < NTOT) st(e, v[u]);
    }
}

// S <= KP (KP a multiple of 16 and of the instruction k), H1 = 16*C1, H2 = 16*C2.
// Workgroup lb of lg takes row tiles lb, lb + lg, ...
// W1G (FP64 inputs wider than 64): W1 is not staged (beyond 128 inputs it would
// not fit in LDS beside W2 and H; at 128 the staged form holds one workgroup
// per CU and measured 9 % slower) and X is not held in registers: both
// operands of the layer-1 product and W1 in the gX product are read from
// memory (L2) per k-step, eight k-steps of loads issued ahead of their
// products.
template <typename T, int C1, int C2, int KP, bool W1G = false>
__device__ __forceinline__ void mlp2_tiles(const MlpArgs &a, int lb, int lg) {
    using MM = Mma<T>;
    using CT = typename MM::CT;    // epilogue arithmetic (and the weights' source type)
    using ACC = typename MM::ACC;
    constexpr bool F64 = sizeof(CT) == 8;
    constexpr int H1 = 16 * C1, H2 = 16 * C2;
    constexpr int NS = KP / 16;                                          // column tiles of gX
    constexpr int XS = KP + MM::PAD;                                      // row stride of X, W1
    constexpr int HK = (H1 > H2 ? H1 : H2);
    constexpr int HKP = ((HK + MM::KSTEP - 1) / MM::KSTEP) * MM::KSTEP;   // padded hidden k
    constexpr int HS = HKP + MM::PAD;                                     // row stride of H, W2, transposes
    constexpr int H1P = ((H1 + MM::KSTEP - 1) / MM::KSTEP) * MM::KSTEP;  // k extent of GEMMs over H1
    constexpr int H2P = ((H2 + MM::KSTEP - 1) / MM::KSTEP) * MM::KSTEP;  // k extent of the GEMM over H2
    constexpr bool TC = MM::kTransposedCopies;
    constexpr bool SP = MM::kSplit;
    static_assert(KP % MM::KSTEP == 0 && KP % 16 == 0, "padded input dim");
    static_assert(!W1G || (std::is_same<T, double>::value && KP % 32 == 0), "W1G: FP64, KP % 32 == 0");
    constexpr int NX = W1G ? 1 : KP / MM::KSTEP * MM::KPL;               // X values per lane (layer-1 A operand)
    __shared__ __align__(16) T W1s[W1G ? 1 : H1 * XS];       // W1[j][k]
    __shared__ __align__(16) T W2s[H2 * HS];                 // W2[j2][j1]
    __shared__ __align__(16) T W1t[TC ? KP * HS : 1];        // W1[j][k] stored [k][j]
    __shared__ __align__(16) T W2t[TC ? H1 * HS : 1];        // W2[j2][j1] stored [j1][j2]
    __shared__ __align__(16) T Hs[64 * HS];
    __shared__ __align__(16) T W1l[SP ? H1 * XS : 1];        // low halves of the forward operands
    __shared__ __align__(16) T W2l[SP ? H2 * HS : 1];
    __shared__ __align__(16) T Hl[SP ? 64 * HS : 1];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int lr = lane & 15;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int S = a.S;
    const int lk = MM::lane_k(lane);
    const CT *w_out_in, *w_in_out, *bias;
    if constexpr (F64) {
        w_out_in = a.w_out_in64;
        w_in_out = a.w_in_out64;
        bias = a.bias64;
    } else {
        w_out_in = a.w_out_in;
        w_in_out = a.w_in_out;
        bias = a.bias;
    }
    const CT *__restrict__ W1g = w_out_in;                     // [H1][S]
    const CT *__restrict__ W2g = w_out_in + (size_t)S * H1;    // [H2][H1]
    const CT *__restrict__ b1 = bias;
    const CT *__restrict__ b2 = bias + H1;
    const CT bo = bias[H1 + H2];
    const CT *__restrict__ wo = w_in_out + (size_t)S * H1 + (size_t)H1 * H2;

    // X of a tile straight from memory into the layer-1 A operand: value u of
    // this lane is X[row 16 wave + lr][(u / KPL) * KSTEP + lk + u % KPL]
    CT xr[NX];
    const CT *xsrc;
    if constexpr (F64) xsrc = a.x64;
    else xsrc = a.x;
    auto load_x = [&](int t) {
        if constexpr (W1G) return;
        const int i = t * 64 + 16 * wave + lr;
        const bool ok = i < a.nrows;
        const CT *__restrict__ xp = xsrc + (size_t)(ok ? i : 0) * S;
#pragma unroll
        for (int u = 0; u < NX; ++u) {
            const int k = (u / MM::KPL) * MM::KSTEP + lk + u % MM::KPL;
            xr[u] = (ok && k < S) ? xp[k] : (CT)0;
        }
    };
    if (lb * 64 < a.nrows) load_x(lb);  // in flight while the weights are staged
    if constexpr (!W1G)
        stage<H1 * XS, CT>(
            tid, [&](int e) { const int j = e / XS, k = e % XS; return k < S ? W1g[(size_t)j * S + k] : (CT)0; },
            [&](int e, CT w) {
                W1s[e] = MM::cvt(w);
                if constexpr (SP) W1l[e] = MM::lo(w);
            });
    stage<H2 * HS, CT>(
        tid, [&](int e) { const int j = e / HS, k = e % HS; return k < H1 ? W2g[(size_t)j * H1 + k] : (CT)0; },
        [&](int e, CT w) {
            W2s[e] = MM::cvt(w);
            if constexpr (SP) W2l[e] = MM::lo(w);
        });
    if constexpr (TC) {
        stage<KP * HS, CT>(
            tid,
            [&](int e) {
                const int k = e / HS, j = e % HS;
                return (k < S && j < H1) ? W1g[(size_t)j * S + k] : (CT)0;
            },
            [&](int e, CT w) { W1t[e] = MM::cvt(w); });
        stage<H1 * HS, CT>(
            tid,
            [&](int e) {
                const int j1 = e / HS, j2 = e % HS;
                return j2 < H2 ? W2g[(size_t)j2 * H1 + j1] : (CT)0;
            },
            [&](int e, CT w) { W2t[e] = MM::cvt(w); });
    }
    __syncthreads();  // the weights are staged; below, each wave touches only its own rows of H
    const int r0 = 16 * wave;
    T *hrow_w = Hs + r0 * HS;                         // this wave's 16 rows of H
    const T *hrow = Hs + (r0 + lr) * HS + lk;         // A-operand pointer into H
    // per-lane constants of the output columns this lane holds
    CT b1r[C1], b2r[C2], wor[C2], gscale[W1G ? 1 : NS];
#pragma unroll
    for (int c = 0; c < C1; ++c) b1r[c] = b1[16 * c + lr];
#pragma unroll
    for (int c = 0; c < C2; ++c) {
        b2r[c] = b2[16 * c + lr];
        wor[c] = wo[16 * c + lr];
    }
#pragma unroll
    for (int cs = 0; cs < (W1G ? 0 : NS); ++cs) {
        const int n = 16 * cs + lr;
        // G = dE/dx / sigma, times 2 4pi/(2l+1) for the L2 power spectrum's backward
        if constexpr (F64) {
            const double l2 = a.l2f ? (n < S ? a.l2f[n] : 0.0)
                                    : (a.radial ? 1.0 : 2.0 * 12.566370614359172954 / (double)(2 * (n % a.L1) + 1));
            gscale[cs] = n < S ? l2 / a.sig64[n] : 0.0;
        } else {
            const float l2 = a.l2f ? (n < S ? (float)a.l2f[n] : 0.f)
                                   : (a.radial ? 1.f : 2.f * kPi4m / (float)(2 * (n % a.L1) + 1));
            gscale[cs] = n < S ? l2 / a.sig[n] : 0.f;
        }
    }
    for (int tile = lb; tile * 64 < a.nrows; tile += lg) {
        const int row0 = tile * 64;
        const int nt = min(64, a.nrows - row0);
        // atoms of the rows whose energies this lane writes (loaded before the next tile's X)
        int atom_r[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = r0 + MM::row(lane, r);
            atom_r[r] = (lr == 0 && i < nt) ? a.rows_atom[row0 + i] : 0;
        }
        // ---- layer 1 ----
        ACC acc1[C1];
#pragma unroll
        for (int c = 0; c < C1; ++c) acc1[c] = ACC{};
        if constexpr (W1G) {
            // A[i][k] = X[row r0 + i][k] from lane i + 16 k', B[k][j] = W1[j][k]; padding
            // columns (k >= S) read column S - 1 against a zero X
            const int i = row0 + r0 + lr;
            const CT *__restrict__ xp = a.x64 + (size_t)(i < a.nrows ? i : 0) * S;
            const bool xok = i < a.nrows;
#pragma unroll 1
            for (int kb = 0; kb < KP; kb += 8 * MM::KSTEP) {
                CT xv[8], wv[8][C1];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int k = kb + u * MM::KSTEP + lk;
                    const int kc = k < S ? k : S - 1;
                    xv[u] = (xok && k < S) ? xp[kc] : (CT)0;
#pragma unroll
                    for (int c = 0; c < C1; ++c) wv[u][c] = W1g[(size_t)(16 * c + lr) * S + kc];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
#pragma unroll
                    for (int c = 0; c < C1; ++c) acc1[c] = MM::mma_r(xv[u], &wv[u][c], acc1[c]);
            }
        }
#pragma unroll
        for (int ks = 0; ks < (W1G ? 0 : KP / MM::KSTEP); ++ks) {
            const int k0 = ks * MM::KSTEP;
            const typename MM::AV xh = MM::a_hi(xr + ks * MM::KPL);
            typename MM::AV xl{};
            if constexpr (SP) xl = MM::a_lo(xr + ks * MM::KPL);
#pragma unroll
            for (int c = 0; c < C1; ++c) {
                const int wofs = (16 * c + lr) * XS + lk + k0;
                acc1[c] = MM::mma_r(xh, W1s + wofs, acc1[c]);
                if constexpr (SP) {
                    acc1[c] = MM::mma_r(xl, W1s + wofs, acc1[c]);
                    acc1[c] = MM::mma_r(xh, W1l + wofs, acc1[c]);
                }
            }
        }
        // the next tile's X is loaded while this tile's other three products run
        if ((tile + lg) * 64 < a.nrows) load_x(tile + lg);
        CT d1[C1][4];
#pragma unroll
        for (int c = 0; c < C1; ++c) {
            const int j = 16 * c + lr;
            const CT bj = b1r[c];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const CT h = silu_grad((CT)(acc1[c][r] + bj), &d1[c][r]);
                hrow_w[MM::row(lane, r) * HS + j] = MM::cvt(h);
                if constexpr (SP) Hl[(r0 + MM::row(lane, r)) * HS + j] = MM::lo(h);
            }
        }
        if constexpr (H1P > H1) {
            for (int e = lane; e < 16 * (H1P - H1); e += 64) {
                hrow_w[(e / (H1P - H1)) * HS + H1 + e % (H1P - H1)] = MM::cvt(0.f);
                if constexpr (SP) Hl[(r0 + e / (H1P - H1)) * HS + H1 + e % (H1P - H1)] = MM::cvt(0.f);
            }
        }
        wave_lds_sync();
        // ---- layer 2 ----
        ACC acc2[C2];
#pragma unroll
        for (int c = 0; c < C2; ++c) acc2[c] = ACC{};
        for (int k0 = 0; k0 < H1P; k0 += MM::KSTEP) {
#pragma unroll
            for (int c = 0; c < C2; ++c) {
                const int wofs = (16 * c + lr) * HS + lk + k0;
                acc2[c] = MM::mma(hrow + k0, W2s + wofs, acc2[c]);
                if constexpr (SP) {
                    const int hofs = (r0 + lr) * HS + lk + k0;
                    acc2[c] = MM::mma(Hl + hofs, W2s + wofs, acc2[c]);
                    acc2[c] = MM::mma(Hs + hofs, W2l + wofs, acc2[c]);
                }
            }
        }
        CT gz2[C2][4];
        CT ysum[4] = {(CT)0, (CT)0, (CT)0, (CT)0};
        const CT sig_e = (CT)a.sig_e;
#pragma unroll
        for (int c = 0; c < C2; ++c) {
            const CT bj = b2r[c], w = wor[c];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                CT dd;
                const CT h = silu_grad((CT)(acc2[c][r] + bj), &dd);
                ysum[r] += w * h;
                gz2[c][r] = sig_e * w * dd;
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            CT v = ysum[r];
            v += __shfl_xor(v, 1);
            v += __shfl_xor(v, 2);
            v += __shfl_xor(v, 4);
            v += __shfl_xor(v, 8);
            ysum[r] = v;
        }
        if (lr == 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = r0 + MM::row(lane, r);
                if (i < nt) a.E[atom_r[r]] = (double)(ysum[r] + bo) * a.sig_e + a.mu_e;
            }
        }
        if (!a.want_grad) continue;
        // ---- gZ1 = (gZ2 W2) * s1'  (k = j2 over H2) ----
        wave_lds_sync();
#pragma unroll
        for (int c = 0; c < C2; ++c) {
            const int j = 16 * c + lr;
#pragma unroll
            for (int r = 0; r < 4; ++r) hrow_w[MM::row(lane, r) * HS + j] = MM::cvt(gz2[c][r]);
        }
        if constexpr (H2P > H2) {
            for (int e = lane; e < 16 * (H2P - H2); e += 64)
                hrow_w[(e / (H2P - H2)) * HS + H2 + e % (H2P - H2)] = MM::cvt(0.f);
        }
        wave_lds_sync();
        ACC acc3[C1];
#pragma unroll
        for (int c = 0; c < C1; ++c) acc3[c] = ACC{};
        for (int k0 = 0; k0 < H2P; k0 += MM::KSTEP) {
#pragma unroll
            for (int c = 0; c < C1; ++c) {
                if constexpr (TC) acc3[c] = MM::mma(hrow + k0, W2t + (16 * c + lr) * HS + lk + k0, acc3[c]);
                else acc3[c] = MM::mma(hrow + k0, W2s + (k0 + lk) * HS + 16 * c + lr, acc3[c]);
            }
        }
        wave_lds_sync();
#pragma unroll
        for (int c = 0; c < C1; ++c) {
            const int j = 16 * c + lr;
#pragma unroll
            for (int r = 0; r < 4; ++r) hrow_w[MM::row(lane, r) * HS + j] = MM::cvt((CT)acc3[c][r] * d1[c][r]);
        }
        if constexpr (H1P > H1) {
            for (int e = lane; e < 16 * (H1P - H1); e += 64)
                hrow_w[(e / (H1P - H1)) * HS + H1 + e % (H1P - H1)] = MM::cvt(0.f);
        }
        wave_lds_sync();
        // ---- gX = gZ1 W1 (k = j over H1) ----
        if constexpr (W1G) {  // B[k][j] = W1[k][16 cs + j] from memory (columns past S: column S - 1)
            static_assert(H1P % (8 * MM::KSTEP) == 0 || !W1G, "W1G: H1 a multiple of 32");
#pragma unroll 1
            for (int cs = 0; cs < NS; ++cs) {
                ACC acc = ACC{};
                const int n = 16 * cs + lr;
                const int nc = n < S ? n : S - 1;
#pragma unroll 1
                for (int kb = 0; kb < H1P; kb += 8 * MM::KSTEP) {
                    CT wv[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) wv[u] = W1g[(size_t)(kb + u * MM::KSTEP + lk) * S + nc];
#pragma unroll
                    for (int u = 0; u < 8; ++u) acc = MM::mma(hrow + kb + u * MM::KSTEP, &wv[u], acc);
                }
                if (n < S) {
                    const CT scale = (a.l2f ? a.l2f[n]
                                            : (a.radial ? 1.0 : 2.0 * 12.566370614359172954 / (double)(2 * (n % a.L1) + 1))) /
                                     a.sig64[n];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int i = r0 + MM::row(lane, r);
                        if (i < nt) a.G[(size_t)(row0 + i) * S + n] = (float)((CT)acc[r] * scale);
                    }
                }
            }
            continue;
        }
#pragma unroll
        for (int cs = 0; cs < NS; ++cs) {
            ACC acc = ACC{};
            const int n = 16 * cs + lr;
            {
                for (int k0 = 0; k0 < H1P; k0 += MM::KSTEP) {
                    if constexpr (TC) acc = MM::mma(hrow + k0, W1t + n * HS + lk + k0, acc);
                    else acc = MM::mma(hrow + k0, W1s + (k0 + lk) * XS + n, acc);
                }
            }
            if (n < S) {
                const CT scale = gscale[cs];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = r0 + MM::row(lane, r);
                    if (i < nt) a.G[(size_t)(row0 + i) * S + n] = (float)((CT)acc[r] * scale);
                }
            }
        }
    }
}

template <typename T, int C1, int C2, int KP, bool W1G>
__global__ __launch_bounds__(256) void k_mlp2_mfma(MlpArgs a) {
    mlp2_tiles<T, C1, C2, KP, W1G>(a, blockIdx.x, gridDim.x);
}
// one launch for all species: workgroups wg0[s] .. wg0[s+1]-1 serve species s
template <typename T, int C1, int C2, int KP, bool W1G>
__global__ __launch_bounds__(256) void k_mlp2_mfma_multi(MlpMulti m) {
    const int b = blockIdx.x;
    int sp = 0;
    while (sp + 1 < m.nsp && b >= m.wg0[sp + 1]) ++sp;
    mlp2_tiles<T, C1, C2, KP, W1G>(m.a[sp], b - m.wg0[sp], m.wg0[sp + 1] - m.wg0[sp]);
}

// ---- dispatch ----
template <typename T, int C1, int C2, int KP>
static void run_mlp2(hipStream_t s, const MlpArgs &a) {
    const int ncu = device_cu_count();
    const int per = a.wg_per_cu > 0 ? a.wg_per_cu : 2;
    const int tiles = (a.nrows + 63) / 64;
    const int grid = tiles < ncu * per ? tiles : ncu * per;
    constexpr bool W1G = std::is_same<T, double>::value && KP > 64;
    hipLaunchKernelGGL((k_mlp2_mfma<T, C1, C2, KP, W1G>), dim3(grid), dim3(256), 0, s, a);
}

template <typename T>
static bool dispatch_mlp2(hipStream_t s, const MlpArgs &a) {
    if (a.nlayers != 2) return false;
    const int h1 = a.hidden[0], h2 = a.hidden[1];
    const int kp = a.S <= 64 ? 64 : (a.S <= 128 ? 128 : (a.S <= 256 ? 256 : (a.S <= 384 ? 384 : 0)));
#define JNN_MLP_SHAPE(H1_, H2_, KP_)                    \
    if (h1 == H1_ && h2 == H2_ && kp == KP_) {          \
        run_mlp2<T, H1_ / 16, H2_ / 16, KP_>(s, a);     \
        return true;                                    \
    }
    JNN_MLP_SHAPE(32, 32, 64)
    JNN_MLP_SHAPE(32, 32, 128)
    JNN_MLP_SHAPE(64, 64, 64)
    JNN_MLP_SHAPE(64, 64, 128)
    if constexpr (!std::is_same<T, __bf16>::value) {  // wider inputs: FP32, and FP64 with W1 and X read
        JNN_MLP_SHAPE(32, 32, 256)                     // from memory (the BF16 transposed copies would not
        JNN_MLP_SHAPE(32, 32, 384)                     // fit in LDS)
        JNN_MLP_SHAPE(64, 64, 256)
        JNN_MLP_SHAPE(64, 64, 384)
    }
#undef JNN_MLP_SHAPE
    return false;
}

template <typename T, int C1, int C2, int KP>
static void run_mlp2_multi(hipStream_t s, const MlpMulti &m) {
    constexpr bool W1G = std::is_same<T, double>::value && KP > 64;
    hipLaunchKernelGGL((k_mlp2_mfma_multi<T, C1, C2, KP, W1G>), dim3(m.wg0[m.nsp]), dim3(256), 0, s, m);
}
template <typename T>
static bool dispatch_mlp2_multi(hipStream_t s, const MlpMulti &m) {
    const MlpArgs &a = m.a[0];
    const int h1 = a.hidden[0], h2 = a.hidden[1];
    const int kp = a.S <= 64 ? 64 : (a.S <= 128 ? 128 : (a.S <= 256 ? 256 : (a.S <= 384 ? 384 : 0)));
#define JNN_MLP_SHAPE(H1_, H2_, KP_)                        \
    if (h1 == H1_ && h2 == H2_ && kp == KP_) {              \
        run_mlp2_multi<T, H1_ / 16, H2_ / 16, KP_>(s, m);   \
        return true;                                        \
    }
    JNN_MLP_SHAPE(32, 32, 64)
    JNN_MLP_SHAPE(32, 32, 128)
    JNN_MLP_SHAPE(64, 64, 64)
    JNN_MLP_SHAPE(64, 64, 128)
    if constexpr (!std::is_same<T, __bf16>::value) {
        JNN_MLP_SHAPE(32, 32, 256)
        JNN_MLP_SHAPE(32, 32, 384)
        JNN_MLP_SHAPE(64, 64, 256)
        JNN_MLP_SHAPE(64, 64, 384)
    }
#undef JNN_MLP_SHAPE
    return false;
}

bool launch_mlp_mfma_multi(hipStream_t s, const MlpArgs *a, int nsp, int mode) {
    if (nsp <= 0 || nsp > kMlpMaxSpecies) return false;
    MlpMulti m{};
    int tiles_tot = 0, first = -1;
    for (int t = 0; t < nsp; ++t) {
        if (a[t].nrows <= 0) continue;
        if (first < 0) first = t;
        const MlpArgs &f = a[first];
        if (a[t].nlayers != 2 || a[t].S != f.S || a[t].hidden[0] != f.hidden[0] || a[t].hidden[1] != f.hidden[1])
            return false;
        tiles_tot += (a[t].nrows + 63) / 64;
    }
    if (first < 0) return true;  // no rows at all
    const int per = a[first].wg_per_cu > 0 ? a[first].wg_per_cu : 2;
    const int total = device_cu_count() * per;
    // workgroups in proportion to each species' tiles, at least one, at most its tiles
    int wg = 0;
    for (int t = 0; t < nsp; ++t) {
        if (a[t].nrows <= 0) continue;
        const int tiles = (a[t].nrows + 63) / 64;
        int w = (int)(((long long)total * tiles + tiles_tot - 1) / tiles_tot);
        w = w < 1 ? 1 : (w > tiles ? tiles : w);
        m.a[m.nsp] = a[t];
        m.wg0[m.nsp] = wg;
        ++m.nsp;
        wg += w;
    }
    m.wg0[m.nsp] = wg;
    if (mode == kMlpBf16) return dispatch_mlp2_multi<__bf16>(s, m);
    if (mode == kMlpFp64) return dispatch_mlp2_multi<double>(s, m);
    return dispatch_mlp2_multi<float>(s, m);
}

bool launch_mlp_mfma_fp32(hipStream_t s, const MlpArgs &a) {
    if (a.nrows <= 0) return true;
    return dispatch_mlp2<float>(s, a);
}
bool launch_mlp_mfma_bf16(hipStream_t s, const MlpArgs &a) {
    if (a.nrows <= 0) return true;
    return dispatch_mlp2<__bf16>(s, a);
}
bool launch_mlp_mfma_fp64(hipStream_t s, const MlpArgs &a) {
    if (a.nrows <= 0) return true;
    return dispatch_mlp2<double>(s, a);
}

}  // namespace jnn
