// Per-species feed-forward network, forward and input-gradient backward, over
// tiles of 64 atoms of one species (reference: nn_FeedForward.hpp:8-21,131-261;
// normalisation and energy scaling nnap_main.hpp:267-323).
//
//   h_1 = silu(W_1 x + b_1), ..., y = w_o . h_L + b_o,  E_i = y sigma_E + (mu_E + ref_E)
//   dE/dx = sigma_E W_1^T (s_1' * W_2^T (s_2' * w_o)),  dE/dfp = dE/dx / sigma
//
// Layout: one workgroup of 4 wavefronts per tile of TR atoms (64, or 32 when
// 64 rows would not fit in LDS); a lane takes atom lane % TR; activations and
// SiLU derivatives live in LDS as [unit][TR]; each wave owns blocks of 8
// output units (with TR = 32 its two halves take two blocks) and reads the
// weights at uniform addresses, so every LDS read feeds 8 FMAs.  The output is Gc = 2*4pi/(2l+1) dE/dfp in
// the [row][S] layout the descriptor backward reads.  T = float (the
// FP32-geometry mode) or double (the plain mode: FP64 inputs, weights,
// activations and sums).
#include "common.h"
#include "kernels.h"

namespace jnn {

constexpr int kMlpTile = 64;  // the largest tile (rows)
constexpr int kUB = 8;
constexpr float kPi4f = 12.566370614359172f;

__device__ __forceinline__ float silu_and_grad(float z, float *g) {
    const float s = 1.f / (1.f + expf(-z));
    *g = s * (1.f + z * (1.f - s));
    return z * s;
}
__device__ __forceinline__ double silu_and_grad(double z, double *g) {
    const double s = 1.0 / (1.0 + exp(-z));
    *g = s * (1.0 + z * (1.0 - s));
    return z * s;
}
// the kernel's inputs in its element type
__device__ __forceinline__ void mlp_src(const MlpArgs &a, const float *&x, const float *&wf, const float *&wb,
                                        const float *&bs) {
    x = a.x; wf = a.w_in_out; wb = a.w_out_in; bs = a.bias;
}
__device__ __forceinline__ void mlp_src(const MlpArgs &a, const double *&x, const double *&wf, const double *&wb,
                                        const double *&bs) {
    x = a.x64; wf = a.w_in_out64; wb = a.w_out_in64; bs = a.bias64;
}

// LDS offset (values) of the activations of layer l (l = 0 is the input x), tiles of TR rows
__device__ __forceinline__ int act_off(const MlpArgs &a, int l, int TR) {
    int off = a.S * TR;
    if (l == 0) return 0;
    for (int k = 0; k < l - 1; ++k) off += 2 * a.hidden[k] * TR;
    return off;
}
// LDS offset of the SiLU derivative of hidden layer l >= 1
__device__ __forceinline__ int der_off(const MlpArgs &a, int l, int TR) { return act_off(a, l, TR) + a.hidden[l - 1] * TR; }

template <typename T, int TR>
__global__ __launch_bounds__(256) void k_mlp_vec(MlpArgs a) {
    extern __shared__ __align__(16) double mlp_smem[];
    T *smem = reinterpret_cast<T *>(mlp_smem);
    constexpr int HV = 64 / TR;  // unit blocks a wave takes at once
    const int lane = threadIdx.x & 63 & (TR - 1);  // the lane's row of the tile
    const int half = (threadIdx.x & 63) / TR;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ub0 = wave * HV + half;
    const int S = a.S, nl = a.nlayers;
    const T *ax, *wf0, *wb0, *bs0;
    mlp_src(a, ax, wf0, wb0, bs0);
    for (int tile = blockIdx.x; tile * TR < a.nrows; tile += gridDim.x) {
        const int row0 = tile * TR;
        const int nt = min(TR, a.nrows - row0);
        for (int idx = threadIdx.x; idx < TR * S; idx += blockDim.x) {
            const int r = idx / S, e = idx % S;
            smem[e * TR + r] = (r < nt) ? ax[(size_t)(row0 + r) * S + e] : (T)0;
        }
        __syncthreads();
        const T *wf = wf0;
        const T *bs = bs0;
        int in = S;
        int ain = 0;  // LDS offset of the layer input
        for (int l = 0; l < nl; ++l) {
            const int out = a.hidden[l];
            const int aout = act_off(a, l + 1, TR), dout = aout + out * TR;
            for (int ub = ub0; ub * kUB < out; ub += 4 * HV) {
                const int o0 = ub * kUB;
                T acc[kUB];
#pragma unroll
                for (int u = 0; u < kUB; ++u) acc[u] = (o0 + u < out) ? bs[o0 + u] : (T)0;
                for (int i = 0; i < in; ++i) {
                    const T v = smem[ain + i * TR + lane];
                    const T *w = wf + (size_t)i * out + o0;
#pragma unroll
                    for (int u = 0; u < kUB; ++u)
                        if (o0 + u < out) acc[u] += w[u] * v;
                }
#pragma unroll
                for (int u = 0; u < kUB; ++u)
                    if (o0 + u < out) {
                        T gd;
                        smem[aout + (o0 + u) * TR + lane] = silu_and_grad(acc[u], &gd);
                        smem[dout + (o0 + u) * TR + lane] = gd;
                    }
            }
            __syncthreads();
            wf += (size_t)in * out;
            bs += out;
            in = out;
            ain = aout;
        }
        // wf now points at the output weight, bs at the output bias
        const T *wo = wf;
        if (wave == 0 && half == 0) {
            T dsum = (T)0;
            for (int i = 0; i < in; ++i) dsum += wo[i] * smem[ain + i * TR + lane];
            const T y = dsum + bs[0];
            if (lane < nt) a.E[a.rows_atom[row0 + lane]] = (double)y * a.sig_e + a.mu_e;
        }
        __syncthreads();
        if (a.want_grad) {
            // gradient at the last hidden pre-activation, written over its activations
            const int dl = der_off(a, nl, TR);
            for (int idx = threadIdx.x; idx < in * TR; idx += blockDim.x) {
                const int u = idx / TR;
                smem[ain + idx] = (T)a.sig_e * wo[u] * smem[dl + idx];
            }
            __syncthreads();
            // per-layer [out][in] weights, walked backwards
            size_t woff = 0;
            {
                int c = S;
                for (int l = 0; l < nl; ++l) { woff += (size_t)c * a.hidden[l]; c = a.hidden[l]; }
            }
            for (int l = nl - 1; l >= 0; --l) {
                const int out = a.hidden[l];
                const int inl = (l == 0) ? S : a.hidden[l - 1];
                woff -= (size_t)inl * out;
                const T *wb = wb0 + woff;
                const int gout = act_off(a, l + 1, TR);
                const int gin = act_off(a, l, TR);
                const int dprev = (l > 0) ? der_off(a, l, TR) : 0;
                for (int ub = ub0; ub * kUB < inl; ub += 4 * HV) {
                    const int i0 = ub * kUB;
                    T acc[kUB];
#pragma unroll
                    for (int u = 0; u < kUB; ++u) acc[u] = (T)0;
                    for (int o = 0; o < out; ++o) {
                        const T v = smem[gout + o * TR + lane];
                        const T *w = wb + (size_t)o * inl + i0;
#pragma unroll
                        for (int u = 0; u < kUB; ++u)
                            if (i0 + u < inl) acc[u] += w[u] * v;
                    }
#pragma unroll
                    for (int u = 0; u < kUB; ++u)
                        if (i0 + u < inl) {
                            const T gd = (l > 0) ? smem[dprev + (i0 + u) * TR + lane] : (T)1;
                            smem[gin + (i0 + u) * TR + lane] = acc[u] * gd;
                        }
                }
                __syncthreads();
            }
            for (int idx = threadIdx.x; idx < TR * S; idx += blockDim.x) {
                const int r = idx / S, e = idx % S;
                if (r < nt) {
                    const int l = e % a.L1;
                    // times 2 4pi/(2l+1) for the L2 power spectrum's backward (not for a radial basis)
                    if constexpr (sizeof(T) == 8) {
                        const double l2 = a.l2f ? a.l2f[e] : (a.radial ? 1.0 : 2.0 * 12.566370614359172954 / (double)(2 * l + 1));
                        a.G[(size_t)(row0 + r) * S + e] = (float)(smem[e * TR + r] / a.sig64[e] * l2);
                    } else {
                        const float l2 = a.l2f ? (float)a.l2f[e] : (a.radial ? 1.f : 2.f * kPi4f / (float)(2 * l + 1));
                        a.G[(size_t)(row0 + r) * S + e] = smem[e * TR + r] / a.sig[e] * l2;
                    }
                }
            }
        }
        __syncthreads();
    }
}

static size_t mlp_lds_bytes(const MlpArgs &a, size_t elem, int rows) {
    size_t f = (size_t)a.S * rows;
    for (int l = 0; l < a.nlayers; ++l) f += 2 * (size_t)a.hidden[l] * rows;
    return f * elem;
}

template <typename T, int TR>
static void run_mlp_vec(hipStream_t s, const MlpArgs &a, size_t lds) {
    static const bool attr_set = [] {  // once per process (thread-safe static initialisation)
        JNN_HIP(hipFuncSetAttribute((const void *)k_mlp_vec<T, TR>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    160 * 1024));
        return true;
    }();
    (void)attr_set;
    const int tiles = (a.nrows + TR - 1) / TR;
    hipLaunchKernelGGL((k_mlp_vec<T, TR>), dim3(tiles), dim3(256), lds, s, a);
}
// tiles of 64 rows, or of 32 when 64 would not fit in LDS
template <typename T>
static bool launch_mlp_vec(hipStream_t s, const MlpArgs &a) {
    if (a.nrows <= 0) return true;
    constexpr size_t cap = 160 * 1024;
    if (mlp_lds_bytes(a, sizeof(T), 64) <= cap) run_mlp_vec<T, 64>(s, a, mlp_lds_bytes(a, sizeof(T), 64));
    else if (mlp_lds_bytes(a, sizeof(T), 32) <= cap) run_mlp_vec<T, 32>(s, a, mlp_lds_bytes(a, sizeof(T), 32));
    else return false;
    return true;
}
bool launch_mlp_fp32(hipStream_t s, const MlpArgs &a) { return launch_mlp_vec<float>(s, a); }
bool launch_mlp_fp64(hipStream_t s, const MlpArgs &a) { return launch_mlp_vec<double>(s, a); }

}  // namespace jnn
