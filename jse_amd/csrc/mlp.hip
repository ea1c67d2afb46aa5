// Per-species feed-forward network, forward and input-gradient backward, over
// tiles of 64 atoms of one species (reference: nn_FeedForward.hpp:8-21,131-261;
// normalisation and energy scaling nnap_main.hpp:267-323).
//
//   h_1 = silu(W_1 x + b_1), ..., y = w_o . h_L + b_o,  E_i = y sigma_E + (mu_E + ref_E)
//   dE/dx = sigma_E W_1^T (s_1' * W_2^T (s_2' * w_o)),  dE/dfp = dE/dx / sigma
//
// Layout: one workgroup of 4 wavefronts per tile, lane = atom; activations and
// SiLU derivatives live in LDS as [unit][64]; each wave owns blocks of 8
// output units and reads the weights at wave-uniform addresses (scalar loads),
// so every LDS read feeds 8 FMAs.  The output is Gc = 2*4pi/(2l+1) dE/dfp in
// the [row][S] layout the descriptor backward reads.
#include "common.h"
#include "kernels.h"

namespace jnn {

constexpr int kMlpTile = 64;
constexpr int kUB = 8;
constexpr float kPi4f = 12.566370614359172f;

__device__ __forceinline__ float silu_and_grad(float z, float *g) {
    const float s = 1.f / (1.f + expf(-z));
    *g = s * (1.f + z * (1.f - s));
    return z * s;
}

// LDS offset (floats) of the activations of layer l (l = 0 is the input x)
__device__ __forceinline__ int act_off(const MlpArgs &a, int l) {
    int off = a.S * kMlpTile;
    if (l == 0) return 0;
    for (int k = 0; k < l - 1; ++k) off += 2 * a.hidden[k] * kMlpTile;
    return off;
}
// LDS offset of the SiLU derivative of hidden layer l >= 1
__device__ __forceinline__ int der_off(const MlpArgs &a, int l) { return act_off(a, l) + a.hidden[l - 1] * kMlpTile; }

__global__ __launch_bounds__(256) void k_mlp_fp32(MlpArgs a) {
    extern __shared__ float smem[];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int S = a.S, nl = a.nlayers;
    for (int tile = blockIdx.x; tile * kMlpTile < a.nrows; tile += gridDim.x) {
        const int row0 = tile * kMlpTile;
        const int nt = min(kMlpTile, a.nrows - row0);
        for (int idx = threadIdx.x; idx < kMlpTile * S; idx += blockDim.x) {
            const int r = idx / S, e = idx % S;
            smem[e * kMlpTile + r] = (r < nt) ? a.x[(size_t)(row0 + r) * S + e] : 0.f;
        }
        __syncthreads();
        const float *wf = a.w_in_out;
        const float *bs = a.bias;
        int in = S;
        int ain = 0;  // LDS offset of the layer input
        for (int l = 0; l < nl; ++l) {
            const int out = a.hidden[l];
            const int aout = act_off(a, l + 1), dout = aout + out * kMlpTile;
            for (int ub = wave; ub * kUB < out; ub += 4) {
                const int o0 = ub * kUB;
                float acc[kUB];
#pragma unroll
                for (int u = 0; u < kUB; ++u) acc[u] = (o0 + u < out) ? bs[o0 + u] : 0.f;
                for (int i = 0; i < in; ++i) {
                    const float v = smem[ain + i * kMlpTile + lane];
                    const float *w = wf + (size_t)i * out + o0;
#pragma unroll
                    for (int u = 0; u < kUB; ++u)
                        if (o0 + u < out) acc[u] += w[u] * v;
                }
#pragma unroll
                for (int u = 0; u < kUB; ++u)
                    if (o0 + u < out) {
                        float gd;
                        smem[aout + (o0 + u) * kMlpTile + lane] = silu_and_grad(acc[u], &gd);
                        smem[dout + (o0 + u) * kMlpTile + lane] = gd;
                    }
            }
            __syncthreads();
            wf += (size_t)in * out;
            bs += out;
            in = out;
            ain = aout;
        }
        // wf now points at the output weight, bs at the output bias
        const float *wo = wf;
        if (wave == 0) {
            float dsum = 0.f;
            for (int i = 0; i < in; ++i) dsum += wo[i] * smem[ain + i * kMlpTile + lane];
            const float y = dsum + bs[0];
            if (lane < nt) a.E[a.rows_atom[row0 + lane]] = (double)y * (double)a.sig_e + (double)a.mu_e;
        }
        __syncthreads();
        if (a.want_grad) {
            // gradient at the last hidden pre-activation, written over its activations
            const int dl = der_off(a, nl);
            for (int idx = threadIdx.x; idx < in * kMlpTile; idx += blockDim.x) {
                const int u = idx / kMlpTile;
                smem[ain + idx] = a.sig_e * wo[u] * smem[dl + idx];
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
                const float *wb = a.w_out_in + woff;
                const int gout = act_off(a, l + 1);
                const int gin = act_off(a, l);
                const int dprev = (l > 0) ? der_off(a, l) : 0;
                for (int ub = wave; ub * kUB < inl; ub += 4) {
                    const int i0 = ub * kUB;
                    float acc[kUB];
#pragma unroll
                    for (int u = 0; u < kUB; ++u) acc[u] = 0.f;
                    for (int o = 0; o < out; ++o) {
                        const float v = smem[gout + o * kMlpTile + lane];
                        const float *w = wb + (size_t)o * inl + i0;
#pragma unroll
                        for (int u = 0; u < kUB; ++u)
                            if (i0 + u < inl) acc[u] += w[u] * v;
                    }
#pragma unroll
                    for (int u = 0; u < kUB; ++u)
                        if (i0 + u < inl) {
                            const float gd = (l > 0) ? smem[dprev + (i0 + u) * kMlpTile + lane] : 1.f;
                            smem[gin + (i0 + u) * kMlpTile + lane] = acc[u] * gd;
                        }
                }
                __syncthreads();
            }
            for (int idx = threadIdx.x; idx < kMlpTile * S; idx += blockDim.x) {
                const int r = idx / S, e = idx % S;
                if (r < nt) {
                    const int l = e % a.L1;
                    // times 2 4pi/(2l+1) for the L2 power spectrum's backward (not for a radial basis)
                    const float l2 = a.radial ? 1.f : 2.f * kPi4f / (float)(2 * l + 1);
                    a.G[(size_t)(row0 + r) * S + e] = smem[e * kMlpTile + r] / a.sig[e] * l2;
                }
            }
        }
        __syncthreads();
    }
}

static size_t mlp_lds_bytes(const MlpArgs &a) {
    size_t f = (size_t)a.S * kMlpTile;
    for (int l = 0; l < a.nlayers; ++l) f += 2 * (size_t)a.hidden[l] * kMlpTile;
    return f * sizeof(float);
}

bool launch_mlp_fp32(hipStream_t s, const MlpArgs &a) {
    if (a.nrows <= 0) return true;
    const size_t lds = mlp_lds_bytes(a);
    if (lds > 160 * 1024) return false;
    static const bool attr_set = [] {  // once per process (thread-safe static initialisation)
        JNN_HIP(hipFuncSetAttribute((const void *)k_mlp_fp32, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        return true;
    }();
    (void)attr_set;
    const int tiles = (a.nrows + kMlpTile - 1) / kMlpTile;
    hipLaunchKernelGGL(k_mlp_fp32, dim3(tiles), dim3(256), lds, s, a);
    return true;
}

}  // namespace jnn
