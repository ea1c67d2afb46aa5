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

namespace jnn {

// sum_i W[i ws] v[i TR] over i < n in order, one accumulator (the plain
// loop's sum, bit for bit), with the weights loaded from global memory eight
// at a time so that eight loads are in flight; dot_gw2 forms two such sums
// over the same weights.
template <int TR>
__device__ __forceinline__ double dot_gw(const double *__restrict__ W, int ws, const double *v, int n) {
    double z = 0.0;
    int i = 0;
    for (; i + 8 <= n; i += 8) {
        double wv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) wv[u] = W[(size_t)(i + u) * ws];
#pragma unroll
        for (int u = 0; u < 8; ++u) z += wv[u] * v[(i + u) * TR];
    }
    for (; i < n; ++i) z += W[(size_t)i * ws] * v[i * TR];
    return z;
}
template <int TR>
__device__ __forceinline__ void dot_gw2(const double *__restrict__ W, int ws, const double *v1, const double *v2,
                                        int n, double &z1, double &z2) {
    z1 = 0.0;
    z2 = 0.0;
    int i = 0;
    for (; i + 8 <= n; i += 8) {
        double wv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) wv[u] = W[(size_t)(i + u) * ws];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            z1 += wv[u] * v1[(i + u) * TR];
            z2 += wv[u] * v2[(i + u) * TR];
        }
    }
    for (; i < n; ++i) {
        z1 += W[(size_t)i * ws] * v1[i * TR];
        z2 += W[(size_t)i * ws] * v2[i * TR];
    }
}

// ---- training derivative: the network parameters' gradient of sum_i w_i E_i ----
// (reference: jse_nnap_backwardEnergy, nnap_interface.cpp:111-134, through
// normedNnBackward with GRAD_PARAM, nnap_main.hpp:297-323, and the layer
// backward of nn_FeedForward.hpp:131-261).  One workgroup per tile of TR atoms
// at a time, FP64 throughout (16-row tiles: the layer products on the FP64
// matrix cores): the forward is formed again from x (act[l][TR]
// in LDS, with the SiLU derivatives), dL/dz of every layer is taken back from
// dL/dy = w_i sigma_E, and the tile's
//   dW_l[o][i] = sum_r dz_l[o][r] act_{l-1}[i][r] (FP64 matrix cores),  db_l[o] = sum_r dz_l[o][r],
//   dw_o[h] = sum_r dy_r act_L[h][r],  db_o = sum_r dy_r
// are added into this workgroup's row of `partial` (its first tile stores), in
// the packed layout [hidden W per layer [out][in] | output W | hidden b | output b].
// Tiles of a workgroup come in a fixed order, so the sums are deterministic.
template <int TR>
__global__ __launch_bounds__(256) void k_mlp_pgrad(MlpArgs a, const double *__restrict__ w, int nparam,
                                                   double *__restrict__ partial) {
    static_assert(TR <= 32 && (TR & (TR - 1)) == 0, "the tile's rows are a power of two up to 32");
    extern __shared__ __align__(16) double pg_smem[];
    const int S = a.S, L = a.nlayers;
    const double *__restrict__ wb = a.w_out_in64;  // per layer [out][in]
    const double *__restrict__ bs = a.bias64;
    // LDS: act_0 = x [S][TR], then per hidden layer l: act [h][TR], der [h][TR], dz [h][TR]; dy [TR]
    int aoff[kMaxLayers + 1], doff[kMaxLayers + 1], goff[kMaxLayers + 1];
    aoff[0] = 0;
    int off = S * TR;
    for (int l = 1; l <= L; ++l) {
        const int h = a.hidden[l - 1];
        aoff[l] = off;
        doff[l] = off + h * TR;
        goff[l] = off + 2 * h * TR;
        off += 3 * h * TR;
    }
    double *dy = pg_smem + off;
    // packed offsets: hidden W of layer l, output W, hidden b of layer l, output b
    int woff[kMaxLayers + 1], boff[kMaxLayers + 1];
    int pw = 0;
    for (int l = 1; l <= L; ++l) {
        woff[l] = pw;
        pw += a.hidden[l - 1] * (l == 1 ? S : a.hidden[l - 2]);
    }
    const int wo_off = pw;
    const int hL = a.hidden[L - 1];
    int pb = wo_off + hL;
    for (int l = 1; l <= L; ++l) {
        boff[l] = pb;
        pb += a.hidden[l - 1];
    }
    const int bo_off = pb;
    // the output weight follows the hidden weights in the forward layout
    const double *__restrict__ wout = a.w_in_out64 + wo_off;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, nwave = blockDim.x >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
    using d4 = __attribute__((__vector_size__(4 * sizeof(double)))) double;
    static_assert(TR % 4 == 0, "the tile's rows are whole k-steps of the matrix instruction");
    double *__restrict__ mine = partial + (size_t)blockIdx.x * nparam;
    bool first = true;
    for (int tile = blockIdx.x; tile * TR < a.nrows; tile += gridDim.x) {
        const int row0 = tile * TR;
        const int nt = min(TR, a.nrows - row0);
        for (int idx = tid; idx < TR * S; idx += blockDim.x) {
            const int r = idx / S, e = idx % S;
            double v = 0.0;
            if (r < nt) v = a.x64 ? a.x64[(size_t)(row0 + r) * S + e] : (double)a.x[(size_t)(row0 + r) * S + e];
            pg_smem[e * TR + r] = v;
        }
        if (tid < TR) dy[tid] = tid < nt ? w[a.rows_atom[row0 + tid]] * a.sig_e : 0.0;
        __syncthreads();
        // forward
        for (int l = 1; l <= L; ++l) {
            const int h = a.hidden[l - 1], in = l == 1 ? S : a.hidden[l - 2];
            const double *__restrict__ W = wb + woff[l];
            const double *__restrict__ bl = bs + (boff[l] - wo_off - hL);  // hidden biases come first
            if constexpr (TR == 16) {
                // z = W act on the FP64 matrix cores: a 16-row block of outputs per
                // wave, the inputs as k-steps (A[o][k] = W from lane o + 16 k, B[k][r]
                // = act from lane r + 16 k, D[o][r] in lane r + 16 (o % 4), element o / 4)
                const double *ac = pg_smem + aoff[l - 1];
                for (int o0 = 16 * wave; o0 < h; o0 += 16 * nwave) {
                    const bool ov = o0 + li < h;
                    const double *wr = W + (size_t)(ov ? o0 + li : 0) * in;
                    d4 acc = {0.0, 0.0, 0.0, 0.0};
                    for (int k0 = 0; k0 < in; k0 += 4) {
                        const int k = k0 + lk;
                        const double av = ov && k < in ? wr[k] : 0.0;
                        const double bv = k < in ? ac[k * TR + li] : 0.0;
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
                    }
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int o = o0 + 4 * e + lk;
                        if (o < h) {
                            double g;
                            pg_smem[aoff[l] + o * TR + li] = silu_and_grad(acc[e] + bl[o], &g);
                            pg_smem[doff[l] + o * TR + li] = g;
                        }
                    }
                }
            } else {
                for (int idx = tid; idx < h * TR; idx += blockDim.x) {
                    const int o = idx / TR, r = idx % TR;
                    const double z = dot_gw<TR>(W + (size_t)o * in, 1, pg_smem + aoff[l - 1] + r, in) + bl[o];
                    double g;
                    pg_smem[aoff[l] + idx] = silu_and_grad(z, &g);
                    pg_smem[doff[l] + idx] = g;
                }
            }
            __syncthreads();
        }
        // dz of the last hidden layer, then back through the layers
        for (int idx = tid; idx < hL * TR; idx += blockDim.x) {
            const int o = idx / TR, r = idx % TR;
            pg_smem[goff[L] + idx] = dy[r] * wout[o] * pg_smem[doff[L] + idx];
        }
        __syncthreads();
        for (int l = L; l >= 2; --l) {
            const int h = a.hidden[l - 1], in = a.hidden[l - 2];
            const double *__restrict__ W = wb + woff[l];
            if constexpr (TR == 16) {
                // dz_{l-1} = (W^T dz_l) s': a 16-row block of inputs per wave, the
                // outputs as k-steps (A[i][k] = W[k][i], B[k][r] = dz_l)
                const double *dz = pg_smem + goff[l];
                for (int i0 = 16 * wave; i0 < in; i0 += 16 * nwave) {
                    const bool iv = i0 + li < in;
                    d4 acc = {0.0, 0.0, 0.0, 0.0};
                    for (int k0 = 0; k0 < h; k0 += 4) {
                        const int k = k0 + lk;
                        const double av = iv && k < h ? W[(size_t)k * in + i0 + li] : 0.0;
                        const double bv = k < h ? dz[k * TR + li] : 0.0;
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
                    }
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int i = i0 + 4 * e + lk;
                        if (i < in) pg_smem[goff[l - 1] + i * TR + li] = acc[e] * pg_smem[doff[l - 1] + i * TR + li];
                    }
                }
            } else {
                for (int idx = tid; idx < in * TR; idx += blockDim.x) {
                    const int i = idx / TR, r = idx % TR;
                    const double s = dot_gw<TR>(W + i, in, pg_smem + goff[l] + r, h);
                    pg_smem[goff[l - 1] + idx] = s * pg_smem[doff[l - 1] + idx];
                }
            }
            __syncthreads();
        }
        // this tile's parameter gradients into the workgroup's partial
        auto put = [&](int p, double v) { mine[p] = first ? v : mine[p] + v; };
        for (int l = 1; l <= L; ++l) {
            const int h = a.hidden[l - 1], in = l == 1 ? S : a.hidden[l - 2];
            // dW_l = dz_l act_{l-1}^T on the FP64 matrix cores: 16 x 16 blocks of
            // (o, i) taken by the waves round-robin, the tile's TR rows as the
            // k-steps (A[o][r] = dz from lane o + 16 r, B[r][i] = act from lane
            // i + 16 r, D[o][i] in lane i + 16 (o % 4), element o / 4)
            const double *dz = pg_smem + goff[l], *ac = pg_smem + aoff[l - 1];
            const int nto = (h + 15) / 16, nti = (in + 15) / 16;
            for (int tt = wave; tt < nto * nti; tt += nwave) {
                const int o0 = (tt / nti) * 16, i0 = (tt % nti) * 16;
                d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int k0 = 0; k0 < TR; k0 += 4) {
                    const int r = k0 + lk;
                    const double av = o0 + li < h ? dz[(o0 + li) * TR + r] : 0.0;
                    const double bv = i0 + li < in ? ac[(i0 + li) * TR + r] : 0.0;
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int o = o0 + 4 * e + lk, i = i0 + li;
                    if (o < h && i < in) put(woff[l] + o * in + i, acc[e]);
                }
            }
            for (int o = tid; o < h; o += blockDim.x) {
                double s = 0.0;
                for (int r = 0; r < TR; ++r) s += pg_smem[goff[l] + o * TR + r];
                put(boff[l] + o, s);
            }
        }
        for (int o = tid; o < hL; o += blockDim.x) {
            double s = 0.0;
            for (int r = 0; r < TR; ++r) s += dy[r] * pg_smem[aoff[L] + o * TR + r];
            put(wo_off + o, s);
        }
        if (tid == 0) {
            double s = 0.0;
            for (int r = 0; r < TR; ++r) s += dy[r];
            put(bo_off, s);
        }
        first = false;
        __syncthreads();
    }
}
// ---- training derivative of the force / virial loss: the networks' part ----
// (reference: jse_nnap_backwardEnergyForce, nnap_interface.cpp:189-232, with
// normedNnBackwardBackward and normedNnBackward, nnap_main.hpp; the layer
// grad-grad of nn_FeedForward.hpp:264-320).  Per atom the network runs with a
// tangent: xd = D_u x (the descriptor's directional derivative along the
// pairs' upstream directions), z_l = W_l a_{l-1} + b_l, a_l = s(z_l),
// zd_l = W_l ad_{l-1}, ad_l = s'(z_l) zd_l, E = sigma_E (w_o a_L + b_o) + mu_E,
// D_u E = sigma_E w_o ad_L, and the loss of the atom is w_i E - D_u E (the
// pairs' g enter the loss as -D_u E).  Reverse, with abar_L = w_i sigma_E w_o,
// adbar_L = -sigma_E w_o:
//   zdbar_l = adbar_l s'(z_l),  zbar_l = abar_l s'(z_l) + adbar_l zd_l s''(z_l),
//   dW_l += zbar_l a_{l-1}^T + zdbar_l ad_{l-1}^T,  db_l += zbar_l,
//   abar_{l-1} = W_l^T zbar_l,  adbar_{l-1} = W_l^T zdbar_l,
//   dw_o += w_i sigma_E a_L - sigma_E ad_L,  db_o += w_i sigma_E,
// and abar_0 / adbar_0 = dL/dx / dL/dxd go to xbar / xdbar for the mixing
// weights' derivative.  Tiles of TR atoms, FP64, partial rows as k_mlp_pgrad.
template <int TR>
__global__ __launch_bounds__(256) void k_mlp_dual_pgrad(MlpArgs a, const double *__restrict__ w,
                                                        const double *__restrict__ xdot, double *__restrict__ xbar,
                                                        double *__restrict__ xdbar, int nparam,
                                                        double *__restrict__ partial) {
    static_assert(TR <= 32 && (TR & (TR - 1)) == 0, "the tile's rows are a power of two up to 32");
    extern __shared__ __align__(16) double pg_smem[];
    const int S = a.S, L = a.nlayers;
    const double *__restrict__ wb = a.w_out_in64;
    const double *__restrict__ bs = a.bias64;
    // LDS: x, xd [S][TR]; per hidden layer: a, s', s'', zd, ad, zbar, zdbar [h][TR]; ybar, ydbar [TR]
    constexpr int NA = 7;
    int aoff[kMaxLayers + 1];
    aoff[0] = 0;
    int off = 2 * S * TR;
    for (int l = 1; l <= L; ++l) {
        aoff[l] = off;
        off += NA * a.hidden[l - 1] * TR;
    }
    double *yb = pg_smem + off, *ydb = yb + TR;
    auto act = [&](int l) { return l == 0 ? pg_smem : pg_smem + aoff[l]; };
    auto actd = [&](int l) { return l == 0 ? pg_smem + S * TR : pg_smem + aoff[l] + 4 * a.hidden[l - 1] * TR; };
    auto der = [&](int l) { return pg_smem + aoff[l] + a.hidden[l - 1] * TR; };
    auto der2 = [&](int l) { return pg_smem + aoff[l] + 2 * a.hidden[l - 1] * TR; };
    auto zd = [&](int l) { return pg_smem + aoff[l] + 3 * a.hidden[l - 1] * TR; };
    auto zb = [&](int l) { return pg_smem + aoff[l] + 5 * a.hidden[l - 1] * TR; };
    auto zdb = [&](int l) { return pg_smem + aoff[l] + 6 * a.hidden[l - 1] * TR; };
    int woff[kMaxLayers + 1], boff[kMaxLayers + 1];
    int pw = 0;
    for (int l = 1; l <= L; ++l) {
        woff[l] = pw;
        pw += a.hidden[l - 1] * (l == 1 ? S : a.hidden[l - 2]);
    }
    const int wo_off = pw;
    const int hL = a.hidden[L - 1];
    int pb = wo_off + hL;
    for (int l = 1; l <= L; ++l) {
        boff[l] = pb;
        pb += a.hidden[l - 1];
    }
    const int bo_off = pb;
    const double *__restrict__ wout = a.w_in_out64 + wo_off;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, nwave = blockDim.x >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
    using d4 = __attribute__((__vector_size__(4 * sizeof(double)))) double;
    static_assert(TR % 4 == 0, "the tile's rows are whole k-steps of the matrix instruction");
    double *__restrict__ mine = partial + (size_t)blockIdx.x * nparam;
    bool first = true;
    for (int tile = blockIdx.x; tile * TR < a.nrows; tile += gridDim.x) {
        const int row0 = tile * TR;
        const int nt = min(TR, a.nrows - row0);
        for (int idx = tid; idx < TR * S; idx += blockDim.x) {
            const int r = idx / S, e = idx % S;
            double v = 0.0, vd = 0.0;
            if (r < nt) {
                v = a.x64 ? a.x64[(size_t)(row0 + r) * S + e] : (double)a.x[(size_t)(row0 + r) * S + e];
                vd = xdot[(size_t)(row0 + r) * S + e];
            }
            act(0)[e * TR + r] = v;
            actd(0)[e * TR + r] = vd;
        }
        if (tid < TR) {
            yb[tid] = tid < nt ? w[a.rows_atom[row0 + tid]] * a.sig_e : 0.0;
            ydb[tid] = tid < nt ? -a.sig_e : 0.0;
        }
        __syncthreads();
        for (int l = 1; l <= L; ++l) {  // forward with the tangent
            const int h = a.hidden[l - 1], in = l == 1 ? S : a.hidden[l - 2];
            const double *__restrict__ W = wb + woff[l];
            const double *ap = act(l - 1), *adp = actd(l - 1);
            for (int idx = tid; idx < h * TR; idx += blockDim.x) {
                const int o = idx / TR, r = idx % TR;
                double z, t;
                dot_gw2<TR>(W + (size_t)o * in, 1, ap + r, adp + r, in, z, t);
                z += bs[boff[l] - wo_off - hL + o];
                const double sg = 1.0 / (1.0 + exp(-z));
                const double d1 = sg * (1.0 + z * (1.0 - sg));
                act(l)[idx] = z * sg;
                der(l)[idx] = d1;
                der2(l)[idx] = sg * (1.0 - sg) * (2.0 + z * (1.0 - 2.0 * sg));
                zd(l)[idx] = t;
                actd(l)[idx] = d1 * t;
            }
            __syncthreads();
        }
        for (int idx = tid; idx < hL * TR; idx += blockDim.x) {  // the last layer's adjoints
            const int o = idx / TR, r = idx % TR;
            const double ab = yb[r] * wout[o], adb = ydb[r] * wout[o];
            zdb(L)[idx] = adb * der(L)[idx];
            zb(L)[idx] = ab * der(L)[idx] + adb * zd(L)[idx] * der2(L)[idx];
        }
        __syncthreads();
        for (int l = L; l >= 1; --l) {
            const int h = a.hidden[l - 1], in = l == 1 ? S : a.hidden[l - 2];
            const double *__restrict__ W = wb + woff[l];
            if (l == 1 && !xbar) break;
            for (int idx = tid; idx < in * TR; idx += blockDim.x) {
                const int i = idx / TR, r = idx % TR;
                double ab, adb;
                dot_gw2<TR>(W + i, in, zb(l) + r, zdb(l) + r, h, ab, adb);
                if (l > 1) {
                    zdb(l - 1)[idx] = adb * der(l - 1)[idx];
                    zb(l - 1)[idx] = ab * der(l - 1)[idx] + adb * zd(l - 1)[idx] * der2(l - 1)[idx];
                } else if (r < nt) {
                    xbar[(size_t)(row0 + r) * S + i] = ab;
                    xdbar[(size_t)(row0 + r) * S + i] = adb;
                }
            }
            __syncthreads();
        }
        auto put = [&](int p, double v) { mine[p] = first ? v : mine[p] + v; };
        for (int l = 1; l <= L; ++l) {
            const int h = a.hidden[l - 1], in = l == 1 ? S : a.hidden[l - 2];
            const double *ap = act(l - 1), *adp = actd(l - 1), *zbl = zb(l), *zdbl = zdb(l);
            // dW_l = zbar act^T + zdbar actd^T on the FP64 matrix cores, in 16 x 16
            // blocks taken by the waves round-robin (lane maps as in k_mlp_pgrad)
            const int nto = (h + 15) / 16, nti = (in + 15) / 16;
            for (int tt = wave; tt < nto * nti; tt += nwave) {
                const int o0 = (tt / nti) * 16, i0 = (tt % nti) * 16;
                const bool ov = o0 + li < h, iv = i0 + li < in;
                d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int k0 = 0; k0 < TR; k0 += 4) {
                    const int r = k0 + lk;
                    const double a1 = ov ? zbl[(o0 + li) * TR + r] : 0.0, b1 = iv ? ap[(i0 + li) * TR + r] : 0.0;
                    const double a2 = ov ? zdbl[(o0 + li) * TR + r] : 0.0, b2 = iv ? adp[(i0 + li) * TR + r] : 0.0;
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, b2, acc, 0, 0, 0);
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int o = o0 + 4 * e + lk, i = i0 + li;
                    if (o < h && i < in) put(woff[l] + o * in + i, acc[e]);
                }
            }
            for (int o = tid; o < h; o += blockDim.x) {
                double s = 0.0;
                for (int r = 0; r < TR; ++r) s += zb(l)[o * TR + r];
                put(boff[l] + o, s);
            }
        }
        for (int o = tid; o < hL; o += blockDim.x) {
            double s = 0.0;
            for (int r = 0; r < TR; ++r) s += yb[r] * act(L)[o * TR + r] + ydb[r] * actd(L)[o * TR + r];
            put(wo_off + o, s);
        }
        if (tid == 0) {
            double s = 0.0;
            for (int r = 0; r < TR; ++r) s += yb[r];
            put(bo_off, s);
        }
        first = false;
        __syncthreads();
    }
}
// out[p] = sum over the workgroups' partial rows: 64 parameters per
// workgroup; wave v sums rows v, v + 8, v + 16, ... in order, then the eight
// sums are added in wave order (a fixed order, whatever the timing).
constexpr int kReduceWaves = 8;
__global__ __launch_bounds__(64 * kReduceWaves) void k_pgrad_reduce(int nwg, int nparam,
                                                                    const double *__restrict__ partial,
                                                                    double *__restrict__ out) {
    __shared__ double s_sum[kReduceWaves][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + lane;
    double s = 0.0;
    if (p < nparam) {
#pragma unroll 8
        for (int g = wv; g < nwg; g += kReduceWaves) s += partial[(size_t)g * nparam + p];
    }
    s_sum[wv][lane] = s;
    __syncthreads();
    if (wv == 0 && p < nparam) {
        double t = 0.0;
        for (int v = 0; v < kReduceWaves; ++v) t += s_sum[v][lane];
        out[p] = t;
    }
}

int mlp_param_count(const MlpArgs &a) {
    int n = 0, in = a.S;
    for (int l = 0; l < a.nlayers; ++l) {
        n += a.hidden[l] * in + a.hidden[l];
        in = a.hidden[l];
    }
    return n + in + 1;
}
void launch_pgrad_reduce(hipStream_t s, int nwg, int np, const double *partial, double *out) {
    hipLaunchKernelGGL(k_pgrad_reduce, dim3((np + 63) / 64), dim3(64 * kReduceWaves), 0, s, nwg, np, partial, out);
}
bool launch_mlp_dual_grad(hipStream_t s, const MlpArgs &a, const double *w, const double *xdot, double *xbar,
                          double *xdbar, double *partial, double *out) {
    if (a.nrows <= 0) return true;
    constexpr int TR = 8;
    size_t f = (size_t)2 * a.S * TR + 2 * TR;
    for (int l = 0; l < a.nlayers; ++l) f += 7 * (size_t)a.hidden[l] * TR;
    const size_t lds = f * sizeof(double);
    if (lds > 160 * 1024) return false;
    static const bool attr_set = [] {
        JNN_HIP(hipFuncSetAttribute((const void *)k_mlp_dual_pgrad<TR>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    160 * 1024));
        return true;
    }();
    (void)attr_set;
    const int np = mlp_param_count(a);
    const int nwg = mlp_dual_groups(a);
    hipLaunchKernelGGL((k_mlp_dual_pgrad<TR>), dim3(nwg), dim3(256), lds, s, a, w, xdot, xbar, xdbar, np, partial);
    launch_pgrad_reduce(s, nwg, np, partial, out);
    return true;
}
// two workgroups per CU (their LDS tiles fit twice); one partial row each
int mlp_dual_groups(const MlpArgs &a) { return std::max(1, std::min((a.nrows + 7) / 8, 2 * device_cu_count())); }
int mlp_pgrad_groups(const MlpArgs &a) {  // two workgroups per CU, one partial row each
    const int tiles = (a.nrows + 15) / 16;
    return std::max(1, std::min(tiles, 2 * device_cu_count()));
}
bool launch_mlp_param_grad(hipStream_t s, const MlpArgs &a, const double *w, double *partial, double *out) {
    if (a.nrows <= 0) return true;
    constexpr int TR = 16;
    size_t f = (size_t)a.S * TR + TR;
    for (int l = 0; l < a.nlayers; ++l) f += 3 * (size_t)a.hidden[l] * TR;
    const size_t lds = f * sizeof(double);
    if (lds > 160 * 1024) return false;
    static const bool attr_set = [] {
        JNN_HIP(hipFuncSetAttribute((const void *)k_mlp_pgrad<16>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    160 * 1024));
        return true;
    }();
    (void)attr_set;
    const int np = mlp_param_count(a), nwg = mlp_pgrad_groups(a);
    hipLaunchKernelGGL((k_mlp_pgrad<TR>), dim3(nwg), dim3(256), lds, s, a, w, np, partial);
    launch_pgrad_reduce(s, nwg, np, partial, out);
    return true;
}

}  // namespace jnn
