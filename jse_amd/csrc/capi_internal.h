// Internals shared by the C-ABI translation units of libjsenn_hip: error
// capture, device allocation helpers and the device-resident model.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include <algorithm>
#include <cmath>

#include "../../include/jse_nnap_hip.h"
#include "common.h"
#include "kernels.h"

namespace jnn {

extern thread_local std::string g_last_error;

// Runs f, mapping exceptions to a status code and the thread's last error.
template <class F>
int guard(F &&f) {
    try {
        f();
        g_last_error.clear();
        return OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host memory exhausted";
        return ERR_MEMORY;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return ERR_ARG;
    }
}

template <class T>
T *dalloc(size_t n) {
    if (n == 0) n = 1;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e != hipSuccess) throw Error(ERR_MEMORY, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return (T *)p;
}
template <class T>
void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}
template <class T>
T *dupload(const std::vector<T> &v) {
    T *p = dalloc<T>(v.size());
    if (!v.empty()) JNN_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
}
template <class T>
T *halloc(size_t n) {  // pinned host memory
    if (n == 0) n = 1;
    void *p = nullptr;
    hipError_t e = hipHostMalloc(&p, n * sizeof(T), hipHostMallocDefault);
    if (e != hipSuccess) throw Error(ERR_MEMORY, std::string("hipHostMalloc failed: ") + hipGetErrorString(e));
    return (T *)p;
}
template <class T>
void hfree(T *&p) {
    if (p) (void)hipHostFree((void *)p);
    p = nullptr;
}

// atoms per chunk of the per-replica total reductions (one workgroup of 256
// threads x 4 atoms each): a replica's totals are summed over the same chunks
// whatever else is evaluated with it
constexpr int kReduceChunk = 1024;
static_assert(kReduceChunk <= kTotalsChunkMax, "a totals chunk must fit one workgroup");

// Cell grid of an orthogonal periodic box for cut-off `rcut`
// (NeighborListGetter.getProperLinkedCell, NeighborListGetter.java:952-1001):
// div = floor(min L / rcut) cells across the smallest edge, cell edge
// min L / div, and per axis max(floor(L_d / edge), div) cells.
inline void make_grid(const double *b, double rcut, ReplicaGrid &g) {
    for (int d = 0; d < 3; ++d)
        JNN_CHECK(std::isfinite(b[d]) && b[d] > 0, ERR_ARG, "box lengths must be positive and finite");
    const double minbox = std::min(b[0], std::min(b[1], b[2]));
    JNN_CHECK(rcut <= minbox, ERR_UNSUPPORTED,
              "rcut larger than the smallest box length (such a box is evaluated as its supercell)");
    const int div = (int)std::floor(minbox / rcut);
    const double cell_len = minbox / (double)div;
    for (int d = 0; d < 3; ++d) {
        const int sz = (int)std::floor(b[d] / cell_len);
        g.size[d] = std::max(sz, div);
        g.cb[d] = b[d] / g.size[d];
        g.box[d] = b[d];
    }
    g.ncell = g.size[0] * g.size[1] * g.size[2];
}

// Copies per axis of the supercell of a box: 1 where the edge reaches rc,
// else the fewest whose total length does.
inline void box_copies(const double *box, double rc, int cp[3]) {
    for (int d = 0; d < 3; ++d) {
        JNN_CHECK(std::isfinite(box[d]) && box[d] > 0, ERR_ARG, "box lengths must be positive and finite");
        JNN_CHECK(rc / box[d] <= 1024.0, ERR_ARG, "a box edge below 1/1024 of the cutoff");
        int c = 1;
        if (box[d] < rc) {
            c = (int)std::ceil(rc / box[d]);
            while (c * box[d] < rc) ++c;
        }
        cp[d] = c;
    }
}

struct SpeciesDev {
    float *W = nullptr, *mu = nullptr, *sig = nullptr;
    double *W64 = nullptr, *mu64 = nullptr, *sig64 = nullptr;
    float *w_fwd = nullptr, *w_bwd = nullptr, *bias = nullptr;
    double *w_fwd64 = nullptr, *w_bwd64 = nullptr, *bias64 = nullptr;
    uint16_t *wb_fwd = nullptr, *wb_bwd = nullptr;
    // a merge basis: each part's mixing (FP32, FP64); the MLP's per-column G factor
    // (a merge basis, or a basis with L3 / L4 invariants)
    std::vector<float *> pW;
    std::vector<double *> pW64;
    double *l2f = nullptr;
    // per basis (basis(k)): the L3 / L4 coupling terms (n34 = 0: none)
    struct Coupling {
        int n34 = 0;
        int4 *idx = nullptr, *gidx = nullptr;
        double *coef = nullptr, *gcoef = nullptr;
        int *off = nullptr, *goff = nullptr;
    };
    std::vector<Coupling> cpl;
};

}  // namespace jnn

struct jse_nnap_model_s {
    jnn::ModelHost host;
    std::vector<jnn::SpeciesDev> dev;
    float *zeros = nullptr, *ones = nullptr;  // identity normalisation for raw descriptors
    double *zeros64 = nullptr, *ones64 = nullptr;
    int smax = 0, pmmax = 0;

    void upload();
    ~jse_nnap_model_s();
};

namespace jnn {
// Precision modes (JSE_NNAP_FP32 / BF16_MLP / FP32_FAST): a system's precision is checked once.
inline void check_precision(int p) {
    JNN_CHECK(p == JSE_NNAP_FP32 || p == JSE_NNAP_BF16_MLP || p == JSE_NNAP_FP32_FAST, ERR_ARG, "unknown precision");
}
// FP64 pair geometry, harmonics and A (the plain mode)
inline bool precision_acc(int p) { return p == JSE_NNAP_FP32; }
// Throws status 2 naming the limit when a species' descriptor shape has no
// kernel that runs it in this precision mode.
void check_model_shapes(const jse_nnap_model_s *m, int precision);
// The descriptor fields of species t's segment shared by every caller (rows,
// outputs and launch shape are the caller's).
SegmentArgs species_segment(const jse_nnap_model_s *m, int t, int precision, const int *types);
// The column layout and L3 / L4 coupling terms of basis k of species t.
void set_segment_basis(SegmentArgs &a, const jse_nnap_model_s *m, int t, int k);
// Species t's segment narrowed to basis k (a merge basis, MergedBasis.java: the
// part's mixing and cutoff, its slice of the x / mu / sigma / G rows (advanced
// by the sizes of parts 0..k-1), its own A block, g added to by parts after
// the first); the segment itself for a single basis.
SegmentArgs part_segment(const SegmentArgs &seg, const jse_nnap_model_s *m, int t, int k);
// The lmax the descriptor kernels are chosen by: max(lmax, l3max, l4max), or
// kRadialLmax for the radial basis.
inline int desc_lmax(const SpeciesHost &b) { return b.radial ? kRadialLmax : b.lmm(); }
// The MlpArgs of species t's segment (its rows, descriptors and gradients).
MlpArgs mlp_args(const jse_nnap_model_s *m, int t, const SegmentArgs &sa, double *E, bool grad, int wg_per_cu = 0);
// Points the segment's descriptor rows at row0 of a buffer of smax values per
// row: FP64 values (the buffer holds 2 smax floats per row) in the plain
// mode, FP32 otherwise.
void set_segment_x(SegmentArgs &a, float *xbuf, size_t row0, int smax);
// Every species' descriptor forward (or backward): one launch over all segments
// when they share one descriptor shape, else one per species.  segs[t] is
// species t's segment (nrows 0: none).
void enqueue_desc(hipStream_t s, const jse_nnap_model_s *m, const SegmentArgs *segs, bool backward,
                  const NeighborList &nl);
// Every species' MLP: one launch when their shapes allow it, else one per species.
void enqueue_mlp_all(hipStream_t s, const jse_nnap_model_s *m, int precision, const MlpArgs *args, int nsp);
// Descriptor forward, MLP and descriptor backward of a plugin hook's rows:
// the rows of species t are rows[seg[t] .. seg[t+1]-1] (list rows, in the
// order of nl); E receives one energy per list row.
void enqueue_hook_rows(hipStream_t s, jse_nnap_model model, int precision, const int *seg, const int *rows,
                       const int *types, float *xs, float *A, float *G, double *E, const NeighborList &nl);
}  // namespace jnn
