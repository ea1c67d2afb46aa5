// The reference's GPUMD hook as a jse-compatible packed-argument entry point
// (jse_nnap_computeGPUMD, nnap_interface_cuda.cu:485-518, driven by
// NNAP.computeGPUMD, NNAP.java:1047-1083).  jse binds native functions as
// `int name(void *args)` with the declared parameters packed back to back in
// declaration order, no padding (SourceScanner.java:167-290).
//
// GPUMD's arrays are all on the device: a full neighbour list, column-major
// [jj * N + i], with FP32 displacements and 0-based types, and FP64 outputs
// that the hook adds into (forces, per-atom virials [9][N], potential).  As
// with the LAMMPS hook, the model is named once (jse_nnap_gpumd_bind) and the
// reference's per-call parameter blocks, merge tables and scratch buffers are
// accepted and not read.
//
// The reference scatters every pair's force and virial with FP64 atomics.
// Rows are sorted by key, so results do not depend on the order of GPUMD's
// list.  When every particle is a centre and the list is symmetric (each pair
// listed from both ends, once), the plain path's atomic-free force gather
// produces per-atom results that are then added to GPUMD's arrays; a pair
// without its reverse, or a neighbour listed twice, sends the call to the
// atomic scatter instead.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "capi_internal.h"
#include "common.h"
#include "kernels.h"

using namespace jnn;

namespace jnn {
void launch_gpumd_types(hipStream_t s, int n, const int *g_type, int *out);
void launch_gpumd_rows(hipStream_t s, int N, int N1, int N2, const int *num, const int *list, const float *nl_dx,
                       const float *nl_dy, const float *nl_dz, float rc2, int tag_rows, const NeighborList &nl,
                       int *overflow);
void launch_gpumd_sort(hipStream_t s, int nrow, const NeighborList &nl, int *key2, float4 *d2, int *flag);
void launch_gpumd_add(hipStream_t s, int N, const double *F, const double *V, const double *E, double *fx, double *fy,
                      double *fz, double *vir, double *pot);
void launch_gpumd_scatter(hipStream_t s, int N, int N1, int nrow, const NeighborList &nl, const double *E, double *fx,
                          double *fy, double *fz, double *vir, double *pot);
void launch_gpumd_group(hipStream_t s, int nrow, const int *g_type, int nt, int *rows, int *seg, int *bad);
}  // namespace jnn

namespace {

struct Args {
    const char *p;
    explicit Args(void *a) : p((const char *)a) {}
    template <class T>
    T get() {
        T v;
        std::memcpy(&v, p, sizeof(T));
        p += sizeof(T);
        return v;
    }
};

// Device state of the hook between calls (one potential per process).
struct GpumdContext {
    jse_nnap_model model = nullptr;
    int precision = JSE_NNAP_FP32;
    hipStream_t stream = nullptr;
    int cap = 64;
    size_t n_alloc = 0, row_alloc = 0, row_n_alloc = 0;
    int *types = nullptr, *rows = nullptr, *overflow = nullptr, *h_overflow = nullptr;
    int *flag = nullptr, *h_flag = nullptr;  // unmatched reverse pairs or repeated neighbours
    float *xs = nullptr, *A = nullptr, *G = nullptr;
    double *E = nullptr, *F = nullptr, *V = nullptr;
    int *key2 = nullptr;
    float4 *d2 = nullptr;
    NeighborList nl{};

    void ensure(int n, int nrow, int cap_) {
        if (!stream) JNN_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        if (!h_overflow) h_overflow = halloc<int>(1);
        if (!overflow) overflow = dalloc<int>(1);
        if (!h_flag) h_flag = halloc<int>(1);
        if (!flag) flag = dalloc<int>(1);
        if ((size_t)n > n_alloc) {
            dfree(types); dfree(F); dfree(V);
            n_alloc = (size_t)n + n / 4 + 64;
            types = dalloc<int>(n_alloc);
            F = dalloc<double>(3 * n_alloc);
            V = dalloc<double>(9 * n_alloc);
        }
        const size_t need = (size_t)nrow * cap_;
        if ((size_t)nrow > row_n_alloc || need > row_alloc) {
            dfree(rows); dfree(xs); dfree(A); dfree(G); dfree(E);
            dfree(nl.cnt); dfree(nl.key); dfree(nl.d); dfree(nl.g); dfree(key2); dfree(d2);
            row_n_alloc = std::max(row_n_alloc, (size_t)nrow + nrow / 4 + 64);
            row_alloc = row_n_alloc * cap_;
            rows = dalloc<int>(row_n_alloc);
            xs = dalloc<float>(row_n_alloc * model->smax * 2);  // room for FP64 rows (the plain mode)
            G = dalloc<float>(row_n_alloc * model->smax);
            A = dalloc<float>(row_n_alloc * model->pmmax);
            E = dalloc<double>(row_n_alloc);
            nl.cnt = dalloc<int>(row_n_alloc);
            nl.key = dalloc<int>(row_alloc);
            nl.d = dalloc<float4>(row_alloc);
            nl.g = dalloc<float4>(row_alloc);
            key2 = dalloc<int>(row_alloc);
            d2 = dalloc<float4>(row_alloc);
        }
        nl.cap = cap_;
    }
};
GpumdContext g_gpumd;
std::mutex g_gpumd_mu;

// ---- the no-allocation variant ----------------------------------------
// Scratch device memory of jse_nnap_computeGPUMD_noalloc: one array per GPU
// the library runs on (it is held until the process ends).  Host code must not
// pass `scratch` itself to a kernel or a copy (in host code the name is a
// stand-in, not a device address): the device address comes from
// hipGetSymbolAddress, once per device, and only that pointer is used.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kNoAllocMaxDevices = 64;
unsigned char *g_scratch_ptr[kNoAllocMaxDevices] = {};
std::mutex g_scratch_mu;

// The device address of this device's scratch array (looked up once per device).
unsigned char *device_scratch() {
    int dev = 0;
    JNN_HIP(hipGetDevice(&dev));
    JNN_CHECK(dev >= 0 && dev < kNoAllocMaxDevices, ERR_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        JNN_CHECK(e == hipSuccess && p, ERR_HIP,
                  std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned char *)p;
    }
    return g_scratch_ptr[dev];
}

// Carves the buffers of one call from the scratch array (256-byte aligned
// pieces); `need` is the total, checked against the array's size by the caller.
struct ScratchLayout {
    unsigned char *base;
    size_t off = 0;
    template <class T>
    T *take(size_t n) {
        T *p = reinterpret_cast<T *>(base ? base + off : nullptr);
        off += ((n * sizeof(T) + 255) / 256) * 256;
        return p;
    }
};
struct NoAllocBufs {
    int *types, *rows, *seg, *overflow, *flag, *bad, *key2;
    float *xs, *A, *G;
    double *E, *F, *V;
    float4 *d2;
    NeighborList nl{};
    size_t bytes;
};
NoAllocBufs carve(unsigned char *base, int N, int nrow, int nt, int cap, int smax, int pmmax) {
    ScratchLayout L{base};
    NoAllocBufs b;
    b.overflow = L.take<int>(1);
    b.flag = L.take<int>(1);
    b.bad = L.take<int>(1);
    b.seg = L.take<int>((size_t)nt + 1);
    b.types = L.take<int>(N);
    b.F = L.take<double>(3 * (size_t)N);
    b.V = L.take<double>(9 * (size_t)N);
    b.rows = L.take<int>(nrow);
    b.xs = L.take<float>((size_t)nrow * smax * 2);  // room for FP64 rows (the plain mode)
    b.G = L.take<float>((size_t)nrow * smax);
    b.A = L.take<float>((size_t)nrow * pmmax);
    b.E = L.take<double>(nrow);
    b.nl.cap = cap;
    b.nl.cnt = L.take<int>(nrow);
    b.nl.key = L.take<int>((size_t)nrow * cap);
    b.nl.d = L.take<float4>((size_t)nrow * cap);
    b.nl.g = L.take<float4>((size_t)nrow * cap);
    b.key2 = L.take<int>((size_t)nrow * cap);
    b.d2 = L.take<float4>((size_t)nrow * cap);
    b.bytes = L.off;
    return b;
}

}  // namespace

extern "C" {

// Binds the model evaluated by the hook: packed (jse_nnap_model model, int precision).
int jse_nnap_gpumd_bind(void *args) {
    return guard([&] {
        JNN_CHECK(args, ERR_ARG, "null argument block");
        Args a(args);
        const jse_nnap_model m = a.get<jse_nnap_model>();
        const int prec = a.get<int>();
        JNN_CHECK(m, ERR_ARG, "null model");
        check_precision(prec);
        check_model_shapes(m, prec);
        std::lock_guard<std::mutex> lk(g_gpumd_mu);
        g_gpumd.model = m;
        g_gpumd.precision = prec;
    });
}

// nnap_interface_cuda.cu:485-493: (int number_of_particles, int N1, int N2, const int *g_neighbor_number,
// const int *g_neighbor_list, const float *nl_dx, const float *nl_dy, const float *nl_dz, const int *g_type,
// const int *nmerges, const int **mergeSorted, const flt_t **cutsq, flt_t **aFpHyperParam, flt_t **aFpParam,
// flt_t **aNnParam, flt_t **aNormParam, double *g_fx, double *g_fy, double *g_fz, double *g_virial,
// double *g_potential, then 9 scratch buffers), flt_t = float.  Centres are the particles N1 .. N2-1;
// forces, virials (on the neighbour) and energies are added into GPUMD's arrays.
int jse_nnap_computeGPUMD(void *args) {
    return guard([&] {
        JNN_CHECK(args, ERR_ARG, "null argument block");
        Args a(args);
        const int N = a.get<int>(), N1 = a.get<int>(), N2 = a.get<int>();
        const int *num = a.get<const int *>();
        const int *list = a.get<const int *>();
        const float *nl_dx = a.get<const float *>();
        const float *nl_dy = a.get<const float *>();
        const float *nl_dz = a.get<const float *>();
        const int *g_type = a.get<const int *>();
        (void)a.get<const int *>();     // nmerges
        (void)a.get<const int **>();    // mergeSorted
        (void)a.get<const float **>();  // cutsq
        for (int k = 0; k < 4; ++k) (void)a.get<float **>();  // parameter blocks (the bound model is used)
        double *g_fx = a.get<double *>();
        double *g_fy = a.get<double *>();
        double *g_fz = a.get<double *>();
        double *g_virial = a.get<double *>();
        double *g_potential = a.get<double *>();
        JNN_CHECK(N >= 0 && N1 >= 0 && N2 >= N1 && N2 <= N, ERR_ARG, "bad GPUMD particle range");
        JNN_CHECK(N < (1 << 26), ERR_ARG, "number_of_particles out of range [0, 2^26)");
        std::lock_guard<std::mutex> lk(g_gpumd_mu);
        GpumdContext &c = g_gpumd;
        JNN_CHECK(c.model, ERR_ARG, "no model bound (jse_nnap_gpumd_bind)");
        const int nrow = N2 - N1;
        if (nrow == 0) return;
        // GPUMD's inputs come from its own stream(s): wait for them, as the reference's
        // legacy-default-stream launches do
        JNN_HIP(hipDeviceSynchronize());
        const ModelHost &mh = c.model->host;
        const int nt = mh.ntypes();
        // species-grouped rows (host) from the centres' types
        std::vector<int> ht(nrow);
        JNN_HIP(hipMemcpy(ht.data(), g_type + N1, sizeof(int) * nrow, hipMemcpyDeviceToHost));
        std::vector<int> seg(nt + 1, 0), rows(nrow);
        for (int r = 0; r < nrow; ++r) {
            const int t = ht[r] + 1;
            JNN_CHECK(t >= 1 && t <= nt, ERR_TYPE, "GPUMD type outside the model's types");
            seg[t]++;
        }
        for (int t = 1; t <= nt; ++t) seg[t] += seg[t - 1];
        {
            std::vector<int> fill(seg.begin(), seg.end() - 1);
            for (int r = 0; r < nrow; ++r) rows[fill[ht[r]]++] = r;
        }
        hipStream_t s = c.stream;
        const float rc = (float)mh.rcut_max();
        const bool all = N1 == 0 && N2 == N;  // every particle a centre: rows by particle, gather possible
        for (int attempt = 0;; ++attempt) {
            c.ensure(N, nrow, c.cap);
            JNN_HIP(hipMemcpyAsync(c.rows, rows.data(), sizeof(int) * nrow, hipMemcpyHostToDevice, s));
            launch_gpumd_types(s, N, g_type, c.types);
            launch_fill_int(s, 1, c.overflow, 0);
            launch_fill_int(s, 1, c.flag, 0);
            c.nl.types = c.types;
            launch_gpumd_rows(s, N, N1, N2, num, list, nl_dx, nl_dy, nl_dz, rc * rc, all ? 1 : 0, c.nl, c.overflow);
            JNN_HIP(hipMemcpyAsync(c.h_overflow, c.overflow, sizeof(int), hipMemcpyDeviceToHost, s));
            JNN_HIP(hipStreamSynchronize(s));
            if (*c.h_overflow <= c.nl.cap) break;
            JNN_CHECK(attempt < 2 && *c.h_overflow <= 4096, ERR_CAPACITY, "neighbour row capacity exceeded");
            c.cap = ((*c.h_overflow + 31) / 32) * 32;
        }
        // rows sorted by key: the results do not depend on the order of GPUMD's list,
        // and the gather's reverse search needs it
        launch_gpumd_sort(s, nrow, c.nl, c.key2, c.d2, c.flag);
        std::swap(c.nl.key, c.key2);
        std::swap(c.nl.d, c.d2);
        enqueue_hook_rows(s, c.model, c.precision, seg.data(), c.rows, c.types, c.xs, c.A, c.G, c.E, c.nl);
        bool gathered = false;
        if (all) {
            launch_gather(s, N, nullptr, c.nl, c.F, c.V, 1, 0, nullptr, c.flag);
            JNN_HIP(hipMemcpyAsync(c.h_flag, c.flag, sizeof(int), hipMemcpyDeviceToHost, s));
            JNN_HIP(hipStreamSynchronize(s));
            if (*c.h_flag == 0) {
                launch_gpumd_add(s, N, c.F, c.V, c.E, g_fx, g_fy, g_fz, g_virial, g_potential);
                gathered = true;
            }
        }
        if (!gathered) launch_gpumd_scatter(s, N, N1, nrow, c.nl, c.E, g_fx, g_fy, g_fz, g_virial, g_potential);
        JNN_HIP(hipGetLastError());
        JNN_HIP(hipStreamSynchronize(s));
    });
}


// The no-allocation variant of jse_nnap_computeGPUMD (same packed arguments,
// same results): every buffer beyond GPUMD's own inputs and outputs is carved
// from this GPU's scratch array, and the call makes no hipMalloc, hipFree or
// host allocation of its own.  It runs on the null stream.
int jse_nnap_computeGPUMD_noalloc(void *args) {
    return guard([&] {
        JNN_CHECK(args, ERR_ARG, "null argument block");
        Args a(args);
        const int N = a.get<int>(), N1 = a.get<int>(), N2 = a.get<int>();
        const int *num = a.get<const int *>();
        const int *list = a.get<const int *>();
        const float *nl_dx = a.get<const float *>();
        const float *nl_dy = a.get<const float *>();
        const float *nl_dz = a.get<const float *>();
        const int *g_type = a.get<const int *>();
        (void)a.get<const int *>();     // nmerges
        (void)a.get<const int **>();    // mergeSorted
        (void)a.get<const float **>();  // cutsq
        for (int k = 0; k < 4; ++k) (void)a.get<float **>();  // parameter blocks (the bound model is used)
        double *g_fx = a.get<double *>();
        double *g_fy = a.get<double *>();
        double *g_fz = a.get<double *>();
        double *g_virial = a.get<double *>();
        double *g_potential = a.get<double *>();
        JNN_CHECK(N >= 0 && N1 >= 0 && N2 >= N1 && N2 <= N, ERR_ARG, "bad GPUMD particle range");
        JNN_CHECK(N < (1 << 26), ERR_ARG, "number_of_particles out of range [0, 2^26)");
        std::lock_guard<std::mutex> lk(g_gpumd_mu);  // calls share this GPU's scratch array
        const jse_nnap_model model = g_gpumd.model;
        const int precision = g_gpumd.precision;
        JNN_CHECK(model, ERR_ARG, "no model bound (jse_nnap_gpumd_bind)");
        const int nrow = N2 - N1;
        if (nrow == 0) return;
        const int nt = model->host.ntypes();
        unsigned char *base = device_scratch();
        JNN_HIP(hipDeviceSynchronize());  // GPUMD's inputs come from its own stream(s)
        hipStream_t s = nullptr;
        const float rc = (float)model->host.rcut_max();
        const bool all = N1 == 0 && N2 == N;
        int cap = 64;
        NoAllocBufs b{};
        int seg[kMaxSpecies + 1];
        for (int attempt = 0;; ++attempt) {
            b = carve(nullptr, N, nrow, nt, cap, model->smax, model->pmmax);
            JNN_CHECK(b.bytes <= kNoAllocScratchBytes, ERR_MEMORY,
                      "jse_nnap_computeGPUMD_noalloc: " + std::to_string(N) + " particles with rows of " +
                          std::to_string(cap) + " neighbours need " + std::to_string(b.bytes) +
                          " bytes of scratch; the array holds " + std::to_string(kNoAllocScratchBytes));
            b = carve(base, N, nrow, nt, cap, model->smax, model->pmmax);
            if (attempt == 0) {
                launch_fill_int(s, 1, b.bad, 0);
                launch_gpumd_group(s, nrow, g_type + N1, nt, b.rows, b.seg, b.bad);
                int bad = 0;
                JNN_HIP(hipMemcpyAsync(seg, b.seg, sizeof(int) * (nt + 1), hipMemcpyDeviceToHost, s));
                JNN_HIP(hipMemcpyAsync(&bad, b.bad, sizeof(int), hipMemcpyDeviceToHost, s));
                JNN_HIP(hipStreamSynchronize(s));
                JNN_CHECK(!bad, ERR_TYPE, "GPUMD type outside the model's types");
            } else {  // the rows sit at another offset of the new layout: group again
                launch_gpumd_group(s, nrow, g_type + N1, nt, b.rows, b.seg, b.bad);
            }
            launch_gpumd_types(s, N, g_type, b.types);
            launch_fill_int(s, 1, b.overflow, 0);
            launch_fill_int(s, 1, b.flag, 0);
            b.nl.types = b.types;
            launch_gpumd_rows(s, N, N1, N2, num, list, nl_dx, nl_dy, nl_dz, rc * rc, all ? 1 : 0, b.nl, b.overflow);
            int ovf = 0;
            JNN_HIP(hipMemcpyAsync(&ovf, b.overflow, sizeof(int), hipMemcpyDeviceToHost, s));
            JNN_HIP(hipStreamSynchronize(s));
            if (ovf <= cap) break;
            JNN_CHECK(attempt < 2 && ovf <= 4096, ERR_CAPACITY, "neighbour row capacity exceeded");
            cap = ((ovf + 31) / 32) * 32;
        }
        launch_gpumd_sort(s, nrow, b.nl, b.key2, b.d2, b.flag);
        std::swap(b.nl.key, b.key2);
        std::swap(b.nl.d, b.d2);
        enqueue_hook_rows(s, model, precision, seg, b.rows, b.types, b.xs, b.A, b.G, b.E, b.nl);
        bool gathered = false;
        if (all) {
            launch_gather(s, N, nullptr, b.nl, b.F, b.V, 1, 0, nullptr, b.flag);
            int flag = 0;
            JNN_HIP(hipMemcpyAsync(&flag, b.flag, sizeof(int), hipMemcpyDeviceToHost, s));
            JNN_HIP(hipStreamSynchronize(s));
            if (flag == 0) {
                launch_gpumd_add(s, N, b.F, b.V, b.E, g_fx, g_fy, g_fz, g_virial, g_potential);
                gathered = true;
            }
        }
        if (!gathered) launch_gpumd_scatter(s, N, N1, nrow, b.nl, b.E, g_fx, g_fy, g_fz, g_virial, g_potential);
        JNN_HIP(hipGetLastError());
        JNN_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"
