// The reference's LAMMPS GPU pair hook as jse-compatible packed-argument
// entry points (nnap_interface_cuda.cu:307-483, driven by
// NNAP.computeLammpsCuda, NNAP.java:949-1030).  jse binds native functions
// as `int name(void *args)` with the declared parameters packed back to back
// in declaration order, no padding (SourceScanner.java:167-290, wrapper
// written at :278-289); every entry point below reads its block that way.
//
// The reference generates a model-specific library, so the model's shapes are
// not among the call arguments.  A prebuilt library is told the model once
// with jse_nnap_lammps_bind (a handle from jse_nnap_model_load of the same
// model file); the per-call parameter pointers (aFpParam, aNnParam, ...),
// the merge tables and the reference's scratch buffers are accepted and not
// read.  Positions are also kept in FP64 on the device, so displacements are
// formed in FP64 (the reference subtracts FP32 positions), and forces and
// virials are accumulated in FP64 before the caller's FP32 buffers are
// written.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "capi_internal.h"
#include "common.h"
#include "kernels.h"

using namespace jnn;

namespace jnn {
void launch_lmp_rows(hipStream_t s, int inum, const int *ilist, const int *numneigh, const int *firstneigh,
                     const double *x, double rc2, const NeighborList &nl, int *overflow);
void launch_lmp_types(hipStream_t s, int n, const int *type, const int *map, int *out);
void launch_zero_double(hipStream_t s, size_t n, double *p);
void launch_lmp_scatter(hipStream_t s, int inum, int nall, const int *ilist, const NeighborList &nl, int vflag,
                        int vflag_atom, int cvflag_atom, double *F, double *V0, double *V1);
void launch_lmp_out(hipStream_t s, int inum, int nall, int eflag, int vflag, int nv1, const double *F, const double *E,
                    const double *V0, const double *V1, float *f, float *eatom0, float *vatom0, float *vatom1);
}  // namespace jnn

namespace {

// Reads the packed argument block in declaration order.
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

// Device state of the hook between its three calls (one pair style per process).
struct LmpContext {
    jse_nnap_model model = nullptr;
    int precision = JSE_NNAP_FP32;
    hipStream_t stream = nullptr;
    int nall = 0, inum = 0, cap = 64;
    size_t nall_alloc = 0, inum_alloc = 0, row_alloc = 0;
    double *x = nullptr, *hx = nullptr;
    int *types = nullptr, *rows = nullptr, *overflow = nullptr, *h_overflow = nullptr;
    float *xs = nullptr, *A = nullptr, *G = nullptr;
    double *E = nullptr, *F = nullptr, *V0 = nullptr, *V1 = nullptr;
    NeighborList nl{};
    std::vector<int> h_type, h_ilist;

    void ensure(int nall_, int inum_, int cap_) {
        if (!stream) JNN_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        if (!h_overflow) h_overflow = halloc<int>(1);
        if (!overflow) overflow = dalloc<int>(1);
        if ((size_t)nall_ > nall_alloc) {
            dfree(x); hfree(hx); dfree(types); dfree(F); dfree(V1);
            nall_alloc = (size_t)nall_ + nall_ / 4 + 64;
            x = dalloc<double>(3 * nall_alloc);
            hx = halloc<double>(3 * nall_alloc);
            types = dalloc<int>(nall_alloc);
            F = dalloc<double>(3 * nall_alloc);
            V1 = dalloc<double>(9 * nall_alloc);
        }
        const size_t need_rows = (size_t)inum_ * cap_;
        if ((size_t)inum_ > inum_alloc || need_rows > row_alloc) {
            dfree(rows); dfree(xs); dfree(A); dfree(G); dfree(E); dfree(V0);
            dfree(nl.cnt); dfree(nl.key); dfree(nl.d); dfree(nl.g);
            inum_alloc = std::max(inum_alloc, (size_t)inum_ + inum_ / 4 + 64);
            row_alloc = inum_alloc * cap_;
            rows = dalloc<int>(inum_alloc);
            xs = dalloc<float>(inum_alloc * model->smax * 2);  // room for FP64 rows (the plain mode)
            G = dalloc<float>(inum_alloc * model->smax);
            A = dalloc<float>(inum_alloc * model->pmmax);
            E = dalloc<double>(inum_alloc);
            V0 = dalloc<double>(6 * inum_alloc);
            nl.cnt = dalloc<int>(inum_alloc);
            nl.key = dalloc<int>(row_alloc);
            nl.d = dalloc<float4>(row_alloc);
            nl.g = dalloc<float4>(row_alloc);
        }
        nl.cap = cap_;
    }
};
LmpContext g_lmp;
std::mutex g_lmp_mu;

}  // namespace

namespace jnn {
void enqueue_hook_rows(hipStream_t s, jse_nnap_model model, int precision, const int *seg, const int *rows,
                       const int *types, float *xs, float *A, float *G, double *E, const NeighborList &nl) {
    const ModelHost &mh = model->host;
    const int nt = mh.ntypes();
    // fixed arrays (no host allocation: the no-alloc GPUMD entry runs through here too)
    SegmentArgs segs[kMaxSpecies];
    MlpArgs mlps[kMaxSpecies];
    for (int t = 0; t < nt; ++t) {
        SegmentArgs &sa = segs[t];
        sa = species_segment(model, t, precision, types);
        sa.nrows = seg[t + 1] - seg[t];
        sa.rows_atom = rows + seg[t];
        set_segment_x(sa, xs, (size_t)seg[t], model->smax);
        sa.A = A + (size_t)seg[t] * model->pmmax;
        sa.G = G + (size_t)seg[t] * model->smax;
        mlps[t] = mlp_args(model, t, sa, E, true);
    }
    // one launch per kernel for all species of one descriptor shape, persistent
    // workgroups as the autotuner picks them for a 32,000-atom liquid (the hooks
    // are not tuned per call)
    for (int t = 0; t < nt; ++t) segs[t].wg_per_cu = 32;
    enqueue_desc(s, model, segs, false, nl);
    enqueue_mlp_all(s, model, precision, mlps, nt);
    for (int t = 0; t < nt; ++t) segs[t].wg_per_cu = 16;
    enqueue_desc(s, model, segs, true, nl);
}
}  // namespace jnn

extern "C" {

// Binds the model evaluated by the hook: packed (jse_nnap_model model, int precision).
int jse_nnap_lammps_bind(void *args) {
    return guard([&] {
        JNN_CHECK(args, ERR_ARG, "null argument block");
        Args a(args);
        const jse_nnap_model m = a.get<jse_nnap_model>();
        const int prec = a.get<int>();
        JNN_CHECK(m, ERR_ARG, "null model");
        check_precision(prec);
        check_model_shapes(m, prec);
        std::lock_guard<std::mutex> lk(g_lmp_mu);
        g_lmp.model = m;
        g_lmp.precision = prec;
    });
}

// nnap_interface_cuda.cu:307-318: (int *ilist, int *numneigh, int inum, int *numneighMax)
int jse_nnap_statNeiNumLammps(void *args) {
    Args a(args);
    const int *ilist = a.get<const int *>();
    const int *numneigh = a.get<const int *>();
    const int inum = a.get<int>();
    int *out = a.get<int *>();
    int mx = 0;
    for (int ii = 0; ii < inum; ++ii) mx = std::max(mx, numneigh[ilist[ii]]);
    out[0] = mx;
    return 0;
}

// nnap_interface_cuda.cu:319-359: (int inum, int nlocalghost, int nlflag, int neighnumMax, double **x,
// int *type, int *ilist, int *numneigh, int **firstneigh, flt_t *fltBuf, int *intBuf, flt_t *cudaX,
// int *cudaType, int *cudaIlist, int *cudaNumneigh, int *cudaFirstneigh), flt_t = float.
int jse_nnap_lammps2cuda(void *args) {
    return guard([&] {
        Args a(args);
        const int inum = a.get<int>(), nall = a.get<int>(), nlflag = a.get<int>(), nmax = a.get<int>();
        double **x = a.get<double **>();
        const int *type = a.get<const int *>();
        const int *ilist = a.get<const int *>();
        const int *numneigh = a.get<const int *>();
        int **firstneigh = a.get<int **>();
        float *fltBuf = a.get<float *>();
        int *intBuf = a.get<int *>();
        float *cudaX = a.get<float *>();
        int *cudaType = a.get<int *>();
        int *cudaIlist = a.get<int *>();
        int *cudaNumneigh = a.get<int *>();
        int *cudaFirstneigh = a.get<int *>();
        JNN_CHECK(inum >= 0 && nall >= inum && nmax >= 0, ERR_ARG, "bad LAMMPS sizes");
        std::lock_guard<std::mutex> lk(g_lmp_mu);
        JNN_CHECK(g_lmp.model, ERR_ARG, "no model bound (jse_nnap_lammps_bind)");
        // the caller's device buffers, as the reference fills them
        for (int i = 0; i < nall; ++i)
            for (int d = 0; d < 3; ++d) fltBuf[(size_t)d * nall + i] = (float)x[i][d];
        JNN_HIP(hipMemcpy(cudaX, fltBuf, sizeof(float) * 3 * nall, hipMemcpyHostToDevice));
        JNN_HIP(hipMemcpy(cudaType, type, sizeof(int) * nall, hipMemcpyHostToDevice));
        if (nlflag) {
            JNN_HIP(hipMemcpy(cudaIlist, ilist, sizeof(int) * inum, hipMemcpyHostToDevice));
            for (int ii = 0; ii < inum; ++ii) intBuf[ii] = numneigh[ilist[ii]];
            JNN_HIP(hipMemcpy(cudaNumneigh, intBuf, sizeof(int) * inum, hipMemcpyHostToDevice));
            for (int ii = 0; ii < inum; ++ii) {
                const int i = ilist[ii];
                for (int jj = 0; jj < numneigh[i]; ++jj)
                    intBuf[(size_t)jj * inum + ii] = firstneigh[i][jj] & 0x1FFFFFFF;  // LAMMPS NEIGHMASK
            }
            JNN_HIP(hipMemcpy(cudaFirstneigh, intBuf, sizeof(int) * (size_t)inum * nmax, hipMemcpyHostToDevice));
            g_lmp.h_ilist.assign(ilist, ilist + inum);
            g_lmp.cap = std::max(32, ((nmax + 31) / 32) * 32);
        }
        // FP64 positions for the displacements
        g_lmp.ensure(nall, inum, g_lmp.cap);
        for (int i = 0; i < nall; ++i)
            for (int d = 0; d < 3; ++d) g_lmp.hx[3 * (size_t)i + d] = x[i][d];
        JNN_HIP(hipMemcpyAsync(g_lmp.x, g_lmp.hx, sizeof(double) * 3 * nall, hipMemcpyHostToDevice, g_lmp.stream));
        JNN_HIP(hipStreamSynchronize(g_lmp.stream));
        g_lmp.h_type.assign(type, type + nall);
        g_lmp.nall = nall;
        g_lmp.inum = inum;
    });
}

// nnap_interface_cuda.cu:431-483: (int inum, int nlocalghost, int eflag, int vflag, int eflagAtom,
// int vflagAtom, int cvflagAtom, flt_t *x, int *type, int *ilist, int *nmerges, int **mergeSorted,
// flt_t **cutsq, int *numneigh, int *firstneigh, int *aLmpType2NNAPType, flt_t **aFpHyperParam,
// flt_t **aFpParam, flt_t **aNnParam, flt_t **aNormParam, flt_t *f, flt_t *eatom0, flt_t *vatom0,
// flt_t *vatom1, then 10 scratch buffers).  Outputs: f[3][nall], eatom0[inum], vatom0[6][inum],
// vatom1[6 or 9][nall] (FP32, overwritten).
int jse_nnap_computeLammpsCuda(void *args) {
    return guard([&] {
        Args a(args);
        const int inum = a.get<int>(), nall = a.get<int>(), eflag = a.get<int>(), vflag = a.get<int>();
        const int eflag_atom = a.get<int>(), vflag_atom = a.get<int>(), cvflag_atom = a.get<int>();
        (void)a.get<float *>();                    // x (FP32 copy; the FP64 one is used)
        const int *type = a.get<const int *>();    // device
        const int *ilist = a.get<const int *>();   // device
        (void)a.get<int *>();                      // nmerges
        (void)a.get<int **>();                     // mergeSorted
        (void)a.get<float **>();                   // cutsq
        const int *numneigh = a.get<const int *>();    // device, in ilist order
        const int *firstneigh = a.get<const int *>();  // device, [jj * inum + ii]
        const int *tmap = a.get<const int *>();        // device, LAMMPS type -> model type
        for (int k = 0; k < 4; ++k) (void)a.get<float **>();  // parameter blocks (the bound model is used)
        float *f = a.get<float *>();
        float *eatom0 = a.get<float *>();
        float *vatom0 = a.get<float *>();
        float *vatom1 = a.get<float *>();
        std::lock_guard<std::mutex> lk(g_lmp_mu);
        LmpContext &c = g_lmp;
        JNN_CHECK(c.model, ERR_ARG, "no model bound (jse_nnap_lammps_bind)");
        JNN_CHECK(inum == c.inum && nall == c.nall, ERR_ARG, "sizes differ from the last jse_nnap_lammps2cuda");
        const ModelHost &mh = c.model->host;
        const int nt = mh.ntypes();
        // model types (device, all atoms) and species-grouped rows (host, listed atoms)
        int maxt = 0;
        for (int t : c.h_type) maxt = std::max(maxt, t);
        std::vector<int> hmap(maxt + 1, 0);
        JNN_HIP(hipMemcpy(hmap.data(), tmap, sizeof(int) * (maxt + 1), hipMemcpyDeviceToHost));
        std::vector<int> seg(nt + 1, 0), rows(inum);
        for (int ii = 0; ii < inum; ++ii) {
            const int t = hmap[c.h_type[c.h_ilist[ii]]];
            JNN_CHECK(t >= 1 && t <= nt, ERR_TYPE, "LAMMPS type maps outside the model's types");
            seg[t]++;
        }
        for (int t = 1; t <= nt; ++t) seg[t] += seg[t - 1];
        {
            std::vector<int> fill(seg.begin(), seg.end() - 1);
            for (int ii = 0; ii < inum; ++ii) rows[fill[hmap[c.h_type[c.h_ilist[ii]]] - 1]++] = ii;
        }
        hipStream_t s = c.stream;
        for (int attempt = 0;; ++attempt) {
            c.ensure(nall, inum, c.cap);
            if (inum) JNN_HIP(hipMemcpyAsync(c.rows, rows.data(), sizeof(int) * inum, hipMemcpyHostToDevice, s));
            launch_lmp_types(s, nall, type, tmap, c.types);
            launch_fill_int(s, 1, c.overflow, 0);
            c.nl.types = c.types;
            launch_lmp_rows(s, inum, ilist, numneigh, firstneigh, c.x, mh.rcut_max() * mh.rcut_max(), c.nl, c.overflow);
            JNN_HIP(hipMemcpyAsync(c.h_overflow, c.overflow, sizeof(int), hipMemcpyDeviceToHost, s));
            JNN_HIP(hipStreamSynchronize(s));
            if (*c.h_overflow <= c.nl.cap) break;
            JNN_CHECK(attempt < 2 && *c.h_overflow <= 4096, ERR_CAPACITY, "neighbour row capacity exceeded");
            c.cap = ((*c.h_overflow + 31) / 32) * 32;
        }
        enqueue_hook_rows(s, c.model, c.precision, seg.data(), c.rows, c.types, c.xs, c.A, c.G, c.E, c.nl);
        launch_zero_double(s, 3 * (size_t)nall, c.F);
        launch_zero_double(s, 9 * (size_t)nall, c.V1);
        launch_lmp_scatter(s, inum, nall, ilist, c.nl, vflag, vflag_atom, cvflag_atom, c.F, c.V0, c.V1);
        const int nv1 = cvflag_atom ? 9 : (vflag_atom ? 6 : 0);
        launch_lmp_out(s, inum, nall, eflag || eflag_atom, vflag, nv1, c.F, c.E, c.V0, c.V1, f, eatom0, vatom0, vatom1);
        JNN_HIP(hipGetLastError());
        JNN_HIP(hipStreamSynchronize(s));
    });
}

// nnap_interface_cuda.cu:361-429: (int inum, int nlocalghost, int eflag, int vflag, int eflagAtom,
// int vflagAtom, int cvflagAtom, double **f, double *engVdwl, double *eatom, double *virial,
// double **vatom, double **cvatom, flt_t *fltBuf, const int *ilist, flt_t *cudaF, flt_t *cudaEatom0,
// flt_t *cudaVatom0, flt_t *cudaVatom1).  Accumulates (+=) into LAMMPS' arrays; the values come from
// the FP64 accumulators of the last computeLammpsCuda.
int jse_nnap_cuda2lammps(void *args) {
    return guard([&] {
        Args a(args);
        const int inum = a.get<int>(), nall = a.get<int>(), eflag = a.get<int>(), vflag = a.get<int>();
        const int eflag_atom = a.get<int>(), vflag_atom = a.get<int>(), cvflag_atom = a.get<int>();
        double **f = a.get<double **>();
        double *eng = a.get<double *>();
        double *eatom = a.get<double *>();
        double *virial = a.get<double *>();
        double **vatom = a.get<double **>();
        double **cvatom = a.get<double **>();
        (void)a.get<float *>();  // fltBuf
        const int *ilist = a.get<const int *>();
        std::lock_guard<std::mutex> lk(g_lmp_mu);
        LmpContext &c = g_lmp;
        JNN_CHECK(inum == c.inum && nall == c.nall, ERR_ARG, "sizes differ from the last jse_nnap_lammps2cuda");
        std::vector<double> F(3 * (size_t)nall), E(inum), V0(6 * (size_t)inum), V1(9 * (size_t)nall);
        hipStream_t s = c.stream;
        if (nall) JNN_HIP(hipMemcpyAsync(F.data(), c.F, sizeof(double) * F.size(), hipMemcpyDeviceToHost, s));
        if (eflag && inum) JNN_HIP(hipMemcpyAsync(E.data(), c.E, sizeof(double) * inum, hipMemcpyDeviceToHost, s));
        if (vflag && inum) JNN_HIP(hipMemcpyAsync(V0.data(), c.V0, sizeof(double) * V0.size(), hipMemcpyDeviceToHost, s));
        if ((cvflag_atom || vflag_atom) && nall)
            JNN_HIP(hipMemcpyAsync(V1.data(), c.V1, sizeof(double) * V1.size(), hipMemcpyDeviceToHost, s));
        JNN_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < nall; ++i)
            for (int d = 0; d < 3; ++d) f[i][d] += F[(size_t)d * nall + i];
        if (eflag)
            for (int ii = 0; ii < inum; ++ii) {
                *eng += E[ii];
                if (eflag_atom) eatom[ilist[ii]] += E[ii];
            }
        if (vflag)
            for (int ii = 0; ii < inum; ++ii)
                for (int q = 0; q < 6; ++q) virial[q] += V0[(size_t)q * inum + ii];
        if (cvflag_atom)
            for (int i = 0; i < nall; ++i)
                for (int q = 0; q < 9; ++q) cvatom[i][q] += V1[(size_t)q * nall + i];
        if (vflag_atom)
            for (int i = 0; i < nall; ++i)
                for (int q = 0; q < 6; ++q) vatom[i][q] += V1[(size_t)q * nall + i];
    });
}

}  // extern "C"
