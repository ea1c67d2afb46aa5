// C-ABI of libjsenn_hip: model and system handles, device memory, the
// evaluation pipeline and its timing.  See include/jse_nnap_hip.h.
#include "../../include/jse_nnap_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <numeric>
#include <sstream>

#include "capi_internal.h"
#include "common.h"
#include "kernels.h"
#include "sh_coupling.h"

using namespace jnn;

namespace jnn {
thread_local std::string g_last_error;
}

namespace {
uint16_t to_bf16(float f) {  // round to nearest even
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

}  // namespace

void jse_nnap_model_s::upload() {
    int nt = host.ntypes();
    dev.resize(nt);
    for (int t = 0; t < nt; ++t) {
        const SpeciesHost &sp = host.species[t];
        SpeciesDev &d = dev[t];
        std::vector<float> W(sp.W.begin(), sp.W.end());
        d.W = dupload(W);
        d.W64 = dupload(sp.W);
        const int S = sp.S();
        std::vector<float> mu(sp.norm.begin() + 2, sp.norm.begin() + 2 + S);
        std::vector<float> sg(sp.norm.begin() + 2 + S, sp.norm.end());
        d.mu = dupload(mu);
        d.sig = dupload(sg);
        d.mu64 = dupload(std::vector<double>(sp.norm.begin() + 2, sp.norm.begin() + 2 + S));
        d.sig64 = dupload(std::vector<double>(sp.norm.begin() + 2 + S, sp.norm.end()));
        // forward layout: per layer [in][out], then output W; backward layout: per layer [out][in]
        std::vector<float> wf, wb, bias;
        size_t off = 0;
        int in = S;
        for (int h : sp.hidden) {
            for (int i = 0; i < in; ++i)
                for (int o = 0; o < h; ++o) wf.push_back((float)sp.nn[off + (size_t)o * in + i]);
            for (size_t q = 0; q < (size_t)in * h; ++q) wb.push_back((float)sp.nn[off + q]);
            off += (size_t)in * h;
            in = h;
        }
        for (int i = 0; i < in; ++i) wf.push_back((float)sp.nn[off + i]);
        off += in;
        for (size_t q = off; q < sp.nn.size(); ++q) bias.push_back((float)sp.nn[q]);
        d.w_fwd = dupload(wf);
        d.w_bwd = dupload(wb);
        d.bias = dupload(bias);
        {  // the same layouts in FP64 (the plain mode's MLP)
            std::vector<double> wf64, wb64, bias64;
            size_t o = 0;
            int ci = S;
            for (int h : sp.hidden) {
                for (int i = 0; i < ci; ++i)
                    for (int q = 0; q < h; ++q) wf64.push_back(sp.nn[o + (size_t)q * ci + i]);
                for (size_t q = 0; q < (size_t)ci * h; ++q) wb64.push_back(sp.nn[o + q]);
                o += (size_t)ci * h;
                ci = h;
            }
            for (int i = 0; i < ci; ++i) wf64.push_back(sp.nn[o + i]);
            o += ci;
            for (size_t q = o; q < sp.nn.size(); ++q) bias64.push_back(sp.nn[q]);
            d.w_fwd64 = dupload(wf64);
            d.w_bwd64 = dupload(wb64);
            d.bias64 = dupload(bias64);
        }
        std::vector<uint16_t> bf(wf.size()), bb(wb.size());
        for (size_t q = 0; q < wf.size(); ++q) bf[q] = to_bf16(wf[q]);
        for (size_t q = 0; q < wb.size(); ++q) bb[q] = to_bf16(wb[q]);
        d.wb_fwd = dupload(bf);
        d.wb_bwd = dupload(bb);
        for (const SpeciesHost &q : sp.parts) {  // a merge basis: the parts' mixings
            d.pW.push_back(dupload(std::vector<float>(q.W.begin(), q.W.end())));
            d.pW64.push_back(dupload(q.W));
        }
        bool any34 = false;
        d.cpl.resize(sp.nbases());
        for (int k = 0; k < sp.nbases(); ++k) {  // L3 / L4 coupling terms per basis
            const SpeciesHost &q = sp.basis(k);
            if (!q.l34()) continue;
            any34 = true;
            const CouplingTerms ct = coupling_terms(q.l3max, q.l4max, q.M());
            SpeciesDev::Coupling &c = d.cpl[k];
            c.n34 = ct.ninv;
            c.idx = reinterpret_cast<int4 *>(dupload(ct.idx));
            c.gidx = reinterpret_cast<int4 *>(dupload(ct.gidx));
            c.coef = dupload(ct.coef);
            c.gcoef = dupload(ct.gcoef);
            c.off = dupload(ct.off);
            c.goff = dupload(ct.goff);
        }
        if (!sp.parts.empty() || any34) {
            // G's per-column factor: 2 4pi/(2l+1) on a spherical basis' L2 columns, 1 on
            // its L3 / L4 columns and on a radial basis'
            std::vector<double> l2;
            for (int k = 0; k < sp.nbases(); ++k) {
                const SpeciesHost &q = sp.basis(k);
                for (int e = 0; e < q.P * q.SL(); ++e) {
                    const int c = e % q.SL();
                    l2.push_back(q.radial || c > q.lmax ? 1.0 : 2.0 * 12.566370614359172954 / (double)(2 * c + 1));
                }
            }
            d.l2f = dupload(l2);
        }
        smax = std::max(smax, S);
        pmmax = std::max(pmmax, sp.pm());
    }
    std::vector<float> z((size_t)smax, 0.f), o((size_t)smax, 1.f);
    zeros = dupload(z);
    ones = dupload(o);
    zeros64 = dupload(std::vector<double>((size_t)smax, 0.0));
    ones64 = dupload(std::vector<double>((size_t)smax, 1.0));
}

jse_nnap_model_s::~jse_nnap_model_s() {
    for (auto &d : dev) {
        dfree(d.W); dfree(d.W64); dfree(d.mu); dfree(d.sig); dfree(d.mu64); dfree(d.sig64); dfree(d.w_fwd); dfree(d.w_bwd); dfree(d.bias); dfree(d.w_fwd64); dfree(d.w_bwd64); dfree(d.bias64);
        dfree(d.wb_fwd); dfree(d.wb_bwd);
        for (float *p : d.pW) dfree(p);
        for (double *p : d.pW64) dfree(p);
        dfree(d.l2f);
        for (SpeciesDev::Coupling &c : d.cpl) {
            dfree(c.idx); dfree(c.gidx); dfree(c.coef); dfree(c.gcoef); dfree(c.off); dfree(c.goff);
        }
    }
    dfree(zeros);
    dfree(ones);
    dfree(zeros64);
    dfree(ones64);
}

namespace jnn {
void check_model_shapes(const jse_nnap_model_s *m, int precision) {
    const int nt = m->host.ntypes();
    for (int t = 0; t < nt; ++t) {
        for (int k = 0; k < m->host.species[t].nbases(); ++k) {
            const SpeciesHost &sp = m->host.species[t].basis(k);
            const int lm = desc_lmax(sp);
            if (!sp.l34() && desc_tiled(lm, sp.K(), sp.P, sp.mix)) continue;
            const std::string why = desc_gen_limit(lm, sp.K(), sp.P, nt, precision_acc(precision));
            JNN_CHECK(why.empty(), ERR_UNSUPPORTED,
                      "species " + std::to_string(t + 1) + " (lmax=" + std::to_string(sp.lmax) + " nmax=" +
                          std::to_string(sp.nmax) + " size_np=" + std::to_string(sp.P) + ", " + std::to_string(nt) +
                          " species): " + why);
        }
    }
}
SegmentArgs species_segment(const jse_nnap_model_s *m, int t, int precision, const int *types) {
    const SpeciesHost &sp = m->host.species[t];
    const SpeciesDev &dv = m->dev[t];
    SegmentArgs a{};
    a.self_t = t + 1;
    a.mirror = sp.mirror;
    a.types = types;
    a.rc = (float)sp.rcut;
    a.rc64 = sp.rcut;
    a.ntypes = m->host.ntypes();
    a.W = dv.W;
    a.W64 = dv.W64;
    a.mu = dv.mu;
    a.sig = dv.sig;
    a.mu64 = dv.mu64;
    a.sig64 = dv.sig64;
    a.acc = precision_acc(precision) ? 1 : 0;
    a.fp32_acc = a.acc ? 0 : 1;
    a.xs = sp.S();
    a.accum_g = 0;
    set_segment_basis(a, m, t, 0);
    return a;
}
SegmentArgs part_segment(const SegmentArgs &seg, const jse_nnap_model_s *m, int t, int k) {
    const SpeciesHost &sp = m->host.species[t];
    SegmentArgs a = seg;
    if (sp.parts.empty()) return a;
    const SpeciesDev &dv = m->dev[t];
    int xoff = 0;
    size_t aoff = 0;
    for (int q = 0; q < k; ++q) {
        xoff += sp.basis(q).P * sp.basis(q).SL();
        aoff += (size_t)seg.nrows * sp.basis(q).P * sp.basis(q).M();
    }
    const SpeciesHost &b = sp.basis(k);
    a.W = dv.pW[k];
    a.W64 = dv.pW64[k];
    a.rc = (float)b.rcut;
    a.rc64 = b.rcut;
    if (a.x) a.x += xoff;
    if (a.x64) a.x64 += xoff;
    a.mu += xoff;
    a.sig += xoff;
    a.mu64 += xoff;
    a.sig64 += xoff;
    a.G += xoff;
    a.A += aoff;
    a.accum_g = k > 0 ? 1 : 0;
    set_segment_basis(a, m, t, k);
    return a;
}
void set_segment_basis(SegmentArgs &a, const jse_nnap_model_s *m, int t, int k) {
    const SpeciesHost &b = m->host.species[t].basis(k);
    const SpeciesDev::Coupling &c = m->dev[t].cpl[k];
    a.l2n = b.lmax + 1;
    a.sl = b.SL();
    a.n34 = c.n34;
    a.cpl_idx = c.idx;
    a.cplg_idx = c.gidx;
    a.cpl_coef = c.coef;
    a.cplg_coef = c.gcoef;
    a.cpl_off = c.off;
    a.cplg_off = c.goff;
}
MlpArgs mlp_args(const jse_nnap_model_s *m, int t, const SegmentArgs &sa, double *E, bool grad, int wg_per_cu) {
    const SpeciesHost &sp = m->host.species[t];
    const SpeciesDev &dv = m->dev[t];
    MlpArgs a{};
    a.nrows = sa.nrows;
    a.S = sp.S();
    a.L1 = sp.lmax + 1;
    a.radial = sp.radial ? 1 : 0;
    a.nlayers = (int)sp.hidden.size();
    for (int l = 0; l < a.nlayers; ++l) a.hidden[l] = sp.hidden[l];
    a.rows_atom = sa.rows_atom;
    a.x = sa.x;
    a.x64 = sa.x64;
    a.w_in_out = dv.w_fwd;
    a.w_out_in = dv.w_bwd;
    a.bias = dv.bias;
    a.sig = dv.sig;
    a.w_in_out64 = dv.w_fwd64;
    a.w_out_in64 = dv.w_bwd64;
    a.bias64 = dv.bias64;
    a.sig64 = dv.sig64;
    a.sig_e = sp.norm[1];
    a.mu_e = sp.norm[0];
    a.E = E;
    a.G = sa.G;
    a.want_grad = grad;
    a.wg_per_cu = wg_per_cu;
    a.l2f = dv.l2f;
    return a;
}
void set_segment_x(SegmentArgs &a, float *xbuf, size_t row0, int smax) {
    if (a.acc) {
        a.x = nullptr;
        a.x64 = reinterpret_cast<double *>(xbuf) + row0 * smax;
    } else {
        a.x = xbuf + row0 * smax;
        a.x64 = nullptr;
    }
}
void enqueue_desc(hipStream_t s, const jse_nnap_model_s *m, const SegmentArgs *segs, bool backward,
                  const NeighborList &nl) {
    const int nt = m->host.ntypes();
    auto same_shape = [&](int t, int u) {
        const SpeciesHost &a = m->host.species[t], &b = m->host.species[u];
        return a.radial == b.radial && a.lmax == b.lmax && a.l3max == b.l3max && a.l4max == b.l4max &&
               a.K() == b.K() && a.P == b.P && a.mix == b.mix;
    };
    int first = -1;
    bool one = nt <= kSegMax, merged = false;
    for (int t = 0; t < nt; ++t) {
        if (segs[t].nrows <= 0) continue;
        if (!m->host.species[t].parts.empty()) merged = true;
        if (first < 0) first = t;
        else if (!same_shape(t, first)) one = false;
    }
    if (first < 0) return;
    auto fail = [&](int t) {
        const SpeciesHost &sp = m->host.species[t];
        throw Error(ERR_UNSUPPORTED, std::string("no descriptor ") + (backward ? "backward " : "") +
                                         "kernel for lmax=" + std::to_string(sp.lmax) + " nmax=" +
                                         std::to_string(sp.nmax) + " size_np=" + std::to_string(sp.P));
    };
    if (merged) {
        // a merge basis (MergedBasis.java): one launch per species and part; part k
        // writes its slice of the species' x row (x, mu, sigma and G advanced by the
        // sizes of parts 0..k-1), keeps its own A, and its backward adds into g
        for (int t = 0; t < nt; ++t) {
            if (segs[t].nrows <= 0) continue;
            const SpeciesHost &sp = m->host.species[t];
            for (int k = 0; k < sp.nbases(); ++k) {
                const SpeciesHost &b = sp.basis(k);
                const SegmentArgs a = part_segment(segs[t], m, t, k);
                const int lm = desc_lmax(b);
                const bool ok = backward ? launch_desc_bwd(s, lm, b.K(), b.P, b.mix, a, nl)
                                         : launch_desc_fwd(s, lm, b.K(), b.P, b.mix, a, nl);
                if (!ok) fail(t);
            }
        }
        return;
    }
    if (one) {
        SegmentSet set{};
        int rows = 0;
        for (int t = 0; t < nt; ++t) {
            if (segs[t].nrows <= 0) continue;
            set.a[set.nseg] = segs[t];
            set.row0[set.nseg] = rows;
            rows += segs[t].nrows;
            ++set.nseg;
        }
        set.row0[set.nseg] = rows;
        const SpeciesHost &sp = m->host.species[first];
        const int lm = desc_lmax(sp);
        const bool ok = backward ? launch_desc_bwd_set(s, lm, sp.K(), sp.P, sp.mix, set, nl)
                                 : launch_desc_fwd_set(s, lm, sp.K(), sp.P, sp.mix, set, nl);
        if (!ok) fail(first);
        return;
    }
    for (int t = 0; t < nt; ++t) {
        if (segs[t].nrows <= 0) continue;
        const SpeciesHost &sp = m->host.species[t];
        const int lm = desc_lmax(sp);
        const bool ok = backward ? launch_desc_bwd(s, lm, sp.K(), sp.P, sp.mix, segs[t], nl)
                                 : launch_desc_fwd(s, lm, sp.K(), sp.P, sp.mix, segs[t], nl);
        if (!ok) fail(t);
    }
}
void enqueue_mlp_all(hipStream_t s, const jse_nnap_model_s *m, int precision, const MlpArgs *args, int nsp) {
    const int mode = precision == JSE_NNAP_BF16_MLP ? kMlpBf16 : (precision_acc(precision) ? kMlpFp64 : kMlpFp32);
    if (launch_mlp_mfma_multi(s, args, nsp, mode)) return;
    for (int t = 0; t < nsp; ++t) {
        const MlpArgs &a = args[t];
        if (a.nrows <= 0) continue;
        if (mode == kMlpBf16) {
            // networks without BF16 tiles (2 hidden layers of 32 or 64) run in FP32, inside the BF16 budget
            if (!launch_mlp_mfma_bf16(s, a) && !launch_mlp_mfma_fp32(s, a) && !launch_mlp_fp32(s, a))
                throw Error(ERR_UNSUPPORTED, kMlpFp32Limit);
        } else if (mode == kMlpFp64) {
            if (!launch_mlp_mfma_fp64(s, a) && !launch_mlp_fp64(s, a)) throw Error(ERR_UNSUPPORTED, kMlpFp64Limit);
        } else if (!launch_mlp_mfma_fp32(s, a) && !launch_mlp_fp32(s, a)) {
            throw Error(ERR_UNSUPPORTED, kMlpFp32Limit);
        }
    }
}
}  // namespace jnn

struct jse_nnap_system_s {
    jse_nnap_model model = nullptr;
    int n = 0;       // atoms over all replicas
    int nrep = 0;    // replicas (independent periodic boxes)
    // Boxes with an edge shorter than the cutoff (the reference's
    // ExpandLinkedCell, NeighborListGetter.java:704-776, 985-995): such a
    // replica is held as its supercell of cp[0] x cp[1] x cp[2] copies of the
    // box, every edge of which reaches the cutoff, so that the cell list finds
    // each pair within the cutoff once per copy.  Copy c = cx + cp[0] (cy +
    // cp[1] cz) of the caller's atom i of a replica is held atom ioff + c * unat
    // + i, at x_i + (cx, cy, cz) * box; copy 0 is the caller's atom and carries
    // its results, and the replica's totals are its supercell's divided by
    // ncopy.  Without copies every map is the identity.
    struct RepCopies {
        int cp[3] = {1, 1, 1};
        int ncopy = 1;
        double box[3] = {0.0, 0.0, 0.0};  // the caller's box
        int unat = 0;                     // the caller's atoms
        int uoff = 0, ioff = 0;           // first caller row, first held row
    };
    std::vector<RepCopies> rcopy;  // per replica
    bool copies = false;           // some replica is held as a supercell
    int nuser = 0;                 // the caller's atoms (n without copies)
    int precision = JSE_NNAP_FP32;
    std::vector<int> types;
    std::vector<int> rep_natoms;
    std::vector<ReplicaGrid> reps;     // host copy of the per-replica grids
    std::vector<int> seg;              // species segment offsets into rows
    int ncell_total = 0;
    double rc2 = 0;
    hipStream_t stream = nullptr;
    bool have_pos = false, have_nl = false, pending = false;
    int *h_overflow = nullptr;         // pinned host copy of the overflow flag
    int last_what = 0;

    double *d_xyz = nullptr, *d_xw = nullptr, *d_sxw = nullptr;
    int *d_cell = nullptr, *d_rank = nullptr, *d_sidx = nullptr, *d_types = nullptr, *d_rows = nullptr;
    int *d_rep_of = nullptr;
    int *d_slot_of = nullptr;  // atom -> its neighbour-list row (cell-sorted slot)
    ReplicaGrid *d_reps = nullptr;
    int *d_cell_cnt = nullptr, *d_cell_start = nullptr, *d_cell_rep = nullptr;
    int ncell_alloc = 0;
    int *d_overflow = nullptr;
    NeighborList nl{};
    float *d_x = nullptr, *d_A = nullptr, *d_G = nullptr;
    double *d_E = nullptr, *d_F = nullptr, *d_V = nullptr, *d_partial = nullptr, *d_tot = nullptr;
    int *d_arrive = nullptr;  // per-replica chunk arrivals of the totals kernel (zero between calls)
    int4 *d_chunks = nullptr;
    int *d_rep_chunk = nullptr;
    int nchunk = 0;
    static constexpr int kChunk = kReduceChunk;
    hipEvent_t ev[8] = {};
    // captured step graphs, one per evaluation kind (E / E+F / E+F+V)
    hipGraphExec_t graph_exec[8] = {};
    bool use_graph = true;
    LaunchTune tune;                   // launch shapes (autotuner)
    bool large_cells = false;          // cell build variant (set with the boxes)
    // Device buffers of the training entry points, kept between calls: the k-th
    // request of a call gets the k-th buffer, reallocated only when it must grow
    // (each call ends with a stream synchronisation, so reuse is ordered).
    std::vector<std::pair<double *, size_t>> tws;
    size_t tws_next = 0;
    double *tws_get(size_t cnt) {
        cnt = std::max<size_t>(cnt, 1);
        if (tws_next == tws.size()) tws.push_back({nullptr, 0});
        auto &b = tws[tws_next++];
        if (b.second < cnt) {
            dfree(b.first);
            b.second = 0;
            b.first = dalloc<double>(cnt);
            b.second = cnt;
        }
        return b.first;
    }

    ~jse_nnap_system_s() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        free_nl();
        dfree(d_xyz); dfree(d_xw); dfree(d_sxw); dfree(d_cell); dfree(d_rank); dfree(d_sidx); dfree(d_types);
        dfree(d_rows); dfree(d_rep_of); dfree(d_slot_of); dfree(d_reps); dfree(d_cell_cnt); dfree(d_cell_start); dfree(d_cell_rep); dfree(d_overflow);
        dfree(d_x); dfree(d_A); dfree(d_G); dfree(d_E); dfree(d_F); dfree(d_V); dfree(d_partial); dfree(d_tot); dfree(d_arrive);
        dfree(d_chunks); dfree(d_rep_chunk);
        for (auto &b : tws) dfree(b.first);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (h_overflow) (void)hipHostFree(h_overflow);
        if (stream) (void)hipStreamDestroy(stream);
    }
    void drop_graphs() {
        for (auto &g : graph_exec)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
    }
    void free_nl() {
        dfree(nl.cnt); dfree(nl.key); dfree(nl.d); dfree(nl.rev); dfree(nl.g); dfree(nl.dlo);
    }
    void alloc_nl(int cap) {
        drop_graphs();
        free_nl();
        JNN_CHECK(cap > 0 && cap % 4 == 0, ERR_ARG, "neighbour capacity must be a positive multiple of 4");
        nl.cap = cap;
        const size_t m = (size_t)n * cap;
        nl.cnt = dalloc<int>(n);
        nl.key = dalloc<int>(m);
        nl.d = dalloc<float4>(m);
        nl.rev = dalloc<int>(m);
        nl.g = dalloc<float4>(m);
        if (precision_acc(precision)) nl.dlo = dalloc<float4>(m);
    }

    void grid_of(const double *b, ReplicaGrid &g) const { make_grid(b, model->host.rcut_max(), g); }
    void set_boxes(const double *boxes) {
        std::vector<ReplicaGrid> r(nrep);
        int coff = 0, aoff = 0;
        for (int q = 0; q < nrep; ++q) {
            grid_of(boxes + 3 * q, r[q]);
            r[q].cell_off = coff;
            r[q].atom_off = aoff;
            r[q].natoms = rep_natoms[q];
            coff += r[q].ncell + 1;   // ncell starts + the end entry
            aoff += rep_natoms[q];
        }
        JNN_HIP(hipStreamSynchronize(stream));
        reps = r;
        ncell_total = coff;
        // the cell build's large-cell variant when some replica's 27 neighbour
        // blocks average 32 atoms or more (a replica of one or two cells per axis)
        large_cells = false;
        for (int q = 0; q < nrep; ++q)
            if (r[q].natoms >= 32 * std::max(1, r[q].ncell)) large_cells = true;
        if (ncell_total + 1 > ncell_alloc) {
            dfree(d_cell_cnt);
            dfree(d_cell_start);
            dfree(d_cell_rep);
            ncell_alloc = ncell_total + 1;
            d_cell_cnt = dalloc<int>(ncell_alloc);
            // zero once: every build's scan clears the counts it consumed
            JNN_HIP(hipMemset(d_cell_cnt, 0, sizeof(int) * ncell_alloc));
            d_cell_start = dalloc<int>(ncell_alloc + 1);
            d_cell_rep = dalloc<int>(ncell_alloc);
        }
        std::vector<int> cell_rep(ncell_total, -1);
        for (int q = 0; q < nrep; ++q)
            std::fill(cell_rep.begin() + r[q].cell_off, cell_rep.begin() + r[q].cell_off + r[q].ncell, q);
        if (!d_reps) d_reps = dalloc<ReplicaGrid>(nrep);
        JNN_HIP(hipMemcpy(d_reps, reps.data(), sizeof(ReplicaGrid) * nrep, hipMemcpyHostToDevice));
        if (ncell_total) JNN_HIP(hipMemcpy(d_cell_rep, cell_rep.data(), sizeof(int) * ncell_total, hipMemcpyHostToDevice));
        drop_graphs();
        have_nl = false;
    }

    void init(jse_nnap_model m, int nr, const int *natoms_per, const int *ty, const double *boxes, int prec) {
        model = m;
        nrep = nr;
        precision = prec;
        JNN_CHECK(nr >= 1, ERR_ARG, "at least one replica is required");
        check_precision(prec);
        check_model_shapes(m, prec);
        long long tot = 0;
        for (int q = 0; q < nr; ++q) {
            JNN_CHECK(natoms_per[q] >= 0, ERR_ARG, "negative atom count");
            tot += natoms_per[q];
        }
        JNN_CHECK(tot < (1 << 26), ERR_ARG, "natoms out of range [0, 2^26)");
        n = (int)tot;
        rep_natoms.assign(natoms_per, natoms_per + nr);
        const int nt = m->host.ntypes();
        types.assign(ty, ty + n);
        for (int i = 0; i < n; ++i)
            JNN_CHECK(types[i] >= 1 && types[i] <= nt, ERR_TYPE,
                      "Exist type (" + std::to_string(types[i]) + ") greater than the input ntypes (" +
                          std::to_string(nt) + ")");
        // species-grouped row order (stable), as the reference's LAMMPS path regroups by type
        std::vector<int> rows(n);
        std::iota(rows.begin(), rows.end(), 0);
        std::stable_sort(rows.begin(), rows.end(), [&](int a, int c) { return types[a] < types[c]; });
        seg.assign(nt + 1, 0);
        for (int i = 0; i < n; ++i) seg[types[i]]++;
        for (int t = 1; t <= nt; ++t) seg[t] += seg[t - 1];
        std::vector<int> rep_of(n);
        std::vector<int4> chunks;
        std::vector<int> rep_chunk(nr + 1, 0);
        for (int q = 0, a = 0; q < nr; ++q) {
            rep_chunk[q] = (int)chunks.size();
            for (int k = 0; k < rep_natoms[q]; ++k) rep_of[a + k] = q;
            // at least one chunk (possibly empty) per replica, so every totals row is written
            for (int c0 = 0; c0 == 0 || c0 < rep_natoms[q]; c0 += kChunk)
                chunks.push_back(make_int4(q, a + c0, a + std::min(c0 + kChunk, rep_natoms[q]), 0));
            a += rep_natoms[q];
        }
        rep_chunk[nr] = (int)chunks.size();
        nchunk = (int)chunks.size();
        rc2 = m->host.rcut_max() * m->host.rcut_max();
        JNN_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        for (auto &e : ev) JNN_HIP(hipEventCreate(&e));
        d_types = dupload(types);
        d_rows = dupload(rows);
        d_rep_of = dupload(rep_of);
        d_chunks = dupload(chunks);
        d_rep_chunk = dupload(rep_chunk);
        d_xyz = dalloc<double>((size_t)3 * n);
        d_xw = dalloc<double>((size_t)3 * n);
        d_sxw = dalloc<double>((size_t)3 * n);
        d_cell = dalloc<int>(n);
        d_rank = dalloc<int>(n);
        d_sidx = dalloc<int>(n);
        d_slot_of = dalloc<int>(n);
        d_overflow = dalloc<int>(1);
        JNN_HIP(hipHostMalloc((void **)&h_overflow, sizeof(int), hipHostMallocDefault));
        *h_overflow = 0;
        d_x = dalloc<float>((size_t)n * m->smax * (precision_acc(prec) ? 2 : 1));  // FP64 rows in the plain mode
        d_G = dalloc<float>((size_t)n * m->smax);
        d_A = dalloc<float>((size_t)n * m->pmmax);
        d_E = dalloc<double>(n);
        d_F = dalloc<double>((size_t)3 * n);
        d_V = dalloc<double>((size_t)9 * n);
        d_partial = dalloc<double>((size_t)std::max(nchunk, 1) * 10);
        d_tot = dalloc<double>((size_t)10 * nr);
        d_arrive = dalloc<int>(nr);
        JNN_HIP(hipMemset(d_arrive, 0, sizeof(int) * nr));
        set_boxes(boxes);
        // initial row capacity from the densest replica, rounded up to whole wavefronts
        double rho = 0.0;
        for (int q = 0; q < nr; ++q)
            if (rep_natoms[q] > 0)
                rho = std::max(rho, rep_natoms[q] / (boxes[3 * q] * boxes[3 * q + 1] * boxes[3 * q + 2]));
        const double rmax = m->host.rcut_max();
        const double est = rho * 4.18879 * rmax * rmax * rmax;
        int cap = (int)std::ceil((est * 1.5 + 16) / 64.0) * 64;
        cap = std::max(64, std::min(cap, 512));
        alloc_nl(cap);
    }

    // Enqueue the neighbour build (no host synchronisation: a row longer
    // than the capacity raises a device flag that finish() checks).
    void build_nl(bool timed, bool copy_flag = true) {
        JNN_CHECK(have_pos, ERR_ARG, "positions not set");
        if (timed) JNN_HIP(hipEventRecord(ev[0], stream));
        // the cell counts start at zero (cleared by the previous build's scan); the
        // scan also clears the overflow flag before the cell-tile build can raise it
        launch_wrap_bin(stream, n, nullptr, d_xyz, d_rep_of, d_reps, nullptr, d_xw, d_cell, d_cell_cnt, d_rank);
        launch_cell_scan(stream, nrep, nullptr, d_reps, nullptr, d_cell_cnt, d_cell_start, true, d_overflow);
        launch_cell_scatter(stream, n, nullptr, d_rep_of, nullptr, d_cell, d_rank, d_cell_start, d_xw, d_sidx, d_sxw,
                            d_slot_of);
        nl.types = d_types;
        launch_nl_build_cells(stream, ncell_total, d_cell_rep, d_reps, rc2, d_cell_start, d_sidx, d_sxw, nl, d_overflow,
                              tune.nl_wg_per_cu, large_cells);
        if (timed) JNN_HIP(hipEventRecord(ev[1], stream));
        // (the force gather finds reverse pairs itself; phase 1 stays empty)
        if (timed) JNN_HIP(hipEventRecord(ev[2], stream));
        if (copy_flag) copy_overflow();
    }
    // the overflow flag to the host (finish() reads it); a full evaluation copies it
    // after its last kernel, so no copy sits between the neighbour build and the descriptor
    void copy_overflow() {
        JNN_HIP(hipMemcpyAsync(h_overflow, d_overflow, sizeof(int), hipMemcpyDeviceToHost, stream));
    }
    // Waits for the enqueued work; when a neighbour row overflowed, grows the
    // capacity and redoes the last request (neighbour build or full compute).
    void finish() {
        JNN_HIP(hipStreamSynchronize(stream));
        if (!pending) return;
        pending = false;
        for (int attempt = 0; *h_overflow > 0; ++attempt) {
            const int ovf = *h_overflow;
            JNN_CHECK(ovf <= 512, ERR_CAPACITY,
                      "an atom has " + std::to_string(ovf) + " neighbours within rcut; at most 512 are supported");
            JNN_CHECK(attempt < 4, ERR_CAPACITY, "neighbour list capacity retries exhausted");
            alloc_nl(((ovf + 16 + 63) / 64) * 64);
            if (last_what) enqueue(last_what, false);
            else build_nl(false);
            JNN_HIP(hipStreamSynchronize(stream));
        }
    }

    // The caller's per-atom rows (width values each; nullptr: zeros) into a
    // device array of n rows: each caller atom on its copy 0, zeros on the
    // other copies of a supercell (their terms are those of the caller's atoms).
    void upload_atom_rows(double *dst, const double *src, int width) {
        const size_t all = (size_t)n * width;
        if (!copies) {
            if (src && all) JNN_HIP(hipMemcpyAsync(dst, src, sizeof(double) * all, hipMemcpyHostToDevice, stream));
            else if (all) JNN_HIP(hipMemsetAsync(dst, 0, sizeof(double) * all, stream));
            return;
        }
        if (!src) {
            if (all) JNN_HIP(hipMemsetAsync(dst, 0, sizeof(double) * all, stream));
            return;
        }
        std::vector<double> h(all, 0.0);  // laid out on the host: one copy for every replica
        for (const RepCopies &r : rcopy)
            std::copy(src + (size_t)r.uoff * width, src + ((size_t)r.uoff + r.unat) * width,
                      h.begin() + (size_t)r.ioff * width);
        if (all) JNN_HIP(hipMemcpyAsync(dst, h.data(), sizeof(double) * all, hipMemcpyHostToDevice, stream));
        JNN_HIP(hipStreamSynchronize(stream));  // h is released on return
    }
    // Per-atom rows of copy 0 of every replica (the caller's atoms) to a host array.
    void download_atom_rows(double *dst, const double *src, int width) {
        if (!copies) {
            if (n) JNN_HIP(hipMemcpyAsync(dst, src, sizeof(double) * n * width, hipMemcpyDeviceToHost, stream));
            return;
        }
        std::vector<double> h((size_t)n * width);  // one copy, then copy 0 of each replica picked on the host
        if (n) JNN_HIP(hipMemcpyAsync(h.data(), src, sizeof(double) * h.size(), hipMemcpyDeviceToHost, stream));
        JNN_HIP(hipStreamSynchronize(stream));
        for (const RepCopies &r : rcopy)
            std::copy(h.begin() + (size_t)r.ioff * width, h.begin() + ((size_t)r.ioff + r.unat) * width,
                      dst + (size_t)r.uoff * width);
    }
    // The caller's row of held atom a, -1 for the other copies of a supercell.
    int user_row(int a) const {
        if (!copies) return a;
        size_t lo = 0, hi = rcopy.size();
        while (hi - lo > 1) {  // the last replica whose first held row is <= a
            const size_t mid = (lo + hi) / 2;
            if (rcopy[mid].ioff <= a) lo = mid;
            else hi = mid;
        }
        const RepCopies &r = rcopy[lo];
        return a - r.ioff < r.unat ? r.uoff + (a - r.ioff) : -1;
    }

    SegmentArgs seg_args(int t, bool raw) const {
        SegmentArgs a = species_segment(model, t, precision, d_types);
        a.nrows = seg[t + 1] - seg[t];
        a.rows_atom = d_rows + seg[t];
        a.nl_of = d_slot_of;  // neighbour rows sit in cell-sorted slot order
        if (raw) {
            a.mu = model->zeros;
            a.sig = model->ones;
            a.mu64 = model->zeros64;
            a.sig64 = model->ones64;
        }
        set_segment_x(a, d_x, (size_t)seg[t], model->smax);
        a.A = d_A + (size_t)seg[t] * model->pmmax;
        a.G = d_G + (size_t)seg[t] * model->smax;
        return a;
    }

    void desc_fwd(bool raw) {
        std::vector<SegmentArgs> segs;
        for (int t = 0; t < model->host.ntypes(); ++t) {
            SegmentArgs a = seg_args(t, raw);
            a.wg_per_cu = tune.desc_fwd_wg_per_cu;
            a.variant = tune.desc_fwd_variant;
            segs.push_back(a);
        }
        enqueue_desc(stream, model, segs.data(), false, nl);
    }

    // the whole evaluation as a sequence of launches on `stream`
    void enqueue(int what, bool timed) {
        build_nl(timed, false);
        const bool grad = (what & (JSE_NNAP_FORCE | JSE_NNAP_VIRIAL)) != 0;
        desc_fwd(false);
        if (timed) JNN_HIP(hipEventRecord(ev[3], stream));
        {
            std::vector<MlpArgs> ma;
            for (int t = 0; t < model->host.ntypes(); ++t) {
                SegmentArgs sa = seg_args(t, false);
                ma.push_back(mlp_args(model, t, sa, d_E, grad, tune.mlp_wg_per_cu));
            }
            enqueue_mlp_all(stream, model, precision, ma.data(), (int)ma.size());
        }
        if (timed) JNN_HIP(hipEventRecord(ev[4], stream));
        if (grad) {
            std::vector<SegmentArgs> segs;
            for (int t = 0; t < model->host.ntypes(); ++t) {
                SegmentArgs a = seg_args(t, false);
                a.wg_per_cu = tune.desc_bwd_wg_per_cu;
                a.variant = tune.desc_bwd_variant;
                segs.push_back(a);
            }
            enqueue_desc(stream, model, segs.data(), true, nl);
        }
        if (timed) JNN_HIP(hipEventRecord(ev[5], stream));
        if (grad)
            launch_gather(stream, n, nullptr, nl, d_F, d_V, (what & JSE_NNAP_VIRIAL) ? 1 : 0, tune.gather_waves, d_sidx);
        launch_replica_totals(stream, nchunk, d_chunks, nrep, d_rep_chunk, d_E, d_V, (what & JSE_NNAP_VIRIAL) ? 1 : 0,
                              d_partial, d_tot, d_arrive);
        if (timed) JNN_HIP(hipEventRecord(ev[6], stream));
        copy_overflow();
    }

    // The neighbour rows by atom: the device rows sit in cell-sorted slot
    // order; perm[i*cap + q] = position of atom i's q-th entry on the device.
    void user_rows(std::vector<int> &cnt_u, std::vector<int> &key_u, std::vector<size_t> &perm) {
        const int c = nl.cap;
        std::vector<int> cnt(n), k((size_t)n * c), slot_h(n);
        if (n) {
            JNN_HIP(hipMemcpyAsync(cnt.data(), nl.cnt, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
            JNN_HIP(hipMemcpyAsync(k.data(), nl.key, sizeof(int) * k.size(), hipMemcpyDeviceToHost, stream));
            JNN_HIP(hipMemcpyAsync(slot_h.data(), d_slot_of, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
        }
        JNN_HIP(hipStreamSynchronize(stream));
        cnt_u.assign(n, 0);
        key_u.assign((size_t)n * c, 0);
        perm.assign((size_t)n * c, 0);
        std::vector<std::pair<int, size_t>> row;
        for (int i = 0; i < n; ++i) {
            const int sl = slot_h[i];
            cnt_u[i] = cnt[sl];
            const int m = std::min(cnt[sl], c);
            row.clear();
            for (int q = 0; q < m; ++q) {
                const size_t src = (size_t)sl * c + q;
                row.emplace_back(k[src], src);
            }
            std::sort(row.begin(), row.end());
            for (int q = 0; q < m; ++q) {
                key_u[(size_t)i * c + q] = row[q].first;
                perm[(size_t)i * c + q] = row[q].second;
            }
        }
    }

    int *tune_field(int k) {
        switch (k) {
            case 0: return &tune.desc_fwd_wg_per_cu;
            case 1: return &tune.desc_bwd_wg_per_cu;
            case 2: return &tune.mlp_wg_per_cu;
            case 3: return &tune.nl_wg_per_cu;
            case 4: return &tune.gather_waves;
            case 5: return &tune.desc_fwd_variant;
            default: return &tune.desc_bwd_variant;
        }
    }
    // Device time per evaluation of `iters` back-to-back eager evaluations.
    double time_eval(int what, int iters) {
        enqueue(what, false);
        JNN_HIP(hipEventRecord(ev[0], stream));
        for (int k = 0; k < iters; ++k) enqueue(what, false);
        JNN_HIP(hipEventRecord(ev[7], stream));
        JNN_HIP(hipEventSynchronize(ev[7]));
        float ms = 0.f;
        JNN_HIP(hipEventElapsedTime(&ms, ev[0], ev[7]));
        return (double)ms / iters;
    }
    // Coordinate descent over the launch-shape knobs, each candidate timed on
    // this system's own data; keeps the fastest shape of every knob.
    double autotune(int what, int iters) {
        JNN_CHECK(have_pos, ERR_ARG, "positions not set");
        JNN_CHECK(iters >= 1, ERR_ARG, "iters must be >= 1");
        static const std::vector<int> cand[kTuneKnobs] = {
            {0, 4, 8, 16, 32}, {0, 4, 8, 16, 32}, {1, 2, 4, 8}, {0, 4, 8, 16}, {1, 2, 4, 8}, {0}, {0, 1, 2}};
        drop_graphs();
        compute(what, true);  // settles the neighbour capacity
        finish();
        double best = time_eval(what, iters);
        for (int k = 0; k < kTuneKnobs; ++k) {
            int *f = tune_field(k);
            int keep = *f;
            for (int v : cand[k]) {
                if (v == keep) continue;
                *f = v;
                const double t = time_eval(what, iters);
                if (t < best) {
                    best = t;
                    keep = v;
                }
            }
            *f = keep;
        }
        drop_graphs();
        return best;
    }

    // One evaluation: replays the captured step graph of this kind when there
    // is one (capturing it on first use), or enqueues the launches directly.
    void compute(int what, bool timed) {
        JNN_CHECK(have_pos, ERR_ARG, "positions not set");
        last_what = what;
        pending = true;
        have_nl = true;
        if (timed || !use_graph || n == 0) {
            enqueue(what, timed);
            return;
        }
        hipGraphExec_t &ge = graph_exec[what & 7];
        if (!ge) {
            hipGraph_t g = nullptr;
            JNN_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            try {
                enqueue(what, false);
            } catch (...) {
                (void)hipStreamEndCapture(stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            JNN_HIP(hipStreamEndCapture(stream, &g));
            hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                ge = nullptr;
                use_graph = false;
                enqueue(what, false);
                return;
            }
        }
        JNN_HIP(hipGraphLaunch(ge, stream));
    }
};

namespace {
// The model's JSON text: the argument itself when it starts with '{', else the
// file it names, read as YAML when its name ends in .yaml / .yml (as NNAP(String)
// chooses, NNAP.java:407-409) and as JSON otherwise.
std::string model_text(const char *json_or_path) {
    std::string text(json_or_path);
    size_t k = text.find_first_not_of(" \t\r\n");
    if (k == std::string::npos || text[k] != '{') {
        const std::string path = text;
        std::ifstream f(path);
        JNN_CHECK(f.good(), ERR_ARG, "cannot open model file: " + path);
        std::stringstream ss;
        ss << f.rdbuf();
        text = ss.str();
        auto ends = [&](const char *suf) {
            const size_t n = std::strlen(suf);
            return path.size() >= n && path.compare(path.size() - n, n, suf) == 0;
        };
        if (ends(".yaml") || ends(".yml")) text = yaml_to_json(text);
    }
    return text;
}
}  // namespace

extern "C" {

const char *jse_nnap_last_error(void) { return g_last_error.c_str(); }

int jse_nnap_model_load(const char *json_or_path, jse_nnap_model *out) {
    return guard([&] {
        JNN_CHECK(json_or_path && out, ERR_ARG, "null argument");
        auto m = std::make_unique<jse_nnap_model_s>();
        m->host = parse_model_json(model_text(json_or_path));
        check_model_shapes(m.get(), JSE_NNAP_FP32_FAST);  // shapes no precision mode can run
        m->upload();
        *out = m.release();
    });
}

int jse_nnap_model_parse(const char *json_or_path, int *ntypes, int *size_np, int *fp_size) {
    return guard([&] {
        JNN_CHECK(json_or_path && ntypes, ERR_ARG, "null argument");
        ModelHost h = parse_model_json(model_text(json_or_path));
        *ntypes = h.ntypes();
        for (int t = 0; t < h.ntypes(); ++t) {
            if (size_np) size_np[t] = h.species[t].P;
            if (fp_size) fp_size[t] = h.species[t].S();
        }
    });
}

int jse_nnap_model_parse_blocks(const char *json_or_path, int ctype, double *fp_param, double *nn_param,
                                double *norm_param, int *nlayers, int *hidden, int *mirror) {
    return guard([&] {
        JNN_CHECK(json_or_path, ERR_ARG, "null argument");
        ModelHost h = parse_model_json(model_text(json_or_path));
        JNN_CHECK(ctype >= 1 && ctype <= h.ntypes(), ERR_ARG, "ctype out of range");
        const SpeciesHost &sp = h.species[ctype - 1];
        if (fp_param) std::copy(sp.W.begin(), sp.W.end(), fp_param);
        if (nn_param) std::copy(sp.nn.begin(), sp.nn.end(), nn_param);
        if (norm_param) std::copy(sp.norm.begin(), sp.norm.end(), norm_param);
        if (nlayers) *nlayers = (int)sp.hidden.size();
        if (hidden) std::copy(sp.hidden.begin(), sp.hidden.end(), hidden);
        if (mirror) *mirror = sp.mirror;
    });
}

int jse_nnap_model_create(int ntypes, const int *nmax, const int *lmax, const int *size_np, const double *rcut,
                          const int *nlayers, const int *hidden, const double *const *fp_param,
                          const double *const *nn_param, const double *const *norm_param, jse_nnap_model *out) {
    return guard([&] {
        JNN_CHECK(out && ntypes >= 1 && nmax && lmax && size_np && rcut && nlayers && hidden && fp_param && nn_param &&
                      norm_param,
                  ERR_ARG, "null argument");
        auto m = std::make_unique<jse_nnap_model_s>();
        for (int t = 0; t < ntypes; ++t) {
            SpeciesHost sp;
            sp.symbol = std::to_string(t + 1);
            sp.nmax = nmax[t];
            sp.lmax = lmax[t];
            sp.P = size_np[t];
            sp.rcut = rcut[t];
            JNN_CHECK(sp.lmax >= 0 && sp.lmax <= 12 && sp.nmax >= 0 && sp.nmax < 32 && sp.P >= 1 && sp.P <= 128 &&
                          sp.rcut > 0 && nlayers[t] >= 1 && nlayers[t] <= kMaxLayers,
                      ERR_ARG, "species shape out of range");
            for (int l = 0; l < nlayers[t]; ++l) sp.hidden.push_back(hidden[t * kMaxLayers + l]);
            const size_t nw = (size_t)ntypes * sp.P * sp.K();
            sp.W.assign(fp_param[t], fp_param[t] + nw);
            size_t nn = 0;
            int c = sp.S();
            for (int h : sp.hidden) { nn += (size_t)c * h + h; c = h; }
            nn += (size_t)c + 1;
            sp.nn.assign(nn_param[t], nn_param[t] + nn);
            sp.norm.assign(norm_param[t], norm_param[t] + 2 + 2 * sp.S());
            m->host.species.push_back(std::move(sp));
        }
        finalize_species(m->host);
        check_model_shapes(m.get(), JSE_NNAP_FP32_FAST);
        m->upload();
        *out = m.release();
    });
}

int jse_nnap_model_info(jse_nnap_model m, int *ntypes, double *rcut_max) {
    return guard([&] {
        JNN_CHECK(m, ERR_ARG, "null model");
        if (ntypes) *ntypes = m->host.ntypes();
        if (rcut_max) *rcut_max = m->host.rcut_max();
    });
}
int jse_nnap_model_fp_size(jse_nnap_model m, int ctype, int *size) {
    return guard([&] {
        JNN_CHECK(m && size, ERR_ARG, "null argument");
        JNN_CHECK(ctype >= 1 && ctype <= m->host.ntypes(), ERR_TYPE, "type out of range");
        *size = m->host.species[ctype - 1].S();
    });
}
void jse_nnap_model_destroy(jse_nnap_model m) { delete m; }

int jse_nnap_sh_invariants(int l3max, int l4max, const double *A, double *out) {
    return guard([&] {
        JNN_CHECK(A && out, ERR_ARG, "null argument");
        const int L = std::max(l3max, l4max);
        const CouplingTerms t = coupling_terms(l3max, l4max, (L + 1) * (L + 1));
        for (int k = 0; k < t.ninv; ++k) {
            double f = 0.0;
            for (int q = t.off[k]; q < t.off[k + 1]; ++q) {
                const int *ix = &t.idx[4 * (size_t)q];
                double v = t.coef[q] * A[ix[0]] * A[ix[1]] * A[ix[2]];
                if (ix[3] >= 0) v *= A[ix[3]];
                f += v;
            }
            out[k] = f;
        }
    });
}

// A system of nrep replicas (natoms[q] atoms of replica q after those of the
// replicas before it, box boxes[3q..3q+2]); a replica with a box edge below
// the cutoff is held as its supercell.
static jse_nnap_system_s *make_system(jse_nnap_model m, int nrep, const int *natoms, const int *types,
                                      const double *boxes, int precision) {
    const double rcut = m->host.rcut_max();
    std::vector<jse_nnap_system_s::RepCopies> rc((size_t)nrep);
    long long nu = 0, ni = 0;
    bool any = false;
    for (int q = 0; q < nrep; ++q) {
        JNN_CHECK(natoms[q] >= 0, ERR_ARG, "negative atom count");
        jse_nnap_system_s::RepCopies &r = rc[q];
        box_copies(boxes + 3 * q, rcut, r.cp);
        const long long nc = (long long)r.cp[0] * r.cp[1] * r.cp[2];
        JNN_CHECK(ni + (long long)natoms[q] * nc < (1 << 26), ERR_ARG,
                  "natoms (times the copies of boxes shorter than the cutoff) out of range [0, 2^26)");
        r.ncopy = (int)nc;
        for (int d = 0; d < 3; ++d) r.box[d] = boxes[3 * q + d];
        r.unat = natoms[q];
        r.uoff = (int)nu;
        r.ioff = (int)ni;
        nu += natoms[q];
        ni += (long long)natoms[q] * nc;
        any = any || nc > 1;
    }
    JNN_CHECK(types || nu == 0, ERR_ARG, "null types");
    auto s = std::make_unique<jse_nnap_system_s>();
    if (!any) {
        s->init(m, nrep, natoms, types, boxes, precision);
    } else {
        std::vector<int> inat((size_t)nrep), ity((size_t)ni);
        std::vector<double> ibox((size_t)3 * nrep);
        for (int q = 0; q < nrep; ++q) {
            const jse_nnap_system_s::RepCopies &r = rc[q];
            inat[q] = r.unat * r.ncopy;
            for (int c = 0; c < r.ncopy; ++c)
                std::copy(types + r.uoff, types + r.uoff + r.unat, ity.begin() + r.ioff + (size_t)c * r.unat);
            for (int d = 0; d < 3; ++d) ibox[3 * q + d] = r.box[d] * r.cp[d];
        }
        s->init(m, nrep, inat.data(), ity.data(), ibox.data(), precision);
    }
    s->rcopy = rc;
    s->copies = any;
    s->nuser = (int)nu;
    return s.release();
}

int jse_nnap_system_create(jse_nnap_model m, int natoms, const int *types, const double *box, int precision,
                           jse_nnap_system *out) {
    return guard([&] {
        JNN_CHECK(m && out && box && (types || natoms == 0), ERR_ARG, "null argument");
        *out = make_system(m, 1, &natoms, types, box, precision);
    });
}

int jse_nnap_batch_create(jse_nnap_model m, int nrep, const int *natoms, const int *types, const double *boxes,
                          int precision, jse_nnap_system *out) {
    return guard([&] {
        JNN_CHECK(m && out && natoms && boxes && nrep >= 1, ERR_ARG, "null argument");
        *out = make_system(m, nrep, natoms, types, boxes, precision);
    });
}

int jse_nnap_system_get_replica_results(jse_nnap_system s, double *energy, double *virial) {
    return guard([&] {
        JNN_CHECK(s, ERR_ARG, "null system");
        JNN_CHECK(s->last_what != 0, ERR_ARG, "nothing computed yet");
        s->finish();
        const bool vir = (s->last_what & JSE_NNAP_VIRIAL) != 0;
        JNN_CHECK(!(virial && !vir), ERR_ARG, "virials were not computed");
        std::vector<double> tot((size_t)10 * s->nrep);
        JNN_HIP(hipMemcpyAsync(tot.data(), s->d_tot, sizeof(double) * tot.size(), hipMemcpyDeviceToHost, s->stream));
        JNN_HIP(hipStreamSynchronize(s->stream));
        for (int q = 0; q < s->nrep; ++q) {
            const double inv = 1.0 / s->rcopy[q].ncopy;  // a supercell's totals per box
            if (energy) energy[q] = tot[(size_t)10 * q] * inv;
            if (virial)
                for (int c = 0; c < 9; ++c) virial[(size_t)9 * q + c] = tot[(size_t)10 * q + 1 + c] * inv;
        }
    });
}

int jse_nnap_system_set_graphs(jse_nnap_system s, int enable) {
    return guard([&] {
        JNN_CHECK(s, ERR_ARG, "null system");
        JNN_HIP(hipStreamSynchronize(s->stream));
        s->drop_graphs();
        s->use_graph = enable != 0;
    });
}
int jse_nnap_system_autotune(jse_nnap_system s, int what, int iters, int *tune, double *ms_step) {
    return guard([&] {
        JNN_CHECK(s, ERR_ARG, "null system");
        const double ms = s->autotune(what, iters);
        if (ms_step) *ms_step = ms;
        if (tune)
            for (int k = 0; k < kTuneKnobs; ++k) tune[k] = *s->tune_field(k);
    });
}
int jse_nnap_system_set_tuning(jse_nnap_system s, const int *tune, int n) {
    return guard([&] {
        JNN_CHECK(s && tune && n >= 0 && n <= kTuneKnobs, ERR_ARG, "bad tuning vector");
        for (int k = 0; k < n; ++k) JNN_CHECK(tune[k] >= 0 && tune[k] <= 64, ERR_ARG, "tuning value out of range");
        if (n > 5) JNN_CHECK(tune[5] <= 2, ERR_ARG, "descriptor forward variant out of range");
        if (n > 6) JNN_CHECK(tune[6] <= 2, ERR_ARG, "descriptor backward variant out of range");
        JNN_HIP(hipStreamSynchronize(s->stream));
        for (int k = 0; k < n; ++k) *s->tune_field(k) = tune[k];
        s->drop_graphs();
    });
}
void jse_nnap_system_destroy(jse_nnap_system s) { delete s; }

int jse_nnap_system_set_box(jse_nnap_system s, const double *box) {
    return guard([&] {
        JNN_CHECK(s && box, ERR_ARG, "null argument");
        std::vector<double> ibox((size_t)3 * s->nrep);
        for (int q = 0; q < s->nrep; ++q) {
            const auto &r = s->rcopy[q];
            int cp[3];
            box_copies(box + 3 * q, s->model->host.rcut_max(), cp);
            JNN_CHECK(cp[0] == r.cp[0] && cp[1] == r.cp[1] && cp[2] == r.cp[2], ERR_UNSUPPORTED,
                      "replica " + std::to_string(q) + ": the new box needs another number of copies of a box "
                      "shorter than the cutoff (a box edge crossed the cutoff): create a new system for it");
            for (int d = 0; d < 3; ++d) ibox[3 * q + d] = box[3 * q + d] * cp[d];
        }
        s->set_boxes(ibox.data());
        for (int q = 0; q < s->nrep; ++q)
            for (int d = 0; d < 3; ++d) s->rcopy[q].box[d] = box[3 * q + d];
    });
}

int jse_nnap_system_set_positions(jse_nnap_system s, const double *xyz) {
    return guard([&] {
        JNN_CHECK(s && (xyz || s->nuser == 0), ERR_ARG, "null argument");
        for (size_t q = 0; q < (size_t)3 * s->nuser; ++q)
            JNN_CHECK(std::isfinite(xyz[q]), ERR_ARG, "non-finite position at atom " + std::to_string(q / 3));
        if (s->copies) {
            std::vector<double> x((size_t)3 * s->n);
            for (const auto &r : s->rcopy)
                for (int c = 0; c < r.ncopy; ++c) {
                    const int cx = c % r.cp[0], cy = (c / r.cp[0]) % r.cp[1], cz = c / (r.cp[0] * r.cp[1]);
                    const double sh[3] = {r.box[0] * cx, r.box[1] * cy, r.box[2] * cz};
                    for (size_t i = 0; i < (size_t)r.unat; ++i)
                        for (int d = 0; d < 3; ++d)
                            x[3 * ((size_t)r.ioff + (size_t)c * r.unat + i) + d] = xyz[3 * ((size_t)r.uoff + i) + d] + sh[d];
                }
            JNN_HIP(hipMemcpyAsync(s->d_xyz, x.data(), sizeof(double) * x.size(), hipMemcpyHostToDevice, s->stream));
            JNN_HIP(hipStreamSynchronize(s->stream));
        } else {
            if (s->n)
                JNN_HIP(hipMemcpyAsync(s->d_xyz, xyz, sizeof(double) * 3 * s->n, hipMemcpyHostToDevice, s->stream));
            JNN_HIP(hipStreamSynchronize(s->stream));
        }
        s->have_pos = true;
        s->have_nl = false;
    });
}
int jse_nnap_system_set_positions_device(jse_nnap_system s, const double *d_xyz) {
    return guard([&] {
        JNN_CHECK(s && (d_xyz || s->nuser == 0), ERR_ARG, "null argument");
        if (s->copies) {
            for (const auto &r : s->rcopy)
                launch_expand_copies(s->stream, r.unat, r.cp, r.box, d_xyz + 3 * (size_t)r.uoff,
                                     s->d_xyz + 3 * (size_t)r.ioff);
        } else if (s->n)
            JNN_HIP(hipMemcpyAsync(s->d_xyz, d_xyz, sizeof(double) * 3 * s->n, hipMemcpyDeviceToDevice, s->stream));
        s->have_pos = true;
        s->have_nl = false;
    });
}

int jse_nnap_system_compute(jse_nnap_system s, int what) {
    return guard([&] {
        JNN_CHECK(s, ERR_ARG, "null system");
        JNN_CHECK((what & ~7) == 0, ERR_ARG, "unknown compute flags");
        s->compute(what | JSE_NNAP_ENERGY, false);
    });
}

int jse_nnap_system_get_results(jse_nnap_system s, double *energy, double *energies, double *forces, double *virial,
                                double *virials) {
    return guard([&] {
        JNN_CHECK(s, ERR_ARG, "null system");
        JNN_CHECK(s->last_what != 0, ERR_ARG, "nothing computed yet");
        s->finish();
        const bool grad = (s->last_what & (JSE_NNAP_FORCE | JSE_NNAP_VIRIAL)) != 0;
        const bool vir = (s->last_what & JSE_NNAP_VIRIAL) != 0;
        JNN_CHECK(!(forces && !grad), ERR_ARG, "forces were not computed");
        JNN_CHECK(!((virial || virials) && !vir), ERR_ARG, "virials were not computed");
        std::vector<double> tot((size_t)10 * s->nrep, 0.0);
        JNN_HIP(hipMemcpyAsync(tot.data(), s->d_tot, sizeof(double) * tot.size(), hipMemcpyDeviceToHost, s->stream));
        // per-atom rows of the caller's atoms (copy 0 of each supercell)
        if (energies) s->download_atom_rows(energies, s->d_E, 1);
        if (forces) s->download_atom_rows(forces, s->d_F, 3);
        if (virials) s->download_atom_rows(virials, s->d_V, 9);
        JNN_HIP(hipStreamSynchronize(s->stream));
        // whole-system totals: replica totals added in replica order (a supercell's per box)
        if (energy) {
            double e = 0.0;
            for (int q = 0; q < s->nrep; ++q) e += tot[(size_t)10 * q] / s->rcopy[q].ncopy;
            *energy = e;
        }
        if (virial)
            for (int c = 0; c < 9; ++c) {
                double v = 0.0;
                for (int q = 0; q < s->nrep; ++q) v += tot[(size_t)10 * q + 1 + c] / s->rcopy[q].ncopy;
                virial[c] = v;
            }
    });
}

int jse_nnap_system_backward_energy(jse_nnap_system s, const double *w, double *const *grad_nn,
                                    double *const *grad_fp) {
    return guard([&] {
        JNN_CHECK(s && grad_nn && (w || s->nuser == 0), ERR_ARG, "null argument");
        const int nt = s->model->host.ntypes();
        for (int t = 0; t < nt; ++t) JNN_CHECK(grad_nn[t] && (!grad_fp || grad_fp[t]), ERR_ARG, "null gradient block");
        // the normalised descriptors of the current positions, as the networks read them
        s->last_what = 0;
        s->pending = true;
        s->build_nl(false);
        s->finish();
        s->have_nl = true;
        s->desc_fwd(false);
        s->tws_next = 0;
        double *d_w = s->tws_get((size_t)s->n);
        try {
            s->upload_atom_rows(d_w, w, 1);
            for (int t = 0; t < nt; ++t) {
                const MlpArgs a = mlp_args(s->model, t, s->seg_args(t, false), s->d_E, false, 0);
                const int np = mlp_param_count(a);
                if (a.nrows <= 0) {
                    std::fill(grad_nn[t], grad_nn[t] + np, 0.0);
                    continue;
                }
                double *partial = s->tws_get((size_t)mlp_pgrad_groups(a) * np);
                double *out = s->tws_get((size_t)np);
                JNN_CHECK(launch_mlp_param_grad(s->stream, a, d_w, partial, out), ERR_UNSUPPORTED,
                          "network too wide for the parameter-gradient kernel (16 rows of its layers in 160 KB)");
                JNN_HIP(hipMemcpyAsync(grad_nn[t], out, sizeof(double) * np, hipMemcpyDeviceToHost, s->stream));
            }
            if (grad_fp) {
                // dE/dfp of every atom (the MLP with gradients), then per basis with a
                // trainable mixing sum_i w_i dE_i/dW; zeros for the other bases
                std::vector<MlpArgs> ma;
                for (int t = 0; t < nt; ++t) ma.push_back(mlp_args(s->model, t, s->seg_args(t, false), s->d_E, true, 0));
                enqueue_mlp_all(s->stream, s->model, s->precision, ma.data(), nt);
                for (int t = 0; t < nt; ++t) {
                    const SpeciesHost &sp = s->model->host.species[t];
                    const SegmentArgs seg = s->seg_args(t, false);
                    size_t goff = 0;
                    for (int k = 0; k < sp.nbases(); ++k) {
                        const SpeciesHost &b = sp.basis(k);
                        const int np = nt * b.P * b.K();
                        if (!b.mix_trainable() || seg.nrows <= 0) {
                            std::fill(grad_fp[t] + goff, grad_fp[t] + goff + np, 0.0);
                        } else {
                            double *partial = s->tws_get((size_t)mix_pgrad_groups(seg.nrows) * np);
                            double *out = s->tws_get((size_t)np);
                            JNN_CHECK(launch_mix_param_grad(s->stream, desc_lmax(b), b.K(), b.P,
                                                            part_segment(seg, s->model, t, k), s->nl, d_w, partial,
                                                            out),
                                      ERR_UNSUPPORTED, "mixing too wide for the parameter-gradient kernel");
                            JNN_HIP(hipMemcpyAsync(grad_fp[t] + goff, out, sizeof(double) * np, hipMemcpyDeviceToHost,
                                                   s->stream));
                        }
                        goff += (size_t)np;
                    }
                }
            }
            JNN_HIP(hipStreamSynchronize(s->stream));
        } catch (...) {
            (void)hipStreamSynchronize(s->stream);
            throw;
        }
    });
}

int jse_nnap_system_backward_energy_force(jse_nnap_system s, const double *w, const double *bgrad_forces,
                                          const double *bgrad_virial, double *const *grad_nn,
                                          double *const *grad_fp) {
    return guard([&] {
        JNN_CHECK(s && grad_nn, ERR_ARG, "null argument");
        const int nt = s->model->host.ntypes();
        for (int t = 0; t < nt; ++t) JNN_CHECK(grad_nn[t] && (!grad_fp || grad_fp[t]), ERR_ARG, "null gradient block");
        jse_nnap_model_s *m = s->model;
        s->last_what = 0;
        s->pending = true;
        s->build_nl(false);
        s->finish();
        s->have_nl = true;
        s->desc_fwd(false);
        const size_t n = (size_t)std::max(s->n, 1);
        s->tws_next = 0;
        auto alloc = [&](size_t cnt) { return s->tws_get(cnt); };
        try {
            double *d_w = alloc(n);
            s->upload_atom_rows(d_w, w, 1);
            TangentArgs base{};
            if (bgrad_forces) {
                double *d_bf = alloc(3 * n);
                s->upload_atom_rows(d_bf, bgrad_forces, 3);
                base.bf = d_bf;
            }
            // a supercell's virial is ncopy times its box's
            if (bgrad_virial && s->nrep > 1) {  // a batch: one dL/dvirial per replica
                std::vector<double> bv((size_t)9 * s->nrep);
                for (int q = 0; q < s->nrep; ++q)
                    for (int c = 0; c < 9; ++c) bv[9 * q + c] = bgrad_virial[9 * q + c] / s->rcopy[q].ncopy;
                double *d_bv = alloc(bv.size());
                JNN_HIP(hipMemcpyAsync(d_bv, bv.data(), sizeof(double) * bv.size(), hipMemcpyHostToDevice,
                                       s->stream));
                base.bv_rep = d_bv;
                base.rep_of = s->d_rep_of;
            } else if (bgrad_virial) {
                for (int q = 0; q < 9; ++q) base.bv[q] = bgrad_virial[q] / s->rcopy[0].ncopy;
            }
            double *d_xdot = alloc(n * m->smax), *d_adot = alloc(n * std::max(m->pmmax, 1));
            double *d_xbar = grad_fp ? alloc(n * m->smax) : nullptr, *d_xdbar = grad_fp ? alloc(n * m->smax) : nullptr;
            // the tangent args of basis k of species t (x slices and A blocks as part_segment narrows them)
            auto tangent = [&](int t, int k, int nrows) {
                const SpeciesHost &sp = m->host.species[t];
                size_t xoff = (size_t)s->seg[t] * m->smax, aoff = (size_t)s->seg[t] * m->pmmax;
                for (int q = 0; q < k; ++q) {
                    xoff += (size_t)sp.basis(q).P * sp.basis(q).SL();
                    aoff += (size_t)nrows * sp.basis(q).P * sp.basis(q).M();
                }
                TangentArgs ta = base;
                ta.xdot = d_xdot + xoff;
                ta.adot = d_adot + aoff;
                ta.xbar = d_xbar ? d_xbar + xoff : nullptr;
                ta.xdbar = d_xdbar ? d_xdbar + xoff : nullptr;
                return ta;
            };
            // D_u x of every atom: each basis' descriptor tangent
            for (int t = 0; t < nt; ++t) {
                const SegmentArgs seg = s->seg_args(t, false);
                if (seg.nrows <= 0) continue;
                const SpeciesHost &sp = m->host.species[t];
                for (int k = 0; k < sp.nbases(); ++k) {
                    const SpeciesHost &b = sp.basis(k);
                    JNN_CHECK(launch_desc_jvp(s->stream, desc_lmax(b), b.K(), b.P, b.mix, part_segment(seg, m, t, k), s->nl,
                                              tangent(t, k, seg.nrows)),
                              ERR_UNSUPPORTED, "basis outside the descriptor tangent kernel's limits");
                }
            }
            // the networks: sum_i (w_i E_i - D_u E_i) through the dual network
            for (int t = 0; t < nt; ++t) {
                const SegmentArgs seg = s->seg_args(t, false);
                const MlpArgs a = mlp_args(m, t, seg, s->d_E, false, 0);
                const int np = mlp_param_count(a);
                if (a.nrows <= 0) {
                    std::fill(grad_nn[t], grad_nn[t] + np, 0.0);
                    continue;
                }
                const TangentArgs ta = tangent(t, 0, seg.nrows);
                double *partial = alloc((size_t)mlp_dual_groups(a) * np), *out = alloc((size_t)np);
                JNN_CHECK(launch_mlp_dual_grad(s->stream, a, d_w, ta.xdot, const_cast<double *>(ta.xbar),
                                               const_cast<double *>(ta.xdbar), partial, out),
                          ERR_UNSUPPORTED, "network too wide for the force-loss gradient kernel (8 rows of its layers in 160 KB)");
                JNN_HIP(hipMemcpyAsync(grad_nn[t], out, sizeof(double) * np, hipMemcpyDeviceToHost, s->stream));
            }
            // the mixing weights
            if (grad_fp) {
                for (int t = 0; t < nt; ++t) {
                    const SpeciesHost &sp = m->host.species[t];
                    const SegmentArgs seg = s->seg_args(t, false);
                    size_t goff = 0;
                    for (int k = 0; k < sp.nbases(); ++k) {
                        const SpeciesHost &b = sp.basis(k);
                        const int np = nt * b.P * b.K();
                        if (!b.mix_trainable() || seg.nrows <= 0) {
                            std::fill(grad_fp[t] + goff, grad_fp[t] + goff + np, 0.0);
                        } else {
                            double *partial = alloc((size_t)mix_pgrad_groups(seg.nrows) * np), *out = alloc((size_t)np);
                            JNN_CHECK(launch_mix_param_grad2(s->stream, desc_lmax(b), b.K(), b.P,
                                                             part_segment(seg, m, t, k), s->nl,
                                                             tangent(t, k, seg.nrows), partial, out),
                                      ERR_UNSUPPORTED, "basis outside the force-loss mixing-gradient kernel's limits");
                            JNN_HIP(hipMemcpyAsync(grad_fp[t] + goff, out, sizeof(double) * np, hipMemcpyDeviceToHost,
                                                   s->stream));
                        }
                        goff += (size_t)np;
                    }
                }
            }
            JNN_HIP(hipStreamSynchronize(s->stream));
        } catch (...) {
            (void)hipStreamSynchronize(s->stream);
            throw;
        }
    });
}

int jse_nnap_system_compute_fp(jse_nnap_system s, double *fp, int smax) {
    return guard([&] {
        JNN_CHECK(s && (fp || s->nuser == 0), ERR_ARG, "null argument");
        JNN_CHECK(smax >= s->model->smax, ERR_ARG, "smax smaller than the largest descriptor");
        s->last_what = 0;
        s->pending = true;
        s->build_nl(false);
        s->finish();
        s->have_nl = true;
        s->desc_fwd(true);
        // FP64 rows in the plain mode, FP32 rows otherwise
        const bool f64 = precision_acc(s->precision);
        std::vector<float> x((size_t)s->n * s->model->smax * (f64 ? 2 : 1));
        std::vector<int> rows(s->n);
        if (s->n) {
            JNN_HIP(hipMemcpyAsync(x.data(), s->d_x, sizeof(float) * x.size(), hipMemcpyDeviceToHost, s->stream));
            JNN_HIP(hipMemcpyAsync(rows.data(), s->d_rows, sizeof(int) * s->n, hipMemcpyDeviceToHost, s->stream));
        }
        JNN_HIP(hipStreamSynchronize(s->stream));
        const double *x64 = reinterpret_cast<const double *>(x.data());
        for (int t = 0; t < s->model->host.ntypes(); ++t) {
            const int S = s->model->host.species[t].S();
            for (int r = s->seg[t]; r < s->seg[t + 1]; ++r) {
                const int u = s->user_row(rows[r]);
                if (u < 0) continue;  // another copy of a supercell
                double *o = fp + (size_t)u * smax;
                const size_t b = (size_t)r * s->model->smax;
                for (int e = 0; e < smax; ++e) o[e] = e < S ? (f64 ? x64[b + e] : (double)x[b + e]) : 0.0;
            }
        }
    });
}

int jse_nnap_system_get_neighbors(jse_nnap_system s, int cap, int *count, int *key, float *d) {
    int rc = guard([&] {
        JNN_CHECK(s && count, ERR_ARG, "null argument");
        JNN_CHECK(!s->copies, ERR_UNSUPPORTED,
                  "neighbour rows of a box shorter than the cutoff name supercell copies: not exported");
        if (!s->have_nl) {
            s->last_what = 0;
            s->pending = true;
            s->build_nl(false);
            s->have_nl = true;
        }
        s->finish();
        const int n = s->n, c = s->nl.cap;
        std::vector<int> cnt, k;
        std::vector<size_t> perm;
        s->user_rows(cnt, k, perm);
        std::vector<float4> dd((size_t)n * c);
        if (n && d) JNN_HIP(hipMemcpyAsync(dd.data(), s->nl.d, sizeof(float4) * dd.size(), hipMemcpyDeviceToHost, s->stream));
        JNN_HIP(hipStreamSynchronize(s->stream));
        bool over = false;
        for (int i = 0; i < n; ++i) {
            count[i] = cnt[i];
            const int m = std::min(std::min(cnt[i], c), cap);
            if (cnt[i] > cap) over = true;
            for (int q = 0; q < m; ++q) {
                const size_t src = (size_t)i * c + q, dst = (size_t)i * cap + q;
                if (key) key[dst] = k[src];
                if (d) {
                    const float4 v = dd[perm[src]];
                    d[3 * dst] = v.x;
                    d[3 * dst + 1] = v.y;
                    d[3 * dst + 2] = v.z;
                }
            }
        }
        if (over) throw Error(ERR_CAPACITY, "a neighbour row is longer than cap");
    });
    return rc;
}

int jse_nnap_system_get_pair_gradients(jse_nnap_system s, int cap, float *g) {
    return guard([&] {
        JNN_CHECK(s && g, ERR_ARG, "null argument");
        JNN_CHECK(!s->copies, ERR_UNSUPPORTED,
                  "pair rows of a box shorter than the cutoff name supercell copies: not exported");
        JNN_CHECK(s->last_what & (JSE_NNAP_FORCE | JSE_NNAP_VIRIAL), ERR_ARG, "forces were not computed");
        s->finish();
        const int n = s->n, c = s->nl.cap;
        std::vector<int> cnt, k;
        std::vector<size_t> perm;
        s->user_rows(cnt, k, perm);
        std::vector<float4> gg((size_t)n * c);
        if (n) JNN_HIP(hipMemcpyAsync(gg.data(), s->nl.g, sizeof(float4) * gg.size(), hipMemcpyDeviceToHost, s->stream));
        JNN_HIP(hipStreamSynchronize(s->stream));
        for (int i = 0; i < n; ++i) {
            const int m = std::min(std::min(cnt[i], c), cap);
            for (int q = 0; q < m; ++q) {
                const size_t src = (size_t)i * c + q, dst = (size_t)i * cap + q;
                const float4 v = gg[perm[src]];
                g[3 * dst] = v.x;
                g[3 * dst + 1] = v.y;
                g[3 * dst + 2] = v.z;
            }
        }
    });
}

int jse_nnap_system_benchmark(jse_nnap_system s, int what, int warmup, int iters, double *ms_step, double *phase_ms,
                              int nphase) {
    return guard([&] {
        JNN_CHECK(s && iters > 0 && warmup >= 0, ERR_ARG, "bad benchmark arguments");
        for (int w = 0; w < warmup; ++w) s->compute(what | JSE_NNAP_ENERGY, false);
        s->finish();
        // pass 1: back-to-back steps, bracketed by two events on the system's stream
        hipEvent_t t0, t1;
        JNN_HIP(hipEventCreate(&t0));
        JNN_HIP(hipEventCreate(&t1));
        JNN_HIP(hipEventRecord(t0, s->stream));
        for (int it = 0; it < iters; ++it) s->compute(what | JSE_NNAP_ENERGY, false);
        JNN_HIP(hipEventRecord(t1, s->stream));
        JNN_HIP(hipEventSynchronize(t1));
        JNN_CHECK(*s->h_overflow == 0, ERR_CAPACITY, "neighbour capacity changed during the benchmark");
        s->pending = false;
        float tot;
        JNN_HIP(hipEventElapsedTime(&tot, t0, t1));
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        if (ms_step) *ms_step = tot / iters;
        // pass 2: per-phase device time, one step at a time
        double acc[6] = {0};
        for (int it = 0; it < iters && phase_ms && nphase > 0; ++it) {
            s->compute(what | JSE_NNAP_ENERGY, true);
            JNN_HIP(hipEventSynchronize(s->ev[6]));
            float ms;
            for (int k = 0; k < 6; ++k) {
                JNN_HIP(hipEventElapsedTime(&ms, s->ev[k], s->ev[k + 1]));
                acc[k] += ms;
            }
        }
        s->finish();
        for (int k = 0; k < nphase && k < 6; ++k) phase_ms[k] = acc[k] / iters;
    });
}

int jse_nnap_system_nl_stats(jse_nnap_system s, int *cap, int *max_count, long long *pairs, long long *unmatched) {
    return guard([&] {
        JNN_CHECK(s, ERR_ARG, "null system");
        if (!s->have_nl) {
            s->last_what = 0;
            s->pending = true;
            s->build_nl(false);
            s->have_nl = true;
        }
        s->finish();
        const int n = s->n, c = s->nl.cap;
        launch_nl_reverse(s->stream, n, nullptr, s->nl, s->d_sidx, s->d_slot_of);
        std::vector<int> cnt(n), rev((size_t)n * c);
        if (n) {
            JNN_HIP(hipMemcpyAsync(cnt.data(), s->nl.cnt, sizeof(int) * n, hipMemcpyDeviceToHost, s->stream));
            JNN_HIP(hipMemcpyAsync(rev.data(), s->nl.rev, sizeof(int) * rev.size(), hipMemcpyDeviceToHost, s->stream));
        }
        JNN_HIP(hipStreamSynchronize(s->stream));
        long long p = 0, u = 0;
        int mx = 0;
        for (int i = 0; i < n; ++i) {
            mx = std::max(mx, cnt[i]);
            const int m = std::min(cnt[i], c);
            p += m;
            for (int q = 0; q < m; ++q)
                if (rev[(size_t)i * c + q] < 0) ++u;
        }
        if (cap) *cap = c;
        if (max_count) *max_count = mx;
        if (pairs) *pairs = p;
        if (unmatched) *unmatched = u;
    });
}

}  // extern "C"
