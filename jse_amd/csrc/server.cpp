// Served evaluation: many independent replicas (MD or forward-flux-sampling
// trajectories of one potential) stay resident in a device arena between
// requests, and requests are evaluated by continuous batching.
//
// * Arena: per-atom slots (positions, reference positions of the last
//   candidate-list build, Verlet candidate rows at rc + skin, rows within rc,
//   energies / forces / virials) and per-replica cell ranges, sized once from
//   the HBM budget.  A replica is admitted on its first request into a free
//   slot range (first fit); when none is free, idle replicas are evicted in
//   least-recently-used order.  An evicted replica is rebuilt on re-admission.
// * Skin rule: a replica's candidate list is rebuilt when some atom moved more
//   than skin / 2 since the last build (max |x - x_build| > skin / 2, checked
//   on the device per request), and on admission.  Otherwise rows within rc
//   are refreshed from the candidates with the current positions.
// * Continuous batching: a worker thread forms a batch from the pending
//   requests (arrival order, at most one request per replica, up to a batch
//   atom budget) as soon as one of two lanes is free; each lane has its own
//   stream, pinned staging and scratch, so the copy-in and evaluation of one
//   batch overlap the other's.  A replica's result depends only on its own
//   atoms and history: rows, reductions and chunked totals never mix
//   replicas, so it is the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "capi_internal.h"
#include "common.h"
#include "kernels.h"

using namespace jnn;

namespace {
using Clock = std::chrono::steady_clock;

// first-fit allocator of index ranges
struct RangeAlloc {
    std::map<long long, long long> free_;  // start -> length
    long long total = 0, used = 0;
    void reset(long long n) {
        free_.clear();
        if (n > 0) free_[0] = n;
        total = n;
        used = 0;
    }
    long long alloc(long long n) {
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n) {
                const long long s = it->first, len = it->second;
                free_.erase(it);
                if (len > n) free_[s + n] = len - n;
                used += n;
                return s;
            }
        return -1;
    }
    void release(long long s, long long n) {
        used -= n;
        auto it = free_.emplace(s, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Request {
    long long ticket = 0;
    int rid = 0, what = 0;
    std::vector<double> xyz;
    double *E = nullptr, *F = nullptr, *V = nullptr;
    Clock::time_point t_arr, t_done;
    int status = OK;
    std::string err;
    bool done = false;
};

struct Replica {
    int natoms = 0;                  // atoms held (the caller's times ncopy)
    // A box edge below rc + skin: the replica is held as its supercell of
    // cp[0] x cp[1] x cp[2] copies of the box (copy c = cx + cp[0] (cy + cp[1] cz)
    // of the caller's atom i is atom c * unat + i at x_i + (cx, cy, cz) * box);
    // copy 0 carries the caller's forces, totals are divided by ncopy.
    int unat = 0, ncopy = 1, cp[3] = {1, 1, 1};
    double ubox[3] = {0.0, 0.0, 0.0};
    std::vector<int> types;
    ReplicaGrid grid{};              // size / cb / box / ncell at rc + skin
    std::vector<int> by_species;     // local atom indices grouped by species (stable)
    std::vector<int> species_off;    // [ntypes + 1] offsets into by_species
    bool resident = false, force = true, in_flight = false;
    long long atom_off = -1, cell_off = -1;
    int pending = 0;
    unsigned long long last_use = 0;
};

struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    std::vector<Request *> reqs;
    std::vector<int> t_off;
    int natoms = 0, nrep = 0;
    // device
    char *d_plan = nullptr;
    double *d_pos = nullptr, *d_Fout = nullptr, *d_partial = nullptr, *d_tot = nullptr;
    int *d_arrive = nullptr;  // per batch replica: chunk arrivals of the totals kernel
    float *d_x = nullptr, *d_A = nullptr, *d_G = nullptr;
    int *d_flags = nullptr;  // [0] candidate-row overflow, [1] row overflow, [2 + b] rebuilt flag of batch replica b
    // pinned host
    char *h_plan = nullptr;
    double *h_pos = nullptr, *h_Fout = nullptr, *h_tot = nullptr;
    int *h_flags = nullptr;
};

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace

struct jse_nnap_server_s {
    jse_nnap_model model = nullptr;
    int precision = JSE_NNAP_FP32;
    double skin = 1.0, rc = 0, rv = 0;
    double budget = 0;
    int max_batch = 0, max_batch_reps = 0, max_replicas = 0;
    int cap_v = 0, cap_w = 0;
    long long slots = 0, cells = 0;
    size_t arena_bytes = 0, lane_bytes = 0, plan_cap = 0;
    int max_chunks = 0;
    int device = 0;

    // arena (device)
    double *x = nullptr, *xb = nullptr, *xwb = nullptr, *xw = nullptr, *sxw = nullptr;
    double *E = nullptr, *F = nullptr, *V = nullptr;
    int *types = nullptr, *rep_of = nullptr, *cell = nullptr, *rank = nullptr, *sidx = nullptr;
    int *cell_cnt = nullptr, *cell_start = nullptr;
    NeighborList vl{}, nl{};
    ReplicaGrid *d_reps = nullptr;
    int *d_need = nullptr;
    RangeAlloc atom_alloc, cell_alloc;

    std::vector<Replica> reps;  // reserved to max_replicas: elements never move
    std::deque<Request *> pending;
    std::unordered_map<long long, Request *> live;
    std::deque<long long> completed;
    Lane lanes[2];
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    bool stop = false;
    long long next_ticket = 1;
    unsigned long long use_clock = 0;
    int grow_v = 0, grow_w = 0;
    std::string fatal;

    // statistics
    long long n_done = 0, n_batches = 0, n_rebuilds = 0, n_admissions = 0, n_evictions = 0, n_growths = 0;
    double atoms_done = 0, batch_atoms = 0;

    // ------------------------------------------------------------------ memory
    size_t slot_bytes() const {
        return sizeof(double) * (15 + 13) + sizeof(int) * 5 + sizeof(int) * (1 + cap_v) +
               (sizeof(int) * 2 + sizeof(float4) * (precision_acc(precision) ? 3 : 2)) * cap_w + sizeof(int);
    }
    void free_arena() {
        dfree(x); dfree(xb); dfree(xwb); dfree(xw); dfree(sxw); dfree(E); dfree(F); dfree(V);
        dfree(types); dfree(rep_of); dfree(cell); dfree(rank); dfree(sidx); dfree(cell_cnt); dfree(cell_start);
        dfree(vl.cnt); dfree(vl.key); dfree(nl.cnt); dfree(nl.key); dfree(nl.d); dfree(nl.rev); dfree(nl.g); dfree(nl.dlo);
    }
    // Sizes the arena from what the budget leaves after lanes and tables.
    void alloc_arena() {
        free_arena();
        const double fixed = (double)lane_bytes + (double)max_replicas * (sizeof(ReplicaGrid) + sizeof(int));
        const double per_slot = (double)slot_bytes() + 2.0 * sizeof(int) / 8.0;  // + cell share (1 cell per 8 slots)
        long long n = (long long)((budget - fixed) / per_slot);
        n = std::min<long long>(n, (1LL << 26) - 1);
        JNN_CHECK(n >= max_batch, ERR_MEMORY,
                  "the HBM budget does not hold one full batch: raise the budget or lower max_batch_atoms");
        slots = n;
        cells = n / 8 + 2LL * max_replicas + 64;
        x = dalloc<double>(3 * n); xb = dalloc<double>(3 * n); xwb = dalloc<double>(3 * n);
        xw = dalloc<double>(3 * n); sxw = dalloc<double>(3 * n);
        E = dalloc<double>(n); F = dalloc<double>(3 * n); V = dalloc<double>(9 * n);
        types = dalloc<int>(n); rep_of = dalloc<int>(n); cell = dalloc<int>(n); rank = dalloc<int>(n);
        sidx = dalloc<int>(n);
        cell_cnt = dalloc<int>(cells); cell_start = dalloc<int>(cells);
        vl.cap = cap_v;
        vl.cnt = dalloc<int>(n);
        vl.key = dalloc<int>((size_t)n * cap_v);
        nl.cap = cap_w;
        nl.cnt = dalloc<int>(n);
        nl.key = dalloc<int>((size_t)n * cap_w);
        nl.d = dalloc<float4>((size_t)n * cap_w);
        nl.g = dalloc<float4>((size_t)n * cap_w);
        if (precision_acc(precision)) nl.dlo = dalloc<float4>((size_t)n * cap_w);
        arena_bytes = (size_t)((double)n * slot_bytes() + (double)cells * 2 * sizeof(int));
        atom_alloc.reset(n);
        cell_alloc.reset(cells);
    }
    size_t plan_bytes(int na, int nr, int nch, int nadm, int ntypes_in) const {
        size_t b = 0;
        b = align_up(b + sizeof(int) * na, 64);           // atoms
        b = align_up(b + sizeof(int) * na, 64);           // rows
        b = align_up(b + sizeof(SrvRep) * nr, 64);        // batch replicas
        b = align_up(b + sizeof(int) * nr, 64);           // replica ids
        b = align_up(b + sizeof(int4) * nch, 64);         // chunks
        b = align_up(b + sizeof(int) * (nr + 1), 64);     // chunk offsets
        b = align_up(b + sizeof(SrvAdmit) * nadm, 64);    // admissions
        b = align_up(b + sizeof(int) * ntypes_in, 64);    // admitted types
        return b;
    }
    void init_lane(Lane &L) {
        JNN_HIP(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
        JNN_HIP(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
        const size_t nb = (size_t)max_batch;
        L.d_plan = dalloc<char>(plan_cap);
        L.d_pos = dalloc<double>(3 * nb);
        L.d_Fout = dalloc<double>(3 * nb);
        L.d_partial = dalloc<double>((size_t)max_chunks * 10);
        L.d_tot = dalloc<double>((size_t)max_batch_reps * 10);
        L.d_arrive = dalloc<int>((size_t)max_batch_reps);
        JNN_HIP(hipMemset(L.d_arrive, 0, sizeof(int) * (size_t)max_batch_reps));
        L.d_x = dalloc<float>(nb * model->smax * 2);  // room for FP64 rows (the plain mode)
        L.d_G = dalloc<float>(nb * model->smax);
        L.d_A = dalloc<float>(nb * model->pmmax);
        L.d_flags = dalloc<int>(2 + (size_t)max_batch_reps);
        L.h_plan = halloc<char>(plan_cap);
        L.h_pos = halloc<double>(3 * nb);
        L.h_Fout = halloc<double>(3 * nb);
        L.h_tot = halloc<double>((size_t)max_batch_reps * 10);
        L.h_flags = halloc<int>(2 + (size_t)max_batch_reps);
    }
    void free_lane(Lane &L) {
        if (L.stream) (void)hipStreamSynchronize(L.stream);
        dfree(L.d_plan); dfree(L.d_pos); dfree(L.d_Fout); dfree(L.d_partial); dfree(L.d_tot); dfree(L.d_arrive);
        dfree(L.d_x); dfree(L.d_G); dfree(L.d_A); dfree(L.d_flags);
        hfree(L.h_plan); hfree(L.h_pos); hfree(L.h_Fout); hfree(L.h_tot); hfree(L.h_flags);
        if (L.done) (void)hipEventDestroy(L.done);
        if (L.stream) (void)hipStreamDestroy(L.stream);
        L.done = nullptr;
        L.stream = nullptr;
    }

    void init(jse_nnap_model m, int prec, double skin_, double budget_, int max_batch_, int max_reps,
              double max_density) {
        JNN_CHECK(m, ERR_ARG, "null model");
        check_precision(prec);
        check_model_shapes(m, prec);
        JNN_CHECK(std::isfinite(skin_) && skin_ >= 0, ERR_ARG, "skin must be finite and >= 0");
        JNN_CHECK(max_batch_ >= 1 && max_batch_ < (1 << 26), ERR_ARG, "max_batch_atoms out of range");
        JNN_CHECK(max_reps >= 1 && max_reps <= (1 << 24), ERR_ARG, "max_replicas out of range");
        JNN_CHECK(std::isfinite(max_density) && max_density > 0, ERR_ARG, "max_density must be positive");
        JNN_HIP(hipGetDevice(&device));
        model = m;
        precision = prec;
        skin = skin_;
        rc = m->host.rcut_max();
        rv = rc + skin;
        max_batch = max_batch_;
        max_replicas = max_reps;
        max_batch_reps = std::min(max_batch_, 8192);
        max_chunks = max_batch_reps + max_batch_ / kReduceChunk + 1;
        const double vol = 4.18879;
        cap_w = (int)std::ceil((1.3 * max_density * vol * rc * rc * rc + 16) / 32.0) * 32;
        cap_v = (int)std::ceil((1.3 * max_density * vol * rv * rv * rv + 16) / 32.0) * 32;
        cap_w = std::min(std::max(cap_w, 32), 512);
        cap_v = std::min(std::max(cap_v, cap_w), 512);
        plan_cap = plan_bytes(max_batch, max_batch_reps, max_chunks, max_batch_reps, max_batch);
        lane_bytes = 2 * (plan_cap + sizeof(double) * 6 * (size_t)max_batch + sizeof(double) * 10 * max_chunks +
                          sizeof(double) * 10 * max_batch_reps +
                          sizeof(float) * (size_t)max_batch * (3 * model->smax + model->pmmax));  // x (FP64 room), G, A
        if (budget_ <= 0) {
            size_t fr = 0, tot = 0;
            JNN_HIP(hipMemGetInfo(&fr, &tot));
            budget_ = 0.5 * (double)fr;
        }
        budget = budget_;
        reps.reserve(max_replicas);
        d_reps = dalloc<ReplicaGrid>(max_replicas);
        d_need = dalloc<int>(max_replicas);
        alloc_arena();
        for (auto &L : lanes) init_lane(L);
        worker = std::thread([this] { run(); });
    }

    ~jse_nnap_server_s() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_work.notify_all();
        if (worker.joinable()) worker.join();
        for (auto &kv : live) delete kv.second;
        for (auto &L : lanes) free_lane(L);
        free_arena();
        dfree(d_reps);
        dfree(d_need);
    }

    // ------------------------------------------------------------- client side
    int add_replica(int natoms, const int *ty, const double *box) {
        const int nt = model->host.ntypes();
        JNN_CHECK(natoms >= 1 && ty && box, ERR_ARG, "a replica needs at least one atom, its types and its box");
        Replica r;
        box_copies(box, rv, r.cp);
        const long long nc = (long long)r.cp[0] * r.cp[1] * r.cp[2];
        JNN_CHECK((long long)natoms * nc <= max_batch, ERR_CAPACITY,
                  "replica (times the copies of a box below rcut + skin) larger than max_batch_atoms (" +
                      std::to_string(max_batch) + ")");
        r.ncopy = (int)nc;
        r.unat = natoms;
        natoms *= r.ncopy;
        r.natoms = natoms;
        for (int k = 0; k < r.unat; ++k)
            JNN_CHECK(ty[k] >= 1 && ty[k] <= nt, ERR_TYPE,
                      "Exist type (" + std::to_string(ty[k]) + ") greater than the input ntypes (" +
                          std::to_string(nt) + ")");
        for (int c = 0; c < r.ncopy; ++c) r.types.insert(r.types.end(), ty, ty + r.unat);
        double sbox[3];
        for (int d = 0; d < 3; ++d) {
            r.ubox[d] = box[d];
            sbox[d] = box[d] * r.cp[d];
        }
        make_grid(sbox, rv, r.grid);
        r.grid.natoms = natoms;
        r.species_off.assign(nt + 1, 0);
        for (int t : r.types) r.species_off[t]++;
        for (int t = 1; t <= nt; ++t) r.species_off[t] += r.species_off[t - 1];
        r.by_species.resize(natoms);
        std::vector<int> fill(r.species_off.begin(), r.species_off.end() - 1);
        for (int k = 0; k < natoms; ++k) r.by_species[fill[r.types[k] - 1]++] = k;
        std::lock_guard<std::mutex> lk(mu);
        JNN_CHECK((int)reps.size() < max_replicas, ERR_CAPACITY, "max_replicas reached");
        JNN_CHECK(natoms <= slots && r.grid.ncell + 1 <= cells, ERR_CAPACITY, "replica larger than the arena");
        reps.push_back(std::move(r));
        return (int)reps.size() - 1;
    }

    long long submit(int rid, const double *xyz, int what, double *e, double *f, double *v) {
        const auto t0 = Clock::now();
        JNN_CHECK(xyz, ERR_ARG, "null positions");
        auto q = std::make_unique<Request>();
        q->t_arr = t0;
        q->what = what;
        q->E = e;
        q->F = f;
        q->V = v;
        std::unique_lock<std::mutex> lk(mu);
        JNN_CHECK(rid >= 0 && rid < (int)reps.size(), ERR_ARG, "unknown replica id");
        JNN_CHECK(fatal.empty(), ERR_HIP, "server stopped after a device error: " + fatal);
        Replica &R = reps[rid];
        lk.unlock();
        const size_t n3 = (size_t)3 * R.unat;
        for (size_t k = 0; k < n3; ++k)
            JNN_CHECK(std::isfinite(xyz[k]), ERR_ARG, "non-finite position in the request");
        q->xyz.assign(xyz, xyz + n3);
        for (int c = 1; c < R.ncopy; ++c) {  // the supercell's other copies
            const int cx = c % R.cp[0], cy = (c / R.cp[0]) % R.cp[1], cz = c / (R.cp[0] * R.cp[1]);
            const double sh[3] = {R.ubox[0] * cx, R.ubox[1] * cy, R.ubox[2] * cz};
            for (size_t k = 0; k < n3; ++k) q->xyz.push_back(xyz[k] + sh[k % 3]);
        }
        q->rid = rid;
        lk.lock();
        q->ticket = next_ticket++;
        Request *p = q.release();
        live[p->ticket] = p;
        pending.push_back(p);
        R.pending++;
        lk.unlock();
        cv_work.notify_one();
        return p->ticket;
    }

    // ------------------------------------------------------------------ worker
    bool any_busy() const { return lanes[0].busy || lanes[1].busy; }

    void evict(Replica &R) {
        atom_alloc.release(R.atom_off, R.natoms);
        cell_alloc.release(R.cell_off, R.grid.ncell + 1);
        R.resident = false;
        R.atom_off = R.cell_off = -1;
        n_evictions++;
    }
    // Places R in the arena, evicting idle replicas (least recently used
    // first, those without pending requests before the others) as needed.
    bool admit(Replica &R) {
        for (;;) {
            const long long a = atom_alloc.alloc(R.natoms);
            const long long c = a >= 0 ? cell_alloc.alloc(R.grid.ncell + 1) : -1;
            if (a >= 0 && c >= 0) {
                R.atom_off = a;
                R.cell_off = c;
                R.resident = true;
                R.force = true;
                n_admissions++;
                return true;
            }
            if (a >= 0) atom_alloc.release(a, R.natoms);
            Replica *victim = nullptr;
            for (auto &C : reps) {
                if (!C.resident || C.in_flight || &C == &R) continue;
                if (!victim || (C.pending == 0) > (victim->pending == 0) ||
                    ((C.pending == 0) == (victim->pending == 0) && C.last_use < victim->last_use))
                    victim = &C;
            }
            if (!victim) return false;
            evict(*victim);
        }
    }

    // Forms and launches one batch on lane L; called with the lock held, which
    // it releases while the host-side plan is built and the work enqueued.
    bool launch(Lane &L, std::unique_lock<std::mutex> &lk) {
        std::vector<Request *> sel;
        std::vector<char> admitted;
        int na = 0;
        for (auto it = pending.begin(); it != pending.end();) {
            Request *q = *it;
            Replica &R = reps[q->rid];
            if (R.in_flight) {
                ++it;
                continue;
            }
            if (na + R.natoms > max_batch || (int)sel.size() >= max_batch_reps) break;
            const bool was = R.resident;
            if (!R.resident && !admit(R)) {
                ++it;
                continue;
            }
            R.in_flight = true;
            sel.push_back(q);
            admitted.push_back(!was);
            na += R.natoms;
            it = pending.erase(it);
        }
        if (sel.empty()) return false;
        std::vector<char> force(sel.size());
        for (size_t b = 0; b < sel.size(); ++b) {
            Replica &R = reps[sel[b]->rid];
            force[b] = R.force;
            R.force = false;
        }
        L.busy = true;
        L.reqs = sel;
        lk.unlock();
        std::string err;
        int code = OK;
        try {
            enqueue(L, sel, admitted, force);
        } catch (const Error &e) {
            err = e.what();
            code = e.code;
        } catch (const std::exception &e) {
            err = e.what();
            code = ERR_ARG;
        }
        lk.lock();
        n_batches++;
        batch_atoms += na;
        if (code != OK) {
            // the stream may hold part of the batch: let it finish before the lane is reused
            (void)hipStreamSynchronize(L.stream);
            L.busy = false;
            for (Request *q : sel) finish_request(q, code, err);
            L.reqs.clear();
            if (code == ERR_HIP) fatal = err;
            cv_done.notify_all();
        }
        return true;
    }

    void enqueue(Lane &L, const std::vector<Request *> &sel, const std::vector<char> &admitted,
                 const std::vector<char> &force) {
        const int nr = (int)sel.size();
        const int nt = model->host.ntypes();
        int na = 0, nadm = 0, ntyp = 0, nch = 0;
        L.t_off.resize(nr);
        for (int b = 0; b < nr; ++b) {
            const Replica &R = reps[sel[b]->rid];
            L.t_off[b] = na;
            na += R.natoms;
            nch += std::max(1, (R.natoms + kReduceChunk - 1) / kReduceChunk);
            if (admitted[b]) {
                nadm++;
                ntyp += R.natoms;
            }
        }
        L.natoms = na;
        L.nrep = nr;
        // plan sections
        char *h = L.h_plan;
        size_t off = 0;
        auto section = [&](size_t bytes) {
            const size_t o = off;
            off = align_up(off + bytes, 64);
            return o;
        };
        const size_t o_atoms = section(sizeof(int) * na), o_rows = section(sizeof(int) * na),
                     o_br = section(sizeof(SrvRep) * nr), o_rl = section(sizeof(int) * nr),
                     o_ch = section(sizeof(int4) * nch), o_rc = section(sizeof(int) * (nr + 1)),
                     o_adm = section(sizeof(SrvAdmit) * nadm), o_ty = section(sizeof(int) * ntyp);
        JNN_CHECK(off <= plan_cap, ERR_CAPACITY, "batch plan exceeds its staging buffer");
        int *atoms = (int *)(h + o_atoms), *rows = (int *)(h + o_rows), *rl = (int *)(h + o_rl);
        SrvRep *br = (SrvRep *)(h + o_br);
        int4 *ch = (int4 *)(h + o_ch);
        int *rch = (int *)(h + o_rc);
        SrvAdmit *adm = (SrvAdmit *)(h + o_adm);
        int *tyin = (int *)(h + o_ty);
        std::vector<int> seg(nt + 1, 0);
        for (int b = 0; b < nr; ++b) {
            const Replica &R = reps[sel[b]->rid];
            for (int t = 0; t < nt; ++t) seg[t + 1] += R.species_off[t + 1] - R.species_off[t];
        }
        for (int t = 0; t < nt; ++t) seg[t + 1] += seg[t];
        std::vector<int> fill(seg.begin(), seg.end() - 1);
        int c = 0, ai = 0, ti = 0;
        for (int b = 0; b < nr; ++b) {
            const Request *q = sel[b];
            const Replica &R = reps[q->rid];
            const int a0 = (int)R.atom_off, n = R.natoms;
            for (int k = 0; k < n; ++k) atoms[L.t_off[b] + k] = a0 + k;
            for (int t = 0; t < nt; ++t)
                for (int s = R.species_off[t]; s < R.species_off[t + 1]; ++s) rows[fill[t]++] = a0 + R.by_species[s];
            SrvRep r{};
            r.rid = q->rid;
            r.atom_off = a0;
            r.natoms = n;
            r.t_off = L.t_off[b];
            r.force = force[b];
            br[b] = r;
            rl[b] = q->rid;
            rch[b] = c;
            // at least one chunk (possibly empty) per replica, so every totals row is written
            for (int c0 = 0; c0 == 0 || c0 < n; c0 += kReduceChunk)
                ch[c++] = make_int4(b, a0 + c0, a0 + std::min(c0 + kReduceChunk, n), 0);
            if (admitted[b]) {
                SrvAdmit A{};
                A.grid = R.grid;
                A.grid.cell_off = (int)R.cell_off;
                A.grid.atom_off = a0;
                A.grid.natoms = n;
                A.rid = q->rid;
                A.types_off = ti;
                adm[ai++] = A;
                std::memcpy(tyin + ti, R.types.data(), sizeof(int) * n);
                ti += n;
            }
            std::memcpy(L.h_pos + 3 * (size_t)L.t_off[b], q->xyz.data(), sizeof(double) * 3 * n);
        }
        rch[nr] = c;
        // device views of the plan
        char *d = L.d_plan;
        const int *d_atoms = (const int *)(d + o_atoms), *d_rows = (const int *)(d + o_rows);
        const SrvRep *d_br = (const SrvRep *)(d + o_br);
        const int *d_rl = (const int *)(d + o_rl);
        const int4 *d_ch = (const int4 *)(d + o_ch);
        const int *d_rch = (const int *)(d + o_rc);
        const SrvAdmit *d_adm = (const SrvAdmit *)(d + o_adm);
        const int *d_ty = (const int *)(d + o_ty);
        hipStream_t s = L.stream;
        JNN_HIP(hipMemcpyAsync(L.d_plan, L.h_plan, off, hipMemcpyHostToDevice, s));
        JNN_HIP(hipMemcpyAsync(L.d_pos, L.h_pos, sizeof(double) * 3 * na, hipMemcpyHostToDevice, s));
        launch_fill_int(s, 2, L.d_flags, 0);
        launch_srv_admit(s, nadm, d_adm, d_ty, d_reps, rep_of, types);
        launch_srv_check(s, nr, d_br, L.d_pos, x, xb, 0.25 * skin * skin, d_need, L.d_flags + 2);
        // candidate lists of the replicas flagged for a rebuild
        launch_srv_zero_cells(s, nr, d_br, d_reps, d_need, cell_cnt);
        launch_wrap_bin(s, na, d_atoms, x, rep_of, d_reps, d_need, xw, cell, cell_cnt, rank);
        launch_cell_scan(s, nr, d_rl, d_reps, d_need, cell_cnt, cell_start);
        launch_cell_scatter(s, na, d_atoms, rep_of, d_need, cell, rank, cell_start, xw, sidx, sxw);
        vl.types = types;
        nl.types = types;
        launch_nl_build(s, na, d_atoms, rep_of, d_reps, d_need, rv * rv, xw, cell, cell_start, sidx, sxw, vl,
                        L.d_flags);
        // rows within rc at the current positions
        launch_srv_update(s, na, d_atoms, rep_of, d_need, x, xb, xwb, xw);
        launch_srv_refresh(s, na, d_atoms, rep_of, d_reps, rc * rc, xw, vl, nl, L.d_flags + 1);
        // descriptor, MLP, backward over the species segments of the batch rows
        {
            std::vector<SegmentArgs> segs;
            std::vector<MlpArgs> ma;
            for (int t = 0; t < nt; ++t) {
                segs.push_back(seg_args(L, t, seg, d_rows));
                const SegmentArgs &a = segs.back();
                ma.push_back(mlp_args(model, t, a, E, true));
            }
            enqueue_desc(s, model, segs.data(), false, nl);
            enqueue_mlp_all(s, model, precision, ma.data(), nt);
            enqueue_desc(s, model, segs.data(), true, nl);
        }
        launch_gather(s, na, d_atoms, nl, F, V, 1);
        launch_replica_totals(s, c, d_ch, nr, d_rch, E, V, 1, L.d_partial, L.d_tot, L.d_arrive);
        launch_srv_pack(s, na, d_atoms, F, L.d_Fout);
        JNN_HIP(hipMemcpyAsync(L.h_Fout, L.d_Fout, sizeof(double) * 3 * na, hipMemcpyDeviceToHost, s));
        JNN_HIP(hipMemcpyAsync(L.h_tot, L.d_tot, sizeof(double) * 10 * nr, hipMemcpyDeviceToHost, s));
        JNN_HIP(hipMemcpyAsync(L.h_flags, L.d_flags, sizeof(int) * (2 + nr), hipMemcpyDeviceToHost, s));
        JNN_HIP(hipGetLastError());
        JNN_HIP(hipEventRecord(L.done, s));
    }

    SegmentArgs seg_args(const Lane &L, int t, const std::vector<int> &seg, const int *d_rows) const {
        SegmentArgs a = species_segment(model, t, precision, types);
        a.nrows = seg[t + 1] - seg[t];
        a.rows_atom = d_rows + seg[t];
        set_segment_x(a, L.d_x, (size_t)seg[t], model->smax);
        a.A = L.d_A + (size_t)seg[t] * model->pmmax;
        a.G = L.d_G + (size_t)seg[t] * model->smax;
        return a;
    }

    void finish_request(Request *q, int code, const std::string &err) {
        Replica &R = reps[q->rid];
        R.in_flight = false;
        R.pending--;
        R.last_use = ++use_clock;
        q->status = code;
        q->err = err;
        q->t_done = Clock::now();
        q->done = true;
        completed.push_back(q->ticket);
        n_done++;
    }

    // lane finished on the device (lock held)
    void complete(Lane &L) {
        L.busy = false;
        const int ov = L.h_flags[0], ow = L.h_flags[1];
        if (ov > 0 || ow > 0) {
            // a row did not fit: redo these requests after the capacities grow
            grow_v = std::max(grow_v, ov);
            grow_w = std::max(grow_w, ow);
            for (auto it = L.reqs.rbegin(); it != L.reqs.rend(); ++it) {
                Replica &R = reps[(*it)->rid];
                R.in_flight = false;
                R.force = true;
                pending.push_front(*it);
            }
            L.reqs.clear();
            return;
        }
        for (int b = 0; b < L.nrep; ++b) {
            Request *q = L.reqs[b];
            const double *tot = L.h_tot + 10 * (size_t)b;
            const Replica &R = reps[q->rid];
            const double inv = 1.0 / R.ncopy;  // a supercell's totals per box
            if (q->E) *q->E = tot[0] * inv;
            if (q->V && (q->what & JSE_NNAP_VIRIAL))
                for (int c = 0; c < 9; ++c) q->V[c] = tot[1 + c] * inv;
            if (q->F && (q->what & JSE_NNAP_FORCE))  // copy 0: the replica's first unat rows
                std::memcpy(q->F, L.h_Fout + 3 * (size_t)L.t_off[b], sizeof(double) * 3 * R.unat);
            n_rebuilds += L.h_flags[2 + b];
            atoms_done += R.unat;
            finish_request(q, OK, "");
        }
        L.reqs.clear();
        cv_done.notify_all();
    }

    void regrow() {
        const int nv = grow_v > 0 ? std::max(cap_v, (int)std::ceil((grow_v * 1.2 + 16) / 32.0) * 32) : cap_v;
        const int nw = grow_w > 0 ? std::max(cap_w, (int)std::ceil((grow_w * 1.2 + 16) / 32.0) * 32) : cap_w;
        grow_v = grow_w = 0;
        n_growths++;
        for (auto &R : reps)
            if (R.resident) evict(R);
        if (nv > 512 || nw > 512) {
            fatal = "an atom has more than 512 neighbours within rcut + skin; at most 512 are supported";
            fail_all(ERR_CAPACITY, fatal);
            return;
        }
        cap_v = std::max(nv, nw);
        cap_w = nw;
        try {
            alloc_arena();
        } catch (const Error &e) {
            fatal = e.what();
            fail_all(e.code, fatal);
        }
    }

    void fail_all(int code, const std::string &err) {
        while (!pending.empty()) {
            Request *q = pending.front();
            pending.pop_front();
            finish_request(q, code, err);
        }
        cv_done.notify_all();
    }

    void run() {
        if (hipSetDevice(device) != hipSuccess) {
            std::lock_guard<std::mutex> lk(mu);
            fatal = "the worker thread cannot select the server's device";
        }
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            bool progressed = false;
            for (auto &L : lanes) {
                if (!L.busy) continue;
                const hipError_t q = hipEventQuery(L.done);
                if (q == hipSuccess) {
                    complete(L);
                    progressed = true;
                } else if (q != hipErrorNotReady) {
                    L.busy = false;
                    fatal = std::string("device error: ") + hipGetErrorString(q);
                    for (Request *r : L.reqs) finish_request(r, ERR_HIP, fatal);
                    L.reqs.clear();
                    fail_all(ERR_HIP, fatal);
                    progressed = true;
                }
            }
            if ((grow_v || grow_w) && !any_busy()) {
                regrow();
                progressed = true;
            }
            if (!fatal.empty() && !pending.empty()) fail_all(ERR_HIP, fatal);
            if (stop && !any_busy()) {
                fail_all(ERR_ARG, "server destroyed before the request ran");
                break;
            }
            if (!stop && !grow_v && !grow_w && fatal.empty())
                for (auto &L : lanes)
                    if (!L.busy && !pending.empty() && launch(L, lk)) progressed = true;
            if (progressed) continue;
            if (any_busy()) {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
                lk.lock();
            } else {
                cv_work.wait_for(lk, std::chrono::milliseconds(2));
            }
        }
        cv_done.notify_all();
    }

    // --------------------------------------------------------------- results
    int wait(long long ticket, double *latency) {
        std::unique_lock<std::mutex> lk(mu);
        auto it = live.find(ticket);
        JNN_CHECK(it != live.end(), ERR_ARG, "unknown or already collected ticket");
        Request *q = it->second;
        cv_done.wait(lk, [&] { return q->done; });
        if (latency) *latency = std::chrono::duration<double>(q->t_done - q->t_arr).count();
        const int st = q->status;
        const std::string err = q->err;
        live.erase(it);
        for (auto c = completed.begin(); c != completed.end(); ++c)
            if (*c == ticket) {
                completed.erase(c);
                break;
            }
        delete q;
        if (st != OK) throw Error(st, err);
        return st;
    }

    int collect(long long *tickets, double *lat, int *status, int cap) {
        std::lock_guard<std::mutex> lk(mu);
        int n = 0;
        while (n < cap && !completed.empty()) {
            const long long t = completed.front();
            completed.pop_front();
            auto it = live.find(t);
            if (it == live.end()) continue;
            Request *q = it->second;
            tickets[n] = t;
            if (lat) lat[n] = std::chrono::duration<double>(q->t_done - q->t_arr).count();
            if (status) status[n] = q->status;
            live.erase(it);
            delete q;
            ++n;
        }
        return n;
    }

    void drain() {
        std::unique_lock<std::mutex> lk(mu);
        cv_done.wait(lk, [&] { return pending.empty() && !any_busy(); });
    }

    void stats(double *out, int n) {
        std::lock_guard<std::mutex> lk(mu);
        long long res_r = 0, res_a = 0;
        for (auto &R : reps)
            if (R.resident) {
                res_r++;
                res_a += R.natoms;
            }
        const double v[] = {(double)n_done, atoms_done, (double)n_batches,
                            n_batches ? batch_atoms / n_batches : 0.0, (double)n_rebuilds, (double)n_admissions,
                            (double)n_evictions, (double)n_growths, (double)res_r, (double)res_a, (double)slots,
                            (double)arena_bytes, (double)cap_v, (double)cap_w, (double)pending.size(),
                            (double)(arena_bytes + lane_bytes)};
        const int nv = (int)(sizeof(v) / sizeof(v[0]));
        for (int k = 0; k < n && k < nv; ++k) out[k] = v[k];
    }
};

extern "C" {

int jse_nnap_server_create(jse_nnap_model model, int precision, double skin, double hbm_budget_bytes,
                           int max_batch_atoms, int max_replicas, double max_density, jse_nnap_server *out) {
    return guard([&] {
        JNN_CHECK(out, ERR_ARG, "null argument");
        auto s = std::make_unique<jse_nnap_server_s>();
        s->init(model, precision, skin, hbm_budget_bytes, max_batch_atoms, max_replicas, max_density);
        *out = s.release();
    });
}

int jse_nnap_server_add_replica(jse_nnap_server s, int natoms, const int *types, const double *box,
                                int *replica_id) {
    return guard([&] {
        JNN_CHECK(s && replica_id, ERR_ARG, "null argument");
        *replica_id = s->add_replica(natoms, types, box);
    });
}

int jse_nnap_server_submit(jse_nnap_server s, int replica_id, const double *xyz, int what, double *energy,
                           double *forces, double *virial, long long *ticket) {
    return guard([&] {
        JNN_CHECK(s && ticket, ERR_ARG, "null argument");
        *ticket = s->submit(replica_id, xyz, what, energy, forces, virial);
    });
}

int jse_nnap_server_wait(jse_nnap_server s, long long ticket, double *latency_s) {
    return guard([&] {
        JNN_CHECK(s, ERR_ARG, "null server");
        s->wait(ticket, latency_s);
    });
}

int jse_nnap_server_collect(jse_nnap_server s, long long *tickets, double *latency_s, int *status, int cap,
                            int *count) {
    return guard([&] {
        JNN_CHECK(s && tickets && count && cap >= 0, ERR_ARG, "null argument");
        *count = s->collect(tickets, latency_s, status, cap);
    });
}

int jse_nnap_server_drain(jse_nnap_server s) {
    return guard([&] {
        JNN_CHECK(s, ERR_ARG, "null server");
        s->drain();
        JNN_CHECK(s->fatal.empty(), ERR_HIP, s->fatal);
    });
}

int jse_nnap_server_stats(jse_nnap_server s, double *out, int n) {
    return guard([&] {
        JNN_CHECK(s && out, ERR_ARG, "null argument");
        s->stats(out, n);
    });
}

void jse_nnap_server_destroy(jse_nnap_server s) { delete s; }

}  // extern "C"
