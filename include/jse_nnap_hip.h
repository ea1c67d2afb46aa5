/*
 * libjsenn_hip: MI355X (gfx950) implementation of jse's NNAP evaluation path.
 *
 * C ABI: plain pointers and sizes, opaque handles, int status (0 = ok) and a
 * thread-local message from jse_nnap_last_error().  Every reference location
 * below is a path relative to the root of the jse repository.
 *
 * Status codes: 0 ok, 1 invalid argument/model, 2 unsupported model shape,
 * 3 HIP runtime error, 4 neighbour capacity, 5 atom type out of range,
 * 6 device memory budget exhausted.  (The reference turns a non-zero native
 * return into IllegalStateException("Exit code: n"), NNAP.java:756,769,787.)
 */
#ifndef JSE_NNAP_HIP_H
#define JSE_NNAP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct jse_nnap_model_s *jse_nnap_model;
typedef struct jse_nnap_system_s *jse_nnap_system;

/* precision of a system (the reference's CUDA path is FP32 only, NNAP.java:145-147).
 * JSE_NNAP_FP32, the plain path, meets 1e-6 eV/atom, 1e-5 eV/A and 1e-5 eV
 * against the reference's FP64 CPU path: the displacement, the pair geometry,
 * the harmonics and radial basis, the contraction A (FP64 matrix cores), the
 * power spectrum, the MLP and the chain rule are FP64, as are the sums; only
 * the backward's dE/dA contraction rounds its inputs to FP32.  The other two
 * modes evaluate the geometry and the descriptor in FP32 (errors near 3e-5
 * eV/A on well conditioned models, more where sigma << mu in the
 * normalisation). */
#define JSE_NNAP_FP32 0       /* plain: FP64 geometry/descriptor/MLP/sums, FP32 dE/dA contraction */
#define JSE_NNAP_BF16_MLP 1   /* FP32 geometry and descriptor, MLP on BF16 MFMA with FP32 accumulation */
#define JSE_NNAP_FP32_FAST 2  /* FP32 geometry and descriptor, FP32 MLP */

/* what jse_nnap_system_compute evaluates */
#define JSE_NNAP_ENERGY 1
#define JSE_NNAP_FORCE 2
#define JSE_NNAP_VIRIAL 4

const char *jse_nnap_last_error(void);

/* ---- model ---------------------------------------------------------------
 * Replaces `new NNAP(String path)` / `new NNAP(Map)` (NNAP.java:402-414,
 * model parsing :129-307).  `json_or_path` is the JSON text itself (a text
 * starting with '{') or a path to a model file: YAML when the name ends in
 * .yaml / .yml (as NNAP(String) chooses, NNAP.java:407-409; block and flow
 * collections, plain and quoted scalars, comments; anchors, tags, block
 * scalars and multi-line scalars are refused with status 1 naming the line),
 * JSON otherwise. */
int jse_nnap_model_load(const char *json_or_path, jse_nnap_model *out);

/* Host-only check of a model file: parses it as jse_nnap_model_load does and
 * reports ntypes and, per species (arrays of >= ntypes entries, may be NULL),
 * the radial channel count P and the descriptor length S.  No device use. */
int jse_nnap_model_parse(const char *json_or_path, int *ntypes, int *size_np, int *fp_size);

/* Host-only: the packed parameter blocks of species `ctype` (1-based) as the
 * loader builds them from a model file, i.e. the contents of the reference's
 * mFpParam / mNnParam / mNormParam entries for that species (NNAP.java:
 * 226-307), in the layouts of jse_nnap_model_create below; hidden[] receives
 * nlayers widths and *mirror the mirrored species (0: none).  Any output may
 * be NULL; the arrays must hold the sizes jse_nnap_model_parse implies. */
int jse_nnap_model_parse_blocks(const char *json_or_path, int ctype, double *fp_param, double *nn_param,
                                double *norm_param, int *nlayers, int *hidden, int *mirror);

/* Host-only: the L3 then L4 invariants of one radial channel's A row, as
 * calSphL3 / calSphL4 form them (basis_SphericalUtil.hpp:395-467), from the
 * coupling terms the descriptor kernels use.  A holds (L+1)^2 values,
 * L = max(l3max, l4max), at index l^2+l+m; out receives 0, 0, 2, 4, 9, 14, 23
 * values for l3max 0..6 followed by 0, 1, 3, 9 for l4max 0..3.  No device use. */
int jse_nnap_sh_invariants(int l3max, int l4max, const double *A, double *out);

/* Replaces the packed parameter blocks the reference hands to its native
 * code (NNAP.java:226-307): per species t (0-based here):
 *   fp_param[t]   = W[ntypes][size_np[t]][nmax[t]+1]      (WTypeBasis.java:202-304)
 *   nn_param[t]   = [hidden W | output W | hidden b | output b] (FeedForward.java:310-316)
 *   norm_param[t] = [muE+refE, sigmaE, mu[S], sigma[S]], S = size_np*(lmax+1)
 *   hidden[t*8 + k] = width of hidden layer k (nlayers[t] <= 8)
 * rcut[t] is the basis hyper-parameter (MergeableBasis.java:24-30). */
int jse_nnap_model_create(int ntypes, const int *nmax, const int *lmax, const int *size_np, const double *rcut,
                          const int *nlayers, const int *hidden, const double *const *fp_param,
                          const double *const *nn_param, const double *const *norm_param, jse_nnap_model *out);
/* ntypes, max rcut (NNAP.rcutMax, NNAP.java:485) and the descriptor length of species t (1-based) */
int jse_nnap_model_info(jse_nnap_model m, int *ntypes, double *rcut_max);
int jse_nnap_model_fp_size(jse_nnap_model m, int ctype, int *size);
void jse_nnap_model_destroy(jse_nnap_model m);

/* ---- system: one periodic configuration, device resident ------------------
 * Replaces the AtomData -> AtomicParameterCalculator step
 * (IPairPotential.calEnergyForceVirials(IAtomData, ...), IPairPotential.java:1137-1141;
 * AtomicParameterCalculator.java:130-157): natoms, 1-based model types (after
 * the symbol type map, IHasSymbol.java:51-99) and an orthogonal box
 * [0,Lx) x [0,Ly) x [0,Lz).  Types are fixed for the life of the system.
 * A box edge shorter than the largest cutoff (the reference's ExpandLinkedCell,
 * NeighborListGetter.java:704-776) is evaluated as the box's supercell whose
 * edges reach the cutoff (so in a batch, replica by replica); per-atom outputs
 * are the caller's atoms', totals per box.  Such a system refuses neighbour and
 * pair-gradient exports (status 2), and set_box returns status 2 when a new box
 * needs another supercell. */
int jse_nnap_system_create(jse_nnap_model m, int natoms, const int *types, const double *box, int precision,
                           jse_nnap_system *out);
void jse_nnap_system_destroy(jse_nnap_system s);

/* A batch of independent periodic replicas evaluated together (no reference
 * counterpart: the reference evaluates one AtomData per call; its
 * forward-flux-sampling driver runs many replicas, ForwardFluxSampling.java:
 * 717-873).  natoms[r] atoms of replica r follow each other in `types` and in
 * the position array; boxes[r*3 + d] is replica r's box.  Every per-atom
 * output keeps this concatenated order; jse_nnap_system_get_replica_results
 * gives each replica's totals, which do not depend on the other replicas in
 * the batch. */
int jse_nnap_batch_create(jse_nnap_model m, int nrep, const int *natoms, const int *types, const double *boxes,
                          int precision, jse_nnap_system *out);
int jse_nnap_system_get_replica_results(jse_nnap_system s, double *energy, double *virial);
/* Replays each evaluation kind from a captured HIP graph (default on); 0 launches kernel by kernel. */
int jse_nnap_system_set_graphs(jse_nnap_system s, int enable);
/* box of every replica, [nrep][3] */
int jse_nnap_system_set_box(jse_nnap_system s, const double *box);
/* positions as an [natoms][3] FP64 row-major host array (the reference's
 * getMatRow layout), copied to the device; wrapping happens on the device */
int jse_nnap_system_set_positions(jse_nnap_system s, const double *xyz);
/* the same from a device pointer (no host round trip) */
int jse_nnap_system_set_positions_device(jse_nnap_system s, const double *d_xyz);

/* Replaces NNAP.calEnergy / NNAP.calEnergyForceVirial driven by
 * IPairPotential.calEnergyForceVirials (NNAP.java:531-594,
 * IPairPotential.java:811-1022): builds the cell list, evaluates every atom
 * and accumulates energies, forces and virials.  Asynchronous on the
 * system's stream; results are read with jse_nnap_system_get_results. */
int jse_nnap_system_compute(jse_nnap_system s, int what);

/* Copies results to host arrays; any pointer may be NULL.
 *   energy: total E;  energies[natoms];  forces[natoms][3];
 *   virial[9] totals and virials[natoms][9] per atom on the neighbour side,
 *   component order xx yy zz xy xz yz yx zx zy (IPairPotential.java:972-1002).
 * Stress follows as -virial/volume (IPotential.java:197-311). */
int jse_nnap_system_get_results(jse_nnap_system s, double *energy, double *energies, double *forces, double *virial,
                                double *virials);

/* Per-atom descriptors fp[natoms][Smax] (Smax = max descriptor length over
 * species) of the last neighbour list: jse_nnap_calFp for every atom
 * (nnap_interface.cpp:11-24).  Runs the neighbour build itself. */
int jse_nnap_system_compute_fp(jse_nnap_system s, double *fp, int smax);

/* Training derivatives: the reference's jse_nnap_backwardEnergy,
 * nnap_interface.cpp:111-134, summed over a system.  The reference's trainer calls it per atom with
 * aGradEng = w[i] and adds rGradNnParam / rGradFpParam; one call here covers
 * every atom at the current positions (FP64 sums in a fixed order):
 *   grad_nn[t] = sum over the atoms i of species t of w[i] dE_i/dtheta, in the
 *     packed nn_param layout [hidden W per layer [out][in] | output W | hidden b
 *     | output b];
 *   grad_fp[t] (may be NULL: not computed) = the same for the mixing weights
 *     W[ntypes][P][K] of each of the species' bases in order (the merge parts'
 *     blocks one after another): sum_i w[i] dE_i/dW for a fuse / rfuse / exfuse
 *     mixing, zeros for the other wtypes, which have no parameters
 *     (basis_SphericalChebyshev.hpp:353-440, basis_Chebyshev.hpp:135-220).
 * A mirror or shared species reports the gradient of its own blocks (a mirror's
 * mixing with its neighbour species swapped, as it evaluates them).  Runs the
 * neighbour build, descriptors and MLP itself; status 2 when a network's
 * 16-row tile of layers exceeds 160 KB of LDS. */
int jse_nnap_system_backward_energy(jse_nnap_system s, const double *w, double *const *grad_nn,
                                    double *const *grad_fp);

/* Training derivatives of the force / stress loss: the reference's
 * jse_nnap_forwardEnergyForce then jse_nnap_backwardEnergyForce,
 * nnap_interface.cpp:141-232, for every atom of a system, as TrainerNNAP drives
 * them (TrainerNNAP.java:1102-1176).  The loss is given by its gradients:
 * w[natoms] by the per-atom energies (aGradEng; NULL: zeros), bgrad_forces
 * [natoms][3] by the forces (NULL: zeros) and bgrad_virial[9] by the total
 * virial, order xx yy zz xy xz yz yx zx zy (NULL: zeros; for a batch of
 * jse_nnap_batch_create, [nrep][9], replica r's by its own virial); the trainer's
 * per-pair upstream is then bgrad_forces[j] - bgrad_forces[i] + sum_a
 * bgrad_virial[ab] d_a for pair (i, j) with displacement d.  grad_nn and
 * grad_fp receive, in the layouts of jse_nnap_system_backward_energy, the
 * gradient of sum_i w_i E_i + sum_k bgrad_forces[k] . F_k + sum bgrad_virial .
 * virial by the network and mixing parameters (second derivatives: the
 * descriptor's directional derivative along the pairs' upstream, the networks
 * carried with that tangent).  FP64 sums in a fixed order.  Status 2 for a
 * network whose 8-row tile of layers exceeds 160 KB of LDS, or a trainable
 * mixing with 16 P (lmax+1)^2 bytes over it. */
int jse_nnap_system_backward_energy_force(jse_nnap_system s, const double *w, const double *bgrad_forces,
                                          const double *bgrad_virial, double *const *grad_nn,
                                          double *const *grad_fp);

/* The neighbour list of the last compute: count[natoms] and, for rows of
 * `cap` slots, key = j*32 + shift_code (shift_code = (sx+1)+3(sy+1)+9(sz+1))
 * sorted ascending and the FP32 displacement d[natoms][cap][3].  Returns
 * status 4 if a row does not fit in `cap` (count[] is still written). */
int jse_nnap_system_get_neighbors(jse_nnap_system s, int cap, int *count, int *key, float *d);

/* Per-pair gradients of the last compute with forces: g[natoms][cap][3] =
 * -dE_i/d(r_j - r_i) in the row order of jse_nnap_system_get_neighbors, the
 * quantity the reference's jse_nnap_calEnergyForce returns in
 * rGradNlDx/Dy/Dz (nnap_interface.cpp:48-89). */
int jse_nnap_system_get_pair_gradients(jse_nnap_system s, int cap, float *g);

/* Timing of `iters` compute() calls after `warmup` untimed ones, with HIP
 * events on the system's stream.  ms_step = device time per call of `iters`
 * back-to-back calls; phase_ms[k] = average device time of phase k per call
 * from a second pass of single calls, k = 0 neighbour build, 1 (empty: the
 * force gather finds reverse pairs itself), 2 descriptor forward, 3 MLP, 4 descriptor backward,
 * 5 forces/virials/totals (nphase <= 6). */
int jse_nnap_system_benchmark(jse_nnap_system s, int what, int warmup, int iters, double *ms_step, double *phase_ms,
                              int nphase);

/* Neighbour-list statistics of the last compute: current row capacity, the
 * largest row and the number of pairs without a reverse partner. */
int jse_nnap_system_nl_stats(jse_nnap_system s, int *cap, int *max_count, long long *pairs, long long *unmatched);

/* Autotuner of a system's launch shapes (no reference counterpart: the
 * reference's CUDA path uses one fixed block size, NNAPGEN's
 * __NNAPGEN_CUDA_BLOCKSIZE__, nnap_interface_cuda.cu:456).  Times `iters`
 * evaluations of kind `what` per candidate on the system's current data and
 * keeps, knob by knob, the fastest: tune[0] descriptor-forward and tune[1]
 * descriptor-backward persistent workgroups per CU (0 = one per atom),
 * tune[2] MLP workgroups per CU, tune[3] neighbour-build workgroups per CU
 * (0 = one per atom), tune[4] force-gather atoms per workgroup, tune[5] the
 * descriptor forward of the FP32 precision mode (0: matrix cores, FP32 sums
 * of 16 neighbours added into FP64; 1 / 2: vector kernel with 6x7 / 9x7 FP64
 * register tiles; the BF16-MLP mode runs the FP32-accumulating matrix-core
 * forward whatever this knob says), tune[6] the
 * descriptor-backward schedule (scheduling barrier every 1 / 2 / no rows).
 * Writes the chosen knobs (7 ints) and the tuned device time per evaluation. */
int jse_nnap_system_autotune(jse_nnap_system s, int what, int iters, int *tune, double *ms_step);
/* Applies a tuning vector (e.g. a cached autotuner result); n <= 7 knobs. */
int jse_nnap_system_set_tuning(jse_nnap_system s, const int *tune, int n);

/* ---- server: device-resident replicas, continuous batching ----------------
 * No reference counterpart as one call: the reference evaluates one AtomData
 * per call, and its forward-flux-sampling driver and LAMMPS generators step
 * many independent replicas (ForwardFluxSampling.java:717-873,
 * lmp/NativeLmpFullPathGenerator.java:40-188), each step an energy/force
 * evaluation (NNAP.java:531-594).  A server keeps every replica's positions,
 * Verlet candidate list (rcut + skin) and outputs on the device between
 * requests, admits replicas into an arena sized from an HBM budget (evicting
 * idle ones, least recently used first), and evaluates pending requests in
 * batches formed as lanes free up, two batches in flight on their own streams.
 * A replica's candidate list is rebuilt when an atom has moved more than
 * skin/2 since the last build, and on admission. */
typedef struct jse_nnap_server_s *jse_nnap_server;

/* hbm_budget_bytes <= 0 takes half of the free device memory; max_density
 * (atoms/A^3) sizes the neighbour rows (they grow, with a full re-admission,
 * if exceeded). */
int jse_nnap_server_create(jse_nnap_model model, int precision, double skin, double hbm_budget_bytes,
                           int max_batch_atoms, int max_replicas, double max_density, jse_nnap_server *out);
/* Registers a replica (host-side only; fixed types and orthogonal box).  A
 * box edge below rcut + skin is held as the box's supercell: results are the
 * caller's atoms' forces and per-box totals. */
int jse_nnap_server_add_replica(jse_nnap_server s, int natoms, const int *types, const double *box,
                                int *replica_id);
/* Queues an evaluation of replica `replica_id` at positions xyz[natoms][3]
 * (copied before return).  On completion the total energy goes to *energy,
 * forces[natoms][3] (if what & JSE_NNAP_FORCE) and the total virial[9] (if
 * what & JSE_NNAP_VIRIAL) are written; these buffers must stay valid until
 * the request is waited for or collected.  Requests of one replica run in
 * submission order. */
int jse_nnap_server_submit(jse_nnap_server s, int replica_id, const double *xyz, int what, double *energy,
                           double *forces, double *virial, long long *ticket);
/* Blocks until the request is done; returns its status and its latency
 * (submission to completion, seconds).  Releases the ticket. */
int jse_nnap_server_wait(jse_nnap_server s, long long ticket, double *latency_s);
/* Releases up to `cap` completed requests: tickets, latencies and statuses. */
int jse_nnap_server_collect(jse_nnap_server s, long long *tickets, double *latency_s, int *status, int cap,
                            int *count);
/* Blocks until no request is pending or running. */
int jse_nnap_server_drain(jse_nnap_server s);
/* out[0..n): completed requests, completed atom-evaluations, batches, mean
 * batch atoms, candidate-list rebuilds, admissions, evictions, capacity
 * growths, resident replicas, resident atoms, arena slots, arena bytes,
 * candidate row capacity, row capacity, pending requests, device bytes. */
int jse_nnap_server_stats(jse_nnap_server s, double *out, int n);
/* Finishes running batches, fails queued requests, frees the server. */
void jse_nnap_server_destroy(jse_nnap_server s);

/* ---- LAMMPS GPU pair hook, jse-compatible packed arguments ----------------
 * jse binds native code as `int name(void *args)`, the declared parameters
 * packed back to back in declaration order with no padding (pointers 8 bytes,
 * int 4, double 8, float 4; SourceScanner.java:167-290) and loaded with
 * dlopen/dlsym (jse_jit_SimpleJIT.c:58-119).  These four replace the
 * reference's generated functions of the same names
 * (nnap_interface_cuda.cu:307-483; driver NNAP.computeLammpsCuda,
 * NNAP.java:949-1030), with flt_t = float and the same argument lists:
 *   jse_nnap_statNeiNumLammps (int *ilist, int *numneigh, int inum, int *numneighMax)
 *   jse_nnap_lammps2cuda      (int inum, int nlocalghost, int nlflag, int neighnumMax, double **x,
 *                              int *type, int *ilist, int *numneigh, int **firstneigh, float *fltBuf,
 *                              int *intBuf, float *cudaX, int *cudaType, int *cudaIlist,
 *                              int *cudaNumneigh, int *cudaFirstneigh)
 *   jse_nnap_computeLammpsCuda(int inum, int nlocalghost, int eflag, int vflag, int eflagAtom,
 *                              int vflagAtom, int cvflagAtom, float *x, int *type, int *ilist,
 *                              int *nmerges, int **mergeSorted, float **cutsq, int *numneigh,
 *                              int *firstneigh, int *aLmpType2NNAPType, float **aFpHyperParam,
 *                              float **aFpParam, float **aNnParam, float **aNormParam, float *f,
 *                              float *eatom0, float *vatom0, float *vatom1, + 10 scratch pointers)
 *   jse_nnap_cuda2lammps      (int inum, int nlocalghost, int eflag, int vflag, int eflagAtom,
 *                              int vflagAtom, int cvflagAtom, double **f, double *engVdwl,
 *                              double *eatom, double *virial, double **vatom, double **cvatom,
 *                              float *fltBuf, const int *ilist, float *cudaF, float *cudaEatom0,
 *                              float *cudaVatom0, float *cudaVatom1)
 * The reference compiles the model's shapes into its generated library; here
 * jse_nnap_lammps_bind (no reference counterpart) names the model once:
 * packed (jse_nnap_model model, int precision).  The per-call parameter
 * blocks, merge tables and scratch pointers are then not read; displacements
 * are formed from FP64 positions and forces / virials accumulated in FP64.
 * Return 0 on success, a status code otherwise. */
int jse_nnap_lammps_bind(void *args);
int jse_nnap_statNeiNumLammps(void *args);
int jse_nnap_lammps2cuda(void *args);
int jse_nnap_computeLammpsCuda(void *args);
int jse_nnap_cuda2lammps(void *args);

/* ---- GPUMD hook, jse-compatible packed arguments --------------------------
 * Replaces the reference's generated jse_nnap_computeGPUMD
 * (nnap_interface_cuda.cu:485-518; driver NNAP.computeGPUMD,
 * NNAP.java:1047-1083), same argument list with flt_t = float:
 *   (int number_of_particles, int N1, int N2, const int *g_neighbor_number,
 *    const int *g_neighbor_list, const float *nl_dx, const float *nl_dy,
 *    const float *nl_dz, const int *g_type, const int *nmerges,
 *    const int **mergeSorted, const float **cutsq, float **aFpHyperParam,
 *    float **aFpParam, float **aNnParam, float **aNormParam, double *g_fx,
 *    double *g_fy, double *g_fz, double *g_virial, double *g_potential,
 *    + 9 scratch pointers)
 * All arrays are GPUMD's device arrays: the neighbour list column-major
 * [jj * number_of_particles + i] with FP32 displacements r_j - r_i, 0-based
 * types; centres are particles N1 .. N2-1, and forces, per-atom virials
 * ([9][number_of_particles], on the neighbour) and potentials are added into.
 * jse_nnap_gpumd_bind (no reference counterpart) names the model once:
 * packed (jse_nnap_model model, int precision); the parameter blocks, merge
 * tables and scratch pointers are then not read.  Returns 0 on success. */
int jse_nnap_gpumd_bind(void *args);
int jse_nnap_computeGPUMD(void *args);

/* No reference counterpart: jse_nnap_computeGPUMD for callers that cannot
 * allocate (tight MD loops, harnesses that forbid hipMalloc).  Same packed
 * arguments and results; every buffer beyond GPUMD's inputs and outputs is
 * carved from a static 64 MiB scratch array per GPU (held until the process
 * ends), and the call makes no hipMalloc, hipFree or host allocation.  A call
 * whose buffers exceed the array returns status 6 naming the limit (about
 * 8.7 KB per centre for the benchmark model with rows of 64 neighbours, i.e.
 * some 7,000 particles).  Calls on one GPU share its array and are serialised
 * inside one process; calls from several processes on one GPU each have
 * their own.  Runs on the null stream. */
int jse_nnap_computeGPUMD_noalloc(void *args);

#ifdef __cplusplus
}
#endif
#endif /* JSE_NNAP_HIP_H */
