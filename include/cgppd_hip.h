/*
 * cgppd_hip.h -- C ABI of the gfx950 REMC energy library (libcgppd_hip.so).
 *
 * Plain C types only: opaque context handle, pointers, sizes, int status.
 * Every entry point returns CGPPD_OK (0) or a negative CGPPD_E* code; the
 * message of the last failure is available from cgppd_last_error().
 *
 * One context = one GPU, one HIP stream, one force field, one system
 * topology (molecules and residue types, shared by every replica) and R
 * replicas whose coordinates, temperatures and energies stay resident in
 * HBM.  A context is driven by one host thread at a time.
 *
 * Each declaration names the reference interface it replaces.  Reference
 * paths are relative to the cgppd source tree.
 */
#ifndef CGPPD_HIP_H
#define CGPPD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CGPPD_OK 0
#define CGPPD_E_ARG (-1)      /* invalid argument or call order   */
#define CGPPD_E_HIP (-2)      /* HIP runtime error                */
#define CGPPD_E_NOMEM (-3)    /* device allocation failed         */
#define CGPPD_E_STATE (-4)    /* context not ready for this call  */

#define CGPPD_AA_COUNT 20

typedef struct cgppd_ctx cgppd_ctx;

/* Library / device set-up.  Replaces the per-thread part of setup_CUDA
 * (src/Simulation.cpp:892-926: cudaSetDevice + stream creation) and
 * teardown_CUDA (src/Simulation.cpp:928-948). */
int cgppd_create(cgppd_ctx **out, int device);
int cgppd_destroy(cgppd_ctx *ctx);
const char *cgppd_last_error(const cgppd_ctx *ctx);
int cgppd_device_count(int *count);

/* Force field.  Replaces LJPotentialDataToDevice (src/cudaExterns.h:19),
 * copyLJPotentialDataToDevice (src/cudaExterns.h:20) and bindLJTexture /
 * unbindLJTexture (src/cudaExterns.h:53-54); the table is owned by the
 * context.  eps[20*20] are the raw contact energies e_ij [RT] of
 * AminoAcids::LJpotentials (src/AminoAcids.h:13); sigma[20] and charge[20]
 * are the per-type van der Waals radius and charge (src/AminoAcids.cpp:29-73),
 * i.e. what the reference packs per residue into meta.y/meta.z
 * (src/Replica.cpp:763-765). */
int cgppd_set_force_field(cgppd_ctx *ctx, const float *eps, const float *sigma, const float *charge);

/* Periodic box edge.  Replaces CUDA_setBoxDimension / CUDA_freeBoxDimension
 * (src/cudaExterns.h:47-48).  box > 0. */
int cgppd_set_box(cgppd_ctx *ctx, float box);

/* System topology, shared by all replicas.  Replaces the residueCount /
 * moleculeCount / moleculePositions device arrays and the metadata float4s
 * built in Replica::ReplicaDataToDevice (src/Replica.cpp:705-830).
 *   n_mol, mol_off[n_mol+1] (mol_off[0] = 0, mol_off[n_mol] = N),
 *   type[N] in 0..19, n_noncrowder = number of leading non-crowder molecules
 *   (the reference requires non-crowders first, src/Replica.cpp:924).
 * Resets all replica state. */
int cgppd_set_system(cgppd_ctx *ctx, int n_mol, const int *mol_off, const int *type, int n_noncrowder);

/* Flexible linkers (the reference's FLEXIBLE_LINKS build) ----------------- */

/* Torsion table: v[20*20*4] in kcal/mol and delta[20*20*4] in radians,
 * indexed [type_j][type_k][n-1] for the middle residues j, k of a torsion.
 * Replaces TorsionalLookupMatrix::loadData (src/TorsionalLookupMatrix.cpp:27-52)
 * and its per-molecule copies (src/Molecule.cpp:30). */
int cgppd_set_torsion_table(cgppd_ctx *ctx, const double *v, const double *delta);

/* Flexible topology of one molecule; call after cgppd_set_system (which
 * makes every molecule rigid) and before cgppd_mc_prepare.
 *   chain_code[n_m]   PDB chain identifier of each residue: consecutive
 *                     residues with equal codes are bonded (src/Geometry.cpp:40-46)
 *   chain_index[n_m]  chain number of each residue (TER records before it,
 *                     src/Molecule.cpp:557)
 *   all_flexible      moldata::all_flexible (src/structures.h:67)
 *   n_segments, seg_off[n_segments+1], seg_res[seg_off[n_segments]]
 *                     moldata::segments, residue indices 0-based within the
 *                     molecule (src/Simulation.cpp:350-374 subtracts 1)
 * The library builds the bonded graph, move sets and exclusion UIDs: replaces
 * Graph::init and Graph::assign_uids (src/Geometry.cpp:34-302, called from
 * src/Molecule.cpp:57 and src/Replica.cpp:30-38) and the pos.w / meta.w
 * packing of the exclusion tags (src/Residue.cpp:41-59, src/Replica.cpp:753-775).
 * With all_flexible = 0 and no segments the molecule is rigid.  Energies
 * then include the molecule's intra-molecular pairs that the GPU kernel of
 * the reference counts (src/CudaFunctions.cu:447-476), and MC steps use the
 * flex and local moves (src/Molecule.cpp:274-466). */
int cgppd_set_topology(cgppd_ctx *ctx, int molecule, const int *chain_code, const int *chain_index,
                       int all_flexible, int n_segments, const int *seg_off, const int *seg_res);

/* internal_molecule_E(false) of one replica (src/Replica.cpp:532-560 over
 * src/Molecule.cpp:636-699), kcal/mol: out[5] = LJ and DH of the indirect
 * neighbours (negated), bond, angle, torsion.  cgppd_energy's total plus the
 * sum of out[] is the replica's total energy, as
 * Replica::EonDevice() + internal_molecule_E(false).total() (src/Replica.cpp:362-371). */
int cgppd_energy_internal(cgppd_ctx *ctx, int replica, double *out);

/* Host-only topology query (no GPU, no context): the lists cgppd_set_topology
 * derives for one molecule of n_res residues.
 *   kind 0 bonds (2 ints each), 1 angles (3), 2 torsions (4), 3 local,
 *   4 crankshaft, 5 flex residues, 6 segment bonds (2), 7 indirect
 *   neighbours (2), 8 rigid-domain index of each residue (-1: none),
 *   9 neighbours of each flex residue, 10 their offsets (n_flex + 1),
 *   11 residues each flex move rotates, 12 their offsets (one per
 *   neighbour entry, + 1), 13 segment-bond index of each residue (-1: none),
 *   14 {vertex count, number of chains}.
 * *count receives the number of ints; out may be NULL to size the buffer.
 * New: the reference's Graph members (src/Geometry.h:137-176). */
int cgppd_topology_list(int n_res, const int *chain_code, const int *chain_index, int all_flexible,
                        int n_segments, const int *seg_off, const int *seg_res, int kind, int *out, int *count);

/* Replica state.  Replaces ReplicaDataToDevice (src/Replica.cpp:705-830) for
 * R replicas at once plus Replica::init_child_replica's ladders
 * (src/Replica.cpp:79-98).  Arrays are row-major by replica:
 *   center[R*M*3], rel[R*N*3]   molecule centres and residue positions
 *                               relative to them (position = rel + center),
 *   temperature[R] [K], translate_step[R] [A], rotate_step[R] [rad],
 *   label[R].  Temperature ranks are assigned by ascending temperature.
 * seed keys the counter-based Philox streams of every move and exchange. */
int cgppd_set_replicas(cgppd_ctx *ctx, int n_rep, const float *center, const float *rel,
                       const double *temperature, const float *translate_step,
                       const float *rotate_step, const int *label, uint64_t seed);

/* Overwrite the absolute coordinates of one replica (xyz[N*3]), as
 * Replica::ReplicaDataToDevice copies positions (src/Replica.cpp:739-760).
 * Each molecule's centre becomes the float mean of its residues and the
 * relative positions follow (src/Molecule.cpp:590-609), so MC may continue
 * from the uploaded conformation after cgppd_mc_prepare. */
int cgppd_upload_positions(cgppd_ctx *ctx, int replica, const float *xyz);

/* Overwrite one molecule's coordinates (xyz[n_m*3]), centre and relative
 * positions recomputed as above.  Replaces Replica::MoleculeDataToDevice
 * (src/Replica.cpp:833-881). */
int cgppd_update_molecule(cgppd_ctx *ctx, int replica, int molecule, const float *xyz);

/* Segment image shift (default on): when two molecules' bounding spheres
 * cannot straddle half a box, one minimum-image shift serves all their
 * residue pairs and the per-pair rint is skipped.  0 forces the per-pair
 * minimum image everywhere (as src/Residue.cpp:68-73). */
int cgppd_set_image_shift(cgppd_ctx *ctx, int enable);

/* Repulsive crowding (default off).  Replaces the reference's compile-time
 * switch REPULSIVE_CROWDING (src/definitions.h:18) in Replica::E
 * (src/Replica.cpp:416-483) and the GPU kernel (src/CudaFunctions.cu:335-338,
 * 492-500).  When on, every residue pair with a crowder molecule (index >=
 * n_noncrowder of cgppd_set_system) contributes (6/r)^12 to LJ for r < 15 A
 * and no Debye-Hueckel term; the energies, the MC rows and the non-crowder
 * energy follow.  Crowders must be rigid: CGPPD_E_ARG when a crowder's
 * topology (cgppd_set_topology) is flexible, whichever of the two calls
 * comes first.  Changing the switch invalidates the MC state: MC calls
 * return CGPPD_E_STATE until cgppd_mc_prepare rebuilds the matrix. */
int cgppd_set_repulsive_crowding(cgppd_ctx *ctx, int enable);

/* Rigid-only override (default off).  Replaces the flag
 * Replica::calculate_rigid_potential_only (src/Replica.h:55) that tests set
 * before ReplicaDataToDevice, which then packs every residue of a molecule
 * into one rigid domain (src/Replica.cpp:753-757): with it on, flexible
 * molecules contribute no intra-molecular pairs to cgppd_energy /
 * cgppd_energy_async / cgppd_energy_all, cgppd_energy_internal returns
 * zeros, and MC steps move flexible molecules as usual but evaluate only
 * the inter-molecular energy (src/Replica.cpp:366-380).  Applies to every
 * replica of the context; changing it requires cgppd_mc_prepare again. */
int cgppd_set_rigid_potential_only(cgppd_ctx *ctx, int enable);

/* Full inter-molecular energy of one replica, synchronous.  Replaces
 * CUDA_EonDevice (src/cudaExterns.h:29) behind Replica::EonDevice
 * (src/Replica.cpp:903-918).  lj and dh are the raw sums (Potential::LJ,
 * Potential::DH, src/Potential.h:34-35); total is kcal/mol as
 * Potential::total (src/Potential.cpp:214-221).  Any output may be NULL. */
int cgppd_energy(cgppd_ctx *ctx, int replica, double *lj, double *dh, double *total);

/* Same over the first n_noncrowder molecules only.  Replaces
 * CUDA_EonDeviceNC (src/cudaExterns.h:30) behind Replica::EonDeviceNC
 * (src/Replica.cpp:920-935). */
int cgppd_energy_nc(cgppd_ctx *ctx, int replica, double *lj, double *dh, double *total);

/* The same for every replica in one launch: total[R] in kcal/mol.  New:
 * Replica::sample calls EonDeviceNC per replica and sample
 * (src/Replica.cpp:1037-1093); the C++ Simulation samples all replicas of a
 * block at once. */
int cgppd_energy_nc_all(cgppd_ctx *ctx, double *total);

/* Asynchronous variant: enqueue on the context stream, then wait.  Replaces
 * CUDA_EonDevice_async (src/cudaExterns.h:42) + the D2H copy in
 * Replica::EonDeviceAsync (src/Replica.cpp:883-901) and
 * Replica::SumGridResults (src/Replica.cpp:627-638). */
int cgppd_energy_async(cgppd_ctx *ctx, int replica);
int cgppd_energy_wait(cgppd_ctx *ctx, int replica, double *lj, double *dh, double *total);

/* Full energies of all replicas in one launch (lj, dh, total: R each; any
 * may be NULL) and, if mat != NULL, the molecule-pair matrix of totals
 * (R*M*M, kcal/mol).  New: the reference evaluates one replica per call. */
int cgppd_energy_all(cgppd_ctx *ctx, double *lj, double *dh, double *total, double *mat);

/* Allocation-free variant of cgppd_energy_all.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It allocates nothing and copies nothing to the host: the
 * full energies of all replicas are enqueued on the context stream and the
 * call returns; lj, dh, total (R doubles each) and mat (R*M*M doubles, the
 * symmetric molecule-pair matrices in kcal/mol) are DEVICE pointers owned
 * by the caller, any of them NULL to skip; cgppd_synchronize waits.  The
 * raw LJ/DH sums go through a scratch array of the library on each GPU
 * (2 doubles per replica, at most CGPPD_NOALLOC_MAX_REPLICAS replicas):
 * CGPPD_E_ARG beyond that, CGPPD_E_STATE when the context needs a set-up
 * upload first (a topology set after the replicas: call cgppd_energy or
 * cgppd_mc_prepare once).  Calls that overlap on one GPU (two contexts,
 * two host threads) share that scratch and race. */
#define CGPPD_NOALLOC_MAX_REPLICAS 1024
int cgppd_energy_all_noalloc(cgppd_ctx *ctx, double *lj, double *dh, double *total, double *mat);

/* Recompute the device molecule-pair energy matrix and running totals of
 * every replica from their coordinates.  Required once after
 * cgppd_set_replicas and before cgppd_mc_steps; may be called at any time
 * to resynchronise the running totals. */
int cgppd_mc_prepare(cgppd_ctx *ctx);

/* n_steps MC steps on every replica, all on the device: molecule pick,
 * rigid translate/rotate (src/Molecule.cpp:146-266,468-504), for flexible
 * molecules also flex / local translate / crankshaft (src/Molecule.cpp:274-466),
 * delta-E of the moved molecule against all others (plus its internal
 * energy when flexible), Metropolis (src/Replica.cpp:385-410) and state
 * update.  Replaces Replica::MCSearch (src/Replica.cpp:350-414)
 * and MCSearchMutate/Evaluate/AcceptReject (src/Replica.cpp:643-701). */
int cgppd_mc_steps(cgppd_ctx *ctx, int n_steps);

/* One replica-exchange round over adjacent temperature ranks, alternating
 * even/odd pairs.  Replaces the exchange loop of Simulation::run
 * (src/Simulation.cpp:770-795) and Replica::exchangeReplicas
 * (src/Replica.cpp:188-210). */
int cgppd_exchange(cgppd_ctx *ctx);

/* n_intervals x (steps_per_interval MC steps + one exchange).  With
 * use_graph != 0 one interval is captured once in a hipGraph and replayed.
 * Replaces the MC/exchange rendezvous of MCthreadableFunction
 * (src/Simulation.cpp:1042-1094) and Simulation::run (src/Simulation.cpp:737-816). */
int cgppd_run(cgppd_ctx *ctx, int n_intervals, int steps_per_interval, int use_graph);

/* Block until all work queued on the context stream has finished. */
int cgppd_synchronize(cgppd_ctx *ctx);

/* State read-back (any pointer may be NULL):
 *   energy[R] running totals, temperature[R], label[R], counters[R*3]
 *   (accept, acceptA, reject as src/Replica.h accept/acceptA/reject),
 *   steps[R], exchange[2] (tests, accepted). */
int cgppd_get_replica_stats(cgppd_ctx *ctx, double *energy, double *temperature, int *label,
                            int64_t *counters, int64_t *steps, int64_t *exchange);

/* Inter-molecular residue pairs evaluated by MC steps since
 * cgppd_set_replicas: the sum over steps of n_m * (N - n_m) for the moved
 * molecule m (the delta-E row).  New: the reference counts none. */
int cgppd_get_pair_evaluations(cgppd_ctx *ctx, int64_t *pairs);

/* Coordinates of replica r: pos[N*3], rel[N*3], center[M*3] (any NULL). */
int cgppd_get_coordinates(cgppd_ctx *ctx, int replica, float *pos, float *rel, float *center);

/* Device molecule-pair matrix of replica r (M*M, kcal/mol). */
int cgppd_get_energy_matrix(cgppd_ctx *ctx, int replica, double *mat);

/* Timing and launch shape of the dominant kernel.
 * cgppd_set_launch: workgroups per CU of the persistent pair kernel
 * (0 = automatic).  cgppd_last_kernel_ms: mean duration in ms of the pair
 * kernel launches of the last cgppd_mc_steps / cgppd_run call, measured with
 * HIP events on the context stream (0 when not measured); with the
 * resident layout (CGPPD_LAYOUT_RESIDENT) active, the span of the resident
 * kernel's launches divided by the MC steps they ran, and `launches` counts
 * those launches (one per 16384 steps).  cgppd_set_timing(ctx, 1) enables
 * the measurement. */
int cgppd_set_launch(cgppd_ctx *ctx, int wg_per_cu);

/* Layout of the pair kernel for systems without flexible molecules.
 * Replaces the reference's block-size choice (Replica::init_first_replica,
 * src/Replica.cpp:59-68: 32 threads below 1024 residues, else 64, or -d).
 *   CGPPD_LAYOUT_AUTO   resident MC steps for systems the resident kernel
 *                       takes with at most 16384 residue pairs per replica
 *                       step (averaged over the molecules), else the
 *                       autotuned choice, else j-lanes
 *   CGPPD_LAYOUT_ILANES lanes hold the moved molecule's residues, the other
 *                       residues are staged through LDS in 256-residue chunks
 *   CGPPD_LAYOUT_JLANES lanes hold two of the other residues each, the moved
 *                       molecule's residues are read as wave-uniform
 *                       records; every wave takes an equal share of the work
 *   CGPPD_LAYOUT_JLANES4 the same with four residues per lane
 *   CGPPD_LAYOUT_RESIDENT MC steps in one workgroup per replica with the
 *                       replica held in LDS, many steps per launch
 *                       (cgppd_mc_steps / cgppd_run): for systems of rigid
 *                       molecules without delayed acceptance or repulsive
 *                       crowding, up to 1024 residues and 256 molecules,
 *                       whose replica fits 64 KB of LDS, with one replica
 *                       group; other systems, cgppd_autotune and timed
 *                       launches use the j-lane kernel.  cgppd_get_pair_launch
 *                       reports CGPPD_LAYOUT_RESIDENT when it applies.
 *                       Chosen explicitly for larger systems it is slower
 *                       than the j-lane kernel (one workgroup per replica). 
 * Systems with flexible molecules always use the i-lane kernel. */
#define CGPPD_LAYOUT_AUTO 0
#define CGPPD_LAYOUT_ILANES 1
#define CGPPD_LAYOUT_JLANES 2
#define CGPPD_LAYOUT_JLANES4 3
#define CGPPD_LAYOUT_RESIDENT 4
int cgppd_set_pair_layout(cgppd_ctx *ctx, int layout);
/* Replica groups of the MC steps (default 1; at most 4).  New: the
 * reference runs replicas on host threads, each waiting on its own GPU
 * stream (src/Simulation.cpp:892-926, src/Replica.cpp:110); here the
 * replicas of one context are split into `groups` contiguous groups whose
 * MC steps (pair kernel, then finish kernel, per step) run on streams of
 * their own, so that one group's finish kernel and launch gaps overlap
 * another group's pair kernel.  Exchanges and read-backs wait for every
 * group.  The trajectories do not depend on the grouping beyond the FP32
 * summation order of the rows (the move streams are keyed by replica). */
int cgppd_set_replica_groups(cgppd_ctx *ctx, int groups);
int cgppd_get_replica_groups(cgppd_ctx *ctx, int *groups);

/* The layout and workgroups per CU the next pair-kernel launch uses. */
int cgppd_get_pair_launch(cgppd_ctx *ctx, int *layout, int *wg_per_cu);

/* Speculative windows of MC steps (new; the reference evaluates one MC step
 * per E_TiledKernel launch, src/Replica.cpp:350-414 with
 * src/CudaFunctions.cu:198-277).  `slots` = 0: automatic (4), 1: off, or 2,
 * 4 or 8 steps per window.  A window is a run of consecutive MC steps of
 * one replica that move distinct molecules (it ends before the first
 * repeated molecule and at the end of the cgppd_mc_steps / cgppd_run
 * interval); the delta-E rows of all its proposals and the energies
 * between the proposals are evaluated in one row-kernel launch against the
 * window's starting state, and the decisions are then taken in step order,
 * so the trajectory is that of one step per launch (up to the FP64
 * summation order of the delta-E).  Windows serve systems of rigid
 * molecules on the j-lane layouts without delayed acceptance, repulsive
 * crowding or replica groups; other systems run one step per launch pair.
 * cgppd_get_window reports the slots the next MC steps use (1: none). */
int cgppd_set_window(cgppd_ctx *ctx, int slots);
int cgppd_get_window(cgppd_ctx *ctx, int *slots);

/* Autotuner (replaces the reference's fixed block size, same lines as
 * above): times `launches` pair-kernel launches on the pending proposals
 * (after cgppd_mc_prepare) for every layout and workgroups-per-CU
 * candidate, keeps the fastest for this context and reports it (any
 * pointer may be NULL).  The MC state is unchanged. */
int cgppd_autotune(cgppd_ctx *ctx, int launches, int *layout, int *wg_per_cu, double *best_ms);
/* The last autotune's mean kernel times in ms per MC step of every replica,
 * ms[3*7]: rows j-lanes, j-lanes x4, i-lanes; columns 2..8 workgroups per
 * CU (0 = not timed).  A layout that runs speculative windows is timed per
 * window launch and divided by the mean window length of the next 1000
 * steps. */
int cgppd_autotune_table(cgppd_ctx *ctx, double *ms);
int cgppd_set_timing(cgppd_ctx *ctx, int enable);
int cgppd_last_kernel_ms(cgppd_ctx *ctx, double *mean_ms, int64_t *launches);

/* The model's float constants as this library evaluates them
 * (src/constants.h:24-25 RT_to_kcalmol and DH_CONVERSION_FACTOR,
 * src/constants.h:16 r0_constant); no context needed. */
int cgppd_model_constants(float *rt_to_kcal, float *dh_factor, float *r0_factor);

/* Sizes of the current system: n_mol, n_res, n_rep (any NULL). */
int cgppd_get_sizes(const cgppd_ctx *ctx, int *n_mol, int *n_res, int *n_rep);

/* Learned surrogate of the pair energy (new: no reference counterpart) ---- *
 * A molecule-pair MLP (80 features -> 128 -> 128 -> 1, SiLU) trained and
 * evaluated by the library's BF16 MFMA kernels with FP32 accumulation; it
 * drives stage one of a delayed-acceptance Metropolis test whose stage two
 * uses the exact pair kernel.  Features of a pair I < J: 32 Gaussian RBFs
 * of the minimum-image centre distance (centres 0, 3, ..., 93 A, width 3 A),
 * the unit centre-to-centre vector in each molecule's body frame (2 x 3),
 * each molecule's residue-type fractions (2 x 20), two zeros.  The
 * surrogate of a pair is 0 beyond centre distance r_I + r_J + 15 A (r: the
 * molecules' bounding radii), so only nearby pairs are trained on and
 * evaluated.
 * Parameters (FP32, cgppd_surrogate_param_count of them): W1[80][128], b1[128],
 * W2[128][128], b2[128], w3[128], b3. */
int cgppd_surrogate_param_count(int *count);
int cgppd_surrogate_init(cgppd_ctx *ctx, uint64_t seed);  /* He-normal weights, zero biases */
int cgppd_surrogate_set_params(cgppd_ctx *ctx, const float *params);  /* resets Adam */
int cgppd_surrogate_get_params(cgppd_ctx *ctx, float *params);
/* Predictions y[n] for feature rows x[n*80] (rounded to BF16 on entry). */
int cgppd_surrogate_predict(cgppd_ctx *ctx, const float *x, int n, float *y);
/* Gradient of 0.5 mean((y - clip(t, -15, 15))^2) over the rows x[n*80],
 * t[n] (parameter layout as above) and the loss; parameters unchanged. */
int cgppd_surrogate_gradient(cgppd_ctx *ctx, const float *x, const float *t, int n, float *grad, double *loss);
/* Training samples held on the device: append host rows, append every
 * molecule pair within the surrogate's range of every replica from the
 * current coordinates with its exact energy (kcal/mol, from the
 * molecule-pair matrix of cgppd_mc_prepare and the MC steps), count, clear. */
int cgppd_surrogate_add_samples(cgppd_ctx *ctx, const float *x, const float *t, int n);
int cgppd_surrogate_collect(cgppd_ctx *ctx);
int cgppd_surrogate_sample_count(cgppd_ctx *ctx, int64_t *count);
int cgppd_surrogate_clear_samples(cgppd_ctx *ctx);
/* steps Adam updates (beta 0.9 / 0.999, eps 1e-8) on minibatches of batch
 * samples (a multiple of 32) drawn with replacement from the sample buffer;
 * mean_loss = mean of the minibatch losses. */
int cgppd_surrogate_train(cgppd_ctx *ctx, int steps, int batch, float lr, uint64_t seed, double *mean_loss);
/* Features of every pair I < J of one replica, x[M(M-1)/2 * 80], pairs in
 * row-major (I, J) order, as the device computes them. */
int cgppd_surrogate_features(cgppd_ctx *ctx, int replica, float *x);
/* Delayed acceptance (Christen & Fox 2005) for cgppd_mc_steps / cgppd_run:
 * stage one accepts a proposal with probability min(1, exp(-dS / kT)) from
 * the surrogate's change dS of the moved molecule's pair energies; only
 * stage-one survivors get the exact delta-E row, and stage two accepts with
 * min(1, exp(-(dE - dS) / kT)).  The sampled distribution is the exact one.
 * Needs an initialised surrogate; enabling recomputes its pair matrix. */
int cgppd_set_delayed_acceptance(cgppd_ctx *ctx, int enable);
/* Stage-one rejections of each replica (R) since cgppd_set_replicas. */
int cgppd_get_stage1_rejections(cgppd_ctx *ctx, int64_t *rejections);

#ifdef __cplusplus
}
#endif
#endif
