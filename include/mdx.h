/*
 * mdx.h -- C-ABI of the MI355X (gfx950) molecular-dynamics engine.
 *
 * The reference (vakuras/moleculardynamics) has no FFI: its cpumd step loop
 * calls free C++ functions linked into the executable.  Each entry point
 * below replaces one of them, with the state kept resident in HBM between
 * calls instead of host arrays:
 *
 *   reference function (file:line)                        entry point here
 *   -----------------------------------------------------  --------------------------
 *   hostMain allocation          cpu/cpumd/Host.cpp:43-48   mdx_create / mdx_destroy
 *   readInput / writeOutput       common/Loader.cpp:88-155   mdx_set_state / mdx_get_state
 *   readyList                     cpu/cpumd/Host.cpp:245-271 mdx_ready_list
 *   buildList/buildNeighborList   cpu/cpumd/Host.cpp:276-291,
 *                                 common/NeighborList.cpp:14-156
 *                                                            mdx_build_list
 *   (the padded int list itself)  common/NeighborList.cpp:88-145
 *                                                            mdx_copy_neighbor_list
 *   predict                       cpu/common/Device.cpp:17-39 mdx_predict
 *   lennardJonesForces            cpu/common/LennardJones.h:35-93
 *                                                            mdx_lj_forces
 *   manyBodyForces                cpu/common/BondOrderFuncs.h:283-397
 *                                                            mdx_bond_forces
 *   correct                       cpu/common/Device.cpp:44-120 mdx_correct
 *   calculateAccelerations        cpu/common/Device.cpp:125-134 mdx_accelerations
 *   the step loop                 cpu/cpumd/Host.cpp:141-195 mdx_run
 *   PerformCalculations           common/DebugCalculations.cpp:46-124
 *                                                            mdx_kinetics
 *   calculatePotentional          common/DebugCalculations.cpp:129-270
 *                                                            mdx_potential_allpairs
 *   the vxcm sweep                cpu/cpumd/Host.cpp:75-216  mdx_sweep_*
 *
 * Arrays crossing the boundary use the reference's host layout: positions as
 * real4 (x, y, z, species code in w) and v/a/b/c/f as real3, all double
 * (the shipped cpumd is built with real = double), in the caller's atom
 * order.  Every entry point returns MDX_OK (0) or a negative MDX_ERR_* code
 * in place of the reference's print-and-exit (common/Global.h:116-117);
 * mdx_last_error() gives the message.  A system is single-threaded: calls on
 * one system must not overlap.  Different systems are independent, each with
 * its own HIP stream, and may be driven from different host threads at once.
 * The members of a sweep (mdx_sweep_*) share the sweep's stream: calls on one
 * sweep and on its members must not overlap.
 */
#ifndef MDX_H
#define MDX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDX_OK 0
#define MDX_ERR_ARG (-1)       /* invalid argument or configuration */
#define MDX_ERR_HIP (-2)       /* HIP runtime error (no device, launch failure) */
#define MDX_ERR_ALLOC (-3)     /* device or host allocation failed */
#define MDX_ERR_GRID (-4)      /* cell grid exceeds its limit */
#define MDX_ERR_CAPACITY (-5)  /* a per-atom list exceeded its capacity */
#define MDX_ERR_STATE (-6)     /* call out of order (e.g. forces before a list build) */
#define MDX_ERR_LIST_STALE (-7) /* an atom moved farther than the list margin allows */

/* box boundary */
#define MDX_BOX_WALLED 0    /* reference walls at +-wall (cpu/common/Device.cpp:86-118) */
#define MDX_BOX_PERIODIC 1  /* cubic periodic box of edge box_length */

/* arithmetic of the pair and bond-order terms; accumulation and integration
 * are FP64 in both */
#define MDX_PREC_MIXED 0    /* FP32 pair / bond-order arithmetic */
#define MDX_PREC_FP64 1     /* FP64 everywhere (the GPU oracle run) */

/* Lennard-Jones pair table */
#define MDX_LJ_REFERENCE 0  /* N/O/Ne rule of cpu/common/LennardJones.h:63-88, reference units */
#define MDX_LJ_UNIFORM 1    /* one species: lj_sigma, lj_eps, mass */

/* bond-order functional form */
#define MDX_BOND_REFERENCE 0  /* the reference's N/O potential, cpu/common/BondOrderFuncs.h */
#define MDX_BOND_TERSOFF 1    /* Tersoff silicon, PRB 38, 9902 (1988), R = 2.7 A, D = 0.3 A */
#define MDX_BOND_SURROGATE 2  /* learned surrogate of MDX_BOND_TERSOFF, network set by mdx_set_surrogate */

/* surrogate network: a flat float blob, in order W1[64][48], b1[64], W2[64][64],
 * b2[64], w3[64], b3, then per descriptor component mean[48], istd[48] and the
 * accepted range lo[48], hi[48] of the standardized value (G - mean) * istd */
#define MDX_SG_NFEAT 48
#define MDX_SG_NTRAIN 7361   /* trainable part: W1 .. b3 */
#define MDX_SG_NPARAM 7553   /* whole blob */

/* list selector for mdx_build_list / mdx_copy_neighbor_list */
#define MDX_LIST_LJ 0
#define MDX_LIST_BOND 1

typedef struct mdx_system mdx_system;

/*
 * Configuration.  The reference's 27 keys (common/Loader.h:25-51) map as:
 * n_atoms = Lennard Jones Particles + Many Body Particles, n_bond = Many Body
 * Particles, use_lj/use_bond = UseLennardJones/UseManyBody, dt = DT,
 * lj_rs/lj_rcut/mb_rs/mb_rcut = LennardJonesRS/RCUT, ManyBodyRS/RCUT,
 * energy_loss = sqrt(EnergyLoss) (cpu/cpumd/Host.cpp:73), wall = the box size
 * N^(1/3) * 52.8 (Host.cpp:31,166).  The remaining fields select the periodic
 * benchmark systems that the reference does not have.
 */
typedef struct mdx_config {
    int32_t n_atoms;
    int32_t n_bond;        /* the first n_bond atoms carry the bond-order potential */
    int32_t box;           /* MDX_BOX_* */
    int32_t precision;     /* MDX_PREC_* */
    int32_t use_lj;
    int32_t use_bond;
    int32_t lj_table;      /* MDX_LJ_* */
    int32_t bond_form;     /* MDX_BOND_* */
    float dt;
    float lj_rs, lj_rcut;  /* list radius (cell edge) and cutoff */
    float mb_rs, mb_rcut;
    float energy_loss;     /* +x wall restitution factor, already square-rooted */
    double wall;           /* walled box: walls at +-wall */
    double box_length;     /* periodic box edge */
    double lj_sigma, lj_eps, mass;  /* MDX_LJ_UNIFORM species */
    double accel_scale;    /* a = F / m * accel_scale (1 in reference and reduced units) */
    int32_t max_neighbors; /* capacity of compact per-atom lists (Tersoff); 0 = default 32 */
    int32_t reserved[7];
} mdx_config;

/* lifetime (cpu/cpumd/Host.cpp:43-48, 218-237) */
int mdx_create(const mdx_config* cfg, mdx_system** out);
int mdx_destroy(mdx_system* sys);
const char* mdx_last_error(const mdx_system* sys);  /* sys NULL: the calling thread's last
                                                       creation error */

/* state transfer (common/Loader.cpp:88-155 layout).  pos4: n*4 doubles (x, y,
 * z, species), vel3/acc3/b3/c3: n*3 doubles; acc3/b3/c3 may be NULL (zero). */
int mdx_set_state(mdx_system* sys, const double* pos4, const double* vel3,
                  const double* acc3, const double* b3, const double* c3);
/* any output pointer may be NULL; force3 is the last computed force */
int mdx_get_state(mdx_system* sys, double* pos4, double* vel3, double* acc3,
                  double* b3, double* c3, double* force3);

/* Forces (n*3 doubles, caller order) into the device state, replacing the
 * last computed force.  The reference hands forceArray to correct and
 * calculateAccelerations (cpu/common/Device.cpp:44,125) and, when
 * Lennard-Jones is on, manyBodyForces adds to it
 * (cpu/common/BondOrderFuncs.h:283-397); callers that keep forces on the host
 * (the reference-signature adapter) upload them here. */
int mdx_set_forces(mdx_system* sys, const double* force3);

/* cpu/cpumd/Host.cpp:245-271: vmax = max |v|, box_size_list = 2 max |coordinate| */
int mdx_ready_list(mdx_system* sys, double* vmax, double* box_size_list);

/* common/NeighborList.cpp:14-156 on the device: bins atoms into cells of edge
 * rs and counting-sorts them into cell order.  Walled box: the reference's
 * grid from box_size_list; periodic box: floor(L/rs) cells per axis
 * (box_size_list ignored).  Reports the reference's nlistsize (min(27 * max
 * cell occupancy, n - 1)) and nlargestsize (longest row). */
int mdx_build_list(mdx_system* sys, int32_t which, double box_size_list,
                   int32_t* nlistsize, int32_t* nlargestsize);

/* The last built list of `which` in the reference's padded format: rows of
 * nlistsize ints in atom order, 27 cells in (dx,dy,dz) lexicographic order,
 * ascending index within a cell, self excluded, -1 padding.  capacity is in
 * ints and must be at least n * nlistsize (n = n_atoms or n_bond). */
int mdx_copy_neighbor_list(mdx_system* sys, int32_t which, int32_t* out, int64_t capacity);

/* The pair lists the periodic force kernels read, as caller atom indices:
 * the atoms found within the list radius (lj_rs; mb_rs for Tersoff and its
 * surrogate) at the last list build, which mdx_run makes.  Row i of out
 * (width ints; -1 after counts[i] entries) holds atom i's partners in the
 * list's own order.  *width receives the longest row; with out NULL nothing
 * else is written (the call that sizes out).  capacity is in ints, at least
 * n * width.  build_pos3 (may be NULL): the positions the lists were built
 * from, 3 doubles per atom.  The reference's list
 * (common/NeighborList.cpp:88-145) is the padded candidate list of
 * mdx_copy_neighbor_list, which its force loops filter by distance
 * (cpu/common/LennardJones.h:35-93); here the filtered list is what the
 * kernels consume, so this export is what parity checks as a set. */
int mdx_copy_pair_list(mdx_system* sys, int32_t* out, int64_t capacity, int32_t* counts, int32_t* width,
                       double* build_pos3);

/* cpu/common/Device.cpp:17-39 */
int mdx_predict(mdx_system* sys);
/* cpu/common/LennardJones.h:35-93: assigns the force */
int mdx_lj_forces(mdx_system* sys);
/* cpu/common/BondOrderFuncs.h:283-397: adds to the force when use_lj, else assigns */
int mdx_bond_forces(mdx_system* sys);
/* cpu/common/Device.cpp:44-120; *wall_hit = 1 when any wall acted (may be NULL) */
int mdx_correct(mdx_system* sys, int32_t* wall_hit);
/* cpu/common/Device.cpp:125-134 */
int mdx_accelerations(mdx_system* sys);

/* The step loop of cpu/cpumd/Host.cpp:141-195 for nsteps steps, device
 * resident.  Walled box: the reference's order and list schedule
 * (next build = t + (int)(((rs - rc)/2)/vmax/|dt|), forces recomputed after a
 * wall hit).  Periodic box: corrector and next predictor fused, lists rebuilt
 * when the largest displacement since the last build reaches (rs - rc)/2.
 * The first call computes the initial forces and a = F/m (Host.cpp:132-138). */
int mdx_run(mdx_system* sys, int32_t nsteps);
/* number of list builds done by mdx_run so far (LJ/periodic grid) */
int mdx_build_count(mdx_system* sys, int64_t* builds);

/* common/DebugCalculations.cpp:46-124 as FP64 device reductions.
 * out8 = {EK, T, COMx, COMy, COMz, Px, Py, Pz}; T uses kB = kboltz. */
int mdx_kinetics(mdx_system* sys, double kboltz, double* out8);
/* mdx_kinetics over caller-owned device arrays, allocating nothing.  No
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).
 * pos3 and vel3 hold component d of atom i at [d * stride + i] (stride >= n),
 * type[i] is the species (0..3), mass4 (host) the mass of each species; out8
 * (device) receives {EK, T, COM, P} as in mdx_kinetics.  stream is a
 * hipStream_t (NULL: the null stream); the call only enqueues work on it.
 * The partial sums live in a static device array of 8 * 1024 + 8 doubles
 * (66 KB on each GPU the library runs on, held until the process ends);
 * n is limited to 2^28 atoms (MDX_ERR_CAPACITY above).  Every call on a GPU
 * shares that GPU's array, so calls that overlap on one GPU race.  The
 * message of a failed call is in mdx_kinetics_noalloc_last_error(). */
int mdx_kinetics_noalloc(const double* pos3, const double* vel3, const int32_t* type, int32_t n, int64_t stride,
                         const double* mass4, double kboltz, double* out8, void* stream);
const char* mdx_kinetics_noalloc_last_error(void);   /* the calling thread's */

/* Potential energy over the current lists with the force cutoffs (LJ shifted
 * to zero at rc in the periodic box; unshifted in the walled box).
 * out2 = {LJ part, bond-order part}. */
int mdx_potential(mdx_system* sys, double* out2);
/* mdx_potential with the virial: out4 = {LJ energy, bond-order energy, LJ
 * virial, bond-order virial}.  The virial is W = sum over pairs i < j of
 * r_ij . F_ij (r_ij = r_i - r_j, minimum image in the periodic box; F_ij the
 * force of j on i) over the current lists with the force cutoffs, summed in
 * FP64 (per atom, then wave-shuffle block sums); for the many-body forms
 * W = sum_a x_a . F_a: over the per-bond gradients of the atom energies in
 * the periodic box (-sum_i sum_j (x_j - x_i) . de_i/d(x_j - x_i)), over the
 * bond-order atoms' bond forces in the isolated walled box.  The pressure is
 * (2 EK + W) / (3 V).  The reference reports no virial (its records hold EK,
 * T, EU, the centre of mass and the momentum,
 * common/DebugCalculations.cpp:46-124); this extends its diagnostics. */
int mdx_potential_virial(mdx_system* sys, double* out4);
/* common/DebugCalculations.cpp:129-270: LJ over all pairs without cutoff and
 * the reference bond-order energy with xi over all atoms (O(N^2) / O(N^3)). */
int mdx_potential_allpairs(mdx_system* sys, double* out2);

/* Per-kernel timing with HIP events on the system's stream.  While enabled,
 * every launch is bracketed by events; mdx_kernel_times() syncs and returns
 * per kernel name the launch count and total milliseconds. */
int mdx_set_timing(mdx_system* sys, int32_t enabled);
/* writes up to cap entries; names is cap * 48 chars; returns count in *n */
int mdx_kernel_times(mdx_system* sys, char* names, int64_t* counts, double* total_ms,
                     int32_t cap, int32_t* n);
/* blocks until all work on the system's stream is done */
int mdx_synchronize(mdx_system* sys);

/*
 * Learned surrogate of Tersoff silicon (bond_form MDX_BOND_SURROGATE, periodic
 * box; same list settings as MDX_BOND_TERSOFF).  No counterpart in the
 * reference: it takes the place of the bond-order force (manyBodyForces,
 * cpu/common/BondOrderFuncs.h:283-397) as E = sum_i NN(G_i), with the
 * descriptor G_i of 32 radial and 16 bond-resolved angular terms inside 3.0 A
 * and a 48-64-64-1 tanh network run on BF16 matrix cores.  Atoms with more
 * than 6 list entries, or a standardized descriptor outside [lo, hi], take
 * the Tersoff per-atom energy 1/2 sum_j V_ij instead (fallback).
 */
int mdx_set_surrogate(mdx_system* sys, const float* params, int64_t len);   /* len = MDX_SG_NPARAM */
/* {atom evaluations that took the Tersoff fallback, all atom evaluations} since creation */
int mdx_surrogate_counts(mdx_system* sys, int64_t* out2);

/*
 * Training of the surrogate on the GPU: n_frames periodic frames of n_atoms
 * atoms in one cubic box (pos3, force3: 3 doubles per atom, frame after
 * frame; energy: one per frame, eV).  Loss = mean over force components of
 * (F - F_ref)^2 + w_energy * mean over frames of ((E - E_ref) / n_atoms)^2.
 * Every atom may have at most 8 neighbours inside 3.0 A.
 */
typedef struct mdx_sg_trainer mdx_sg_trainer;
int mdx_sg_trainer_create(int32_t n_frames, int32_t n_atoms, double box_length, const double* pos3,
                          const double* force3, const double* energy, mdx_sg_trainer** out);
int mdx_sg_trainer_destroy(mdx_sg_trainer* t);
const char* mdx_sg_trainer_last_error(const mdx_sg_trainer* t);   /* t NULL: this thread's */
/* per descriptor component over all atoms: mean[48], std[48], min[48], max[48] */
int mdx_sg_trainer_feature_stats(mdx_sg_trainer* t, double* out4x48);
/* descriptors of one frame, [atom][48] */
int mdx_sg_trainer_descriptors(mdx_sg_trainer* t, int32_t frame, float* out);
/* the network being trained (MDX_SG_NPARAM floats); set clears the Adam state */
int mdx_sg_trainer_set_params(mdx_sg_trainer* t, const float* params, int64_t len);
int mdx_sg_trainer_get_params(mdx_sg_trainer* t, float* params, int64_t len);
/* loss of the minibatch frames[0..n) into out3 = {loss, force mse, energy mse per atom}
 * (may be NULL) and its gradient in the MDX_SG_NTRAIN trainable weights (may be NULL) */
int mdx_sg_trainer_gradient(mdx_sg_trainer* t, const int32_t* frames, int32_t n, double w_energy, float* grad,
                            double* out3);
/* one Adam step (beta 0.9, 0.999, eps 1e-8) on the minibatch; out3 as above, before the step */
int mdx_sg_trainer_step(mdx_sg_trainer* t, const int32_t* frames, int32_t n, double w_energy, double lr,
                        double* out3);

/*
 * Batched replicas.  The reference runs independent systems one after
 * another (the vxcm sweep restarts the step loop per velocity offset,
 * cpu/cpumd/Host.cpp:75-216); a batch holds n_replicas independent periodic
 * Lennard-Jones boxes of different sizes in one set of device arrays and
 * advances all of them with one graph launch per two steps.  From `common`:
 * precision, dt, lj_rs, lj_rcut, lj_sigma, lj_eps, mass, accel_scale,
 * max_neighbors (0: from the densest replica); the other fields are ignored.
 * Every box must be at least 2 lj_rs wide (minimum image) and hold at most
 * 16384 atoms.  Lists are rebuilt
 * per replica, decided on the device from the predicted positions: exactly
 * when an atom moved (lj_rs - lj_rcut)/2 since its replica's last build.
 * State arrays hold the replicas one after another, 3 doubles per atom.
 */
typedef struct mdx_batch mdx_batch;
int mdx_batch_create(const mdx_config* common, int32_t n_replicas, const int32_t* n_atoms,
                     const double* box_length, mdx_batch** out);
int mdx_batch_destroy(mdx_batch* batch);
const char* mdx_batch_last_error(const mdx_batch* batch);  /* batch NULL: this thread's */
/* (re)starts the batch: positions, velocities; Gear history cleared */
int mdx_batch_set_state(mdx_batch* batch, const double* pos3, const double* vel3);
/* any output may be NULL; force3 is the last computed force */
int mdx_batch_get_state(mdx_batch* batch, double* pos3, double* vel3, double* force3);
/* nsteps steps of every replica; the first call after set_state computes the
 * initial forces and a = F/m (cpu/cpumd/Host.cpp:132-138) */
int mdx_batch_run(mdx_batch* batch, int32_t nsteps);
/* list rebuilds per replica since creation (n_replicas entries) */
int mdx_batch_build_counts(mdx_batch* batch, int64_t* builds);
/* the path of each replica (n_replicas entries), fixed at creation: 1 for a
 * box of at least 3 cells of lj_rs per axis whose lists and steps run through
 * the single box's LDS-staged kernels (its atoms kept in cell order inside the
 * batch; get_state returns them in the caller's order), 2 for a narrower box
 * of at most 2048 atoms, stepped by one workgroup per 64 atoms over all pairs
 * staged in LDS, 0 for the per-atom gather kernels over neighbour lists.  No
 * counterpart in the reference, which runs one system at a time; for callers
 * that size or account work per replica. */
int mdx_batch_staged(mdx_batch* batch, int32_t* path);
/* per replica {kinetic, potential (LJ shifted to zero at rc)}: 2 * n_replicas */
int mdx_batch_energies(mdx_batch* batch, double* out2);
/* per replica the virial W = sum over pairs i < j of r_ij . F_ij (minimum
 * image), n_replicas doubles; per-atom FP64 terms summed by wave shuffles.
 * No reference counterpart (mdx_potential_virial is the single box's). */
int mdx_batch_virials(mdx_batch* batch, double* out);
/* per-kernel timing with events; while enabled, mdx_batch_run launches the
 * kernels one by one instead of through the graph */
int mdx_batch_set_timing(mdx_batch* batch, int32_t enabled);
int mdx_batch_kernel_times(mdx_batch* batch, char* names, int64_t* counts, double* total_ms,
                           int32_t cap, int32_t* n);
int mdx_batch_synchronize(mdx_batch* batch);

/*
 * The vxcm sweep as one batch.  The reference runs its sweep's members one
 * after another (cpu/cpumd/Host.cpp:75-216): the same walled system, re-read
 * for each centre-of-mass velocity offset, through the whole step loop.  A
 * sweep holds `members` walled systems of one configuration (cfg as for
 * mdx_create, box MDX_BOX_WALLED) on one stream; mdx_sweep_run advances all
 * of them with one launch of each step kernel for every member (predict,
 * Lennard-Jones and bond-order forces, corrector with walls, and the forces
 * and a = F/m again for the members whose corrector hit a wall), each member
 * keeping its own list schedule (Host.cpp:143-152, 276-291) and wall flag.
 * The rebuilds of all members due at a step share two read-backs.  Each
 * member's trajectory is the one mdx_run gives the same system, bit for bit.
 */
typedef struct mdx_sweep mdx_sweep;
int mdx_sweep_create(const mdx_config* cfg, int32_t members, mdx_sweep** out);
int mdx_sweep_destroy(mdx_sweep* sweep);
const char* mdx_sweep_last_error(const mdx_sweep* sweep);  /* sweep NULL: this thread's */
/* The system of one member (0 <= member < members), for every per-system
 * call: state, readyList and list builds, diagnostics, mdx_run of that member
 * alone.  It belongs to the sweep: mdx_destroy refuses it. */
int mdx_sweep_member(mdx_sweep* sweep, int32_t member, mdx_system** out);
/* nsteps steps of every member (mdx_run's loop); a member's first step
 * computes its missing lists, its initial forces and a = F/m (Host.cpp:117-138) */
int mdx_sweep_run(mdx_sweep* sweep, int32_t nsteps);
/* per-kernel timing of the members' launches (mdx_set_timing on every member);
 * the sweep's step kernels are counted in member 0's mdx_kernel_times */
int mdx_sweep_set_timing(mdx_sweep* sweep, int32_t enabled);

/*
 * Launch shapes of the LDS-staged kernels.  No counterpart in the reference,
 * whose CUDA path hard-wires its block shapes (cuda/cudamd/Host.cpp:86-88,373);
 * here they come from a tuning table written by the project's tuner
 * (python -m moleculardynamics_amd.tune) and read by mdx_create (the periodic
 * Lennard-Jones box, kind "lj", by its atom count) and mdx_batch_create (kind
 * "replicas", by the batch's total atom count).  The table is the file named
 * by MDX_TUNING_FILE when that is set (empty: none), else data/tuning.json
 * beside libmdx.so; without a table or a matching entry the built-in shape
 * applies.  The shape changes only how the work is cut into workgroups: every
 * atom's candidates, and so its force, come out the same.
 */
typedef struct mdx_shape {
    int32_t block;     /* threads a workgroup of the single box's LDS kernels: 512 or 1024 */
    float fill;        /* home atoms a segment aims at, as a fraction of one pass (block / 4) */
    float stage_aim;   /* replicas: staged atoms a segment's neighbourhood aims at */
    int32_t source;    /* MDX_SHAPE_DEFAULT or MDX_SHAPE_TUNED */
} mdx_shape;
#define MDX_SHAPE_DEFAULT 0   /* built in: block 512, fill 0.85, stage_aim 1250 */
#define MDX_SHAPE_TUNED 1     /* from the tuning table */
/* the shape a table at path gives a system of kind on a device of
 * architecture arch ("gfx950"); host code only, no device needed */
int mdx_tuning_lookup(const char* path, const char* arch, const char* kind, int64_t atoms, mdx_shape* out);
const char* mdx_tuning_last_error(void);   /* the calling thread's */
/* the shape a system or batch runs with */
int mdx_launch_shape(mdx_system* sys, mdx_shape* out);
int mdx_batch_launch_shape(mdx_batch* batch, mdx_shape* out);

#ifdef __cplusplus
}
#endif
#endif
