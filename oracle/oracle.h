/*
 * oracle.h -- FP64 CPU restatement of the vakuras/moleculardynamics cpumd hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine in
 * moleculardynamics_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  The product path never links or calls it.
 *
 * Two families of functions:
 *
 *  orc_ref_*   the reference's own semantics (walled box, N/O/Ne species, the
 *              Morse-like bond-order potential, hashed 27-cell candidate lists,
 *              Gear 5th-order predictor-corrector), AoS arrays exactly as the
 *              reference lays them out (real4 position with the species code
 *              in w, real3 for v/a/b/c/f).  Each function cites the reference
 *              file:line it restates.
 *
 *  orc_pbc_*   the periodic, reduced-unit Lennard-Jones and Tersoff-Si systems
 *              that the benchmark configurations use.  The reference has no
 *              periodic boundaries and no Tersoff-Si; these are restated from
 *              the textbook forms (12-6 LJ; Tersoff, PRB 38, 9902 (1988)) and
 *              pinned by closed forms, brute force and finite differences in
 *              tests/.
 *
 * Parity status per function (what pins it) is listed in tests/ and in the
 * design notes; every function here has at least one independent check.
 */
#ifndef MDX_ORACLE_H
#define MDX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference-mode types (common/Global.h:58-72, common/NeighborList.h:19-28) ---- */
typedef struct { double x, y, z, w; } orc_real4;
typedef struct { double x, y, z; } orc_real3;

typedef struct {
    float maxnlmove;
    float maxnlmovesq;
    float rcut;
    float rs;
    float rcutsq;
    int32_t nlistsize;
    int32_t nlargestsize;
} orc_list_settings;

/* species codes stored in pos.w (common/Constants.h:13-15) */
enum { ORC_N = 0, ORC_O = 1, ORC_NG = 2 };

/* common/NeighborList.cpp:14-156.  Returns a malloc'ed padded list of
 * n * ls->nlistsize ints (free with orc_free); updates nlistsize/nlargestsize. */
int32_t* orc_ref_build_neighbor_list(const orc_real4* pos, orc_list_settings* ls,
                                     double box_size, int32_t buckets, int32_t n);
void orc_free(void* p);

/* cpu/cpumd/Host.cpp:245-271 */
void orc_ref_ready_list(const orc_real4* pos, const orc_real3* vel, int32_t n,
                        double* vmax, double* box_size_list);
/* cpu/cpumd/Host.cpp:288-290: next build step (INT32_MAX == never) */
int32_t orc_ref_next_build(float maxnlmove, double vmax, double dt, int32_t t);

/* cpu/common/Device.cpp:17-39 */
void orc_ref_predict(orc_real4* r, orc_real3* v, orc_real3* a, orc_real3* b, orc_real3* c,
                     double dt, int32_t n);
/* cpu/common/Device.cpp:44-120 */
void orc_ref_correct(orc_real4* r, orc_real3* v, const orc_real3* f, orc_real3* a,
                     orc_real3* b, orc_real3* c, double dt, int32_t n, double npos,
                     int32_t* flag, float energy_loss);
/* cpu/common/Device.cpp:125-134 */
void orc_ref_accelerations(const orc_real4* r, orc_real3* acc, const orc_real3* f, int32_t n);
/* cpu/common/LennardJones.h:35-93 (assigns f) */
void orc_ref_lj_forces(const orc_real4* r, orc_real3* f, int32_t n, int32_t nlargestsize,
                       int32_t nlistsize, const int32_t* list, double rcutsq);
/* cpu/common/BondOrderFuncs.h:283-397 (adds to f when ljexist, else assigns) */
void orc_ref_mb_forces(const orc_real4* r, orc_real3* f, int32_t n, int32_t nlargestsize,
                       int32_t nlistsize, const int32_t* list, double rcutsq, int32_t ljexist);
/* common/DebugCalculations.cpp:46-124; out[0..6] = ek, T, comx, comy, comz, px, py, pz */
void orc_ref_kinetics(const orc_real4* r, const orc_real3* v, int32_t n, double* out8);
/* common/DebugCalculations.cpp:129-270; out[0]=LJ part, out[1]=bond-order part */
void orc_ref_potential(const orc_real4* r, int32_t n, double* out2);

/* arvdwfc scalar helpers (common/arvdwfc.cpp:16-59), exported for derivative tests */
void orc_ref_attractive(double r, double de, double beta, double re, double* va, double* dva);
void orc_ref_repulsive(double r, double de, double beta, double re, double* vr, double* dvr);
void orc_ref_vdw(double r, double r0v, double* vdw, double* dvdw);
void orc_ref_cutoff(double r, double* fc, double* dfc);
/* parameter getters (common/Getters.cpp:13-78); which: 0 mass,1 re,2 de,3 beta,4 r0v */
double orc_ref_param(int32_t which, int32_t ti, int32_t tj);

/* Reference step loop for one vxcm value (cpu/cpumd/Host.cpp:78-198), run
 * on caller arrays (clamped into +-box/2 and given +vxcm first, as at
 * Host.cpp:94-106); a/b/c start at zero as at Host.cpp:86-88.  Writes
 * the energy record (t, T, E, EK, EU, COM, P) at each output step into
 * records (11 doubles each, capacity max_records); returns the number of
 * records written, or -1 on allocation failure.
 * wall: 0 puts the walls at +-box (box = N^(1/3) * 52.8, Host.cpp:31,166);
 * a positive value puts them there instead (tests that drive atoms into the
 * walls).  stats (may be NULL) receives ORC_RUN_STATS counts: [0..6] atoms
 * turned back per wall branch (x+ reflect, x+ teleport, x-, y+, y-, z+, z-),
 * [7] steps with a wall hit (forces recomputed), [8] Lennard-Jones list
 * builds in the loop, [9] of those the builds that came after a teleport. */
#define ORC_RUN_STATS 10
typedef struct {
    int32_t n_lj, n_mb, timesteps, output_timesteps;
    float dt, lj_rs, lj_rcut, mb_rs, mb_rcut, energy_loss /* raw, sqrt applied inside */;
    float vxcm;
    int32_t use_lj, use_mb;
    double wall;
    int64_t* stats;
} orc_ref_run_cfg;
int32_t orc_ref_run(const orc_ref_run_cfg* cfg, orc_real4* r, orc_real3* v,
                    double* records, int32_t max_records, double* elapsed_s);

/* the compiler and flags the oracle was built with */
const char* orc_build_info(void);

/* ---- periodic reduced-unit Lennard-Jones (benchmark configurations) ---- */

/* Candidate list in the reference's padded format built on a periodic cell
 * grid: ncell = floor(L/rs) per axis (>= 3), cell size L/ncell, 27-cell
 * stencil in (dx,dy,dz) lexicographic order from -1, ascending index within a
 * cell, self excluded.  Returns malloc'ed n*nlistsize ints (pad -1). */
int32_t* orc_pbc_build_list(const double* pos3, int32_t n, double L, double rs,
                            int32_t* nlistsize, int32_t* nlargestsize);

/* Check of an exported pair list against the exact FP64 sets within rs and
 * rc at the build positions; see oracle_pbc.c for the 8 counts in stats. */
int32_t orc_pbc_pair_list_check(const double* pos3, int32_t n, double L, double rs, double rc, double slack2,
                                const int32_t* rows, const int32_t* counts, int32_t width, int64_t* stats);

/* LJ forces over a candidate list with the minimum image; single species
 * (sigma, eps).  f is assigned.  Returns the truncated, shifted potential
 * energy (sum over pairs with r < rc of 4eps[(s/r)^12-(s/r)^6] - shift);
 * *virial receives sum over pairs of r.F. */
double orc_pbc_lj_forces(const double* pos3, double* f3, int32_t n, double L,
                         double sigma, double eps, double rc, int32_t nlistsize,
                         const int32_t* list, double* virial);

/* Gear 5th-order predictor / corrector on AoS double3 arrays with
 * accel = F / mass * scale; same coefficients and operation order as the
 * reference (cpu/common/Device.cpp:17-120, common/Constants.h:64-67), no walls. */
void orc_gear_predict(double* r3, double* v3, double* a3, double* b3, const double* c3,
                      double dt, int32_t n);
void orc_gear_correct(double* r3, double* v3, const double* f3, double* a3, double* b3,
                      double* c3, double dt, int32_t n, double mass, double scale);

/* Full periodic LJ run: Gear PC, list rebuilt when the max displacement since
 * the last build exceeds skin/2 (skin = rs - rc) or every step if
 * rebuild_every_step.  energies (if non-null) receives 3 doubles per step
 * (potential, kinetic, total) for steps 0..nsteps.  Returns the number of
 * list builds, or -1 on allocation failure.  *elapsed_s covers the step loop. */
int32_t orc_pbc_lj_run(double* r3, double* v3, int32_t n, double L, double rc, double rs,
                       double dt, int32_t nsteps, int32_t rebuild_every_step,
                       double* energies, double* elapsed_s);
/* the same loop with its force, Gear and displacement loops split over
 * `threads` OpenMP threads (1: the serial loop above, bit for bit) */
int32_t orc_pbc_lj_run_threads(double* r, double* v, int32_t n, double L, double rc, double rs,
                               double dt, int32_t nsteps, int32_t rebuild_every_step,
                               double* energies, double* elapsed_s, int32_t threads);

/* ---- Tersoff silicon (Tersoff 1988) ---- */
typedef struct {
    double A, B, lambda1, lambda2, lambda3, beta, n, c, d, h, R, D;
    int32_t m;
} orc_tersoff_params;
void orc_tersoff_si_params(orc_tersoff_params* p);  /* Si(C) set with R=2.7, D=0.3 */

/* Energy (eV) and forces (eV/A) of a periodic (L > 0) or free (L <= 0) system,
 * brute force over all pairs with the minimum image.  f may be NULL. */
double orc_tersoff_energy_forces(const orc_tersoff_params* p, const double* pos3, int32_t n,
                                 double L, double* f3);
/* Per-centre form: only centres with centre[i] != 0 (all if NULL) contribute
 * their e_i = 1/2 sum_j V_ij; f3 = minus the gradient of that partial sum,
 * e_atom (if non-null) = e_i.  The surrogate's plain-Tersoff fallback. */
double orc_tersoff_centres(const orc_tersoff_params* p, const double* pos3, int32_t n, double L,
                           const uint8_t* centre, double* f3, double* e_atom);

#ifdef __cplusplus
}
#endif
#endif
