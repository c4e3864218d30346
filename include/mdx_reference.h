/*
 * mdx_reference.h -- the reference's own free-function interface, served by
 * the MI355X engine (libmdx_ref.so).
 *
 * vakuras/moleculardynamics has no FFI: the cpumd (and equi) hostMain calls
 * C++ functions compiled into the executable from cpu/common/Device.cpp,
 * cpu/common/LennardJones.h, cpu/common/BondOrderFuncs.h and
 * common/NeighborList.cpp.  libmdx_ref.so defines functions with the same
 * names, parameter types and C++ linkage, so an executable linked against it
 * in place of those four sources resolves the same symbols:
 *
 *   declared here                  replaces (file:line)
 *   ------------------------------ -------------------------------------------
 *   predict                        cpu/common/Device.h:19, Device.cpp:17-39
 *   correct                        cpu/common/Device.h:20, Device.cpp:44-120
 *   lennardJonesForces             cpu/common/Device.h:21, LennardJones.h:35-93
 *   manyBodyForces                 cpu/common/Device.h:22, BondOrderFuncs.h:283-397
 *   calculateAccelerations         cpu/common/Device.h:23, Device.cpp:125-134
 *   buildNeighborList              common/NeighborList.h:30, NeighborList.cpp:14-156
 *
 * The types have the reference's layout (common/Global.h:25-72,
 * common/NeighborList.h:19-28) with real = double, the type the shipped cpumd
 * is built with (its project defines EQUI).  computeBucket
 * (common/NeighborList.h:31) is the reference list builder's private hash
 * helper and is not provided.
 *
 * Each call copies the caller's host arrays to the device, runs the
 * corresponding C-ABI entry point of include/mdx.h in FP64 (bit-identical to
 * the reference's arithmetic; MDX_REF_PRECISION=mixed selects the FP32 pair
 * path) and copies the results back.  This is the parity path: the fast path
 * keeps the state on the device and calls mdx_run.  Errors print a message and
 * exit(EXIT_FAILURE), as the reference does (common/Global.h:116-117).
 */
#ifndef MDX_REFERENCE_H
#define MDX_REFERENCE_H

typedef double real;

struct real3 {
    real x, y, z;
};
struct real4 {
    real x, y, z, w;  /* w: species code */
};
struct int3 {
    int x, y, z;
};
struct listSettings {
    float maxnlmove, maxnlmovesq, rcut, rs, rcutsq;
    int nlistsize, nlargestsize;
};

void predict(real4* pos, real3* vel, real3* acc, real3* acc1, real3* acc2, real dt, int n);
void correct(real4* pos, real3* vel, real3* force, real3* acc, real3* acc1, real3* acc2, real dt, int n,
             real wall, bool& hit, float energyLoss);
void lennardJonesForces(real4* pos, real3* force, int n, int nlargestsize, int nlistsize, int* list,
                        real rcutsq);
void manyBodyForces(real4* pos, real3* force, int n, int nlargestsize, int nlistsize, int* list,
                    real rcutsq, bool addToLennardJones);
void calculateAccelerations(real4* pos, real3* acc, real3* force, int n);
/* returns a new[] array of n * nlistsize ints (the caller delete[]s it) */
int* buildNeighborList(real4* pos, listSettings* settings, real boxSize, int buckets, int n);

/*
 * Nothing besides these calls is needed: which listSettings object is the
 * Lennard-Jones one and which the bond-order one, the atom counts, and whether
 * Lennard-Jones is on are read from the calls themselves (a list handed to
 * lennardJonesForces or manyBodyForces names its settings' role, manyBodyForces
 * passes ManyBodyParticles and useLennardJones, predict and correct the atom
 * count, dt, the wall and the restitution).  A list built before its role is
 * known is built by a list-only device system of the same atoms and radii,
 * which gives the same rows.
 *
 * No counterpart in the reference: releases the device state (optional; also
 * done at exit).
 */
void mdxRefRelease();

#endif
