"""Headline benchmark: atom-steps/s of the device-resident MD step loop.

Workload (one MI355X): periodic reduced-unit Lennard-Jones, 1,048,576 atoms
(64^3 FCC cells x 4) at density 0.8442, T* = 0.728 seeded Maxwell-Boltzmann
velocities, cutoff 2.5 sigma with 0.3 sigma skin (cells of edge >= 2.8),
FP32 pair arithmetic with FP64 accumulation and FP64 Gear 5th-order
predictor-corrector, dt = 0.005.  One step = predictor, [cell-list rebuild
when due], LJ forces, corrector (fused with the next predictor).

Prints one JSON line (see the fields below).  --gpus > 1 is not supported:
the engine runs on one GPU.

--workload tersoff measures the bond-order configuration instead: Tersoff
silicon (R = 2.7 A, D = 0.3 A), 262,144 atoms of diamond (32^3 cells of
5.431 A) with 0.05 A Gaussian displacements and 1000 K velocities, list
radius 3.3 A, dt = 1 fs, mixed precision.

--workload surrogate measures the learned surrogate of the Tersoff
configuration (same system and initial state): descriptors, the 48-64-64-1
network on BF16 MFMA and its analytic forces, with plain-Tersoff fallback for
out-of-range atoms; plain Tersoff is timed beside it on the same state, and
the surrogate's force error is taken against plain Tersoff on the same
configuration.

--workload replicas measures the batched configuration: 512 independent
periodic LJ replicas of 256-8788 atoms (log-uniform sizes, density
U[0.70, 0.95], T* U[0.5, 2.0]; 1,234,468 atoms together), rc 2.5, rs 2.8,
dt 0.005, mixed precision, advanced by graph replays.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
DT = float(np.float32(0.005))  # DT is a float in the configuration (common/Global.h:79)
# bytes per atom per step of the whole step, FP64 state: Gear state r,v,a,b,c
# read+written by the integrator (240), force written by the force kernel and
# read by the integrator (48), positions read by the force kernel (24)
STEP_BYTES_PER_ATOM = 312
# the dominant kernel (md_step) does the whole step: LJ force, corrector and the
# next predictor; its algorithmic bytes per atom are the step's 312 B
# measured HBM bytes per md_step launch (FETCH_SIZE + WRITE_SIZE, copy-calibrated)
# of this workload, as recorded by the committed rocprofv3 summary
TRAFFIC_PROFILE = ROOT / "profiles" / "md_step_lj1m.md"


def measured_traffic(profile: Path = TRAFFIC_PROFILE) -> int | None:
    try:
        for line in profile.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--ncell", type=int, default=64, help="FCC cells per edge (4 ncell^3 atoms)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-ncell", type=int, default=32)
    p.add_argument("--cpu-steps", type=int, default=20)
    p.add_argument("--workload", choices=["lj", "tersoff", "replicas", "surrogate"], default="lj")
    p.add_argument("--weights", default=None, help="surrogate network (default: the shipped one)")
    p.add_argument("--mb-rs", type=float, default=3.3, help="Tersoff / surrogate list radius (A)")
    p.add_argument("--copy-bytes", type=int, default=0,
                   help="only run device-to-device copies of this many bytes (counter calibration)")
    return p.parse_args()


def copy_calibration(nbytes: int, reps: int = 3) -> None:
    """Device-to-device copies of a known size: what FETCH_SIZE / WRITE_SIZE
    read for them scales the counters of the MD kernels into bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_bytes": nbytes, "reps": reps}))


def cpu_baseline(ncell: int, steps: int) -> dict:
    """The FP64 oracle's single-threaded periodic LJ step loop (the reference
    algorithm: 27-cell candidate lists, Gear PC) on a sub-box sample."""
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle as O
    from moleculardynamics_amd import initial

    pos, vel, L = initial.lj_fluid(ncell, 0.8442, 0.728, seed=12345)
    _, _, _, builds, elapsed = O.pbc_lj_run(pos, vel, L, 2.5, 2.8, DT, steps)
    return {
        "value": len(pos) * steps / elapsed,
        "unit": "atom-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{len(pos)} atoms ({ncell}^3 FCC x4, same density/T/cutoffs), {steps} steps, "
                  f"{builds} list builds, {elapsed:.1f} s on {platform.processor() or platform.machine()}",
    }


# Tersoff's per-atom kernel (k_ts_atoms) per atom: positions read (24 B;
# neighbour positions are cache hits), the list (4 x 4 B) and its count (4 B),
# the per-slot vectors written (4 x 16 B)
TS_ATOMS_BYTES_PER_ATOM = 24 + 16 + 4 + 64


def cpu_baseline_tersoff(ncell: int, steps: int) -> dict:
    """The FP64 oracle's Tersoff forces (cell search, O(N L^2)) with its Gear
    predictor-corrector, single-threaded, on a sub-box sample."""
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle as O
    from moleculardynamics_amd import initial

    pos, vel, L = initial.si_diamond(ncell, 1000.0, seed=12345, disp=0.05)
    t0 = time.perf_counter()
    O.tersoff_run(pos, vel, L, 1.0, steps, initial.SI_MASS, initial.ACCEL_EV_A_AMU)
    elapsed = time.perf_counter() - t0
    return {
        "value": len(pos) * steps / elapsed,
        "unit": "atom-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{len(pos)} atoms ({ncell}^3 diamond x8, same T/displacements), {steps} steps, "
                  f"forces every step, {elapsed:.1f} s on {platform.processor() or platform.machine()}",
    }


def bench_tersoff(a) -> None:
    from moleculardynamics_amd import BOND_TERSOFF, BOX_PERIODIC, LJ_UNIFORM, PREC_MIXED, System, initial

    ncell = 32 if a.ncell == 64 else a.ncell
    pos, vel, L = initial.si_diamond(ncell, 1000.0, seed=2024, disp=0.05)
    n = len(pos)
    print(f"bench: Tersoff Si {n} atoms, L={L:.4f}; warmup {a.warmup}, steps {a.steps}", file=sys.stderr, flush=True)
    s = System(n_atoms=n, n_bond=n, box=BOX_PERIODIC, precision=PREC_MIXED, use_lj=0, use_bond=1,
               lj_table=LJ_UNIFORM, bond_form=BOND_TERSOFF, dt=1.0, mb_rs=a.mb_rs, mb_rcut=3.0, box_length=L,
               mass=initial.SI_MASS, accel_scale=initial.ACCEL_EV_A_AMU)
    s.set_state(np.c_[pos, np.zeros(n)], vel)
    s.run(a.warmup)
    s.synchronize()
    b0 = s.build_count()
    t0 = time.perf_counter()
    s.run(a.steps)
    s.synchronize()
    t1 = time.perf_counter()
    builds = s.build_count() - b0
    # per-kernel times from a separate pass with events around each launch
    s.set_timing(True)
    s.run(min(a.steps, 50))
    s.synchronize()
    times = s.kernel_times()
    s.set_timing(False)
    s.close()
    elapsed = t1 - t0
    value = n * a.steps / elapsed
    print(f"bench: {elapsed:.3f} s, {builds} list builds; kernels {times}", file=sys.stderr, flush=True)
    cnt, ms = times.get("ts_atoms", (0, 0.0))
    k_ms = ms / max(cnt, 1)
    achieved = n * TS_ATOMS_BYTES_PER_ATOM / (k_ms * 1e-3) / 1e9 if cnt else None
    out = {
        "metric": "atom-steps/s per GPU (Tersoff Si 262k, mixed precision)",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-terms/fp64-accum",
        "data": "synthetic (seeded diamond lattice + Gaussian displacements + Maxwell-Boltzmann velocities)",
        "config": {"workload": f"periodic Tersoff Si {n} atoms (diamond {ncell}^3 x8, a 5.431 A), 1000 K, "
                               f"R 2.7 D 0.3, list radius {a.mb_rs} A, dt 1 fs", "atoms": n, "list_builds": builds},
        "roofline": {
            "kernel": "k_ts_atoms (per-atom Tersoff terms and per-slot gradients)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            # copy-calibrated FETCH_SIZE + WRITE_SIZE per launch (profiles/tersoff_si262k.md)
            "traffic": measured_traffic(ROOT / "profiles" / "tersoff_si262k.md"),
            "avg_launch_ms": k_ms,
            "bytes_per_atom": TS_ATOMS_BYTES_PER_ATOM,
        },
        "kernel_ms": {k: v[1] / max(v[0], 1) for k, v in times.items()},
    }
    if not a.no_cpu:
        print("bench: cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline_tersoff(16, 200)
    print(json.dumps(out))


# the surrogate's force kernel (k_sg_eval) per atom: positions read (24 B;
# neighbour positions are cache hits), the list (4 x 4 B) and its count (4 B)
# read, the per-slot gradients written (4 slots x 12 B)
SG_EVAL_BYTES_PER_ATOM = 24 + 16 + 4 + 48
# network work per atom as the kernel issues it (16x16x32 BF16 MFMA tiles,
# K of the first layer padded 48 -> 64): forward 2 layers, backward 2 layers
SG_MFMA_FLOP_PER_ATOM = 2 * (64 * 64 + 64 * 64 + 64 * 64 + 48 * 64)
MFMA_BF16_PEAK_TFLOPS = 2500.0


def cpu_baseline_surrogate(ncell: int, steps: int, blob) -> dict:
    """The FP64 oracle's surrogate (numpy: descriptors, network, analytic
    forces) with the oracle's Gear predictor-corrector, on a sub-box sample."""
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle as O
    from moleculardynamics_amd import initial
    from oracle import surrogate as S

    p = S.unpack(np.asarray(blob, dtype=np.float64))
    pos, vel, L = initial.si_diamond(ncell, 1000.0, seed=12345, disp=0.05)
    m, sc = initial.SI_MASS, initial.ACCEL_EV_A_AMU
    r, v = pos.copy(), vel.copy()

    def fb(mask):   # plain-Tersoff per-atom terms of out-of-range atoms
        _, f_t, e_t = O.tersoff_centres(r, L, mask.astype(np.uint8))
        return e_t, f_t

    a_, b_, c_ = np.zeros_like(r), np.zeros_like(r), np.zeros_like(r)
    from threadpoolctl import threadpool_limits

    with threadpool_limits(1):   # one core, as stated below
        t0 = time.perf_counter()
        _, f, _, _ = S.energy_forces(p, r, L, tersoff_fallback=fb)
        a_[:] = f / m * sc
        for _ in range(steps):
            r, v, a_, b_, c_ = O.gear_predict(r, v, a_, b_, c_, 1.0)
            _, f, _, _ = S.energy_forces(p, np.mod(r, L), L, tersoff_fallback=fb)
            r, v, a_, b_, c_ = O.gear_correct(r, v, f, a_, b_, c_, 1.0, m, sc)
        elapsed = time.perf_counter() - t0
    return {
        "value": len(pos) * steps / elapsed,
        "unit": "atom-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{len(pos)} atoms ({ncell}^3 diamond x8, 1000 K), {steps} steps of the numpy FP64 surrogate "
                  f"(brute-force neighbours), {elapsed:.1f} s on {platform.processor() or platform.machine()}",
    }


def bench_surrogate(a) -> None:
    from moleculardynamics_amd import BOND_SURROGATE, BOND_TERSOFF, LIST_BOND, initial
    from moleculardynamics_amd import surrogate as SG

    blob = SG.load_weights(a.weights) if a.weights else SG.load_weights()
    ncell = 32 if a.ncell == 64 else a.ncell
    pos, vel, L = initial.si_diamond(ncell, 1000.0, seed=2024, disp=0.05)
    n = len(pos)
    print(f"bench: surrogate Si {n} atoms, L={L:.4f}; warmup {a.warmup}, steps {a.steps}", file=sys.stderr,
          flush=True)

    def timed(form):
        s = SG.silicon_system(n, L, bond_form=form, mb_rs=a.mb_rs)
        if form == BOND_SURROGATE:
            s.set_surrogate(blob)
        s.set_state(np.c_[pos, np.zeros(n)], vel)
        s.run(a.warmup)
        s.synchronize()
        t0 = time.perf_counter()
        s.run(a.steps)
        s.synchronize()
        elapsed = time.perf_counter() - t0
        s.set_timing(True)
        s.run(min(a.steps, 50))
        s.synchronize()
        times = s.kernel_times()
        s.set_timing(False)
        return s, elapsed, times

    sp, el_plain, t_plain = timed(BOND_TERSOFF)
    final = np.mod(sp.get_state()["pos"][:, :3], L)
    sp.close()
    print(f"bench: plain Tersoff {el_plain:.3f} s; kernels {t_plain}", file=sys.stderr, flush=True)
    ss, elapsed, times = timed(BOND_SURROGATE)
    fallback, evals = ss.surrogate_counts()
    ss.close()
    print(f"bench: surrogate {elapsed:.3f} s; kernels {times}", file=sys.stderr, flush=True)
    # force error on the configuration plain Tersoff reached
    def forces(form):
        with SG.silicon_system(n, L, bond_form=form) as s:
            if form == BOND_SURROGATE:
                s.set_surrogate(blob)
            s.set_state(np.c_[final, np.zeros(n)], np.zeros((n, 3)))
            s.build_list(LIST_BOND)
            s.bond_forces()
            return s.get_state()["force"]
    f_plain, f_sg = forces(BOND_TERSOFF), forces(BOND_SURROGATE)
    rmse = float(np.sqrt(np.mean((f_sg - f_plain) ** 2)))
    value = n * a.steps / elapsed
    value_plain = n * a.steps / el_plain
    cnt, ms = times.get("sg_eval", (0, 0.0))
    k_ms = ms / max(cnt, 1)
    achieved = n * SG_EVAL_BYTES_PER_ATOM / (k_ms * 1e-3) / 1e9 if cnt else None
    out = {
        "metric": "atom-steps/s per GPU (Tersoff-Si surrogate 262k, BF16 MFMA network) vs plain Tersoff",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16-mfma-network/fp32-descriptors/fp64-accum",
        "data": "synthetic (seeded diamond lattice + Gaussian displacements + Maxwell-Boltzmann velocities); "
                "network trained on GPU plain-Tersoff frames",
        "config": {"workload": f"periodic Tersoff-Si surrogate {n} atoms (diamond {ncell}^3 x8, a 5.431 A), "
                               f"1000 K, list radius {a.mb_rs} A, dt 1 fs", "atoms": n},
        "plain_tersoff": {"value": value_plain, "ms_per_step": 1e3 * el_plain / a.steps,
                          "kernel_ms": {k: v[1] / max(v[0], 1) for k, v in t_plain.items()}},
        "speedup_vs_plain": value / value_plain,
        "force_rmse_vs_plain_eV_per_A": rmse,
        "force_rms_plain_eV_per_A": float(np.sqrt(np.mean(f_plain ** 2))),
        "fallback_fraction": fallback / max(evals, 1),
        "roofline": {
            "kernel": "k_sg_eval (descriptors, BF16 MFMA network, per-slot gradients)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            # copy-calibrated FETCH_SIZE + WRITE_SIZE per launch (profiles/sg_eval_si262k.md)
            "traffic": measured_traffic(ROOT / "profiles" / "sg_eval_si262k.md"),
            "avg_launch_ms": k_ms,
            "bytes_per_atom": SG_EVAL_BYTES_PER_ATOM,
        },
        "mfma": {"flop_per_atom": SG_MFMA_FLOP_PER_ATOM,
                 "achieved_tflops": n * SG_MFMA_FLOP_PER_ATOM / (k_ms * 1e-3) / 1e12 if cnt else None,
                 "peak_tflops": MFMA_BF16_PEAK_TFLOPS},
        "kernel_ms": {k: v[1] / max(v[0], 1) for k, v in times.items()},
    }
    if not a.no_cpu:
        print("bench: cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline_surrogate(4, 240, blob)
    print(json.dumps(out))


# The batch's step kernel of the staged replicas (rep_step_lds, k_step_lds
# over their z-segments) does their whole step like md_step: its algorithmic
# bytes per atom are the step's 312 B.  The other replicas' step (rep_step,
# k_rep_step) is priced apart.


def cpu_baseline_replicas(ns, Ls, pos, vel, max_atoms: int, steps: int) -> dict:
    """The FP64 oracle's periodic LJ loop (27-cell lists, Gear PC) on the
    first replicas wide enough for its 3-cell grid, about max_atoms atoms."""
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle as O

    off, done, atoms, elapsed = 0, 0, 0, 0.0
    for n, L in zip(ns, Ls):
        if atoms >= max_atoms:
            break
        if int(np.floor(L / 2.8)) >= 3:
            _, _, _, _, el = O.pbc_lj_run(pos[off: off + n], vel[off: off + n], L, 2.5, 2.8, DT, steps)
            elapsed += el
            atoms += n
            done += 1
        off += n
    return {
        "value": atoms * steps / elapsed,
        "unit": "atom-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{done} replicas of the batch ({atoms} atoms), {steps} steps each, {elapsed:.1f} s on "
                  f"{platform.processor() or platform.machine()}",
    }


def bench_replicas(a) -> None:
    from moleculardynamics_amd import PREC_MIXED, Batch, initial

    ns, Ls, pos, vel = initial.replica_batch(512, seed=2024)
    n = int(ns.sum())
    print(f"bench: {len(ns)} LJ replicas, {n} atoms; warmup {a.warmup}, steps {a.steps}", file=sys.stderr, flush=True)
    b = Batch(ns, Ls, precision=PREC_MIXED, dt=0.005, lj_rs=2.8, lj_rcut=2.5)
    staged = b.staged()
    n_staged = int(ns[staged].sum())
    n_small = int(ns[b.paths() == 2].sum())
    b.set_state(pos, vel)
    b.run(a.warmup)
    b.synchronize()
    builds0 = b.build_counts().sum()
    t0 = time.perf_counter()
    b.run(a.steps)                      # graph replays
    b.synchronize()
    t1 = time.perf_counter()
    builds = int(b.build_counts().sum() - builds0)
    elapsed = t1 - t0
    value = n * a.steps / elapsed
    # the same steps launched one by one with events around each kernel, for the per-kernel record
    b.set_timing(True)
    b.run(a.steps)
    b.synchronize()
    times = b.kernel_times()
    b.set_timing(False)
    b.close()
    print(f"bench: {elapsed:.3f} s (graph), {builds} replica list builds; kernels {times}", file=sys.stderr, flush=True)
    cnt, ms = times.get("rep_step_lds", (0, 0.0))
    k_ms = ms / max(cnt, 1)
    achieved = n_staged * STEP_BYTES_PER_ATOM / (k_ms * 1e-3) / 1e9 if cnt else None
    out = {
        "metric": "atom-steps/s per GPU (512 LJ replicas, mixed precision)",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-pair/fp64-accum",
        "data": "synthetic (seeded FCC blocks + Maxwell-Boltzmann velocities per replica)",
        "config": {"workload": f"{len(ns)} periodic LJ replicas, {n} atoms ({int(ns.min())}-{int(ns.max())} per replica), "
                               f"rho U[0.70,0.95], T* U[0.5,2.0], rc 2.5 skin 0.3, dt 0.005", "atoms": n,
                   "replica_list_builds": builds, "staged_replicas": int(staged.sum()), "staged_atoms": n_staged,
                   "all_pairs_atoms": n_small},
        "roofline": {
            "kernel": "rep_step_lds (k_step_lds over the staged replicas' segments)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": None,
            "avg_launch_ms": k_ms,
            "atoms": n_staged,
            "bytes_per_atom": STEP_BYTES_PER_ATOM,
        },
        "kernel_ms": {k: v[1] / max(v[0], 1) for k, v in times.items()},
    }
    if not a.no_cpu:
        print("bench: cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline_replicas(ns, Ls, pos, vel, 60000, 20)
    print(json.dumps(out))


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "the engine runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if a.copy_bytes:
        copy_calibration(a.copy_bytes)
        return
    if a.workload == "tersoff":
        bench_tersoff(a)
        return
    if a.workload == "replicas":
        bench_replicas(a)
        return
    if a.workload == "surrogate":
        bench_surrogate(a)
        return
    from moleculardynamics_amd import BOX_PERIODIC, LJ_UNIFORM, PREC_MIXED, System, initial

    pos, vel, L = initial.lj_fluid(a.ncell, 0.8442, 0.728, seed=2024)
    n = len(pos)
    print(f"bench: LJ {n} atoms, L={L:.4f}; warmup {a.warmup}, steps {a.steps}", file=sys.stderr, flush=True)
    s = System(n_atoms=n, box=BOX_PERIODIC, precision=PREC_MIXED, use_lj=1, lj_table=LJ_UNIFORM,
               dt=0.005, lj_rs=2.8, lj_rcut=2.5, box_length=L, lj_sigma=1.0, lj_eps=1.0, mass=1.0)
    s.set_state(np.c_[pos, np.zeros(n)], vel)
    s.run(a.warmup)
    s.synchronize()
    b0 = s.build_count()
    s.set_timing(True)
    t0 = time.perf_counter()
    s.run(a.steps)
    s.synchronize()
    t1 = time.perf_counter()
    times = s.kernel_times()
    s.set_timing(False)
    builds = s.build_count() - b0
    elapsed = t1 - t0
    value = n * a.steps / elapsed
    print(f"bench: {elapsed:.3f} s, {builds} list builds; kernels {times}", file=sys.stderr, flush=True)

    cnt, ms = times.get("md_step", (0, 0.0))
    k_ms = ms / max(cnt, 1)
    achieved = n * STEP_BYTES_PER_ATOM / (k_ms * 1e-3) / 1e9 if cnt else None
    step_gbs = value * STEP_BYTES_PER_ATOM / 1e9
    out = {
        "metric": "atom-steps/s per GPU (LJ 1M, mixed precision)" if n == 1 << 20 else
                  f"atom-steps/s per GPU (LJ {n} atoms, mixed precision)",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-pair/fp64-accum",
        "data": "synthetic (seeded FCC lattice + Maxwell-Boltzmann velocities)",
        "config": {"workload": f"periodic LJ {n} atoms (FCC {a.ncell}^3 x4), rho 0.8442, T* 0.728, "
                               f"rc 2.5 skin 0.3, dt 0.005", "atoms": n, "list_builds": builds},
        "roofline": {
            "kernel": "md_step",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": measured_traffic(),
            "traffic_source": str(TRAFFIC_PROFILE.relative_to(ROOT)),
            "avg_launch_ms": k_ms,
            "bytes_per_atom": STEP_BYTES_PER_ATOM,
            "step_gbs_at_312B_per_atom": step_gbs,
            "step_frac": step_gbs / HBM_PEAK_GBS,
        },
        "kernel_ms": {k: v[1] / max(v[0], 1) for k, v in times.items()},
    }
    s.close()
    if not a.no_cpu:
        print("bench: cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(a.cpu_ncell, a.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
