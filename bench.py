"""Headline benchmark: atom-steps/s of the device-resident MD step loop.

Default run (what `python bench.py` measures, one JSON line):

  * headline: periodic reduced-unit Lennard-Jones, 1,048,576 atoms (64^3 FCC
    cells x4) at density 0.8442, T* = 0.728 seeded Maxwell-Boltzmann
    velocities, cutoff 2.5 sigma with 0.3 sigma skin (cells of edge >= 2.8),
    FP32 pair arithmetic with FP64 accumulation and FP64 Gear 5th-order
    predictor-corrector, dt = 0.005.  One step = [cell-list rebuild when
    due], LJ forces, corrector fused with the next predictor.  --steps K
    steps (default 1000, the configuration's run length, so the list
    rebuilds are sampled over the whole run) are timed by the host clock
    around the run with no per-launch events; per-kernel times come from a
    second pass of K steps with HIP events around every launch.
  * subrecords (the rest of the metric, each with its own roofline and CPU
    baseline): Tersoff silicon 262,144 atoms (32^3 diamond cells x8, 0.05 A
    displacements, 1000 K, R 2.7 A, D 0.3 A, list radius 3.3 A, dt 1 fs,
    500 steps); the learned surrogate on the same system against plain
    Tersoff (speed ratio, force RMSE, fallback fraction); 512 batched LJ
    replicas (256-8788 atoms, log-uniform, rho U[0.70, 0.95], T*
    U[0.5, 2.0], 2000 steps).

Roofline: every configuration is priced at the algorithmic 312 B per
atom-step (the FP64 Gear state r, v, a, b, c read and written, the force
written and read back, the positions read once more) over the mean duration
of the kernels that make up one step.

--workload tersoff | surrogate | replicas prints only that record (for
profiling); --no-sub skips the subrecords, --no-cpu the CPU legs.  --gpus
above 1 is refused: the engine runs on one GPU.
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

METRIC = "atom-steps/s per GPU: LJ 1M, Tersoff 262k; surrogate vs plain at force RMSE"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
MFMA_BF16_PEAK_TFLOPS = 2500.0
DT = float(np.float32(0.005))  # DT is a float in the configuration (common/Global.h:79)
# algorithmic bytes per atom-step: Gear state r, v, a, b, c read and written
# (240), force written and read back (48), positions read by the force pass (24)
STEP_BYTES_PER_ATOM = 312
PROFILES = ROOT / "profiles"


def measured_traffic(name: str) -> int | None:
    """HBM bytes per launch recorded by a committed rocprofv3 summary."""
    try:
        for line in (PROFILES / name).read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1000)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--ncell", type=int, default=64, help="FCC cells per edge (4 ncell^3 atoms)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    p.add_argument("--no-sub", action="store_true", help="headline only")
    p.add_argument("--workload", choices=["lj", "tersoff", "replicas", "surrogate", "sweep"], default="lj")
    p.add_argument("--sub-steps", type=int, default=0,
                   help="steps of the subrecords (0: their configurations' 500 / 500 / 2000)")
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


def log(msg: str) -> None:
    print(f"bench: {msg}", file=sys.stderr, flush=True)


# ---------------- host facts for the CPU legs ----------------
def host_facts() -> dict:
    model = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": avail, "cpu_count": os.cpu_count()}


def omp_threads() -> int:
    """Threads for the all-cores variant: OMP_NUM_THREADS when set, else the
    cores this process may use, at most 16 (one GPU's share of a host)."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, min(16, host_facts()["nproc"]))


def oracle():
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle as O
    return O


# ---------------- timing helpers ----------------
def timed_run(obj, steps: int) -> float:
    """Host-clock seconds of `steps` steps, bracketed by synchronizations,
    with no per-launch events."""
    obj.synchronize()
    t0 = time.perf_counter()
    obj.run(steps)
    obj.synchronize()
    return time.perf_counter() - t0


def kernel_pass(obj, steps: int) -> dict:
    """Mean ms per launch of each kernel over `steps` more steps, with HIP
    events recorded on the system's stream around every launch."""
    obj.set_timing(True)
    obj.run(steps)
    obj.synchronize()
    times = obj.kernel_times()
    obj.set_timing(False)
    return {k: (v[0], v[1] / max(v[0], 1)) for k, v in times.items()}


def roofline(kernels: list[str], times: dict, atoms: int, traffic: int | None, what: str) -> dict:
    """312 B per atom-step over the summed mean duration of the step's kernels."""
    ms = sum(times.get(k, (0, 0.0))[1] for k in kernels)
    ok = all(k in times for k in kernels) and ms > 0
    achieved = atoms * STEP_BYTES_PER_ATOM / (ms * 1e-3) / 1e9 if ok else None
    return {"kernel": what, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if ok else None, "traffic": traffic, "avg_launch_ms": ms if ok else None,
            "atoms": atoms, "bytes_per_atom": STEP_BYTES_PER_ATOM}


def record(value, steps, warmup, elapsed, dtype, data, config) -> dict:
    return {"metric": METRIC, "value": value, "unit": "atom-steps/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": 1e3 * elapsed / steps, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": data, "config": config}


# ---------------- LJ 1M (headline) ----------------
def cpu_baseline_lj(pos, vel, L, steps_1t: int, steps_omp: int) -> dict:
    """The FP64 oracle's periodic LJ loop (27-cell candidate lists in the
    reference's format, Gear PC) on the headline system itself: one thread
    for 8 steps (about one list rebuild; ~3.7 s a step at 1M atoms), and its OpenMP
    variant on all of this process's cores for the full step count."""
    O = oracle()
    n = len(pos)
    facts = host_facts()
    _, _, _, b1, el1 = O.pbc_lj_run(pos, vel, L, 2.5, 2.8, DT, steps_1t, threads=1)
    th = omp_threads()
    _, _, _, bt, elt = O.pbc_lj_run(pos, vel, L, 2.5, 2.8, DT, steps_omp, threads=th)
    return {
        "value": n * steps_1t / el1, "unit": "atom-steps/s", "cores": 1, "kind": "port",
        "sample": f"the headline system ({n} atoms, same lattice, seed, density, T*, cutoffs), {steps_1t} steps "
                  f"({b1 - 1} list rebuilds), {el1:.1f} s on one thread of {facts['cpu_model']}",
        **facts, "compiler": O.build_info(),
        "openmp": {"value": n * steps_omp / elt, "unit": "atom-steps/s", "cores": th,
                   "sample": f"the same system, {steps_omp} steps ({bt - 1} list rebuilds), {elt:.1f} s on {th} threads"},
    }


def bench_lj(a) -> dict:
    from moleculardynamics_amd import BOX_PERIODIC, LJ_UNIFORM, PREC_MIXED, System, initial

    pos, vel, L = initial.lj_fluid(a.ncell, 0.8442, 0.728, seed=2024)
    n = len(pos)
    log(f"LJ {n} atoms, L={L:.4f}; warmup {a.warmup}, steps {a.steps}")
    s = System(n_atoms=n, box=BOX_PERIODIC, precision=PREC_MIXED, use_lj=1, lj_table=LJ_UNIFORM,
               dt=0.005, lj_rs=2.8, lj_rcut=2.5, box_length=L, lj_sigma=1.0, lj_eps=1.0, mass=1.0)
    shape = s.launch_shape()
    s.set_state(np.c_[pos, np.zeros(n)], vel)
    s.run(a.warmup)
    b0 = s.build_count()
    elapsed = timed_run(s, a.steps)
    builds = s.build_count() - b0
    times = kernel_pass(s, a.steps)
    s.close()
    value = n * a.steps / elapsed
    log(f"{elapsed:.4f} s, {builds} list builds; kernels {times}")
    out = record(value, a.steps, a.warmup, elapsed, "fp32-pair/fp64-accum",
                 "synthetic (seeded FCC lattice + Maxwell-Boltzmann velocities)",
                 {"workload": f"periodic LJ {n} atoms (FCC {a.ncell}^3 x4), rho 0.8442, T* 0.728, "
                              f"rc 2.5 skin 0.3, dt 0.005", "atoms": n, "list_builds": builds,
                  "launch_shape": shape})
    out["roofline"] = roofline(["md_step"], times, n, measured_traffic("md_step_lj1m.md"),
                               "k_step_lds (md_step: LJ force, corrector, next predictor)")
    out["roofline"]["traffic_source"] = "profiles/md_step_lj1m.md"
    out["roofline"]["step_gbs_at_312B_per_atom"] = value * STEP_BYTES_PER_ATOM / 1e9
    out["roofline"]["step_frac"] = value * STEP_BYTES_PER_ATOM / 1e9 / HBM_PEAK_GBS
    out["kernel_ms"] = {k: v[1] for k, v in times.items()}
    out["kernel_launches"] = {k: v[0] for k, v in times.items()}
    if not a.no_cpu:
        log("cpu baseline (LJ) ...")
        out["cpu_baseline"] = cpu_baseline_lj(pos, vel, L, 8, 20)
    return out


# ---------------- Tersoff silicon 262k ----------------
def cpu_baseline_tersoff(ncell: int, steps: int) -> dict:
    """The FP64 oracle's Tersoff forces (cell search, O(N L^2) over each
    atom's neighbours: the reference's N x N scratch cannot hold 262k atoms)
    with its Gear predictor-corrector, single-threaded, on a sub-box sample."""
    O = oracle()
    from moleculardynamics_amd import initial

    pos, vel, L = initial.si_diamond(ncell, 1000.0, seed=12345, disp=0.05)
    t0 = time.perf_counter()
    O.tersoff_run(pos, vel, L, 1.0, steps, initial.SI_MASS, initial.ACCEL_EV_A_AMU)
    elapsed = time.perf_counter() - t0
    return {"value": len(pos) * steps / elapsed, "unit": "atom-steps/s", "cores": 1, "kind": "port",
            "sample": f"{len(pos)} atoms ({ncell}^3 diamond x8, same T/displacements), {steps} steps, forces every "
                      f"step by the O(N L^2) cell-search restatement, {elapsed:.1f} s on one thread",
            **host_facts(), "compiler": O.build_info()}


def silicon(a):
    from moleculardynamics_amd import initial
    return initial.si_diamond(32, 1000.0, seed=2024, disp=0.05)


def bench_tersoff(a, steps: int) -> dict:
    from moleculardynamics_amd import BOND_TERSOFF
    from moleculardynamics_amd import surrogate as SG

    pos, vel, L = silicon(a)
    n = len(pos)
    log(f"Tersoff Si {n} atoms, L={L:.4f}; warmup {a.warmup}, steps {steps}")
    s = SG.silicon_system(n, L, bond_form=BOND_TERSOFF, mb_rs=a.mb_rs)
    s.set_state(np.c_[pos, np.zeros(n)], vel)
    s.run(a.warmup)
    b0 = s.build_count()
    elapsed = timed_run(s, steps)
    builds = s.build_count() - b0
    times = kernel_pass(s, min(steps, 100))
    s.close()
    log(f"Tersoff {elapsed:.4f} s, {builds} list builds; kernels {times}")
    out = record(n * steps / elapsed, steps, a.warmup, elapsed, "fp32-terms/fp64-accum",
                 "synthetic (seeded diamond lattice + Gaussian displacements + Maxwell-Boltzmann velocities)",
                 {"workload": f"periodic Tersoff Si {n} atoms (diamond 32^3 x8, a 5.431 A), 1000 K, R 2.7 D 0.3, "
                              f"list radius {a.mb_rs} A, dt 1 fs", "atoms": n, "list_builds": builds})
    out["metric"] = "atom-steps/s per GPU (Tersoff Si 262k, mixed precision)"
    out["roofline"] = roofline(["ts_atoms", "gather_gear_step"], times, n, None,
                               "k_ts_atoms + k_correct_predict_track (the step's two kernels)")
    out["kernel_ms"] = {k: v[1] for k, v in times.items()}
    if not a.no_cpu:
        log("cpu baseline (Tersoff) ...")
        out["cpu_baseline"] = cpu_baseline_tersoff(16, 150)
    return out


# ---------------- surrogate vs plain Tersoff ----------------
SG_MFMA_FLOP_PER_ATOM = 2 * (64 * 64 + 64 * 64 + 64 * 64 + 48 * 64)   # as issued: 4 layers, K 48 padded to 64


def cpu_baseline_surrogate(ncell: int, steps: int, blob) -> dict:
    """The FP64 oracle's surrogate (numpy: descriptors, network, analytic
    forces) with the oracle's Gear predictor-corrector, on a sub-box sample."""
    O = oracle()
    from threadpoolctl import threadpool_limits

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
    with threadpool_limits(1):   # one core, as stated below
        t0 = time.perf_counter()
        _, f, _, _ = S.energy_forces(p, r, L, tersoff_fallback=fb)
        a_[:] = f / m * sc
        for _ in range(steps):
            r, v, a_, b_, c_ = O.gear_predict(r, v, a_, b_, c_, 1.0)
            _, f, _, _ = S.energy_forces(p, np.mod(r, L), L, tersoff_fallback=fb)
            r, v, a_, b_, c_ = O.gear_correct(r, v, f, a_, b_, c_, 1.0, m, sc)
        elapsed = time.perf_counter() - t0
    return {"value": len(pos) * steps / elapsed, "unit": "atom-steps/s", "cores": 1, "kind": "port",
            "sample": f"{len(pos)} atoms ({ncell}^3 diamond x8, 1000 K), {steps} steps of the numpy FP64 surrogate "
                      f"(brute-force neighbours), {elapsed:.1f} s on one thread", **host_facts()}


def bench_surrogate(a, steps: int) -> dict:
    from moleculardynamics_amd import BOND_SURROGATE, BOND_TERSOFF, LIST_BOND
    from moleculardynamics_amd import surrogate as SG

    blob = SG.load_weights(a.weights) if a.weights else SG.load_weights()
    pos, vel, L = silicon(a)
    n = len(pos)
    log(f"surrogate Si {n} atoms; warmup {a.warmup}, steps {steps}")

    def timed(form):
        s = SG.silicon_system(n, L, bond_form=form, mb_rs=a.mb_rs)
        if form == BOND_SURROGATE:
            s.set_surrogate(blob)
        s.set_state(np.c_[pos, np.zeros(n)], vel)
        s.run(a.warmup)
        elapsed = timed_run(s, steps)
        times = kernel_pass(s, min(steps, 100))
        return s, elapsed, times

    sp, el_plain, t_plain = timed(BOND_TERSOFF)
    final = np.mod(sp.get_state()["pos"][:, :3], L)
    sp.close()
    ss, elapsed, times = timed(BOND_SURROGATE)
    fallback, evals = ss.surrogate_counts()
    ss.close()
    log(f"plain {el_plain:.4f} s, surrogate {elapsed:.4f} s; kernels {times}")

    def forces(form):   # on the configuration plain Tersoff reached
        with SG.silicon_system(n, L, bond_form=form) as s:
            if form == BOND_SURROGATE:
                s.set_surrogate(blob)
            s.set_state(np.c_[final, np.zeros(n)], np.zeros((n, 3)))
            s.build_list(LIST_BOND)
            s.bond_forces()
            return s.get_state()["force"]

    f_plain, f_sg = forces(BOND_TERSOFF), forces(BOND_SURROGATE)
    value, value_plain = n * steps / elapsed, n * steps / el_plain
    out = record(value, steps, a.warmup, elapsed, "bf16-mfma-network/fp32-descriptors/fp64-accum",
                 "synthetic (seeded diamond lattice + Gaussian displacements + Maxwell-Boltzmann velocities); "
                 "network trained on GPU plain-Tersoff frames",
                 {"workload": f"periodic Tersoff-Si surrogate {n} atoms (diamond 32^3 x8), 1000 K, "
                              f"list radius {a.mb_rs} A, dt 1 fs", "atoms": n})
    out["metric"] = "atom-steps/s per GPU (Tersoff-Si surrogate 262k) vs plain Tersoff at force RMSE"
    out["plain_tersoff"] = {"value": value_plain, "ms_per_step": 1e3 * el_plain / steps,
                            "kernel_ms": {k: v[1] for k, v in t_plain.items()}}
    out["speedup_vs_plain"] = value / value_plain
    out["force_rmse_vs_plain_eV_per_A"] = float(np.sqrt(np.mean((f_sg - f_plain) ** 2)))
    out["force_rms_plain_eV_per_A"] = float(np.sqrt(np.mean(f_plain ** 2)))
    out["fallback_fraction"] = fallback / max(evals, 1)
    out["roofline"] = roofline(["sg_eval", "gather_gear_step"], times, n, None,
                               "k_sg_eval + k_correct_predict_track (the step's two kernels)")
    cnt, k_ms = times.get("sg_eval", (0, 0.0))
    out["mfma"] = {"kernel": "k_sg_eval", "flop_per_atom": SG_MFMA_FLOP_PER_ATOM,
                   "achieved_tflops": n * SG_MFMA_FLOP_PER_ATOM / (k_ms * 1e-3) / 1e12 if cnt else None,
                   "peak_tflops": MFMA_BF16_PEAK_TFLOPS}
    out["kernel_ms"] = {k: v[1] for k, v in times.items()}
    if not a.no_cpu:
        log("cpu baseline (surrogate) ...")
        out["cpu_baseline"] = cpu_baseline_surrogate(4, 200, blob)
    return out


# ---------------- 512 replicas ----------------
def cpu_baseline_replicas(ns, Ls, pos, vel, max_atoms: int, steps: int) -> dict:
    """The FP64 oracle's periodic LJ loop (27-cell lists, Gear PC) run replica
    after replica, as the reference's sweep runs its systems, on the first
    replicas wide enough for its 3-cell grid, about max_atoms atoms."""
    O = oracle()
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
    return {"value": atoms * steps / elapsed, "unit": "atom-steps/s", "cores": 1, "kind": "port",
            "sample": f"{done} replicas of the batch ({atoms} atoms), {steps} steps each, one after another, "
                      f"{elapsed:.1f} s on one thread", **host_facts(), "compiler": O.build_info()}


def bench_replicas(a, steps: int) -> dict:
    from moleculardynamics_amd import PREC_MIXED, Batch, initial

    ns, Ls, pos, vel = initial.replica_batch(512, seed=2024)
    n = int(ns.sum())
    log(f"{len(ns)} LJ replicas, {n} atoms; warmup {a.warmup}, steps {steps}")
    b = Batch(ns, Ls, precision=PREC_MIXED, dt=0.005, lj_rs=2.8, lj_rcut=2.5)
    shape = b.launch_shape()
    staged = b.staged()
    n_staged = int(ns[staged].sum())
    n_small = int(ns[b.paths() == 2].sum())
    b.set_state(pos, vel)
    b.run(a.warmup)
    builds0 = b.build_counts().sum()
    elapsed = timed_run(b, steps)              # graph replays
    builds = int(b.build_counts().sum() - builds0)
    times = kernel_pass(b, min(steps, 200))    # the same steps launched one by one with events
    b.close()
    log(f"replicas {elapsed:.4f} s, {builds} replica list builds; kernels {times}")
    out = record(n * steps / elapsed, steps, a.warmup, elapsed, "fp32-pair/fp64-accum",
                 "synthetic (seeded FCC blocks + Maxwell-Boltzmann velocities per replica)",
                 {"workload": f"{len(ns)} periodic LJ replicas, {n} atoms ({int(ns.min())}-{int(ns.max())} per "
                              f"replica), rho U[0.70,0.95], T* U[0.5,2.0], rc 2.5 skin 0.3, dt 0.005", "atoms": n,
                  "replica_list_builds": builds, "staged_replicas": int(staged.sum()), "staged_atoms": n_staged,
                  "all_pairs_atoms": n_small, "launch_shape": shape})
    out["metric"] = "atom-steps/s per GPU (512 LJ replicas, mixed precision)"
    out["roofline"] = roofline(["rep_step_lds"], times, n_staged, None,
                               "k_step_lds over the staged replicas' segments (their whole step)")
    out["kernel_ms"] = {k: v[1] for k, v in times.items()}
    if not a.no_cpu:
        log("cpu baseline (replicas) ...")
        out["cpu_baseline"] = cpu_baseline_replicas(ns, Ls, pos, vel, 60000, 20)
    return out


# ---------------- the vxcm sweep of a walled N/O/Ne gas ----------------
SWEEP_VXCM = [-0.003 + 0.001 * k for k in range(8)]   # VXCMFROM -0.003, VXCMTO 0.004, VXCMSTEP 0.001


def sweep_config(n: int, n_bond: int) -> dict:
    from moleculardynamics_amd import BOND_REFERENCE, BOX_WALLED, LJ_REFERENCE, PREC_FP64
    return dict(n_atoms=n, n_bond=n_bond, box=BOX_WALLED, precision=PREC_FP64, use_lj=1, use_bond=1,
                lj_table=LJ_REFERENCE, bond_form=BOND_REFERENCE, dt=1.0, lj_rs=7.6, lj_rcut=6.8, mb_rs=3.6,
                mb_rcut=3.0, energy_loss=1.0, wall=n ** (1 / 3) * 52.8)


def cpu_baseline_sweep(pos4, vel, n_bond: int, steps: int) -> dict:
    """The FP64 oracle's run of the reference loop (cpu/cpumd/Host.cpp:141-195)
    for the sweep's first member."""
    O = oracle()
    n = len(pos4)
    _, _, _, el = O.ref_run(pos4, vel, n - n_bond, n_bond, steps, 10 ** 9, 1.0, 7.6, 6.8, 3.6, 3.0,
                            vxcm=float(np.float32(SWEEP_VXCM[0])), use_lj=True, use_mb=True)
    return {"value": n * steps / el, "unit": "atom-steps/s", "cores": 1, "kind": "port",
            "sample": f"the first member ({n} atoms, vxcm {SWEEP_VXCM[0]}), {steps} steps of the reference loop, "
                      f"{el:.1f} s on one thread", **host_facts(), "compiler": O.build_info()}


def bench_sweep(a, steps: int) -> dict:
    from moleculardynamics_amd import Sweep, System, initial

    n_mol, n_ne = 30, 4000
    pos4, vel = initial.walled_mixture(n_mol, n_ne, 300.0, seed=7)
    n, nb, R = len(pos4), 2 * n_mol, len(SWEEP_VXCM)
    cfg = sweep_config(n, nb)
    starts = []
    for vx in SWEEP_VXCM:
        v = vel.copy()
        v[:, 0] += float(np.float32(vx))
        starts.append(v)
    log(f"vxcm sweep: {R} members of {n} atoms ({nb} bond order), warmup {a.warmup}, steps {steps}")
    # the members one after another, as the reference runs them (one system each)
    seq = 0.0
    seq_builds = 0
    for v in starts:
        with System(**cfg) as s:
            s.set_state(pos4, v)
            s.run(a.warmup)
            seq += timed_run(s, steps)
            seq_builds += s.build_count()
    w = Sweep(R, **cfg)
    for m, v in enumerate(starts):
        w.member(m).set_state(pos4, v)
    w.run(a.warmup)
    b0 = sum(w.member(m).build_count() for m in range(R))
    w.member(0).synchronize()
    t0 = time.perf_counter()
    w.run(steps)
    w.member(0).synchronize()
    elapsed = time.perf_counter() - t0
    builds = sum(w.member(m).build_count() for m in range(R)) - b0
    w.set_timing(True)
    w.run(min(steps, 100))
    times = w.member(0).kernel_times()
    w.set_timing(False)
    w.close()
    times = {k: (v[0], v[1] / max(v[0], 1)) for k, v in times.items()}
    log(f"sweep {elapsed:.4f} s ({builds} member list builds), one after another {seq:.4f} s; kernels {times}")
    out = record(R * n * steps / elapsed, steps, a.warmup, elapsed, "fp64",
                 "synthetic (seeded N2/O2/NO + Ne gas on a jittered grid, Maxwell-Boltzmann velocities at 300 K)",
                 {"workload": f"vxcm sweep of {R} members (vxcm -0.003 .. 0.004 step 0.001), walled N/O/Ne gas of "
                              f"{n} atoms ({nb} bond order), LJ rs 7.6 rc 6.8, bond order rs 3.6 rc 3.0, dt 1, walls "
                              f"at N^(1/3) 52.8", "atoms": n, "members": R, "member_list_builds": builds})
    out["metric"] = "atom-steps/s per GPU (vxcm sweep of a walled N/O/Ne gas, FP64)"
    out["one_after_another"] = {"value": R * n * steps / seq, "ms_per_step_all_members": 1e3 * seq / steps,
                                "member_list_builds": seq_builds, "speedup_of_batch": seq / elapsed}
    step_kernels = ["sweep_predict", "sweep_lj", "sweep_bond", "sweep_correct", "sweep_accelerations"]
    out["roofline"] = roofline(step_kernels, times, R * n, None,
                               "the batched step's kernels (predict, LJ, bond order, corrector, gated a = F/m)")
    out["kernel_ms"] = {k: v[1] for k, v in times.items()}
    if not a.no_cpu:
        log("cpu baseline (sweep) ...")
        out["cpu_baseline"] = cpu_baseline_sweep(pos4, vel, nb, 2000)
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "the engine runs on one GPU: bench.py measures --gpus 1 only"}))
        sys.exit(2)
    if a.copy_bytes:
        copy_calibration(a.copy_bytes)
        return
    sub = {"tersoff": a.sub_steps or 500, "surrogate": a.sub_steps or 500, "replicas": a.sub_steps or 2000,
           "sweep": a.sub_steps or 1000}
    if a.workload == "tersoff":
        print(json.dumps(bench_tersoff(a, a.steps)))
        return
    if a.workload == "surrogate":
        print(json.dumps(bench_surrogate(a, a.steps)))
        return
    if a.workload == "replicas":
        print(json.dumps(bench_replicas(a, a.steps)))
        return
    if a.workload == "sweep":
        print(json.dumps(bench_sweep(a, a.steps)))
        return
    out = bench_lj(a)
    if not a.no_sub:
        out["subrecords"] = {"tersoff_262k": bench_tersoff(a, sub["tersoff"]),
                             "surrogate_vs_plain_262k": bench_surrogate(a, sub["surrogate"]),
                             "replicas_512": bench_replicas(a, sub["replicas"]),
                             "vxcm_sweep_8": bench_sweep(a, sub["sweep"])}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
