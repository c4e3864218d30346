"""Launch-shape tuner for the LDS-staged kernels.

    python -m moleculardynamics_amd.tune [--out PATH] [--steps K] [--reps R]

runs on the GPU and writes the tuning table (default: data/tuning.json beside
libmdx.so, which mdx_create and mdx_batch_create read).  For each size bucket
of each kind it builds one probe system and times every candidate shape on it:

  * kind "lj" (the periodic Lennard-Jones box, FCC at density 0.8442, T* 0.728,
    rc 2.5, rs 2.8): workgroups of 512 threads with segment fill 0.7, 0.75,
    0.8, 0.85, 0.9 or 1.0 of one pass, or 1024 threads with 0.7, 0.85 or 1.0.  Every candidate starts from the same state and
    the candidates advance in lockstep; the shape does not change any atom's
    force, so they follow the same trajectory with the same list rebuilds.
  * kind "replicas" (512 LJ replicas of 256-8192 atoms, the batch benchmark's
    configuration and seed): staging aim 900, 1250 or 1600 atoms x fill 0.7 or 0.85.

A candidate is timed through the same path a user's system takes: a one-entry
table in a temporary file named by MDX_TUNING_FILE.  Each timing is the host
clock over K steps after a warm-up, repeated R times round robin over the
candidates, keeping each candidate's fastest.  The built-in shape is one of the
candidates, and a bucket's entry takes another shape only when that shape is
faster than the built-in one by more than 2 %; so a tuned shape is never
slower than the built-in shape on the probe.  Each entry records the timings.
"""
from __future__ import annotations

import argparse
import json
import os
import tempfile
import time
from pathlib import Path

import numpy as np

from . import BOX_PERIODIC, LJ_UNIFORM, PREC_MIXED, Batch, System, initial

ARCH = "gfx950"
TABLE = Path(__file__).resolve().parent / "data" / "tuning.json"
MARGIN = 0.02
DT = 0.005

# (atoms_min, atoms_max, FCC cells per edge of the probe)
LJ_BUCKETS = [(0, 131072, 20), (131072, 2097152, 64), (2097152, 2 ** 31 - 1, 100)]
LJ_CANDIDATES = ([{"block": 512, "fill": f} for f in (0.7, 0.75, 0.8, 0.85, 0.9, 1.0)] +
                 [{"block": 1024, "fill": f} for f in (0.7, 0.85, 1.0)])
REPLICA_CANDIDATES = [{"stage_aim": a, "fill": f} for a in (900.0, 1250.0, 1600.0) for f in (0.7, 0.85)]
DEFAULT = {"block": 512, "fill": 0.85, "stage_aim": 1250.0}


def log(msg: str) -> None:
    print(f"tune: {msg}", flush=True)


class TableFile:
    """A temporary one-entry table named by MDX_TUNING_FILE while in use."""

    def __init__(self, entry: dict | None):
        self.entry = entry

    def __enter__(self):
        fd, self.path = tempfile.mkstemp(suffix=".json")
        with os.fdopen(fd, "w") as f:
            json.dump({"format": "mdx-tuning-1", "entries": [self.entry] if self.entry else []}, f)
        self.old = os.environ.get("MDX_TUNING_FILE")
        os.environ["MDX_TUNING_FILE"] = self.path
        return self

    def __exit__(self, *exc):
        if self.old is None:
            os.environ.pop("MDX_TUNING_FILE", None)
        else:
            os.environ["MDX_TUNING_FILE"] = self.old
        os.unlink(self.path)


def timed(obj, steps: int) -> float:
    obj.synchronize()
    t0 = time.perf_counter()
    obj.run(steps)
    obj.synchronize()
    return time.perf_counter() - t0


def pick(default_key, results: list[tuple[dict, float]]):
    """The fastest candidate when it beats the built-in shape by the margin,
    else the built-in shape."""
    base = next(t for c, t in results if default_key(c))
    best_c, best_t = min(results, key=lambda ct: ct[1])
    if best_t < base * (1.0 - MARGIN):
        return best_c, best_t, base
    return next(c for c, _ in results if default_key(c)), base, base


def tune_lj(steps: int, reps: int, warmup: int) -> list[dict]:
    entries = []
    for lo, hi, ncell in LJ_BUCKETS:
        pos, vel, L = initial.lj_fluid(ncell, 0.8442, 0.728, seed=2024)
        n = len(pos)
        log(f"lj bucket [{lo}, {hi}): probe {n} atoms")
        systems = []
        for cand in LJ_CANDIDATES:
            entry = {"arch": ARCH, "kind": "lj", "atoms_min": 0, "atoms_max": 2 ** 31 - 1, **cand}
            with TableFile(entry):
                s = System(n_atoms=n, box=BOX_PERIODIC, precision=PREC_MIXED, use_lj=1, lj_table=LJ_UNIFORM,
                           dt=DT, lj_rs=2.8, lj_rcut=2.5, box_length=L, lj_sigma=1.0, lj_eps=1.0, mass=1.0)
            assert s.launch_shape()["source"] == "tuned"
            s.set_state(np.c_[pos, np.zeros(n)], vel)
            s.run(warmup)
            systems.append(s)
        best = [float("inf")] * len(systems)
        for _ in range(reps):
            for k, s in enumerate(systems):
                best[k] = min(best[k], timed(s, steps))
        for s in systems:
            s.close()
        results = [(c, 1e3 * t / steps) for c, t in zip(LJ_CANDIDATES, best)]
        for c, ms in results:
            log(f"  block {c['block']:4d} fill {c['fill']:.2f}: {ms:.4f} ms/step")
        chosen, ms, base = pick(lambda c: c["block"] == DEFAULT["block"] and c["fill"] == DEFAULT["fill"], results)
        entries.append({"arch": ARCH, "kind": "lj", "atoms_min": lo, "atoms_max": hi, **chosen,
                        "ms_per_step": round(ms, 5), "default_ms_per_step": round(base, 5),
                        "probe": f"periodic LJ {n} atoms (FCC {ncell}^3 x4), rho 0.8442, T* 0.728, rc 2.5, rs 2.8, "
                                 f"mixed precision, {reps} x {steps} steps after {warmup}",
                        "candidates": [{**c, "ms_per_step": round(t, 5)} for c, t in results]})
    return entries


def tune_replicas(steps: int, reps: int, warmup: int) -> list[dict]:
    n_atoms, boxes, pos, vel = initial.replica_batch(512, seed=2024)
    total = int(np.sum(n_atoms))
    log(f"replicas: probe 512 replicas, {total} atoms")
    batches = []
    for cand in REPLICA_CANDIDATES:
        entry = {"arch": ARCH, "kind": "replicas", "atoms_min": 0, "atoms_max": 2 ** 31 - 1, **cand}
        with TableFile(entry):
            b = Batch(n_atoms, boxes, precision=PREC_MIXED, dt=DT, lj_rs=2.8, lj_rcut=2.5)
        assert b.launch_shape()["source"] == "tuned"
        b.set_state(pos, vel)
        b.run(warmup)
        batches.append(b)
    best = [float("inf")] * len(batches)
    for _ in range(reps):
        for k, b in enumerate(batches):
            best[k] = min(best[k], timed(b, steps))
    for b in batches:
        b.close()
    results = [(c, 1e3 * t / steps) for c, t in zip(REPLICA_CANDIDATES, best)]
    for c, ms in results:
        log(f"  stage_aim {c['stage_aim']:6.0f} fill {c['fill']:.2f}: {ms:.4f} ms/step")
    chosen, ms, base = pick(lambda c: c["stage_aim"] == DEFAULT["stage_aim"] and c["fill"] == DEFAULT["fill"],
                            results)
    return [{"arch": ARCH, "kind": "replicas", "atoms_min": 0, "atoms_max": 2 ** 31 - 1, **chosen,
             "ms_per_step": round(ms, 5), "default_ms_per_step": round(base, 5),
             "probe": f"512 LJ replicas ({total} atoms, 256-8192 each, log-uniform, seed 2024), rho U[0.70, 0.95], "
                      f"T* U[0.5, 2.0], mixed precision, {reps} x {steps} steps after {warmup}",
             "candidates": [{**c, "ms_per_step": round(t, 5)} for c, t in results]}]


def main() -> None:
    p = argparse.ArgumentParser()
    p.add_argument("--out", default=str(TABLE))
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--reps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--kinds", default="lj,replicas")
    a = p.parse_args()
    kinds = a.kinds.split(",")
    entries = []
    if "lj" in kinds:
        entries += tune_lj(a.steps, a.reps, a.warmup)
    if "replicas" in kinds:
        entries += tune_replicas(a.steps, a.reps, a.warmup)
    table = {"format": "mdx-tuning-1",
             "note": "written by python -m moleculardynamics_amd.tune on one MI355X; a shape differs from the "
                     "built-in one only where it was faster by more than 2 % on the probe",
             "entries": entries}
    out = Path(a.out)
    out.parent.mkdir(parents=True, exist_ok=True)
    tmp = out.with_suffix(".tmp")
    tmp.write_text(json.dumps(table, indent=1) + "\n")
    os.replace(tmp, out)
    log(f"wrote {out}")


if __name__ == "__main__":
    main()
