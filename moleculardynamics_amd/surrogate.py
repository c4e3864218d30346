"""Learned surrogate of Tersoff silicon: training set, training loop and the
shipped network.

Everything that computes runs on the GPU through the C-ABI:
  * frames come from plain-Tersoff trajectories of the engine (FP64 Tersoff,
    periodic 512-atom diamond boxes, 48 temperatures over 300-2000 K, three
    of them held out);
  * training is minibatch Adam on the surrogate's own FP32 forward/backward
    kernels (mdx_sg_trainer_*), loss = force MSE + w_E * (energy/atom)^2;
  * inference (bond_form BOND_SURROGATE) runs the network on BF16 MFMA.

    python -m moleculardynamics_amd.surrogate --out moleculardynamics_amd/data/surrogate_si.npy
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

from . import initial
from ._lib import (BOND_SURROGATE, BOND_TERSOFF, BOX_PERIODIC, LIST_BOND, LJ_UNIFORM, PREC_FP64, PREC_MIXED,
                   SG_NFEAT, SG_NPARAM, SgTrainer, System)

PKG = Path(__file__).resolve().parent
WEIGHTS = PKG / "data" / "surrogate_si.npy"
NHID = 64
NTRAIN = NHID * SG_NFEAT + NHID + NHID * NHID + NHID + NHID + 1


def silicon_system(n, L, bond_form=BOND_TERSOFF, precision=PREC_MIXED, mb_rs=3.3):
    """Periodic silicon box in A, fs, amu, eV units (Tersoff or its surrogate)."""
    return System(n_atoms=n, n_bond=n, box=BOX_PERIODIC, precision=precision, use_lj=0, use_bond=1,
                  lj_table=LJ_UNIFORM, bond_form=bond_form, dt=1.0, mb_rs=mb_rs, mb_rcut=3.0, box_length=L,
                  mass=initial.SI_MASS, accel_scale=initial.ACCEL_EV_A_AMU)


def tersoff_forces(pos, L, precision=PREC_FP64):
    """Plain-Tersoff energy and forces of one configuration on the GPU."""
    n = len(pos)
    with silicon_system(n, L, precision=precision) as s:
        s.set_state(np.c_[np.mod(pos, L), np.zeros(n)], np.zeros((n, 3)))
        s.build_list(LIST_BOND)
        s.bond_forces()
        f = s.get_state()["force"]
        e = s.potential()[1]
    return e, f


def tersoff_frames(n_traj=48, frames_per_traj=84, t_min=300.0, t_max=2000.0, ncell=4, stride=20, equil=200,
                   seed=1000, log=None):
    """Frames of plain-Tersoff MD (Gear PC, dt = 1 fs): trajectory k starts
    from diamond at twice its target temperature T_k (equipartition halves
    it), runs `equil` steps, then gives one frame every `stride` steps.
    Returns positions [F, n, 3], forces [F, n, 3], energies [F], the target
    temperature of each frame and the box edge."""
    temps = np.linspace(t_min, t_max, n_traj)
    P, Fo, E, T = [], [], [], []
    L = None
    for k, tk in enumerate(temps):
        pos, vel, L = initial.si_diamond(ncell, 2.0 * tk, seed + k, disp=0.05)
        n = len(pos)
        with silicon_system(n, L, precision=PREC_FP64) as s, silicon_system(n, L, precision=PREC_FP64) as ev:
            s.set_state(np.c_[pos, np.zeros(n)], vel)
            s.run(equil)
            for _ in range(frames_per_traj):
                s.run(stride)
                p = np.mod(s.get_state()["pos"][:, :3], L)
                ev.set_state(np.c_[p, np.zeros(n)], np.zeros((n, 3)))
                ev.build_list(LIST_BOND)
                ev.bond_forces()
                P.append(p)
                Fo.append(ev.get_state()["force"])
                E.append(ev.potential()[1])
                T.append(tk)
        if log:
            log(f"trajectory {k + 1}/{n_traj} T={tk:.0f} K done")
    return np.array(P), np.array(Fo), np.array(E), np.array(T), L


def init_params(stats, e_per_atom, seed=0, margin=0.1):
    """Network initialisation: standardisation and accepted range from the
    descriptor statistics of the training set, small random weights."""
    rng = np.random.default_rng(seed)
    mean, std = stats["mean"], stats["std"]
    istd = 1.0 / (std + 1e-4)
    gmin, gmax = (stats["min"] - mean) * istd, (stats["max"] - mean) * istd
    span = gmax - gmin
    blob = np.concatenate([
        rng.normal(0, 1 / np.sqrt(SG_NFEAT), NHID * SG_NFEAT), np.zeros(NHID),
        rng.normal(0, 1 / np.sqrt(NHID), NHID * NHID), np.zeros(NHID),
        rng.normal(0, 0.1 / np.sqrt(NHID), NHID), [e_per_atom],
        mean, istd, gmin - margin * span - 0.5, gmax + margin * span + 0.5,
    ])
    assert blob.shape == (SG_NPARAM,)
    return blob.astype(np.float32)


def train(trainer, blob, train_frames, epochs=60, batch=16, lr_max=3e-3, lr_min=2e-5, w_energy=1.0, seed=0,
          log=None, eval_frames=None):
    """Minibatch Adam with a linear warm-up and a cosine decay of the rate."""
    trainer.set_params(blob)
    rng = np.random.default_rng(seed)
    nb = len(train_frames) // batch
    total = epochs * nb
    step = 0
    for ep in range(epochs):
        order = rng.permutation(train_frames)
        for b in range(nb):
            warm = min(1.0, (step + 1) / (0.03 * total))
            lr = lr_min + (lr_max - lr_min) * warm * 0.5 * (1 + np.cos(np.pi * step / total))
            trainer.step(order[b * batch:(b + 1) * batch], lr, w_energy)
            step += 1
        if log and (ep % 5 == 4 or ep == epochs - 1):
            msg = f"epoch {ep + 1}/{epochs}"
            if eval_frames is not None:
                f_mse, e_mse = evaluate(trainer, eval_frames)
                msg += f" held-out force rmse {np.sqrt(f_mse):.4f} eV/A, energy rmse {np.sqrt(e_mse) * 1e3:.2f} meV/atom"
            log(msg)
    return trainer.get_params()


def evaluate(trainer, frames, chunk=64):
    """Force MSE and per-atom energy MSE of the FP32 network over frames."""
    f_sum, e_sum, n = 0.0, 0.0, 0
    frames = np.asarray(frames)
    for k in range(0, len(frames), chunk):
        sel = frames[k:k + chunk]
        out, _ = trainer.gradient(sel, want_grad=False)
        f_sum += out[1] * len(sel)
        e_sum += out[2] * len(sel)
        n += len(sel)
    return f_sum / n, e_sum / n


def load_weights(path=WEIGHTS):
    blob = np.load(path, allow_pickle=False).astype(np.float32)
    if blob.shape != (SG_NPARAM,):
        raise ValueError(f"{path}: expected {SG_NPARAM} floats, found {blob.shape}")
    return blob


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=str(WEIGHTS))
    ap.add_argument("--report", default=None, help="JSON summary path")
    ap.add_argument("--trajectories", type=int, default=48)
    ap.add_argument("--frames", type=int, default=84, help="frames per trajectory")
    ap.add_argument("--epochs", type=int, default=120)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    t0 = time.time()

    def log(msg):
        print(f"[{time.time() - t0:7.1f}s] {msg}", flush=True)

    P, F, E, T, L = tersoff_frames(a.trajectories, a.frames, log=log)
    log(f"{len(P)} frames of {P.shape[1]} atoms, L = {L:.3f} A")
    temps = np.unique(T)
    held = temps[[len(temps) // 8, len(temps) // 2 - 1, (7 * len(temps)) // 8]] if len(temps) >= 8 else temps[:0]
    test = np.nonzero(np.isin(T, held))[0]
    trainf = np.nonzero(~np.isin(T, held))[0]
    with SgTrainer(P, F, E, L) as t:
        stats = t.feature_stats()
        blob = init_params(stats, float(np.mean(E[trainf]) / P.shape[1]), seed=a.seed)
        log("training set on the device")
        blob = train(t, blob, trainf, epochs=a.epochs, batch=a.batch, lr_max=a.lr, seed=a.seed, log=log,
                     eval_frames=test if len(test) else None)
        report = {"frames": int(len(P)), "atoms_per_frame": int(P.shape[1]), "train_frames": int(len(trainf)),
                  "held_out_temperatures_K": [float(x) for x in held]}
        f_mse, e_mse = evaluate(t, trainf)
        report["train_force_rmse"] = float(np.sqrt(f_mse))
        report["train_energy_rmse_per_atom"] = float(np.sqrt(e_mse))
        for tv in held:
            f_mse, e_mse = evaluate(t, np.nonzero(T == tv)[0])
            report[f"held_out_force_rmse_{tv:.0f}K"] = float(np.sqrt(f_mse))
    out = Path(a.out)
    out.parent.mkdir(parents=True, exist_ok=True)
    np.save(out, blob)
    report["seconds"] = time.time() - t0
    log(json.dumps(report))
    if a.report:
        Path(a.report).write_text(json.dumps(report, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
