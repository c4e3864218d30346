"""The learned Tersoff surrogate on the GPU against the oracle
(oracle/surrogate.py): descriptors and the training gradient of the FP32
training kernels, the BF16 MFMA inference forces and energies (the oracle
emulates the kernel's BF16 rounding points), and the per-atom plain-Tersoff
fallback (against the oracle's per-centre Tersoff)."""
import numpy as np
import pytest

import _oracle as O
from moleculardynamics_amd import (BOND_SURROGATE, BOND_TERSOFF, BOX_PERIODIC, LIST_BOND, LJ_UNIFORM, PREC_FP64,
                                   PREC_MIXED, SgTrainer, System, initial)
from oracle import surrogate as S

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def frames(n_frames=4, ncell=2, seed=20):
    ps, fs, es = [], [], []
    for k in range(n_frames):
        pos, _, L = initial.si_diamond(ncell, 300.0, seed + k, disp=0.06 + 0.02 * k)
        e, f = O.tersoff(pos, L)
        ps.append(pos)
        fs.append(f)
        es.append(e)
    return np.array(ps), np.array(fs), np.array(es), L


def random_params(G, seed=0):
    rng = np.random.default_rng(seed)
    p = {
        "W1": rng.normal(0, 0.3, (S.NHID, S.NFEAT)), "b1": rng.normal(0, 0.1, S.NHID),
        "W2": rng.normal(0, 0.15, (S.NHID, S.NHID)), "b2": rng.normal(0, 0.1, S.NHID),
        "w3": rng.normal(0, 0.3, S.NHID), "b3": np.array([-4.6]),
        "mean": G.mean(0), "istd": 1.0 / (G.std(0) + 1e-2),
        "lo": np.full(S.NFEAT, -np.inf), "hi": np.full(S.NFEAT, np.inf),
    }
    # the blob is float32: the oracle works with the same rounded values
    return S.unpack(S.pack(p).astype(np.float32).astype(np.float64))


def surrogate_system(n, L):
    return System(n_atoms=n, n_bond=n, box=BOX_PERIODIC, precision=PREC_MIXED, use_lj=0, use_bond=1,
                  lj_table=LJ_UNIFORM, bond_form=BOND_SURROGATE, dt=1.0, mb_rs=3.3, mb_rcut=3.0, box_length=L,
                  mass=initial.SI_MASS, accel_scale=initial.ACCEL_EV_A_AMU)


def sg_forces(pos, L, blob):
    """Forces and energy of the surrogate system (two evaluations: forces,
    then the energy) and its fallback counts."""
    n = len(pos)
    with surrogate_system(n, L) as s:
        s.set_surrogate(blob.astype(np.float32))
        s.set_state(np.c_[pos, np.zeros(n)], np.zeros((n, 3)))
        s.build_list(LIST_BOND)
        s.bond_forces()
        f = s.get_state()["force"]
        e = s.potential()[1]
        counts = s.surrogate_counts()
    return f, e, counts


@pytest.fixture(scope="module")
def silicon216():
    return initial.si_diamond(3, 1000.0, seed=31, disp=0.09)


def test_trainer_descriptors_match_oracle():
    P, F, E, L = frames()
    with SgTrainer(P, F, E, L) as t:
        for k in range(len(P)):
            G = t.descriptors(k).astype(np.float64)
            Go = S.frame_descriptors(P[k], L)[0]
            scale = np.abs(Go).max(0) + 1e-3
            # FP32 descriptors; the radial Gaussians come from a recurrence whose
            # float32 emulation stays within 2.4e-5 relative on every term above 1e-8
            assert np.abs((G - Go) / scale).max() < 5e-5, k
        st = t.feature_stats()
    Gall = np.concatenate([S.frame_descriptors(P[k], L)[0] for k in range(len(P))])
    np.testing.assert_allclose(st["mean"], Gall.mean(0), rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(st["std"], Gall.std(0), rtol=1e-3, atol=1e-5)


def test_trainer_gradient_matches_oracle():
    P, F, E, L = frames()
    G = np.concatenate([S.frame_descriptors(P[k], L)[0] for k in range(len(P))])
    p = random_params(G, seed=3)
    blob = S.pack(p).astype(np.float32)
    sel = [2, 0, 3]
    fr = []
    for k in sel:
        Gk, Jk, idx, _ = S.frame_descriptors(P[k], L, slots=8)
        fr.append((Gk, Jk, idx, F[k], E[k]))
    loss, g = S.train_loss_gradient(p, fr, w_energy=0.5)
    with SgTrainer(P, F, E, L) as t:
        t.set_params(blob)
        out, gg = t.gradient(sel, w_energy=0.5)
    assert out[0] == pytest.approx(loss, rel=1e-4)
    flat = np.concatenate([g[k].reshape(-1) for k in ("W1", "b1", "W2", "b2", "w3", "b3")])
    assert np.abs(gg - flat).max() < 1e-4 * np.abs(flat).max()
    assert rel_rms(gg, flat) < 1e-5


def test_trainer_steps_reduce_loss():
    P, F, E, L = frames(6)
    G = np.concatenate([S.frame_descriptors(P[k], L)[0] for k in range(len(P))])
    p = random_params(G, seed=4)
    with SgTrainer(P, F, E, L) as t:
        t.set_params(S.pack(p).astype(np.float32))
        first = t.gradient(range(6), want_grad=False)[0][0]
        for _ in range(60):
            t.step(range(6), lr=3e-3)
        last = t.gradient(range(6), want_grad=False)[0][0]
    assert last < 0.3 * first


def test_surrogate_forces_match_bf16_oracle(silicon216):
    pos, _, L = silicon216
    G = S.frame_descriptors(pos, L)[0]
    p = random_params(G, seed=5)
    f, e, counts = sg_forces(pos, L, S.pack(p))
    Eo, Fo, _, fl = S.energy_forces(p, pos, L, emulate_bf16=True)
    E64, F64, _, _ = S.energy_forces(p, pos, L)
    assert not fl.any() and counts == (0, 2 * len(pos))
    assert rel_rms(f, Fo) < 2e-3
    assert e == pytest.approx(Eo, rel=1e-5, abs=1e-3)
    # BF16 inference stays near the FP64 network, and the emulation is closer
    assert rel_rms(f, F64) < 0.05
    assert rel_rms(f, Fo) < rel_rms(f, F64)
    np.testing.assert_allclose(f.sum(0), 0, atol=1e-6 * np.abs(f).max() * len(f))


def test_surrogate_full_fallback_is_plain_tersoff(silicon216):
    pos, _, L = silicon216
    G = S.frame_descriptors(pos, L)[0]
    p = random_params(G, seed=6)
    p["hi"] = np.full(S.NFEAT, -np.inf)      # every atom out of range
    f, e, counts = sg_forces(pos, L, S.pack(p))
    eo, fo = O.tersoff(pos, L)
    assert counts == (2 * len(pos), 2 * len(pos))
    assert rel_rms(f, fo) < 1e-5
    # FP32 Tersoff terms: the float32-rounded parameters alone shift this
    # configuration's energy by +1.0e-6 relative (a float32 emulation of the
    # pair terms on the CPU gives the same +9.4e-4 eV)
    assert e == pytest.approx(eo, rel=2e-6)


def test_surrogate_partial_fallback_matches_oracle(silicon216):
    pos, _, L = silicon216
    G = S.frame_descriptors(pos, L)[0]
    p = random_params(G, seed=7)
    gs = np.sort(S.standardize(p, G)[:, 1])
    k = int(0.7 * len(gs))
    gap = np.argmax(np.diff(gs[k - 20: k + 20])) + k - 20     # a threshold well away from every value
    p["hi"][1] = 0.5 * (gs[gap] + gs[gap + 1])
    assert gs[gap + 1] - gs[gap] > 1e-4

    def fallback(mask):
        _, ft, e_atom = O.tersoff_centres(pos, L, mask.astype(np.uint8))
        return e_atom, ft

    Eo, Fo, _, fl = S.energy_forces(p, pos, L, emulate_bf16=True, tersoff_fallback=fallback)
    f, e, counts = sg_forces(pos, L, S.pack(p))
    assert 0 < fl.sum() < len(pos)
    assert counts == (2 * int(fl.sum()), 2 * len(pos))
    assert rel_rms(f, Fo) < 2e-3
    assert e == pytest.approx(Eo, rel=1e-5, abs=1e-3)


def test_shipped_network_against_plain_tersoff():
    """The shipped network (trained on GPU plain-Tersoff frames) against plain
    Tersoff on a 4096-atom box at 1000 K it never saw: force RMSE within
    0.05 eV/A and more than 20x below the same network untrained (random
    weights, same standardisation), no fallbacks, energy conserved over 400
    steps.  The step times of both are printed, not asserted: the per-atom
    plain Tersoff runs as fast as the surrogate on this hardware."""
    import time
    from moleculardynamics_amd import surrogate as SG

    blob = SG.load_weights()
    rng = np.random.default_rng(3)
    untrained = blob.copy()
    untrained[: 64 * 48] = rng.normal(0, 1 / np.sqrt(48), 64 * 48)
    untrained[64 * 48 + 64: 64 * 48 + 64 + 64 * 64] = rng.normal(0, 1 / 8, 64 * 64)
    untrained[64 * 48 + 64 + 64 * 64 + 64: 64 * 48 + 64 + 64 * 64 + 128] = rng.normal(0, 0.1 / 8, 64)
    pos, vel, L = initial.si_diamond(8, 1000.0, seed=77, disp=0.05)
    n = len(pos)

    def energy(s):   # eV: kinetic from amu A^2 / fs^2
        ek = 0.5 * initial.SI_MASS * (s.get_state()["vel"] ** 2).sum() / initial.ACCEL_EV_A_AMU
        return ek + s.potential()[1]

    def run(form, steps):
        with SG.silicon_system(n, L, bond_form=form) as s:
            if form == BOND_SURROGATE:
                s.set_surrogate(blob)
            s.set_state(np.c_[pos, np.zeros(n)], vel)
            s.run(20)
            s.synchronize()
            e0 = energy(s)
            t0 = time.perf_counter()
            s.run(steps)
            s.synchronize()
            dt = time.perf_counter() - t0
            e1 = energy(s)
            st = s.get_state()
            counts = s.surrogate_counts() if form == BOND_SURROGATE else (0, 1)
        return dt, e0, e1, np.mod(st["pos"][:, :3], L), counts

    def forces(form, params, at):
        with SG.silicon_system(n, L, bond_form=form) as s:
            if params is not None:
                s.set_surrogate(params)
            s.set_state(np.c_[at, np.zeros(n)], np.zeros((n, 3)))
            s.build_list(LIST_BOND)
            s.bond_forces()
            return s.get_state()["force"]

    t_plain, _, _, final, _ = run(BOND_TERSOFF, 400)
    t_sg, e0, e1, _, counts = run(BOND_SURROGATE, 400)
    f_plain = forces(BOND_TERSOFF, None, final)
    rmse = np.sqrt(np.mean((forces(BOND_SURROGATE, blob, final) - f_plain) ** 2))
    rmse_untrained = np.sqrt(np.mean((forces(BOND_SURROGATE, untrained, final) - f_plain) ** 2))
    print(f"4096 atoms, 400 steps: surrogate {t_sg:.4f} s, plain Tersoff {t_plain:.4f} s; force RMSE "
          f"{rmse:.4f} eV/A trained, {rmse_untrained:.3f} untrained")
    assert rmse < 0.05
    assert rmse_untrained > 20 * rmse
    assert counts[0] == 0
    assert abs(e1 - e0) < 1e-4 * abs(e0)        # NVE drift over 400 fs, as plain Tersoff's test


@pytest.mark.gpu
def test_shipped_network_at_1850K():
    """The hot end of the training range on a trajectory the network never
    saw: plain-Tersoff FP64 MD of a 512-atom box at 1850 K from a seed outside
    the training set's (1000-1047), 40 frames 20 fs apart.  The network does
    not reach the 0.05 eV/A of the 1000 K configuration here: its error grows
    with temperature (held out in training: 0.014 eV/A at 517 K, 0.026 at
    1132 K, 0.044 at 1819 K) and along this trajectory (0.071 over its first
    10 frames, 0.096 over all 40), about 4% of the 2.3 eV/A force RMS.
    Training five times longer (on the earlier 24-temperature set) left the
    hot end's held-out error unchanged, which points at the 48-64-64-1
    network and its descriptor rather than the training.  The bound here, 5%
    of the force RMS, guards that level."""
    from moleculardynamics_amd import surrogate as SG

    blob = SG.load_weights()
    P, F, _, _, L = SG.tersoff_frames(n_traj=1, frames_per_traj=40, t_min=1850.0, t_max=1850.0, seed=4242)
    n = P.shape[1]
    err2 = 0.0
    with SG.silicon_system(n, L, bond_form=BOND_SURROGATE) as s:
        s.set_surrogate(blob)
        for p, f in zip(P, F):
            s.set_state(np.c_[p, np.zeros(n)], np.zeros((n, 3)))
            s.build_list(LIST_BOND)
            s.bond_forces()
            err2 += float(np.sum((s.get_state()["force"] - f) ** 2))
        fallbacks = int(s.surrogate_counts()[0])
    rmse = np.sqrt(err2 / (3 * n * len(P)))
    print(f"1850 K, {len(P)} frames of {n} atoms: force RMSE {rmse:.4f} eV/A, force RMS "
          f"{np.sqrt(np.mean(F ** 2)):.3f}, fallback atom-evaluations {fallbacks}", flush=True)
    assert rmse < 0.05 * np.sqrt(np.mean(F ** 2))
