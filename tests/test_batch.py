"""Batched periodic Lennard-Jones replicas (mdx_batch_*): each replica must
behave as the same system run alone.  Boxes narrower than 3 cells of rs take
the all-pairs list path, the others the cell path; both are checked against
the oracle's forces and Gear trajectory of each replica."""
import numpy as np
import pytest

import _oracle as O
from moleculardynamics_amd import PREC_FP64, PREC_MIXED, Batch, initial

DT = float(np.float32(0.005))
RC, RS = 2.5, 2.8


def rel_rms(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def make_batch():
    """Five replicas: 256 atoms at rho 0.95 (L 6.46, all pairs), 500 at 0.80
    (L 8.55, 3 cells), 256 at 0.7 (L 7.15, all pairs), 864 at 0.9 (L 9.86,
    3 cells) and 6912 at 0.8 (L 20.5, 7 cells)."""
    rng = np.random.default_rng(8)
    spec = [(4, 0.95, 1.5), (5, 0.80, 0.728), (4, 0.70, 2.0), (6, 0.90, 1.0), (12, 0.80, 1.2)]
    ns, Ls, ps, vs = [], [], [], []
    for k, rho, T in spec:
        p, L = initial.fcc(k, rho)
        p = np.mod(p + rng.normal(0, 0.03, p.shape), L)
        ns.append(len(p)); Ls.append(L); ps.append(p)
        vs.append(initial.maxwell_boltzmann(len(p), T, int(rng.integers(1 << 30))))
    return np.array(ns, np.int32), np.array(Ls), ps, vs


def batch(ns, Ls, precision):
    return Batch(ns, Ls, precision=precision, dt=0.005, lj_rs=RS, lj_rcut=RC)


def full_list(n):
    return np.array([[j for j in range(n) if j != i] for i in range(n)], dtype=np.int32)


def oracle_list(p, L):
    if int(np.floor(L / RS)) >= 3:
        return O.pbc_list(p, L, RS)[0]
    return full_list(len(p))


def oracle_run(p, v, L, steps, with_force=False):
    """Gear PC with the oracle's forces over a list rebuilt from the current
    positions at every step (the exact pair set within rc); with_force also
    returns the last force (of the last step's predicted positions)."""
    r, vel = p.copy(), v.copy()
    a, b, c = np.zeros_like(r), np.zeros_like(r), np.zeros_like(r)
    f, _, _ = O.pbc_lj_forces(r, L, RC, oracle_list(np.mod(r, L), L))
    a[:] = f
    for _ in range(steps):
        r, vel, a, b, c = O.gear_predict(r, vel, a, b, c, DT)
        f, _, _ = O.pbc_lj_forces(r, L, RC, oracle_list(np.mod(r, L), L))
        r, vel, a, b, c = O.gear_correct(r, vel, f, a, b, c, DT)
    return (r, vel, f) if with_force else (r, vel)


@pytest.mark.gpu
def test_batch_paths():
    """Boxes of 3 or more cells of rs take the LDS-staged path (1), narrower
    ones the all-pairs path (2); the gather path (0) keeps what neither holds."""
    ns, Ls, _, _ = make_batch()
    with batch(ns, Ls, PREC_MIXED) as b:
        assert b.paths().tolist() == [2, 1, 2, 1, 1]
        assert b.staged().tolist() == [False, True, False, True, True]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED])
def test_batch_forces_match_oracle(precision):
    ns, Ls, ps, vs = make_batch()
    with batch(ns, Ls, precision) as b:
        b.set_state(np.concatenate(ps), np.concatenate(vs))
        b.run(0)
        f = b.get_state()["force"]
    off = 0
    for n, L, p in zip(ns, Ls, ps):
        fo, _, _ = O.pbc_lj_forces(p, L, RC, oracle_list(p, L))
        got = f[off: off + n]
        off += n
        assert rel_rms(got, fo) < (1e-13 if precision == PREC_FP64 else 1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED])
def test_batch_virials_match_oracle(precision):
    """mdx_batch_virials: per replica the sum over pairs of r_ij . F_ij (FP64
    terms, wave-shuffle sums) against the oracle's pair virial."""
    ns, Ls, ps, vs = make_batch()
    with batch(ns, Ls, precision) as b:
        b.set_state(np.concatenate(ps), np.concatenate(vs))
        b.run(0)
        w = b.virials()
        e = b.energies()
    for r, (n, L, p, v) in enumerate(zip(ns, Ls, ps, vs)):
        _, eo, wo = O.pbc_lj_forces(p, L, RC, oracle_list(p, L))
        assert w[r] == pytest.approx(wo, rel=1e-10)
        assert e[r, 1] == pytest.approx(eo, rel=1e-10)
        assert e[r, 0] == pytest.approx(0.5 * np.sum(v ** 2), rel=1e-12)


@pytest.mark.gpu
def test_batch_run_matches_oracle():
    ns, Ls, ps, vs = make_batch()
    steps = 30
    with batch(ns, Ls, PREC_FP64) as b:
        b.set_state(np.concatenate(ps), np.concatenate(vs))
        b.run(steps)
        st = b.get_state()
    off = 0
    for n, L, p, v in zip(ns, Ls, ps, vs):
        ro, vo = oracle_run(p, v, L, steps)
        d = st["pos"][off: off + n] - ro
        d -= L * np.round(d / L)
        assert np.abs(d).max() < 1e-10
        assert rel_rms(st["vel"][off: off + n], vo) < 1e-10
        off += n


@pytest.mark.gpu
def test_batch_mixed_steps_match_oracle():
    """The mixed-precision step kernels of every path (the staged replicas'
    k_step_lds, the small replicas' k_rep_small; a step and a final step)
    against the FP64 oracle: after 2 steps the force of each replica within
    1e-5 relative RMS of the oracle's at the same step and the positions within
    1e-9; after 30 steps the mixed trajectory within 1e-7 sigma of the FP64
    batch's."""
    ns, Ls, ps, vs = make_batch()
    pos, vel = np.concatenate(ps), np.concatenate(vs)
    with batch(ns, Ls, PREC_MIXED) as b:
        b.set_state(pos, vel)
        b.run(2)
        st2 = b.get_state()
        b.run(28)
        mixed30 = b.get_state()
    with batch(ns, Ls, PREC_FP64) as b:
        b.set_state(pos, vel)
        b.run(30)
        fp64_30 = b.get_state()
    off = 0
    for n, L, p, v in zip(ns, Ls, ps, vs):
        ro, vo, fo = oracle_run(p, v, L, 2, with_force=True)
        sl = slice(off, off + n)
        off += n
        assert rel_rms(st2["force"][sl], fo) < 1e-5
        d = st2["pos"][sl] - ro
        d -= L * np.round(d / L)
        assert np.abs(d).max() < 1e-9
        d = mixed30["pos"][sl] - fp64_30["pos"][sl]
        d -= L * np.round(d / L)
        assert np.abs(d).max() < 1e-7
        assert rel_rms(mixed30["vel"][sl], fp64_30["vel"][sl]) < 1e-6


@pytest.mark.gpu
def test_batch_graph_equals_eager_and_split_runs():
    ns, Ls, ps, vs = make_batch()
    pos, vel = np.concatenate(ps), np.concatenate(vs)
    out = []
    for timing, chunks in [(False, [41]), (True, [41]), (False, [20, 21])]:
        with batch(ns, Ls, PREC_MIXED) as b:
            b.set_state(pos, vel)
            b.set_timing(timing)
            for k in chunks:
                b.run(k)
            st = b.get_state()
            out.append((st["pos"], st["vel"], b.build_counts()))
    for o in out[1:]:
        np.testing.assert_array_equal(o[0], out[0][0])
        np.testing.assert_array_equal(o[1], out[0][1])
        np.testing.assert_array_equal(o[2], out[0][2])


@pytest.mark.gpu
def test_batch_energy_conservation_and_rebuilds():
    ns, Ls, ps, vs = make_batch()
    with batch(ns, Ls, PREC_MIXED) as b:
        b.set_state(np.concatenate(ps), np.concatenate(vs))
        b.run(1)
        e0 = b.energies().sum(1)
        b.run(600)
        e1 = b.energies().sum(1)
        builds = b.build_counts()
    assert np.all(builds >= 1)          # every replica rebuilt its list at least once
    assert np.all(np.abs(e1 - e0) < 2e-3 * np.abs(e0) + 0.05)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED])
def test_batch_gather_path_matches_oracle(precision):
    """A long cutoff (rc 4.5, rs 5.0) puts ~150 atoms in a cell of the 3-cell box,
    more than the LDS staging takes, so that box runs on the gather kernels
    (k_rep_list, k_rep_step); the 2-cell box takes the all-pairs kernel.  20
    steps of each against the oracle's Gear run with the exact pair set: FP64
    within 1e-10, mixed within 1e-7 (positions) and 1e-6 (velocities)."""
    rc, rs, steps = 4.5, 5.0, 20
    rng = np.random.default_rng(21)
    ns, Ls, ps, vs = [], [], [], []
    for k, T in [(10, 1.0), (6, 1.5)]:
        p, L = initial.fcc(k, 0.8)
        p = np.mod(p + rng.normal(0, 0.03, p.shape), L)
        ns.append(len(p)); Ls.append(L); ps.append(p)
        vs.append(initial.maxwell_boltzmann(len(p), T, int(rng.integers(1 << 30))))
    with Batch(np.array(ns, np.int32), np.array(Ls), precision=precision, dt=0.005, lj_rs=rs, lj_rcut=rc) as b:
        assert b.paths().tolist() == [0, 2]
        b.set_state(np.concatenate(ps), np.concatenate(vs))
        b.run(steps)
        st = b.get_state()
    off = 0
    for n, L, p, v in zip(ns, Ls, ps, vs):
        def lst(x):
            return O.pbc_list(np.mod(x, L), L, rs)[0] if int(np.floor(L / rs)) >= 3 else full_list(n)
        r, vel = p.copy(), v.copy()
        a, b3, c = np.zeros_like(r), np.zeros_like(r), np.zeros_like(r)
        a[:] = O.pbc_lj_forces(r, L, rc, lst(r))[0]
        for _ in range(steps):
            r, vel, a, b3, c = O.gear_predict(r, vel, a, b3, c, DT)
            f = O.pbc_lj_forces(r, L, rc, lst(r))[0]
            r, vel, a, b3, c = O.gear_correct(r, vel, f, a, b3, c, DT)
        d = st["pos"][off: off + n] - r
        d -= L * np.round(d / L)
        tol = (1e-10, 1e-10) if precision == PREC_FP64 else (1e-7, 1e-6)
        assert np.abs(d).max() < tol[0]
        assert rel_rms(st["vel"][off: off + n], vel) < tol[1]
        off += n
