"""Parity of the HIP engine (through the C-ABI) with the FP64 CPU oracle.

Bars: candidate lists identical to the oracle's (bit-exact, the reference's
row format); FP64 kernels equal the oracle bit for bit where the arithmetic
is the same operations in the same order (LJ, Gear, walls) and within 1e-12
relative where device libm differs (bond order); mixed-precision forces
within 1e-5 relative RMS of the FP64 oracle.
"""
import numpy as np
import pytest

import _oracle as O
from moleculardynamics_amd import (BOND_REFERENCE, BOND_TERSOFF, BOX_PERIODIC, BOX_WALLED, LIST_BOND, LIST_LJ,
                                   LJ_REFERENCE, LJ_UNIFORM, PREC_FP64, PREC_MIXED, System, initial)

pytestmark = pytest.mark.gpu

N, OX, NE = 0, 1, 2
SIG_NE = float(np.float32(2.72))
DT = float(np.float32(0.005))   # the engine takes DT as a float, like the reference


def rel_rms(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def gas_mixture(n_mol, n_ne, seed, spacing=3.3):
    """Ne atoms and N2/O2/NO molecules on a jittered grid (reference units, A).
    The bond-order atoms come first, as the reference requires."""
    rng = np.random.default_rng(seed)
    sites = n_mol + n_ne
    k = int(np.ceil(sites ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(k)] * 3, indexing="ij"), -1).reshape(-1, 3)[:sites].astype(float)
    g = (g - (k - 1) / 2) * spacing + rng.uniform(-0.3, 0.3, g.shape)
    rng.shuffle(g)
    mol, ne = g[:n_mol], g[n_mol:]
    pos = []
    for c in mol:
        u = rng.normal(size=3); u /= np.linalg.norm(u)
        kinds = rng.choice([(N, N), (OX, OX), (N, OX)])
        pos.append([*(c + 0.57 * u), kinds[0]])
        pos.append([*(c - 0.57 * u), kinds[1]])
    for c in ne:
        pos.append([*c, NE])
    pos4 = np.array(pos, float)
    masses = np.array([O.param(0, int(t), 0) for t in pos4[:, 3]])
    vel = rng.normal(size=(len(pos4), 3)) * np.sqrt(8.314e-7 * 300 / masses)[:, None]
    return pos4, vel


def walled(pos4, n_bond, precision, use_lj=True, use_bond=True, lj=(7.6, 6.8), mb=(3.6, 3.0), dt=1.0,
           wall=None, eloss=1.0):
    n = len(pos4)
    return System(n_atoms=n, n_bond=n_bond, box=BOX_WALLED, precision=precision, use_lj=int(use_lj),
                  use_bond=int(use_bond), lj_table=LJ_REFERENCE, bond_form=BOND_REFERENCE, dt=dt,
                  lj_rs=lj[0], lj_rcut=lj[1], mb_rs=mb[0], mb_rcut=mb[1], energy_loss=eloss,
                  wall=wall if wall is not None else n ** (1 / 3) * 52.8)


@pytest.fixture(scope="module")
def mixture():
    return gas_mixture(30, 500, seed=1)


def test_lists_bit_exact_walled(mixture):
    pos4, vel = mixture
    nb = 60
    with walled(pos4, nb, PREC_FP64) as s:
        s.set_state(pos4, vel)
        vmax, box = s.ready_list()
        vo, bo = O.ready_list(pos4, vel)
        assert vmax == vo and box == bo
        nls, nlg = s.build_list(LIST_LJ, box)
        ref, ls = O.build_neighbor_list(pos4, 7.6, 6.8, box)
        assert (nls, nlg) == (ls.nlistsize, ls.nlargestsize)
        np.testing.assert_array_equal(s.neighbor_list(LIST_LJ, nls), ref)
        nls, nlg = s.build_list(LIST_BOND, box)
        ref, ls = O.build_neighbor_list(pos4[:nb], 3.6, 3.0, box)
        assert (nls, nlg) == (ls.nlistsize, ls.nlargestsize)
        np.testing.assert_array_equal(s.neighbor_list(LIST_BOND, nls, items=nb), ref)


def test_lists_bit_exact_walled_both_builds():
    """Walled grids of few atoms and cells are built by one workgroup in LDS
    (at most 6,144 atoms and 12,287 cells), larger ones by the multi-kernel
    build.  Two Ne atoms moved out to +-75 A widen the grids: the bond-order
    grid (cells of 3.6 A) past the one-workgroup limit, the Lennard-Jones grid
    (7.6 A) still under it.  Both lists and both sets of forces of the
    mixture must equal the oracle's."""
    pos4, vel = gas_mixture(30, 500, seed=1)
    pos4 = pos4.copy()
    pos4[-1, :3] = (75.0, 0.5, -0.5)
    pos4[-2, :3] = (-75.0, -0.5, 0.5)
    nb = 60
    _, box = O.ready_list(pos4, vel)
    assert (int(box / 3.6 + 0.99) + 1) ** 3 > 12287 >= (int(box / 7.6 + 0.99) + 1) ** 3
    with walled(pos4, nb, PREC_FP64) as s:
        s.set_state(pos4, vel)
        _, box_d = s.ready_list()
        assert box_d == box
        nls, nlg = s.build_list(LIST_LJ, box)
        ref, ls = O.build_neighbor_list(pos4, 7.6, 6.8, box)
        assert (nls, nlg) == (ls.nlistsize, ls.nlargestsize)
        np.testing.assert_array_equal(s.neighbor_list(LIST_LJ, nls), ref)
        nls, nlg = s.build_list(LIST_BOND, box)
        ref_mb, lsb = O.build_neighbor_list(pos4[:nb], 3.6, 3.0, box)
        assert (nls, nlg) == (lsb.nlistsize, lsb.nlargestsize)
        np.testing.assert_array_equal(s.neighbor_list(LIST_BOND, nls, items=nb), ref_mb)
        s.lj_forces()
        s.bond_forces()
        f = s.get_state()["force"]
    fo = O.lj_forces(pos4, ref, ls.nlargestsize, float(ls.rcutsq))
    np.testing.assert_array_equal(f[nb:], fo[nb:])
    fo = O.mb_forces(pos4, ref_mb, lsb.nlargestsize, float(lsb.rcutsq), f_in=fo, ljexist=True)
    assert rel_rms(f[:nb], fo[:nb]) < 1e-12


def test_list_edge_cell_walled():
    pos4 = np.array([[3.005, 0, 0, 2], [-3.005, 0, 0, 2], [2.2, 0.1, 0.1, 2], [0, 0, 0, 2], [1.0, -1, 0.5, 2]])
    with walled(pos4, 0, PREC_FP64, use_bond=False, lj=(2.0, 1.5)) as s:
        s.set_state(pos4, np.zeros((5, 3)))
        _, box = s.ready_list()
        nls, _ = s.build_list(LIST_LJ, box)
        ref, _ = O.build_neighbor_list(pos4, 2.0, 1.5, box)
        np.testing.assert_array_equal(s.neighbor_list(LIST_LJ, nls), ref)


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED])
def test_lj_forces_walled(mixture, precision):
    pos4, vel = mixture
    with walled(pos4, 60, precision, use_bond=False) as s:
        s.set_state(pos4, vel)
        _, box = s.ready_list()
        s.build_list(LIST_LJ, box)
        s.lj_forces()
        f = s.get_state()["force"]
    ref, ls = O.build_neighbor_list(pos4, 7.6, 6.8, box)
    fo = O.lj_forces(pos4, ref, ls.nlargestsize, float(ls.rcutsq))
    if precision == PREC_FP64:
        np.testing.assert_array_equal(f, fo)
    else:
        assert rel_rms(f, fo) < 1e-5


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED])
def test_bond_forces_walled(mixture, precision):
    pos4, vel = mixture
    nb = 60
    with walled(pos4, nb, precision) as s:
        s.set_state(pos4, vel)
        _, box = s.ready_list()
        s.build_list(LIST_LJ, box)
        s.build_list(LIST_BOND, box)
        s.lj_forces()
        s.bond_forces()
        f = s.get_state()["force"]
    lj, ls = O.build_neighbor_list(pos4, 7.6, 6.8, box)
    fo = O.lj_forces(pos4, lj, ls.nlargestsize, float(ls.rcutsq))
    mb, lsb = O.build_neighbor_list(pos4[:nb], 3.6, 3.0, box)
    fo = O.mb_forces(pos4, mb, lsb.nlargestsize, float(lsb.rcutsq), f_in=fo, ljexist=True)
    tol = 1e-12 if precision == PREC_FP64 else 1e-5
    assert rel_rms(f[:nb], fo[:nb]) < tol
    assert rel_rms(f, fo) < tol


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED])
def test_lj_forces_walled_dense_rows(precision):
    """Ne atoms 2.0 A apart: about 110 pairs within rc per atom (up to 171),
    so the wavefront-per-atom kernel fills its 64-entry buffer of in-cutoff
    pairs and evaluates it mid-row, for some atoms twice.  Forces against the oracle
    (FP64 bit for bit, mixed 1e-5) and, in FP64, the virial (the energy
    variant of the kernel) against sum_i r_i . F_i."""
    pos4, vel = gas_mixture(0, 1000, seed=3, spacing=2.0)
    with walled(pos4, 0, precision, use_bond=False) as s:
        s.set_state(pos4, vel)
        _, box = s.ready_list()
        s.build_list(LIST_LJ, box)
        s.lj_forces()
        f = s.get_state()["force"]
        w_lj = s.potential_virial()[2] if precision == PREC_FP64 else None
    ref, ls = O.build_neighbor_list(pos4, 7.6, 6.8, box)
    fo = O.lj_forces(pos4, ref, ls.nlargestsize, float(ls.rcutsq))
    d = pos4[:, None, :3] - pos4[None, :, :3]
    pairs = ((d ** 2).sum(-1) < 6.8 ** 2).sum(1) - 1
    assert pairs.max() > 128, pairs.max()   # at least two mid-row evaluations for some atoms
    if precision == PREC_FP64:
        np.testing.assert_array_equal(f, fo)
        assert w_lj == pytest.approx(float(np.sum(pos4[:, :3] * fo)), rel=1e-9)
    else:
        assert rel_rms(f, fo) < 1e-5


def test_lj_forces_walled_thread_per_atom():
    """33,000 Ne atoms: above the size where the walled LJ force runs one
    wavefront per atom, so this covers the thread-per-atom kernel (same bar:
    FP64 bit for bit)."""
    pos4, vel = gas_mixture(0, 33000, seed=8)
    with walled(pos4, 0, PREC_FP64, use_bond=False) as s:
        s.set_state(pos4, vel)
        _, box = s.ready_list()
        s.build_list(LIST_LJ, box)
        s.lj_forces()
        f = s.get_state()["force"]
    ref, ls = O.build_neighbor_list(pos4, 7.6, 6.8, box)
    fo = O.lj_forces(pos4, ref, ls.nlargestsize, float(ls.rcutsq))
    np.testing.assert_array_equal(f, fo)


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED])
def test_bond_forces_walled_long_rows(precision):
    """Molecules 2 A apart: the bond-order rows exceed the 256 entries one
    wavefront stages, so the thread-per-atom bond kernels run."""
    pos4, vel = gas_mixture(400, 0, seed=9, spacing=2.0)
    nb = len(pos4)
    with walled(pos4, nb, precision, use_lj=False) as s:
        s.set_state(pos4, vel)
        _, box = s.ready_list()
        nls, _ = s.build_list(LIST_BOND, box)
        assert nls > 256
        s.bond_forces()
        f = s.get_state()["force"]
    mb, lsb = O.build_neighbor_list(pos4, 3.6, 3.0, box)
    fo = O.mb_forces(pos4, mb, lsb.nlargestsize, float(lsb.rcutsq), f_in=np.zeros((nb, 3)), ljexist=False)
    assert rel_rms(f, fo) < (1e-12 if precision == PREC_FP64 else 1e-5)


def test_bond_only_assigns(mixture):
    pos4, vel = mixture
    nb = 60
    with walled(pos4[:nb], nb, PREC_FP64, use_lj=False) as s:
        s.set_state(pos4[:nb], vel[:nb])
        _, box = s.ready_list()
        s.build_list(LIST_BOND, box)
        s.bond_forces()
        f1 = s.get_state()["force"]
        s.bond_forces()
        f2 = s.get_state()["force"]
    np.testing.assert_array_equal(f1, f2)


def test_gear_and_walls_bit_exact():
    rng = np.random.default_rng(4)
    n = 700
    pos4 = np.c_[rng.uniform(-12, 12, (n, 3)), rng.integers(0, 3, n)]
    v, a, b, c, f = (rng.normal(size=(n, 3)) for _ in range(5))
    npos, eloss, dt = 10.0, np.float32(0.9), float(np.float32(0.5))
    with walled(pos4, 0, PREC_FP64, use_lj=True, use_bond=False, dt=dt, wall=npos, eloss=float(eloss)) as s:
        s.set_state(pos4, v, a, b, c)
        s.predict()
        st = s.get_state()
        ro = O.predict(pos4, v, a, b, c, dt)
        for got, want in zip((st["pos"], st["vel"], st["acc"], st["b"], st["c"]), ro):
            np.testing.assert_array_equal(got, want)
        # corrector with a chosen force: upload it through the state (f = m a_new)
        s.set_state(pos4, v, a, b, c)
    # corrector parity uses the oracle's correct on the same inputs; the device
    # force is produced by the LJ kernel, so compare on a force computed there
    with walled(pos4, 0, PREC_FP64, use_lj=True, use_bond=False, dt=dt, wall=npos, eloss=float(eloss),
                lj=(7.6, 6.8)) as s:
        s.set_state(pos4, v, a, b, c)
        _, box = s.ready_list()
        s.build_list(LIST_LJ, box)
        s.lj_forces()
        st0 = s.get_state()
        hit = s.correct()
        st = s.get_state()
    ro = O.correct(pos4, v, st0["force"], a, b, c, dt, npos, eloss)
    for got, want in zip((st["pos"], st["vel"], st["acc"], st["b"], st["c"]), ro[:5]):
        np.testing.assert_array_equal(got, want)
    assert hit == ro[5] == 1


def test_walled_run_matches_reference_loop(mixture):
    pos4, vel = mixture
    n, nb, steps = len(pos4), 60, 60
    ro, vo, rec, _ = O.ref_run(pos4, vel, n - nb, nb, steps, 1000, 1.0, 7.6, 6.8, 3.6, 3.0, use_lj=True, use_mb=True)
    with walled(pos4, nb, PREC_FP64) as s:
        s.set_state(pos4, vel)
        _, box = s.ready_list()
        s.build_list(LIST_LJ, box)
        s.build_list(LIST_BOND, box)
        s.run(steps)
        st = s.get_state()
    assert np.abs(st["pos"][:, :3] - ro[:, :3]).max() < 1e-9
    assert rel_rms(st["vel"], vo) < 1e-9


WALL_ELOSS = 0.81   # the EnergyLoss key; the engine takes its square root (cpu/cpumd/Host.cpp:73)


@pytest.mark.parametrize("use_mb", [False, True], ids=["lj", "lj+bond"])
def test_walled_run_wall_hits_match_reference_loop(use_mb):
    """150 steps of a hot mixture in walls at +-15.5 A: atoms reach every wall
    branch of cpu/common/Device.cpp:86-118 (x+ reflected with sqrt(EnergyLoss)
    when y > 0, x+ teleported to -nPos when y < 0, x-, +-y, +-z) on many steps,
    so the loop's recompute of the forces and a = F/m after a hit
    (cpu/cpumd/Host.cpp:166-178) runs, and lists are rebuilt after teleports.
    The oracle's run counts each branch; the engine must follow its trajectory
    bit for bit with Lennard-Jones alone and within 1e-9 with the bond order
    (device exp/tanh differ from glibc in the last bits), and must rebuild its
    list as often."""
    pos4, vel = gas_mixture(30, 500, seed=10)
    vel = vel * 12
    n, steps, wall = len(pos4), 150, 15.5
    nb = 60 if use_mb else 0
    ro, vo, _, _, st = O.ref_run(pos4, vel, n - nb, nb, steps, 1000, 1.0, 7.6, 6.8, 3.6, 3.0,
                                 energy_loss=WALL_ELOSS, use_lj=True, use_mb=use_mb, wall=wall, stats=True)
    branches = [st[k] for k in O.RUN_STATS[:7]]
    assert min(branches) >= 3, st
    assert st["hit steps"] >= 50 and st["builds after teleport"] >= 2, st
    eloss = float(np.sqrt(np.float32(WALL_ELOSS)))
    print(f"walled run with wall hits: {n} atoms, bond={use_mb}, {steps} steps, oracle counts {st}", flush=True)
    with walled(pos4, nb, PREC_FP64, use_bond=use_mb, wall=wall, eloss=eloss) as s:
        s.set_state(pos4, vel)
        s.run(steps)
        got = s.get_state()
        builds = s.build_count()
    assert builds == 1 + st["lj builds"]
    if use_mb:
        assert np.abs(got["pos"][:, :3] - ro[:, :3]).max() < 1e-9
        assert rel_rms(got["vel"], vo) < 1e-9
    else:
        np.testing.assert_array_equal(got["pos"], ro)
        np.testing.assert_array_equal(got["vel"], vo)


def test_walled_virials_match_oracle(mixture):
    """The walled box is isolated: its virial is sum_i r_i . F_i.  The LJ part
    (summed pair by pair on the device) and the bond-order part (from the
    bond-order atoms' bond forces) against that sum over the oracle's forces."""
    pos4, vel = mixture
    nb = 60
    with walled(pos4, nb, PREC_FP64) as s:
        s.set_state(pos4, vel)
        _, box = s.ready_list()
        s.build_list(LIST_LJ, box)
        s.build_list(LIST_BOND, box)
        e_lj, e_b, w_lj, w_b = s.potential_virial()
    ref, ls = O.build_neighbor_list(pos4, 7.6, 6.8, box)
    f_lj = O.lj_forces(pos4, ref, ls.nlargestsize, float(ls.rcutsq))
    mb, lsb = O.build_neighbor_list(pos4[:nb], 3.6, 3.0, box)
    f_b = O.mb_forces(pos4, mb, lsb.nlargestsize, float(lsb.rcutsq), f_in=np.zeros((len(pos4), 3)), ljexist=False)
    assert w_lj == pytest.approx(float(np.sum(pos4[:, :3] * f_lj)), rel=1e-9)
    assert w_b == pytest.approx(float(np.sum(pos4[:nb, :3] * f_b[:nb])), rel=1e-10)


def test_kinetics_and_allpairs(mixture):
    pos4, vel = mixture
    nb = 60
    with walled(pos4, nb, PREC_FP64) as s:
        s.set_state(pos4, vel)
        k = s.kinetics(float(np.float32(8.314e-7)))
        u = s.potential_allpairs()
    ko = O.kinetics(pos4, vel)
    np.testing.assert_allclose(k, ko, rtol=1e-12, atol=1e-15)
    uo = O.potential(pos4)
    np.testing.assert_allclose(u, uo, rtol=1e-11)


# ---------------- periodic reduced-unit LJ ----------------
def periodic(n, L, precision, rc=2.5, rs=2.8, dt=0.005):
    return System(n_atoms=n, box=BOX_PERIODIC, precision=precision, use_lj=1, lj_table=LJ_UNIFORM,
                  dt=dt, lj_rs=rs, lj_rcut=rc, box_length=L, lj_sigma=1.0, lj_eps=1.0, mass=1.0)


@pytest.fixture(scope="module")
def fluid():
    pos, vel, L = initial.lj_fluid(10, 0.8442, 0.728, seed=3)
    rng = np.random.default_rng(9)
    pos = np.mod(pos + rng.normal(0, 0.02, pos.shape), L)
    return pos, vel, L


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED])
def test_periodic_list_and_forces(fluid, precision):
    pos, vel, L = fluid
    n = len(pos)
    with periodic(n, L, precision) as s:
        s.set_state(np.c_[pos, np.zeros(n)], vel)
        nls, nlg = s.build_list(LIST_LJ)
        lst = s.neighbor_list(LIST_LJ, nls)
        s.lj_forces()
        f = s.get_state()["force"]
        e = s.potential()[0]
    ref, nlg_o = O.pbc_list(pos, L, 2.8)
    assert nlg == nlg_o and nls == ref.shape[1]
    np.testing.assert_array_equal(lst, ref)
    fo, eo, _ = O.pbc_lj_forces(pos, L, 2.5, ref)
    if precision == PREC_FP64:
        np.testing.assert_array_equal(f, fo)
        assert e == pytest.approx(eo, rel=1e-13)
    else:
        assert rel_rms(f, fo) < 1e-5
        assert e == pytest.approx(eo, rel=1e-6)


def test_periodic_fused_steps_match_oracle(fluid):
    """The fused step kernel (FP64) walks the list in the oracle's candidate
    order and performs the oracle's operations; the only difference is that
    periodic images are formed as x_j + shift in the staged neighbourhood
    rather than (x_i - x_j) - shift, so a few steps agree to rounding."""
    pos, vel, L = fluid
    n, steps = len(pos), 4
    ro, vo, _, builds, _ = O.pbc_lj_run(pos, vel, L, 2.5, 2.8, DT, steps)
    assert builds == 1
    with periodic(n, L, PREC_FP64) as s:
        s.set_state(np.c_[pos, np.zeros(n)], vel)
        s.run(steps)
        st = s.get_state()
        assert s.build_count() == 1
    np.testing.assert_allclose(st["pos"][:, :3], ro, rtol=0, atol=1e-12)
    np.testing.assert_allclose(st["vel"], vo, rtol=0, atol=1e-11)


def test_periodic_run_matches_oracle(fluid):
    pos, vel, L = fluid
    n, steps = len(pos), 200
    ro, vo, en, builds, _ = O.pbc_lj_run(pos, vel, L, 2.5, 2.8, DT, steps)
    with periodic(n, L, PREC_FP64) as s:
        s.set_state(np.c_[pos, np.zeros(n)], vel)
        s.run(steps)
        st = s.get_state()
        assert s.build_count() >= 2
    d = st["pos"][:, :3] - ro
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-8
    assert rel_rms(st["vel"], vo) < 1e-8


# The band above rs^2 an exported list may reach: the Lennard-Jones list build
# tests FP32 positions relative to the segment corner against rs^2 + 1e-4
# (sigma^2); the Tersoff build tests the FP64 minimum image rounded to FP32
# against rs^2 (1 + 1e-6).
LJ_LIST_SLACK2 = 2e-4
TS_LIST_SLACK2 = 1e-5 * 3.3 ** 2


def assert_pair_sets(s, L, rs, rc, slack2, pos=None):
    """The list the periodic force kernels read (caller indices, at its build
    positions) against the exact FP64 sets of the oracle's checker: it holds
    every pair within rs, nothing at rs^2 + slack2 or beyond, no repeats; so
    its pairs within rc are exactly the oracle's (common/NeighborList.cpp:88-145
    with the distance filter of cpu/common/LennardJones.h:51-56)."""
    rows, counts, bpos = s.pair_list()
    if pos is not None:   # built from these positions (wrapped into the box)
        d = bpos - pos
        d -= L * np.round(d / L)
        assert np.abs(d).max() < 1e-12
    st = O.pair_list_check(bpos, L, rs, rc, rows, counts, slack2)
    assert st["missing within rs"] == 0 and st["beyond rs + slack"] == 0 and st["bad entries"] == 0, st
    assert st["rc pairs listed"] == st["pairs within rc"] > 0, st
    assert st["entries"] == st["pairs within rs"] + st["in slack band"], st
    assert st["in slack band"] <= 1e-4 * st["entries"], st
    return st


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED])
def test_periodic_pair_lists_are_the_exact_sets(fluid, precision):
    """k_nlist_lds's lists (16-bit LDS indices, decoded through each segment's
    staging) as sets: at the first build and at the last one of a 200-step
    run with rebuilds."""
    pos, vel, L = fluid
    n = len(pos)
    with periodic(n, L, precision) as s:
        s.set_state(np.c_[pos, np.zeros(n)], vel)
        s.run(0)
        st = assert_pair_sets(s, L, 2.8, 2.5, LJ_LIST_SLACK2, pos)
        s.run(200)
        assert s.build_count() >= 3
        assert_pair_sets(s, L, 2.8, 2.5, LJ_LIST_SLACK2)
    print(f"pair list of {n} atoms: {st}", flush=True)


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED])
def test_periodic_virial_matches_oracle(fluid, precision):
    """mdx_potential_virial: the LJ virial (sum over pairs of r_ij . F_ij)
    against the oracle's, which is pinned by the scaling derivative of its
    energy; FP64 to summation order, mixed to 1e-6."""
    pos, vel, L = fluid
    n = len(pos)
    with periodic(n, L, precision) as s:
        s.set_state(np.c_[pos, np.zeros(n)], vel)
        s.build_list(LIST_LJ)
        e_lj, e_b, w_lj, w_b = s.potential_virial()
    ref, _ = O.pbc_list(pos, L, 2.8)
    _, eo, wo = O.pbc_lj_forces(pos, L, 2.5, ref)
    tol = 1e-12 if precision == PREC_FP64 else 1e-6
    assert w_lj == pytest.approx(wo, rel=tol)
    assert e_lj == pytest.approx(eo, rel=tol)
    assert e_b == 0 and w_b == 0


def test_periodic_virial_of_perfect_fcc():
    """Perfect FCC: the virial per atom is the closed-form lattice sum."""
    from test_oracle_pbc import fcc_lattice_virial
    pos, vel, L = initial.lj_fluid(10, 0.8442, 0.728, seed=1)
    n = len(pos)
    with periodic(n, L, PREC_FP64) as s:
        s.set_state(np.c_[pos, np.zeros(n)], vel)
        w = s.potential_virial()[2]
    assert w / n == pytest.approx(fcc_lattice_virial(L / 10, 2.5), rel=1e-12)


def fast_atom_lattice(v0=40.0):
    """A quiet FCC lattice (864 atoms) with one atom moving at v0 along x
    (0.2 sigma per step at dt = 0.005, past the list margin (rs - rc)/2 = 0.15
    sigma), which knocks lattice atoms out of place within single steps."""
    pos, _, L = initial.lj_fluid(6, 0.8442, 0.728, seed=1)
    vel = np.zeros_like(pos)
    vel[0, 0] = v0
    return pos, vel, L


def test_periodic_list_gate_is_exact():
    """The Lennard-Jones step kernels gate themselves: the last workgroup of
    each step's kernel finds the largest displacement of the next positions
    since the list build, and a step whose positions have moved (rs - rc)/2 or
    more runs as a no-op; the host rebuilds from those positions and launches
    it again.  That is the oracle's own rule (rebuild when any atom moved
    skin/2 or more), so with the same FP32 radii the runs make the same builds
    and follow the same trajectory, in one call and in calls of one step, and
    no step ever uses a stale list.  The fast atom needs a new list every step
    and the atoms it strikes are thrown past the margin within a step, which
    no decision made from earlier positions could foresee."""
    pos, vel, L = fast_atom_lattice()
    n, steps = len(pos), 30
    rs32, rc32 = float(np.float32(2.8)), float(np.float32(2.5))
    ro, vo, _, builds_o, _ = O.pbc_lj_run(pos, vel, L, rc32, rs32, DT, steps)
    assert builds_o >= steps // 2

    def system():
        s = periodic(n, L, PREC_FP64)
        s.set_state(np.c_[pos, np.zeros(n)], vel)
        return s

    for calls in ([steps], [1] * steps, [7, 1, 22]):
        with system() as s:
            for k in calls:
                s.run(k)
            st = s.get_state()
            assert s.build_count() == builds_o, calls
        d = st["pos"][:, :3] - ro
        d -= L * np.round(d / L)
        assert np.abs(d).max() < 1e-9, calls
        assert rel_rms(st["vel"], vo) < 1e-9, calls


def stale_step(err):
    import re
    m = re.search(r"step (\d+) used a stale list", str(err))
    assert m, str(err)
    return int(m.group(1))


def test_periodic_rs_equal_rc_rebuilds_every_step(fluid):
    """rs = rc leaves no margin: the reference's schedule then rebuilds before
    every step (cpu/cpumd/Host.cpp:288-290 with maxnlmove = 0), and so must the
    periodic loop, matching the oracle that rebuilds every step."""
    pos, vel, L = fluid
    n, steps = len(pos), 20
    ro, vo, _, builds, _ = O.pbc_lj_run(pos, vel, L, 2.5, 2.5, DT, steps, rebuild_every_step=True)
    with periodic(n, L, PREC_FP64, rc=2.5, rs=2.5) as s:
        s.set_state(np.c_[pos, np.zeros(n)], vel)
        s.run(steps)
        st = s.get_state()
        assert s.build_count() == steps + 1
    d = st["pos"][:, :3] - ro
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-10
    assert rel_rms(st["vel"], vo) < 1e-10


def test_periodic_split_runs_match_oracle(fluid):
    """The same 200 steps as above in calls of 1, 2, 3, 47 and 147 steps: each
    call starts with the predictor's flags (copied to the host) and hands the
    steps' flags over from the step kernel, so the rebuild decisions made
    across calls must keep the trajectory on the oracle's."""
    pos, vel, L = fluid
    n, chunks = len(pos), [1, 2, 3, 47, 147]
    ro, vo, _, _, _ = O.pbc_lj_run(pos, vel, L, 2.5, 2.8, DT, sum(chunks))
    with periodic(n, L, PREC_FP64) as s:
        s.set_state(np.c_[pos, np.zeros(n)], vel)
        for k in chunks:
            s.run(k)
        st = s.get_state()
        assert s.build_count() >= 2
    d = st["pos"][:, :3] - ro
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-8
    assert rel_rms(st["vel"], vo) < 1e-8


def drift_rate(steps, energies, n):
    """Least-squares slope of the total energy per atom per step."""
    return float(np.polyfit(np.asarray(steps, float), np.asarray(energies) / n, 1)[0])


def test_periodic_mixed_energy_drift(fluid):
    """1000 steps of the 4000-atom fluid: the total energy (shifted LJ; the
    force is cut at rc without a shift, so it is not exactly conserved) drifts
    in the mixed-precision run no faster than in the FP64 oracle's run, to 5%
    of that rate (slopes of the energy sampled every 50 steps)."""
    pos, vel, L = fluid
    n, steps, every = len(pos), 1000, 50
    _, _, en, _, _ = O.pbc_lj_run(pos, vel, L, 2.5, 2.8, DT, steps)
    ts = np.arange(0, steps + 1, every)
    rate_o = drift_rate(ts, en[ts, 2], n)
    e = []
    with periodic(n, L, PREC_MIXED) as s:
        s.set_state(np.c_[pos, np.zeros(n)], vel)
        for k in range(len(ts)):
            if k:
                s.run(every)
            e.append(s.potential()[0] + s.kinetics(1.0)[0])
    rate = drift_rate(ts, e, n)
    print(f"energy drift per atom-step over {steps} steps: oracle FP64 {rate_o:.3e}, mixed {rate:.3e}", flush=True)
    assert e[0] == pytest.approx(en[0, 2], rel=1e-6)
    assert abs(rate) <= 1.05 * abs(rate_o) + 1e-9


# ---------------- periodic Tersoff silicon ----------------
def tersoff_system(n, L, precision, dt=1.0, mb_rs=3.3):
    return System(n_atoms=n, n_bond=n, box=BOX_PERIODIC, precision=precision, use_lj=0, use_bond=1,
                  lj_table=LJ_UNIFORM, bond_form=BOND_TERSOFF, dt=dt, mb_rs=mb_rs, mb_rcut=3.0, box_length=L,
                  mass=initial.SI_MASS, accel_scale=initial.ACCEL_EV_A_AMU)


@pytest.fixture(scope="module")
def silicon():
    # 3^3 diamond cells (216 atoms, L = 16.3 A: 4 cells of 3.3 A per axis), 1000 K
    return initial.si_diamond(3, 1000.0, seed=5, disp=0.08)


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED])
def test_tersoff_forces_and_energy(silicon, precision):
    pos, vel, L = silicon
    n = len(pos)
    with tersoff_system(n, L, precision) as s:
        s.set_state(np.c_[pos, np.zeros(n)], vel)
        s.build_list(LIST_BOND)
        s.bond_forces()
        f = s.get_state()["force"]
        e = s.potential()[1]
    eo, fo = O.tersoff(pos, L)
    if precision == PREC_FP64:
        assert rel_rms(f, fo) < 1e-12
        assert e == pytest.approx(eo, rel=1e-13)
    else:
        assert rel_rms(f, fo) < 1e-5
        assert e == pytest.approx(eo, rel=1e-6)


def test_tersoff_run_matches_oracle(silicon):
    pos, vel, L = silicon
    n, steps = len(pos), 40
    ro, vo = O.tersoff_run(pos, vel, L, 1.0, steps, initial.SI_MASS, initial.ACCEL_EV_A_AMU)
    with tersoff_system(n, L, PREC_FP64) as s:
        s.set_state(np.c_[pos, np.zeros(n)], vel)
        s.run(steps)
        st = s.get_state()
        assert s.build_count() >= 1
    d = st["pos"][:, :3] - ro
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-9
    assert rel_rms(st["vel"], vo) < 1e-9


def test_tersoff_split_runs_match_oracle(silicon):
    """The same 40 steps in calls of 1, 2, 3, 6 and 28 steps: each call starts
    from the predictor's flags (copied to the host), each step's gather + Gear
    kernel hands the previous step's flags over, and the last step's are
    copied, so the rebuild decisions made across calls must keep the
    trajectory on the oracle's."""
    pos, vel, L = silicon
    n, chunks = len(pos), [1, 2, 3, 6, 28]
    ro, vo = O.tersoff_run(pos, vel, L, 1.0, sum(chunks), initial.SI_MASS, initial.ACCEL_EV_A_AMU)
    with tersoff_system(n, L, PREC_FP64) as s:
        s.set_state(np.c_[pos, np.zeros(n)], vel)
        for k in chunks:
            s.run(k)
        st = s.get_state()
        assert s.build_count() >= 1
    d = st["pos"][:, :3] - ro
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-9
    assert rel_rms(st["vel"], vo) < 1e-9


def test_tersoff_pair_lists_are_the_exact_sets(silicon):
    """k_ts_list's compact lists as sets within R + D = 3.0 A and the list
    radius 3.3 A, at a build and after 200 steps with rebuilds."""
    pos, vel, L = silicon
    n = len(pos)
    with tersoff_system(n, L, PREC_MIXED) as s:
        s.set_state(np.c_[pos, np.zeros(n)], vel)
        s.build_list(LIST_BOND)
        assert_pair_sets(s, L, 3.3, 3.0, TS_LIST_SLACK2, pos)
        s.run(200)
        assert s.build_count() >= 2
        assert_pair_sets(s, L, 3.3, 3.0, TS_LIST_SLACK2)


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED])
def test_tersoff_virial_is_the_scaling_derivative(silicon, precision):
    """The many-body virial -sum_i sum_j (x_j - x_i) . de_i/d(x_j - x_i) from the
    per-bond gradients equals -dE/dlambda of the oracle's Tersoff energy with
    positions and box scaled by lambda (central differences): FP64 to 1e-7;
    mixed (FP32 bond terms, whose float32-rounded parameters alone move the
    energy by ~1e-6; the virial is a sum with strong cancellation, measured
    1.3e-5 off) to 5e-5."""
    pos, vel, L = silicon
    n = len(pos)
    with tersoff_system(n, L, precision) as s:
        s.set_state(np.c_[pos, np.zeros(n)], vel)
        s.build_list(LIST_BOND)
        e_lj, e_b, w_lj, w_b = s.potential_virial()
    h = 1e-6
    fd = -(O.tersoff(pos * (1 + h), L * (1 + h), forces=False)[0] -
           O.tersoff(pos * (1 - h), L * (1 - h), forces=False)[0]) / (2 * h)
    assert w_b == pytest.approx(fd, rel=1e-7 if precision == PREC_FP64 else 5e-5)
    assert e_b == pytest.approx(O.tersoff(pos, L, forces=False)[0], rel=1e-12 if precision == PREC_FP64 else 2e-6)
    assert e_lj == 0 and w_lj == 0


def test_tersoff_mixed_energy_drift(silicon):
    """Mixed precision, 400 steps of 1 fs at 1000 K: total energy holds to
    1e-4 of the potential energy scale."""
    pos, vel, L = silicon
    n = len(pos)
    with tersoff_system(n, L, PREC_MIXED) as s:
        s.set_state(np.c_[pos, np.zeros(n)], vel)
        s.run(1)
        e = []
        for _ in range(4):
            st = s.get_state()
            ek = 0.5 * initial.SI_MASS * (st["vel"] ** 2).sum() / initial.ACCEL_EV_A_AMU
            e.append(s.potential()[1] + ek)
            s.run(100)
    assert max(e) - min(e) < 1e-4 * abs(e[0])


def squeezed_silicon(delta=1.4):
    """3^3 diamond cells at rest with atom 0 pushed delta A towards its nearest
    neighbour (a 2.35 A bond squeezed to 0.95 A): the pair flies apart from
    rest, speeding up steeply over the first femtoseconds."""
    pos, _, L = initial.si_diamond(3, 1000.0, seed=5, disp=0.0)
    d = pos - pos[0]
    d -= L * np.round(d / L)
    r = np.sqrt((d ** 2).sum(1))
    r[0] = np.inf
    j = int(np.argmin(r))
    pos[0] = np.mod(pos[0] + delta * d[j] / r[j], L)
    return pos, np.zeros_like(pos), L


def test_tersoff_stale_list_is_reported():
    """Tersoff silicon decides its rebuilds on the host from the flags of the
    previous step's positions (largest displacement D and single-step move S:
    rebuild at D + 1.2 S >= 0.9 of the margin); each force kernel checks its
    own positions against the full margin and counts the atoms beyond it.  The
    squeezed pair accelerates faster than that estimate allows for, and the
    step that uses the stale list must fail the run with MDX_ERR_LIST_STALE
    naming that step, whether the host finds it while later steps run or in
    the check at the end of the call (a call that ends with the stale step).
    A single step per call is always exact: each call's first decision reads
    the flags of the positions it steps from, so no step goes stale."""
    from moleculardynamics_amd._lib import MdxError
    pos, vel, L = squeezed_silicon()
    n = len(pos)

    def system():
        s = tersoff_system(n, L, PREC_FP64)
        s.set_state(np.c_[pos, np.zeros(n)], vel)
        return s

    with system() as s, pytest.raises(MdxError) as e:
        s.run(60)
    assert e.value.code == -7, str(e.value)
    k = stale_step(e.value)
    print(f"stale list at step {k}", flush=True)
    assert k >= 1
    with system() as s, pytest.raises(MdxError) as e:
        s.run(k + 1)
    assert e.value.code == -7 and stale_step(e.value) == k, str(e.value)
    with system() as s:
        for _ in range(k + 3):
            s.run(1)
