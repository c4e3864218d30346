"""The vxcm sweep as one batch (mdx_sweep_*) against the single-system loop
and the oracle's reference loop (cpu/cpumd/Host.cpp:75-216).

Each member of a sweep runs the reference's loop body with its own list
schedule and wall flag, but every step kernel is launched once for all
members.  Bars: every member's state equals, bit for bit, what mdx_run gives
the same system on its own (the same kernels over the same data); and the
members follow the oracle's run of the reference loop for their vxcm -- bit
for bit with Lennard-Jones alone, within 1e-9 with the bond order (device
exp/tanh differ from glibc in the last bits) -- through steps where atoms hit
every wall branch and lists are rebuilt at different steps in each member.
"""
import ctypes as C

import numpy as np
import pytest

import _oracle as O
from moleculardynamics_amd import (BOND_REFERENCE, BOX_WALLED, LJ_REFERENCE, PREC_FP64, PREC_MIXED, MdxError,
                                   Sweep, System)
from test_gpu_parity import gas_mixture

pytestmark = pytest.mark.gpu

WALL, ELOSS = 15.5, 0.81
# centre-of-mass x velocities of the members, exact in float32 (the
# reference's vxcm is a float).  Larger offsets towards +x pile this hot gas
# against the +x wall until a bond-order pair collapses (|v| ~ 1e3 within 100
# steps, then NaN in the oracle itself); past such an event the last-bit
# differences of device exp/tanh grow without bound, so no 1e-9 bar holds.
# (Lennard-Jones alone is compared bit for bit, which holds through any event.)
VXCM = [0.0, 0.0009765625, -0.00146484375, -0.0029296875]


def config(n, n_bond, precision, use_bond, wall=WALL):
    return dict(n_atoms=n, n_bond=n_bond, box=BOX_WALLED, precision=precision, use_lj=1, use_bond=int(use_bond),
                lj_table=LJ_REFERENCE, bond_form=BOND_REFERENCE, dt=1.0, lj_rs=7.6, lj_rcut=6.8, mb_rs=3.6,
                mb_rcut=3.0, energy_loss=float(np.sqrt(np.float32(ELOSS))), wall=wall)


def hot_gas(seed=10):
    pos4, vel = gas_mixture(30, 500, seed=seed)
    return pos4, vel * 12


def member_velocities(vel, vx):
    v = vel.copy()
    v[:, 0] += vx
    return v


def state_of(s):
    st = s.get_state()
    return {k: st[k] for k in ("pos", "vel", "acc", "b", "c")}


@pytest.mark.parametrize("use_bond", [False, True], ids=["lj", "lj+bond"])
@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED], ids=["fp64", "mixed"])
def test_sweep_members_equal_single_systems(use_bond, precision):
    """Four members with different vxcm, run in calls of 40, 1 and 69 steps,
    against four independent systems run with mdx_run: the same bits in every
    Gear array and the same number of list builds."""
    pos4, vel = hot_gas()
    n, nb = len(pos4), (60 if use_bond else 0)
    cfg = config(n, nb, precision, use_bond)
    print(f"sweep vs single systems: {len(VXCM)} members of {n} atoms, bond={use_bond}, precision={precision}",
          flush=True)
    with Sweep(len(VXCM), **cfg) as w:
        for m, vx in enumerate(VXCM):
            w.member(m).set_state(pos4, member_velocities(vel, vx))
        for k in (40, 1, 69):
            w.run(k)
        got = [state_of(w.member(m)) for m in range(len(VXCM))]
        builds = [w.member(m).build_count() for m in range(len(VXCM))]
    for m, vx in enumerate(VXCM):
        with System(**cfg) as s:
            s.set_state(pos4, member_velocities(vel, vx))
            s.run(110)
            ref = state_of(s)
            assert s.build_count() == builds[m], m
        for k in ref:
            np.testing.assert_array_equal(got[m][k], ref[k], err_msg=f"member {m} {k}")


@pytest.mark.parametrize("use_bond", [False, True], ids=["lj", "lj+bond"])
def test_sweep_matches_reference_loop(use_bond):
    """150 steps of the hot mixture in walls at +-15.5 A for every member: the
    oracle's run of each member counts hits on every wall branch and list
    rebuilds after teleports; each member must follow it (FP64)."""
    pos4, vel = hot_gas()
    n, steps = len(pos4), 150
    nb = 60 if use_bond else 0
    cfg = config(n, nb, PREC_FP64, use_bond)
    print(f"sweep vs oracle: {len(VXCM)} members of {n} atoms, bond={use_bond}, {steps} steps", flush=True)
    with Sweep(len(VXCM), **cfg) as w:
        for m, vx in enumerate(VXCM):
            w.member(m).set_state(pos4, member_velocities(vel, vx))
        w.run(steps)
        got = [w.member(m).get_state() for m in range(len(VXCM))]
        builds = [w.member(m).build_count() for m in range(len(VXCM))]
    for m, vx in enumerate(VXCM):
        ro, vo, _, _, st = O.ref_run(pos4, vel, n - nb, nb, steps, 1000, 1.0, 7.6, 6.8, 3.6, 3.0,
                                     energy_loss=ELOSS, vxcm=vx, use_lj=True, use_mb=use_bond, wall=WALL,
                                     stats=True)
        assert min(st[k] for k in O.RUN_STATS[:7]) >= 1, (m, st)
        assert st["hit steps"] >= 50, (m, st)
        assert builds[m] == 1 + st["lj builds"], m
        if use_bond:
            assert np.abs(vo).max() < 1.0, m   # the oracle's gas stayed bounded (see VXCM)
            assert np.abs(got[m]["pos"][:, :3] - ro[:, :3]).max() < 1e-9, m
            assert np.sqrt(((got[m]["vel"] - vo) ** 2).sum() / (vo ** 2).sum()) < 1e-9, m
        else:
            np.testing.assert_array_equal(got[m]["pos"], ro, err_msg=f"member {m}")
            np.testing.assert_array_equal(got[m]["vel"], vo, err_msg=f"member {m}")


def test_sweep_member_with_long_rows_uses_its_own_kernels():
    """A member whose bond-order rows hold more than 64 candidates (its 100
    molecules packed in a block with centres 2.6 A apart, the Ne gas moved
    aside), so that mdx_run takes the thread-per-atom bond-order kernels for
    it, beside members whose rows fit the wavefront-per-atom kernels: each
    member runs on the kernels mdx_run would pick for it, so every member
    still equals its single-system run bit for bit."""
    pos4, vel = gas_mixture(100, 400, seed=3)
    n, nb = len(pos4), 200
    cfg = config(n, nb, PREC_FP64, True, wall=40.0)
    dense = pos4.copy()
    dense[:nb] = gas_mixture(100, 0, seed=5, spacing=2.6)[0]
    dense[nb:, 0] += 20.0
    starts = [pos4, dense, pos4]
    # the kernels each member takes: the xi table holds the longest bond-order
    # row, and rows of more than 64 leave the wavefront-per-atom kernels
    for p, long_rows in zip(starts, (False, True, False)):
        _, box = O.ready_list(p, vel)
        _, ls = O.build_neighbor_list(p[:nb], 3.6, 3.0, box)
        assert (ls.nlargestsize > 64) == long_rows
    print(f"sweep with a long-row member: 3 members of {n} atoms ({nb} bond order), 20 steps", flush=True)
    with Sweep(3, **cfg) as w:
        for m, p in enumerate(starts):
            w.member(m).set_state(p, vel)
        w.run(20)
        got = [state_of(w.member(m)) for m in range(3)]
    assert all(np.abs(g["vel"]).max() < 1.0 for g in got)
    for m, p in enumerate(starts):
        with System(**cfg) as s:
            s.set_state(p, vel)
            s.run(20)
            ref = state_of(s)
        for k in ref:
            np.testing.assert_array_equal(got[m][k], ref[k], err_msg=f"member {m} {k}")


def test_sweep_member_belongs_to_the_sweep():
    pos4, vel = hot_gas()
    cfg = config(len(pos4), 0, PREC_FP64, False)
    with Sweep(2, **cfg) as w:
        m0 = w.member(0)
        rc = m0.lib.mdx_destroy(m0.h)
        assert rc == -6   # MDX_ERR_STATE
        assert b"sweep" in m0.lib.mdx_last_error(m0.h)
        h = C.c_void_p()
        assert w.lib.mdx_sweep_member(w.h, 2, C.byref(h)) == -1
        with pytest.raises(MdxError):
            w.run(-1)
