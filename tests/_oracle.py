"""ctypes access to the FP64 CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: tests compare the HIP engine against it; nothing
in the product imports this module.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE = ROOT / "oracle" / "liboracle.so"


class ListSettings(C.Structure):
    _fields_ = [("maxnlmove", C.c_float), ("maxnlmovesq", C.c_float), ("rcut", C.c_float),
                ("rs", C.c_float), ("rcutsq", C.c_float), ("nlistsize", C.c_int32),
                ("nlargestsize", C.c_int32)]


class RunCfg(C.Structure):
    _fields_ = [("n_lj", C.c_int32), ("n_mb", C.c_int32), ("timesteps", C.c_int32),
                ("output_timesteps", C.c_int32), ("dt", C.c_float), ("lj_rs", C.c_float),
                ("lj_rcut", C.c_float), ("mb_rs", C.c_float), ("mb_rcut", C.c_float),
                ("energy_loss", C.c_float), ("vxcm", C.c_float), ("use_lj", C.c_int32),
                ("use_mb", C.c_int32), ("wall", C.c_double), ("stats", C.c_void_p)]


class TersoffParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in "A B lambda1 lambda2 lambda3 beta n c d h R D".split()] + [("m", C.c_int32)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE.exists():
            subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
        L = C.CDLL(str(ORACLE))
        P, I, D, F = C.c_void_p, C.c_int32, C.c_double, C.c_float
        L.orc_ref_build_neighbor_list.restype = C.POINTER(C.c_int32)
        L.orc_ref_build_neighbor_list.argtypes = [P, C.POINTER(ListSettings), D, I, I]
        L.orc_free.argtypes = [P]
        L.orc_ref_ready_list.argtypes = [P, P, I, C.POINTER(D), C.POINTER(D)]
        L.orc_ref_next_build.argtypes = [F, D, D, I]
        L.orc_ref_next_build.restype = I
        L.orc_ref_predict.argtypes = [P, P, P, P, P, D, I]
        L.orc_ref_correct.argtypes = [P, P, P, P, P, P, D, I, D, C.POINTER(I), F]
        L.orc_ref_accelerations.argtypes = [P, P, P, I]
        L.orc_ref_lj_forces.argtypes = [P, P, I, I, I, P, D]
        L.orc_ref_mb_forces.argtypes = [P, P, I, I, I, P, D, I]
        L.orc_ref_kinetics.argtypes = [P, P, I, P]
        L.orc_ref_potential.argtypes = [P, I, P]
        for nm in ("orc_ref_attractive", "orc_ref_repulsive"):
            getattr(L, nm).argtypes = [D, D, D, D, C.POINTER(D), C.POINTER(D)]
        L.orc_ref_vdw.argtypes = [D, D, C.POINTER(D), C.POINTER(D)]
        L.orc_ref_cutoff.argtypes = [D, C.POINTER(D), C.POINTER(D)]
        L.orc_ref_param.argtypes = [I, I, I]
        L.orc_ref_param.restype = D
        L.orc_ref_run.argtypes = [C.POINTER(RunCfg), P, P, P, I, C.POINTER(D)]
        L.orc_pbc_build_list.restype = C.POINTER(C.c_int32)
        L.orc_pbc_build_list.argtypes = [P, I, D, D, C.POINTER(I), C.POINTER(I)]
        L.orc_pbc_pair_list_check.argtypes = [P, I, D, D, D, D, P, P, I, P]
        L.orc_pbc_pair_list_check.restype = I
        L.orc_pbc_lj_forces.restype = D
        L.orc_pbc_lj_forces.argtypes = [P, P, I, D, D, D, D, I, P, C.POINTER(D)]
        L.orc_gear_predict.argtypes = [P, P, P, P, P, D, I]
        L.orc_gear_correct.argtypes = [P, P, P, P, P, P, D, I, D, D]
        L.orc_pbc_lj_run.argtypes = [P, P, I, D, D, D, D, I, I, P, C.POINTER(D)]
        L.orc_pbc_lj_run.restype = I
        L.orc_pbc_lj_run_threads.argtypes = [P, P, I, D, D, D, D, I, I, P, C.POINTER(D), I]
        L.orc_pbc_lj_run_threads.restype = I
        L.orc_build_info.argtypes = []
        L.orc_build_info.restype = C.c_char_p
        L.orc_tersoff_si_params.argtypes = [C.POINTER(TersoffParams)]
        L.orc_tersoff_energy_forces.restype = D
        L.orc_tersoff_energy_forces.argtypes = [C.POINTER(TersoffParams), P, I, D, P]
        L.orc_tersoff_centres.restype = D
        L.orc_tersoff_centres.argtypes = [C.POINTER(TersoffParams), P, I, D, P, P, P]
        _lib = L
    return _lib


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        assert a.shape == shape, (a.shape, shape)
    return a


# ---------------- reference mode ----------------
def list_settings(rs: float, rc: float) -> ListSettings:
    """cpu/cpumd/Host.cpp:55-71 (the fields are floats)."""
    ls = ListSettings()
    ls.maxnlmove = np.float32((np.float32(rs) - np.float32(rc)) * np.float32(0.5))
    ls.maxnlmovesq = np.float32(float(ls.maxnlmove) ** 2)
    ls.rcutsq = np.float32(float(np.float32(rc)) ** 2)
    ls.rcut = rc
    ls.rs = rs
    return ls


def build_neighbor_list(pos4, rs, rc, box_size):
    pos4 = f64(pos4)
    n = len(pos4)
    ls = list_settings(rs, rc)
    p = lib().orc_ref_build_neighbor_list(ptr(pos4), C.byref(ls), box_size, 3 * n, n)
    nls = ls.nlistsize
    arr = np.ctypeslib.as_array(p, shape=(max(n * nls, 1),))[: n * nls].copy().reshape(n, nls)
    lib().orc_free(C.cast(p, C.c_void_p))
    return arr, ls


def ready_list(pos4, vel3):
    vm, bs = C.c_double(), C.c_double()
    lib().orc_ref_ready_list(ptr(f64(pos4)), ptr(f64(vel3)), len(pos4), C.byref(vm), C.byref(bs))
    return vm.value, bs.value


def next_build(maxnlmove, vmax, dt, t):
    return lib().orc_ref_next_build(maxnlmove, vmax, dt, t)


def lj_forces(pos4, nlist, nlargest, rcutsq):
    pos4 = f64(pos4)
    n = len(pos4)
    f = np.zeros((n, 3))
    nlist = np.ascontiguousarray(nlist, dtype=np.int32)
    lib().orc_ref_lj_forces(ptr(pos4), ptr(f), n, nlargest, nlist.shape[1], ptr(nlist), rcutsq)
    return f


def mb_forces(pos4, nlist, nlargest, rcutsq, f_in=None, ljexist=False):
    pos4 = f64(pos4)
    n = nlist.shape[0]
    f = np.zeros((len(pos4), 3)) if f_in is None else f64(f_in).copy()
    nlist = np.ascontiguousarray(nlist, dtype=np.int32)
    lib().orc_ref_mb_forces(ptr(pos4), ptr(f), n, nlargest, nlist.shape[1], ptr(nlist), rcutsq, 1 if ljexist else 0)
    return f


def predict(pos4, v, a, b, c, dt):
    arrs = [f64(x).copy() for x in (pos4, v, a, b, c)]
    lib().orc_ref_predict(*[ptr(x) for x in arrs], dt, len(arrs[0]))
    return arrs


def correct(pos4, v, f, a, b, c, dt, npos, energy_loss):
    r, v, a, b, c = [f64(x).copy() for x in (pos4, v, a, b, c)]
    flag = C.c_int32()
    lib().orc_ref_correct(ptr(r), ptr(v), ptr(f64(f)), ptr(a), ptr(b), ptr(c), dt, len(r), npos,
                          C.byref(flag), energy_loss)
    return r, v, a, b, c, flag.value


def accelerations(pos4, f):
    a = np.zeros((len(pos4), 3))
    lib().orc_ref_accelerations(ptr(f64(pos4)), ptr(a), ptr(f64(f)), len(pos4))
    return a


def kinetics(pos4, vel3):
    out = np.zeros(8)
    lib().orc_ref_kinetics(ptr(f64(pos4)), ptr(f64(vel3)), len(pos4), ptr(out))
    return out


def potential(pos4):
    out = np.zeros(2)
    lib().orc_ref_potential(ptr(f64(pos4)), len(pos4), ptr(out))
    return out


def scalar2(fn, *args):
    a, b = C.c_double(), C.c_double()
    getattr(lib(), fn)(*args, C.byref(a), C.byref(b))
    return a.value, b.value


def build_info() -> str:
    return lib().orc_build_info().decode()


def param(which, ti, tj):
    return lib().orc_ref_param(which, ti, tj)


RUN_STATS = ("x+ reflect", "x+ teleport", "x-", "y+", "y-", "z+", "z-", "hit steps", "lj builds",
             "builds after teleport")


def ref_run(pos4, vel3, n_lj, n_mb, timesteps, output_timesteps, dt, lj_rs, lj_rcut, mb_rs, mb_rcut,
            energy_loss=1.0, vxcm=0.0, use_lj=True, use_mb=False, max_records=64, wall=0.0, stats=False):
    """The reference loop; wall > 0 moves the walls to +-wall.  With
    stats=True a fifth value holds the loop's counts keyed by RUN_STATS."""
    st = np.zeros(len(RUN_STATS), dtype=np.int64)
    cfg = RunCfg(n_lj, n_mb, timesteps, output_timesteps, dt, lj_rs, lj_rcut, mb_rs, mb_rcut,
                 energy_loss, vxcm, 1 if use_lj else 0, 1 if use_mb else 0, wall, st.ctypes.data)
    r, v = f64(pos4).copy(), f64(vel3).copy()
    rec = np.zeros((max_records, 11))
    el = C.c_double()
    k = lib().orc_ref_run(C.byref(cfg), ptr(r), ptr(v), ptr(rec), max_records, C.byref(el))
    assert k >= 0
    if stats:
        return r, v, rec[:k], el.value, dict(zip(RUN_STATS, st.tolist()))
    return r, v, rec[:k], el.value


# ---------------- periodic ----------------
def pbc_list(pos3, L, rs):
    pos3 = f64(pos3)
    n = len(pos3)
    nls, nlg = C.c_int32(), C.c_int32()
    p = lib().orc_pbc_build_list(ptr(pos3), n, L, rs, C.byref(nls), C.byref(nlg))
    assert p, "periodic list needs at least 3 cells per axis"
    arr = np.ctypeslib.as_array(p, shape=(max(n * nls.value, 1),))[: n * nls.value].copy().reshape(n, nls.value)
    lib().orc_free(C.cast(p, C.c_void_p))
    return arr, nlg.value


PAIR_CHECK = ("entries", "missing within rs", "beyond rs + slack", "in slack band", "bad entries",
              "pairs within rc", "rc pairs listed", "pairs within rs")


def pair_list_check(pos3, L, rs, rc, rows, counts, slack2):
    """Counts (keyed by PAIR_CHECK) of an exported pair list against the
    exact FP64 sets within rs and rc at the build positions pos3."""
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    counts = np.ascontiguousarray(counts, dtype=np.int32)
    st = np.zeros(8, dtype=np.int64)
    rc_ = lib().orc_pbc_pair_list_check(ptr(f64(pos3)), len(pos3), L, rs, rc, slack2, ptr(rows), ptr(counts),
                                        rows.shape[1], ptr(st))
    assert rc_ == 0
    return dict(zip(PAIR_CHECK, st.tolist()))


def pbc_lj_forces(pos3, L, rc, nlist, sigma=1.0, eps=1.0):
    pos3 = f64(pos3)
    f = np.zeros_like(pos3)
    nlist = np.ascontiguousarray(nlist, dtype=np.int32)
    w = C.c_double()
    e = lib().orc_pbc_lj_forces(ptr(pos3), ptr(f), len(pos3), L, sigma, eps, rc, nlist.shape[1], ptr(nlist), C.byref(w))
    return f, e, w.value


def pbc_lj_run(pos3, vel3, L, rc, rs, dt, nsteps, rebuild_every_step=False, threads=1):
    """The oracle's periodic LJ loop; threads > 1 runs its OpenMP variant."""
    r, v = f64(pos3).copy(), f64(vel3).copy()
    en = np.zeros((nsteps + 1, 3))
    el = C.c_double()
    builds = lib().orc_pbc_lj_run_threads(ptr(r), ptr(v), len(r), L, rc, rs, dt, nsteps,
                                          1 if rebuild_every_step else 0, ptr(en), C.byref(el), threads)
    assert builds >= 0
    return r, v, en, builds, el.value


def gear_predict(r, v, a, b, c, dt):
    arrs = [f64(x).copy() for x in (r, v, a, b, c)]
    lib().orc_gear_predict(*[ptr(x) for x in arrs], dt, len(arrs[0]))
    return arrs


def gear_correct(r, v, f, a, b, c, dt, mass=1.0, scale=1.0):
    r, v, a, b, c = [f64(x).copy() for x in (r, v, a, b, c)]
    lib().orc_gear_correct(ptr(r), ptr(v), ptr(f64(f)), ptr(a), ptr(b), ptr(c), dt, len(r), mass, scale)
    return r, v, a, b, c


def tersoff_params() -> TersoffParams:
    p = TersoffParams()
    lib().orc_tersoff_si_params(C.byref(p))
    return p


def tersoff(pos3, L, p=None, forces=True):
    p = p or tersoff_params()
    pos3 = f64(pos3)
    f = np.zeros_like(pos3) if forces else None
    e = lib().orc_tersoff_energy_forces(C.byref(p), ptr(pos3), len(pos3), L, ptr(f))
    return e, f


def tersoff_centres(pos3, L, centre, p=None):
    """Per-centre Tersoff: (energy of the selected centres, minus its
    gradient, per-atom energies e_i of the selected centres)."""
    p = p or tersoff_params()
    pos3 = f64(pos3)
    mask = np.ascontiguousarray(np.asarray(centre, dtype=np.uint8))
    f = np.zeros_like(pos3)
    e_atom = np.zeros(len(pos3))
    e = lib().orc_tersoff_centres(C.byref(p), ptr(pos3), len(pos3), L, ptr(mask), ptr(f), ptr(e_atom))
    return e, f, e_atom


def tersoff_run(pos3, vel3, L, dt, nsteps, mass, scale):
    """Tersoff silicon, FP64 Gear PC with forces recomputed by the oracle at
    every step (no list): the trajectory the engine's periodic run must follow."""
    r, v = f64(pos3).copy(), f64(vel3).copy()
    a, b, c = np.zeros_like(r), np.zeros_like(r), np.zeros_like(r)
    _, f = tersoff(r, L)
    a[:] = f / mass * scale
    for _ in range(nsteps):
        r, v, a, b, c = gear_predict(r, v, a, b, c, dt)
        _, f = tersoff(r, L)
        r, v, a, b, c = gear_correct(r, v, f, a, b, c, dt, mass, scale)
    return r, v
