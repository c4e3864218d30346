"""The C-ABI library loads and exports every entry point include/mdx.h declares
(no device calls: these run without a GPU)."""
import ctypes as C
import subprocess

import pytest

from moleculardynamics_amd import _lib


def test_header_declares_entry_points():
    syms = _lib.header_symbols()
    for must in ("mdx_create", "mdx_destroy", "mdx_run", "mdx_build_list", "mdx_copy_neighbor_list",
                 "mdx_predict", "mdx_correct", "mdx_lj_forces", "mdx_bond_forces", "mdx_accelerations",
                 "mdx_ready_list", "mdx_kinetics", "mdx_potential", "mdx_potential_allpairs"):
        assert must in syms


def test_library_exports_every_declared_symbol():
    assert _lib.LIB_PATH.exists(), "run __graft_entry__.build() first"
    lib = C.CDLL(str(_lib.LIB_PATH))
    missing = [s for s in _lib.header_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(_lib.header_symbols()) <= exported


def test_library_is_built_for_gfx950_only():
    # the embedded offload bundle names its target triple
    blob = _lib.LIB_PATH.read_bytes()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob
    for other in (b"--gfx942", b"--gfx90a", b"--gfx1100"):
        assert other not in blob


def test_binding_signatures_cover_header():
    lib = _lib.load()
    for s in _lib.header_symbols():
        assert getattr(lib, s).restype is not None


def test_create_without_device_fails_cleanly():
    # on a machine without a GPU, mdx_create reports MDX_ERR_HIP instead of crashing;
    # on a GPU machine it succeeds
    lib = _lib.load()
    cfg = _lib.MdxConfig()
    cfg.n_atoms = 8
    cfg.box = _lib.BOX_PERIODIC
    cfg.use_lj = 1
    cfg.lj_table = _lib.LJ_UNIFORM
    cfg.lj_rs, cfg.lj_rcut, cfg.box_length, cfg.dt = 2.8, 2.5, 10.0, 0.005
    h = C.c_void_p()
    rc = lib.mdx_create(C.byref(cfg), C.byref(h))
    if rc == 0:
        lib.mdx_destroy(h)
    else:
        assert rc in (-2,), lib.mdx_last_error(None)
    # invalid configuration is rejected before any device call
    cfg.lj_rcut = 3.0
    assert lib.mdx_create(C.byref(cfg), C.byref(h)) == -1
    assert b"RCUT" in lib.mdx_last_error(None)


def test_sweep_create_checks_arguments_before_the_device():
    # a sweep holds walled members only, at least one; both are checked before
    # any device call, and a valid sweep on a machine without a GPU reports
    # MDX_ERR_HIP instead of crashing
    lib = _lib.load()
    cfg = _lib.MdxConfig()
    cfg.n_atoms, cfg.n_bond = 8, 2
    cfg.box = _lib.BOX_WALLED
    cfg.use_lj, cfg.use_bond = 1, 1
    cfg.lj_rs, cfg.lj_rcut, cfg.mb_rs, cfg.mb_rcut, cfg.dt = 7.6, 6.8, 3.6, 3.0, 1.0
    cfg.wall, cfg.energy_loss = 100.0, 1.0
    h = C.c_void_p()
    assert lib.mdx_sweep_create(C.byref(cfg), 0, C.byref(h)) == -1
    assert b"member" in lib.mdx_sweep_last_error(None)
    rc = lib.mdx_sweep_create(C.byref(cfg), 3, C.byref(h))
    if rc == 0:
        lib.mdx_sweep_destroy(h)
    else:
        assert rc == -2, lib.mdx_sweep_last_error(None)
    cfg.box = _lib.BOX_PERIODIC
    cfg.box_length = 20.0
    cfg.use_bond = 0
    assert lib.mdx_sweep_create(C.byref(cfg), 3, C.byref(h)) == -1
    assert b"walled" in lib.mdx_sweep_last_error(None)
