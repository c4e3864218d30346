"""ctypes binding of the C-ABI in include/mdx.h (libmdx.so, built in-tree).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB_PATH = PKG / "libmdx.so"
HEADER = ROOT / "include" / "mdx.h"

MDX_OK = 0
ERRORS = {
    -1: "MDX_ERR_ARG", -2: "MDX_ERR_HIP", -3: "MDX_ERR_ALLOC", -4: "MDX_ERR_GRID",
    -5: "MDX_ERR_CAPACITY", -6: "MDX_ERR_STATE", -7: "MDX_ERR_LIST_STALE",
}
BOX_WALLED, BOX_PERIODIC = 0, 1
PREC_MIXED, PREC_FP64 = 0, 1
LJ_REFERENCE, LJ_UNIFORM = 0, 1
BOND_REFERENCE, BOND_TERSOFF, BOND_SURROGATE = 0, 1, 2
SG_NFEAT, SG_NTRAIN, SG_NPARAM = 48, 7361, 7553
LIST_LJ, LIST_BOND = 0, 1


class MdxConfig(C.Structure):
    _fields_ = [
        ("n_atoms", C.c_int32), ("n_bond", C.c_int32), ("box", C.c_int32), ("precision", C.c_int32),
        ("use_lj", C.c_int32), ("use_bond", C.c_int32), ("lj_table", C.c_int32), ("bond_form", C.c_int32),
        ("dt", C.c_float), ("lj_rs", C.c_float), ("lj_rcut", C.c_float), ("mb_rs", C.c_float),
        ("mb_rcut", C.c_float), ("energy_loss", C.c_float),
        ("wall", C.c_double), ("box_length", C.c_double),
        ("lj_sigma", C.c_double), ("lj_eps", C.c_double), ("mass", C.c_double),
        ("accel_scale", C.c_double), ("max_neighbors", C.c_int32), ("reserved", C.c_int32 * 7),
    ]


SHAPE_DEFAULT, SHAPE_TUNED = 0, 1


class MdxShape(C.Structure):
    _fields_ = [("block", C.c_int32), ("fill", C.c_float), ("stage_aim", C.c_float), ("source", C.c_int32)]

    def as_dict(self) -> dict:
        return {"block": self.block, "fill": round(float(self.fill), 6), "stage_aim": round(float(self.stage_aim), 3),
                "source": "tuned" if self.source == SHAPE_TUNED else "default"}


class MdxError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} -> {ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def header_symbols() -> list[str]:
    """Entry points declared in include/mdx.h."""
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(mdx_\w+)\s*\(", text, re.M)))


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(str(LIB_PATH))
    P, I32, I64, D = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    sig = {
        "mdx_create": [C.POINTER(MdxConfig), C.POINTER(P)],
        "mdx_destroy": [P],
        "mdx_set_state": [P, P, P, P, P, P],
        "mdx_get_state": [P, P, P, P, P, P, P],
        "mdx_ready_list": [P, C.POINTER(D), C.POINTER(D)],
        "mdx_build_list": [P, I32, D, C.POINTER(I32), C.POINTER(I32)],
        "mdx_copy_neighbor_list": [P, I32, P, I64],
        "mdx_copy_pair_list": [P, P, I64, P, C.POINTER(I32), P],
        "mdx_predict": [P], "mdx_lj_forces": [P], "mdx_bond_forces": [P],
        "mdx_correct": [P, C.POINTER(I32)], "mdx_accelerations": [P],
        "mdx_run": [P, I32], "mdx_build_count": [P, C.POINTER(I64)],
        "mdx_kinetics": [P, D, P], "mdx_potential": [P, P], "mdx_potential_allpairs": [P, P],
        "mdx_potential_virial": [P, P],
        "mdx_set_timing": [P, I32], "mdx_kernel_times": [P, P, P, P, I32, C.POINTER(I32)],
        "mdx_synchronize": [P], "mdx_set_forces": [P, P],
        "mdx_batch_create": [C.POINTER(MdxConfig), I32, P, P, C.POINTER(P)],
        "mdx_batch_destroy": [P], "mdx_batch_set_state": [P, P, P], "mdx_batch_get_state": [P, P, P, P],
        "mdx_batch_run": [P, I32], "mdx_batch_build_counts": [P, P], "mdx_batch_staged": [P, P],
        "mdx_batch_energies": [P, P], "mdx_batch_virials": [P, P],
        "mdx_batch_set_timing": [P, I32], "mdx_batch_kernel_times": [P, P, P, P, I32, C.POINTER(I32)],
        "mdx_batch_synchronize": [P],
        "mdx_launch_shape": [P, C.POINTER(MdxShape)], "mdx_batch_launch_shape": [P, C.POINTER(MdxShape)],
        "mdx_tuning_lookup": [C.c_char_p, C.c_char_p, C.c_char_p, I64, C.POINTER(MdxShape)],
        "mdx_set_surrogate": [P, P, I64], "mdx_surrogate_counts": [P, P],
        "mdx_sg_trainer_create": [I32, I32, D, P, P, P, C.POINTER(P)], "mdx_sg_trainer_destroy": [P],
        "mdx_sg_trainer_feature_stats": [P, P], "mdx_sg_trainer_descriptors": [P, I32, P],
        "mdx_sg_trainer_set_params": [P, P, I64], "mdx_sg_trainer_get_params": [P, P, I64],
        "mdx_sg_trainer_gradient": [P, P, I32, D, P, P], "mdx_sg_trainer_step": [P, P, I32, D, D, P],
        "mdx_sweep_create": [C.POINTER(MdxConfig), I32, C.POINTER(P)], "mdx_sweep_destroy": [P],
        "mdx_sweep_member": [P, I32, C.POINTER(P)], "mdx_sweep_run": [P, I32], "mdx_sweep_set_timing": [P, I32],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    lib.mdx_last_error.argtypes = [P]
    lib.mdx_last_error.restype = C.c_char_p
    lib.mdx_batch_last_error.argtypes = [P]
    lib.mdx_batch_last_error.restype = C.c_char_p
    lib.mdx_sweep_last_error.argtypes = [P]
    lib.mdx_sweep_last_error.restype = C.c_char_p
    lib.mdx_sg_trainer_last_error.argtypes = [P]
    lib.mdx_sg_trainer_last_error.restype = C.c_char_p
    lib.mdx_tuning_last_error.argtypes = []
    lib.mdx_tuning_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def tuning_lookup(path, arch: str, kind: str, atoms: int) -> dict:
    """The launch shape a tuning table gives a system of `kind` ("lj" or
    "replicas") of `atoms` atoms on a device of architecture `arch` (host code,
    no device needed)."""
    lib = load()
    sh = MdxShape()
    rc = lib.mdx_tuning_lookup(str(path).encode(), arch.encode(), kind.encode(), int(atoms), C.byref(sh))
    if rc != 0:
        raise MdxError("mdx_tuning_lookup", rc, lib.mdx_tuning_last_error().decode())
    return sh.as_dict()


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def make_config(**cfg) -> MdxConfig:
    c = MdxConfig()
    c.accel_scale = 1.0
    c.lj_sigma = c.lj_eps = c.mass = 1.0
    c.energy_loss = 1.0
    for k, v in cfg.items():
        setattr(c, k, v)
    return c


class System:
    """Device-resident system; a thin owner of an mdx_system handle (or a
    borrowed handle of a sweep member, which the sweep destroys)."""

    def __init__(self, _handle=None, _cfg=None, **cfg):
        self.lib = load()
        if _handle is not None:
            self.cfg, self.n, self.h, self.owned = _cfg, _cfg.n_atoms, _handle, False
            return
        c = make_config(**cfg)
        self.cfg = c
        self.n = c.n_atoms
        self.owned = True
        h = C.c_void_p()
        rc = self.lib.mdx_create(C.byref(c), C.byref(h))
        if rc != MDX_OK:
            raise MdxError("mdx_create", rc, self.lib.mdx_last_error(None).decode())
        self.h = h

    def _chk(self, fn: str, rc: int):
        if rc != MDX_OK:
            raise MdxError(fn, rc, self.lib.mdx_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            if self.owned:
                self.lib.mdx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state --
    def set_state(self, pos4, vel3, acc3=None, b3=None, c3=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (pos4, vel3, acc3, b3, c3)]
        assert arrs[0].shape == (self.n, 4) and arrs[1].shape == (self.n, 3)
        self._chk("mdx_set_state", self.lib.mdx_set_state(self.h, *[_ptr(a) for a in arrs]))

    def get_state(self):
        out = {"pos": np.zeros((self.n, 4)), "vel": np.zeros((self.n, 3)), "acc": np.zeros((self.n, 3)),
               "b": np.zeros((self.n, 3)), "c": np.zeros((self.n, 3)), "force": np.zeros((self.n, 3))}
        self._chk("mdx_get_state", self.lib.mdx_get_state(
            self.h, *[_ptr(out[k]) for k in ("pos", "vel", "acc", "b", "c", "force")]))
        return out

    # -- reference entry points --
    def ready_list(self):
        vmax, box = C.c_double(), C.c_double()
        self._chk("mdx_ready_list", self.lib.mdx_ready_list(self.h, C.byref(vmax), C.byref(box)))
        return vmax.value, box.value

    def build_list(self, which=LIST_LJ, box_size_list=0.0):
        nls, nlg = C.c_int32(), C.c_int32()
        self._chk("mdx_build_list", self.lib.mdx_build_list(self.h, which, box_size_list, C.byref(nls), C.byref(nlg)))
        return nls.value, nlg.value

    def neighbor_list(self, which=LIST_LJ, nlistsize=None, items=None):
        items = self.n if items is None else items
        out = np.zeros(max(items * nlistsize, 1), dtype=np.int32)
        self._chk("mdx_copy_neighbor_list", self.lib.mdx_copy_neighbor_list(self.h, which, _ptr(out), out.size))
        return out[: items * nlistsize].reshape(items, nlistsize)

    def pair_list(self):
        """(rows, counts, build positions) of the list the periodic force
        kernels read: rows[i, :counts[i]] are atom i's partners (caller
        indices), -1 after them."""
        w = C.c_int32()
        pos = np.zeros((self.n, 3))
        self._chk("mdx_copy_pair_list", self.lib.mdx_copy_pair_list(self.h, None, 0, None, C.byref(w), _ptr(pos)))
        rows = np.zeros((self.n, max(w.value, 1)), dtype=np.int32)
        counts = np.zeros(self.n, dtype=np.int32)
        self._chk("mdx_copy_pair_list", self.lib.mdx_copy_pair_list(self.h, _ptr(rows), rows.size, _ptr(counts),
                                                                    C.byref(w), None))
        return rows[:, : w.value], counts, pos

    def predict(self):
        self._chk("mdx_predict", self.lib.mdx_predict(self.h))

    def lj_forces(self):
        self._chk("mdx_lj_forces", self.lib.mdx_lj_forces(self.h))

    def bond_forces(self):
        self._chk("mdx_bond_forces", self.lib.mdx_bond_forces(self.h))

    def correct(self):
        hit = C.c_int32()
        self._chk("mdx_correct", self.lib.mdx_correct(self.h, C.byref(hit)))
        return hit.value

    def accelerations(self):
        self._chk("mdx_accelerations", self.lib.mdx_accelerations(self.h))

    def run(self, nsteps):
        self._chk("mdx_run", self.lib.mdx_run(self.h, int(nsteps)))

    def build_count(self):
        b = C.c_int64()
        self._chk("mdx_build_count", self.lib.mdx_build_count(self.h, C.byref(b)))
        return b.value

    def launch_shape(self) -> dict:
        """The LDS kernels' launch shape and where it came from."""
        sh = MdxShape()
        self._chk("mdx_launch_shape", self.lib.mdx_launch_shape(self.h, C.byref(sh)))
        return sh.as_dict()

    def kinetics(self, kboltz):
        out = np.zeros(8)
        self._chk("mdx_kinetics", self.lib.mdx_kinetics(self.h, kboltz, _ptr(out)))
        return out

    def potential(self):
        out = np.zeros(2)
        self._chk("mdx_potential", self.lib.mdx_potential(self.h, _ptr(out)))
        return out

    def potential_virial(self):
        """{LJ energy, bond-order energy, LJ virial, bond-order virial}."""
        out = np.zeros(4)
        self._chk("mdx_potential_virial", self.lib.mdx_potential_virial(self.h, _ptr(out)))
        return out

    def potential_allpairs(self):
        out = np.zeros(2)
        self._chk("mdx_potential_allpairs", self.lib.mdx_potential_allpairs(self.h, _ptr(out)))
        return out

    def synchronize(self):
        self._chk("mdx_synchronize", self.lib.mdx_synchronize(self.h))

    def set_surrogate(self, params):
        """Upload the surrogate network (MDX_SG_NPARAM floats, layout of include/mdx.h)."""
        b = np.ascontiguousarray(params, dtype=np.float32)
        assert b.shape == (SG_NPARAM,)
        self._chk("mdx_set_surrogate", self.lib.mdx_set_surrogate(self.h, _ptr(b), SG_NPARAM))

    def surrogate_counts(self):
        """(atom evaluations by the Tersoff fallback, all atom evaluations)."""
        out = np.zeros(2, dtype=np.int64)
        self._chk("mdx_surrogate_counts", self.lib.mdx_surrogate_counts(self.h, _ptr(out)))
        return int(out[0]), int(out[1])

    def set_timing(self, on: bool):
        self._chk("mdx_set_timing", self.lib.mdx_set_timing(self.h, 1 if on else 0))

    def kernel_times(self):
        cap = 64
        names = C.create_string_buffer(48 * cap)
        counts = np.zeros(cap, dtype=np.int64)
        tot = np.zeros(cap)
        n = C.c_int32()
        self._chk("mdx_kernel_times", self.lib.mdx_kernel_times(self.h, names, _ptr(counts), _ptr(tot), cap, C.byref(n)))
        res = {}
        for k in range(n.value):
            nm = names.raw[48 * k: 48 * k + 48].split(b"\0", 1)[0].decode()
            res[nm] = (int(counts[k]), float(tot[k]))
        return res


def _kernel_table(names, counts, tot, n):
    res = {}
    for k in range(n):
        nm = names.raw[48 * k: 48 * k + 48].split(b"\0", 1)[0].decode()
        res[nm] = (int(counts[k]), float(tot[k]))
    return res


class Batch:
    """Independent periodic Lennard-Jones replicas advanced together
    (mdx_batch_*).  State arrays hold the replicas one after another."""

    def __init__(self, n_atoms, box_length, **cfg):
        self.lib = load()
        c = MdxConfig()
        c.accel_scale = 1.0
        c.lj_sigma = c.lj_eps = c.mass = 1.0
        for k, v in cfg.items():
            setattr(c, k, v)
        self.n_atoms = np.ascontiguousarray(n_atoms, dtype=np.int32)
        self.box = np.ascontiguousarray(box_length, dtype=np.float64)
        self.R = len(self.n_atoms)
        self.n = int(self.n_atoms.sum())
        h = C.c_void_p()
        rc = self.lib.mdx_batch_create(C.byref(c), self.R, _ptr(self.n_atoms), _ptr(self.box), C.byref(h))
        if rc != MDX_OK:
            raise MdxError("mdx_batch_create", rc, self.lib.mdx_batch_last_error(None).decode())
        self.h = h

    def _chk(self, fn: str, rc: int):
        if rc != MDX_OK:
            raise MdxError(fn, rc, self.lib.mdx_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.mdx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_state(self, pos3, vel3):
        p = np.ascontiguousarray(pos3, dtype=np.float64)
        v = np.ascontiguousarray(vel3, dtype=np.float64)
        assert p.shape == (self.n, 3) and v.shape == (self.n, 3)
        self._chk("mdx_batch_set_state", self.lib.mdx_batch_set_state(self.h, _ptr(p), _ptr(v)))

    def get_state(self):
        p, v, f = (np.zeros((self.n, 3)) for _ in range(3))
        self._chk("mdx_batch_get_state", self.lib.mdx_batch_get_state(self.h, _ptr(p), _ptr(v), _ptr(f)))
        return {"pos": p, "vel": v, "force": f}

    def run(self, nsteps):
        self._chk("mdx_batch_run", self.lib.mdx_batch_run(self.h, int(nsteps)))

    def build_counts(self):
        b = np.zeros(self.R, dtype=np.int64)
        self._chk("mdx_batch_build_counts", self.lib.mdx_batch_build_counts(self.h, _ptr(b)))
        return b

    def paths(self):
        """Per replica, fixed at creation: 1 the LDS-staged list and step
        kernels, 2 the one-workgroup all-pairs kernel of narrow boxes, 0 the
        gather kernels."""
        s = np.zeros(self.R, dtype=np.int32)
        self._chk("mdx_batch_staged", self.lib.mdx_batch_staged(self.h, _ptr(s)))
        return s

    def staged(self):
        """True for the replicas on the LDS-staged path."""
        return self.paths() == 1

    def launch_shape(self) -> dict:
        """The staged replicas' launch shape and where it came from."""
        sh = MdxShape()
        self._chk("mdx_batch_launch_shape", self.lib.mdx_batch_launch_shape(self.h, C.byref(sh)))
        return sh.as_dict()

    def energies(self):
        e = np.zeros((self.R, 2))
        self._chk("mdx_batch_energies", self.lib.mdx_batch_energies(self.h, _ptr(e)))
        return e

    def virials(self):
        """Per replica the virial sum over pairs of r_ij . F_ij."""
        out = np.zeros(self.R)
        self._chk("mdx_batch_virials", self.lib.mdx_batch_virials(self.h, _ptr(out)))
        return out

    def synchronize(self):
        self._chk("mdx_batch_synchronize", self.lib.mdx_batch_synchronize(self.h))

    def set_timing(self, on: bool):
        self._chk("mdx_batch_set_timing", self.lib.mdx_batch_set_timing(self.h, 1 if on else 0))

    def kernel_times(self):
        cap = 64
        names = C.create_string_buffer(48 * cap)
        counts = np.zeros(cap, dtype=np.int64)
        tot = np.zeros(cap)
        n = C.c_int32()
        self._chk("mdx_batch_kernel_times",
                  self.lib.mdx_batch_kernel_times(self.h, names, _ptr(counts), _ptr(tot), cap, C.byref(n)))
        return _kernel_table(names, counts, tot, n.value)


class Sweep:
    """Walled systems of one configuration advanced together (mdx_sweep_*):
    the reference's vxcm sweep as one batch.  member(m) is the m-th member's
    System (state, lists, diagnostics), owned by the sweep."""

    def __init__(self, members, **cfg):
        self.lib = load()
        self.cfg = make_config(**cfg)
        self.R = int(members)
        h = C.c_void_p()
        rc = self.lib.mdx_sweep_create(C.byref(self.cfg), self.R, C.byref(h))
        if rc != MDX_OK:
            raise MdxError("mdx_sweep_create", rc, self.lib.mdx_sweep_last_error(None).decode())
        self.h = h
        self.members = []
        for m in range(self.R):
            mh = C.c_void_p()
            self._chk("mdx_sweep_member", self.lib.mdx_sweep_member(self.h, m, C.byref(mh)))
            self.members.append(System(_handle=mh, _cfg=self.cfg))

    def _chk(self, fn: str, rc: int):
        if rc != MDX_OK:
            raise MdxError(fn, rc, self.lib.mdx_sweep_last_error(self.h).decode())

    def member(self, m) -> System:
        return self.members[m]

    def run(self, nsteps):
        self._chk("mdx_sweep_run", self.lib.mdx_sweep_run(self.h, int(nsteps)))

    def set_timing(self, on: bool):
        self._chk("mdx_sweep_set_timing", self.lib.mdx_sweep_set_timing(self.h, 1 if on else 0))

    def close(self):
        if getattr(self, "h", None):
            for m in self.members:
                m.h = None
            self.lib.mdx_sweep_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class SgTrainer:
    """Training set of the Tersoff surrogate on the device (mdx_sg_trainer_*):
    frames of n_atoms atoms in one periodic box with reference forces and
    energies; minibatch gradients and Adam steps run on the GPU."""

    def __init__(self, pos, force, energy, box_length):
        self.lib = load()
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        force = np.ascontiguousarray(force, dtype=np.float64)
        energy = np.ascontiguousarray(energy, dtype=np.float64)
        self.nf, self.na = pos.shape[:2]
        assert pos.shape == (self.nf, self.na, 3) and force.shape == pos.shape and energy.shape == (self.nf,)
        h = C.c_void_p()
        rc = self.lib.mdx_sg_trainer_create(self.nf, self.na, float(box_length), _ptr(pos), _ptr(force),
                                            _ptr(energy), C.byref(h))
        if rc != MDX_OK:
            raise MdxError("mdx_sg_trainer_create", rc, self.lib.mdx_sg_trainer_last_error(None).decode())
        self.h = h

    def _chk(self, fn: str, rc: int):
        if rc != MDX_OK:
            raise MdxError(fn, rc, self.lib.mdx_sg_trainer_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.mdx_sg_trainer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def feature_stats(self):
        out = np.zeros((4, SG_NFEAT))
        self._chk("mdx_sg_trainer_feature_stats", self.lib.mdx_sg_trainer_feature_stats(self.h, _ptr(out)))
        return {"mean": out[0], "std": out[1], "min": out[2], "max": out[3]}

    def descriptors(self, frame):
        out = np.zeros((self.na, SG_NFEAT), dtype=np.float32)
        self._chk("mdx_sg_trainer_descriptors", self.lib.mdx_sg_trainer_descriptors(self.h, int(frame), _ptr(out)))
        return out

    def set_params(self, params):
        b = np.ascontiguousarray(params, dtype=np.float32)
        assert b.shape == (SG_NPARAM,)
        self._chk("mdx_sg_trainer_set_params", self.lib.mdx_sg_trainer_set_params(self.h, _ptr(b), SG_NPARAM))

    def get_params(self):
        b = np.zeros(SG_NPARAM, dtype=np.float32)
        self._chk("mdx_sg_trainer_get_params", self.lib.mdx_sg_trainer_get_params(self.h, _ptr(b), SG_NPARAM))
        return b

    def gradient(self, frames, w_energy=1.0, want_grad=True):
        """(loss, force mse, energy mse per atom) and the gradient of the trainable weights."""
        fr = np.ascontiguousarray(frames, dtype=np.int32)
        out = np.zeros(3)
        g = np.zeros(SG_NTRAIN, dtype=np.float32) if want_grad else None
        self._chk("mdx_sg_trainer_gradient",
                  self.lib.mdx_sg_trainer_gradient(self.h, _ptr(fr), len(fr), float(w_energy), _ptr(g), _ptr(out)))
        return out, g

    def step(self, frames, lr, w_energy=1.0, report=False):
        fr = np.ascontiguousarray(frames, dtype=np.int32)
        out = np.zeros(3) if report else None
        self._chk("mdx_sg_trainer_step",
                  self.lib.mdx_sg_trainer_step(self.h, _ptr(fr), len(fr), float(w_energy), float(lr), _ptr(out)))
        return out
