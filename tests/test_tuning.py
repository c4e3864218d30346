"""Launch shapes from the tuning table (moleculardynamics_amd/csrc/tuning.hip,
moleculardynamics_amd/tune.py).

CPU: mdx_tuning_lookup's reading of a table (first matching entry by
architecture prefix, kind and atom range; built-in shape otherwise; malformed
files and out-of-range shapes rejected), and the committed table itself.
GPU: mdx_create and mdx_batch_create take their shape from the table named by
MDX_TUNING_FILE; the single box's trajectory does not depend on the shape
(bit for bit, with the same list rebuilds), the batch's within FP64 rounding.
"""
import json
import os
from contextlib import contextmanager
from pathlib import Path

import numpy as np
import pytest

from moleculardynamics_amd import (BOX_PERIODIC, LJ_UNIFORM, PREC_FP64, Batch, MdxError, System, initial,
                                   tuning_lookup)

TABLE = Path(__file__).resolve().parent.parent / "moleculardynamics_amd" / "data" / "tuning.json"
BIG = 2 ** 31 - 1


def table(tmp_path, entries, name="t.json", fmt="mdx-tuning-1"):
    p = tmp_path / name
    p.write_text(json.dumps({"format": fmt, "entries": entries}))
    return p


def test_lookup_defaults_and_matching(tmp_path):
    p = table(tmp_path, [
        {"arch": "gfx942", "kind": "lj", "atoms_min": 0, "atoms_max": BIG, "block": 1024},
        {"arch": "gfx950", "kind": "replicas", "atoms_min": 0, "atoms_max": BIG, "stage_aim": 900},
        {"arch": "gfx950", "kind": "lj", "atoms_min": 1000, "atoms_max": 5000, "block": 1024, "fill": 0.7,
         "ms_per_step": 0.1, "candidates": [{"block": 512}]},
        {"arch": "gfx950", "kind": "lj", "atoms_min": 1000, "atoms_max": 9000, "block": 512, "fill": 1.0},
    ])
    default = {"block": 512, "fill": 0.85, "stage_aim": 1250.0, "source": "default"}
    # no entry below atoms_min, nor for another architecture
    assert tuning_lookup(p, "gfx950:sramecc+:xnack-", "lj", 999) == default
    assert tuning_lookup(p, "gfx90a", "lj", 2000) == default
    # the first match wins; the architecture matches by prefix; unknown keys are ignored
    assert tuning_lookup(p, "gfx950:sramecc+:xnack-", "lj", 1000) == \
        {"block": 1024, "fill": 0.7, "stage_aim": 1250.0, "source": "tuned"}
    assert tuning_lookup(p, "gfx950", "lj", 5000) == {"block": 512, "fill": 1.0, "stage_aim": 1250.0,
                                                      "source": "tuned"}
    # atoms_max is exclusive
    assert tuning_lookup(p, "gfx950", "lj", 9000) == default
    assert tuning_lookup(p, "gfx950", "replicas", 10 ** 6) == {"block": 512, "fill": 0.85, "stage_aim": 900.0,
                                                               "source": "tuned"}
    # an empty entry list is a valid table
    assert tuning_lookup(table(tmp_path, [], "e.json"), "gfx950", "lj", 10) == default


@pytest.mark.parametrize("bad, why", [
    ('{"format": "mdx-tuning-1", "entries": [', "unexpected end"),
    ('{"format": "other", "entries": []}', "not an mdx-tuning-1"),
    ('{"format": "mdx-tuning-1"}', "no entries"),
    ('{"format": "mdx-tuning-1", "entries": []} x', "trailing"),
    ('{"format": "mdx-tuning-1", "entries": [{"arch": "gfx950", "kind": "lj", "atoms_min": 0, '
     '"atoms_max": 10, "block": 640}]}', "block must be 512 or 1024"),
    ('{"format": "mdx-tuning-1", "entries": [{"arch": "gfx950", "kind": "lj", "atoms_min": 0, '
     '"atoms_max": 10, "fill": 1.5}]}', "fill must lie"),
    ('{"format": "mdx-tuning-1", "entries": [{"arch": "gfx950", "kind": "replicas", "atoms_min": 0, '
     '"atoms_max": 10, "stage_aim": 100}]}', "stage_aim must lie"),
])
def test_lookup_rejects_bad_tables(tmp_path, bad, why):
    p = tmp_path / "bad.json"
    p.write_text(bad)
    with pytest.raises(MdxError) as e:
        tuning_lookup(p, "gfx950", "lj" if "replicas" not in bad else "replicas", 5)
    assert why in str(e.value)


def test_lookup_missing_file(tmp_path):
    with pytest.raises(MdxError) as e:
        tuning_lookup(tmp_path / "none.json", "gfx950", "lj", 5)
    assert "cannot read" in str(e.value)


def test_committed_table():
    """The shipped table parses and gives a shape for the benchmark systems;
    each tuned entry records timings that show it beat the built-in shape."""
    data = json.loads(TABLE.read_text())
    assert data["format"] == "mdx-tuning-1"
    for n in (4000, 1048576, 8_000_000):
        sh = tuning_lookup(TABLE, "gfx950", "lj", n)
        assert sh["block"] in (512, 1024) and 0.25 <= sh["fill"] <= 1.0
    tuning_lookup(TABLE, "gfx950", "replicas", 1234468)
    for e in data["entries"]:
        assert e["ms_per_step"] <= e["default_ms_per_step"]
        base = {"lj": {"block": 512, "fill": 0.85}, "replicas": {"stage_aim": 1250.0, "fill": 0.85}}[e["kind"]]
        if any(e[k] != v for k, v in base.items()):
            assert e["ms_per_step"] < 0.98 * e["default_ms_per_step"]


@contextmanager
def tuning_file(path):
    old = os.environ.get("MDX_TUNING_FILE")
    os.environ["MDX_TUNING_FILE"] = str(path)
    try:
        yield
    finally:
        if old is None:
            os.environ.pop("MDX_TUNING_FILE", None)
        else:
            os.environ["MDX_TUNING_FILE"] = old


def lj_system(n, L):
    return System(n_atoms=n, box=BOX_PERIODIC, precision=PREC_FP64, use_lj=1, lj_table=LJ_UNIFORM, dt=0.005,
                  lj_rs=2.8, lj_rcut=2.5, box_length=L, lj_sigma=1.0, lj_eps=1.0, mass=1.0)


@pytest.mark.gpu
def test_single_box_shape_from_table_is_exact(tmp_path):
    """Workgroups of 1024 threads and segments of 0.7 of a pass against the
    built-in 512 and 0.85: each atom's candidates come in the same order, so
    the runs agree bit for bit and rebuild their lists at the same steps."""
    pos, vel, L = initial.lj_fluid(16, 0.8442, 0.728, seed=11)
    n = len(pos)
    out = []
    for entries in ([], [{"arch": "gfx950", "kind": "lj", "atoms_min": 0, "atoms_max": BIG, "block": 1024,
                          "fill": 0.7}]):
        with tuning_file(table(tmp_path, entries, f"lj{len(entries)}.json")):
            s = lj_system(n, L)
        with s:
            shape = s.launch_shape()
            s.set_state(np.c_[pos, np.zeros(n)], vel)
            s.run(150)
            st = s.get_state()
            out.append((shape, st, s.build_count()))
    (sh0, st0, b0), (sh1, st1, b1) = out
    assert sh0 == {"block": 512, "fill": 0.85, "stage_aim": 1250.0, "source": "default"}
    assert sh1 == {"block": 1024, "fill": 0.7, "stage_aim": 1250.0, "source": "tuned"}
    assert b0 == b1 >= 3
    np.testing.assert_array_equal(st0["pos"], st1["pos"])
    np.testing.assert_array_equal(st0["vel"], st1["vel"])


@pytest.mark.gpu
def test_batch_shape_from_table(tmp_path):
    """A smaller staging aim and fill cut the staged replicas into more
    segments (and may move a replica between paths); the trajectories agree to
    FP64 rounding."""
    ns, Ls, pos, vel = initial.replica_batch(24, seed=5, n_max=4096)
    out = []
    for entries in ([], [{"arch": "gfx950", "kind": "replicas", "atoms_min": 0, "atoms_max": BIG,
                          "stage_aim": 900, "fill": 0.7}]):
        with tuning_file(table(tmp_path, entries, f"r{len(entries)}.json")):
            b = Batch(ns, Ls, precision=PREC_FP64, dt=0.005, lj_rs=2.8, lj_rcut=2.5)
        with b:
            shape = b.launch_shape()
            b.set_state(pos, vel)
            b.run(100)
            out.append((shape, b.get_state()))
    (sh0, st0), (sh1, st1) = out
    assert sh0["source"] == "default" and sh1 == {"block": 512, "fill": 0.7, "stage_aim": 900.0,
                                                  "source": "tuned"}
    L = np.repeat(Ls, ns)[:, None]   # positions may be wrapped at different builds
    d = st0["pos"] - st1["pos"]
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-9
    assert np.abs(st0["vel"] - st1["vel"]).max() < 1e-9


@pytest.mark.gpu
def test_bad_table_fails_create(tmp_path):
    p = tmp_path / "bad.json"
    p.write_text('{"format": "mdx-tuning-1", "entries": [{"arch": "gfx950", "kind": "lj", "atoms_min": 0, '
                 '"atoms_max": 2147483647, "block": 768}]}')
    pos, vel, L = initial.lj_fluid(10, 0.8442, 0.728, seed=1)
    with tuning_file(p), pytest.raises(MdxError) as e:
        lj_system(len(pos), L)
    assert "block must be 512 or 1024" in str(e.value)
