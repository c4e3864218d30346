"""The C++ host layer keeps the reference's configuration and file formats
(common/ConfigParser.cpp, common/Loader.cpp, common/Results.cpp).  No GPU."""
import ctypes as C
import json
import struct

import numpy as np
import pytest

from moleculardynamics_amd import _lib

HOSTLIB = _lib.PKG / "libmdx_host.so"

KEYS = {
    "Lennard Jones Particles": "400", "Many Body Particles": "40", "Timesteps": "1000", "DT": "1",
    "Temperature": "300", "LennardJonesRS": "7.6", "LennardJonesRCUT": "6.8", "ManyBodyRS": "3.6",
    "ManyBodyRCUT": "3.0", "Input": "in.txt", "Output": "out.txt", "Filename": "res", "CUDA Blocks": "64",
    "Use Cuda": "1", "UseLennardJones": "1", "UseManyBody": "1", "LennardJonesBPP": "0", "ManyBodyBPP": "0",
    "OutputTimesteps": "100", "Debug": "0", "Fallback": "1", "VXCMFROM": "0", "VXCMTO": "0.02",
    "VXCMSTEP": "0.01", "EnergyLoss": "0.81", "AnimationTimesteps": "-1", "AnimationData": "anim",
}


@pytest.fixture(scope="module")
def host():
    if not HOSTLIB.exists():
        pytest.skip("libmdx_host.so not built")
    L = C.CDLL(str(HOSTLIB))
    L.mdxh_read_configuration.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
    L.mdxh_config_lookup.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
    L.mdxh_write_output.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_void_p]
    L.mdxh_read_input.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_void_p]
    L.mdxh_sweep_values.argtypes = [C.c_float, C.c_float, C.c_float, C.c_void_p, C.c_int]
    L.mdxh_write_results.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_float, C.c_int, C.c_float, C.c_float,
                                     C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_void_p, C.c_int,
                                     C.c_void_p, C.c_int]
    return L


def lookup(host, text, key):
    buf = C.create_string_buffer(4096)
    rc = host.mdxh_config_lookup(text.encode(), key.encode(), buf, 4096)
    return rc, buf.value.decode()


def test_parser_rules(host):
    text = ("// comment line\n"
            "\n"
            "dt=0.5 // trailing comment\n"
            "Input=my file.txt   \n"
            "X=5    \n"
            "A=1\n"
            "a=2\n"
            "Spaced Key=7\n"
            "K =3\n")
    assert lookup(host, text, "DT") == (0, "0.5")           # comment cut, then trailing blanks
    assert lookup(host, text, "input") == (0, "my file.txt")  # keys are case-insensitive, blanks trimmed
    assert lookup(host, text, "x") == (0, "5    ")           # only the first character is non-blank: kept
    assert lookup(host, text, "A") == (0, "1")               # first occurrence wins
    assert lookup(host, text, "SPACED KEY") == (0, "7")
    assert lookup(host, text, "K")[0] == 1                   # key "K " keeps its blank
    assert lookup(host, "no equals sign\n", "A")[0] == 2
    assert lookup(host, "=5\n", "A")[0] == 2


def write_cfg(path, keys):
    path.write_text("".join(f"{k}={v}\n" for k, v in keys.items()))


def test_read_configuration_all_keys(host, tmp_path):
    p = tmp_path / "default.cfg"
    write_cfg(p, {**KEYS, "BoxLength": "107.49", "Precision": "0", "Unknown Key": "ignored"})
    buf = C.create_string_buffer(8192)
    assert host.mdxh_read_configuration(str(p).encode(), buf, 8192) == 0
    c = json.loads(buf.value.decode())
    assert c["LennardJonesParticles"] == 400 and c["ManyBodyParticles"] == 40
    assert c["useLennardJones"] == 1 and c["Fallback"] == 1 and c["animts"] == -1
    assert c["vxcmTo"] == pytest.approx(0.02) and c["Input"] == "in.txt" and c["AnimFile"] == "anim"
    assert c["BoxLength"] == pytest.approx(107.49) and c["Precision"] == 0
    assert c["SweepBatch"] == 8                    # the default: members advanced together


def test_sweep_batch_key(host, tmp_path):
    p = tmp_path / "default.cfg"
    write_cfg(p, {**KEYS, "SweepBatch": "2"})
    buf = C.create_string_buffer(8192)
    assert host.mdxh_read_configuration(str(p).encode(), buf, 8192) == 0
    assert json.loads(buf.value.decode())["SweepBatch"] == 2


def float32_sweep(a, b, step):
    """cpu/cpumd/Host.cpp:75: for (float vxcm = from; vxcm <= to; vxcm += step)"""
    a, b, step = np.float32(a), np.float32(b), np.float32(step)
    out, v = [], a
    while v <= b:
        out.append(float(v))
        if not step > 0:
            break
        v = np.float32(v + step)
    return out


@pytest.mark.parametrize("a,b,step", [(0, 0.003, 0.0015), (0, 0.3, 0.1), (-0.05, 0.05, 0.01), (0.2, 0.1, 0.05),
                                      (0.1, 0.1, 0.0), (0, 1, 0.07)])
def test_sweep_members_follow_the_float_loop(host, a, b, step):
    out = np.zeros(64, dtype=np.float32)
    k = host.mdxh_sweep_values(a, b, step, out.ctypes.data, len(out))
    want = float32_sweep(a, b, step)
    assert k == len(want)
    np.testing.assert_array_equal(out[:k], np.array(want, dtype=np.float32))


@pytest.mark.parametrize("drop", ["DT", "AnimationData", "Many Body Particles"])
def test_missing_key_is_an_error(host, tmp_path, drop):
    p = tmp_path / "bad.cfg"
    write_cfg(p, {k: v for k, v in KEYS.items() if k != drop})
    buf = C.create_string_buffer(512)
    assert host.mdxh_read_configuration(str(p).encode(), buf, 512) == 1
    assert "Unable to read configuration file" in buf.value.decode()


def test_boolean_accepts_only_0_and_1(host, tmp_path):
    p = tmp_path / "b.cfg"
    write_cfg(p, {**KEYS, "Debug": "true"})
    buf = C.create_string_buffer(512)
    assert host.mdxh_read_configuration(str(p).encode(), buf, 512) == 1


def test_output_then_input_round_trip(host, tmp_path):
    rng = np.random.default_rng(0)
    n = 5
    pos = np.c_[rng.normal(size=(n, 3)) * 10, rng.integers(0, 3, n)]
    vel = rng.normal(size=(n, 3)) * 1e-3
    f = tmp_path / "out.txt"
    assert host.mdxh_write_output(str(f).encode(), n, pos.ctypes.data, vel.ctypes.data) == 0
    text = f.read_text()
    assert "\n" not in text                       # the reference writes no newlines
    toks = text.split(" ")
    assert toks[-1] == "" and len(toks) == 7 * n + 1
    assert toks[0] == f"{pos[0, 0]:.6e}"          # scientific, precision 6
    p2, v2 = np.zeros((n, 4)), np.zeros((n, 3))
    assert host.mdxh_read_input(str(f).encode(), n, p2.ctypes.data, v2.ctypes.data) == 0
    np.testing.assert_allclose(p2, pos, rtol=1e-6)
    np.testing.assert_allclose(v2, vel, rtol=1e-6)
    # a file that ends right after the last number fails the reference's good() check
    g = tmp_path / "tight.txt"
    g.write_text(text.rstrip())
    assert host.mdxh_read_input(str(g).encode(), n, p2.ctypes.data, v2.ctypes.data) == 1


def test_results_files(host, tmp_path):
    rec = np.arange(33, dtype=np.float64).reshape(3, 11) * 0.5
    pos = np.array([[0, 0, 0, 0], [1.0, 0, 0, 1], [5, 5, 5, 1], [1.2, 1.2, 0, 0]], float)
    stem = str(tmp_path / "res")
    assert host.mdxh_write_results(stem.encode(), 4, 100, 1.0, 1, 7.6, 6.8, 0, 0, 0, 0.5, 12.0,
                                   rec.ctypes.data, 3, pos.ctypes.data, 4) == 0
    dat = (tmp_path / "res(vxcm-0.5).dat").read_bytes()
    nop, steps, dt = struct.unpack_from("<iif", dat, 0)
    assert (nop, steps, dt) == (4, 100, 1.0)
    off = 12
    assert dat[off] == 0xFF
    rs, rc = struct.unpack_from("<ff", dat, off + 1)
    assert rs == pytest.approx(7.6) and rc == pytest.approx(6.8)
    off += 9
    assert dat[off] == 0x00
    (vx,) = struct.unpack_from("<f", dat, off + 1)
    assert vx == 0.5
    body = np.frombuffer(dat[off + 5:], dtype="<f8").reshape(-1, 11)
    np.testing.assert_array_equal(body, rec)
    txt = (tmp_path / "res(vxcm-0.5).txt").read_text()
    assert txt.startswith("Number Of Particles = 4, Timesteps = 100\nDT = 1.000000, LennardJonesRS = 7.600000")
    assert "Time took: 12.000000 ms" in txt
    # unlike bond-order pairs closer than 2 reNO = 2.3: (0,1) and (1,3)
    assert txt.endswith("ManyBody Pairs:\nAmount: 2\n[0, 1] [1, 3] ")
