"""Pins for the oracle of the learned Tersoff surrogate (oracle/surrogate.py).

The reference has no surrogate; these checks fix the oracle independently of
its own formulas: central finite differences (descriptor Jacobian, energy vs
forces, loss vs weight gradient), closed forms of an ideal tetrahedron,
rotation / translation / permutation invariance, torch autograd for the
double back-propagation of the force-matching gradient, torch.bfloat16 for
the rounding emulation, and the per-centre Tersoff split that the fallback
uses (all centres == the whole Tersoff sum)."""
import numpy as np
import pytest
import torch

import _oracle as O
from moleculardynamics_amd import initial
from oracle import surrogate as S


def si64(seed=3, disp=0.08):
    pos, _, L = initial.si_diamond(2, 300.0, seed, disp=disp)
    return pos, L


def random_params(G, seed=0, scale=0.4):
    rng = np.random.default_rng(seed)
    p = {
        "W1": rng.normal(0, scale, (S.NHID, S.NFEAT)), "b1": rng.normal(0, 0.1, S.NHID),
        "W2": rng.normal(0, scale / 2, (S.NHID, S.NHID)), "b2": rng.normal(0, 0.1, S.NHID),
        "w3": rng.normal(0, 0.3, S.NHID), "b3": np.array([-4.6]),
        "mean": G.mean(0), "istd": 1.0 / (G.std(0) + 1e-3),
        "lo": np.full(S.NFEAT, -1e30), "hi": np.full(S.NFEAT, 1e30),
    }
    return p


def test_pack_roundtrip_and_sizes():
    G = np.random.default_rng(1).normal(size=(10, S.NFEAT))
    p = random_params(G)
    blob = S.pack(p)
    assert blob.shape == (S.NBLOB,) and S.NTRAIN == 64 * 48 + 64 + 64 * 64 + 64 + 64 + 1
    q = S.unpack(blob)
    for k in p:
        np.testing.assert_array_equal(q[k], p[k])


def test_bf16_rounding_matches_torch():
    x = np.random.default_rng(2).normal(size=4096).astype(np.float32) * 10.0 ** np.random.default_rng(3).integers(-6, 6, 4096)
    x[:4] = [1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -(1.0 + 2 ** -8), 0.0]      # ties go to even
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(S.bf16(x), ref)


def test_fcut_matches_tersoff_cutoff():
    r = np.linspace(2.0, 3.2, 121)
    fc, dfc = S.fcut(r)
    p = O.tersoff_params()
    assert (p.R, p.D) == (2.7, 0.3)
    x = np.pi / 2 * (r - 2.7) / 0.3
    exp = np.where(r < 2.4, 1, np.where(r > 3.0, 0, 0.5 - 0.5 * np.sin(x)))
    np.testing.assert_allclose(fc, exp, atol=1e-15)
    num = (S.fcut(r + 1e-7)[0] - S.fcut(r - 1e-7)[0]) / 2e-7
    np.testing.assert_allclose(dfc, num, atol=1e-6)


def test_tetrahedron_closed_form():
    r0 = 2.35
    dirs = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]) / np.sqrt(3)
    dv = (dirs * r0)[None]
    G, _ = S.descriptors(dv, np.ones((1, 4), bool))
    rad = 4 * np.exp(-S.ETA * (r0 - S.MU) ** 2)
    np.testing.assert_allclose(G[0, :32], rad, rtol=1e-13)
    c = -1.0 / 3.0
    A0, A1, A2, A3 = 3.0, 3 * c, 3 * (2 * c * c - 1), 3 * (4 * c ** 3 - 3 * c)
    mono = np.array([A1, A2, A3, A1 * A1, A1 * A2, A2 * A2, A0 * A1, A1 ** 3])
    for m in range(2):
        psi = np.exp(-S.ETA_B * (r0 - S.NU[m]) ** 2)
        np.testing.assert_allclose(G[0, 32 + 8 * m: 40 + 8 * m], 4 * psi * mono, rtol=1e-12)


def test_descriptor_jacobian_finite_differences():
    pos, L = si64()
    G, J, idx, cnt = S.frame_descriptors(pos, L)
    assert cnt.max() <= S.SLOTS
    h = 1e-6
    for a in (0, 17, 40):
        for q in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[a, q] += h
            p2[a, q] -= h
            num = (S.frame_descriptors(p1, L)[0] - S.frame_descriptors(p2, L)[0]) / (2 * h)
            an = np.zeros_like(G)
            an[a] -= J[a, :, :, q].sum(-1)
            for i, s in zip(*np.nonzero(idx == a)):
                an[i] += J[i, :, s, q]
            assert np.abs(num - an).max() < 1e-6 * max(1.0, np.abs(num).max())


def test_descriptors_invariant_under_rotation_translation_permutation():
    pos, L = si64(disp=0.1)
    idx, dv, cnt = S.neighbours(pos, L)
    G, _ = S.descriptors(dv, idx >= 0)
    rng = np.random.default_rng(4)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    perm = rng.permutation(dv.shape[1])
    dv2 = (dv @ Q.T)[:, perm]
    G2, _ = S.descriptors(dv2, (idx >= 0)[:, perm])
    np.testing.assert_allclose(G2, G, rtol=1e-11, atol=1e-12)
    shift = rng.normal(size=3) * 3.0
    G3, _, _, _ = S.frame_descriptors(np.mod(pos + shift, L), L)
    np.testing.assert_allclose(G3, G, rtol=1e-11, atol=1e-12)


def test_energy_forces_finite_differences_and_momentum():
    pos, L = si64()
    G = S.frame_descriptors(pos, L)[0]
    p = random_params(G)
    E, F, e, fl = S.energy_forces(p, pos, L)
    assert not fl.any()
    assert abs(e.sum() - E) < 1e-9
    np.testing.assert_allclose(F.sum(0), 0, atol=1e-9)
    h = 1e-5
    for a in (3, 31):
        for q in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[a, q] += h
            p2[a, q] -= h
            num = -(S.energy_forces(p, p1, L)[0] - S.energy_forces(p, p2, L)[0]) / (2 * h)
            assert abs(num - F[a, q]) < 1e-6 * max(1.0, abs(num))


def test_mlp_gradient_matches_torch_autograd():
    rng = np.random.default_rng(5)
    Gs = rng.normal(size=(32, S.NFEAT))
    p = random_params(Gs)
    E, h = S.mlp(p, Gs)
    t = {k: torch.tensor(v) for k, v in p.items()}
    x = torch.tensor(Gs, requires_grad=True)
    Et = torch.tanh(torch.tanh(x @ t["W1"].T + t["b1"]) @ t["W2"].T + t["b2"]) @ t["w3"] + t["b3"][0]
    (ht,) = torch.autograd.grad(Et.sum(), x)
    np.testing.assert_allclose(E, Et.detach().numpy(), rtol=1e-13)
    np.testing.assert_allclose(h, ht.numpy(), rtol=1e-11, atol=1e-13)


def test_bf16_emulation_is_close_to_fp64():
    rng = np.random.default_rng(6)
    Gs = rng.normal(size=(64, S.NFEAT))
    p = random_params(Gs)
    E, h = S.mlp(p, Gs)
    Eb, hb = S.mlp(p, Gs, emulate_bf16=True)
    assert np.abs(Eb - E).max() < 0.05
    assert np.sqrt(((hb - h) ** 2).sum() / (h ** 2).sum()) < 0.02


def _frames(n_frames=2, seed=7):
    out = []
    for k in range(n_frames):
        pos, L = si64(seed=seed + k, disp=0.09)
        G, J, idx, _ = S.frame_descriptors(pos, L)
        e, f = O.tersoff(pos, L)
        out.append((G, J, idx, f, e))
    return out


def _torch_loss(p, frames, w_energy):
    """The same loss built from torch autograd (create_graph for the forces)."""
    t = {k: torch.tensor(v, requires_grad=k in ("W1", "b1", "W2", "b2", "w3", "b3")) for k, v in p.items()}
    ncomp = sum(3 * len(f[0]) for f in frames)
    loss = 0
    for G, J, idx, Fref, Eref in frames:
        n = len(G)
        Gs = ((torch.tensor(G) - t["mean"]) * t["istd"]).requires_grad_(True)
        E = torch.tanh(torch.tanh(Gs @ t["W1"].T + t["b1"]) @ t["W2"].T + t["b2"]) @ t["w3"] + t["b3"][0]
        (hs,) = torch.autograd.grad(E.sum(), Gs, create_graph=True)
        w = torch.einsum("nf,nfsd->nsd", hs * t["istd"], torch.tensor(J))
        F = w.sum(1)
        valid = torch.tensor(idx >= 0)
        F = F.index_add(0, torch.tensor(idx)[valid], -w[valid])
        loss = loss + ((F - torch.tensor(Fref)) ** 2).sum() / ncomp + w_energy * ((E.sum() - Eref) / n) ** 2 / len(frames)
    return loss, t


def test_train_gradient_matches_torch_double_backprop():
    frames = _frames()
    p = random_params(np.concatenate([f[0] for f in frames]))
    loss, g = S.train_loss_gradient(p, frames, w_energy=0.7)
    lt, t = _torch_loss(p, frames, 0.7)
    lt.backward()
    assert abs(loss - float(lt.detach())) < 1e-10 * abs(loss)
    for k in g:
        ref = t[k].grad.numpy()
        assert np.abs(g[k] - ref).max() < 1e-9 * max(1.0, np.abs(ref).max()), k


def test_train_gradient_finite_differences():
    frames = _frames(1, seed=11)
    p = random_params(frames[0][0], seed=2)
    _, g = S.train_loss_gradient(p, frames, w_energy=1.0)
    rng = np.random.default_rng(9)
    for k in ("W1", "W2", "b1", "w3", "b3"):
        for _ in range(2):
            ix = tuple(rng.integers(0, d) for d in p[k].shape)
            h = 1e-6
            p1 = {a: b.copy() for a, b in p.items()}
            p2 = {a: b.copy() for a, b in p.items()}
            p1[k][ix] += h
            p2[k][ix] -= h
            num = (S.train_loss_gradient(p1, frames)[0] - S.train_loss_gradient(p2, frames)[0]) / (2 * h)
            assert abs(num - g[k][ix]) < 1e-5 * max(1.0, abs(num)), (k, ix)


def test_tersoff_centres_split():
    pos, L = si64(seed=12)
    e_all, f_all = O.tersoff(pos, L)
    mask = np.zeros(len(pos), np.uint8)
    mask[::3] = 1
    e1, f1, ea1 = O.tersoff_centres(pos, L, mask)
    e2, f2, ea2 = O.tersoff_centres(pos, L, 1 - mask)
    assert abs(e1 + e2 - e_all) < 1e-9 * abs(e_all)
    np.testing.assert_allclose(f1 + f2, f_all, atol=1e-10)
    np.testing.assert_allclose(ea1.sum(), e1, rtol=1e-12)
    assert np.all(ea1[mask == 0] == 0)
    h = 1e-5
    for a, q in ((0, 0), (5, 2)):
        p1, p2 = pos.copy(), pos.copy()
        p1[a, q] += h
        p2[a, q] -= h
        num = -(O.tersoff_centres(p1, L, mask)[0] - O.tersoff_centres(p2, L, mask)[0]) / (2 * h)
        assert abs(num - f1[a, q]) < 1e-6 * max(1, abs(num))


def test_fallback_for_all_atoms_is_plain_tersoff():
    pos, L = si64(seed=13)
    G = S.frame_descriptors(pos, L)[0]
    p = random_params(G)
    p["hi"] = np.full(S.NFEAT, -1e30)          # every atom out of range

    def fallback(mask):
        _, f, e_atom = O.tersoff_centres(pos, L, mask.astype(np.uint8))
        return e_atom, f

    E, F, e, fl = S.energy_forces(p, pos, L, tersoff_fallback=fallback)
    e_t, f_t = O.tersoff(pos, L)
    assert fl.all()
    assert abs(E - e_t) < 1e-9 * abs(e_t)
    np.testing.assert_allclose(F, f_t, atol=1e-10)
