"""FP64 restatement of the learned Tersoff surrogate (numpy).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module as the checker of the HIP surrogate kernels
in moleculardynamics_amd/csrc/surrogate.hip.  The product never imports it.

The reference has no surrogate.  What it replaces is the bond-order force
(manyBodyForces, cpu/common/BondOrderFuncs.h:283-397; here the Tersoff-Si form
of moleculardynamics_amd/csrc/tersoff.hip).  The model:

  E = sum_i NN(Gs_i),   Gs_i = (G_i - mean) * istd
  NN(x) = w3 . tanh(W2 tanh(W1 x + b1) + b2) + b3          (48 -> 64 -> 64 -> 1)

Descriptor G_i (48 values) of atom i over its neighbours j with r_ij < 3.0 A,
fc the Tersoff cut-off (1 below 2.4 A, 1/2 - 1/2 sin(pi/2 (r - 2.7)/0.3) to 3.0):
  radial   G[k]        = sum_j exp(-eta (r_ij - mu_k)^2) fc(r_ij),
                          mu_k = 2.0 + k/31 (k < 32), eta = 1/2 (31/1.5)^2
  angular  G[32+8m+q]  = sum_j psi_m(r_ij) mono_q(A_j),
                          psi_m(r) = exp(-(r - nu_m)^2 / (2 * 0.2^2)) fc(r), nu = 2.25, 2.55
    with the bond-resolved angular moments of bond j
      A_j0 = sum_{k != j} fc(r_ik),  A_jl = sum_{k != j} fc(r_ik) T_l(cos theta_jik)  (l = 1..3,
      Chebyshev T_l) and mono = (A1, A2, A3, A1^2, A1 A2, A2^2, A0 A1, A1^3).

Forces come from the per-slot gradients w_is = dE_i/d(x_j - x_i):
  F_a = sum_s w_{a s} - sum_{(i, s): j(i, s) = a} w_{i s}.

Atoms flagged as out of range (any Gs outside [lo, hi], or more than SLOTS
neighbours inside 3.0 A) take their energy from the Tersoff per-atom term
e_i = 1/2 sum_j V_ij instead (plain-Tersoff fallback).

What pins each function (tests/test_surrogate_oracle.py):
  descriptors/Jacobian  central finite differences; closed forms of an ideal
                        tetrahedron; rotation, translation and permutation invariance
  mlp / energy_forces   finite differences of the total energy; torch autograd
  train_gradient        torch autograd of the same loss (an independent
                        implementation of the double back-propagation); finite
                        differences of the loss in the weights
  bf16 emulation        round-to-nearest-even checked against torch.bfloat16
"""
from __future__ import annotations

import numpy as np

RC = 3.0
RIN = 2.4
NRAD = 32
NFEAT = 48
NHID = 64
SLOTS = 6                       # neighbours inside RC handled by the network
MU = 2.0 + np.arange(NRAD) / 31.0
ETA = 0.5 * (31.0 / 1.5) ** 2
NU = np.array([2.25, 2.55])
ETA_B = 0.5 / 0.2 ** 2

# flat parameter blob, float32, in this order (shared with the C-ABI)
LAYOUT = [("W1", (NHID, NFEAT)), ("b1", (NHID,)), ("W2", (NHID, NHID)), ("b2", (NHID,)),
          ("w3", (NHID,)), ("b3", (1,)),
          ("mean", (NFEAT,)), ("istd", (NFEAT,)), ("lo", (NFEAT,)), ("hi", (NFEAT,))]
NTRAIN = 7361                   # W1 .. b3
NBLOB = NTRAIN + 4 * NFEAT


def unpack(blob):
    blob = np.asarray(blob, dtype=np.float64)
    assert blob.shape == (NBLOB,)
    out, o = {}, 0
    for name, shape in LAYOUT:
        k = int(np.prod(shape))
        out[name] = blob[o:o + k].reshape(shape)
        o += k
    return out


def pack(p) -> np.ndarray:
    return np.concatenate([np.asarray(p[name], dtype=np.float64).reshape(-1) for name, _ in LAYOUT])


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest, ties to even) and
    return them as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


# ---------------------------------------------------------------- geometry

def fcut(r):
    r = np.asarray(r, dtype=np.float64)
    x = np.clip((r - 2.7) / 0.3, -1.0, 1.0) * (np.pi / 2)
    fc = np.where(r <= RIN, 1.0, np.where(r >= RC, 0.0, 0.5 - 0.5 * np.sin(x)))
    dfc = np.where((r > RIN) & (r < RC), -0.25 * np.pi / 0.3 * np.cos(x), 0.0)
    return fc, dfc


def neighbours(pos, L, rc=RC, slots=None):
    """Brute-force minimum-image neighbours within rc, ascending index order.
    Returns idx [n, S] (-1 padded), dv [n, S, 3] (x_j - x_i), count [n]."""
    pos = np.asarray(pos, dtype=np.float64)
    n = len(pos)
    d = pos[None, :, :] - pos[:, None, :]
    if L > 0:
        d -= L * np.round(d / L)
    r2 = (d ** 2).sum(-1)
    np.fill_diagonal(r2, np.inf)
    hit = r2 < rc * rc
    cnt = hit.sum(1)
    S = int(cnt.max()) if slots is None else slots
    S = max(S, 1)
    idx = -np.ones((n, S), dtype=np.int64)
    dv = np.zeros((n, S, 3))
    for i in range(n):
        js = np.nonzero(hit[i])[0][:S]
        idx[i, :len(js)] = js
        dv[i, :len(js)] = d[i, js]
    return idx, dv, cnt


def _cheb(c):
    T = np.stack([np.ones_like(c), c, 2 * c * c - 1, 4 * c ** 3 - 3 * c])
    dT = np.stack([np.zeros_like(c), np.ones_like(c), 4 * c, 12 * c * c - 3])
    return T, dT


def descriptors(dv, valid):
    """G [n, 48] and its Jacobian J [n, 48, S, 3] with respect to the slot
    displacements dv [n, S, 3] (entries with valid False are ignored)."""
    dv = np.asarray(dv, dtype=np.float64)
    n, S, _ = dv.shape
    r = np.sqrt((dv ** 2).sum(-1))
    r = np.where(valid, r, 10.0)
    u = np.where(valid[..., None], dv / r[..., None], 0.0)
    fc, dfc = fcut(r)
    fc = fc * valid
    dfc = dfc * valid
    G = np.zeros((n, NFEAT))
    J = np.zeros((n, NFEAT, S, 3))
    # radial
    e = np.exp(-ETA * (r[..., None] - MU) ** 2)                          # n,S,32
    phi = e * fc[..., None]
    dphi = e * (dfc[..., None] - 2 * ETA * (r[..., None] - MU) * fc[..., None])
    G[:, :NRAD] = phi.sum(1)
    J[:, :NRAD] = np.einsum("nsk,nsd->nksd", dphi, u)
    # bond-resolved angular moments
    off = 1.0 - np.eye(S)[None]
    c = np.einsum("nsd,ntd->nst", u, u)
    T, dT = _cheb(c)                                                     # 4,n,S,S
    A = np.einsum("lnst,nt->nsl", T * off[None], fc)                     # n,S,4 (s: bond, t: third)
    # dA_sl/d dv_q: q = s via c; q = t via fc(r_t) and c
    dcs_s = (u[:, None, :, :] - c[..., None] * u[:, :, None, :]) / r[:, :, None, None]   # d c_st / d dv_s  n,s,t,3
    dcs_t = (u[:, :, None, :] - c[..., None] * u[:, None, :, :]) / r[:, None, :, None]   # d c_st / d dv_t  n,s,t,3
    dA = np.zeros((n, S, 4, S, 3))                                       # n, s, l, q, 3
    for l in range(4):
        wt = (dT[l] * off * fc[:, None, :])[..., None]                   # n,s,t,1
        ds = (wt * dcs_s).sum(2)                                         # n,s,3  (q = s)
        for s in range(S):
            dA[:, s, l, s, :] += ds[:, s]
        dq = wt * dcs_t + (T[l] * off * dfc[:, None, :])[..., None] * u[:, None, :, :]   # n,s,t,3 (q = t)
        dA[:, :, l, :, :] += dq
    A0, A1, A2, A3 = A[..., 0], A[..., 1], A[..., 2], A[..., 3]
    mono = np.stack([A1, A2, A3, A1 * A1, A1 * A2, A2 * A2, A0 * A1, A1 ** 3], -1)          # n,S,8
    z = np.zeros_like(A1)
    o = np.ones_like(A1)
    dmono = np.stack([                                                     # n,S,8,4 (d mono / d A_l)
        np.stack([z, o, z, z], -1), np.stack([z, z, o, z], -1), np.stack([z, z, z, o], -1),
        np.stack([z, 2 * A1, z, z], -1), np.stack([z, A2, A1, z], -1), np.stack([z, z, 2 * A2, z], -1),
        np.stack([A1, A0, z, z], -1), np.stack([z, 3 * A1 * A1, z, z], -1)], -2)
    for m in range(2):
        g = np.exp(-ETA_B * (r - NU[m]) ** 2)
        psi = g * fc
        dpsi = g * (dfc - 2 * ETA_B * (r - NU[m]) * fc)
        G[:, NRAD + 8 * m: NRAD + 8 * m + 8] = np.einsum("ns,nsq->nq", psi, mono)
        Jm = np.einsum("ns,nsq,nsd,sp->nqpd", dpsi, mono, u, np.eye(S))
        Jm += np.einsum("ns,nsqa,nsapd->nqpd", psi, dmono, dA)
        J[:, NRAD + 8 * m: NRAD + 8 * m + 8] = Jm
    return G, J


def frame_descriptors(pos, L, slots=SLOTS):
    """Descriptors of a frame: (G, J, idx, count) with J over `slots` slots."""
    idx, dv, cnt = neighbours(pos, L, RC, slots=max(slots, 1))
    valid = idx >= 0
    G, J = descriptors(dv, valid)
    return G, J, idx, cnt


# ---------------------------------------------------------------- network

def mlp(p, Gs, emulate_bf16=False):
    """Per-atom energy and h = dE/dGs for standardized descriptors Gs [n, 48].
    With emulate_bf16 the inputs, weights, the first activations and both
    back-propagated vectors are rounded to bfloat16 where the MFMA kernel
    rounds them (float32 accumulation, tanh in float32)."""
    W1, W2 = p["W1"], p["W2"]
    if emulate_bf16:
        f32 = lambda x: np.asarray(x, dtype=np.float32).astype(np.float64)
        W1b, W2b = bf16(W1), bf16(W2)
        x = bf16(Gs)
        z1 = f32(x @ W1b.T + p["b1"])
        a1 = f32(np.tanh(z1))
        z2 = f32(bf16(a1) @ W2b.T + p["b2"])
        a2 = f32(np.tanh(z2))
        E = a2 @ p["w3"] + p["b3"][0]
        g2 = bf16(p["w3"] * (1 - a2 * a2))
        q = f32(g2 @ W2b)
        g1 = bf16(q * (1 - a1 * a1))
        h = f32(g1 @ W1b)
        return E, h
    z1 = Gs @ W1.T + p["b1"]
    a1 = np.tanh(z1)
    z2 = a1 @ W2.T + p["b2"]
    a2 = np.tanh(z2)
    E = a2 @ p["w3"] + p["b3"][0]
    g2 = p["w3"] * (1 - a2 * a2)
    g1 = (g2 @ W2) * (1 - a1 * a1)
    return E, g1 @ W1


def standardize(p, G):
    return (G - p["mean"]) * p["istd"]


def flags(p, Gs, count):
    """Atoms the network does not cover: a standardized descriptor outside
    [lo, hi] or more than SLOTS neighbours inside RC."""
    return np.any((Gs < p["lo"]) | (Gs > p["hi"]), axis=1) | (count > SLOTS)


def gather_forces(w, idx):
    """F_a = sum_s w_as - sum over the slots (i, s) that point at a of w_is."""
    n = len(w)
    F = w.sum(1).copy()
    valid = idx >= 0
    np.add.at(F, idx[valid], -w[valid])
    return F


def energy_forces(p, pos, L, emulate_bf16=False, tersoff_fallback=None):
    """Surrogate energy (eV), forces (eV/A), per-atom energies and the
    fallback flags of one configuration.  tersoff_fallback(mask) -> (e_atoms,
    F) supplies the plain-Tersoff part for the flagged atoms (their per-atom
    energies and the forces of those energies); without it flagged atoms raise."""
    pos = np.asarray(pos, dtype=np.float64)
    idx, dv, cnt = neighbours(pos, L, RC, slots=None)
    S = idx.shape[1]
    valid = idx >= 0
    G, J = descriptors(dv, valid)
    Gs = standardize(p, G)
    fl = flags(p, Gs, cnt)
    e, h = mlp(p, Gs, emulate_bf16)
    h = h * p["istd"]                     # dE/dG
    h[fl] = 0.0
    e = np.where(fl, 0.0, e)
    w = np.einsum("nf,nfsd->nsd", h, J)
    F = gather_forces(w, idx)
    if fl.any():
        if tersoff_fallback is None:
            raise ValueError("atoms outside the network's range and no fallback given")
        et, Ft = tersoff_fallback(fl)
        e = np.where(fl, et, e)
        F = F + Ft
    return float(e.sum()), F, e, fl


# ---------------------------------------------------------------- training

def train_loss_gradient(p, frames, w_energy=1.0):
    """Loss and its gradient in the trainable weights (dict of arrays).

    frames: list of (G, J, idx, F_ref, E_ref) per configuration, J over the
    slots of idx.  Loss = mean over all force components of (F - F_ref)^2
    + w_energy * mean over frames of ((E - E_ref) / n)^2.

    The force term's gradient uses the directional derivative of dE/dGs along
    u_i = istd * dL/d(dE_i/dG_i) (double back-propagation written out)."""
    W1, b1, W2, b2, w3, b3 = (p[k] for k in ("W1", "b1", "W2", "b2", "w3", "b3"))
    ncomp = sum(3 * len(f[0]) for f in frames)
    nfr = len(frames)
    loss = 0.0
    grads = {k: np.zeros_like(p[k]) for k in ("W1", "b1", "W2", "b2", "w3", "b3")}
    for G, J, idx, Fref, Eref in frames:
        n = len(G)
        Gs = standardize(p, G)
        z1 = Gs @ W1.T + b1
        a1 = np.tanh(z1)
        z2 = a1 @ W2.T + b2
        a2 = np.tanh(z2)
        E = a2 @ w3 + b3[0]
        g2 = w3 * (1 - a2 * a2)
        q = g2 @ W2
        g1 = q * (1 - a1 * a1)
        hs = g1 @ W1
        w = np.einsum("nf,nfsd->nsd", hs * p["istd"], J)
        F = gather_forces(w, idx)
        rF = F - Fref
        Etot = E.sum()
        loss += (rF ** 2).sum() / ncomp + w_energy * ((Etot - Eref) / n) ** 2 / nfr
        rho = 2 * rF / ncomp                                  # dL/dF
        eps = 2 * w_energy * (Etot - Eref) / (n * n) / nfr    # dL/dE_i (every atom of the frame)
        # dL/dw_is = rho_i - rho_j(i,s); dL/dh_raw = sum_s J_is (rho_i - rho_j)
        valid = idx >= 0
        rj = np.where(valid[..., None], rho[np.where(valid, idx, 0)], 0.0)
        dw = np.where(valid[..., None], rho[:, None, :] - rj, 0.0)
        u = np.einsum("nfsd,nsd->nf", J, dw) * p["istd"]     # tangent in standardized space
        # tangent pass: t1 = W1 u, t1a = (1 - a1^2) t1, t2 = W2 t1a
        t1 = u @ W1.T
        d1 = 1 - a1 * a1
        d2 = 1 - a2 * a2
        t1a = d1 * t1
        t2 = t1a @ W2.T
        # s = w3 . (d2 * t2); back-propagate s and the energy term
        c2 = -2 * a2 * d2 * w3 * t2 + eps * g2                # d/dz2
        qq = g2 @ W2                                          # = q
        c1 = (c2 @ W2 - 2 * a1 * qq * t1) * d1                # d/dz1 (the -2 a1 q t1 term: via d1 in t1a)
        g1t = qq * d1                                         # d s / d t1
        grads["w3"] += (d2 * t2).sum(0) + eps * a2.sum(0)
        grads["b3"] += np.array([eps * n])
        grads["W2"] += g2.T @ t1a + c2.T @ a1
        grads["b2"] += c2.sum(0)
        grads["W1"] += g1t.T @ u + c1.T @ Gs
        grads["b1"] += c1.sum(0)
    return loss, grads
