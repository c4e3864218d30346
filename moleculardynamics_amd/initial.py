"""Seeded initial conditions (lattices + Maxwell-Boltzmann velocities).

The reference's equilibrator seeds with the wall clock (cpu/equi/Host.cpp:83);
these generators are seeded so two engines can read identical inputs.  They
write the reference's text input format (common/Loader.cpp:125-155:
per atom "x y z type vx vy vz").
"""
from __future__ import annotations

import numpy as np

FCC_BASIS = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]])
DIAMOND_BASIS = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0],
                          [0.25, 0.25, 0.25], [0.25, 0.75, 0.75], [0.75, 0.25, 0.75], [0.75, 0.75, 0.25]])


def lattice(ncell: int, a: float, basis: np.ndarray) -> np.ndarray:
    g = np.stack(np.meshgrid(np.arange(ncell), np.arange(ncell), np.arange(ncell), indexing="ij"), -1).reshape(-1, 3)
    return ((g[:, None, :] + basis[None]) * a).reshape(-1, 3)


def fcc(ncell: int, density: float) -> tuple[np.ndarray, float]:
    """FCC block of 4 ncell^3 atoms at number density `density`; returns (pos, box edge)."""
    a = (4.0 / density) ** (1.0 / 3.0)
    return lattice(ncell, a, FCC_BASIS), a * ncell


def diamond(ncell: int, a: float) -> tuple[np.ndarray, float]:
    return lattice(ncell, a, DIAMOND_BASIS), a * ncell


def maxwell_boltzmann(n: int, kT_over_m: float, seed: int) -> np.ndarray:
    """Gaussian velocities with zero net momentum, rescaled so that
    sum v^2 / (3 n) equals kT/m exactly (single species)."""
    rng = np.random.default_rng(seed)
    v = rng.normal(0.0, np.sqrt(kT_over_m), size=(n, 3))
    v -= v.mean(axis=0)
    v *= np.sqrt(3 * n * kT_over_m / np.sum(v * v))
    return v


def lj_fluid(ncell: int, density: float, temperature: float, seed: int):
    """Reduced-unit LJ start: FCC at density, MB velocities at T* (m = kB = 1)."""
    pos, L = fcc(ncell, density)
    vel = maxwell_boltzmann(len(pos), temperature, seed)
    return pos, vel, L


def write_reference_input(path, pos3: np.ndarray, types: np.ndarray, vel3: np.ndarray) -> None:
    """common/Loader.cpp:125-155 reads x y z type vx vy vz per atom with >>."""
    with open(path, "w") as f:
        for p, t, v in zip(pos3, types, vel3):
            f.write(f"{p[0]:.17g} {p[1]:.17g} {p[2]:.17g} {int(t)} {v[0]:.17g} {v[1]:.17g} {v[2]:.17g}\n")


def read_reference_input(path, n: int):
    vals = np.array(open(path).read().split(), dtype=np.float64)[: 7 * n].reshape(n, 7)
    return vals[:, :4].copy(), vals[:, 4:7].copy()


# silicon in metal-style units: A, fs, amu, eV
SI_LATTICE = 5.431          # A
SI_MASS = 28.0855           # amu
KB_EV = 8.617333262e-5      # eV / K
ACCEL_EV_A_AMU = 9.6485332e-3   # 1 eV / (A amu) in A / fs^2: a = F / m * this


def si_diamond(ncell: int, temperature_k: float, seed: int, disp: float = 0.05):
    """Diamond silicon, 8 ncell^3 atoms: Gaussian displacements of `disp` A per
    axis and Maxwell-Boltzmann velocities (A/fs) at temperature_k, from one seed.
    Returns (positions wrapped into [0, L), velocities, L)."""
    pos, L = diamond(ncell, SI_LATTICE)
    rng = np.random.default_rng(seed)
    pos = np.mod(pos + rng.normal(0.0, disp, pos.shape), L)
    kT_over_m = KB_EV * temperature_k / SI_MASS * ACCEL_EV_A_AMU   # A^2 / fs^2
    vel = maxwell_boltzmann(len(pos), kT_over_m, seed + 1)
    return pos, vel, L


def replica_batch(n_replicas: int, seed: int, n_min: int = 256, n_max: int = 8192):
    """Replica sizes log-uniform in [n_min, n_max], each rounded to an FCC block
    of 4 k^3 atoms, density U[0.70, 0.95] and T* U[0.5, 2.0], all from one seed.
    Returns (n_atoms, box_lengths, positions, velocities) with the replicas'
    atoms one after another."""
    rng = np.random.default_rng(seed)
    ns, boxes, pos, vel = [], [], [], []
    for r in range(n_replicas):
        target = np.exp(rng.uniform(np.log(n_min), np.log(n_max)))
        k = max(4, int(round((target / 4) ** (1 / 3))))
        rho = rng.uniform(0.70, 0.95)
        temp = rng.uniform(0.5, 2.0)
        p, L = fcc(k, rho)
        v = maxwell_boltzmann(len(p), temp, int(rng.integers(1 << 30)))
        ns.append(len(p))
        boxes.append(L)
        pos.append(p)
        vel.append(v)
    return np.array(ns, dtype=np.int32), np.array(boxes), np.concatenate(pos), np.concatenate(vel)


# the reference's walled N/O/Ne gas: species codes (common/Constants.h:13-15),
# masses in amu (common/Constants.h:36-38) and kB in its energy units
# (common/DebugCalculations.cpp:122)
SPECIES_N, SPECIES_O, SPECIES_NE = 0, 1, 2
REF_MASS = {SPECIES_N: 14.00674, SPECIES_O: 15.9994, SPECIES_NE: 20.1797}
REF_KB = 8.314e-7


def walled_mixture(n_mol: int, n_ne: int, temperature_k: float, seed: int, spacing: float = 3.3,
                   bond: float = 1.14):
    """n_mol diatomic molecules (N2, O2 or NO, bond length `bond` A) and n_ne Ne
    atoms on a jittered cubic grid of `spacing` A around the origin, with
    Maxwell-Boltzmann velocities at temperature_k (reference units), all from
    one seed.  The molecules' atoms come first, as the reference's bond-order
    atoms must (cpu/cpumd/Host.cpp:120).  Returns (pos4 with the species in
    column 3, vel3)."""
    rng = np.random.default_rng(seed)
    sites = n_mol + n_ne
    k = int(np.ceil(sites ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(k)] * 3, indexing="ij"), -1).reshape(-1, 3)[:sites].astype(float)
    g = (g - (k - 1) / 2) * spacing + rng.uniform(-0.1, 0.1, g.shape) * spacing
    rng.shuffle(g)
    kinds = np.array([(SPECIES_N, SPECIES_N), (SPECIES_O, SPECIES_O), (SPECIES_N, SPECIES_O)])
    u = rng.normal(size=(n_mol, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    mk = kinds[rng.integers(0, 3, n_mol)]
    a = np.c_[g[:n_mol] + 0.5 * bond * u, mk[:, 0]]
    b = np.c_[g[:n_mol] - 0.5 * bond * u, mk[:, 1]]
    mol = np.stack([a, b], 1).reshape(-1, 4)
    ne = np.c_[g[n_mol:], np.full(n_ne, SPECIES_NE)]
    pos4 = np.concatenate([mol, ne]).astype(np.float64)
    m = np.array([REF_MASS[int(t)] for t in pos4[:, 3]])
    vel = rng.normal(size=(len(pos4), 3)) * np.sqrt(REF_KB * temperature_k / m)[:, None]
    return pos4, vel
