"""Seeded synthetic globules for the GB-integral workloads.

The workloads are specified as a self-avoiding walk (bond 0.153 nm, non-bonded
separation >= 0.28 nm) compacted to 100 atoms/nm^3.  Those constraints cannot hold
together: 0.28 nm hard spheres reach at most ~64 atoms/nm^3 even at close packing.
The generator therefore keeps the density and the chemistry-like distributions and
relaxes the separation:

* N points uniform in the sphere of radius R_s = (3N / (4 pi 100 nm^-3))^(1/3);
* soft-sphere relaxation until every pair is >= ``dmin`` = 0.20 nm apart (or the
  iteration budget ends; the achieved minimum is reported);
* atoms numbered along a serpentine path through the globule (slab / row / column),
  so consecutive indices are spatial neighbours as along a chain;
* charges U[-0.8, 0.8] e shifted to zero net charge;
* radii drawn uniformly from {0.100, 0.170, 0.185, 0.200} nm.

Everything is a pure function of (n_atoms, seed).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

DENSITY = 100.0  # atoms / nm^3
RADII = np.array([0.100, 0.170, 0.185, 0.200])


@dataclass
class System:
    pos: np.ndarray      # (N, 3) nm
    charge: np.ndarray   # (N,) e
    radius: np.ndarray   # (N,) nm
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return self.pos.shape[0]


def _relax(pos: np.ndarray, rs: float, dmin: float, iters: int) -> np.ndarray:
    from scipy.spatial import cKDTree

    for _ in range(iters):
        tree = cKDTree(pos)
        pairs = tree.query_pairs(dmin, output_type="ndarray")
        if len(pairs) == 0:
            break
        d = pos[pairs[:, 1]] - pos[pairs[:, 0]]
        r = np.linalg.norm(d, axis=1)
        r = np.maximum(r, 1e-6)
        push = (0.55 * (dmin - r) / r)[:, None] * d
        disp = np.zeros_like(pos)
        np.add.at(disp, pairs[:, 0], -push)
        np.add.at(disp, pairs[:, 1], push)
        pos = pos + disp
        rr = np.linalg.norm(pos, axis=1)
        out = rr > rs
        pos[out] *= (rs / rr[out])[:, None]
    return pos


def _serpentine_order(pos: np.ndarray, slab: float = 0.45) -> np.ndarray:
    ix = np.floor((pos[:, 0] - pos[:, 0].min()) / slab).astype(np.int64)
    iy = np.floor((pos[:, 1] - pos[:, 1].min()) / slab).astype(np.int64)
    z = pos[:, 2]
    # alternate direction of rows and columns so the path stays connected
    y_key = np.where(ix % 2 == 0, iy, -iy)
    z_key = np.where((ix + iy) % 2 == 0, z, -z)
    return np.lexsort((z_key, y_key, ix))


def globule(n_atoms: int, seed: int, dmin: float = 0.20, iters: int = 200,
            center: tuple = (0.0, 0.0, 0.0)) -> System:
    """A compact synthetic globule of ``n_atoms`` atoms at 100 atoms/nm^3."""
    if n_atoms < 1:
        raise ValueError("n_atoms must be positive")
    rng = np.random.default_rng(seed)
    rs = (3.0 * n_atoms / (4.0 * np.pi * DENSITY)) ** (1.0 / 3.0)
    dirs = rng.normal(size=(n_atoms, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    pos = dirs * (rs * rng.random(n_atoms) ** (1.0 / 3.0))[:, None]
    if n_atoms > 1:
        pos = _relax(pos, rs, dmin, iters)
        pos = pos[_serpentine_order(pos)]
    charge = rng.uniform(-0.8, 0.8, n_atoms)
    charge -= charge.mean()
    radius = RADII[rng.integers(0, len(RADII), n_atoms)]
    pos = pos + np.asarray(center, dtype=np.float64)
    if n_atoms > 1:
        from scipy.spatial import cKDTree

        dmin_achieved = float(cKDTree(pos).query(pos, k=2)[0][:, 1].min())
    else:
        dmin_achieved = float("inf")
    meta = {"n_atoms": n_atoms, "seed": seed, "sphere_radius_nm": rs,
            "density_per_nm3": DENSITY, "dmin_target_nm": dmin, "dmin_achieved_nm": dmin_achieved}
    return System(np.ascontiguousarray(pos), charge, radius.astype(np.float64), meta)


def liquid(box, seed: int, density: float = None, dmin: float = 0.20, iters: int = 200) -> System:
    """Atoms filling a rectangular periodic box (edges ``box``, nm) at ``density`` atoms/nm^3
    (default: the globule density), pushed apart to ``dmin`` under minimum-image distances
    and wrapped into [0, box); serpentine order, charges and radii drawn as for ``globule``."""
    from scipy.spatial import cKDTree

    L = np.asarray(box, dtype=np.float64)
    rng = np.random.default_rng(seed)
    n = int(round((density or DENSITY) * float(np.prod(L))))
    pos = rng.uniform(0.0, 1.0, (n, 3)) * L
    for _ in range(iters):
        pairs = cKDTree(pos, boxsize=L).query_pairs(dmin, output_type="ndarray")
        if len(pairs) == 0:
            break
        d = pos[pairs[:, 1]] - pos[pairs[:, 0]]
        d -= L * np.round(d / L)
        r = np.maximum(np.linalg.norm(d, axis=1), 1e-6)
        push = (0.55 * (dmin - r) / r)[:, None] * d
        disp = np.zeros_like(pos)
        np.add.at(disp, pairs[:, 0], -push)
        np.add.at(disp, pairs[:, 1], push)
        pos = np.mod(pos + disp, L)
    pos = pos[_serpentine_order(pos)]
    charge = rng.uniform(-0.8, 0.8, n)
    charge -= charge.mean()
    radius = RADII[rng.integers(0, len(RADII), n)]
    meta = {"n_atoms": n, "seed": seed, "box_nm": L.tolist(), "density_per_nm3": density or DENSITY,
            "dmin_target_nm": dmin}
    return System(np.ascontiguousarray(pos), charge, radius.astype(np.float64), meta)


def conformations(base: System, count: int, seed: int, sigma: float = 0.02) -> list:
    """``count`` perturbed copies of ``base``: x + N(0, sigma^2) per coordinate."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        pos = base.pos + rng.normal(scale=sigma, size=base.pos.shape)
        out.append(System(np.ascontiguousarray(pos), base.charge.copy(), base.radius.copy(),
                          dict(base.meta, conformation=k, sigma_nm=sigma)))
    return out


def batch_sizes(count: int = 64, lo: int = 300, hi: int = 6000, seed: int = 0) -> np.ndarray:
    """Atom counts of the batched workload, U[lo, hi] (seeded)."""
    return np.random.default_rng(seed).integers(lo, hi + 1, count)


def globules(specs, workers: int = 0) -> list:
    """``globule(n, seed)`` for every (n, seed) of ``specs``, in order; with ``workers`` > 1 on
    that many threads (the k-d tree and array work release the GIL).  Results are identical
    either way."""
    specs = [(int(n), int(s)) for n, s in specs]
    if workers <= 1 or len(specs) < 2:
        return [globule(n, s) for n, s in specs]
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(min(workers, len(specs))) as ex:
        return list(ex.map(lambda a: globule(*a), specs))


def batch64(seed: int = 0, workers: int = 0):
    """The batched workload: 64 globules with N ~ U[300, 6000] (``batch_sizes(64, 300, 6000,
    seed)``), system k seeded 1000 + k.  Returns (sizes, systems)."""
    sizes = batch_sizes(64, 300, 6000, seed=seed)
    return sizes, globules([(n, 1000 + k) for k, n in enumerate(sizes)], workers)
