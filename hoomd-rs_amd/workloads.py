"""Seeded synthetic initial states for the sweep workloads.

All generators return (box_edges, positions) with positions as an (N, D)
float64 array centred on the origin, inside [-L/2, L/2).
"""
import math

import numpy as np


def square_lattice_disks(n_side=128, packing_fraction=0.70, diameter=1.0):
    """Hard disks on an n_side x n_side square lattice scaled to the packing fraction."""
    n = n_side * n_side
    area = n * math.pi * (diameter / 2) ** 2 / packing_fraction
    L = math.sqrt(area)
    a = L / n_side
    i, j = np.meshgrid(np.arange(n_side), np.arange(n_side), indexing="ij")
    pos = np.stack([(i.ravel() + 0.5) * a - L / 2, (j.ravel() + 0.5) * a - L / 2], axis=1)
    return np.array([L, L]), pos


def fcc_spheres(n_cells=(63, 63, 63), packing_fraction=0.50, diameter=1.0):
    """Spheres on an FCC lattice of n_cells conventional cells (4 sites each)."""
    nx, ny, nz = n_cells
    n = 4 * nx * ny * nz
    volume = n * math.pi * diameter**3 / 6 / packing_fraction
    return fcc_at_volume(n_cells, volume)


def fcc_number_density(n_cells, number_density):
    nx, ny, nz = n_cells
    return fcc_at_volume(n_cells, 4 * nx * ny * nz / number_density)


def fcc_at_volume(n_cells, volume):
    nx, ny, nz = n_cells
    a = (volume / (nx * ny * nz)) ** (1.0 / 3.0)
    L = np.array([nx * a, ny * a, nz * a])
    basis = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]]) + 0.25
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cells = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
    pos = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3) * a - L / 2
    return L, pos


def random_dilute(n, L, dim, min_distance, seed):
    """Random non-overlapping points by sequential rejection (small systems only)."""
    rng = np.random.default_rng(seed)
    L = np.asarray(L, np.float64)
    pts = []
    tries = 0
    while len(pts) < n:
        tries += 1
        if tries > 1000 * n:
            raise RuntimeError("could not place points")
        p = rng.uniform(-L / 2, L / 2, size=dim)
        if pts:
            d = np.asarray(pts) - p
            d -= L * np.round(d / L)
            if np.min((d * d).sum(1)) < min_distance**2:
                continue
        pts.append(p)
    return np.asarray(pts)


def batched_replicas(n_replicas=256, seed=2024, k_range=(8, 25), phi_range=(0.30, 0.54)):
    """Replicas of hard spheres on FCC lattices: N per replica log-uniform in
    [4 k_min^3, 4 k_max^3] rounded to an FCC count 4 k^3, packing fractions
    evenly spaced over phi_range.  Returns a list of (box_edges, positions)."""
    rng = np.random.default_rng(seed)
    lo, hi = 4 * k_range[0] ** 3, 4 * k_range[1] ** 3
    out = []
    for r in range(n_replicas):
        n = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        k = int(round((n / 4.0) ** (1.0 / 3.0)))
        k = min(max(k, k_range[0]), k_range[1])
        phi = phi_range[0] + (phi_range[1] - phi_range[0]) * (r / max(n_replicas - 1, 1))
        out.append(fcc_spheres((k, k, k), packing_fraction=phi))
    return out


def wrap(box_edges, positions):
    """Periodic<Hypercuboid>::wrap (hoomd-microstate/src/boundary/periodic/
    cuboid.rs:90-103), the form in which Microstate::add_body stores body
    positions: per axis lambda = x / L, lambda - round(lambda) with halves
    rounded away from zero (Rust's f64::round), +1/2 mapped to -1/2, times L.
    Not idempotent in floating point."""
    L = np.asarray(box_edges, np.float64)[: np.shape(positions)[1]]
    lam = np.asarray(positions, np.float64) / L
    t = np.trunc(lam)
    r = np.where(np.abs(lam - t) >= 0.5, t + np.sign(lam), t)  # keeps round(-0.0) = -0.0
    lam = lam - r
    lam = np.where(lam == 0.5, -0.5, lam)
    return lam * L


# ConvexPolytope vertices of the unit cube, in the order of the reference's
# test fixtures
UNIT_CUBE = np.array([[-0.5, -0.5, -0.5], [0.5, -0.5, -0.5], [0.5, 0.5, -0.5], [-0.5, 0.5, -0.5],
                      [-0.5, -0.5, 0.5], [0.5, -0.5, 0.5], [0.5, 0.5, 0.5], [-0.5, 0.5, 0.5]])


def simple_cubic(n_side=64, packing_fraction=0.55, particle_volume=1.0):
    """n_side^3 bodies of the given volume on a simple cubic lattice at the
    packing fraction, centred in a cubic box.  Returns (box_edges, positions)."""
    L = (n_side**3 * particle_volume / packing_fraction) ** (1.0 / 3.0)
    a = L / n_side
    i, j, k = np.meshgrid(np.arange(n_side), np.arange(n_side), np.arange(n_side), indexing="ij")
    pos = (np.stack([i.ravel(), j.ravel(), k.ravel()], 1) + 0.5) * a - L / 2
    return np.array([L, L, L]), pos
