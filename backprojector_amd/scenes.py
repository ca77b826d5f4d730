"""Seeded synthetic NLOS captures for tests, benchmarks and training (no reference
counterpart: the reference reads Zaragoza HDF5 datasets, which are not available offline).

Geometry of every workload: relay wall z = 0, wall grid uniform over [-1, 1]^2 m, laser and
SPAD origins at (0, -1, 0.5) m, hidden volume [-0.5, 0.5]^2 x [0.3, 1.3] m.  Hidden geometry
is a list of surface samples (x, y, z, nx, ny, nz, area); captures are rendered by
libbpx (bpx_render_transient) with 1/r^4 falloff, Lambertian cosines and Poisson noise.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import core

ORIGIN = np.array([0.0, -1.0, 0.5], np.float32)
VOLUME_POSITION = np.array([0.0, 0.0, 0.8], np.float32)
VOLUME_SIZE = np.array([1.0, 1.0, 1.0], np.float32)
POINT_AREA = 1e-3  # weight of an isotropic hidden point (a ~3 cm patch)


def wall_grid(side: int) -> np.ndarray:
    """[side][side][3] float32 wall points, row-major (x index first)."""
    g = np.linspace(-1.0, 1.0, side, dtype=np.float64)
    X, Y = np.meshgrid(g, g, indexing="ij")
    return np.stack([X, Y, np.zeros_like(X)], -1).astype(np.float32)


def wall_pairs(laser_grid, camera_grid, confocal, origin=ORIGIN):
    """Pair table for rendering (laser/camera points, wall-to-origin distances)."""
    lg = laser_grid.reshape(-1, 3).astype(np.float64)
    cg = camera_grid.reshape(-1, 3).astype(np.float64)
    o = np.asarray(origin, np.float64)
    if confocal:
        pr = np.zeros(len(cg), core.PAIR_DTYPE)
        pr["camera_point"] = cg
        pr["laser_point"] = cg
        w = np.linalg.norm(cg - o, axis=1)
        pr["camera_wall"] = w
        pr["laser_wall"] = w
        return pr
    L, C = len(lg), len(cg)
    pr = np.zeros(L * C, core.PAIR_DTYPE)
    pr["laser_point"] = np.repeat(lg, C, axis=0)
    pr["camera_point"] = np.tile(cg, (L, 1))
    pr["laser_wall"] = np.repeat(np.linalg.norm(lg - o, axis=1), C)
    pr["camera_wall"] = np.tile(np.linalg.norm(cg - o, axis=1), L)
    return pr


def points(rng, n, lo=(-0.4, -0.4, 0.5), hi=(0.4, 0.4, 1.2)):
    p = rng.uniform(lo, hi, size=(n, 3))
    out = np.zeros((n, 7), np.float32)
    out[:, :3] = p
    out[:, 6] = POINT_AREA
    return out


def patch(rng, size=0.3, depth=(0.6, 1.1), spacing=0.002, centre_xy=None):
    """Fronto-parallel square patch facing the wall (normal -z), sampled every `spacing`."""
    z = rng.uniform(*depth)
    c = rng.uniform(-0.5 + size / 2, 0.5 - size / 2, size=2) if centre_xy is None else centre_xy
    k = max(2, int(round(size / spacing)))
    u = (np.arange(k) + 0.5) / k * size - size / 2
    X, Y = np.meshgrid(c[0] + u, c[1] + u, indexing="ij")
    out = np.zeros((k * k, 7), np.float32)
    out[:, 0] = X.ravel()
    out[:, 1] = Y.ravel()
    out[:, 2] = z
    out[:, 5] = -1.0
    out[:, 6] = (size / k) ** 2
    return out


def sphere(rng, radius=None, spacing=0.004, lo=(-0.35, -0.35, 0.55), hi=(0.35, 0.35, 1.15)):
    """Sphere sampled on a Fibonacci lattice, outward normals."""
    r = rng.uniform(0.05, 0.15) if radius is None else radius
    c = rng.uniform(lo, hi)
    n = max(16, int(4 * np.pi * r * r / spacing ** 2))
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    th = np.pi * (1 + 5 ** 0.5) * i
    d = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], -1)
    out = np.zeros((n, 7), np.float32)
    out[:, :3] = c + r * d
    out[:, 3:6] = d
    out[:, 6] = 4 * np.pi * r * r / n
    return out


@dataclass
class Capture:
    transient: np.ndarray          # confocal [X][Y][T] / exhaustive [Lx][Ly][Cx][Cy][T]
    camera_grid: np.ndarray
    laser_grid: np.ndarray
    camera_position: np.ndarray
    laser_position: np.ndarray
    t0: float
    deltaT: float
    confocal: bool
    volume_position: np.ndarray = field(default_factory=lambda: VOLUME_POSITION.copy())
    volume_size: np.ndarray = field(default_factory=lambda: VOLUME_SIZE.copy())
    hidden: np.ndarray | None = None

    def nlos_args(self, voxels_per_side):
        """Positional arguments of nlosbpy.backproject up to voxels_per_side."""
        return (self.transient, self.camera_grid, self.laser_grid, self.camera_position,
                self.laser_position, self.t0, self.deltaT, self.confocal, self.volume_position,
                self.volume_size, np.asarray(voxels_per_side, np.uint32))


def render(hidden, side, T, t0=0.0, deltaT=0.001, confocal=True, laser_side=None, peak=1e3,
           poisson=True, seed=0) -> Capture:
    cg = wall_grid(side)
    lg = wall_grid(laser_side or side)
    pr = wall_pairs(lg if not confocal else cg, cg, confocal)
    H = core.render_transient(pr, T, t0, deltaT, hidden, peak_count=peak, poisson=poisson,
                              seed=seed)
    shape = (side, side, T) if confocal else (lg.shape[0], lg.shape[1], side, side, T)
    return Capture(H.reshape(shape), cg, cg if confocal else lg, ORIGIN.copy(), ORIGIN.copy(),
                   float(t0), float(deltaT), confocal, hidden=hidden)


# ----------------------------------------------------------------- the workload configs
def config_scene(k: int, rng=None):
    """Hidden geometry of workload k (seed 1000 + k)."""
    rng = rng or np.random.default_rng(1000 + k)
    if k == 1:
        return points(rng, 3)
    if k == 2:
        return np.concatenate([points(rng, 5), patch(rng)])
    if k in (3, 5, 0):
        return np.concatenate([points(rng, 8), patch(rng), patch(rng)])
    raise ValueError(k)


def metric_capture(T=4096, side=128, seed=1000, patch_spacing=0.002) -> Capture:
    """128 x 128 confocal, T = 4096 bins of 1 mm, 8 points + 2 patches (headline workload)."""
    rng = np.random.default_rng(seed)
    hidden = np.concatenate([points(rng, 8), patch(rng, spacing=patch_spacing),
                             patch(rng, spacing=patch_spacing)])
    return render(hidden, side, T, 0.0, 0.001, True, seed=seed)


def batch_captures(count=64, seed=1004, sides=(16, 32, 64, 96, 128), Ts=(1024, 2048, 4096),
                   voxels=(16, 32, 64, 96)):
    """Uneven confocal captures for batched reconstruction: capture i draws its wall side, bin
    count T (bin width 4.096 m / T) and voxels per side from SeedSequence(seed).spawn(count)[i];
    hidden scene 2-8 points and one 0.3 m patch.  Returns [(Capture, voxels_per_side)]."""
    out = []
    for ss in np.random.SeedSequence(seed).spawn(count):
        rng = np.random.default_rng(ss)
        side = int(rng.choice(sides))
        T = int(rng.choice(Ts))
        N = int(rng.choice(voxels))
        hidden = np.concatenate([points(rng, int(rng.integers(2, 9))), patch(rng, spacing=0.004)])
        cap = render(hidden, side, T, 0.0, 4.096 / T, True, seed=int(rng.integers(1 << 31)))
        out.append((cap, N))
    return out
