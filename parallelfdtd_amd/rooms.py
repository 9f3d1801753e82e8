"""Seeded synthetic rooms, sources and receivers for the measured workloads.

Everything is driven by one documented generator, SplitMix64 (Steele, Lea & Flood 2014), so
the GPU path, the CPU oracle and the fixtures see identical inputs.

A room is a *label grid* of shape (Z, Y, X), uint8: 0 = inside (air), m+1 = solid voxel of
material m. Every room has a one-voxel solid shell. Classification of labels into position
and material bytes happens in ``pfdtd_set_domain_labels`` (GPU) and in the oracle (CPU).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MASK64 = (1 << 64) - 1
LAMBDA = math.sqrt(1.0 / 3.0)  # SimulationParameters::setUpdateType, SimulationParameters.cpp:123-137


class SplitMix64:
    """SplitMix64 generator; uniform doubles use the top 53 bits."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi] (inclusive)."""
        return lo + self.next_u64() % (hi - lo + 1)

    def fork(self, tag: int) -> "SplitMix64":
        """Independent stream derived from this one's seed and a tag."""
        return SplitMix64(self.next_u64() ^ ((tag * 0xD1342543DE82EF95) & MASK64))


def box_room(X: int, Y: int, Z: int, floor: int = 0, ceiling: int = 0, walls: int = 0) -> np.ndarray:
    """Empty box room with a one-voxel solid shell; materials per shell face."""
    lab = np.zeros((Z, Y, X), dtype=np.uint8)
    lab[:, :, 0] = lab[:, :, X - 1] = walls + 1
    lab[:, 0, :] = lab[:, Y - 1, :] = walls + 1
    lab[0, :, :] = floor + 1
    lab[Z - 1, :, :] = ceiling + 1
    return lab


def add_boxes(lab: np.ndarray, rng: SplitMix64, n: int, smin: int, smax: int,
              materials: list[int]) -> list[tuple]:
    """Paint n solid boxes with sides in [smin, smax] fully inside the shell."""
    Z, Y, X = lab.shape
    boxes = []
    for b in range(n):
        sx = min(rng.randint(smin, smax), X - 4)
        sy = min(rng.randint(smin, smax), Y - 4)
        sz = min(rng.randint(smin, smax), Z - 4)
        x0 = rng.randint(2, X - 2 - sx)
        y0 = rng.randint(2, Y - 2 - sy)
        z0 = rng.randint(2, Z - 2 - sz)
        m = materials[b % len(materials)]
        lab[z0:z0 + sz, y0:y0 + sy, x0:x0 + sx] = m + 1
        boxes.append((x0, y0, z0, sx, sy, sz, m))
    return boxes


def is_air_k6(lab: np.ndarray, x: int, y: int, z: int) -> bool:
    Z, Y, X = lab.shape
    if not (1 <= x < X - 1 and 1 <= y < Y - 1 and 1 <= z < Z - 1):
        return False
    return (lab[z, y, x] == 0 and lab[z, y, x - 1] == 0 and lab[z, y, x + 1] == 0
            and lab[z, y - 1, x] == 0 and lab[z, y + 1, x] == 0
            and lab[z - 1, y, x] == 0 and lab[z + 1, y, x] == 0)


def seeded_points(lab: np.ndarray, rng: SplitMix64, n: int, margin: int = 2) -> np.ndarray:
    """n distinct inside points whose six face neighbours are inside too (K = 6)."""
    Z, Y, X = lab.shape
    pts: list[tuple[int, int, int]] = []
    seen = set()
    tries = 0
    while len(pts) < n:
        tries += 1
        if tries > 1000000:
            raise RuntimeError("could not place points")
        x = rng.randint(margin, X - 1 - margin)
        y = rng.randint(margin, Y - 1 - margin)
        z = rng.randint(margin, Z - 1 - margin)
        if (x, y, z) in seen or not is_air_k6(lab, x, y, z):
            continue
        seen.add((x, y, z))
        pts.append((x, y, z))
    return np.asarray(pts, dtype=np.int32).reshape(n, 3)


def gaussian_pulse(n_steps: int, t0: float, sigma: float) -> np.ndarray:
    """Unit-peak Gaussian exp(-((n - t0)/sigma)^2 / 2), float64."""
    n = np.arange(n_steps, dtype=np.float64)
    return np.exp(-0.5 * ((n - t0) / sigma) ** 2)


def gaussian_derivative_pulse(n_steps: int, t0: float, sigma: float) -> np.ndarray:
    """Unit-peak band-limited Gaussian-derivative pulse, float64."""
    n = np.arange(n_steps, dtype=np.float64)
    u = (n - t0) / sigma
    return -u * np.exp(0.5 - 0.5 * u * u)


@dataclass
class Scene:
    """One workload: room labels, materials, sources, receivers and run length."""
    name: str
    labels: np.ndarray
    betas: np.ndarray                 # admittance per material, float32
    src_xyz: np.ndarray               # (n_src, 3) int32 grid indices
    src_kind: np.ndarray              # (n_src,) int32
    signals: np.ndarray               # (n_src, n_samples) float64
    rec_xyz: np.ndarray               # (n_rec, 3) int32
    n_steps: int
    meta: dict = field(default_factory=dict)

    @property
    def shape_xyz(self) -> tuple[int, int, int]:
        Z, Y, X = self.labels.shape
        return X, Y, Z

    @property
    def n_nodes(self) -> int:
        return int(self.labels.size)

    def material_table(self, n_coef: int = 20) -> np.ndarray:
        """[n_mat][n_coef] table with the admittance in every octave slot."""
        return np.repeat(self.betas.astype(np.float32)[:, None], n_coef, axis=1)


def scene_c1(n_steps: int = 2000, seed: int = 1) -> Scene:
    """130x98x66 box (128x96x64 air, one-voxel shell), one material beta=0.1, one soft
    Gaussian source (sigma 10 steps, unit peak), one receiver."""
    rng = SplitMix64(seed)
    lab = box_room(130, 98, 66)
    src = seeded_points(lab, rng.fork(1), 1)
    rec = seeded_points(lab, rng.fork(2), 1)
    return Scene("c1_box_130x98x66", lab, np.array([0.1], np.float32), src,
                 np.array([1], np.int32), gaussian_pulse(n_steps, 50.0, 10.0)[None, :],
                 rec, n_steps)


def scene_c2(n_steps: int = 4000, seed: int = 2, scale: int = 1) -> Scene:
    """L-shaped hall 1024x768x384 (corner block removed), 8 solid boxes (sides 8..96),
    5 materials (floor, ceiling, walls, cut walls, boxes), 16 receivers. `scale` divides
    every dimension (for CPU-sized variants of the same scene)."""
    rng = SplitMix64(seed)
    X, Y, Z = 1024 // scale, 768 // scale, 384 // scale
    betas = np.array([rng.uniform(0.02, 0.5) for _ in range(5)], np.float32)
    lab = box_room(X, Y, Z, floor=0, ceiling=1, walls=2)
    x0 = int(X * rng.uniform(0.4, 0.7))
    y0 = int(Y * rng.uniform(0.4, 0.7))
    lab[1:Z - 1, y0:, x0:] = 3 + 1  # cut corner: solid with the "cut walls" material
    add_boxes(lab, rng.fork(3), 8, max(2, 8 // scale), max(3, 96 // scale), [4])
    src = seeded_points(lab, rng.fork(4), 1)
    rec = seeded_points(lab, rng.fork(5), 16)
    return Scene(f"c2_lhall_{X}x{Y}x{Z}", lab, betas, src, np.array([1], np.int32),
                 gaussian_pulse(n_steps, 50.0, 10.0)[None, :], rec, n_steps)


def scene_c3(n_steps: int = 3000, seed: int = 3, scale: int = 1) -> Scene:
    """Box 2048x1536x768 (2.4 G voxels) with 32 seeded boxes (sides 16..192), 8 materials,
    a Gaussian-derivative source (sigma 8, t0 48, unit peak), 64 receivers."""
    rng = SplitMix64(seed)
    X, Y, Z = 2048 // scale, 1536 // scale, 768 // scale
    betas = np.array([rng.uniform(0.01, 0.6) for _ in range(8)], np.float32)
    lab = box_room(X, Y, Z, floor=0, ceiling=1, walls=2)
    add_boxes(lab, rng.fork(3), 32, max(2, 16 // scale), max(3, 192 // scale), [3, 4, 5, 6, 7])
    src = seeded_points(lab, rng.fork(4), 1)
    rec = seeded_points(lab, rng.fork(5), 64)
    return Scene(f"c3_box_{X}x{Y}x{Z}", lab, betas, src, np.array([1], np.int32),
                 gaussian_derivative_pulse(n_steps, 48.0, 8.0)[None, :], rec, n_steps)


def scene_c4(n_sources: int = 32, n_steps: int = 2000, seed: int = 4, scale: int = 1) -> Scene:
    """512x384x256 with 12 boxes, 4 materials; n_sources seeded sources, 8 receivers each."""
    rng = SplitMix64(seed)
    X, Y, Z = 512 // scale, 384 // scale, 256 // scale
    betas = np.array([rng.uniform(0.02, 0.4) for _ in range(4)], np.float32)
    lab = box_room(X, Y, Z, floor=0, ceiling=1, walls=2)
    add_boxes(lab, rng.fork(3), 12, max(2, 8 // scale), max(3, 64 // scale), [3])
    src = seeded_points(lab, rng.fork(4), n_sources)
    rec = seeded_points(lab, rng.fork(5), 8 * n_sources)
    return Scene(f"c4_{X}x{Y}x{Z}_B{n_sources}", lab, betas, src,
                 np.ones(n_sources, np.int32),
                 np.repeat(gaussian_pulse(n_steps, 50.0, 10.0)[None, :], n_sources, axis=0),
                 rec, n_steps)


def scene_random_small(seed: int, dims=(12, 10, 9), n_mat: int = 3, n_boxes: int = 2,
                       n_src: int = 2, n_rec: int = 3, n_steps: int = 60) -> Scene:
    """Tiny random room for parity tests: random boxes, materials, mixed hard/soft sources."""
    rng = SplitMix64(seed)
    X, Y, Z = dims
    betas = np.array([rng.uniform(0.01, 0.6) for _ in range(n_mat)], np.float32)
    lab = box_room(X, Y, Z, floor=0, ceiling=min(1, n_mat - 1), walls=min(2, n_mat - 1))
    add_boxes(lab, rng.fork(3), n_boxes, 1, 3, list(range(n_mat)))
    # sources/receivers anywhere strictly inside the shell (inside or solid voxels)
    def pts(r, n):
        return np.array([[r.randint(1, X - 2), r.randint(1, Y - 2), r.randint(1, Z - 2)]
                         for _ in range(n)], np.int32).reshape(n, 3)
    src = pts(rng.fork(4), n_src)
    rec = pts(rng.fork(5), n_rec)
    kinds = np.array([rng.randint(0, 1) for _ in range(n_src)], np.int32)
    r6 = rng.fork(6)
    sig = np.array([[r6.uniform(-1, 1) for _ in range(n_steps)] for _ in range(n_src)])
    return Scene(f"rand{seed}", lab, betas, src, kinds, sig.reshape(n_src, n_steps), rec, n_steps)
