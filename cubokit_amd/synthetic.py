"""Seeded synthetic cubes of the benchmark workloads, generated on the device (cubo_hip_cube_fill_synthetic).

Workloads (scan rows x scan columns x detector rows x detector columns):
- ``small_f32``: 20x20x512x512 Float32; disk r=30 with a 2 px sigmoid edge, 1000 counts/px, centre (256, 256)
  deflected by (3 sin(2 pi x/20), 3 cos(2 pi y/20)); background 2 counts/px; Poisson.
- ``bandwidth_u16``: 256x256x128x128 uint16; disk r=12 at 500 counts/px with a +-2 px sinusoidal deflection;
  Gaussian rings at q=30 and q=45 px, amplitude 20; Poisson.
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib as L


def params(seed: int, center, radius: float, counts: float, edge: float = 0.0, background: float = 0.0,
           rings=(), waves=(), poisson: bool = True) -> L.SynthParams:
    p = L.SynthParams()
    p.seed = seed
    p.center_row, p.center_column = center
    p.radius, p.edge, p.counts, p.background = radius, edge, counts, background
    p.n_rings = len(rings)
    for k, (q, a, s) in enumerate(rings):
        p.ring_q[k], p.ring_amplitude[k], p.ring_sigma[k] = q, a, s
    p.n_waves = len(waves)
    for k, (ar, ac, fr, fc, ph) in enumerate(waves):
        p.wave_row_amp[k], p.wave_col_amp[k], p.wave_fr[k], p.wave_fc[k], p.wave_phase[k] = ar, ac, fr, fc, ph
    p.poisson = int(poisson)
    return p


def small_f32(seed: int = 20261101):
    """The 20x20x512x512 Float32 parity workload. The deflection is 3 sin(2 pi x/20) along rows and
    3 cos(2 pi y/20) along columns: two waves, one per scan axis (the other amplitude zero)."""
    shape = (20, 20, 512, 512)
    waves = [(3.0, 0.0, 1 / 20, 0.0, 0.0), (0.0, 3.0, 0.0, 1 / 20, 0.0)]
    return shape, np.float32, params(seed, (256.0, 256.0), 30.0, 1000.0, edge=2.0, background=2.0, waves=waves)


def bandwidth_u16(seed: int = 20261102, scan=(256, 256)):
    """The 256x256x128x128 uint16 bandwidth workload (2 GiB)."""
    shape = (scan[0], scan[1], 128, 128)
    waves = [(2.0, 0.0, 1 / 64, 0.0, 0.0), (0.0, 2.0, 0.0, 1 / 64, 0.0)]
    rings = [(30.0, 20.0, 1.5), (45.0, 20.0, 1.5)]
    return shape, np.uint16, params(seed, (64.0, 64.0), 12.0, 500.0, edge=0.5, rings=rings, waves=waves)


def random_waves(rng: np.random.Generator, n: int, amplitude: float, scan):
    """n deflection waves whose summed amplitude stays within +-amplitude."""
    out = []
    for _ in range(n):
        ar, ac = rng.uniform(-1, 1, 2) * amplitude / n
        fr, fc = rng.integers(1, 4, 2) / np.array(scan, float)
        out.append((float(ar), float(ac), float(fr), float(fc), float(rng.uniform(0, 2 * math.pi))))
    return out


def stack_masks(qy: int, qx: int, seed: int = 20261103, n_sectors: int = 29, n_dense: int = 32) -> np.ndarray:
    """The K = 3 + n_sectors + n_dense mask stack: {1, q_row, q_col}, annular sectors with seeded radii in
    [0, 64) and angles, and dense uniform [0, 1) masks. Float32, K x (qy * qx)."""
    rng = np.random.default_rng(seed)
    ii, jj = np.indices((qy, qx), dtype=np.float64)
    cy, cx = qy / 2.0, qx / 2.0
    d = np.hypot(ii - cy, jj - cx)
    ang = np.arctan2(ii - cy, jj - cx)
    rows = [np.ones(qy * qx), ii.reshape(-1), jj.reshape(-1)]
    for _ in range(n_sectors):
        r0, r1 = np.sort(rng.uniform(0, 64, 2))
        a0 = rng.uniform(-math.pi, math.pi)
        width = rng.uniform(0.3, 2 * math.pi)
        da = np.mod(ang - a0, 2 * math.pi)
        rows.append(((d >= r0) & (d < r1) & (da < width)).astype(np.float64).reshape(-1))
    for _ in range(n_dense):
        rows.append(rng.uniform(0, 1, qy * qx))
    return np.ascontiguousarray(np.stack(rows).astype(np.float32))


BATCH_SCANS = [(32, 32), (64, 64), (96, 128), (128, 128), (256, 256)]
BATCH_DETECTORS = [64, 96, 128, 256]


def batch_items(n: int = 48, seed: int = 20261104):
    """A seeded screening batch of n uneven uint16 cubes: scan in {32x32, 64x64, 96x128, 128x128, 256x256},
    square detectors of side {64, 96, 128, 256} (up to 256x256x256x256, 8 GiB), 8..64 masks each (the first k
    rows of stack_masks for that detector). Returns [(shape, SynthParams, k)]; batch_bytes gives their sum (47.4 GiB
    for the default seed), which the caller checks against the device (StackBatch.availability)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        scan = BATCH_SCANS[rng.integers(len(BATCH_SCANS))]
        det = BATCH_DETECTORS[rng.integers(len(BATCH_DETECTORS))]
        shape = (scan[0], scan[1], det, det)
        k = int(rng.integers(8, 65))
        p = params(int(rng.integers(1 << 62)), (det / 2, det / 2), det / 10, float(rng.uniform(100, 3000)),
                   edge=0.5, background=2.0, rings=[(det / 4, 20.0, 1.5)],
                   waves=random_waves(rng, 2, 3.0, scan))
        out.append((shape, p, k))
    return out


def batch_bytes(items) -> int:
    """The uint16 bytes of every cube of a batch_items list."""
    return int(sum(2 * int(np.prod(shape, dtype=np.int64)) for shape, _, _ in items))


def differentiation_u16(seed: int = 20261105, scan=(512, 512)):
    """The 512x512x128x128 uint16 differentiation workload (8 GiB): disk r=12 at 500 counts/px, background
    2/px, and a smooth deflection field, the sum of 8 seeded sinusoids within +-2 px."""
    rng = np.random.default_rng(seed)
    shape = (scan[0], scan[1], 128, 128)
    waves = random_waves(rng, 8, 2.0, scan)
    return shape, np.uint16, params(seed, (64.0, 64.0), 12.0, 500.0, edge=0.5, background=2.0, waves=waves)


def differentiation_masks(qy: int = 128, qx: int = 128, seed: int = 20261106) -> np.ndarray:
    """K = 16 masks (8 seeded annular sectors, 8 dense uniform [0, 1)) followed by the centre-of-mass rows
    {1, q_row, q_col}: 19 rows, Float32."""
    m = stack_masks(qy, qx, seed=seed, n_sectors=8, n_dense=8)       # {1, q_row, q_col, 8 sectors, 8 dense}
    return np.ascontiguousarray(np.concatenate([m[3:], m[:3]]), np.float32)
