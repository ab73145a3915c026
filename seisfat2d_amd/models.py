"""Seeded synthetic models and acquisition geometries of the five workloads.

Arrays are (nx, nz) float32 in the reference's z-fastest layout.  All random
draws use numpy.random.default_rng(seed) so every workload is reproducible.

  C1  iso, 201 x 1001, dh 10 m: v(z) = 1500 + 1.5 z with three +-10 %
      Gaussian anomalies (sigma 200 m), source (z=0, x=5000), 1001 surface
      receivers; constant 2000 m/s variant for the t = r/v check.
  C2  iso, 1001 x 4001, dh 5 m: 12-layer model (thickness U(200,600) m,
      velocities 1500 -> 5000 m/s with U(-300,300) jitter, sorted increasing),
      Gaussian-smoothed with sigma 3 cells; surface source at x = 10000 m.
  C3  the C2 model with 256 surface shots 20 m apart, each on a window of
      U{1200..4001} columns holding its source.
  C4  TTI qP, 801 x 3001, dh 5 m: Vp from the C2 generator, Thomsen epsilon
      in [0,0.3], delta in [-0.1,0.2], tilt in [-30,30] deg (white noise
      smoothed with sigma 50 cells, rescaled), rho = 1000, Vs = Vp/sqrt(3);
      128 surface shots.
  C5  iso, 301 x 1201, dh 10 m: v(z) = 1500 + 1.5 z with four +-5 % anomalies
      (true model) and v(z) alone (starting model); 128 shots, 401 receivers.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy.ndimage import gaussian_filter

from .stiffness import compute_stiffness


@dataclass
class Workload:
    name: str
    nz: int
    nx: int
    dz: float
    dx: float
    s: np.ndarray                      # slowness (nx, nz)
    shots: list = field(default_factory=list)   # (sx, sz, x0, width, rec_offset, nrec)
    rx: np.ndarray = None
    rz: np.ndarray = None
    stiffness: list = None             # six (nx, nz) arrays for anisotropic workloads
    extra: dict = field(default_factory=dict)


def _grid(nz, nx, dh):
    z = np.arange(nz, dtype=np.float64) * dh
    x = np.arange(nx, dtype=np.float64) * dh
    return x[:, None], z[None, :]


def _gaussian_anomalies(rng, nz, nx, dh, n, amp, sigma):
    x, z = _grid(nz, nx, dh)
    f = np.ones((nx, nz))
    for _ in range(n):
        cx = rng.uniform(0, (nx - 1) * dh)
        cz = rng.uniform(0, (nz - 1) * dh)
        a = amp * rng.choice([-1.0, 1.0])
        f += a * np.exp(-((x - cx) ** 2 + (z - cz) ** 2) / (2.0 * sigma ** 2))
    return f


def vz_model(nz, nx, dh, seed=1, n_anom=3, amp=0.10, sigma=200.0):
    rng = np.random.default_rng(seed)
    x, z = _grid(nz, nx, dh)
    v = (1500.0 + 1.5 * z) * np.ones_like(x)
    if n_anom:
        v = v * _gaussian_anomalies(rng, nz, nx, dh, n_anom, amp, sigma)
    return v


def layered_model(nz, nx, dh, seed=2, nlayers=12, smooth=3.0):
    rng = np.random.default_rng(seed)
    thick = rng.uniform(200.0, 600.0, nlayers)
    tops = np.concatenate([[0.0], np.cumsum(thick)[:-1]])
    vel = np.linspace(1500.0, 5000.0, nlayers) + rng.uniform(-300.0, 300.0, nlayers)
    vel = np.sort(vel)
    z = np.arange(nz) * dh
    idx = np.searchsorted(tops, z, side="right") - 1
    col = vel[np.clip(idx, 0, nlayers - 1)]
    v = np.broadcast_to(col[None, :], (nx, nz)).astype(np.float64)
    if smooth:
        v = gaussian_filter(v, smooth, mode="nearest")
    return v


def c1(constant=False, seed=1):
    nz, nx, dh = 201, 1001, 10.0
    v = np.full((nx, nz), 2000.0) if constant else vz_model(nz, nx, dh, seed)
    rx = np.arange(nx, dtype=np.float32) * dh
    rz = np.zeros(nx, np.float32)
    return Workload("C1" + ("-const" if constant else ""), nz, nx, dh, dh, (1.0 / v).astype(np.float32),
                    [(5000.0, 0.0, 0, nx, 0, nx)], rx, rz)


def c2(seed=2):
    nz, nx, dh = 1001, 4001, 5.0
    v = layered_model(nz, nx, dh, seed)
    rx = np.arange(0, nx, 4, dtype=np.float32) * dh
    rz = np.zeros(len(rx), np.float32)
    return Workload("C2", nz, nx, dh, dh, (1.0 / v).astype(np.float32),
                    [(10000.0, 0.0, 0, nx, 0, len(rx))], rx, rz)


def c3(nshots=256, seed=3, model_seed=2, rec_step=4):
    base = c2(model_seed)
    nz, nx, dh = base.nz, base.nx, base.dx
    rng = np.random.default_rng(seed)
    span = (nshots - 1) * 20.0
    x_first = 0.5 * ((nx - 1) * dh - span)
    widths = rng.integers(1200, 4002, nshots)
    shots, rxs = [], []
    off = 0
    for k in range(nshots):
        sx = x_first + 20.0 * k
        w = int(widths[k])
        col = int(round(sx / dh))
        x0 = int(np.clip(col - w // 2, 0, nx - w))
        cols = np.arange(x0, x0 + w, rec_step)
        rxs.append(cols.astype(np.float32) * dh)
        shots.append((sx, 0.0, x0, w, off, len(cols)))
        off += len(cols)
    rx = np.concatenate(rxs)
    return Workload("C3", nz, nx, dh, dh, base.s, shots, rx, np.zeros(len(rx), np.float32),
                    extra={"widths": widths})


def c4(nshots=128, seed=4, rec_step=10):
    nz, nx, dh = 801, 3001, 5.0
    vp = layered_model(nz, nx, dh, seed=2)
    rng = np.random.default_rng(seed)

    def field_in(lo, hi):
        f = gaussian_filter(rng.standard_normal((nx, nz)), 50.0, mode="nearest")
        f = (f - f.min()) / (f.max() - f.min())
        return lo + (hi - lo) * f

    eps = field_in(0.0, 0.3)
    dlt = field_in(-0.1, 0.2)
    tht = field_in(-30.0, 30.0)
    rho = np.full((nx, nz), 1000.0)
    vs = vp / np.sqrt(3.0)
    cs = compute_stiffness(vp, vs, rho, eps, dlt, tht)
    s = np.sqrt(rho / cs[3])  # makes Ro = c33 * S^2 equal rho in the qP kernel
    spacing = (nx - 1) * dh / nshots
    cols = np.arange(0, nx, rec_step)
    rx = cols.astype(np.float32) * dh
    shots = []
    rxa = []
    for k in range(nshots):
        sx = (k + 0.5) * spacing
        shots.append((sx, 0.0, 0, nx, k * len(cols), len(cols)))
        rxa.append(rx)
    rx = np.concatenate(rxa)
    return Workload("C4", nz, nx, dh, dh, s.astype(np.float32), shots, rx, np.zeros(len(rx), np.float32),
                    stiffness=[c.astype(np.float32) for c in cs],
                    extra={"eps": eps, "delta": dlt, "theta": tht, "vp": vp, "shot_spacing_m": spacing})


def c5(nshots=128, seed=5):
    nz, nx, dh = 301, 1201, 10.0
    v_true = vz_model(nz, nx, dh, seed, n_anom=4, amp=0.05, sigma=300.0)
    v_start = vz_model(nz, nx, dh, seed, n_anom=0)
    rx1 = np.arange(0.0, 12001.0, 30.0).astype(np.float32)
    spacing = (nx - 1) * dh / nshots
    shots = []
    for k in range(nshots):
        shots.append(((k + 0.5) * spacing, 0.0, 0, nx, k * len(rx1), len(rx1)))
    rx = np.tile(rx1, nshots)
    return Workload("C5", nz, nx, dh, dh, (1.0 / v_start).astype(np.float32), shots, rx,
                    np.zeros(len(rx), np.float32),
                    extra={"s_true": (1.0 / v_true).astype(np.float32), "shot_spacing_m": spacing})


def migration(seed=2, src_step=200.0, rec_step=50.0):
    """Kirchhoff migration workload with the reference's sample parameters
    (run/parameters.txt:5-12,72-90): a 1001 x 1001 model at 10 m (layered,
    seed 2), time_samples 1001 at 1 ms, max_frequency 30 Hz, mig_aperture
    30 deg, angles to 60 deg every 2 deg, offsets below 2000 m; surface
    sources every 200 m and receivers every 50 m (the sample's geometry files
    are not shipped)."""
    nz = nx = 1001
    dh = 10.0
    v = layered_model(nz, nx, dh, seed)
    xs = np.arange(0.0, (nx - 1) * dh + 1.0, src_step).astype(np.float32)
    xr = np.arange(0.0, (nx - 1) * dh + 1.0, rec_step).astype(np.float32)
    return Workload("MIG", nz, nx, dh, dh, (1.0 / v).astype(np.float32), [], xr, np.zeros(len(xr), np.float32),
                    extra={"xs": xs, "zs": np.zeros(len(xs), np.float32),
                           "params": dict(nt=1001, dt=1e-3, fmax=30.0, aperture=30.0, max_angle=60.0, da=2.0,
                                          max_offset=2000.0)})


def rough(nz, nx, dz, dx, seed=0, vlo=1500.0, vhi=6000.0, sigma=0.0):
    """Rough, high-contrast model for parity tests: white-noise velocity
    U(vlo, vhi) per node (optionally Gaussian-smoothed with `sigma` cells and
    rescaled to the same range), source at 37 % of the width and 20 % of the
    depth.  With dz:dx far from 1 the local solver's one-edge branches
    (B when dz > dx, C when dx > dz) decide many nodes."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.0, 1.0, (nx, nz))
    if sigma:
        u = gaussian_filter(u, sigma, mode="nearest")
        u = (u - u.min()) / max(u.max() - u.min(), 1e-30)
    v = vlo + (vhi - vlo) * u
    sx = float(np.float32(0.37 * (nx - 1) * dx))
    sz = float(np.float32(0.2 * (nz - 1) * dz))
    rx = rng.uniform(0, (nx - 1) * dx, 23).astype(np.float32)
    rz = rng.uniform(0, (nz - 1) * dz, 23).astype(np.float32)
    w = Workload(f"rough{nz}x{nx}", nz, nx, dz, dx, (1.0 / v).astype(np.float32),
                 [(sx, sz, 0, nx, 0, len(rx))], rx, rz)
    w.sx, w.sz = sx, sz
    return w


def cell_updates(nz, widths):
    """Reference-schedule work of one pass: 4 sweeps x interior padded nodes."""
    widths = np.asarray(widths, np.int64)
    return int(np.sum(4 * (nz + 4) * (widths + 4)))
