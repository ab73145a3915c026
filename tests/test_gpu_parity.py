"""GPU parity of the eikonal path against the CPU oracle.

The contract is bit-exactness: both schedules evaluate the reference's local
solver with the same correctly rounded FP32 operations in a dependency-
equivalent order, so traveltimes and picks must equal the oracle's bit for
bit (tolerance 0).  Full-size workloads are compared node by node where the
oracle finishes in seconds (C2 single shot, every C3 window with the
shot-parallel oracle) and through properties otherwise.
"""
import ctypes as C

import numpy as np
import pytest
from scipy.ndimage import gaussian_filter

from seisfat2d_amd import ANI, ISO, SCHED_PLAIN, SCHED_TILED, SCHED_TILED_DIAG, Eikonal, capi, models

SCHEDULES = (SCHED_TILED, SCHED_TILED_DIAG, SCHED_PLAIN)

pytestmark = pytest.mark.gpu
NB = 3


def smooth_slowness(rng, nz, nx, dh, sigma=3.0):
    v = 2000 + 500 * gaussian_filter(rng.standard_normal((nx, nz)), sigma) + 1.0 * np.arange(nz)[None, :] * dh
    return (1.0 / v).astype(np.float32)


def solve_both(oracle, s, dx, dz, sx, sz, rx=None, rz=None):
    nx, nz = s.shape
    e = Eikonal(nz, nx, dz, dx)
    e.set_slowness(s)
    Tref = oracle.solve_iso(oracle.expand_boundary(s), dx, dz, sx, sz)
    out = {}
    for sched in SCHEDULES:
        out[sched] = e.solve(sx, sz, schedule=sched, rx=rx, rz=rz)
    e.close()
    return Tref, out


def test_c1_bit_exact_both_schedules(oracle):
    w = models.c1()
    Tref, out = solve_both(oracle, w.s, w.dx, w.dz, 5000.0, 0.0, w.rx, w.rz)
    pref = oracle.picks(Tref, w.dx, w.dz, w.rx, w.rz)
    for sched, (T, pk) in out.items():
        assert np.array_equal(T, Tref), sched
        assert np.array_equal(pk, pref), sched


def test_c1_constant_velocity_vs_r_over_v(oracle):
    w = models.c1(constant=True)
    Tref, out = solve_both(oracle, w.s, w.dx, w.dz, 5000.0, 0.0)
    T = out[SCHED_TILED][0]
    assert np.array_equal(T, Tref)
    x = (np.arange(w.nx + 6) - NB) * w.dx
    z = (np.arange(w.nz + 6) - NB) * w.dz
    r = np.hypot(x[:, None] - 5000.0, z[None, :])[NB:-NB, NB:-NB]
    m = r > 0
    rel = np.abs(T[NB:-NB, NB:-NB][m] - r[m] / 2000.0) / (r[m] / 2000.0)
    print(f"C1 constant velocity: max rel err vs r/v = {rel.max():.4f}, mean {rel.mean():.5f}")
    assert rel.max() < 0.03


@pytest.mark.parametrize("nz,nx,dz,dx,sxf,szf", [
    (2, 2, 10.0, 10.0, 0.0, 0.0),
    (3, 97, 5.0, 5.0, 1.0, 1.0),
    (59, 3, 5.0, 5.0, 0.5, 0.0),
    (64, 64, 10.0, 10.0, 0.3, 0.7),
    (65, 130, 10.0, 8.0, 0.0, 1.0),
    (123, 66, 7.0, 10.0, 1.0, 0.0),
    (130, 200, 4.0, 4.0, 0.37, 0.61),
    (200, 129, 10.0, 10.0, 0.999, 0.5),
])
def test_shapes_sources_and_spacings(oracle, rng, nz, nx, dz, dx, sxf, szf):
    s = smooth_slowness(rng, nz, nx, min(dx, dz))
    sx, sz = sxf * (nx - 1) * dx, szf * (nz - 1) * dz
    rx = rng.uniform(0, (nx - 1) * dx, 17).astype(np.float32)
    rz = rng.uniform(0, (nz - 1) * dz, 17).astype(np.float32)
    Tref, out = solve_both(oracle, s, dx, dz, sx, sz, rx, rz)
    pref = oracle.picks(Tref, dx, dz, rx, rz)
    for sched, (T, pk) in out.items():
        assert np.array_equal(T, Tref), (sched, nz, nx)
        assert np.array_equal(pk, pref), (sched, nz, nx)


FDIV_FALLBACK_H = float.fromhex("0x1.94f5a6p+2")  # dz whose two-instruction division is inexact


@pytest.mark.parametrize("nz,nx,dz,dx,seed,branch", [
    (64, 96, 3.0, 10.0, 11, "C"),
    (96, 64, 10.0, 3.0, 12, "B"),
    (123, 66, 7.0, 10.0, 13, None),
    (140, 230, FDIV_FALLBACK_H, 2.0, 14, "B"),
    (70, 150, 2.5, 9.0, 15, "C"),
])
@pytest.mark.parametrize("pairs", ["1", "0"])
def test_rough_unequal_spacing_all_schedules_bit_exact(oracle, monkeypatch, nz, nx, dz, dx, seed, branch, pairs):
    """Rough, high-contrast models (white-noise velocity 1500-6000 m/s, no
    smoothing) on grids with dz:dx far from 1, where the local solver's
    one-edge branches B and C (src/modeling/modeling.cu:344-351) and branch
    A's causality guard decide many nodes (counted by the oracle): every
    schedule — per-diagonal launches, tile diagonals, the persistent tiles as
    fused sweep pairs or one launch per sweep — and a batch of windows,
    bit-exact against the oracle."""
    monkeypatch.setenv("SEISFAT_SWEEP_PAIRS", pairs)
    m = models.rough(nz, nx, dz, dx, seed=seed)
    S = oracle.expand_boundary(m.s)
    counts = oracle.branch_counts(S, dx, dz, m.sx, m.sz)
    print(f"{nz}x{nx} dz={dz} dx={dx}: deciding candidates {counts}")
    if branch:
        assert counts[branch] >= 20, counts
    Tref = oracle.solve_iso(S, dx, dz, m.sx, m.sz)
    pref = oracle.picks(Tref, dx, dz, m.rx, m.rz)
    e = Eikonal(nz, nx, dz, dx)
    e.set_slowness(m.s)
    for sched in SCHEDULES:
        T, pk = e.solve(m.sx, m.sz, schedule=sched, rx=m.rx, rz=m.rz)
        assert np.array_equal(T, Tref), sched
        assert np.array_equal(pk, pref), sched
    rng = np.random.default_rng(seed)
    shots = []
    for k in range(9):
        wd = int(rng.integers(1, nx + 1)) if k else nx
        x0 = int(rng.integers(0, nx - wd + 1))
        shots.append((float(np.float32((x0 + rng.uniform(0.0, wd - 1)) * dx)),
                      float(np.float32(rng.uniform(0.0, (nz - 1) * dz))), x0, wd, 0, 0))
    b = e.batch(shots, schedule=SCHED_TILED)
    assert b.num_sweep_launches == (2 if pairs == "1" else 4)
    Trefs = oracle.solve_windows(m.s, dx, dz, [x[0] for x in shots], [x[1] for x in shots],
                                 [x[2] for x in shots], [x[3] for x in shots], nthreads=16)
    for run in range(2):
        b.run()
        b.status()
        for k, Tr in enumerate(Trefs):
            assert np.array_equal(b.T(k), Tr), (run, k)
    b.close()
    e.close()


def test_rough_anisotropic_unequal_spacing_bit_exact(oracle):
    """qP solves on a rough TTI model with dz:dx = 1:3.3 (the qP slowness
    pass feeds one-edge branches too), every schedule, bit-exact."""
    from seisfat2d_amd.stiffness import compute_stiffness
    nz, nx, dz, dx = 70, 110, 3.0, 10.0
    rng = np.random.default_rng(21)
    vp = rng.uniform(1500.0, 6000.0, (nx, nz))
    cs = compute_stiffness(vp, vp / np.sqrt(3), 1000.0, rng.uniform(0.0, 0.3, (nx, nz)),
                           rng.uniform(-0.1, 0.2, (nx, nz)), rng.uniform(-30.0, 30.0, (nx, nz)))
    s = np.sqrt(1000.0 / cs[3]).astype(np.float32)
    cs = [c.astype(np.float32) for c in cs]
    S = oracle.expand_boundary(s)
    Cq, mn, mx = [], [], []
    for c in cs:
        u, a, b_ = oracle.compress(oracle.expand_boundary(c))
        Cq.append(u)
        mn.append(a)
        mx.append(b_)
    src = [(0.37 * (nx - 1) * dx, 0.2 * (nz - 1) * dz), (0.9 * (nx - 1) * dx, 0.0)]
    src = [(float(np.float32(x)), float(np.float32(z))) for x, z in src]
    Trefs, Sq = [], []
    for x, z in src:
        T, Sw = oracle.solve_ani(S, dx, dz, x, z, Cq, mn, mx)
        Trefs.append(T)
        Sq.append(Sw)
    c = oracle.branch_counts(Sq[0], dx, dz, src[0][0], src[0][1])
    print("qP solve deciding candidates", c)
    assert c["C"] >= 20, c
    e = Eikonal(nz, nx, dz, dx)
    e.set_slowness(s)
    e.set_stiffness(cs)
    for sched in SCHEDULES:
        b = e.batch([(x, z, 0, nx, 0, 0) for x, z in src], type=ANI, schedule=sched)
        b.run()
        b.status()
        for k, Tr in enumerate(Trefs):
            assert np.array_equal(b.T(k), Tr), (sched, k)
        b.close()
    e.close()


def test_c2_full_size_bit_exact(oracle):
    w = models.c2()
    Tref, out = solve_both(oracle, w.s, w.dx, w.dz, 10000.0, 0.0)
    for sched, (T, _) in out.items():
        assert np.array_equal(T, Tref), sched


def test_c3_all_256_windows_bit_exact(oracle):
    w = models.c3()
    e = Eikonal(w.nz, w.nx, w.dz, w.dx)
    e.set_slowness(w.s)
    b = e.batch(w.shots, w.rx, w.rz)
    b.run()
    picks = b.picks()
    b.run()  # graph replay must reproduce the same bits
    picks2 = b.picks()
    b.status()
    assert np.array_equal(picks, picks2)
    sx = [s[0] for s in w.shots]
    x0 = [s[2] for s in w.shots]
    wd = [s[3] for s in w.shots]
    chunk = 32
    for k0 in range(0, len(w.shots), chunk):
        ks = range(k0, min(k0 + chunk, len(w.shots)))
        Trefs = oracle.solve_windows(w.s, w.dx, w.dz, [sx[k] for k in ks], [0.0] * len(ks),
                                     [x0[k] for k in ks], [wd[k] for k in ks], nthreads=16)
        for k, Tr in zip(ks, Trefs):
            assert np.array_equal(b.T(k), Tr), k
            _, _, xk, _, ro, nr = w.shots[k]
            pref = oracle.picks(Tr, w.dx, w.dz, w.rx[ro:ro + nr] - np.float32(xk * w.dx), w.rz[ro:ro + nr])
            assert np.array_equal(picks[ro:ro + nr], pref), k
    b.close()
    e.close()


def test_batch_rerun_after_new_slowness(oracle, rng):
    """A batch reads the handle's model at run time: after set_slowness with a
    new model, the next run (same graph) solves the new model; after a rejected
    model it refuses to run."""
    nz, nx, dh = 90, 300, 10.0
    shots = [(600.0, 0.0, 0, 160, 0, 0), (1500.0, 20.0, 100, 200, 0, 0), (2500.0, 0.0, 0, nx, 0, 0)]
    e = Eikonal(nz, nx, dh, dh)
    e.set_slowness(smooth_slowness(rng, nz, nx, dh))
    for sched in (SCHED_TILED, SCHED_PLAIN):
        b = e.batch(shots, schedule=sched)
        b.run()
        for it in range(2):
            s = smooth_slowness(rng, nz, nx, dh)
            e.set_slowness(s)
            b.run()
            b.status()
            for k, (sx, sz, x0, wd, _, _) in enumerate(shots):
                Tr = oracle.solve_windows(s, dh, dh, [sx], [sz], [x0], [wd], nthreads=1)[0]
                assert np.array_equal(b.T(k), Tr), (sched, it, k)
        bad = smooth_slowness(rng, nz, nx, dh)
        bad[7, 7] = -1.0
        with pytest.raises(capi.SeisfatError):
            e.set_slowness(bad)
        with pytest.raises(capi.SeisfatError) as ex:
            b.run()
        assert ex.value.code == capi.ESTATE
        e.set_slowness(s)
        b.close()
    e.close()


def test_batch_many_tiles_per_workgroup_bit_exact(oracle, rng):
    """More tiles than resident workgroups (every workgroup runs several, with
    the next tile's traveltimes staged while the current one runs): windows
    of a shallow model, every shot against the oracle."""
    nz, nx, dh = 130, 1500, 5.0
    s = smooth_slowness(rng, nz, nx, dh)
    nshots = 160  # ~9k tiles, about 3.5 per resident workgroup
    shots = []
    for k in range(nshots):
        wd = int(rng.integers(700, nx + 1))
        x0 = int(rng.integers(0, nx - wd + 1))
        sx = float((x0 + rng.uniform(0.0, wd - 1)) * dh)
        shots.append((sx, float(rng.uniform(0.0, (nz - 1) * dh)), x0, wd, 0, 0))
    e = Eikonal(nz, nx, dh, dh)
    e.set_slowness(s)
    b = e.batch(shots, schedule=SCHED_TILED)
    for run in range(2):
        b.run()
        b.status()
        Trefs = oracle.solve_windows(s, dh, dh, [x[0] for x in shots], [x[1] for x in shots],
                                     [x[2] for x in shots], [x[3] for x in shots], nthreads=16)
        for k, Tr in enumerate(Trefs):
            assert np.array_equal(b.T(k), Tr), (run, k)
    b.close()
    e.close()


def test_batch_fused_sweep_pairs_bit_exact(oracle, rng):
    """The tiled isotropic schedule runs the sweeps as two fused pairs with the
    shots spread over the XCDs (a second-sweep tile starts once the first-sweep
    tiles under it are written back): two sweep launches per solve, and every
    shot bit-exact against the oracle on three consecutive runs, with odd
    strip counts and windows of 1 to 3 columns among them."""
    nz, nx, dh = 97, 900, 5.0
    s = smooth_slowness(rng, nz, nx, dh)
    shots = []
    for k in range(75):
        wd = [1, 2, 3][k] if k < 3 else int(rng.integers(4, nx + 1))
        x0 = int(rng.integers(0, nx - wd + 1))
        sx = float((x0 + rng.uniform(0.0, wd - 1)) * dh)
        shots.append((sx, float(rng.uniform(0.0, (nz - 1) * dh)), x0, wd, 0, 0))
    e = Eikonal(nz, nx, dh, dh)
    e.set_slowness(s)
    b = e.batch(shots, schedule=SCHED_TILED)
    assert b.num_sweep_launches == 2
    Trefs = oracle.solve_windows(s, dh, dh, [x[0] for x in shots], [x[1] for x in shots],
                                 [x[2] for x in shots], [x[3] for x in shots], nthreads=16)
    for run in range(3):
        b.run()
        b.status()
        for k, Tr in enumerate(Trefs):
            assert np.array_equal(b.T(k), Tr), (run, k)
    b.close()
    e.close()


def test_batch_per_sweep_launches_bit_exact(oracle, rng, monkeypatch):
    """The schedule used where workgroups do not go round robin over 8 XCDs
    (SEISFAT_SWEEP_PAIRS=0 forces it): one persistent launch per sweep, every
    shot bit-exact against the oracle, isotropic and anisotropic."""
    monkeypatch.setenv("SEISFAT_SWEEP_PAIRS", "0")
    nz, nx, dh = 97, 700, 5.0
    s = smooth_slowness(rng, nz, nx, dh)
    shots = []
    for k in range(24):
        wd = int(rng.integers(1, nx + 1))
        x0 = int(rng.integers(0, nx - wd + 1))
        sx = float((x0 + rng.uniform(0.0, wd - 1)) * dh)
        shots.append((sx, float(rng.uniform(0.0, (nz - 1) * dh)), x0, wd, 0, 0))
    e = Eikonal(nz, nx, dh, dh)
    e.set_slowness(s)
    b = e.batch(shots, schedule=SCHED_TILED)
    assert b.num_sweep_launches == 4
    Trefs = oracle.solve_windows(s, dh, dh, [x[0] for x in shots], [x[1] for x in shots],
                                 [x[2] for x in shots], [x[3] for x in shots], nthreads=16)
    for run in range(2):
        b.run()
        b.status()
        for k, Tr in enumerate(Trefs):
            assert np.array_equal(b.T(k), Tr), (run, k)
    b.close()
    e.close()
    nz, nx = 161, 257
    s, cs, sxs = ani_setup(nz, nx, dh, 3)
    e = Eikonal(nz, nx, dh, dh)
    e.set_slowness(s)
    e.set_stiffness(cs)
    b = e.batch([(float(x), 0.0, 0, nx, 0, 0) for x in sxs], type=ANI, schedule=SCHED_TILED)
    assert b.num_sweep_launches == 8
    b.run()
    b.status()
    for k, x in enumerate(sxs):
        assert np.array_equal(b.T(k), oracle_ani(oracle, s, cs, dh, float(x), 0.0)), k
    b.close()
    e.close()


# FDIV_FALLBACK_H (dz = dx = 0x1.94f5a6p+2): the two-instruction division by
# dz rounds one dividend per binade away from IEEE division, so this geometry
# must run the five-instruction kernels


@pytest.mark.parametrize("dz,dx,fast", [
    (5.0, 5.0, True),
    (10.0, 8.0, None),
    (7.0, 10.0, None),
    (FDIV_FALLBACK_H, FDIV_FALLBACK_H, False),
    (FDIV_FALLBACK_H, 5.0, False),
])
def test_batch_division_variants_bit_exact(oracle, rng, dz, dx, fast):
    """Batched isotropic sweeps divide with the two-instruction known-divisor
    sequence only when all four divisors of the geometry passed the exhaustive
    check at create; both variants reproduce the oracle bit for bit."""
    nz, nx = 150, 420
    s = smooth_slowness(rng, nz, nx, min(dz, dx))
    shots = [(0.3 * (nx - 1) * dx, 0.0, 0, nx, 0, 0), (0.7 * (nx - 1) * dx, 0.6 * (nz - 1) * dz, 60, 300, 0, 0)]
    e = Eikonal(nz, nx, dz, dx)
    if fast is not None:
        assert e.fast_division == fast
    print(f"dz={dz!r} dx={dx!r}: fast division {e.fast_division}")
    e.set_slowness(s)
    b = e.batch(shots, schedule=SCHED_TILED)
    b.run()
    b.status()
    for k, (sx, sz, x0, wd, _, _) in enumerate(shots):
        Tr = oracle.solve_windows(s, dx, dz, [sx], [sz], [x0], [wd], nthreads=1)[0]
        assert np.array_equal(b.T(k), Tr), k
    b.close()
    e.close()


def ani_setup(nz, nx, dh, nshots, seed=7):
    from seisfat2d_amd.stiffness import compute_stiffness
    rng = np.random.default_rng(seed)
    vp = models.layered_model(nz, nx, dh, seed=2, nlayers=6)

    def field(lo, hi):
        g = gaussian_filter(rng.standard_normal((nx, nz)), 10.0)
        return lo + (hi - lo) * (g - g.min()) / (g.max() - g.min())

    cs = compute_stiffness(vp, vp / np.sqrt(3), 1000.0, field(0.0, 0.3), field(-0.1, 0.2), field(-30.0, 30.0))
    s = np.sqrt(1000.0 / cs[3]).astype(np.float32)
    cs = [c.astype(np.float32) for c in cs]
    sxs = (np.arange(nshots) + 0.5) * (nx - 1) * dh / nshots
    return s, cs, sxs


def oracle_ani(oracle, s, cs, dh, sx, sz):
    S = oracle.expand_boundary(s)
    Cq, mn, mx = [], [], []
    for c in cs:
        u, a, b = oracle.compress(oracle.expand_boundary(c))
        Cq.append(u)
        mn.append(a)
        mx.append(b)
    return oracle.solve_ani(S, dh, dh, sx, sz, Cq, mn, mx)[0]


def test_ani_batch_bit_exact(oracle):
    nz, nx, dh = 161, 257, 5.0
    s, cs, sxs = ani_setup(nz, nx, dh, 5)
    e = Eikonal(nz, nx, dh, dh)
    e.set_slowness(s)
    e.set_stiffness(cs)
    rx = np.arange(0, nx, 7, dtype=np.float32) * dh
    shots = [(float(x), 0.0, 0, nx, 0, 0) for x in sxs]
    shots[2] = (float(sxs[2]), 0.0, 0, nx, 0, len(rx))
    for sched in SCHEDULES:
        b = e.batch(shots, rx, np.zeros_like(rx), type=ANI, schedule=sched)
        b.run()
        pk = b.picks()
        for k, x in enumerate(sxs):
            Tr = oracle_ani(oracle, s, cs, dh, float(x), 0.0)
            assert np.array_equal(b.T(k), Tr), (sched, k)
            if k == 2:
                assert np.array_equal(pk, oracle.picks(Tr, dh, dh, rx, np.zeros_like(rx)))
        b.close()
    T1, _ = e.solve(float(sxs[1]), 0.0, type=ANI)
    assert np.array_equal(T1, oracle_ani(oracle, s, cs, dh, float(sxs[1]), 0.0))
    e.close()


@pytest.mark.parametrize("nz,nx,dz,dx", [(37, 45, 5.0, 5.0), (70, 131, 10.0, 8.0), (130, 66, 7.0, 10.0)])
def test_ani_batch_odd_shapes_bit_exact(oracle, nz, nx, dz, dx):
    """Anisotropic batches on grids narrower than one 64-column block, one
    strip, partial strips and blocks, unequal spacings and sources down to the
    bottom row: most of the qP solve's staged-slowness tiles are edge tiles."""
    s, cs, sxs = ani_setup(nz, nx, dx, 3)
    szs = [0.0, 0.4 * (nz - 1) * dz, (nz - 1) * dz]
    S = oracle.expand_boundary(s)
    Cq, mn, mx = [], [], []
    for c in cs:
        u, a, b = oracle.compress(oracle.expand_boundary(c))
        Cq.append(u)
        mn.append(a)
        mx.append(b)
    Trefs = [oracle.solve_ani(S, dx, dz, float(x), float(z), Cq, mn, mx)[0] for x, z in zip(sxs, szs)]
    e = Eikonal(nz, nx, dz, dx)
    e.set_slowness(s)
    e.set_stiffness(cs)
    for sched in SCHEDULES:
        b = e.batch([(float(x), float(z), 0, nx, 0, 0) for x, z in zip(sxs, szs)], type=ANI, schedule=sched)
        b.run()
        b.status()
        for k, Tr in enumerate(Trefs):
            assert np.array_equal(b.T(k), Tr), (sched, k)
        b.close()
    e.close()


def test_ani_vti_constant_components_bit_exact(oracle):
    """A VTI model (no tilt): C15 and C35 are zero everywhere, so their
    quantisation range is empty; every code is 1 and decodes to 0, and the
    qP solve matches the oracle bit for bit."""
    nz, nx, dh = 90, 140, 5.0
    s, cs, sxs = ani_setup(nz, nx, dh, 2)
    cs[2] = np.zeros_like(cs[2])
    cs[4] = np.zeros_like(cs[4])
    e = Eikonal(nz, nx, dh, dh)
    e.set_slowness(s)
    e.set_stiffness(cs)
    b = e.batch([(float(x), 0.0, 0, nx, 0, 0) for x in sxs], type=ANI, schedule=SCHED_TILED)
    b.run()
    b.status()
    for k, x in enumerate(sxs):
        assert np.array_equal(b.T(k), oracle_ani(oracle, s, cs, dh, float(x), 0.0)), k
    b.close()
    e.close()


def test_ani_full_size_c4_shot(oracle):
    w = models.c4(nshots=2)
    e = Eikonal(w.nz, w.nx, w.dz, w.dx)
    e.set_slowness(w.s)
    e.set_stiffness(w.stiffness)
    sx = w.shots[1][0]
    T, _ = e.solve(sx, 0.0, type=ANI)
    assert np.array_equal(T, oracle_ani(oracle, w.s, w.stiffness, w.dx, sx, 0.0))
    e.close()


def tinf_fixed(dx, dz):
    """Whether the geometry keeps an all-1e6 neighbourhood at 1e6: branch A's
    value there is at least RN(RN(RN(1e6/dz^2) + RN(1e6/dx^2)) / sum)."""
    f = np.float32
    dz2i, dx2i = f(1) / (f(dz) * f(dz)), f(1) / (f(dx) * f(dx))
    a = f(1e6) * dz2i + f(1e6) * dx2i
    return bool(a / (dz2i + dx2i) >= f(1e6))


@pytest.mark.parametrize("dh", [5.0, 7.0])
@pytest.mark.parametrize("pairs", ["1", "0"])
def test_unreached_tiles_skipped_only_where_the_geometry_keeps_1e6(oracle, monkeypatch, dh, pairs):
    """Tiles whose every input is 1e6 ("not reached": in a first sweep the
    region against the front's direction) are not run where the geometry
    provably keeps such nodes at 1e6 (dh = 5 m), and run where it does not
    (dh = 7 m: RN(RN(RN(1e6/dh^2) * 2) / (2/dh^2)) = 999999.94, so the
    reference's own arithmetic lowers them); every window bit-exact against
    the oracle either way, with the fused sweep pairs (whose second sweep also
    skips above skipped first-sweep tiles) and with one launch per sweep."""
    assert tinf_fixed(5.0, 5.0) and not tinf_fixed(7.0, 7.0)
    monkeypatch.setenv("SEISFAT_SWEEP_PAIRS", pairs)
    rng = np.random.default_rng(31)
    nz, nx = 300, 700
    s = smooth_slowness(rng, nz, nx, dh)
    shots = []
    for k in range(6):
        w = int(rng.integers(200, nx + 1))
        x0 = int(rng.integers(0, nx - w + 1))
        shots.append((float(np.float32((x0 + rng.uniform(0.3, 0.7) * (w - 1)) * dh)),
                      float(np.float32(rng.uniform(0.0, 0.5) * (nz - 1) * dh)), x0, w, 0, 0))
    e = Eikonal(nz, nx, dh, dh)
    e.set_slowness(s)
    b = e.batch(shots, schedule=SCHED_TILED)
    Trefs = oracle.solve_windows(s, dh, dh, [x[0] for x in shots], [x[1] for x in shots],
                                 [x[2] for x in shots], [x[3] for x in shots], nthreads=16)
    for run in range(2):
        b.run()
        b.status()
        total, skipped = b.tile_counts()
        print(f"dh={dh} pairs={pairs}: {skipped} of {total} tiles skipped")
        if tinf_fixed(dh, dh):
            assert skipped > 0.05 * total
        else:
            assert skipped == 0
        for k, Tr in enumerate(Trefs):
            assert np.array_equal(b.T(k), Tr), (run, k)
    b.close()
    e.close()


def test_window_at_the_far_end_of_a_model_with_skewed_copies_over_2_gib(oracle, monkeypatch):
    """A 8000 x 70000 model, whose skewed slowness copies (one per sweep
    orientation) each exceed 2^31 bytes: a narrow window at the model's right
    edge reads them at their largest offsets (beyond a 32-bit byte offset from
    the copy's start), bit-exact against the oracle with the fused sweep pairs
    and with one persistent launch per sweep."""
    nz, nx, dh = 8000, 70000, 10.0
    nzz, nxx = nz + 6, nx + 6
    KP = 64 * ((nzz - 2 + 63) // 64) + 64
    NS = nxx + 2 * 160
    assert NS * KP * 4 > 2 ** 31
    z = np.arange(nz, dtype=np.float32) * dh
    x = np.arange(nx, dtype=np.float32) * dh
    s = np.empty((nx, nz), np.float32)
    for j0 in range(0, nx, 10000):          # 1/v = 1/(1500 + 0.4 z + 300 sin(x / 2 km)), in column blocks
        xs = x[j0:j0 + 10000, None]
        s[j0:j0 + 10000] = (1.0 / (1500.0 + 0.4 * z[None, :] + 300.0 * np.sin(xs / 2000.0))).astype(np.float32)
    w, x0 = 300, nx - 300
    sx, sz = float(np.float32((x0 + 123) * dh)), 2500.0
    Tr = oracle.solve_windows(s, dh, dh, [sx], [sz], [x0], [w], nthreads=1)[0]
    for pairs in ("1", "0"):
        monkeypatch.setenv("SEISFAT_SWEEP_PAIRS", pairs)
        e = Eikonal(nz, nx, dh, dh)
        e.set_slowness(s)
        b = e.batch([(sx, sz, x0, w, 0, 0)], schedule=SCHED_TILED)
        b.run()
        b.status()
        T = b.T(0)
        b.close()
        e.close()
        assert np.array_equal(T, Tr), (pairs, np.abs(T - Tr).max())


def test_ani_c4_benched_batch_all_128_shots_bit_exact(oracle):
    """The benchmarked anisotropic batch itself: 128 full-model TTI shots on
    801x3001, every shot's traveltimes against the oracle's
    Eikonal_ANI::time_propagation (shots solved on host threads)."""
    from concurrent.futures import ThreadPoolExecutor
    w = models.c4()
    e = Eikonal(w.nz, w.nx, w.dz, w.dx)
    e.set_slowness(w.s)
    e.set_stiffness(w.stiffness)
    b = e.batch(w.shots, w.rx, w.rz, type=ANI, schedule=SCHED_TILED)
    b.run()
    picks = b.picks()
    b.status()
    S = oracle.expand_boundary(w.s)
    Cq, mn, mx = [], [], []
    for c in w.stiffness:
        u, a, bb = oracle.compress(oracle.expand_boundary(c))
        Cq.append(u)
        mn.append(a)
        mx.append(bb)

    def ref(k):
        return oracle.solve_ani(S, w.dx, w.dz, w.shots[k][0], 0.0, Cq, mn, mx)[0]

    with ThreadPoolExecutor(max_workers=16) as ex:
        for k, Tr in enumerate(ex.map(ref, range(len(w.shots)))):
            assert np.array_equal(b.T(k), Tr), k
            _, _, _, _, ro, nr = w.shots[k]
            assert np.array_equal(picks[ro:ro + nr], oracle.picks(Tr, w.dx, w.dz, w.rx[ro:ro + nr], w.rz[ro:ro + nr])), k
    b.close()
    e.close()


def test_invalid_inputs_rejected():
    e = Eikonal(20, 30, 10.0, 10.0)
    s = np.full((30, 20), 5e-4, np.float32)
    s[3, 4] = 0.0
    with pytest.raises(capi.SeisfatError) as ex:
        e.set_slowness(s)
    assert ex.value.code == capi.EINVAL
    s[3, 4] = np.nan
    with pytest.raises(capi.SeisfatError):
        e.set_slowness(s)
    s[3, 4] = 1e-20  # slowness*spacing below the solver's operating range
    with pytest.raises(capi.SeisfatError):
        e.set_slowness(s)
    with pytest.raises(capi.SeisfatError) as ex:  # no valid slowness yet
        e.solve(10.0, 10.0)
    assert ex.value.code == capi.ESTATE
    e.set_slowness(np.full((30, 20), 5e-4, np.float32))
    for sx, sz in ((-1.0, 0.0), (290.5, 0.0), (10.0, 191.0)):
        with pytest.raises(capi.SeisfatError) as ex:
            e.solve(sx, sz)
        assert ex.value.code == capi.EINVAL
    with pytest.raises(capi.SeisfatError) as ex:
        e.solve(10.0, 10.0, rx=[400.0], rz=[0.0])
    assert ex.value.code == capi.EINVAL
    with pytest.raises(capi.SeisfatError) as ex:
        e.solve(10.0, 10.0, type=ANI)
    assert ex.value.code == capi.ESTATE
    with pytest.raises(capi.SeisfatError) as ex:  # window outside the model
        e.batch([(10.0, 0.0, 25, 10, 0, 0)])
    assert ex.value.code == capi.EINVAL
    with pytest.raises(capi.SeisfatError) as ex:  # source outside its window
        e.batch([(10.0, 0.0, 5, 10, 0, 0)])
    assert ex.value.code == capi.EINVAL
    e.close()


def test_device_slowness_from_torch(oracle, rng):
    import torch
    nz, nx, dh = 70, 90, 10.0
    s = smooth_slowness(rng, nz, nx, dh)
    e = Eikonal(nz, nx, dh, dh)
    e.set_slowness(torch.from_numpy(s).cuda())
    torch.cuda.synchronize()
    T, _ = e.solve(200.0, 50.0)
    assert np.array_equal(T, oracle.solve_iso(oracle.expand_boundary(s), dh, dh, 200.0, 50.0))
    bad = torch.from_numpy(s).cuda()
    bad[5, 5] = -1.0
    with pytest.raises(capi.SeisfatError):
        e.set_slowness(bad)
    e.close()


def test_device_T_pointer_matches_copy():
    nz, nx, dh = 40, 50, 10.0
    e = Eikonal(nz, nx, dh, dh)
    e.set_slowness(np.full((nx, nz), 4e-4, np.float32))
    T, _ = e.solve(100.0, 0.0)
    p = e.device_T_ptr()
    assert p
    import torch
    buf = torch.empty(T.size, dtype=torch.float32, device="cuda")
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    assert hip.hipMemcpy(C.c_void_p(buf.data_ptr()), C.c_void_p(p), T.nbytes, 3) == 0  # device to device
    assert np.array_equal(buf.cpu().numpy().reshape(T.shape), T)
    e.close()
