"""Schedule options on the GPU: every combination of seisfat_sched_opts
gives the oracle's traveltimes bit for bit (they change only how tiles are
launched and ordered), the batch reports the options it runs with, and a
tuning run on a small batch picks options from the search space, caches them
and serves them to opts="tuned"."""
import itertools

import numpy as np
import pytest

from seisfat2d_amd import SCHED_TILED, Eikonal, models, tune

pytestmark = pytest.mark.gpu


def window_batch(seed=5, nz=96, nx=160, dz=4.0, dx=6.0, n=7):
    m = models.rough(nz, nx, dz, dx, seed=seed)
    rng = np.random.default_rng(seed)
    shots = []
    for k in range(n):
        wd = int(rng.integers(40, nx + 1)) if k else nx
        x0 = int(rng.integers(0, nx - wd + 1))
        shots.append((float(np.float32((x0 + rng.uniform(0.0, wd - 1)) * dx)),
                      float(np.float32(rng.uniform(0.0, 0.4) * (nz - 1) * dz)), x0, wd, 0, 0))
    return m, shots


@pytest.mark.parametrize("dh", [None, 5.0])
def test_every_option_combination_bit_exact(oracle, dh):
    m, shots = window_batch()
    s, dx, dz = m.s, m.dx, m.dz
    if dh:   # equal spacing with the not-reached skip in effect
        dx = dz = dh
        shots = [(float(np.float32(x / m.dx * dh)), float(np.float32(z / m.dz * dh)), x0, w, o, n)
                 for x, z, x0, w, o, n in shots]
    nx, nz = s.shape
    Trefs = oracle.solve_windows(s, dx, dz, [x[0] for x in shots], [x[1] for x in shots],
                                 [x[2] for x in shots], [x[3] for x in shots], nthreads=16)
    e = Eikonal(nz, nx, dz, dx)
    e.set_slowness(s)
    for pairs, wpcu, skip, prio in itertools.product((1, 0), (0, 3), (1, 0), (1, 0)):
        opts = {"pairs": pairs, "wpcu": wpcu, "skip": skip, "priority": prio}
        b = e.batch(shots, schedule=SCHED_TILED, opts=opts)
        assert b.opts == opts, (b.opts, opts)
        assert b.num_sweep_launches == (2 if pairs else 4)
        b.run()
        b.status()
        total, skipped = b.tile_counts()
        assert skipped <= total and (skip or skipped == 0), (opts, skipped)
        for k, Tr in enumerate(Trefs):
            assert np.array_equal(b.T(k), Tr), (opts, k)
        b.close()
    e.close()


def test_tuning_run_caches_and_serves_the_options(oracle, tmp_path, monkeypatch):
    monkeypatch.setenv("SEISFAT_TUNE_CACHE", str(tmp_path / "schedule.json"))
    m, shots = window_batch(seed=9, n=12)
    nx, nz = m.s.shape
    e = Eikonal(nz, nx, m.dz, m.dx)
    e.set_slowness(m.s)
    r = tune.tune(e, shots, runs=3, warmup=1)
    print("tuned:", r["opts"], f"{r['ms']:.3f} ms after {len(r['trials'])} trials")
    assert not r["cached"] and len(r["trials"]) >= 2
    assert all(r["opts"][k] in tune.SPACE[k] for k in tune.ORDER)
    assert r["ms"] == min(t["ms"] for t in r["trials"] if t["opts"] == r["opts"])
    r2 = tune.tune(e, shots)
    assert r2["cached"] and r2["opts"] == r["opts"]
    b = e.batch(shots, schedule=SCHED_TILED, opts="tuned")
    want = dict(r["opts"])
    assert b.opts == want
    b.run()
    b.status()
    Trefs = oracle.solve_windows(m.s, m.dx, m.dz, [x[0] for x in shots], [x[1] for x in shots],
                                 [x[2] for x in shots], [x[3] for x in shots], nthreads=16)
    for k, Tr in enumerate(Trefs):
        assert np.array_equal(b.T(k), Tr), k
    b.close()
    e.close()


def test_every_option_combination_bit_exact_anisotropic(oracle):
    """The qP solves (slowness staged per shot, both solves through the same
    schedule) under every option combination, against the oracle's
    Eikonal_ANI::time_propagation."""
    from seisfat2d_amd import ANI
    from seisfat2d_amd.stiffness import compute_stiffness
    nz, nx, dh = 70, 110, 5.0
    rng = np.random.default_rng(23)
    vp = rng.uniform(1500.0, 4500.0, (nx, nz))
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
    src = [(float(np.float32(f * (nx - 1) * dh)), float(np.float32(g * (nz - 1) * dh)))
           for f, g in ((0.37, 0.2), (0.9, 0.0), (0.05, 0.6))]
    Trefs = [oracle.solve_ani(S, dh, dh, x, z, Cq, mn, mx)[0] for x, z in src]
    e = Eikonal(nz, nx, dh, dh)
    e.set_slowness(s)
    e.set_stiffness(cs)
    for pairs, wpcu, skip, prio in itertools.product((1, 0), (0, 3), (1, 0), (1, 0)):
        opts = {"pairs": pairs, "wpcu": wpcu, "skip": skip, "priority": prio}
        b = e.batch([(x, z, 0, nx, 0, 0) for x, z in src], type=ANI, schedule=SCHED_TILED, opts=opts)
        assert b.opts == opts
        b.run()
        b.status()
        for k, Tr in enumerate(Trefs):
            assert np.array_equal(b.T(k), Tr), (opts, k)
        b.close()
    e.close()
