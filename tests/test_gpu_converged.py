"""Converged mode on the GPU (seisfat_batch_run_converged): passes of the
four sweeps repeat until one leaves every traveltime unchanged.  Every shot's
traveltimes equal the oracle's converged mode (orc_fsm mode 1: passes until
the grid is unchanged) bit for bit, the pass count equals the oracle's
(the batch stops when its slowest shot has converged), and a pass limit
stops it early with the limited passes' traveltimes."""
import numpy as np
import pytest

from seisfat2d_amd import SCHED_TILED, Eikonal, models

pytestmark = pytest.mark.gpu
NB = 3


def oracle_conv(oracle, s, dx, dz, shot, max_passes):
    sx, sz, x0, w = shot[:4]
    S = oracle.expand_boundary(np.ascontiguousarray(s[x0:x0 + w]))
    return oracle.solve_iso(S, dx, dz, sx - x0 * dx, sz, mode=1, max_passes=max_passes)


@pytest.mark.parametrize("case", ["c1", "rough_dz3_dx10", "rough_windows"])
def test_converged_mode_matches_oracle(oracle, case):
    if case == "c1":
        w = models.c1()
        s, dx, dz = w.s, w.dx, w.dz
        shots = [(5000.0, 0.0, 0, w.nx, 0, 0), (1234.5, 700.0, 0, w.nx, 0, 0)]
    else:
        m = models.rough(64, 96, 3.0, 10.0, seed=11)
        s, dx, dz = m.s, m.dx, m.dz
        if case == "rough_dz3_dx10":
            shots = [(m.sx, m.sz, 0, m.nx, 0, 0)]
        else:
            rng = np.random.default_rng(3)
            shots = []
            for k in range(7):
                wd = int(rng.integers(2, m.nx + 1))
                x0 = int(rng.integers(0, m.nx - wd + 1))
                shots.append((float(np.float32((x0 + rng.uniform(0, wd - 1)) * dx)),
                              float(np.float32(rng.uniform(0, (m.nz - 1) * dz))), x0, wd, 0, 0))
    nx, nz = s.shape
    e = Eikonal(nz, nx, dz, dx)
    e.set_slowness(s)
    b = e.batch(shots, schedule=SCHED_TILED)
    refs = [oracle_conv(oracle, s, dx, dz, sh, 200) for sh in shots]
    want = max(p for _, p in refs)
    for chunk in (4, 7):
        passes = b.run_converged(max_passes=200, chunk=chunk)
        print(f"{case}: {passes} passes (oracle per shot {[p for _, p in refs]})")
        assert passes == want
        for k, (Tr, _) in enumerate(refs):
            assert np.array_equal(b.T(k), Tr), (case, chunk, k)
    # a pass limit: the traveltimes after exactly that many passes
    lim = 2
    assert b.run_converged(max_passes=lim, chunk=4) == lim
    for k, sh in enumerate(shots):
        Tr, p = oracle_conv(oracle, s, dx, dz, sh, lim)
        assert np.array_equal(b.T(k), Tr), (case, "limit", k)
    # and the single-pass run still gives the reference's one pass
    b.run()
    b.status()
    for k, sh in enumerate(shots):
        sx, sz, x0, wd = sh[:4]
        Tr = oracle.solve_windows(s, dx, dz, [sx], [sz], [x0], [wd], nthreads=1)[0]
        assert np.array_equal(b.T(k), Tr), (case, "one pass", k)
    b.close()
    e.close()


def test_converged_mode_rejects_other_schedules():
    from seisfat2d_amd import SCHED_PLAIN, capi
    e = Eikonal(20, 30, 10.0, 10.0)
    e.set_slowness(np.full((30, 20), 5e-4, np.float32))
    b = e.batch([(100.0, 0.0, 0, 30, 0, 0)], schedule=SCHED_PLAIN)
    with pytest.raises(capi.SeisfatError) as ex:
        b.run_converged()
    assert ex.value.code == capi.EINVAL
    b.close()
    e.close()
