"""Host side of the migration module (CPU): the gather sizes, the trace list
and the Ricker wavelet the library evaluates on the host, against the
reference's definitions (src/migration/migration.cu:54-99, KDM.cu:14-53) on
worked examples and against the oracle."""
import numpy as np
import pytest

from seisfat2d_amd import migration as mig


def test_gathers_of_the_sample_parameters():
    # sources every 100 m, receivers every 25 m: dCMP = 12.5 m; CMPs 0..350 m
    xs = np.arange(4) * 100.0
    xr = np.arange(17) * 25.0
    nang, dcmp, min_cmp, ncmp = mig.gathers(xs, xr, da=2.0, max_angle=60.0)
    assert nang == 31                      # run/parameters.txt: mig_max_angle 60, mig_angle_spacing 2
    assert dcmp == np.float32(12.5) and min_cmp == 0.0
    assert ncmp == 29                      # (350 - 0) / 12.5 + 1


def test_gathers_single_source_and_rejections():
    nang, dcmp, min_cmp, ncmp = mig.gathers([500.0], np.arange(5) * 40.0 + 10.0, da=3.0, max_angle=10.0)
    assert (nang, dcmp, ncmp) == (4, np.float32(20.0), 5)     # CMPs 255..335 m every 20 m
    assert min_cmp == np.float32(255.0)
    with pytest.raises(ValueError):
        mig.gathers([0.0], [0.0], 1.0, 10.0)
    with pytest.raises(ValueError):
        mig.gathers([100.0, 0.0], [0.0, 10.0, 20.0], 1.0, 10.0)   # decreasing sources: negative spacing


def test_trace_list_order_offsets_and_bins():
    xs = np.array([0.0, 100.0, 200.0])
    xr = np.arange(9) * 25.0
    nang, dcmp, min_cmp, ncmp = mig.gathers(xs, xr, 2.0, 60.0)
    src, rec, cmp, cid = mig.trace_list(xs, xr, 60.0, min_cmp, dcmp)
    want = [(s, r) for s in range(3) for r in range(9) if abs(xs[s] - xr[r]) < 60.0]
    assert list(zip(src.tolist(), rec.tolist())) == want          # sources outer, receivers inner
    assert np.array_equal(cmp, (0.5 * (xs[src] + xr[rec])).astype(np.float32))
    assert np.array_equal(cid, np.floor(cmp / 12.5).astype(np.int32))
    assert cid.min() >= 0 and cid.max() < ncmp
    # an offset equal to max_offset is left out (strict <)
    s2, _, _, _ = mig.trace_list(xs, xr, 50.0, min_cmp, dcmp)
    assert len(s2) == sum(1 for s in range(3) for r in range(9) if abs(xs[s] - xr[r]) < 50.0)


@pytest.mark.parametrize("dt,fmax", [(1e-3, 30.0), (5e-4, 45.0), (2e-3, 12.5), (1e-3, 400.0)])
def test_library_ricker_equals_the_oracle(oracle, dt, fmax):
    """seisfat_ricker runs on the host (no GPU call) with the reference's
    float arithmetic: bit-identical to the oracle."""
    w = mig.ricker(dt, fmax)
    assert np.array_equal(w, oracle.ricker(dt, fmax))
    assert len(w) == 2 * int((2 * np.sqrt(np.pi) / fmax - 0.5 * dt) / dt) or len(w) == 0


def test_library_ricker_rejects_bad_arguments():
    from seisfat2d_amd.capi import SeisfatError
    with pytest.raises(SeisfatError):
        mig.ricker(0.0, 30.0)
    with pytest.raises(SeisfatError):
        mig.ricker(1e-3, -1.0)
