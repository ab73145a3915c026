"""Kirchhoff depth migration on the GPU against the oracle.

Bars (the GPU forms each weight from per-table amplitudes and half-angle
cosines and sines, within a few ulp of the reference's order of roundings,
and its acosf / expf differ from the host's by an ulp or two; traveltime sums,
time indices, the convolutions and the trace order of the image-domain
adjoint are exact):
  - convolutions: bit-identical (the same double sums in the same order);
  - image-domain adjoint: |m_gpu - m_oracle| <= 1e-5 * max |m_oracle|;
  - forward (LDS accumulation order differs from the reference's atomics):
    per sample |d_gpu - d_oracle| <= 1e-5 * (L |m|) + 1e-30, L |m| the
    oracle's forward of |m|;
  - angle domain: the same bars on all but 0.1 % of the entries (an angle
    within an ulp of a bin edge may land in the neighbouring bin), and the
    gather summed over angles within the image-domain bar;
  - the migration loops: KDM as the adjoint; LSKDM residual norms within
    1e-4 relative (float64 sums here, float32 sums in trace order in the
    reference), the image within 1e-4 of its largest value, the angle
    gather within 1e-3 in norm after three iterations.
"""
import numpy as np
import pytest

from seisfat2d_amd import Eikonal, migration as mig, models

pytestmark = pytest.mark.gpu

NB = 3


@pytest.fixture(scope="module")
def torch():
    import torch
    return torch


def survey(nz=48, nx=84, dz=10.0, dx=10.0, nsrc=4, nrec=14, seed=1):
    m = models.rough(nz, nx, dz, dx, seed=seed, vlo=1800.0, vhi=3500.0, sigma=3.0)
    xs = np.linspace(40.0, (nx - 1) * dx - 40.0, nsrc).astype(np.float32)
    xr = np.linspace(0.0, (nx - 1) * dx, nrec).astype(np.float32)
    return m, xs, xr


def make(e, m, xs, xr, domain, nt=420, dt=2e-3, max_offset=500.0, aperture=30.0, da=3.0, max_angle=60.0,
         tables=None):
    return mig.Kirchhoff(e, xs, np.zeros_like(xs), xr, np.zeros_like(xr), nt, dt, aperture, max_offset, 30.0,
                         domain=domain, da=da, max_angle=max_angle, tables=tables)


@pytest.fixture(scope="module")
def setup():
    m, xs, xr = survey()
    e = Eikonal(m.s.shape[1], m.s.shape[0], m.dz, m.dx)
    e.set_slowness(m.s)
    ts = e.traveltime_tables(xs, np.zeros_like(xs))
    tr = e.traveltime_tables(xr, np.zeros_like(xr))
    yield dict(m=m, xs=xs, xr=xr, e=e, ts=ts, tr=tr)
    e.close()


def orc_args(oracle, K, st):
    S = oracle.expand_boundary(st["m"].s)
    return S, dict(dt=float(K.dt), dx=st["m"].dx, dz=st["m"].dz, aperture=30.0, da=float(K.da), nang=K.nang)


def test_tables_are_the_oracle_solves(oracle, setup):
    st = setup
    S = oracle.expand_boundary(st["m"].s)
    for k in (0, 3):
        T = oracle.solve_iso(S, st["m"].dx, st["m"].dz, float(st["xs"][k]), 0.0)
        assert np.array_equal(st["ts"][k], T)


@pytest.mark.parametrize("adjoint", [False, True])
@pytest.mark.parametrize("nt,dt", [(1001, 1e-3), (64, 2e-3)])
def test_convolution_bit_exact(oracle, setup, torch, adjoint, nt, dt):
    st = setup
    K = make(st["e"], st["m"], st["xs"], st["xr"], mig.IMAGE, nt=nt, dt=dt, tables=(st["ts"], st["tr"]))
    rng = np.random.default_rng(nt)
    x = rng.standard_normal((K.ntr, nt)).astype(np.float32)
    got = K.convolve(torch.from_numpy(x).cuda(), adjoint).cpu().numpy()
    assert np.array_equal(got, oracle.trace_convolve(x, K.wavelet, adjoint))
    K.close()


@pytest.mark.parametrize("adjoint", [False, True])
@pytest.mark.parametrize("nt,dt,fmax", [(1001, 1e-3, 2.0),       # wavelet (3542) longer than the trace, in LDS
                                        (16000, 1e-3, 30.0),     # trace and wavelet just fit LDS
                                        (17000, 1e-3, 30.0)])    # neither: the global-memory kernel
def test_convolution_bit_exact_every_kernel(oracle, torch, adjoint, nt, dt, fmax):
    import ctypes as C
    from seisfat2d_amd import capi
    w = mig.ricker(dt, fmax)
    rng = np.random.default_rng(nt)
    x = rng.standard_normal((3, nt)).astype(np.float32)
    xi = torch.from_numpy(x).cuda()
    out = torch.empty_like(xi)
    capi.check(capi.lib().seisfat_trace_convolve(int(adjoint), 3, nt, capi._ptr(w), len(w), C.c_void_p(xi.data_ptr()),
                                                 C.c_void_p(out.data_ptr()), None))
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), oracle.trace_convolve(x, w, adjoint))


# (aperture, dt, nt): the sample case; a negative aperture (negative taper
# divisors: IEEE division for (x - cmp) / sigma); a time step above 1 s (the
# time index by IEEE division below 2^-100); the longest record
@pytest.mark.parametrize("aperture,dt,nt", [(30.0, 2e-3, 420), (-30.0, 2e-3, 420), (30.0, 1.25, 3),
                                              (30.0, 5e-5, 16384)])
def test_image_adjoint_matches_oracle(oracle, setup, torch, aperture, dt, nt):
    st = setup
    K = make(st["e"], st["m"], st["xs"], st["xr"], mig.IMAGE, nt=nt, dt=dt, aperture=aperture,
             tables=(st["ts"], st["tr"]))
    assert K.ntr > 20
    S, op = orc_args(oracle, K, st)
    op["aperture"] = aperture
    rng = np.random.default_rng(5)
    d = rng.standard_normal((K.ntr, K.nt)).astype(np.float32)
    m0 = rng.standard_normal(K.model_shape).astype(np.float32) * 1e-3
    got = K.adjoint(torch.from_numpy(d).cuda(), torch.from_numpy(m0.copy()).cuda()).cpu().numpy()
    want = oracle.kdm(0, True, S, st["ts"], st["tr"], K.src, K.rec, K.cmp, None, d, m0.copy(), **op)
    assert np.abs(want).max() > 1e-3
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
    K.close()


@pytest.mark.parametrize("nt", [420, 5000, 9000, 16384])   # 16384: the largest record (a 64 KB LDS histogram)
@pytest.mark.parametrize("domain", [mig.IMAGE, mig.ANGLE])
def test_forward_matches_oracle(oracle, setup, torch, domain, nt):
    st = setup
    K = make(st["e"], st["m"], st["xs"], st["xr"], domain, nt=nt, tables=(st["ts"], st["tr"]))
    S, op = orc_args(oracle, K, st)
    rng = np.random.default_rng(7 + domain)
    m = rng.standard_normal(K.model_shape).astype(np.float32)
    d0 = rng.standard_normal((K.ntr, K.nt)).astype(np.float32) * 1e-4
    cid = K.cmp_id if domain == mig.ANGLE else None
    got = K.forward(torch.from_numpy(m).cuda(), torch.from_numpy(d0.copy()).cuda()).cpu().numpy()
    want = oracle.kdm(domain, False, S, st["ts"], st["tr"], K.src, K.rec, K.cmp, cid, d0.copy(), m, **op)
    bound = oracle.kdm(domain, False, S, st["ts"], st["tr"], K.src, K.rec, K.cmp, cid, np.abs(d0), np.abs(m), **op)
    bad = np.abs(got - want) > 1e-5 * bound + 1e-30
    assert np.abs(want).max() > 0
    if domain == mig.IMAGE:
        assert not bad.any(), np.argwhere(bad)[:5]
    else:
        assert bad.mean() <= 1e-3, bad.sum()
    K.close()


def test_angle_adjoint_matches_oracle(oracle, setup, torch):
    st = setup
    K = make(st["e"], st["m"], st["xs"], st["xr"], mig.ANGLE, tables=(st["ts"], st["tr"]))
    S, op = orc_args(oracle, K, st)
    assert K.model_shape == (K.ncmp, K.nang, 48) and K.nang == 21
    rng = np.random.default_rng(9)
    d = rng.standard_normal((K.ntr, K.nt)).astype(np.float32)
    got = K.adjoint(torch.from_numpy(d).cuda()).cpu().numpy()
    want = oracle.kdm(1, True, S, st["ts"], st["tr"], K.src, K.rec, K.cmp, K.cmp_id, d,
                      np.zeros(K.model_shape, np.float32), **op)
    bound = oracle.kdm(1, True, S, st["ts"], st["tr"], K.src, K.rec, K.cmp, K.cmp_id, np.abs(d),
                       np.zeros(K.model_shape, np.float32), **op)
    assert np.abs(want).max() > 0
    bad = np.abs(got - want) > 1e-5 * bound + 1e-30
    assert bad.mean() <= 1e-3, bad.sum()
    g, w = got.sum(axis=1, dtype=np.float64), want.sum(axis=1, dtype=np.float64)
    assert np.abs(g - w).max() <= 1e-5 * bound.sum(axis=1).max()
    K.close()


@pytest.mark.parametrize("domain", [mig.IMAGE, mig.ANGLE])
def test_dot_product(setup, torch, domain):
    """<L m, d> = <m, L^T d>: the reference's adjointness test of its
    operators (src/migration/migration.cu:282-396, without the convolutions).
    Both sides sum products of random signs, so they cancel: the float32
    rounding of L m and L^T d is bounded against <L|m|, |d|> (the sum of the
    terms' magnitudes), not against |a| + |b|."""
    st = setup
    K = make(st["e"], st["m"], st["xs"], st["xr"], domain, tables=(st["ts"], st["tr"]))
    g = torch.Generator(device="cuda").manual_seed(3)
    m = torch.randn(K.model_shape, device="cuda", generator=g)
    d = torch.randn((K.ntr, K.nt), device="cuda", generator=g)
    Lm = K.forward(m)
    Ltd = K.adjoint(d)
    a = torch.sum(Lm.double() * d.double()).item()
    b = torch.sum(m.double() * Ltd.double()).item()
    scale = torch.sum(K.forward(m.abs()).double() * d.abs().double()).item()
    assert scale > 0 and abs(a - b) <= 1e-5 * scale, (a, b, scale)
    K.close()


@pytest.mark.parametrize("domain", [mig.IMAGE, mig.ANGLE])
def test_migration_matches_oracle(oracle, setup, torch, domain):
    """KDM::kirchhoff_depth_migration: adjoint convolution of every trace,
    then the adjoint operator (IDKDM / ADKDM)."""
    st = setup
    K = make(st["e"], st["m"], st["xs"], st["xr"], domain, tables=(st["ts"], st["tr"]))
    S, op = orc_args(oracle, K, st)
    rng = np.random.default_rng(13)
    seismic = rng.standard_normal((K.nsrc, K.nrec, K.nt)).astype(np.float32)
    got = K.migrate(seismic).cpu().numpy()
    d = np.ascontiguousarray(oracle.trace_convolve(seismic[K.src, K.rec], K.wavelet, True))
    cid = K.cmp_id if domain == mig.ANGLE else None
    want = oracle.kdm(domain, True, S, st["ts"], st["tr"], K.src, K.rec, K.cmp, cid, d,
                      np.zeros(K.model_shape, np.float32), **op)
    bound = oracle.kdm(domain, True, S, st["ts"], st["tr"], K.src, K.rec, K.cmp, cid, np.abs(d),
                       np.zeros(K.model_shape, np.float32), **op)
    bad = np.abs(got - want) > 1e-5 * bound + 1e-30
    assert bad.mean() <= (0 if domain == mig.IMAGE else 1e-3), bad.sum()
    K.close()


@pytest.mark.parametrize("domain,smooth", [(mig.IMAGE, 0.0), (mig.ANGLE, 1e-3)])
def test_least_squares_migration_matches_oracle(oracle, setup, torch, domain, smooth):
    st = setup
    K = make(st["e"], st["m"], st["xs"], st["xr"], domain, tables=(st["ts"], st["tr"]))
    S, op = orc_args(oracle, K, st)
    # consistent data: the forward model of a reflector
    cid = K.cmp_id if domain == mig.ANGLE else None
    m_true = np.zeros(K.model_shape, np.float32)
    m_true[..., 30] = 1.0
    d = oracle.kdm(domain, False, S, st["ts"], st["tr"], K.src, K.rec, K.cmp, cid,
                   np.zeros((K.ntr, K.nt), np.float32), m_true, **op)
    dobs = oracle.trace_convolve(d, K.wavelet, False)
    seismic = np.zeros((K.nsrc, K.nrec, K.nt), np.float32)
    seismic[K.src, K.rec] = dobs
    got, res = K.lsm(seismic, max_it=3, smooth=smooth)
    want, wres = oracle.lskdm(domain, S, st["ts"], st["tr"], K.src, K.rec, K.cmp, cid, dobs, K.wavelet, 3, smooth,
                              op["dt"], op["dx"], op["dz"], op["aperture"], K.model_shape, op["da"], op["nang"])
    got = got.cpu().numpy()
    assert len(res) == len(wres) == 4
    assert np.allclose(res, wres, rtol=1e-4, atol=0), (res, wres)
    assert res[-1] < res[0]
    if domain == mig.IMAGE:
        assert np.abs(got - want).max() <= 1e-4 * np.abs(want).max()
    else:   # a contribution moved to the neighbouring angle bin spreads over the later iterations
        assert np.linalg.norm(got - want) <= 1e-3 * np.linalg.norm(want)
    K.close()


def test_argument_errors(setup, torch):
    import ctypes as C
    from seisfat2d_amd import capi
    st = setup
    K = make(st["e"], st["m"], st["xs"], st["xr"], mig.IMAGE, tables=(st["ts"], st["tr"]))
    with pytest.raises(ValueError):
        K.adjoint(torch.zeros((K.ntr, K.nt + 1), device="cuda"))
    with pytest.raises(ValueError):
        K.forward(torch.zeros((3, 3), device="cuda"))
    with pytest.raises(ValueError):
        make(st["e"], st["m"], st["xs"], st["xr"], mig.IMAGE, tables=(st["ts"][:2], st["tr"]))
    # trace indices out of range, a table index out of range, operators before any table
    bad = np.array([0, 99], np.int32)
    rc = capi.lib().seisfat_kdm_set_traces(K._k, 2, capi._ptr(bad), capi._ptr(np.zeros(2, np.int32)),
                                           capi._ptr(np.zeros(2, np.float32)), None)
    assert rc == capi.EINVAL and "out of range" in capi.lib().seisfat_last_error().decode()
    assert capi.lib().seisfat_kdm_set_table(K._k, 1, K.nrec, C.c_void_p(1), None) == capi.EINVAL
    K.close()
    e = st["e"]
    for params, msg in (((mig.IMAGE, 0, 1e-3, 30.0, 1.0, 1, 1), "nt"), ((7, 100, 1e-3, 30.0, 1.0, 1, 1), "domain"),
                        ((mig.ANGLE, 100, 1e-3, 30.0, 1.0, 0, 4), "angle domain"),
                        ((mig.IMAGE, 100, -1e-3, 30.0, 1.0, 1, 1), "dt")):
        h = C.c_void_p()
        p = capi.KdmParams(*params)
        assert capi.lib().seisfat_kdm_create(e._h, C.byref(p), 2, 3, C.byref(h)) == capi.EINVAL
        assert msg in capi.lib().seisfat_last_error().decode()
    h = C.c_void_p()
    p = capi.KdmParams(mig.IMAGE, 100, 1e-3, 30.0, 1.0, 1, 1)
    capi.check(capi.lib().seisfat_kdm_create(e._h, C.byref(p), 2, 3, C.byref(h)))
    d = torch.zeros((1, 100), device="cuda")
    m = torch.zeros((e.nx, e.nz), device="cuda")
    assert capi.lib().seisfat_kdm_adjoint(h, C.c_void_p(d.data_ptr()), C.c_void_p(m.data_ptr()), None) == capi.ESTATE
    capi.lib().seisfat_kdm_destroy(h)


@pytest.mark.parametrize("domain", [mig.IMAGE, mig.ANGLE])
def test_adjoint_over_several_block_parts(oracle, setup, torch, domain):
    """13 sources: four source blocks, which the adjoint spreads over several
    waves per chunk (partial images added in order in the image domain)."""
    st = setup
    e = st["e"]
    xs = np.linspace(20.0, 810.0, 13).astype(np.float32)
    ts = e.traveltime_tables(xs, np.zeros_like(xs))
    K = make(e, st["m"], xs, st["xr"], domain, tables=(ts, st["tr"]))
    S, op = orc_args(oracle, K, st)
    cid = K.cmp_id if domain == mig.ANGLE else None
    rng = np.random.default_rng(41)
    d = rng.standard_normal((K.ntr, K.nt)).astype(np.float32)
    m0 = (rng.standard_normal(K.model_shape) * 1e-3).astype(np.float32) if domain == mig.IMAGE else \
        np.zeros(K.model_shape, np.float32)
    got = K.adjoint(torch.from_numpy(d).cuda(), torch.from_numpy(m0.copy()).cuda()).cpu().numpy()
    want = oracle.kdm(domain, True, S, ts, st["tr"], K.src, K.rec, K.cmp, cid, d, m0.copy(), **op)
    bound = oracle.kdm(domain, True, S, ts, st["tr"], K.src, K.rec, K.cmp, cid, np.abs(d), np.abs(m0), **op)
    assert np.abs(want).max() > 0
    bad = np.abs(got - want) > 1e-5 * bound + 1e-30
    assert bad.mean() <= (0 if domain == mig.IMAGE else 1e-3), bad.sum()
    # the same adjoint twice: the image domain is deterministic
    again = K.adjoint(torch.from_numpy(d).cuda(), torch.from_numpy(m0.copy()).cuda()).cpu().numpy()
    if domain == mig.IMAGE:
        assert np.array_equal(again, got)
    K.close()


@pytest.mark.parametrize("domain", [mig.IMAGE, mig.ANGLE])
def test_sums_outside_the_fast_division_range(oracle, setup, torch, domain):
    """Traveltime sums the known-divisor division does not serve: the first
    source table shifted down by 4 ms (negative sums near that source, nodes
    the reference's negative index leaves out) and a receiver on every source
    (zero-offset traces: sum 0 at the other sources' nodes, index 0).  The
    waves holding such a sum take the IEEE path; the results match the
    oracle."""
    st = setup
    xr = np.sort(np.concatenate([st["xr"], st["xs"]])).astype(np.float32)
    e = st["e"]
    ts = st["ts"].copy()
    ts[0] -= np.float32(4e-3)
    assert (ts[1:, 3:-3, 3:-3] == 0).sum() == len(ts) - 1      # sources on grid nodes
    tr = e.traveltime_tables(xr, np.zeros_like(xr))
    K = make(e, st["m"], st["xs"], xr, domain, tables=(ts, tr))
    assert (K.src.size > 0) and np.any(st["xs"][K.src] == xr[K.rec])
    S, op = orc_args(oracle, K, st)
    cid = K.cmp_id if domain == mig.ANGLE else None
    rng = np.random.default_rng(31)
    d = rng.standard_normal((K.ntr, K.nt)).astype(np.float32)
    zero = np.zeros(K.model_shape, np.float32)
    got = K.adjoint(torch.from_numpy(d).cuda()).cpu().numpy()
    want = oracle.kdm(domain, True, S, ts, tr, K.src, K.rec, K.cmp, cid, d, zero.copy(), **op)
    bound = oracle.kdm(domain, True, S, ts, tr, K.src, K.rec, K.cmp, cid, np.abs(d), zero.copy(), **op)
    assert np.abs(want).max() > 0
    bad = np.abs(got - want) > 1e-5 * bound + 1e-30
    assert bad.mean() <= (0 if domain == mig.IMAGE else 1e-3), bad.sum()
    m = rng.standard_normal(K.model_shape).astype(np.float32)
    d0 = np.zeros((K.ntr, K.nt), np.float32)
    got = K.forward(torch.from_numpy(m).cuda()).cpu().numpy()
    want = oracle.kdm(domain, False, S, ts, tr, K.src, K.rec, K.cmp, cid, d0.copy(), m, **op)
    bound = oracle.kdm(domain, False, S, ts, tr, K.src, K.rec, K.cmp, cid, d0.copy(), np.abs(m), **op)
    bad = np.abs(got - want) > 1e-5 * bound + 1e-30
    assert bad.mean() <= (0 if domain == mig.IMAGE else 1e-3), bad.sum()
    K.close()


def test_tables_follow_the_slowness(oracle, setup, torch):
    """The table quantities hold the slowness they were evaluated with: after
    a slowness upload the operators refuse to run until every table is set
    again, and then weigh with the new slowness (the oracle on the new S)."""
    import ctypes as C
    from seisfat2d_amd import capi
    from seisfat2d_amd.capi import SeisfatError
    st = setup
    m = st["m"]
    e = Eikonal(m.s.shape[1], m.s.shape[0], m.dz, m.dx)
    try:
        e.set_slowness(m.s)
        K = make(e, m, st["xs"], st["xr"], mig.IMAGE, tables=(st["ts"], st["tr"]))
        d = torch.from_numpy(np.random.default_rng(21).standard_normal((K.ntr, K.nt)).astype(np.float32)).cuda()
        m1 = K.adjoint(d).cpu().numpy()
        s2 = (m.s * 1.25).astype(np.float32)
        e.set_slowness(s2)
        with pytest.raises(SeisfatError) as ex:
            K.adjoint(d)
        assert ex.value.code == capi.ESTATE and "slowness changed" in str(ex.value)
        ts = torch.from_numpy(np.ascontiguousarray(st["ts"], np.float32)).cuda()
        tr = torch.from_numpy(np.ascontiguousarray(st["tr"], np.float32)).cuda()
        # all but the last receiver table: still refused
        for which, tabs, n in ((0, ts, K.nsrc), (1, tr, K.nrec - 1)):
            for k in range(n):
                capi.check(capi.lib().seisfat_kdm_set_table(K._k, which, k, C.c_void_p(tabs[k].data_ptr()), None))
        with pytest.raises(SeisfatError) as ex:
            K.adjoint(d)
        assert ex.value.code == capi.ESTATE and "tables not set" in str(ex.value)
        capi.check(capi.lib().seisfat_kdm_set_table(K._k, 1, K.nrec - 1, C.c_void_p(tr[-1].data_ptr()), None))
        got = K.adjoint(d).cpu().numpy()
        _, op = orc_args(oracle, K, st)
        S2 = oracle.expand_boundary(s2)
        want = oracle.kdm(0, True, S2, st["ts"], st["tr"], K.src, K.rec, K.cmp, None, d.cpu().numpy(),
                          np.zeros(K.model_shape, np.float32), **op)
        assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
        assert np.abs(got - m1).max() > 1e-3 * np.abs(m1).max()   # the slowness entered the weights
        # back to the first slowness, every table again (Kirchhoff.set_tables):
        # the first image, bit for bit
        e.set_slowness(m.s)
        K.set_tables(ts, tr)
        assert np.array_equal(K.adjoint(d).cpu().numpy(), m1)
        K.close()
    finally:
        e.close()


def test_no_traces_within_the_offset(setup, torch):
    """A maximum offset that keeps no pair: no traces, a zero image, a
    zero-trace forward."""
    st = setup
    K = make(st["e"], st["m"], st["xs"], st["xr"], mig.IMAGE, max_offset=1.0, tables=(st["ts"], st["tr"]))
    assert K.ntr == 0
    seismic = np.ones((K.nsrc, K.nrec, K.nt), np.float32)
    img = K.migrate(seismic)
    assert img.shape == K.model_shape and not img.any()
    assert tuple(K.forward(torch.ones(K.model_shape, device="cuda")).shape) == (0, K.nt)
    K.close()
