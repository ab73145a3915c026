"""The oracle's Kirchhoff migration restatement, pinned (CPU).

The reference ships no migration fixtures (its dot-product test is commented
out, src/migration/migration.cu:282-396) and its migration does not build
here (CUDA and FFTW), so each oracle function is pinned by what the
mathematics fixes:
  - ricker: the textbook Ricker of peak frequency fmax / 3 (pi^3 fc^2 =
    pi^2 (fmax/3)^2) and its spectrum's peak; the reference's sample count;
  - trace_convolve: numpy's FFT product with the reference's conventions
    (unnormalised inverse of size nfft, conjugate for the adjoint, shift
    nw/2 + nw/10);
  - kdm (the four kernels): adjointness <L m, d> = <m, L^T d> in both
    domains; the isochron support of a spike trace; the weights on plane and
    quadratic traveltime tables, whose 9-point differences are exact,
    against each published factor evaluated in float64 from the tables'
    closed-form gradient and Hessian; the angle bin against the half angle
    between two plane-wave gradients; the angle gather summed over angles
    against the image.
"""
import math

import numpy as np
import pytest

NB = 3
EPS = 1e-6


@pytest.fixture(scope="module")
def o(oracle):
    return oracle


# -- Ricker wavelet ------------------------------------------------------------------

@pytest.mark.parametrize("dt,fmax", [(1e-3, 30.0), (5e-4, 45.0), (2e-3, 12.5)])
def test_ricker_is_the_textbook_wavelet_of_peak_frequency_fmax_over_3(o, dt, fmax):
    w = o.ricker(dt, fmax)
    t0 = 2.0 * math.sqrt(math.pi) / fmax
    assert len(w) == 2 * int((t0 - 0.5 * dt) / dt)
    fp = fmax / 3.0
    t = np.arange(len(w)) * dt - t0
    a = (math.pi * fp * t) ** 2
    assert np.allclose(w, (1.0 - 2.0 * a) * np.exp(-a), rtol=0, atol=2e-6)
    # its amplitude spectrum peaks at fp
    n = 1 << 20
    spec = np.abs(np.fft.rfft(w.astype(np.float64), n))
    f = np.fft.rfftfreq(n, dt)
    assert abs(f[np.argmax(spec)] - fp) < 0.01 * fp


# -- wavelet convolutions --------------------------------------------------------------

def fft_convolve(x, w, adjoint):
    """The reference's FFTW method, in numpy: spectra of size nfft, product
    (conjugate wavelet for the adjoint), unnormalised inverse, shift."""
    nt, nw = x.shape[-1], len(w)
    nfft = 1
    while nfft < nt + nw - 1:
        nfft *= 2
    sh = nw // 2 + nw // 10
    X = np.fft.rfft(x.astype(np.float64), nfft)
    W = np.fft.rfft(w.astype(np.float64), nfft)
    y = np.fft.irfft(X * (np.conj(W) if adjoint else W), nfft) * nfft
    out = np.zeros(x.shape, np.float64)
    if adjoint:
        if sh < nt:
            out[..., sh:] = y[..., : nt - sh]
    else:
        out[...] = y[..., sh: sh + nt]
    return out


@pytest.mark.parametrize("adjoint", [False, True])
@pytest.mark.parametrize("nt,dt,fmax", [(301, 1e-3, 30.0), (64, 2e-3, 25.0), (1001, 1e-3, 30.0)])
def test_convolution_is_the_fft_product(o, adjoint, nt, dt, fmax):
    rng = np.random.default_rng(nt)
    w = o.ricker(dt, fmax)
    x = rng.standard_normal((5, nt)).astype(np.float32)
    x[1, : nt // 3] = 0.0   # a trace with a silent start
    got = o.trace_convolve(x, w, adjoint)
    want = fft_convolve(x, w, adjoint)
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 1e-6 * scale
    if adjoint:
        sh = len(w) // 2 + len(w) // 10
        assert not got[:, :sh].any()


# -- the Kirchhoff operators -------------------------------------------------------------

def node_xz(nz, nx, dz, dx):
    """Physical (x, z) of every padded node, arrays of shape (nxx, nzz)."""
    j = np.arange(nx + 2 * NB) - NB
    i = np.arange(nz + 2 * NB) - NB
    return np.meshgrid(j * dx, i * dz, indexing="ij")


def factors(Ts, Tr, g_s, g_r, det_s, det_r, S, x, z, cmp, aperture):
    """Each published factor of the Kirchhoff weight (migration.cu:436-475)
    in float64, from closed-form gradients g = (dT/dx, dT/dz) and Hessian
    determinants."""
    ns = np.hypot(*g_s) + EPS
    nr = np.hypot(*g_r) + EPS
    jterm = np.sqrt((np.abs(det_s) + EPS) * (np.abs(det_r) + EPS)) * S / (ns * nr)
    cs, cr = np.abs(g_s[1] / ns), np.abs(g_r[1] / nr)            # unit gradient . (0, -1)
    obliq = np.sqrt(cs * cr)
    G = 1.0 / np.sqrt(np.maximum(Ts / S, EPS) * np.maximum(Tr / S, EPS))
    th = 0.5 * (np.arccos(np.clip(-g_s[1] / ns, -1, 1)) + np.arccos(np.clip(-g_r[1] / nr, -1, 1)))
    R = 1.0 + 0.2 * np.cos(th)
    sigma = math.tan(math.radians(aperture)) * z
    taper = np.exp(-0.5 * ((x - cmp) / (sigma + EPS)) ** 2)
    return jterm * obliq * G * R * taper / (2.0 * math.pi)


def run_adjoint(o, domain, S, ts, tr, cmp, data, dt, dx, dz, aperture, da=1.0, nang=1, cid=None):
    nxx, nzz = S.shape
    ntr = data.shape[0]
    src = np.arange(ntr) % len(ts)
    rec = np.arange(ntr) % len(tr)
    if domain == 0:
        model = np.zeros((nxx - 2 * NB, nzz - 2 * NB), np.float32)
    else:
        cid = np.zeros(ntr, np.int32) if cid is None else cid
        model = np.zeros((cid.max() + 1, nang, nzz - 2 * NB), np.float32)
    return o.kdm(domain, True, S, ts, tr, src, rec, np.float32(cmp) * np.ones(ntr, np.float32), cid, data, model,
                 dt, dx, dz, aperture, da, nang)


def test_weights_on_plane_wave_tables(o):
    """Linear tables: the differences are exact, the Hessians vanish (the
    determinant is EPS), and the weight is every other factor."""
    nz, nx, dz, dx = 40, 56, 7.0, 11.0
    X, Z = node_xz(nz, nx, dz, dx)
    gs, gr = (2.0e-4, 3.5e-4), (-1.5e-4, 2.5e-4)
    Ts = (0.3 + gs[0] * X + gs[1] * Z).astype(np.float32)
    Tr = (0.35 + gr[0] * X + gr[1] * Z).astype(np.float32)
    rng = np.random.default_rng(3)
    S = rng.uniform(2e-4, 5e-4, X.shape).astype(np.float32)
    dt, nt, cmp, ap = 1e-3, 1200, 230.0, 30.0
    m = run_adjoint(o, 0, S, Ts[None], Tr[None], cmp, np.ones((1, nt), np.float32), dt, dx, dz, ap)
    inner = (slice(NB, -NB), slice(NB, -NB))
    want = factors(Ts[inner].astype(np.float64), Tr[inner].astype(np.float64), gs, gr, 0.0, 0.0,
                   S[inner].astype(np.float64), X[inner], Z[inner], cmp, ap)
    assert np.all(np.isfinite(m)) and np.abs(m).max() > 0
    rel = np.abs(m - want) / np.abs(want).max()
    assert rel.max() < 2e-4, rel.max()


def test_weights_on_quadratic_tables(o):
    """Quadratic tables: central differences are exact for the gradient and
    the Hessian, which is constant: |det H| = |4pq - r^2| dominates EPS."""
    nz, nx, dz, dx = 30, 36, 0.5, 1.0
    X, Z = node_xz(nz, nx, dz, dx)
    ps, qs, rs, ls = 2e-3, 3e-3, 1e-3, (0.05, 0.08)
    pr, qr, rr, lr = 1.5e-3, 4e-3, -2e-3, (-0.04, 0.06)
    Ts = 1.0 + ls[0] * X + ls[1] * Z + ps * X * X + qs * Z * Z + rs * X * Z
    Tr = 1.2 + lr[0] * X + lr[1] * Z + pr * X * X + qr * Z * Z + rr * X * Z
    Ts32, Tr32 = Ts.astype(np.float32), Tr.astype(np.float32)
    S = np.full(X.shape, 3e-4, np.float32)
    dt, nt, cmp, ap = 1e-2, 2000, 14.0, 40.0
    m = run_adjoint(o, 0, S, Ts32[None], Tr32[None], cmp, np.ones((1, nt), np.float32), dt, dx, dz, ap)
    inner = (slice(NB, -NB), slice(NB, -NB))
    x, z = X[inner], Z[inner]
    g_s = (ls[0] + 2 * ps * x + rs * z, ls[1] + 2 * qs * z + rs * x)
    g_r = (lr[0] + 2 * pr * x + rr * z, lr[1] + 2 * qr * z + rr * x)
    det_s, det_r = 4 * ps * qs - rs * rs, 4 * pr * qr - rr * rr
    assert min(abs(det_s), abs(det_r)) > 10 * EPS
    want = factors(Ts32[inner].astype(np.float64), Tr32[inner].astype(np.float64), g_s, g_r, det_s, det_r,
                   3e-4, x, z, cmp, ap)
    rel = np.abs(m - want) / np.abs(want).max()
    assert rel.max() < 1e-3, rel.max()


def homogeneous_tables(nz, nx, dz, dx, v, pos):
    X, Z = node_xz(nz, nx, dz, dx)
    return np.stack([np.hypot(X - px, Z - pz) / v for px, pz in pos]).astype(np.float32)


def test_spike_trace_images_its_isochron(o):
    nz, nx, dz, dx, v = 50, 80, 10.0, 10.0, 2000.0
    ts = homogeneous_tables(nz, nx, dz, dx, v, [(200.0, 0.0)])
    tr = homogeneous_tables(nz, nx, dz, dx, v, [(560.0, 0.0)])
    S = np.full(ts.shape[1:], 1.0 / v, np.float32)
    dt, nt, k0 = 1e-2, 100, 40
    d = np.zeros((1, nt), np.float32)
    d[0, k0] = 1.0
    m = run_adjoint(o, 0, S, ts, tr, 380.0, d, dt, dx, dz, 60.0)
    T = (ts[0] + tr[0])[NB:-NB, NB:-NB]
    on = np.floor(T / np.float32(dt)) == k0
    assert on.sum() > 20
    assert not m[~on].any()
    deep = on.copy()
    deep[:, :10] = False        # shallower, the narrow taper underflows to 0
    assert deep.sum() > 10 and np.all(m[deep] != 0.0)


@pytest.mark.parametrize("domain", [0, 1])
def test_forward_and_adjoint_are_adjoint(o, domain):
    nz, nx, dz, dx = 36, 60, 10.0, 12.0
    rng = np.random.default_rng(11 + domain)
    v = rng.uniform(1800.0, 2600.0)
    ts = homogeneous_tables(nz, nx, dz, dx, v, [(100.0, 0.0), (300.0, 5.0), (520.0, 0.0)])
    tr = homogeneous_tables(nz, nx, dz, dx, v, [(40.0, 0.0), (260.0, 0.0), (400.0, 10.0), (650.0, 0.0)])
    ts += rng.uniform(0, 1e-3, ts.shape).astype(np.float32)   # rough tables: every weight factor varies
    S = rng.uniform(3.5e-4, 5.5e-4, ts.shape[1:]).astype(np.float32)
    src = np.array([0, 0, 1, 1, 1, 2, 2], np.int32)
    rec = np.array([0, 2, 1, 2, 3, 0, 3], np.int32)
    cmp = rng.uniform(0, nx * dx, len(src)).astype(np.float32)
    nt, dt, nang, da = 450, 2e-3, 16, 3.0
    cid = np.array([0, 1, 1, 2, 3, 3, 4], np.int32)
    mshape = (nx, nz) if domain == 0 else (5, nang, nz)
    m = rng.standard_normal(mshape).astype(np.float32)
    d = rng.standard_normal((len(src), nt)).astype(np.float32)
    Lm = o.kdm(domain, False, S, ts, tr, src, rec, cmp, cid, np.zeros_like(d), m, dt, dx, dz, 35.0, da, nang)
    Ltd = o.kdm(domain, True, S, ts, tr, src, rec, cmp, cid, d, np.zeros_like(m), dt, dx, dz, 35.0, da, nang)
    a = np.dot(Lm.ravel().astype(np.float64), d.ravel())
    b = np.dot(m.ravel().astype(np.float64), Ltd.ravel())
    assert np.abs(Lm).max() > 0 and np.abs(Ltd).max() > 0
    assert abs(a - b) <= 1e-5 * (abs(a) + abs(b)), (a, b)


def test_angle_bin_is_half_the_angle_between_the_gradients(o):
    nz, nx, dz, dx = 24, 30, 8.0, 8.0
    X, Z = node_xz(nz, nx, dz, dx)
    for gs, gr in [((3e-4, 4e-4), (-3e-4, 4e-4)), ((1e-4, 4.5e-4), (-2e-4, 3e-4)), ((0.0, 5e-4), (4e-4, 3e-4))]:
        ts = (0.2 + gs[0] * X + gs[1] * Z).astype(np.float32)[None]
        tr = (0.2 + gr[0] * X + gr[1] * Z).astype(np.float32)[None]
        S = np.full(X.shape, 4e-4, np.float32)
        half = 0.5 * math.degrees(math.acos(np.dot(gs, gr) / (np.hypot(*gs) * np.hypot(*gr))))
        da = 1.7
        b = int(half // da)
        assert 0.1 < half / da - b < 0.9          # safely inside its bin
        nang = b + 3
        m = run_adjoint(o, 1, S, ts, tr, 100.0, np.ones((1, 800), np.float32), 1e-3, dx, dz, 50.0, da, nang)
        assert m.shape == (1, nang, nz)
        used = np.nonzero(np.abs(m[0]).sum(axis=1))[0]
        assert list(used) == [b], (half, used)


def test_angle_gather_summed_over_angles_is_the_image_summed_over_x(o):
    nz, nx, dz, dx, v = 30, 50, 10.0, 10.0, 2200.0
    rng = np.random.default_rng(2)
    ts = homogeneous_tables(nz, nx, dz, dx, v, [(120.0, 0.0)])
    tr = homogeneous_tables(nz, nx, dz, dx, v, [(330.0, 0.0)])
    S = np.full(ts.shape[1:], 1.0 / v, np.float32)
    d = rng.standard_normal((1, 500)).astype(np.float32)
    img = run_adjoint(o, 0, S, ts, tr, 225.0, d, 1e-3, dx, dz, 30.0)
    ang = run_adjoint(o, 1, S, ts, tr, 225.0, d, 1e-3, dx, dz, 30.0, da=1.0, nang=91)
    a, b = ang[0].sum(axis=0, dtype=np.float64), img.sum(axis=0, dtype=np.float64)
    assert np.allclose(a, b, rtol=1e-5, atol=1e-6 * np.abs(b).max())


# -- the least-squares loop --------------------------------------------------------------

def ls_problem(o, domain, nz=24, nx=40, dz=10.0, dx=10.0, nt=300, seed=4):
    v = 2000.0
    xs = [(60.0, 0.0), (200.0, 0.0)]
    xr = [(20.0, 0.0), (120.0, 0.0), (240.0, 0.0), (330.0, 0.0)]
    ts = homogeneous_tables(nz, nx, dz, dx, v, xs)
    tr = homogeneous_tables(nz, nx, dz, dx, v, xr)
    S = np.full(ts.shape[1:], 1.0 / v, np.float32)
    src = np.repeat(np.arange(2, dtype=np.int32), 4)
    rec = np.tile(np.arange(4, dtype=np.int32), 2)
    cmp = (0.5 * (np.array([p[0] for p in xs])[src] + np.array([p[0] for p in xr])[rec])).astype(np.float32)
    cid = ((cmp - cmp.min()) / np.float32(20.0)).astype(np.int32)
    nang, da = 12, 5.0
    mshape = (nx, nz) if domain == 0 else (cid.max() + 1, nang, nz)
    rng = np.random.default_rng(seed)
    m_true = np.zeros(mshape, np.float32)
    if domain == 0:
        m_true[:, nz // 2] = 1.0
    else:
        m_true[:, :, nz // 2] = rng.uniform(0.5, 1.0, mshape[:2])
    w = o.ricker(2e-3, 30.0)
    op = dict(dt=2e-3, dx=dx, dz=dz, aperture=30.0, da=da, nang=nang)
    d = o.kdm(domain, False, S, ts, tr, src, rec, cmp, cid, np.zeros((len(src), nt), np.float32), m_true, **op)
    dobs = o.trace_convolve(d, w, False)
    return dict(S=S, ts=ts, tr=tr, src=src, rec=rec, cmp=cmp, cid=cid, dobs=dobs, w=w, op=op, mshape=mshape)


def run_ls(o, domain, P, max_it, smooth=0.0):
    op = P["op"]
    return o.lskdm(domain, P["S"], P["ts"], P["tr"], P["src"], P["rec"], P["cmp"], P["cid"], P["dobs"], P["w"],
                   max_it, smooth, op["dt"], op["dx"], op["dz"], op["aperture"], P["mshape"], op["da"], op["nang"])


@pytest.mark.parametrize("domain", [0, 1])
def test_least_squares_first_residual_is_the_data_norm_and_the_misfit_falls(o, domain):
    P = ls_problem(o, domain)
    model, res = run_ls(o, domain, P, 4)
    assert len(res) == 5
    d = P["dobs"]
    norm = math.sqrt(float(np.sum(np.where(np.abs(d) > EPS, d, 0).astype(np.float64) ** 2)))
    assert abs(res[0] - norm) <= 1e-5 * norm
    # a consistent problem: every conjugate-gradient step lowers the misfit
    assert all(b < a for a, b in zip(res, res[1:])), res


def test_least_squares_first_step_is_the_scaled_gradient(o):
    """One iteration from m = 0: direction = gradient (beta = 0), model =
    alpha * gradient with alpha = |g|^2 / (|A g|^2 + 1e-6) (the step length
    of LSKDM.cu:128-196, its EPS included)."""
    P = ls_problem(o, 0)
    op = P["op"]
    m1, _ = run_ls(o, 0, P, 1)
    r = np.where(np.abs(P["dobs"]) > EPS, P["dobs"], 0).astype(np.float32)
    g = o.kdm(0, True, P["S"], P["ts"], P["tr"], P["src"], P["rec"], P["cmp"], P["cid"],
              np.ascontiguousarray(o.trace_convolve(r, P["w"], True)), np.zeros(P["mshape"], np.float32), **op)
    Ag = o.trace_convolve(o.kdm(0, False, P["S"], P["ts"], P["tr"], P["src"], P["rec"], P["cmp"], P["cid"],
                                np.zeros_like(r), g, **op), P["w"], False)
    alpha = np.sum(g.astype(np.float64) ** 2) / (np.sum(Ag.astype(np.float64) ** 2) + EPS)
    assert np.allclose(m1, alpha * g, rtol=1e-4, atol=1e-6 * np.abs(alpha * g).max())


def test_angle_smoothing_only_changes_the_update_along_angle(o):
    """ADLSKDM::regularization adds smooth * d2m/da2 to the direction; a
    smoothing weight of 0 is the plain loop, and the smoothed model differs
    only where the previous model varied along angle."""
    P = ls_problem(o, 1)
    m0, r0 = run_ls(o, 1, P, 2, smooth=0.0)
    m1, r1 = run_ls(o, 1, P, 2, smooth=1e-2)
    assert r0[:2] == r1[:2]             # the first update sees a zero model: no smoothing term
    assert not np.array_equal(m0, m1)


def test_threaded_image_adjoint_is_the_serial_one(o):
    """Image columns on several threads: every image point still receives its
    contributions in trace order, so the image is bit for bit the serial one."""
    nz, nx, dz, dx = 30, 44, 10.0, 10.0
    rng = np.random.default_rng(8)
    ts = homogeneous_tables(nz, nx, dz, dx, 2100.0, [(50.0, 0.0), (250.0, 0.0)])
    tr = homogeneous_tables(nz, nx, dz, dx, 2100.0, [(10.0, 0.0), (180.0, 0.0), (400.0, 0.0)])
    S = rng.uniform(4e-4, 5e-4, ts.shape[1:]).astype(np.float32)
    src = np.array([0, 0, 1, 1, 1], np.int32)
    rec = np.array([0, 1, 0, 1, 2], np.int32)
    cmp = rng.uniform(0, 400, 5).astype(np.float32)
    d = rng.standard_normal((5, 400)).astype(np.float32)
    args = (S, ts, tr, src, rec, cmp, None)
    m1 = o.kdm(0, True, *args, d, np.zeros((nx, nz), np.float32), 1e-3, dx, dz, 30.0)
    m8 = o.kdm(0, True, *args, d, np.zeros((nx, nz), np.float32), 1e-3, dx, dz, 30.0, nthreads=8)
    assert np.abs(m1).max() > 0 and np.array_equal(m1, m8)


def test_folded_weight_within_a_few_ulp(o):
    """The GPU folds the weight's per-table factors into one amplitude per
    table node, amp = sqrt(det cos) / (norm sqrt(max(T/S, EPS))) in double,
    and the angle factor's cos((th_s + th_r) / 2) into the half angles'
    cosines and sines, sqrt((1 +- cos th) / 2) in float32.  Restated here in float32 next to the reference's order
    of roundings (src/migration/migration.cu:417-475) on solved tables of a
    rough model with unequal spacing: every weight above 1e-30 within 8 ulp.
    (The GPU's weights themselves are checked against the oracle's on the
    GPU.)"""
    from seisfat2d_amd import models
    f32 = np.float32
    nz, nx, dz, dx = 40, 72, 7.0, 10.0
    m = models.rough(nz, nx, dz, dx, seed=3, vlo=1500.0, vhi=4500.0, sigma=1.0)
    S = o.expand_boundary(m.s)
    sn = m.s.astype(f32)

    def table(xpos):
        T = o.solve_iso(S, dx, dz, float(xpos), 0.0)
        c = T[3:-3, 3:-3]
        up, dn, lf, rt = T[3:-3, 2:-4], T[3:-3, 4:-2], T[2:-4, 3:-3], T[4:-2, 3:-3]
        d_z = f32(0.5) * (dn - up) / f32(dz)
        d_x = f32(0.5) * (rt - lf) / f32(dx)
        d2z = (dn - f32(2) * c + up) / f32(dz * dz)
        d2x = (rt - f32(2) * c + lf) / f32(dx * dx)
        dxz = (T[4:-2, 4:-2] - T[4:-2, 2:-4] - T[2:-4, 4:-2] + T[2:-4, 2:-4]) / (f32(4) * f32(dx) * f32(dz))
        norm = np.sqrt(d_x * d_x + d_z * d_z) + f32(EPS)
        det = np.abs(d2x * d2z - dxz * dxz) + f32(EPS)
        dot = (d_z / norm) * f32(-1)
        th = np.arccos(np.clip(dot, f32(-1), f32(1)))
        rs = np.maximum(c / sn, f32(EPS))
        amp = (np.sqrt(det.astype(np.float64) * np.abs(dot)) / (norm.astype(np.float64) * np.sqrt(rs))).astype(f32)
        cth = np.clip(dot, f32(-1), f32(1))
        return dict(det=det, norm=norm, cos=np.abs(dot), th=th, rs=rs, amp=amp,
                    ch=np.sqrt(f32(0.5) * (f32(1) + cth)), sh=np.sqrt(f32(0.5) * (f32(1) - cth)))

    x = (np.arange(nx, dtype=f32) * f32(dx))[:, None] * np.ones((1, nz), f32)
    z = np.ones((nx, 1), f32) * (np.arange(nz, dtype=f32) * f32(dz))[None, :]
    tan_ap = f32(math.tan(math.radians(30.0)))
    scale = f32(1.0 / (2.0 * math.pi))
    rels = []
    for xs in (60.0, 400.0):
        s = table(xs)
        for xr in (0.0, 250.0, 700.0):
            r = table(xr)
            cmp = f32(0.5) * (f32(xs) + f32(xr))
            a = (x - cmp) / (tan_ap * z + f32(EPS))
            taper = np.exp(f32(-0.5) * a * a)
            ref = (scale * (np.sqrt(s["det"] * r["det"]) * sn / (s["norm"] * r["norm"]))
                   * np.sqrt(s["cos"] * r["cos"]) * (f32(1) / np.sqrt(s["rs"] * r["rs"]))
                   * (f32(1) + f32(0.2) * np.cos(f32(0.5) * (s["th"] + r["th"])))) * taper
            rf = f32(1) + f32(0.2) * (s["ch"] * r["ch"] - s["sh"] * r["sh"])
            fold = scale * sn * (s["amp"] * r["amp"]) * rf * taper
            keep = np.abs(ref) > 1e-30
            rels.append(np.abs(fold[keep].astype(np.float64) - ref[keep]) / np.abs(ref[keep]))
    rel = np.concatenate(rels)
    assert rel.size > 10000
    assert rel.max() <= 8 * 2.0 ** -23, rel.max() / 2.0 ** -23
