"""Pins of the CPU oracle (oracle/oracle.cpp) against closed forms, textbook
formulas, brute force and the reference's literal schedule.  The reference
ships no tests or golden vectors for this path, so every function of the
oracle is pinned here by something other than its own formula."""
import numpy as np
import pytest
from scipy.ndimage import gaussian_filter

NB = 3


def smooth_model(rng, nz, nx, dh, v0=2000.0, amp=400.0, grad=1.0, sigma=3.0):
    v = v0 + amp * gaussian_filter(rng.standard_normal((nx, nz)), sigma) + grad * np.arange(nz)[None, :] * dh
    return (1.0 / v)


# ---------------------------------------------------------------- padding (A1)
def test_expand_boundary_hand_fixture(oracle):
    # model nz=2, nx=2: column x=0 holds (1, 2), column x=1 holds (3, 4)
    s = np.array([[1.0, 2.0], [3.0, 4.0]], np.float32)
    col0 = [1, 1, 1, 1, 2, 2, 2, 2]
    col1 = [3, 3, 3, 3, 4, 4, 4, 4]
    want = np.array([col0] * 4 + [col1] * 4, np.float32)
    assert np.array_equal(oracle.expand_boundary(s), want)


def test_expand_reduce_round_trip_and_clamp_form(oracle, rng):
    s = rng.random((7, 5)).astype(np.float32)
    S = oracle.expand_boundary(s)
    assert S.shape == (13, 11)
    jj = np.clip(np.arange(13) - NB, 0, 6)
    ii = np.clip(np.arange(11) - NB, 0, 4)
    assert np.array_equal(S, s[jj][:, ii])  # nearest-edge replication
    assert np.array_equal(oracle.reduce_boundary(S), s)


# ------------------------------------------------------------------ init (A3)
def test_init_closed_form(oracle):
    nz, nx, dx, dz = 20, 30, 10.0, 7.0
    s = np.linspace(1e-4, 5e-4, nz * nx).reshape(nx, nz).astype(np.float32)
    S = oracle.expand_boundary(s)
    sx, sz = 123.4, 61.0
    T, sIdx, sIdz = oracle.init_time(S, dx, dz, sx, sz)
    assert (sIdx, sIdz) == (int(round(sx / dx)) + NB, int(round(sz / dz)) + NB)
    mask = np.zeros(T.shape, bool)
    for j in range(sIdx - 1, sIdx + 2):
        for i in range(sIdz - 1, sIdz + 2):
            mask[j, i] = True
            d = np.hypot((j - NB) * dx - sx, (i - NB) * dz - sz)
            assert abs(T[j, i] - S[j, i] * d) <= 1e-6 * S[j, i] * d
    assert np.all(T[~mask] == np.float32(1e6))


# ------------------------------------------------------------- sweeps (A4/A5)
@pytest.mark.parametrize("nz,nx", [(37, 53), (60, 9), (5, 70), (2, 2)])
def test_raster_equals_reference_diagonal_schedule(oracle, rng, nz, nx):
    s = smooth_model(rng, nz, nx, 10.0).astype(np.float32)
    S = oracle.expand_boundary(s)
    sx, sz = rng.uniform(0, (nx - 1) * 10.0), rng.uniform(0, (nz - 1) * 10.0)
    T0, _, _ = oracle.init_time(S, 10.0, 10.0, sx, sz)
    a, b = T0.copy(), T0.copy()
    oracle.fsm(a, S, 10.0, 10.0, schedule=0)
    oracle.fsm(b, S, 10.0, 10.0, schedule=1)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("dx,dz", [(10.0, 10.0), (10.0, 7.0), (4.0, 9.0)])
@pytest.mark.parametrize("phi_deg", [20.0, 45.0, 70.0])
def test_sweep_reproduces_plane_wave(oracle, dx, dz, phi_deg):
    # Linear traveltime fields solve the eikonal equation exactly and branch A
    # of the local solver must return them to rounding; boundary row/column 0
    # carry the exact values, one (+z,+x) sweep fills the interior.
    nz, nx, s = 30, 40, 1.0 / 2000.0
    S = np.full((nx + 6, nz + 6), s, np.float32)
    phi = np.radians(phi_deg)
    j = np.arange(nx + 6)[:, None]
    i = np.arange(nz + 6)[None, :]
    Tex = s * (j * dx * np.cos(phi) + i * dz * np.sin(phi))
    T = np.full(S.shape, 1e6, np.float32)
    T[0, :] = Tex[0, :]
    T[:, 0] = Tex[:, 0]
    oracle.sweep(T, S, dx, dz, 0)
    inner = (slice(1, nx + 5), slice(1, nz + 5))
    assert np.max(np.abs(T[inner] - Tex[inner]) / Tex[inner]) < 2e-6


def _plane_wave(nx, nz, dx, dz, s, phi):
    j = np.arange(nx + 6)[:, None]
    i = np.arange(nz + 6)[None, :]
    return (s * (j * dx * np.cos(phi) + i * dz * np.sin(phi))).astype(np.float64)


# Branches B and C (src/modeling/modeling.cu:344-351) are the 2-D updates from
# one cell edge alone: B through te and tev (the x-upwind column), C through tv
# and tev (the z-upwind row).  With the other upwind neighbour not reached
# (1e6), branch A's causality guard (:337) fails, and the exact time of a plane
# wave through the two seeded neighbours must come out of B (or C) at node
# (1, 1); the rest of the first column (row) then follows by branch A.  The
# angles put the wave inside B's (C's) window but outside the window the other
# spacing would give (dz^2/hyp vs dx^2/hyp), so a swapped threshold fails too.
@pytest.mark.parametrize("dx,dz,phi_deg", [(4.0, 9.0, 30.0), (10.0, 7.0, 20.0), (10.0, 3.0, 10.0), (3.0, 10.0, 55.0)])
def test_branch_b_from_x_edge_plane_wave(oracle, dx, dz, phi_deg):
    nz, nx, s = 20, 12, 1.0 / 2500.0
    phi = np.radians(phi_deg)
    hyp = np.hypot(dx, dz)
    assert np.sin(phi) <= dz / hyp  # inside B's window
    S = np.full((nx + 6, nz + 6), s, np.float32)
    Tex = _plane_wave(nx, nz, dx, dz, s, phi)
    T = np.full(S.shape, 1e6, np.float32)
    T[0, :] = Tex[0, :]                  # x-boundary column exact, z-boundary row unreached
    counts = oracle.sweep_counted(T, S, dx, dz, 0)
    assert counts["B"] >= 1
    col = T[1, 1:nz + 5].astype(np.float64)
    assert np.max(np.abs(col - Tex[1, 1:nz + 5]) / Tex[1, 1:nz + 5]) < 2e-6


@pytest.mark.parametrize("dx,dz,phi_deg", [(9.0, 4.0, 60.0), (7.0, 10.0, 70.0), (3.0, 10.0, 80.0), (10.0, 3.0, 35.0)])
def test_branch_c_from_z_edge_plane_wave(oracle, dx, dz, phi_deg):
    nz, nx, s = 12, 20, 1.0 / 2500.0
    phi = np.radians(phi_deg)
    hyp = np.hypot(dx, dz)
    assert np.cos(phi) <= dx / hyp  # inside C's window
    S = np.full((nx + 6, nz + 6), s, np.float32)
    Tex = _plane_wave(nx, nz, dx, dz, s, phi)
    T = np.full(S.shape, 1e6, np.float32)
    T[:, 0] = Tex[:, 0]                  # z-boundary row exact, x-boundary column unreached
    counts = oracle.sweep_counted(T, S, dx, dz, 0)
    assert counts["C"] >= 1
    row = T[1:nx + 5, 1].astype(np.float64)
    assert np.max(np.abs(row - Tex[1:nx + 5, 1]) / Tex[1:nx + 5, 1]) < 2e-6


def _sweep_dirs(sweep):
    # sgnt = {{1,1},{-1,1},{1,-1},{-1,-1}} (src/modeling/modeling.cu:70)
    return (1 if sweep in (0, 2) else -1), (1 if sweep in (0, 1) else -1)


# 1-D edge candidates (src/modeling/modeling.cu:328-331).  A single reached
# node on the boundary and everything else unreached: along its row (column)
# the sweep can only propagate in 1-D, each step adding spacing x the smaller
# slowness of the two cells beside the edge it crosses (cell rows i-1 and i
# for a row, cell columns j-1 and j for a column; the reference reads cell 1
# in place of cell 0).  The traveltimes are the float32 running sums of those
# increments, exactly.
@pytest.mark.parametrize("sweep", [0, 1, 2, 3])
@pytest.mark.parametrize("row", [1, 5])
def test_1d_x_edge_running_sum(oracle, rng, sweep, row):
    nz, nx, dx, dz = 10, 15, 10.0, 3.0
    a, b = _sweep_dirs(sweep)
    nzz, nxx = nz + 6, nx + 6
    S = rng.uniform(1 / 6000, 1 / 1500, (nxx, nzz)).astype(np.float32)
    i0 = row if a > 0 else nzz - 1 - row
    j_start = 0 if b > 0 else nxx - 1
    T = np.full((nxx, nzz), 1e6, np.float32)
    T[j_start, i0] = np.float32(0.25)
    oracle.sweep(T, S, dx, dz, sweep)
    t = np.float32(0.25)
    r_lo, r_hi = max(i0 - 1, 1), min(i0, nzz - 1)
    for step in range(1, nxx - 1):
        j = j_start + b * step
        c = min(j, j - b)                         # cell column between j-b and j
        t = np.float32(t + np.float32(dx) * min(S[c, r_lo], S[c, r_hi]))
        assert T[j, i0] == t, (sweep, row, j)


@pytest.mark.parametrize("sweep", [0, 1, 2, 3])
@pytest.mark.parametrize("col", [1, 6])
def test_1d_z_edge_running_sum(oracle, rng, sweep, col):
    nz, nx, dx, dz = 15, 10, 3.0, 10.0
    a, b = _sweep_dirs(sweep)
    nzz, nxx = nz + 6, nx + 6
    S = rng.uniform(1 / 6000, 1 / 1500, (nxx, nzz)).astype(np.float32)
    j0 = col if b > 0 else nxx - 1 - col
    i_start = 0 if a > 0 else nzz - 1
    T = np.full((nxx, nzz), 1e6, np.float32)
    T[j0, i_start] = np.float32(0.25)
    oracle.sweep(T, S, dx, dz, sweep)
    t = np.float32(0.25)
    c_lo, c_hi = max(j0 - 1, 1), min(j0, nxx - 1)
    for step in range(1, nzz - 1):
        i = i_start + a * step
        r = min(i, i - a)                         # cell row between i-a and i
        t = np.float32(t + np.float32(dz) * min(S[c_lo, r], S[c_hi, r]))
        assert T[j0, i] == t, (sweep, col, i)


def test_rough_unequal_spacing_models_reach_branches_b_and_c(oracle):
    # The GPU parity models with dz:dx far from 1 (white-noise velocity, no
    # smoothing) must really let branches B and C decide nodes.
    from seisfat2d_amd import models
    for m, want in ((models.rough(64, 96, 3.0, 10.0, seed=11), "C"), (models.rough(96, 64, 10.0, 3.0, seed=12), "B")):
        S = oracle.expand_boundary(m.s)
        c = oracle.branch_counts(S, m.dx, m.dz, m.sx, m.sz)
        assert c[want] >= 20, c


def test_constant_velocity_first_order_source_error(oracle):
    nz, nx, dh, v = 201, 1001, 10.0, 2000.0
    S = oracle.expand_boundary(np.full((nx, nz), 1.0 / v, np.float32))
    T = oracle.solve_iso(S, dh, dh, 5000.0, 0.0)
    x = (np.arange(nx + 6) - NB) * dh
    z = (np.arange(nz + 6) - NB) * dh
    r = np.hypot(x[:, None] - 5000.0, z[None, :])
    inner = (slice(NB, nx + NB), slice(NB, nz + NB))
    rel = np.abs(T[inner] - r[inner] / v)[r[inner] > 0] / (r[inner] / v)[r[inner] > 0]
    assert rel.max() < 0.03                       # first-order error near the source
    far = r[inner] > 1500.0
    assert np.max(np.abs(T[inner][far] - r[inner][far] / v) / (r[inner][far] / v)) < 0.01
    assert np.all(T[inner] >= r[inner] / v * (1 - 1e-6))  # upwind scheme never undershoots r/v


@pytest.mark.parametrize("nz,nx,dz,dx", [(60, 200, 3.0, 10.0), (200, 60, 10.0, 3.0), (80, 120, 4.0, 9.0)])
def test_constant_velocity_unequal_spacing_far_field(oracle, nz, nx, dz, dx):
    # A point source in a constant model on a strongly anisotropic grid: past
    # 20 cells from the source the single pass stays within 2 % of r/v (the
    # scheme's first-order source error).  Branch A's causality guard decides
    # which updates use the three-point plane wave; a guard that tested the
    # z-neighbour against dz (and the x-neighbour against dx) would fall back
    # to one-sided updates across most of the grid (errors near 40 %).
    v = 2000.0
    S = oracle.expand_boundary(np.full((nx, nz), 1.0 / v, np.float32))
    sx, sz = 0.43 * (nx - 1) * dx, 0.31 * (nz - 1) * dz
    T = oracle.solve_iso(S, dx, dz, sx, sz)
    x = (np.arange(nx + 6) - NB) * dx
    z = (np.arange(nz + 6) - NB) * dz
    r = np.hypot(x[:, None] - sx, z[None, :] - sz)[NB:-NB, NB:-NB]
    far = r > 20 * max(dx, dz)
    rel = np.abs(T[NB:-NB, NB:-NB][far] - r[far] / v) / (r[far] / v)
    assert rel.max() < 0.02


def test_two_layer_head_wave(oracle):
    # Flat two-layer model: surface first arrivals equal min(direct, head wave)
    # with the textbook head-wave time x/v2 + 2 h cos(ic)/v1.
    nz, nx, dh, h, v1, v2 = 151, 801, 10.0, 400.0, 2000.0, 4000.0
    v = np.where(np.arange(nz)[None, :] * dh < h, v1, v2) * np.ones((nx, 1))
    S = oracle.expand_boundary((1.0 / v).astype(np.float32))
    T = oracle.solve_iso(S, dh, dh, 0.0, 0.0)
    x = np.arange(nx) * dh
    ic = np.arcsin(v1 / v2)
    t_head = x / v2 + 2 * h * np.cos(ic) / v1
    t_first = np.minimum(x / v1, t_head)
    surf = T[NB:nx + NB, NB]
    sel = x >= 1000.0
    assert np.max(np.abs(surf[sel] - t_first[sel]) / t_first[sel]) < 0.02
    beyond = x > 3000.0  # well past the crossover the head wave must win
    assert np.all(surf[beyond] < x[beyond] / v1)


def test_fp32_matches_fp64(oracle, rng):
    s = smooth_model(rng, 81, 121, 10.0)
    S32 = oracle.expand_boundary(s.astype(np.float32))
    S64 = oracle.expand_boundary(s.astype(np.float64))
    T32 = oracle.solve_iso(S32, 10.0, 10.0, 600.0, 0.0)
    T64 = oracle.solve_iso(S64, 10.0, 10.0, 600.0, 0.0)
    inner = (slice(NB, -NB), slice(NB, -NB))
    m = T64[inner] > 0
    assert np.max(np.abs(T32[inner][m] - T64[inner][m]) / T64[inner][m]) < 1e-5


def test_extra_passes_only_lower_and_converge(oracle, rng):
    s = smooth_model(rng, 60, 90, 10.0, amp=800.0)
    S = oracle.expand_boundary(s.astype(np.float32))
    T1 = oracle.solve_iso(S, 10.0, 10.0, 450.0, 300.0)
    Tc, passes = oracle.solve_iso(S, 10.0, 10.0, 450.0, 300.0, mode=1, max_passes=50)
    assert passes < 50
    assert np.all(Tc <= T1)
    again = Tc.copy()
    oracle.fsm(again, S, 10.0, 10.0)
    assert np.array_equal(again, Tc)  # converged state is a fixed point


def test_mirror_symmetry(oracle):
    nz, nx, dh = 41, 61, 10.0
    v = (1500 + 1.5 * np.arange(nz) * dh)[None, :] * np.ones((nx, 1))
    S = oracle.expand_boundary((1.0 / v).astype(np.float32))
    T = oracle.solve_iso(S, dh, dh, 300.0, 0.0)   # source on the centre column
    assert np.array_equal(T, T[::-1, :])


# ------------------------------------------------------- anisotropy (A7, A8, A12)
def test_compress_decode_bounds(oracle, rng):
    c = (1e9 + 5e8 * rng.random((50, 40))).astype(np.float32)
    u, mn, mx = oracle.compress(c)
    assert mn == c.min() and mx == c.max()
    assert u[np.unravel_index(np.argmin(c), c.shape)] == 1
    assert u[np.unravel_index(np.argmax(c), c.shape)] == 65535
    dec = mn + (u.astype(np.float64) - 1.0) * (mx - mn) / 65534.0
    step = (mx - mn) / 65534.0
    # truncation toward the lower level; FP32 rounding of the scaled value can
    # push an exact level boundary up by one step, never more
    assert np.all(np.abs(dec - c) <= step * 1.001)
    assert np.mean(dec <= c + 1e-6 * step) > 0.99


def test_compress_constant_component(oracle):
    # Every VTI model has C15 = C35 = 0 everywhere: the quantiser's range is
    # empty, so every code is 1 and decodes to the constant exactly.
    for v in (0.0, 3.25e9):
        u, mn, mx = oracle.compress(np.full((9, 7), v, np.float32))
        assert mn == mx == np.float32(v)
        assert np.all(u == 1)
        assert np.float32(mn + (np.float32(u[0, 0]) - 1.0) * (mx - mn) / 65534.0) == np.float32(v)


def test_stiffness_thomsen_relations():
    from seisfat2d_amd.stiffness import compute_stiffness
    vp, vs, rho, eps, dlt = 3000.0, 1700.0, 2200.0, 0.2, 0.1
    c11, c13, c15, c33, c35, c55 = compute_stiffness(vp, vs, rho, eps, dlt, 0.0)
    assert np.isclose(c33, rho * vp ** 2) and np.isclose(c55, rho * vs ** 2)
    assert np.isclose((c11 - c33) / (2 * c33), eps)
    assert np.isclose(((c13 + c55) ** 2 - (c33 - c55) ** 2) / (2 * c33 * (c33 - c55)), dlt)
    assert abs(c15) < 1e-6 * c33 and abs(c35) < 1e-6 * c33
    # an isotropic medium is unchanged by any tilt
    iso0 = compute_stiffness(vp, vs, rho, 0.0, 0.0, 0.0)
    iso = compute_stiffness(vp, vs, rho, 0.0, 0.0, 23.0)
    assert np.allclose(iso, iso0, rtol=1e-12, atol=1e-3)


def _thomsen_qp_velocity(vp, vs, eps, dlt, theta):
    # exact qP phase velocity of a VTI medium (Tsvankin 1996), theta from the axis
    f = 1.0 - (vs / vp) ** 2
    s2 = np.sin(theta) ** 2
    root = np.sqrt((1 + 2 * eps * s2 / f) ** 2 - 2 * (eps - dlt) * np.sin(2 * theta) ** 2 / f)
    return vp * np.sqrt(1 + eps * s2 - f / 2 + f / 2 * root)


@pytest.mark.parametrize("tilt", [0.0, 25.0, -30.0])
def test_quasi_slowness_is_thomsen_qp(oracle, tilt):
    # A plane-wave traveltime field along direction phi (from vertical) and a
    # TTI medium; the qP slowness must be 1/v_qP(phi - tilt).  The right half
    # of the grid holds a 10 % different medium so that the uint16 range is
    # non-degenerate.
    from seisfat2d_amd.stiffness import compute_stiffness
    nz, nx, dh = 20, 24, 10.0
    vp, vs, rho, eps, dlt = 3000.0, 3000.0 / np.sqrt(3), 1000.0, 0.25, 0.08
    VP = np.full((nx, nz), vp)
    VP[nx // 2:] *= 1.1
    cs = compute_stiffness(VP, VP / np.sqrt(3), rho, eps, dlt, tilt)
    s = np.sqrt(rho / cs[3]).astype(np.float32)
    S = oracle.expand_boundary(s)
    Cq, mn, mx = [], [], []
    for c in cs:
        u, a, b = oracle.compress(oracle.expand_boundary(c.astype(np.float32)))
        Cq.append(u)
        mn.append(a)
        mx.append(b)
    for phi_deg in (0.0, 30.0, 60.0, 85.0):
        phi = np.radians(phi_deg)
        j = np.arange(nx + 6)[:, None]
        i = np.arange(nz + 6)[None, :]
        T = (np.sin(phi) * j * dh + np.cos(phi) * i * dh).astype(np.float32) * np.float32(1e-3)
        T = np.ascontiguousarray(T + np.float32(0.1))
        Sw = S.copy()
        oracle.quasi_slowness(T, Sw, dh, dh, 0, 0, Cq, mn, mx)
        want = 1.0 / _thomsen_qp_velocity(vp, vs, eps, dlt, phi - np.radians(tilt))
        got = Sw[NB + 2:nx // 2 + NB - 2, NB + 2:nz + NB - 2]
        assert np.max(np.abs(got - want) / want) < 1e-4, (phi_deg, tilt)


# ------------------------------------------------------------ picks (A10)
def test_picks_on_node_exact_and_quadratic(oracle, rng):
    nz, nx, dh = 30, 40, 10.0
    j = (np.arange(nx + 6) - NB)[:, None] * dh
    i = (np.arange(nz + 6) - NB)[None, :] * dh
    q = lambda x, z: 0.5 + 1e-3 * x - 2e-3 * z + 3e-6 * x * x + 2e-6 * z * z - 1e-6 * x * z
    T = q(j, i)
    rx = rng.uniform(0, (nx - 1) * dh, 200)
    rz = rng.uniform(0, (nz - 1) * dh, 200)
    p = oracle.picks(T.astype(np.float64), dh, dh, rx, rz)
    assert np.max(np.abs(p - q(rx, rz))) < 1e-12   # Catmull-Rom is exact on quadratics
    T32 = T.astype(np.float32)
    nodes_x = np.arange(0, nx) * dh
    p32 = oracle.picks(T32, dh, dh, nodes_x, np.full(nx, 50.0))
    assert np.array_equal(p32, T32[NB:nx + NB, NB + 5])  # on-node receivers are exact


# ---------------------------------------------------------- ray paths (A11)
def test_ray_lengths_constant_velocity(oracle):
    nz, nx, dh = 50, 120, 10.0
    S = oracle.expand_boundary(np.full((nx, nz), 1 / 2000.0, np.float32))
    sx, sz = 300.0, 0.0
    T = oracle.solve_iso(S, dh, dh, sx, sz)
    rx = np.array([700.0, 900.0, 1100.0], np.float32)
    rz = np.array([0.0, 200.0, 400.0], np.float32)
    rec, cell, ln = oracle.ray_paths(T, nz, nx, dh, dh, sx, sz, rx, rz)
    for r in range(3):
        dist = np.hypot(rx[r] - sx, rz[r] - sz)
        assert abs(ln[rec == r].sum() - dist) < 3 * 0.2 * dh + 0.02 * dist


# ------------------------------------------------------- discrete adjoint
def test_adjoint_matches_central_differences(oracle, rng):
    nz, nx, dh = 21, 31, 10.0
    s = smooth_model(rng, nz, nx, dh, grad=2.0)
    S = oracle.expand_boundary(s.astype(np.float64))
    rx = np.concatenate([np.arange(2, nx - 2, 3) * dh, [55.0, 123.0]])
    rz = np.concatenate([np.zeros(len(np.arange(2, nx - 2, 3))), [100.0, 57.0]])
    obs = oracle.picks(oracle.solve_iso(S * 1.02, dh, dh, 150.0, 40.0), dh, dh, rx, rz)
    J, g, _ = oracle.adjoint(S, dh, dh, 150.0, 40.0, rx, rz, obs)

    def Jf(Sx):
        p = oracle.picks(oracle.solve_iso(Sx, dh, dh, 150.0, 40.0), dh, dh, rx, rz)
        return 0.5 * np.sum((p - obs) ** 2)

    assert np.isclose(J, Jf(S), rtol=1e-14)
    for _ in range(3):
        V = rng.standard_normal(S.shape) * S
        h = 1e-6
        fd = (Jf(S + h * V) - Jf(S - h * V)) / (2 * h)
        assert abs(fd - np.sum(g * V)) <= 1e-6 * abs(fd)


def test_batch_misfit_window_fold_matches_central_differences(oracle, rng):
    # Window shots fold their padded gradient onto the window's model cells;
    # FP64 central differences of the batch misfit pin that fold.
    nz, nx, dh = 17, 40, 10.0
    s = smooth_model(rng, nz, nx, dh, grad=2.0).astype(np.float64)
    shots, rx, rz = [], [], []
    for sx, sz, x0, w in ((200.0, 0.0, 0, 40), (120.0, 30.0, 5, 20), (300.0, 0.0, 18, 22)):
        xs = np.arange(x0 + 1, x0 + w - 1, 4) * dh + 3.0
        shots.append((sx, sz, x0, w, len(rx), len(xs)))
        rx.extend(xs)
        rz.extend([12.0] * len(xs))
    rx, rz = np.array(rx), np.array(rz)
    _, _, obs = oracle.batch_misfit(s * 1.03, dh, dh, shots, rx, rz, np.zeros(len(rx)), grad=False)
    J, g, _ = oracle.batch_misfit(s, dh, dh, shots, rx, rz, obs)
    assert np.isclose(J, oracle.batch_misfit(s, dh, dh, shots, rx, rz, obs, grad=False)[0], rtol=1e-14)
    for _ in range(3):
        V = rng.standard_normal(s.shape) * s
        h = 1e-6
        fd = (oracle.batch_misfit(s + h * V, dh, dh, shots, rx, rz, obs, grad=False)[0]
              - oracle.batch_misfit(s - h * V, dh, dh, shots, rx, rz, obs, grad=False)[0]) / (2 * h)
        assert abs(fd - np.sum(g * V)) <= 1e-6 * abs(fd)
    # one full-model shot reproduces the single-shot adjoint folded by reduce_boundary_adjoint
    s32 = s.astype(np.float32)
    sel = [shots[0]]
    n0 = shots[0][5]
    _, g1, _ = oracle.batch_misfit(s32, dh, dh, sel, rx[:n0], rz[:n0], obs[:n0], nthreads=1)
    _, gp, _ = oracle.adjoint(oracle.expand_boundary(s32), dh, dh, 200.0, 0.0, rx[:n0].astype(np.float32),
                              rz[:n0].astype(np.float32), obs[:n0].astype(np.float32))
    assert np.array_equal(g1, oracle.reduce_boundary_adjoint(gp))


# ---------------------------------------------- tomography operators (f1)
@pytest.mark.parametrize("order", [1, 2, 3])
def test_tikhonov_operator_on_polynomials(oracle, order):
    # The reference's regularisation matrix (inversion.cpp:209-262) is the
    # order-k forward difference of the flattened model: it annihilates
    # polynomials of degree < k and maps n^k to the constant k!.
    from math import factorial
    n = 40
    t = np.arange(n, dtype=np.float64)
    for deg in range(order):
        y = oracle.tikhonov(order, 3.0 - 0.5 * t ** deg)
        assert y.shape == (n - order,)
        assert np.max(np.abs(y)) < 1e-9
    assert np.allclose(oracle.tikhonov(order, t ** order), factorial(order), rtol=0, atol=1e-6)
    assert np.array_equal(oracle.tikhonov(0, t), t)               # order 0: identity
    # the adjoint is the transpose: <D x, y> = <x, D^T y>
    rng = np.random.default_rng(order)
    x, y = rng.standard_normal(n), rng.standard_normal(n - order)
    assert np.isclose(oracle.tikhonov(order, x) @ y, x @ oracle.tikhonov_adjoint(order, y, n), rtol=1e-12)
    assert np.array_equal(oracle.tikhonov(1, np.eye(3)[1]), [1.0, -1.0])  # df = [-1, 1]


def test_model_smoothing_closed_forms(oracle, rng):
    # Gaussian smoothing of the model update (inversion.cpp:347-447): the
    # kernel sums to one and is symmetric, so a linear field is unchanged
    # wherever the kernel stays inside the model; the model is embedded in
    # zeros, so near an edge a constant c becomes c times the kernel mass
    # inside the model; stdv 0 is the identity.
    nz, nx, stdv, ns = 30, 40, 2.0, 7
    mid = ns // 2
    x = np.arange(nx)[:, None].astype(np.float32)
    z = np.arange(nz)[None, :].astype(np.float32)
    lin = (1.0 + 0.01 * x - 0.02 * z).astype(np.float32) * np.ones((nx, nz), np.float32)
    out = oracle.smooth_model(lin, stdv, ns)
    inner = (slice(mid, nx - mid), slice(mid, nz - mid))
    assert np.max(np.abs(out[inner] - lin[inner])) < 2e-6
    c = np.full((nx, nz), 3.0, np.float32)
    out = oracle.smooth_model(c, stdv, ns)
    k1 = np.exp(-np.arange(-mid, mid + 1) ** 2 / (2 * stdv ** 2))     # separable Gaussian weights
    k1 = k1 / k1.sum()
    for j, i in ((0, 0), (1, 5), (nx - 1, nz - 2), (20, 0)):
        wx = sum(k1[q + mid] for q in range(-mid, mid + 1) if 0 <= j + q < nx)
        wz = sum(k1[q + mid] for q in range(-mid, mid + 1) if 0 <= i + q < nz)
        assert abs(out[j, i] - 3.0 * wx * wz) < 1e-5, (j, i)
    m = rng.standard_normal((nx, nz)).astype(np.float32)
    assert np.array_equal(oracle.smooth_model(m, 0.0, ns), m)
