"""The C++ modeling driver (seisfat2d_amd/host): parameter-file parsing with
the reference's conventions, geometry and model I/O, error behaviour on bad
inputs (CPU), and end-to-end per-shot / batched runs checked against the
oracle (GPU)."""
import os
import subprocess

import numpy as np
import pytest

from seisfat2d_amd import build

EXE = build.MODELING_EXE


def write_case(tmp, nz=41, nx=61, dh=10.0, nsrc=3, ani=False, rec_extra=""):
    rng = np.random.default_rng(5)
    v = 1800.0 + 3.0 * np.arange(nz)[None, :] * dh + 50.0 * rng.standard_normal((nx, nz))
    s = (1.0 / v).astype(np.float32)
    s.tofile(os.path.join(tmp, "slowness.bin"))
    if ani:
        from seisfat2d_amd.stiffness import compute_stiffness
        cs = compute_stiffness(v, v / np.sqrt(3), 1000.0, 0.1, 0.05, 15.0)
        s = np.sqrt(1000.0 / cs[3]).astype(np.float32)
        s.tofile(os.path.join(tmp, "slowness.bin"))
        for name, c in zip(("C11", "C13", "C15", "C33", "C35", "C55"), cs):
            c.astype(np.float32).tofile(os.path.join(tmp, f"model_{name}.bin"))
    with open(os.path.join(tmp, "SPS.txt"), "w") as f:
        f.write("# sources: x, z\n")
        for k in range(nsrc):
            f.write(f"{100.0 + 150.0 * k}, 0.0\n")
    with open(os.path.join(tmp, "RPS.txt"), "w") as f:
        f.write("# receivers\n")
        for x in np.arange(0, nx, 3) * dh:
            f.write(f"{x}, 0.0\n")
        f.write(rec_extra)
    os.makedirs(os.path.join(tmp, "out"), exist_ok=True)
    params = f"""#-----------------------------------------------
# Model paramenters
#-----------------------------------------------

x_samples = {nx}                       # <int>
z_samples = {nz}                       # <int>

x_spacing = {dh}                       # [m] <float>
z_spacing = {dh}                       # [m] <float>

slowness_file = {tmp}/slowness.bin
Cijkl_folder = {tmp}/model_

SPS = {tmp}/SPS.txt
RPS = {tmp}/RPS.txt

modeling_type = {1 if ani else 0}

modeling_output_folder = {tmp}/out/
"""
    p = os.path.join(tmp, "parameters.txt")
    with open(p, "w") as f:
        f.write(params)
    return p, s


def run(args):
    return subprocess.run([EXE, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)


def test_check_accepts_reference_style_file(tmp_path):
    p, _ = write_case(str(tmp_path))
    r = run([p, "--check"])
    assert r.returncode == 0, r.stderr
    assert "41 x 61" in r.stdout and "3 sources" in r.stdout and "21 receivers" in r.stdout


@pytest.mark.parametrize("mutate,msg", [
    ("short_model", "fewer than"),
    ("missing_key", "z_samples"),
    ("bad_receiver_line", "malformed"),
    ("receiver_outside", "outside the model"),
    ("bad_type", "modeling_type"),
])
def test_check_rejects_bad_inputs(tmp_path, mutate, msg):
    tmp = str(tmp_path)
    p, s = write_case(tmp, rec_extra="oops\n" if mutate == "bad_receiver_line" else
                      ("9999.0, 0.0\n" if mutate == "receiver_outside" else ""))
    txt = open(p).read()
    if mutate == "short_model":
        s[:-5].tofile(os.path.join(tmp, "slowness.bin"))
    if mutate == "missing_key":
        txt = txt.replace("z_samples", "# z_samples")
    if mutate == "bad_type":
        txt = txt.replace("modeling_type = 0", "modeling_type = 7")
    open(p, "w").write(txt)
    r = run([p, "--check"])
    assert r.returncode == 1
    assert msg in r.stderr


def _oracle_picks(oracle, s, dh, sx, rx):
    T = oracle.solve_iso(oracle.expand_boundary(s), dh, dh, sx, 0.0)
    return oracle.picks(T, dh, dh, rx, np.zeros_like(rx))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [[], ["--batch"], ["--schedule", "plain"]])
def test_driver_iso_outputs_match_oracle(tmp_path, oracle, mode):
    tmp = str(tmp_path)
    p, s = write_case(tmp)
    r = run([p, "--quiet", *mode])
    assert r.returncode == 0, r.stderr
    rx = (np.arange(0, 61, 3) * 10.0).astype(np.float32)
    for k in range(3):
        got = np.fromfile(os.path.join(tmp, "out", f"eikonal_iso_nStations{len(rx)}_shot_{k + 1}.bin"), np.float32)
        assert np.array_equal(got, _oracle_picks(oracle, s, 10.0, 100.0 + 150.0 * k, rx)), (mode, k)


@pytest.mark.gpu
def test_driver_ani_outputs_match_oracle(tmp_path, oracle):
    tmp = str(tmp_path)
    p, s = write_case(tmp, ani=True)
    for mode in ([], ["--batch"]):
        r = run([p, "--quiet", *mode])
        assert r.returncode == 0, r.stderr
        S = oracle.expand_boundary(s)
        Cq, mn, mx = [], [], []
        for name in ("C11", "C13", "C15", "C33", "C35", "C55"):
            c = np.fromfile(os.path.join(tmp, f"model_{name}.bin"), np.float32).reshape(61, 41)
            u, a, b = oracle.compress(oracle.expand_boundary(c))
            Cq.append(u)
            mn.append(a)
            mx.append(b)
        rx = (np.arange(0, 61, 3) * 10.0).astype(np.float32)
        for k in range(3):
            T = oracle.solve_ani(S, 10.0, 10.0, 100.0 + 150.0 * k, 0.0, Cq, mn, mx)[0]
            want = oracle.picks(T, 10.0, 10.0, rx, np.zeros_like(rx))
            got = np.fromfile(os.path.join(tmp, "out", f"eikonal_ani_nStations{len(rx)}_shot_{k + 1}.bin"),
                              np.float32)
            assert np.array_equal(got, want), (mode, k)


def _tables_case(tmp, ani):
    p, s = write_case(tmp, ani=ani)
    os.makedirs(os.path.join(tmp, "tables"), exist_ok=True)
    with open(p, "a") as f:
        f.write(f"\nmig_tables_folder = {tmp}/tables/\nanisotropy = {'true' if ani else 'false'}\n")
    return p, s


@pytest.mark.gpu
@pytest.mark.parametrize("ani", [False, True])
def test_tables_driver_matches_oracle(tmp_path, oracle, ani):
    """seisfat2d_tables writes the migration tables the reference writes
    (src/migration/migration.cu:116-171): one padded traveltime grid per
    source (eikonal_src_<k>.bin) and per receiver (eikonal_rec_<k>.bin), k from
    1, each equal to the oracle's single-pass solve from that position; the
    positions run in batches (chunk 8 here: several batches and a partial one)."""
    tmp = str(tmp_path)
    p, s = _tables_case(tmp, ani)
    r = subprocess.run([build.TABLES_EXE, p, "--chunk", "8"], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    nz, nx, dh = 41, 61, 10.0
    S = oracle.expand_boundary(s)
    if ani:
        Cq, mn, mx = [], [], []
        for name in ("C11", "C13", "C15", "C33", "C35", "C55"):
            c = np.fromfile(os.path.join(tmp, f"model_{name}.bin"), np.float32).reshape(nx, nz)
            u, a, b = oracle.compress(oracle.expand_boundary(c))
            Cq.append(u)
            mn.append(a)
            mx.append(b)
        solve = lambda x, z: oracle.solve_ani(S, dh, dh, x, z, Cq, mn, mx)[0]
    else:
        solve = lambda x, z: oracle.solve_iso(S, dh, dh, x, z)
    srcs = [(100.0 + 150.0 * k, 0.0) for k in range(3)]
    recs = [(float(x), 0.0) for x in np.arange(0, nx, 3) * dh]
    for prefix, pos in (("eikonal_src_", srcs), ("eikonal_rec_", recs)):
        for k, (x, z) in enumerate(pos):
            got = np.fromfile(os.path.join(tmp, "tables", f"{prefix}{k + 1}.bin"), np.float32)
            assert got.size == (nz + 6) * (nx + 6)
            assert np.array_equal(got.reshape(nx + 6, nz + 6), solve(x, z)), (prefix, k)
    assert not os.path.exists(os.path.join(tmp, "tables", f"eikonal_rec_{len(recs) + 1}.bin"))


# -- the migration driver (src/migration_main.cpp) ------------------------------------

MIG_KEYS = """
mig_type = {t}
anisotropy = false
time_samples = {nt}
time_spacing = 2e-3
max_frequency = 30.0
mig_aperture = 30
mig_max_angle = 60
mig_angle_spacing = 3.0
mig_max_offset = 300
mig_max_iteration = 2
mig_smooth_lambda = 0.001
mig_data_folder = {tmp}/data/
mig_data_prefix = mig_input_shot_
mig_tables_folder = {tmp}/tables/
mig_seismic_folder = {tmp}/seismic/
mig_residuo_folder = {tmp}/residuo/
"""


def _mig_case(tmp, t, nt=300, drop=None, ani=False):
    p, s = write_case(tmp, ani=ani)
    keys = MIG_KEYS.format(t=t, nt=nt, tmp=tmp)
    if ani:
        keys = keys.replace("anisotropy = false", "anisotropy = true")
    if drop:
        keys = "\n".join(l for l in keys.splitlines() if not l.startswith(drop))
    with open(p, "a") as f:
        f.write(keys)
    for d in ("data", "tables", "seismic", "residuo"):
        os.makedirs(os.path.join(tmp, d), exist_ok=True)
    nrec = len(np.arange(0, 61, 3))
    rng = np.random.default_rng(t)
    for k in range(3):
        rng.standard_normal((nrec, nt)).astype(np.float32).tofile(os.path.join(tmp, "data", f"mig_input_shot_{k + 1}.bin"))
    return p, s


@pytest.mark.parametrize("drop,msg", [("mig_data_folder", "mig_data_folder"), ("mig_type", "mig_type"),
                                      ("mig_max_iteration", "mig_max_iteration")])
def test_migration_driver_rejects_missing_keys(tmp_path, drop, msg):
    """A missing key fails before any device work, naming the key (LS keys
    only for the least-squares types)."""
    tmp = str(tmp_path)
    p, _ = _mig_case(tmp, 2, drop=drop)
    r = subprocess.run([build.MIGRATION_EXE, p], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=60)
    assert r.returncode != 0 and msg in r.stderr, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("t", [0, 1, 2, 3])
def test_migration_driver_matches_oracle(tmp_path, oracle, t):
    """seisfat2d_migration runs the reference's four migrations from a
    parameter file: tables solved on the device (and written as
    eikonal_{src,rec}_<k>.bin), the seismic files read per source, the result
    written under the reference's name; checked against the oracle's
    restatement on the oracle's own tables."""
    from seisfat2d_amd import migration as mig
    tmp = str(tmp_path)
    nt = 300
    p, s = _mig_case(tmp, t, nt)
    r = subprocess.run([build.MIGRATION_EXE, p], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr + r.stdout
    nz, nx, dh = 41, 61, 10.0
    S = oracle.expand_boundary(s)
    xs = np.array([100.0 + 150.0 * k for k in range(3)], np.float32)
    xr = (np.arange(0, nx, 3) * dh).astype(np.float32)
    ts = np.stack([oracle.solve_iso(S, dh, dh, float(x), 0.0) for x in xs])
    tr = np.stack([oracle.solve_iso(S, dh, dh, float(x), 0.0) for x in xr])
    for k in range(len(xr)):
        got = np.fromfile(os.path.join(tmp, "tables", f"eikonal_rec_{k + 1}.bin"), np.float32)
        assert np.array_equal(got.reshape(nx + 6, nz + 6), tr[k])
    nang, dcmp, min_cmp, ncmp = mig.gathers(xs, xr, 3.0, 60.0)
    src, rec, cmp, cid = mig.trace_list(xs, xr, 300.0, min_cmp, dcmp)
    seismic = np.stack([np.fromfile(os.path.join(tmp, "data", f"mig_input_shot_{k + 1}.bin"), np.float32)
                        .reshape(len(xr), nt) for k in range(3)])
    dobs = np.ascontiguousarray(seismic[src, rec])
    w = oracle.ricker(2e-3, 30.0)
    angle = t in (1, 3)
    domain = 1 if angle else 0
    shape = (ncmp, nang, nz) if angle else (nx, nz)
    name = ["IDKDM", "ADKDM", "IDLSKDM", "ADLSKDM"][t]
    size = f"{nz}x{ncmp}x{nang}" if angle else f"{nz}x{nx}"
    out = os.path.join(tmp, "seismic", f"{name}_result_{size}{'_iteration_2' if t >= 2 else ''}.bin")
    got = np.fromfile(out, np.float32).reshape(shape)
    op = dict(dt=2e-3, dx=dh, dz=dh, aperture=30.0, da=3.0, nang=nang)
    if t < 2:
        d = np.ascontiguousarray(oracle.trace_convolve(dobs, w, True))
        want = oracle.kdm(domain, True, S, ts, tr, src, rec, cmp, cid, d, np.zeros(shape, np.float32), **op)
        bound = oracle.kdm(domain, True, S, ts, tr, src, rec, cmp, cid, np.abs(d), np.zeros(shape, np.float32), **op)
        bad = np.abs(got - want) > 1e-5 * bound + 1e-30
        assert np.abs(want).max() > 0 and bad.mean() <= (0 if not angle else 1e-3), bad.sum()
    else:
        want, res = oracle.lskdm(domain, S, ts, tr, src, rec, cmp, cid, dobs, w, 2, 0.001, 2e-3, dh, dh, 30.0,
                                 shape, 3.0, nang)
        got_res = np.loadtxt(os.path.join(tmp, "residuo", f"{name}_convergence_2_iterations.txt"))
        assert np.allclose(got_res, res, rtol=1e-4, atol=0), (got_res, res)
        assert np.linalg.norm(got - want) <= 1e-3 * np.linalg.norm(want)


@pytest.mark.gpu
def test_migration_driver_anisotropic_tables(tmp_path, oracle):
    """anisotropy = true: the tables are qP solves (Eikonal_ANI, as
    src/migration/migration.cu:26-27 selects), each equal to the oracle's, and
    the image the oracle's migration on those tables."""
    from seisfat2d_amd import migration as mig
    tmp = str(tmp_path)
    nt = 300
    p, s = _mig_case(tmp, 0, nt, ani=True)
    r = subprocess.run([build.MIGRATION_EXE, p], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr + r.stdout
    nz, nx, dh = 41, 61, 10.0
    S = oracle.expand_boundary(s)
    Cq, mn, mx = [], [], []
    for name in ("C11", "C13", "C15", "C33", "C35", "C55"):
        c = np.fromfile(os.path.join(tmp, f"model_{name}.bin"), np.float32).reshape(nx, nz)
        u, a, b = oracle.compress(oracle.expand_boundary(c))
        Cq.append(u)
        mn.append(a)
        mx.append(b)
    xs = np.array([100.0 + 150.0 * k for k in range(3)], np.float32)
    xr = (np.arange(0, nx, 3) * dh).astype(np.float32)
    ts = np.stack([oracle.solve_ani(S, dh, dh, float(x), 0.0, Cq, mn, mx)[0] for x in xs])
    tr = np.stack([oracle.solve_ani(S, dh, dh, float(x), 0.0, Cq, mn, mx)[0] for x in xr])
    for k in range(len(xs)):
        got = np.fromfile(os.path.join(tmp, "tables", f"eikonal_src_{k + 1}.bin"), np.float32)
        assert np.array_equal(got.reshape(nx + 6, nz + 6), ts[k])
    nang, dcmp, min_cmp, ncmp = mig.gathers(xs, xr, 3.0, 60.0)
    src, rec, cmp, cid = mig.trace_list(xs, xr, 300.0, min_cmp, dcmp)
    seismic = np.stack([np.fromfile(os.path.join(tmp, "data", f"mig_input_shot_{k + 1}.bin"), np.float32)
                        .reshape(len(xr), nt) for k in range(3)])
    d = np.ascontiguousarray(oracle.trace_convolve(np.ascontiguousarray(seismic[src, rec]), oracle.ricker(2e-3, 30.0),
                                                   True))
    op = dict(dt=2e-3, dx=dh, dz=dh, aperture=30.0, da=3.0, nang=nang)
    want = oracle.kdm(0, True, S, ts, tr, src, rec, cmp, None, d, np.zeros((nx, nz), np.float32), **op)
    bound = oracle.kdm(0, True, S, ts, tr, src, rec, cmp, None, np.abs(d), np.zeros((nx, nz), np.float32), **op)
    got = np.fromfile(os.path.join(tmp, "seismic", f"IDKDM_result_{nz}x{nx}.bin"), np.float32).reshape(nx, nz)
    assert np.abs(want).max() > 0 and not (np.abs(got - want) > 1e-5 * bound + 1e-30).any()
