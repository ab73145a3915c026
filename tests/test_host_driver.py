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
