"""First-arrival tomography driver (seisfat2d_amd/host/inversion.cpp, the
seisfat2d_inversion executable): the reference's Inversion / Tomography_ISO
loop (src/inversion_main.cpp, src/inversion/inversion.cpp,
tomography_iso.cpp) on the adjoint-state gradient.

CPU: parameter errors are reported before any device work.  GPU: five
iterations of the 128-shot x 401-receiver tomography workload (301 x 1201,
dh 10 m; observed picks from the FP64 oracle on the true model) lower the
residual at every iteration, and every model update equals the oracle's step
(Tikhonov gradient + model smoothing of the reference, restated in the
oracle) computed from the driver's own gradient, coverage and step length."""
import os
import subprocess

import numpy as np
import pytest

from seisfat2d_amd import build, models

NB = 3


def write_case(tmp, w, s_start, obs, it=5, tk_param="1e-3", extra=""):
    s_start.astype(np.float32).tofile(os.path.join(tmp, "start.bin"))
    nsrc = len(w.shots)
    nrec = w.shots[0][5]
    with open(os.path.join(tmp, "SPS.txt"), "w") as f:
        for sh in w.shots:
            f.write(f"{sh[0]!r}, {sh[1]!r}\n")
    with open(os.path.join(tmp, "RPS.txt"), "w") as f:
        for x, z in zip(w.rx[:nrec], w.rz[:nrec]):
            f.write(f"{float(x)!r}, {float(z)!r}\n")
    for d in ("obs", "res", "out", "grad"):
        os.makedirs(os.path.join(tmp, d), exist_ok=True)
    for k in range(nsrc):
        obs[k * nrec:(k + 1) * nrec].astype(np.float32).tofile(os.path.join(tmp, "obs", f"obs_shot_{k + 1}.bin"))
    p = os.path.join(tmp, "parameters.txt")
    with open(p, "w") as f:
        f.write(f"""# model
x_samples = {w.nx}
z_samples = {w.nz}
x_spacing = {w.dx}
z_spacing = {w.dz}
slowness_file = {tmp}/start.bin
SPS = {tmp}/SPS.txt
RPS = {tmp}/RPS.txt

# inversion (the keys of the reference's parameter file)
inv_type = 0
tk_order = 2
tk_param = {tk_param}
inv_max_it = {it}
smooth_per_iteration = true
gaussian_filter_stdv = 5.0
gaussian_filter_samples = 11
inv_residuo_folder = {tmp}/res/
inv_output_folder = {tmp}/out/
inv_data_folder = {tmp}/obs/
inv_data_prefix = obs_shot_
inv_gradient_folder = {tmp}/grad/
{extra}""")
    return p


def test_parameter_errors_before_device_work(tmp_path):
    w = models.c5(nshots=2)
    p = write_case(str(tmp_path), w, w.s, np.zeros(2 * 401, np.float32))
    txt = open(p).read()
    open(p, "w").write(txt.replace("inv_max_it", "# inv_max_it"))
    r = subprocess.run([build.INVERSION_EXE, p], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "inv_max_it" in r.stderr
    open(p, "w").write(txt.replace("gaussian_filter_samples = 11", "gaussian_filter_samples = 10"))
    r = subprocess.run([build.INVERSION_EXE, p], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "odd" in r.stderr
    open(p, "w").write(txt.replace("inv_type = 0", "inv_type = 1"))
    r = subprocess.run([build.INVERSION_EXE, p], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "inversion_type" in r.stderr


@pytest.mark.gpu
def test_c5_tomography_loop(tmp_path, oracle):
    tmp = str(tmp_path)
    w = models.c5()
    s_true = w.extra["s_true"]
    nrec = w.shots[0][5]
    # observed picks: the FP64 oracle on the true model
    _, _, obs = oracle.batch_misfit(s_true.astype(np.float64), w.dx, w.dz, w.shots, w.rx.astype(np.float64),
                                    w.rz.astype(np.float64), np.zeros(len(w.rx)), grad=False)
    niter = 5
    p = write_case(tmp, w, w.s, obs.astype(np.float32), it=niter)
    r = subprocess.run([build.INVERSION_EXE, p], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    res = np.loadtxt(os.path.join(tmp, "res", f"tomography_iso_convergence_{niter}_iterations.txt"))
    print("residuo per iteration:", res)
    assert len(res) == niter + 1
    assert np.all(np.diff(res) < 0), res                   # the residual falls at every iteration
    tk = float(np.float32(1e-3))
    s0 = w.s.astype(np.float32)
    s = s0.copy()
    for it in range(1, niter + 1):
        g = np.fromfile(os.path.join(tmp, "grad", f"gradient_{it}.bin"), np.float64).reshape(w.nx, w.nz)
        R = np.fromfile(os.path.join(tmp, "grad", f"coverage_{it}.bin"), np.float64).reshape(w.nx, w.nz)
        dS = np.fromfile(os.path.join(tmp, "grad", f"update_{it}.bin"), np.float32).reshape(w.nx, w.nz)
        step = dict(l.split() for l in open(os.path.join(tmp, "grad", f"step_{it}.txt")))
        alpha = np.float32(float(step["alpha"]))
        assert alpha > 0 and float(step["phi_after"]) < float(step["phi_before"])
        # the oracle's step on the same gradient: Tikhonov term of the
        # reference's operator on the coverage-weighted model change, then the
        # reference's smoothing of the negative gradient
        n = w.nx * w.nz
        m = (R * (s.astype(np.float64) - s0.astype(np.float64))).ravel()
        x = oracle.tikhonov_adjoint(2, oracle.tikhonov(2, m), n).reshape(w.nx, w.nz)
        gt = g + tk * tk * R * x
        pdir = oracle.smooth_model((-gt).astype(np.float32), 5.0, 11)
        assert np.sum(gt * pdir) < 0                      # a descent direction
        want = (alpha * pdir).astype(np.float32)
        assert np.array_equal(dS, want), (it, np.abs(dS - want).max())
        s = (s + dS).astype(np.float32)
    vp = np.fromfile(os.path.join(tmp, "out", f"tomography_iso_final_model_vp_{w.nz}x{w.nx}.bin"), np.float32)
    assert np.array_equal(vp.reshape(w.nx, w.nz), (1.0 / s).astype(np.float32))
    # the model moved towards the true one
    assert np.linalg.norm(s - s_true) < np.linalg.norm(s0 - s_true)
