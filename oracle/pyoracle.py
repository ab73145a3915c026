"""ctypes front end of the CPU oracle (test infrastructure only).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker for the HIP path; the product package never imports it.  The
library is built by ``make -C oracle`` (or ``__graft_entry__.build()``).

All grids are numpy arrays in the reference's z-fastest layout: a padded grid
of nzz x nxx nodes is held as an array of shape (nxx, nzz) in C order, so that
``A[j, i]`` is node (z=i, x=j) and ``A.ravel()[i + j*nzz]`` is the reference's
``idx = i + j*nzz`` (src/modeling/modeling.cu:209-212,324).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

NB = 3
_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
ip = C.POINTER(C.c_int)


def lib():
    global _LIB
    if _LIB is None:
        # SEISFAT_ORACLE_LIB points the checker at another build of oracle.cpp
        # (the mutation tests load deliberately broken builds this way)
        path = os.environ.get("SEISFAT_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        i, f, d = C.c_int, C.c_float, C.c_double
        sig = {
            "orc_expand_boundary_f32": (None, [f32p, f32p, i, i]),
            "orc_expand_boundary_f64": (None, [f64p, f64p, i, i]),
            "orc_reduce_boundary_f32": (None, [f32p, f32p, i, i]),
            "orc_init_f32": (None, [f32p, f32p, i, i, f, f, f, f, ip, ip]),
            "orc_init_f64": (None, [f64p, f64p, i, i, d, d, d, d, ip, ip]),
            "orc_fsm_f32": (i, [f32p, f32p, i, i, f, f, i, i, i]),
            "orc_fsm_f64": (i, [f64p, f64p, i, i, d, d, i, i, i]),
            "orc_sweep_f32": (None, [f32p, f32p, i, i, f, f, i]),
            "orc_solve_iso_f32": (i, [f32p, f32p, i, i, f, f, f, f, i, i]),
            "orc_solve_iso_f64": (i, [f64p, f64p, i, i, d, d, d, d, i, i]),
            "orc_solve_windows_f32": (None, [f32p, i, i, f, f, i, f32p, f32p, i32p, i32p, i64p, f32p, i]),
            "orc_compress": (None, [f32p, u16p, C.c_longlong, C.POINTER(f), C.POINTER(f)]),
            "orc_quasi_slowness": (None, [f32p, f32p, i, i, f, f, i, i, u16p, u16p, u16p, u16p, u16p, u16p, f32p, f32p]),
            "orc_solve_ani_f32": (None, [f32p, f32p, f32p, i, i, f, f, f, f, u16p, u16p, u16p, u16p, u16p, u16p, f32p, f32p]),
            "orc_picks_f32": (None, [f32p, i, f, f, f32p, f32p, i, f32p]),
            "orc_picks_f64": (None, [f64p, i, d, d, f64p, f64p, i, f64p]),
            "orc_adjoint_f32": (d, [f32p, i, i, f, f, f, f, f32p, f32p, i, f32p, f64p, f32p]),
            "orc_adjoint_f64": (d, [f64p, i, i, d, d, d, d, f64p, f64p, i, f64p, f64p, f64p]),
            "orc_branch_counts_f32": (None, [f32p, i, i, f, f, f, f, i64p]),
            "orc_sweep_counted_f32": (None, [f32p, f32p, i, i, f, f, i, i64p]),
            "orc_tikhonov_apply": (None, [i, f64p, C.c_longlong, f64p]),
            "orc_tikhonov_adjoint": (None, [i, f64p, C.c_longlong, f64p]),
            "orc_smooth_model": (None, [f32p, f32p, i, i, f, i]),
            "orc_ray_paths": (i, [f32p, i, i, f, f, f, f, f32p, f32p, i, i, i32p, i32p, f32p, i]),
            "orc_batch_misfit_f32": (d, [f32p, i, i, f, f, i, f32p, f32p, i32p, i32p, i32p, i32p, f32p, f32p,
                                         f32p, C.c_void_p, C.c_void_p, i]),
            "orc_batch_misfit_f64": (d, [f64p, i, i, d, d, i, f64p, f64p, i32p, i32p, i32p, i32p, f64p, f64p,
                                         f64p, C.c_void_p, C.c_void_p, i]),
            "orc_ricker": (i, [f, f, f32p, i]),
            "orc_trace_convolve": (None, [i, i, i, f32p, i, f32p, f32p]),
            "orc_kdm": (None, [i, i, f32p, f32p, f32p, i, i32p, i32p, f32p, C.c_void_p, f32p, f32p, i, f, f, f,
                               i, i, f, f, i, i]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _suffix(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def expand_boundary(s):
    """Pad an (nx, nz) model by nb=3 with edge replication (modeling.cu:204-232)."""
    s = np.ascontiguousarray(s)
    nx, nz = s.shape
    out = np.zeros((nx + 2 * NB, nz + 2 * NB), s.dtype)
    getattr(lib(), "orc_expand_boundary_" + _suffix(s.dtype))(s, out, nz, nx)
    return out


def reduce_boundary(S):
    S = np.ascontiguousarray(S, np.float32)
    nxx, nzz = S.shape
    out = np.zeros((nxx - 2 * NB, nzz - 2 * NB), np.float32)
    lib().orc_reduce_boundary_f32(S, out, nzz - 2 * NB, nxx - 2 * NB)
    return out


def reduce_boundary_adjoint(G):
    """Adjoint of expand_boundary: sum each padded cell into the model cell it copies."""
    nxx, nzz = G.shape
    nx, nz = nxx - 2 * NB, nzz - 2 * NB
    jj = np.clip(np.arange(nxx) - NB, 0, nx - 1)
    ii = np.clip(np.arange(nzz) - NB, 0, nz - 1)
    out = np.zeros((nx, nz), np.float64)
    np.add.at(out, (jj[:, None], ii[None, :]), G)
    return out


def init_time(S, dx, dz, sx, sz):
    S = np.ascontiguousarray(S)
    nxx, nzz = S.shape
    T = np.empty_like(S)
    a, b = C.c_int(), C.c_int()
    getattr(lib(), "orc_init_" + _suffix(S.dtype))(T, S, nzz, nxx, dx, dz, sx, sz, C.byref(a), C.byref(b))
    return T, a.value, b.value


def fsm(T, S, dx, dz, mode=0, max_passes=1, schedule=0):
    """In-place sweeps on T.  mode 0 = one pass (reference), 1 = until unchanged."""
    assert T.dtype == S.dtype and T.flags.c_contiguous and S.flags.c_contiguous
    nxx, nzz = S.shape
    return getattr(lib(), "orc_fsm_" + _suffix(S.dtype))(T, S, nzz, nxx, dx, dz, mode, max_passes, schedule)


def sweep(T, S, dx, dz, s):
    nxx, nzz = S.shape
    lib().orc_sweep_f32(T, S, nzz, nxx, dx, dz, s)


KINDS = ("old", "1d_z", "1d_x", "A", "B", "C", "const")


def branch_counts(S, dx, dz, sx, sz):
    """Which candidate decided each node update of one single-pass solve:
    a dict kind -> count (kinds as KINDS)."""
    S = np.ascontiguousarray(S, np.float32)
    nxx, nzz = S.shape
    out = np.zeros(7, np.int64)
    lib().orc_branch_counts_f32(S, nzz, nxx, dx, dz, sx, sz, out)
    return dict(zip(KINDS, out.tolist()))


def sweep_counted(T, S, dx, dz, s):
    """One raster sweep on T in place; returns the deciding-candidate counts."""
    nxx, nzz = S.shape
    out = np.zeros(7, np.int64)
    lib().orc_sweep_counted_f32(T, S, nzz, nxx, dx, dz, s, out)
    return dict(zip(KINDS, out.tolist()))


def solve_iso(S, dx, dz, sx, sz, mode=0, max_passes=1):
    """Eikonal_ISO::time_propagation on a padded slowness grid; returns padded T."""
    S = np.ascontiguousarray(S)
    nxx, nzz = S.shape
    T = np.empty_like(S)
    passes = getattr(lib(), "orc_solve_iso_" + _suffix(S.dtype))(T, S, nzz, nxx, dx, dz, sx, sz, mode, max_passes)
    return T if mode == 0 else (T, passes)


def solve_windows(s, dx, dz, sx, sz, x0, w, nthreads=1):
    """Shot-parallel iso solves on column windows of an unpadded (nx, nz) model."""
    s = np.ascontiguousarray(s, np.float32)
    nx, nz = s.shape
    sx = np.ascontiguousarray(sx, np.float32)
    sz = np.ascontiguousarray(sz, np.float32)
    x0 = np.ascontiguousarray(x0, np.int32)
    w = np.ascontiguousarray(w, np.int32)
    sizes = (nz + 2 * NB) * (w.astype(np.int64) + 2 * NB)
    toff = np.ascontiguousarray(np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64))
    T = np.empty(int(sizes.sum()), np.float32)
    lib().orc_solve_windows_f32(s, nz, nx, dx, dz, len(w), sx, sz, x0, w, toff, T, nthreads)
    return [T[o:o + n].reshape(int(wk) + 2 * NB, nz + 2 * NB) for o, n, wk in zip(toff, sizes, w)]


def compress(c):
    c = np.ascontiguousarray(c, np.float32)
    out = np.empty(c.shape, np.uint16)
    mx, mn = C.c_float(), C.c_float()
    lib().orc_compress(c.ravel(), out.ravel(), c.size, C.byref(mx), C.byref(mn))
    return out, mn.value, mx.value


def quasi_slowness(T, S, dx, dz, sIdx, sIdz, Cq, mn, mx):
    """get_quasi_slowness on padded T/S (S updated in place); Cq = 6 uint16 grids."""
    nxx, nzz = S.shape
    mn = np.ascontiguousarray(mn, np.float32)
    mx = np.ascontiguousarray(mx, np.float32)
    lib().orc_quasi_slowness(T, S, nzz, nxx, dx, dz, sIdx, sIdz, *[np.ascontiguousarray(c) for c in Cq], mn, mx)


def solve_ani(S_iso, dx, dz, sx, sz, Cq, mn, mx):
    S_iso = np.ascontiguousarray(S_iso, np.float32)
    nxx, nzz = S_iso.shape
    T = np.empty_like(S_iso)
    S_work = np.empty_like(S_iso)
    lib().orc_solve_ani_f32(T, S_iso, S_work, nzz, nxx, dx, dz, sx, sz,
                            *[np.ascontiguousarray(c) for c in Cq],
                            np.ascontiguousarray(mn, np.float32), np.ascontiguousarray(mx, np.float32))
    return T, S_work


def picks(T, dx, dz, rx, rz):
    T = np.ascontiguousarray(T)
    nxx, nzz = T.shape
    dt = T.dtype
    rx = np.ascontiguousarray(rx, dt)
    rz = np.ascontiguousarray(rz, dt)
    out = np.empty(len(rx), dt)
    getattr(lib(), "orc_picks_" + _suffix(dt))(T, nzz, dx, dz, rx, rz, len(rx), out)
    return out


def adjoint(S, dx, dz, sx, sz, rx, rz, obs):
    """Discrete adjoint of one single-pass solve.  Returns (J, dJ/dS padded, picks)."""
    S = np.ascontiguousarray(S)
    dt = S.dtype
    nxx, nzz = S.shape
    g = np.zeros(S.shape, np.float64)
    pk = np.empty(len(rx), dt)
    J = getattr(lib(), "orc_adjoint_" + _suffix(dt))(
        S, nzz, nxx, dx, dz, sx, sz, np.ascontiguousarray(rx, dt), np.ascontiguousarray(rz, dt),
        len(rx), np.ascontiguousarray(obs, dt), g, pk)
    return J, g, pk


def batch_misfit(s, dx, dz, shots, rx, rz, obs, grad=True, nthreads=16):
    """Window-shot batch (GPU batch semantics) on an unpadded (nx, nz) model of
    dtype float32 or float64.  shots: (sx, sz, x0, width, rec_offset, nrec)
    tuples.  Returns (J, dJ/ds as (nx, nz) float64 or None, picks)."""
    s = np.ascontiguousarray(s)
    dt = s.dtype
    nx, nz = s.shape
    col = lambda k, t: np.ascontiguousarray([sh[k] for sh in shots], t)
    g = np.zeros((nx, nz), np.float64) if grad else None
    pk = np.zeros(len(rx), dt)
    J = getattr(lib(), "orc_batch_misfit_" + _suffix(dt))(
        s, nz, nx, dx, dz, len(shots), col(0, dt), col(1, dt), col(2, np.int32), col(3, np.int32),
        col(4, np.int32), col(5, np.int32), np.ascontiguousarray(rx, dt), np.ascontiguousarray(rz, dt),
        np.ascontiguousarray(obs, dt), None if g is None else g.ctypes.data_as(C.c_void_p),
        pk.ctypes.data_as(C.c_void_p), nthreads)
    return J, g, pk


def ray_paths(T, nz, nx, dx, dz, sx, sz, rx, rz, max_steps=200000, cap=None):
    """Ray-traced sensitivity rows for one shot (inversion.cpp:95-177) as COO."""
    T = np.ascontiguousarray(T, np.float32)
    rx = np.ascontiguousarray(rx, np.float32)
    rz = np.ascontiguousarray(rz, np.float32)
    cap = cap or 64 * (nx + nz) * max(1, len(rx))
    rec = np.empty(cap, np.int32)
    cell = np.empty(cap, np.int32)
    ln = np.empty(cap, np.float32)
    n = lib().orc_ray_paths(T, nz, nx, dx, dz, sx, sz, rx, rz, len(rx), max_steps, rec, cell, ln, cap)
    if n < 0:
        raise RuntimeError(f"ray from receiver {-1 - n} did not reach the source cell")
    if n > cap:
        raise RuntimeError("COO capacity exceeded")
    return rec[:n], cell[:n], ln[:n]


def tikhonov(order, x):
    """The reference's Tikhonov operator D (inversion.cpp:209-262) on the
    flattened model x: n - order rows."""
    x = np.ascontiguousarray(x, np.float64).ravel()
    y = np.zeros(len(x) - max(order, 0))
    lib().orc_tikhonov_apply(order, x, len(x), y)
    return y


def tikhonov_adjoint(order, y, n):
    y = np.ascontiguousarray(y, np.float64).ravel()
    x = np.zeros(n)
    lib().orc_tikhonov_adjoint(order, y, n, x)
    return x


def smooth_model(m, stdv, samples):
    """Inversion::model_smoothing (inversion.cpp:347-447) of an (nx, nz) model."""
    m = np.ascontiguousarray(m, np.float32)
    nx, nz = m.shape
    out = np.empty_like(m)
    lib().orc_smooth_model(m, out, nz, nx, stdv, samples)
    return out


# ---------------------------------------------------------------------------
# Kirchhoff depth migration (src/migration/)

def ricker(dt, fmax):
    """Migration::set_wavelet (src/migration/migration.cu:54-73)."""
    n = lib().orc_ricker(dt, fmax, np.zeros(1, np.float32), 0)
    out = np.zeros(max(n, 1), np.float32)
    lib().orc_ricker(dt, fmax, out, n)
    return out[:n]


def trace_convolve(traces, wavelet, adjoint):
    """forward_convolution (adjoint=False) / adjoint_convolution (adjoint=True)
    of src/migration/migration.cu:197-261 on every row of traces (ntr, nt)."""
    x = np.ascontiguousarray(traces, np.float32)
    x2 = x.reshape(-1, x.shape[-1])
    w = np.ascontiguousarray(wavelet, np.float32)
    out = np.empty_like(x2)
    lib().orc_trace_convolve(int(bool(adjoint)), x2.shape[0], x2.shape[1], w, len(w), x2, out)
    return out.reshape(x.shape)


def kdm(domain, adjoint, S, src_tables, rec_tables, isrc, irec, cmp, cmp_id, data, model, dt, dx, dz,
        aperture, da=1.0, nang=1, nthreads=1):
    """The reference's per-trace Kirchhoff kernels (src/migration/migration.cu:398-756)
    for every trace t in order: adjoint adds into model, forward into data
    (both updated in place and returned).  S: padded slowness (nxx, nzz);
    tables (n, nxx, nzz); data (ntr, nt); model (nx, nz) for domain 0,
    (ncmp, nang, nz) for domain 1.  nthreads: threads over image columns for
    the image-domain adjoint (the same result)."""
    S = np.ascontiguousarray(S, np.float32)
    ts = np.ascontiguousarray(src_tables, np.float32)
    tr = np.ascontiguousarray(rec_tables, np.float32)
    nxx, nzz = S.shape
    assert ts.shape[1:] == S.shape and tr.shape[1:] == S.shape
    isrc = np.ascontiguousarray(isrc, np.int32)
    irec = np.ascontiguousarray(irec, np.int32)
    cmp = np.ascontiguousarray(cmp, np.float32)
    ntr = len(isrc)
    assert data.dtype == np.float32 and data.flags.c_contiguous and data.shape[0] == ntr
    assert model.dtype == np.float32 and model.flags.c_contiguous
    assert isrc.max(initial=0) < len(ts) and irec.max(initial=0) < len(tr)
    cid = None
    if domain == 1:
        cid = np.ascontiguousarray(cmp_id, np.int32)
        assert model.ndim == 3 and model.shape[1:] == (nang, nzz - 2 * NB) and model.shape[0] > cid.max(initial=0)
        assert cid.min(initial=0) >= 0
    else:
        assert model.shape == (nxx - 2 * NB, nzz - 2 * NB)
    lib().orc_kdm(domain, int(bool(adjoint)), S, ts, tr, ntr, isrc, irec, cmp,
                  None if cid is None else cid.ctypes.data_as(C.c_void_p), data, model, data.shape[1], dt, dx, dz,
                  nzz, nxx, aperture, da, nang, nthreads)
    return model if adjoint else data


def _seqsum(a):
    """float32 sum in index order (the reference's host loops)."""
    a = np.ascontiguousarray(a, np.float32).ravel()
    return np.cumsum(a, dtype=np.float32)[-1] if a.size else np.float32(0)


def lskdm(domain, S, src_tables, rec_tables, isrc, irec, cmp, cmp_id, dobs, wavelet, max_it, smooth, dt, dx, dz,
          aperture, model_shape, da=1.0, nang=1):
    """LSKDM::kirchhoff_depth_migration (src/migration/LSKDM.cu:3-226) with
    ADLSKDM::regularization (src/migration/ADLSKDM.cu:13-41) for domain 1:
    float32 vectors, float32 sums in the reference's order.  dobs: (ntr, nt)
    observed traces in trace order.  Returns (model, residual norms)."""
    f = np.float32
    eps = f(1e-6)
    dobs = np.ascontiguousarray(dobs, np.float32)
    mask = np.abs(dobs) > eps
    model = np.zeros(model_shape, np.float32)
    direction = np.zeros_like(model)
    grad_old = np.zeros_like(model)
    op = dict(dt=dt, dx=dx, dz=dz, aperture=aperture, da=da, nang=nang)

    def fwd(m):
        d = kdm(domain, False, S, src_tables, rec_tables, isrc, irec, cmp, cmp_id, np.zeros_like(dobs), m, **op)
        return trace_convolve(d, wavelet, False)

    def adj(d):
        d = np.ascontiguousarray(trace_convolve(d, wavelet, True))
        return kdm(domain, True, S, src_tables, rec_tables, isrc, irec, cmp, cmp_id, d, np.zeros_like(model), **op)

    residuo, it = [], 0
    while True:
        r = np.where(mask, dobs - fwd(model), f(0)).astype(np.float32)
        residuals = _seqsum(r * r)
        grad = adj(r)
        it += 1
        residuo.append(float(np.sqrt(residuals)))
        if it > max_it:
            break
        beta = _seqsum(grad * grad) / (_seqsum(grad_old * grad_old) + eps)
        direction = (grad + f(beta) * direction).astype(np.float32)
        alpha_num = _seqsum(grad * direction)
        grad_old = grad
        dd = fwd(direction)
        alpha = alpha_num / (_seqsum(dd * dd) + eps)
        if domain == 1 and smooth:
            m = model
            fw = np.concatenate([m[:, 1:], m[:, -2:-1]], axis=1)
            bw = np.concatenate([m[:, 1:2], m[:, :-1]], axis=1)
            direction = direction + (f(smooth) * ((fw - f(2) * m) + bw)) / (f(da) * f(da))
        model = (model + f(alpha) * direction).astype(np.float32)
    return model, residuo
