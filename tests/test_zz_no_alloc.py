"""The allocation-free per-shot solve (seisfat_eik_solve_noalloc) through the
C ABI: traveltimes and picks equal the allocating seisfat_eik_solve's bit for
bit, on two calls in a row, the second on a smaller model with fewer
receivers (the fixed device state keeps the first call's contents), and a
model beyond the state's fixed size is refused with an error, not run."""
import ctypes as C

import numpy as np
import pytest

from seisfat2d_amd import Eikonal, capi, models

pytestmark = pytest.mark.gpu


def _noalloc(e, sx, sz, rx, rz):
    import torch
    L = capi.lib()
    L.seisfat_eik_solve_noalloc.restype = C.c_int
    L.seisfat_eik_solve_noalloc.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    T = torch.full((e.nxx * e.nzz,), -1.0, dtype=torch.float32, device="cuda")
    d_rx = torch.from_numpy(np.ascontiguousarray(rx, np.float32)).cuda()
    d_rz = torch.from_numpy(np.ascontiguousarray(rz, np.float32)).cuda()
    picks = torch.full((len(rx),), -1.0, dtype=torch.float32, device="cuda")
    err = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    rc = L.seisfat_eik_solve_noalloc(e._h, float(sx), float(sz), T.data_ptr(), d_rx.data_ptr(), d_rz.data_ptr(),
                                     len(rx), picks.data_ptr(), err.data_ptr(), C.c_void_p(stream))
    capi.check(rc)
    torch.cuda.synchronize()
    return T.cpu().numpy().reshape(e.nxx, e.nzz), picks.cpu().numpy(), int(err.item())


def test_noalloc_matches_allocating_solve_twice_then_smaller():
    rng = np.random.default_rng(41)
    w = models.c1()
    e1 = Eikonal(w.nz, w.nx, w.dz, w.dx)
    e1.set_slowness(w.s)
    rx1 = rng.uniform(0, (w.nx - 1) * w.dx, 300).astype(np.float32)
    rz1 = rng.uniform(0, (w.nz - 1) * w.dz, 300).astype(np.float32)
    T_ref, p_ref = e1.solve(5000.0, 0.0, rx=rx1, rz=rz1)
    T, p, err = _noalloc(e1, 5000.0, 0.0, rx1, rz1)
    assert err == 0
    assert np.array_equal(T, T_ref)
    assert np.array_equal(p, p_ref)

    m = models.rough(64, 96, 3.0, 10.0, seed=11)  # smaller model, fewer receivers
    e2 = Eikonal(m.nz, m.nx, m.dz, m.dx)
    e2.set_slowness(m.s)
    T_ref, p_ref = e2.solve(m.sx, m.sz, rx=m.rx[:9], rz=m.rz[:9])
    T, p, err = _noalloc(e2, m.sx, m.sz, m.rx[:9], m.rz[:9])
    assert err == 0
    assert np.array_equal(T, T_ref)
    assert np.array_equal(p, p_ref)

    # and the first geometry again after the smaller one
    T_ref, p_ref = e1.solve(1234.5, 600.0, rx=rx1, rz=rz1)
    T, p, err = _noalloc(e1, 1234.5, 600.0, rx1, rz1)
    assert err == 0 and np.array_equal(T, T_ref) and np.array_equal(p, p_ref)
    e1.close()
    e2.close()


def test_noalloc_refuses_what_it_cannot_hold():
    # 63 strips x 20032 columns of exchange words: beyond the fixed device state
    big = Eikonal(4001, 20001, 5.0, 5.0)
    L = capi.lib()
    L.seisfat_eik_solve_noalloc.restype = C.c_int
    L.seisfat_eik_solve_noalloc.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    rc = L.seisfat_eik_solve_noalloc(big._h, 10.0, 0.0, C.c_void_p(1), None, None, 0, None, None, None)
    assert rc == capi.EINVAL
    assert b"too large" in L.seisfat_last_error()
    big.close()
    small = Eikonal(20, 30, 10.0, 10.0)
    small.set_slowness(np.full((30, 20), 5e-4, np.float32))
    rc = L.seisfat_eik_solve_noalloc(small._h, -1.0, 0.0, C.c_void_p(1), None, None, 0, None, None, None)
    assert rc == capi.EINVAL  # source outside the model: refused before any launch
    small.close()
