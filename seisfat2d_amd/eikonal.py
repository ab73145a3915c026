"""Python front end of the C ABI (used by the tests, the benchmark and the
autograd wrapper; C++ callers use seisfat2d_amd/host/ instead).

Arrays follow the reference's z-fastest layout: a model of nz x nx samples is
a numpy array of shape (nx, nz) in C order (``a.ravel()[i + j*nz]`` is sample
z=i, x=j); a padded traveltime grid has shape (nx+6, nz+6).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import ANI, ISO, SCHED_PLAIN, SCHED_TILED, SCHED_TILED_DIAG, Shot, check, lib

__all__ = ["Eikonal", "Batch", "ISO", "ANI", "SCHED_PLAIN", "SCHED_TILED", "SCHED_TILED_DIAG", "NB"]

NB = 3


class Eikonal:
    """One model on the GPU (Modeling::set_parameters + set_properties)."""

    def __init__(self, nz: int, nx: int, dz: float, dx: float):
        self.nz, self.nx, self.dz, self.dx = int(nz), int(nx), float(dz), float(dx)
        self.nzz, self.nxx = self.nz + 2 * NB, self.nx + 2 * NB
        h = C.c_void_p()
        check(lib().seisfat_eik_create(self.nz, self.nx, self.dz, self.dx, C.byref(h)))
        self._h = h

    # -- model -------------------------------------------------------------
    def _model(self, a):
        a = np.ascontiguousarray(a, np.float32)
        if a.shape != (self.nx, self.nz):
            raise ValueError(f"model must have shape (nx, nz) = {(self.nx, self.nz)}, got {a.shape}")
        return a

    def set_slowness(self, s):
        """s: (nx, nz) numpy array, or a CUDA tensor of nx*nz float32 (device copy)."""
        if hasattr(s, "data_ptr") and getattr(s, "is_cuda", False):
            import torch
            t = s.detach().contiguous().to(torch.float32)
            if t.numel() != self.nx * self.nz:
                raise ValueError("slowness tensor has the wrong number of elements")
            stream = torch.cuda.current_stream(t.device).cuda_stream
            check(lib().seisfat_eik_set_slowness_device(self._h, C.c_void_p(t.data_ptr()), C.c_void_p(stream)))
        else:
            check(lib().seisfat_eik_set_slowness(self._h, self._model(s)))

    def set_slowness_padded(self, S):
        """S: padded (nx+6, nz+6) slowness uploaded verbatim, border included
        (Modeling::copy_slowness_to_device); whole-model solves read its
        border as given."""
        S = np.ascontiguousarray(S, np.float32)
        if S.shape != (self.nxx, self.nzz):
            raise ValueError(f"padded model must have shape (nx+6, nz+6) = {(self.nxx, self.nzz)}, got {S.shape}")
        check(lib().seisfat_eik_set_slowness_padded(self._h, S))

    def device_S_ptr(self):
        """Device address of the padded slowness (the reference's d_S)."""
        return lib().seisfat_eik_device_S(self._h)

    def set_stiffness(self, cs):
        """cs: six (nx, nz) arrays C11, C13, C15, C33, C35, C55."""
        if len(cs) != 6:
            raise ValueError("need six stiffness components")
        arrs = [self._model(c) for c in cs]
        ptrs = (C.c_void_p * 6)(*[a.ctypes.data_as(C.c_void_p) for a in arrs])
        check(lib().seisfat_eik_set_stiffness(self._h, ptrs))

    # -- solves -------------------------------------------------------------
    def solve(self, sx, sz, type=ISO, schedule=SCHED_TILED, rx=None, rz=None, return_T=True):
        """One shot on the whole model; returns (padded T (nxx, nzz) or None, picks or None)."""
        T = np.empty((self.nxx, self.nzz), np.float32) if return_T else None
        rx = None if rx is None else np.ascontiguousarray(rx, np.float32)
        rz = None if rz is None else np.ascontiguousarray(rz, np.float32)
        nrec = 0 if rx is None else len(rx)
        if nrec and (rz is None or len(rz) != nrec):
            raise ValueError("rx and rz must have the same length")
        picks = np.empty(nrec, np.float32) if nrec else None
        check(lib().seisfat_eik_solve(self._h, type, schedule, float(sx), float(sz), capi._ptr(T),
                                      capi._ptr(rx), capi._ptr(rz), nrec, capi._ptr(picks)))
        return T, picks

    def device_T_ptr(self):
        return lib().seisfat_eik_device_T(self._h)

    @property
    def fast_division(self):
        """True when the batched tiled sweeps use the two-instruction
        known-divisor division (results are bit-identical either way)."""
        rc = lib().seisfat_eik_fast_division(self._h)
        if rc < 0:
            check(rc)
        return rc == 1

    def batch(self, shots, rx=None, rz=None, type=ISO, schedule=SCHED_TILED, opts=None):
        """A batch of shots.  opts: schedule options of the tiled schedule
        (pairs, wpcu, skip, priority; bit-identical results), a dict or
        "tuned" for the options seisfat2d_amd.tune found for this workload
        (measured now and cached when the cache has none)."""
        if isinstance(opts, str):
            if opts != "tuned":
                raise ValueError('opts must be a dict, None or "tuned"')
            from . import tune
            opts = tune.tune(self, shots, rx, rz, type=type)["opts"]
        return Batch(self, shots, rx, rz, type, schedule, opts)

    def traveltime_tables(self, x, z, type=ISO, schedule=SCHED_TILED, chunk=64):
        """Padded traveltime grids (n, nx+6, nz+6) of whole-model solves from
        each position (x[k], z[k]): the tables of Migration::set_src_travel_times
        and set_rec_travel_times (src/migration/migration.cu:116-171), solved
        `chunk` positions per batch."""
        x = np.asarray(x, np.float64).ravel()
        z = np.asarray(z, np.float64).ravel()
        if x.shape != z.shape:
            raise ValueError("x and z must have the same length")
        out = np.empty((len(x), self.nxx, self.nzz), np.float32)
        for k0 in range(0, len(x), chunk):
            shots = [(float(x[k]), float(z[k]), 0, self.nx, 0, 0) for k in range(k0, min(k0 + chunk, len(x)))]
            b = Batch(self, shots, None, None, type, schedule)
            try:
                b.run()
                b.status()
                for q in range(len(shots)):
                    out[k0 + q] = b.T(q)
            finally:
                b.close()
        return out

    def close(self):
        if getattr(self, "_h", None):
            lib().seisfat_eik_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many shots solved together (tiled schedule: one HIP graph replay per run).

    shots: sequence of (sx, sz, x0, width, rec_offset, nrec) tuples.
    """

    def __init__(self, eik: Eikonal, shots, rx, rz, type=ISO, schedule=SCHED_TILED, opts=None):
        self.eik = eik
        n = len(shots)
        arr = (Shot * n)(*[Shot(float(s[0]), float(s[1]), int(s[2]), int(s[3]), int(s[4]), int(s[5]))
                           for s in shots])
        self.shots = [tuple(s) for s in shots]
        self.rx = np.ascontiguousarray(rx if rx is not None else np.zeros(0), np.float32)
        self.rz = np.ascontiguousarray(rz if rz is not None else np.zeros(0), np.float32)
        self.nrec = len(self.rx)
        b = C.c_void_p()
        so = capi.SchedOpts.of(opts)
        check(lib().seisfat_batch_create_opts(eik._h, type, schedule, n, arr, self.nrec, capi._ptr(self.rx),
                                              capi._ptr(self.rz), C.byref(so), C.byref(b)))
        self._b = b
        self.records = 0  # recorded runs so far; the adjoint linearises the latest one

    def run(self, stream=None):
        check(lib().seisfat_batch_run(self._b, None if stream is None else C.c_void_p(stream)))

    def run_converged(self, max_passes=100, chunk=8, stream=None):
        """Passes until one changes nothing (at most max_passes); returns the
        number of passes (synchronous)."""
        n = C.c_int()
        check(lib().seisfat_batch_run_converged(self._b, int(max_passes), int(chunk), C.byref(n),
                                                None if stream is None else C.c_void_p(stream)))
        return n.value

    def picks(self):
        out = np.empty(self.nrec, np.float32)
        check(lib().seisfat_batch_picks(self._b, capi._ptr(out)))
        return out

    def device_picks_ptr(self):
        return lib().seisfat_batch_device_picks(self._b)

    def T(self, shot):
        nzz, nxx = C.c_int(), C.c_int()
        lib().seisfat_batch_device_T(self._b, shot, C.byref(nzz), C.byref(nxx))
        out = np.empty((nxx.value, nzz.value), np.float32)
        check(lib().seisfat_batch_copy_T(self._b, shot, capi._ptr(out)))
        return out

    def device_T_ptr(self, shot):
        nzz, nxx = C.c_int(), C.c_int()
        p = lib().seisfat_batch_device_T(self._b, shot, C.byref(nzz), C.byref(nxx))
        return p, nzz.value, nxx.value

    @property
    def num_launches(self):
        return lib().seisfat_batch_num_launches(self._b)

    @property
    def num_sweep_launches(self):
        return lib().seisfat_batch_num_sweep_launches(self._b)

    def sweep_ms(self):
        """Device time of the sweep phase of the last completed run (HIP events)."""
        ms = C.c_float()
        check(lib().seisfat_batch_sweep_ms(self._b, C.byref(ms)))
        return ms.value

    @property
    def opts(self):
        """The schedule options the batch runs with (pairs reads 0 where the
        fused pairs are unavailable)."""
        so = capi.SchedOpts()
        check(lib().seisfat_batch_get_opts(self._b, C.byref(so)))
        return so.as_dict()

    def tile_counts(self):
        """(total, skipped) tiles of the last completed run of the persistent
        tiled schedule; skipped tiles had only "not reached" (1e6) inputs and
        provably leave their nodes at 1e6, so their steps did not run."""
        total, skipped = C.c_longlong(), C.c_longlong()
        check(lib().seisfat_batch_tile_counts(self._b, C.byref(total), C.byref(skipped)))
        return total.value, skipped.value

    def status(self):
        """Synchronise and raise if the device flagged a failure during the runs."""
        check(lib().seisfat_batch_status(self._b))

    # -- adjoint (isotropic batches) ---------------------------------------
    def run_record(self, stream=None):
        """Forward solve that keeps what the adjoint needs (same picks as run)."""
        check(lib().seisfat_batch_run_record(self._b, None if stream is None else C.c_void_p(stream)))
        self.records += 1

    def vjp_device(self, w_ptr, grad_ptr, stream=None):
        """grad (nz*nx float64, device) = sum_r w[r] dpick_r/ds for the recorded run."""
        check(lib().seisfat_batch_vjp_device(self._b, C.c_void_p(w_ptr), C.c_void_p(grad_ptr),
                                             None if stream is None else C.c_void_p(stream)))

    def misfit_gradient(self, obs):
        """(J, dJ/ds as (nx, nz) float64, picks) for J = 1/2 sum (pick - obs)^2."""
        obs = np.ascontiguousarray(obs, np.float32)
        if obs.size != self.nrec:
            raise ValueError("obs must hold one value per receiver")
        J = C.c_double()
        g = np.empty((self.eik.nx, self.eik.nz), np.float64)
        pk = np.empty(self.nrec, np.float32)
        check(lib().seisfat_batch_misfit_gradient(self._b, capi._ptr(obs), C.byref(J), capi._ptr(g),
                                                  capi._ptr(pk)))
        self.records += 1  # it records a run of its own
        return J.value, g, pk

    def close(self):
        if getattr(self, "_b", None):
            lib().seisfat_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
