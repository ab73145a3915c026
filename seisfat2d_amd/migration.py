"""Kirchhoff depth migration on the GPU: the operators, gathers and migration
loops of SeisFAT2D's migration module (src/migration/), over the C ABI of
include/seisfat2d_kdm.h.

Layouts follow the reference: a shot's record is nrec x nt samples
(receiver-major, src/migration/KDM.cu:17-18,48-49), so a survey is an array of
shape (nsrc, nrec, nt); the image-domain model is (nx, nz) z-fastest
(IDKDM.cu:7); the angle-domain model is (ncmp, nang, nz), index
z + angle*nz + cmp*nang*nz (migration.cu:659, ADKDM.cu:7).  Device arrays are
torch CUDA tensors and every operator runs on torch's current stream.  There
is no CPU path: the operators raise if the HIP library is missing.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import capi
from .capi import KDM_ANGLE, KDM_IMAGE, KdmParams, check, lib
from .eikonal import ISO

__all__ = ["IMAGE", "ANGLE", "ricker", "gathers", "trace_list", "Kirchhoff"]

IMAGE, ANGLE = KDM_IMAGE, KDM_ANGLE
EPS = np.float32(1e-6)  # src/migration/migration.cuh:7
f32 = np.float32


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the migration operators need a GPU (no CPU path)")
    return torch


def _stream(torch):
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dptr(t):
    return C.c_void_p(t.data_ptr())


def ricker(dt, fmax):
    """Migration::set_wavelet (src/migration/migration.cu:54-73): the Ricker
    wavelet of maximum frequency fmax sampled at dt, as float32."""
    n = lib().seisfat_ricker(float(dt), float(fmax), None, 0)
    if n < 0:
        check(n)
    out = np.zeros(n, np.float32)
    if n:
        rc = lib().seisfat_ricker(float(dt), float(fmax), capi._ptr(out), n)
        if rc < 0:
            check(rc)
    return out


def gathers(xsrc, xrec, da, max_angle):
    """Migration::set_gathers (src/migration/migration.cu:75-99) in the
    reference's float32 arithmetic: returns (nang, dcmp, min_cmp, ncmp).
    dcmp is half the smaller of the source and receiver spacings (first two
    positions of each); a geometry with a single source or receiver uses the
    other spacing.  The largest CMP starts from -1e6 (the reference leaves it
    unset)."""
    xs = np.asarray(xsrc, f32).ravel()
    xr = np.asarray(xrec, f32).ravel()
    if len(xs) == 0 or len(xr) == 0:
        raise ValueError("need at least one source and one receiver")
    if not (f32(da) > 0):
        raise ValueError("angle spacing must be positive")
    nang = int(f32(max_angle) / f32(da)) + 1
    steps = [d[1] - d[0] for d in (xs, xr) if len(d) > 1]
    if not steps:
        raise ValueError("one source and one receiver: no CMP spacing")
    dcmp = f32(0.5) * f32(min(steps))
    if not (dcmp > 0):
        raise ValueError("sources and receivers must be listed with increasing x")
    cmp = f32(0.5) * (xs[:, None] + xr[None, :])
    min_cmp = min(f32(1e6), cmp.min())
    max_cmp = max(f32(-1e6), cmp.max())
    ncmp = int((max_cmp - min_cmp) / dcmp) + 1
    return nang, dcmp, f32(min_cmp), ncmp


def trace_list(xsrc, xrec, max_offset, min_cmp, dcmp):
    """The traces the migration loops visit, in their order
    (src/migration/KDM.cu:14-53, LSKDM.cu:60-100): every (source, receiver)
    pair with |sx - rx| < max_offset, sources outer.  Returns int32 source and
    receiver indices, float32 CMP positions and int32 CMP bins."""
    xs = np.asarray(xsrc, f32).ravel()
    xr = np.asarray(xrec, f32).ravel()
    keep = np.abs(xs[:, None] - xr[None, :]) < f32(max_offset)
    src, rec = np.nonzero(keep)
    cmp = f32(0.5) * (xs[src] + xr[rec])
    cid = ((cmp - f32(min_cmp)) / f32(dcmp)).astype(np.int32)
    return src.astype(np.int32), rec.astype(np.int32), cmp.astype(np.float32), cid


class Kirchhoff:
    """Kirchhoff depth migration of one acquisition on one model
    (Migration::set_parameters, src/migration/migration.cu:3-52, with the
    KDM / LSKDM loops): the traveltime tables of every source and receiver
    position are solved by the eikonal handle `eik` (or given), and the
    traces are the source-receiver pairs within max_offset.

    Parameters use the reference's names: nt, dt (time_samples,
    time_spacing), aperture (mig_aperture, degrees), max_offset
    (mig_max_offset), da and max_angle (mig_angle_spacing, mig_max_angle; the
    angle domain), fmax (max_frequency: the Ricker wavelet).
    """

    def __init__(self, eik, xsrc, zsrc, xrec, zrec, nt, dt, aperture, max_offset, fmax, domain=IMAGE,
                 da=1.0, max_angle=90.0, type=ISO, tables=None):
        torch = _torch()
        if domain not in (IMAGE, ANGLE):
            raise ValueError("domain must be IMAGE or ANGLE")
        self.eik, self.domain = eik, domain
        self.nt, self.dt = int(nt), f32(dt)
        self.xsrc, self.xrec = np.asarray(xsrc, f32).ravel(), np.asarray(xrec, f32).ravel()
        zsrc, zrec = np.asarray(zsrc, f32).ravel(), np.asarray(zrec, f32).ravel()
        self.nsrc, self.nrec = len(self.xsrc), len(self.xrec)
        self.nang, self.dcmp, self.min_cmp, self.ncmp = gathers(self.xsrc, self.xrec, da, max_angle)
        self.da = f32(da)
        self.src, self.rec, self.cmp, self.cmp_id = trace_list(self.xsrc, self.xrec, max_offset, self.min_cmp,
                                                               self.dcmp)
        self.ntr = len(self.src)
        self.wavelet = ricker(dt, fmax)
        if domain == IMAGE:
            self.model_shape = (eik.nx, eik.nz)
        else:
            self.model_shape = (self.ncmp, self.nang, eik.nz)
        p = KdmParams(domain, self.nt, float(dt), float(aperture), float(da), self.nang, self.ncmp)
        h = C.c_void_p()
        check(lib().seisfat_kdm_create(eik._h, C.byref(p), self.nsrc, self.nrec, C.byref(h)))
        self._k = h
        if tables is None:
            ts = eik.traveltime_tables(self.xsrc, zsrc, type=type)
            tr = eik.traveltime_tables(self.xrec, zrec, type=type)
        else:
            ts, tr = tables
        self.set_tables(ts, tr)
        torch.cuda.current_stream().synchronize()
        cid = self.cmp_id if domain == ANGLE else None
        check(lib().seisfat_kdm_set_traces(self._k, self.ntr, capi._ptr(self.src), capi._ptr(self.rec),
                                           capi._ptr(self.cmp), capi._ptr(cid)))

    def set_tables(self, ts, tr):
        """The source and receiver traveltime tables, (nsrc, nxx, nzz) and
        (nrec, nxx, nzz) padded grids (numpy or CUDA tensors): their weight
        quantities are evaluated on the device with the eikonal handle's
        current slowness (Migration::set_src/rec_travel_times,
        src/migration/migration.cu:116-171).  After a slowness upload to the
        handle the operators refuse to run until the tables are set again.
        Returns the float32 device grids used (the inputs themselves when they
        already are)."""
        torch = _torch()
        shape = (self.eik.nxx, self.eik.nzz)
        ts_d = torch.as_tensor(np.asarray(ts, f32) if not torch.is_tensor(ts) else ts, device="cuda")
        tr_d = torch.as_tensor(np.asarray(tr, f32) if not torch.is_tensor(tr) else tr, device="cuda")
        if tuple(ts_d.shape) != (self.nsrc, *shape) or tuple(tr_d.shape) != (self.nrec, *shape):
            raise ValueError(f"tables must be (nsrc, {shape[0]}, {shape[1]}) and (nrec, ...) padded grids")
        ts_d, tr_d = ts_d.to(torch.float32).contiguous(), tr_d.to(torch.float32).contiguous()
        check(lib().seisfat_kdm_set_tables(self._k, _dptr(ts_d), _dptr(tr_d), _stream(torch)))
        return ts_d, tr_d

    # -- the operators --------------------------------------------------------
    def _traces(self, torch, d):
        if not (torch.is_tensor(d) and d.is_cuda and d.dtype == torch.float32 and d.is_contiguous()
                and tuple(d.shape) == (self.ntr, self.nt)):
            raise ValueError(f"traces must be a contiguous float32 CUDA tensor of shape ({self.ntr}, {self.nt})")
        return d

    def _model(self, torch, m):
        if not (torch.is_tensor(m) and m.is_cuda and m.dtype == torch.float32 and m.is_contiguous()
                and tuple(m.shape) == self.model_shape):
            raise ValueError(f"model must be a contiguous float32 CUDA tensor of shape {self.model_shape}")
        return m

    def adjoint(self, traces, model=None):
        """model += L^T traces: perform_adjoint for every trace in order
        (src/migration/migration.cu:398-483,572-663).  traces: (ntr, nt)
        after the adjoint convolution.  Returns model (zeros when None)."""
        torch = _torch()
        if model is None:
            model = torch.zeros(self.model_shape, dtype=torch.float32, device="cuda")
        check(lib().seisfat_kdm_adjoint(self._k, _dptr(self._traces(torch, traces)), _dptr(self._model(torch, model)),
                                        _stream(torch)))
        return model

    def forward(self, model, traces=None):
        """traces += L model: perform_forward for every trace
        (src/migration/migration.cu:485-570,665-756).  Returns traces (zeros
        when None), before the forward convolution."""
        torch = _torch()
        if traces is None:
            traces = torch.zeros((self.ntr, self.nt), dtype=torch.float32, device="cuda")
        check(lib().seisfat_kdm_forward(self._k, _dptr(self._model(torch, model)), _dptr(self._traces(torch, traces)),
                                        _stream(torch)))
        return traces

    def convolve(self, traces, adjoint):
        """forward_convolution / adjoint_convolution (src/migration/migration.cu:197-261)
        of every trace with the Ricker wavelet; returns a new tensor."""
        torch = _torch()
        self._traces(torch, traces)
        out = torch.empty_like(traces)
        w = self.wavelet
        check(lib().seisfat_trace_convolve(int(bool(adjoint)), self.ntr, self.nt, capi._ptr(w), len(w),
                                           _dptr(traces), _dptr(out), _stream(torch)))
        return out

    def gather(self, seismic):
        """The migrated traces of a survey (nsrc, nrec, nt) (numpy or tensor),
        in trace order, as a (ntr, nt) CUDA tensor."""
        torch = _torch()
        s = torch.as_tensor(seismic, dtype=torch.float32, device="cuda")
        if tuple(s.shape) != (self.nsrc, self.nrec, self.nt):
            raise ValueError(f"seismic must have shape (nsrc, nrec, nt) = {(self.nsrc, self.nrec, self.nt)}")
        idx = torch.as_tensor(self.src.astype(np.int64) * self.nrec + self.rec, device="cuda")
        return s.reshape(-1, self.nt).index_select(0, idx).contiguous()

    # -- the migration loops --------------------------------------------------
    def migrate(self, seismic):
        """KDM::kirchhoff_depth_migration (src/migration/KDM.cu:3-55): the
        adjoint convolution of every trace, then the adjoint operator into a
        zero model.  Returns the model (IDKDM / ADKDM result)."""
        d = self.convolve(self.gather(seismic), adjoint=True)
        return self.adjoint(d)

    def _regularize(self, torch, direction, model, smooth):
        """ADLSKDM::regularization (src/migration/ADLSKDM.cu:13-41): the
        direction gains smooth * the second difference of the model along
        angle, the end bins mirrored."""
        m = model
        fw = torch.cat([m[:, 1:], m[:, -2:-1]], dim=1)
        bw = torch.cat([m[:, 1:2], m[:, :-1]], dim=1)
        lap = (fw - 2.0 * m) + bw
        direction += (float(f32(smooth)) * lap) / float(self.da * self.da)

    def lsm(self, seismic, max_it, smooth=0.0):
        """LSKDM::kirchhoff_depth_migration (src/migration/LSKDM.cu:3-226):
        max_it conjugate-gradient iterations on the masked residual
        (samples with |d_obs| <= 1e-6 do not count), Fletcher-Reeves
        direction, exact step length, the angle-domain smoothing; returns
        (model, residuals) with max_it + 1 residual norms.  Reductions are
        taken in float64 and rounded to float32, the step's scalars formed in
        float32, all on the device: no host round trip until the residual
        norms are returned."""
        torch = _torch()
        if self.domain == ANGLE and smooth and self.nang < 2:
            raise ValueError("angle-domain smoothing needs at least two angle bins")
        dobs = self.gather(seismic)
        mask = dobs.abs() > float(EPS)
        model = torch.zeros(self.model_shape, dtype=torch.float32, device="cuda")
        direction = torch.zeros_like(model)
        grad_old = torch.zeros_like(model)

        def sumsq(a, b):  # float64 sum of a*b, rounded to float32 (a 0-d device tensor)
            return torch.sum(a.double() * b.double()).float()

        residuo = []
        it = 0
        while True:
            dcal = self.convolve(self.forward(model), adjoint=False)
            r = torch.where(mask, dobs - dcal, torch.zeros_like(dcal))
            residuals = sumsq(r, r)
            grad = self.adjoint(self.convolve(r, adjoint=True))
            it += 1
            residuo.append(torch.sqrt(residuals))
            if it > max_it:
                break
            beta = sumsq(grad, grad) / (sumsq(grad_old, grad_old) + float(EPS))
            direction = grad + beta * direction
            alpha_num = sumsq(grad, direction)
            grad_old = grad
            dd = self.convolve(self.forward(direction), adjoint=False)
            alpha = alpha_num / (sumsq(dd, dd) + float(EPS))
            if self.domain == ANGLE and smooth:
                self._regularize(torch, direction, model, smooth)
            model = model + alpha * direction
        return model, [float(v) for v in torch.stack(residuo).cpu().numpy()]

    def close(self):
        if getattr(self, "_k", None):
            lib().seisfat_kdm_destroy(self._k)
            self._k = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
