"""Python mirror of the engine interface.

`Engine` wraps the extension ABI (include/spme_ext.h): device-resident
inputs, FP64 or FP32 mesh, batches of right-hand sides, full or partial M.
`DropIn` wraps the reference's own surface (include/spme.h:
create_spme_engine / compute_spme / destroy_spme_engine, old-version/spme.h:
85-97) with host numpy arrays, the way the reference's driver calls it
(old-version/main.c:215-250).

torch is used only to hold device memory and to name the current stream.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib as _L
from .lib import (SPME_FP32, SPME_FP32_SUM64, SPME_FP64, SPME_FP64_PLAIN, SPME_FULL, SPME_REAL,
                  SPME_RECIP, SPME_SELF, SpmeError, SpmeInfo, SpmeParams, SpmeTiles,
                  check, load)

__all__ = ["Engine", "Group", "DropIn", "cutoff2xi", "SPME_FP32", "SPME_FP32_SUM64", "SPME_FP64",
           "SPME_FP64_PLAIN", "SPME_FULL", "SPME_REAL", "SPME_RECIP", "SPME_SELF",
           "SpmeError"]


def cutoff2xi(rmax: float, ewald_eps: float, box_size: float) -> float:
    """old-version/spme.h:85-86 (declared only in the reference)."""
    xi = ctypes.c_double()
    load().cutoff2xi(float(rmax), float(ewald_eps), float(box_size),
                     ctypes.byref(xi))
    return xi.value


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise SpmeError("no GPU visible: the SPME engine runs only on the GPU")
    return torch


STREAM_MODES = {"read": 0, "write": 1, "copy": 2, "scale": 3}
# HBM bytes one launch moves, per buffer byte
STREAM_TRAFFIC = {"read": 1, "write": 1, "copy": 2, "scale": 2}


def stream_probe(mode, a, b, value=1.0, nontemporal=False, blocks=0, reps=5):
    """Mean ms of one launch of a hand-written streaming kernel (no
    counterpart in the reference): 'read' a (b[k] gets block k's sum),
    'write' b = value, 'copy' b = a, 'scale' b *= value.  a and b: device
    tensors of the same byte size (a multiple of 16); runs on torch's current
    stream, synchronously."""
    torch = _torch()
    nbytes = b.numel() * b.element_size()
    if a is not None and a.numel() * a.element_size() != nbytes:
        raise SpmeError("stream_probe: a and b differ in size")
    ms = ctypes.c_float()
    stream = torch.cuda.current_stream().cuda_stream
    check(load().spme_stream_probe(STREAM_MODES[mode], int(bool(nontemporal)),
                                   None if a is None else a.data_ptr(), b.data_ptr(),
                                   nbytes, float(value), int(blocks), int(reps), stream,
                                   ctypes.byref(ms)), "spme_stream_probe")
    return ms.value


class Engine:
    """One SPME engine (positions, P data, mesh) on the current GPU."""

    def __init__(self, xi, dim, porder, np_, box_size, rc=0.0,
                 precision=SPME_FP64, max_rhs=1, radii=False):
        self._lib = load()
        prm = SpmeParams(float(xi), int(dim), int(porder), int(np_),
                         float(box_size), float(rc), int(precision),
                         int(max_rhs), int(bool(radii)))
        h = ctypes.c_void_p()
        check(self._lib.spme_engine_create(ctypes.byref(prm), ctypes.byref(h)),
              "spme_engine_create")
        self._h = h
        self.N = int(np_)
        self.precision = int(precision)
        self.max_rhs = int(max_rhs)

    @property
    def dtype(self):
        """torch dtype of the input vectors (and of the mesh path)."""
        torch = _torch()
        return torch.float32 if self.precision in (SPME_FP32, SPME_FP32_SUM64) else torch.float64

    @property
    def out_dtype(self):
        """torch dtype of the output vectors: double for SPME_FP32_SUM64,
        whose final real + reciprocal sum is formed in FP64."""
        torch = _torch()
        return torch.float32 if self.precision == SPME_FP32 else torch.float64

    def close(self):
        if getattr(self, "_h", None):
            self._lib.spme_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self) -> SpmeInfo:
        inf = SpmeInfo()
        check(self._lib.spme_engine_info(self._h, ctypes.byref(inf)), "info")
        return inf

    def set_positions(self, pos, stream=None):
        """pos: torch CUDA tensor [N, 3] float64, or a host numpy array."""
        if isinstance(pos, np.ndarray):
            a = np.ascontiguousarray(pos, dtype=np.float64)
            check(self._lib.spme_engine_set_positions(
                self._h, a.ctypes.data_as(ctypes.c_void_p), 0, None),
                "set_positions")
            return
        torch = _torch()
        assert pos.dtype == torch.float64 and pos.is_cuda and pos.is_contiguous()
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(self._lib.spme_engine_set_positions(
            self._h, ctypes.c_void_p(pos.data_ptr()), 1, ctypes.c_void_p(s)),
            "set_positions")

    def set_radii(self, radii, stream=None):
        """Per-particle radii [N] (float64 numpy array or CUDA tensor) of an
        engine created with radii=True; positions must be set afterwards."""
        if isinstance(radii, np.ndarray):
            a = np.ascontiguousarray(radii, dtype=np.float64)
            check(self._lib.spme_engine_set_radii(
                self._h, a.ctypes.data_as(ctypes.c_void_p), 0, None), "set_radii")
            return
        torch = _torch()
        assert radii.dtype == torch.float64 and radii.is_cuda and radii.is_contiguous()
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(self._lib.spme_engine_set_radii(
            self._h, ctypes.c_void_p(radii.data_ptr()), 1, ctypes.c_void_p(s)), "set_radii")

    def apply(self, f, u=None, alpha=1.0, beta=0.0, parts=SPME_RECIP,
              stream=None):
        """u <- beta*u + alpha*M f for f [nrhs, 3N] (torch CUDA tensor of the
        engine's precision, last dim contiguous).  Returns u."""
        torch = _torch()
        if f.dim() == 1:
            f = f.view(1, -1)
        if u is not None and u.dim() == 1:
            u = u.view(1, -1)
        if u is None:
            u = torch.zeros(f.shape, dtype=self.out_dtype, device=f.device)
            beta = 0.0
        nrhs = f.shape[0]
        assert f.dtype == self.dtype and u.dtype == self.out_dtype
        assert f.stride(1) == 1 and u.stride(1) == 1
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(self._lib.spme_engine_apply(
            self._h, int(nrhs), float(alpha), ctypes.c_void_p(f.data_ptr()),
            int(f.stride(0)), float(beta), ctypes.c_void_p(u.data_ptr()),
            int(u.stride(0)), int(parts), ctypes.c_void_p(s)), "apply")
        return u

    def set_tiles(self, tiles=None):
        """tiles: dict of spme_tiles_t fields (missing = 0 = default), or None."""
        t = _tiles(tiles)
        check(self._lib.spme_engine_set_tiles(
            self._h, ctypes.byref(t) if t is not None else None), "set_tiles")

    def tiles(self, nrhs=1):
        t = SpmeTiles()
        check(self._lib.spme_engine_get_tiles(self._h, int(nrhs), ctypes.byref(t)),
              "get_tiles")
        return t.as_dict()

    def autotune(self, nrhs=1, reps=5, stream=None):
        """Time the candidate tile shapes, keep and return the fastest."""
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        t = SpmeTiles()
        check(self._lib.spme_engine_autotune(self._h, int(nrhs), int(reps),
                                             ctypes.byref(t), ctypes.c_void_p(s)),
              "autotune")
        return t.as_dict()

    def set_timing(self, enable=True):
        check(self._lib.spme_engine_set_timing(self._h, int(bool(enable))),
              "set_timing")

    def timing(self):
        """{stage: (total ms, launches)} since timing was enabled."""
        n = len(_L.STAGES)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int * n)()
        check(self._lib.spme_engine_get_timing(self._h, ms, cnt), "get_timing")
        return {name: (ms[i], cnt[i]) for i, name in enumerate(_L.STAGES)}

    def debug_stage(self, stage, f):
        """Mesh after spreading (stage 1) or after FFT, influence and inverse
        FFT (stage 2), as numpy [3*nrhs, K, K, RS]."""
        inf = self.info()
        f = f.view(1, -1) if f.dim() == 1 else f
        nrhs = f.shape[0]
        dt = np.float32 if self.precision in (SPME_FP32, SPME_FP32_SUM64) else np.float64
        out = np.empty(3 * nrhs * inf.comp_stride, dtype=dt)
        check(self._lib.spme_engine_debug_stage(
            self._h, int(stage), int(nrhs), ctypes.c_void_p(f.data_ptr()),
            int(f.stride(0)), out.ctypes.data_as(ctypes.c_void_p)), "debug_stage")
        return out.reshape(3 * nrhs, inf.K, inf.K, inf.RS)


def _tiles(d):
    if d is None:
        return None
    t = SpmeTiles()
    for k, v in d.items():
        setattr(t, k, int(v))
    return t


class Group:
    """Independent systems applied together: one launch per stage over all
    of them, optionally replayed as one captured hipGraph.  The engines stay
    owned by the caller and must outlive the group."""

    def __init__(self, engines, nrhs):
        self._lib = load()
        self.engines = list(engines)
        self.nrhs = int(nrhs)
        arr = (ctypes.c_void_p * len(self.engines))(*[e._h.value for e in self.engines])
        h = ctypes.c_void_p()
        check(self._lib.spme_group_create(len(self.engines), arr, self.nrhs,
                                          ctypes.byref(h)), "spme_group_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.spme_group_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tiles(self, tiles=None):
        t = _tiles(tiles)
        check(self._lib.spme_group_set_tiles(
            self._h, ctypes.byref(t) if t is not None else None), "group set_tiles")

    def tiles(self):
        t = SpmeTiles()
        check(self._lib.spme_group_get_tiles(self._h, ctypes.byref(t)), "group get_tiles")
        return t.as_dict()

    def autotune(self, parts=SPME_RECIP, reps=5, stream=None):
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        t = SpmeTiles()
        check(self._lib.spme_group_autotune(self._h, int(parts), int(reps),
                                            ctypes.byref(t), ctypes.c_void_p(s)),
              "group autotune")
        return t.as_dict()

    def apply(self, fs, us, alpha=1.0, beta=0.0, parts=SPME_RECIP, graph=True,
              stream=None):
        """us[s] <- beta*us[s] + alpha*M_s fs[s]; fs[s], us[s]: torch CUDA
        tensors [nrhs, 3 N_s] (last dim contiguous)."""
        torch = _torch()
        n = len(self.engines)
        assert len(fs) == n and len(us) == n
        vp, lg = ctypes.c_void_p, ctypes.c_long
        fin = (vp * n)(*[f.data_ptr() for f in fs])
        fout = (vp * n)(*[u.data_ptr() for u in us])
        ldi = (lg * n)(*[f.stride(0) for f in fs])
        ldo = (lg * n)(*[u.stride(0) for u in us])
        for f, u, e in zip(fs, us, self.engines):
            assert f.dtype == e.dtype and u.dtype == e.dtype
            assert f.stride(1) == 1 and u.stride(1) == 1 and f.shape[0] == self.nrhs
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(self._lib.spme_group_apply(self._h, float(alpha), fin, ldi, float(beta), fout,
                                         ldo, int(parts), int(bool(graph)), vp(s)),
              "group apply")
        return us


class DropIn:
    """The reference's C surface, called with host arrays."""

    def __init__(self, xi, dim, porder, np_, box_size, nthreads=1):
        self._lib = load()
        h = ctypes.c_void_p()
        self._lib.create_spme_engine(float(xi), int(dim), int(porder),
                                     int(np_), float(box_size), int(nthreads),
                                     ctypes.byref(h))
        self._h = h
        self.N = int(np_)

    def compute(self, pos, nrhs, alpha, vec_in, ldin, beta, vec_out, ldout):
        p = None
        if pos is not None:
            pos = np.ascontiguousarray(pos, dtype=np.float64)
            p = pos.ctypes.data_as(ctypes.c_void_p)
        assert vec_in.dtype == np.float64 and vec_out.dtype == np.float64
        assert vec_in.flags.c_contiguous and vec_out.flags.c_contiguous
        self._lib.compute_spme(self._h, p, int(nrhs), float(alpha),
                               vec_in.ctypes.data_as(ctypes.c_void_p), int(ldin),
                               float(beta),
                               vec_out.ctypes.data_as(ctypes.c_void_p),
                               int(ldout))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.destroy_spme_engine(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
