"""lcurve_re_amd -- MI355X light-curve engine (Python mirror of the C ABI).

Mirrors the reference's interface for the light-curve path:

* ``Model`` -- Lcurve::Model (lcurve.h:317-421): built from the JSON
  configuration (``Model.from_json``), with ``nvary``/``set_param``/
  ``get_param``/``get_range``/``get_dstep``/``get_limit``/``is_not_legal``/
  ``get_name`` (model.cpp:363-1660);
* ``make_data`` -- Lcurve::Data (lcurve.h:427-472) as an array of lc_datum;
* ``Engine.light_curve_comp`` / ``light_curve_comp_fast`` -- the evaluation
  calls (lcurve.h:474-483, 669-674), raising ``LcurveError`` where the
  reference throws Lcurve_Error/Roche_Error;
* ``Engine.light_curve_comp_batch`` -- many parameter vectors in one launch
  sequence, with a status per instance.

All computation happens in the native library (``_lib/liblcurve_amd.so``);
this module only marshals arguments.  Nothing here falls back to a CPU path:
without the library or a GPU the calls raise.
"""
import ctypes as C
import json

import numpy as np

from . import _abi
from ._abi import (LC_OK, LC_E_DEVICE, LC_PREC_FP64, LC_PREC_MIXED, LC_LDC_CLARET, LC_LDC_POLY,
                   lc_datum, lc_eval_opts, lc_ldc, lc_model, lc_scalars)

__all__ = ["Model", "ModelBatch", "Engine", "DeviceBuffer", "LcurveError", "make_data", "read_data", "write_data", "lc_model", "lc_datum", "lc_ldc",
           "LC_PREC_MIXED", "LC_PREC_FP64", "LC_LDC_POLY", "LC_LDC_CLARET"]


class LcurveError(RuntimeError):
    """Lcurve::Lcurve_Error / Roche::Roche_Error raised by the engine."""

    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


def _lib():
    return _abi.load()


_DP = C.POINTER(C.c_double)


def _dp(a):
    return a.ctypes.data_as(_DP)


_hip_rt = None


def _hip():
    """The HIP runtime the library is linked against (loaded with it), for
    the few device-memory calls DeviceBuffer needs."""
    global _hip_rt
    if _hip_rt is None:
        _lib()
        rt = C.CDLL("libamdhip64.so.7")
        rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        rt.hipFree.argtypes = [C.c_void_p]
        rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        rt.hipSetDevice.argtypes = [C.c_int]
        for f in (rt.hipMalloc, rt.hipFree, rt.hipMemcpy, rt.hipSetDevice):
            f.restype = C.c_int
        _hip_rt = rt
    return _hip_rt


class DeviceBuffer:
    """A float64 array of `shape` in GPU memory (hipMalloc on `device`), for
    lc_light_curve_comp_batch_dev: `ptr` is its device address; `fill`
    writes a host array or scalar into it, `to_host` copies it back."""

    def __init__(self, shape, device=0):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.nbytes = int(np.prod(self.shape)) * 8
        self.device = int(device)
        rt = _hip()
        p = C.c_void_p()
        if rt.hipSetDevice(self.device) or rt.hipMalloc(C.byref(p), max(self.nbytes, 8)):
            raise LcurveError(LC_E_DEVICE, "DeviceBuffer: hipMalloc of %d bytes failed" % self.nbytes)
        self.ptr = p.value

    def fill(self, values):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(values, np.float64), self.shape))
        if _hip().hipMemcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), self.nbytes, 1):
            raise LcurveError(LC_E_DEVICE, "DeviceBuffer: copy to the device failed")
        return self

    def to_host(self):
        a = np.empty(self.shape, np.float64)
        if _hip().hipMemcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes, 2):
            raise LcurveError(LC_E_DEVICE, "DeviceBuffer: copy to the host failed")
        return a

    def close(self):
        if self.ptr:
            _hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Model:
    """Lcurve::Model over the C struct lc_model."""

    def __init__(self, pod=None):
        self.pod = pod if pod is not None else lc_model()

    @classmethod
    def from_json(cls, config):
        """Model(const json&) (model.cpp:10-360); config: str, bytes or dict."""
        if isinstance(config, dict):
            config = json.dumps(config)
        if isinstance(config, str):
            config = config.encode()
        pod = lc_model()
        st = _lib().lc_model_from_json(config, len(config), C.byref(pod))
        if st != LC_OK:
            raise LcurveError(st, _lib().lc_thread_error().decode())
        return cls(pod)

    def copy(self):
        m = lc_model()
        C.memmove(C.byref(m), C.byref(self.pod), C.sizeof(lc_model))
        return Model(m)

    def __getattr__(self, name):
        pod = self.__dict__.get("pod")
        if pod is not None and hasattr(pod, name):
            v = getattr(pod, name)
            return v.value if isinstance(v, _abi.lc_pparam) else v
        raise AttributeError(name)

    def set(self, name, value):
        """Sets the value of a physical parameter or a computational key."""
        v = getattr(self.pod, name)
        if isinstance(v, _abi.lc_pparam):
            v.value = float(value)
        else:
            setattr(self.pod, name, value)
        return self

    def nvary(self):
        return _lib().lc_model_nvary(C.byref(self.pod))

    def set_param(self, vpar):
        v = np.ascontiguousarray(vpar, np.float64)
        st = _lib().lc_model_set_param(C.byref(self.pod), _dp(v), len(v))
        if st != LC_OK:
            raise LcurveError(st, _lib().lc_thread_error().decode())

    def _get(self, fn):
        out = np.zeros(max(self.nvary(), 1), np.float64)
        st = fn(C.byref(self.pod), _dp(out), len(out))
        if st != LC_OK:
            raise LcurveError(st, _lib().lc_thread_error().decode())
        return out[: self.nvary()]

    def get_param(self):
        return self._get(_lib().lc_model_get_param)

    def get_range(self):
        return self._get(_lib().lc_model_get_range)

    def get_dstep(self):
        return self._get(_lib().lc_model_get_dstep)

    def get_limit(self):
        n = max(self.nvary(), 1)
        lo, hi = np.zeros(n), np.zeros(n)
        _lib().lc_model_get_limit(C.byref(self.pod), _dp(lo), _dp(hi), n)
        return list(zip(lo[: self.nvary()], hi[: self.nvary()]))

    def is_not_legal(self, vpar):
        v = np.ascontiguousarray(vpar, np.float64)
        return bool(_lib().lc_model_is_not_legal(C.byref(self.pod), _dp(v), len(v)))

    def is_not_legal_batch(self, P):
        """is_not_legal of each row of P: a bool array."""
        P = np.ascontiguousarray(np.atleast_2d(P), np.float64)
        out = np.zeros(P.shape[0], np.int32)
        if P.shape[0]:
            st = _lib().lc_model_is_not_legal_batch(C.byref(self.pod), _dp(P), P.shape[0], P.shape[1],
                                                    out.ctypes.data_as(C.POINTER(C.c_int)))
            if st != LC_OK:
                raise LcurveError(st, "lc_model_is_not_legal_batch")
        return out.astype(bool)

    def get_name(self, i):
        buf = C.create_string_buffer(64)
        _lib().lc_model_get_name(C.byref(self.pod), int(i), buf, 64)
        return buf.value.decode()

    def grid_size(self, star, fine=True):
        n = C.c_size_t()
        st = _lib().lc_star_grid_size(C.byref(self.pod), int(star), int(fine), C.byref(n))
        if st != LC_OK:
            raise LcurveError(st, _lib().lc_thread_error().decode())
        return n.value


def make_data(time, expose=0.0, flux=0.0, ferr=1.0, weight=1.0, ndiv=1):
    """Lcurve::Data as a ctypes array of lc_datum (broadcasting scalars)."""
    t = np.asarray(time, np.float64)
    n = t.size
    cols = [np.broadcast_to(np.asarray(x, np.float64), (n,)) for x in (expose, flux, ferr, weight)]
    nd = np.broadcast_to(np.asarray(ndiv, np.int32), (n,))
    arr = (lc_datum * n)()
    buf = np.frombuffer(arr, dtype=np.dtype([("time", "f8"), ("expose", "f8"), ("flux", "f8"), ("ferr", "f8"),
                                             ("weight", "f8"), ("ndiv", "i4"), ("pad", "i4")]))
    buf["time"], buf["expose"], buf["flux"], buf["ferr"], buf["weight"] = t, *cols
    buf["ndiv"] = nd
    buf["pad"] = 0
    return arr


def data_view(arr):
    """numpy structured view of an lc_datum array."""
    return np.frombuffer(arr, dtype=np.dtype([("time", "f8"), ("expose", "f8"), ("flux", "f8"), ("ferr", "f8"),
                                              ("weight", "f8"), ("ndiv", "i4"), ("pad", "i4")]))


class ModelBatch:
    """Models packed once into the C layout (lc_model[m]) so that repeated
    batch calls on the same models skip the per-call packing."""

    def __init__(self, models):
        self.m = len(models)
        self.pods = (lc_model * max(self.m, 1))()
        for i, md in enumerate(models):
            C.memmove(C.byref(self.pods[i]), C.byref(md.pod), C.sizeof(lc_model))

    def __len__(self):
        return self.m


def _packed(models):
    if isinstance(models, ModelBatch):
        return models.pods, models.m
    b = ModelBatch(models)
    return b.pods, b.m


class Engine:
    """One lc_ctx: device buffers, stream and profile counters."""

    def __init__(self, device=0):
        self._ctx = C.c_void_p()
        st = _lib().lc_ctx_create(int(device), C.byref(self._ctx))
        if st != LC_OK:
            raise LcurveError(st, "lc_ctx_create failed: " + _lib().lc_thread_error().decode())

    def close(self):
        if self._ctx:
            _lib().lc_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self, st):
        raise LcurveError(st, _lib().lc_last_error(self._ctx).decode())

    @staticmethod
    def _opts(scale, rdata, info, diagnostics, max_model_points, precision):
        return lc_eval_opts(int(scale), int(rdata), int(info), int(diagnostics), int(max_model_points),
                            int(precision))

    def light_curve_comp(self, model, data, scale=False, rdata=True, info=False, sfac=None, diagnostics=True,
                         precision=LC_PREC_MIXED, max_model_points=0):
        """Returns (calc, scalars); raises LcurveError like the reference."""
        n = len(data)
        calc = np.zeros(n, np.float64)
        io = lc_scalars()
        for k in range(5):
            io.sfac[k] = 1.0 if sfac is None else (sfac[k] if hasattr(sfac, "__len__") else sfac)
        o = self._opts(scale, rdata, info, diagnostics, max_model_points, precision)
        st = _lib().lc_light_curve_comp(self._ctx, C.byref(model.pod), data, n, C.byref(o), _dp(calc),
                                        C.byref(io))
        if st != LC_OK:
            self._err(st)
        return calc, io

    def light_curve_comp_fast(self, model, data, max_model_points=500, **kw):
        return self.light_curve_comp(model, data, max_model_points=max_model_points, **kw)

    def light_curve_comp_batch(self, models, data, scale=False, rdata=True, diagnostics=False, sfac=1.0,
                               precision=LC_PREC_MIXED, max_model_points=0, out=None):
        """Evaluates many models against one data set; returns (calc[m, n], scalars, status[m]).

        `models` is a sequence of Model or a ModelBatch (packed once for
        repeated calls); `out`, if given, is a C-contiguous float64 [m, n]
        array that receives calc (rows of failed instances are left as they
        were)."""
        pods, m = _packed(models)
        n = len(data)
        if out is None:
            calc = np.zeros((m, n), np.float64)
        else:
            calc = out
            if calc.shape != (m, n) or calc.dtype != np.float64 or not calc.flags.c_contiguous:
                raise ValueError("out must be a C-contiguous float64 array of shape (%d, %d)" % (m, n))
        io = (lc_scalars * m)()
        if m:
            np.ctypeslib.as_array(io)["sfac"][:] = sfac
        status = np.zeros(m, np.int32)
        o = self._opts(scale, rdata, 0, diagnostics, max_model_points, precision)
        st = _lib().lc_light_curve_comp_batch(self._ctx, pods, m, data, n, C.byref(o), _dp(calc), io,
                                              status.ctypes.data_as(C.POINTER(C.c_int32)))
        if st != LC_OK:
            self._err(st)
        return calc, io, status

    def light_curve_comp_batch_dev(self, models, data, dev_calc, scale=False, rdata=True, diagnostics=False,
                                   sfac=1.0, precision=LC_PREC_MIXED, max_model_points=0):
        """As light_curve_comp_batch, with calc written to GPU memory: dev_calc
        is the device address (an int, e.g. a tensor's data_ptr()) of m x n
        doubles on this engine's GPU.  Returns (scalars, status)."""
        pods, m = _packed(models)
        n = len(data)
        io = (lc_scalars * m)()
        if m:
            np.ctypeslib.as_array(io)["sfac"][:] = sfac
        status = np.zeros(m, np.int32)
        o = self._opts(scale, rdata, 0, diagnostics, max_model_points, precision)
        st = _lib().lc_light_curve_comp_batch_dev(self._ctx, pods, m, data, n, C.byref(o), C.c_void_p(int(dev_calc)),
                                                  io, status.ctypes.data_as(C.POINTER(C.c_int32)))
        if st != LC_OK:
            self._err(st)
        return io, status

    def flat_sum(self, star, grid_view, phase_view, ldc, beam=0.0, spin=1.0, VFAC=0.0, XCOFM=0.0, glens1=0,
                 rlens1=0.0, precision=LC_PREC_MIXED):
        """flat_sum_star{1,2}_multi on lc_flat_grid / lc_phase_batch views."""
        out = np.zeros(phase_view.n, np.float64)
        if star == 0:
            st = _lib().lc_flat_sum_star1_multi(self._ctx, C.byref(grid_view), C.byref(phase_view), C.byref(ldc),
                                                beam, spin, VFAC, XCOFM, int(precision), _dp(out))
        else:
            st = _lib().lc_flat_sum_star2_multi(self._ctx, C.byref(grid_view), C.byref(phase_view), C.byref(ldc),
                                                beam, spin, VFAC, XCOFM, int(glens1), rlens1, int(precision),
                                                _dp(out))
        if st != LC_OK:
            self._err(st)
        return out

    def set_star_grid(self, model, star, fine, faces_view):
        st = _lib().lc_set_star_grid(self._ctx, C.byref(model.pod), int(star), int(fine), C.byref(faces_view))
        if st != LC_OK:
            self._err(st)

    def set_star_continuum(self, model, view1, view2):
        st = _lib().lc_set_star_continuum(self._ctx, C.byref(model.pod), C.byref(view1), C.byref(view2))
        if st != LC_OK:
            self._err(st)

    def profile(self):
        p = _abi.lc_profile()
        _lib().lc_ctx_get_profile(self._ctx, C.byref(p))
        out = {}
        for k, _ in p._fields_:
            v = getattr(p, k)
            out[k] = list(v) if isinstance(v, C.Array) else v
        return out

    def reset_profile(self):
        _lib().lc_ctx_reset_profile(self._ctx)

    def autotune(self, models, data, precision=LC_PREC_MIXED, max_model_points=0):
        """Picks the flux tile shape (sub-phases per lane, face chunk) for this
        workload by timing every candidate; returns (R, chunk)."""
        pods, m = _packed(models)
        o = self._opts(False, True, 0, False, max_model_points, precision)
        r, c = C.c_int32(), C.c_int32()
        st = _lib().lc_ctx_autotune(self._ctx, pods, m, data, len(data), C.byref(o), C.byref(r), C.byref(c))
        if st != LC_OK:
            self._err(st)
        return r.value, c.value

    def set_stage_cache(self, max_faces):
        """Faces of star grids kept on the device for reuse across calls (0: off)."""
        st = _lib().lc_ctx_set_stage_cache(self._ctx, int(max_faces))
        if st != LC_OK:
            self._err(st)

    def set_graphs(self, on):
        """Replay repeated batched calls of one launch shape as a captured
        graph (on by default); the profile's stage split covers only the
        calls that were not replayed."""
        st = _lib().lc_ctx_set_graphs(self._ctx, int(bool(on)))
        if st != LC_OK:
            self._err(st)


from .data_io import read_data, write_chain, write_config_and_model, write_data  # noqa: E402
