"""Python mirror of the reference's operator / solver API over the C ABI of
include/born_op.h.

* ``DeviceVector``  -- my_vector (Born/my_vector.h:10-41): HBM storage,
  FP64 dot products reduced on the GPU;
* ``RtmZeroOp``     -- rtm_zero_op (Born/rtm_zero_op_3d.h:9-38):
  ``forward(add, model, data)`` / ``adjoint(add, model, data)`` /
  ``dot_test()``;
* ``cgstep`` / ``SimpleSolver`` -- Born/cgstep.h, Born/lin_solver.h:
  ``SimpleSolver(cgstep(), data, op).solve(niter)`` runs the CG of
  Born/lin_solver.cpp:115-143 on the device, with an optional relative
  data-residual stop and an optional preconditioner of the gradient.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _lib
from ._lib import call, f32, i32, ptr

_P, _I, _L, _D, _F = C.c_void_p, C.c_int, C.c_long, C.c_double, C.c_float


class OpDesc(C.Structure):
    """brtm_op_desc of include/born_op.h."""

    _fields_ = [("m1", _I), ("m2", _I), ("m3", _I), ("o1", _I), ("o2", _I), ("o3", _I),
                ("nshots", _I), ("nsrc", _I), ("src_locs", _P), ("src_traces", _P),
                ("ntr_s", _I), ("nrx", _I), ("nry", _I), ("rec_locs", _P), ("nt_data", _I),
                ("dt_data", _F), ("jt", _I), ("jimg", _I), ("store", _I), ("ckpt", _I),
                ("mute", _P)]


PRECOND_FN = C.CFUNCTYPE(_I, _P, _P, _P)


class CgParams(C.Structure):
    _fields_ = [("niter", _I), ("rel_tol", _D), ("precond", PRECOND_FN), ("precond_user", _P),
                ("verbose", _I), ("offset", _P), ("history", _I)]


class CgStats(C.Structure):
    _fields_ = [("iterations", _I), ("fallbacks", _I), ("converged", _I), ("breakdown", _I),
                ("seconds", _D), ("rel_resid", _D)]


_SIGS = {
    "brtm_vec_create": [_P, _L, C.POINTER(_P)],
    "brtm_vec_destroy": [_P],
    "brtm_vec_size": [_P],
    "brtm_vec_data": [_P],
    "brtm_vec_upload": [_P, _P],
    "brtm_vec_download": [_P, _P],
    "brtm_vec_zero": [_P],
    "brtm_vec_copy": [_P, _P],
    "brtm_vec_scale": [_P, _D],
    "brtm_vec_scale_add": [_P, _D, _P, _D],
    "brtm_vec_add": [_P, _P],
    "brtm_vec_mult": [_P, _P],
    "brtm_vec_dot": [_P, _P, C.POINTER(_D)],
    "brtm_vec_random": [_P, C.c_ulonglong],
    "brtm_op_create": [_P, C.POINTER(OpDesc), C.POINTER(_P)],
    "brtm_op_destroy": [_P],
    "brtm_op_prepare": [_P],
    "brtm_op_model_size": [_P],
    "brtm_op_data_size": [_P],
    "brtm_op_forward": [_P, _I, _P, _P],
    "brtm_op_adjoint": [_P, _I, _P, _P],
    "brtm_op_dot_test": [_P, C.c_ulonglong, C.POINTER(_D)],
    "brtm_op_illumination": [_P, _P],
    "brtm_cg_solve": [_P, _P, _P, C.POINTER(CgParams), _P, _P, C.POINTER(CgStats)],
    "brtm_sinc_table": [_I, _I, _P],
}
_RES = {"brtm_vec_size": _L, "brtm_op_model_size": _L, "brtm_op_data_size": _L,
        "brtm_vec_data": _P}
_registered = False


def _ensure():
    global _registered
    if not _registered:
        _lib.register(_SIGS, _RES)
        _registered = True
    return _lib.load()


def sinc_table(ntab: int, nsinc: int = 8) -> np.ndarray:
    """Hale's least-squares sinc tables (Born/sinc_bob.cpp) computed by the library."""
    _ensure()
    out = np.zeros((ntab, nsinc), np.float32)
    call("brtm_sinc_table", int(ntab), int(nsinc), ptr(out))
    return out


class DeviceVector:
    """A float vector in HBM (my_vector)."""

    def __init__(self, dev, n: int):
        _ensure()
        self.dev = dev
        self.n = int(n)
        h = _P()
        call("brtm_vec_create", dev.h, self.n, C.byref(h))
        self.h = h
        dev.adopt(self)

    @classmethod
    def view(cls, dev, handle, n: int):
        """Non-owning wrapper of a library vector handle (e.g. the g and z a
        preconditioner callback receives)."""
        v = cls.__new__(cls)
        v.dev = dev
        v.n = int(n)
        v.h = _P(handle)
        v._view = True
        return v

    @classmethod
    def from_array(cls, dev, a):
        a = f32(a).reshape(-1)
        v = cls(dev, a.size)
        v.upload(a)
        return v

    def close(self):
        if getattr(self, "_view", False):
            self.h = None
            return
        if getattr(self, "h", None) is not None and self.h.value:
            if not self.dev.alive():
                self.h = None
                return
            call("brtm_vec_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, a):
        a = f32(a).reshape(-1)
        assert a.size == self.n
        call("brtm_vec_upload", self.h, ptr(a))

    def numpy(self):
        out = np.zeros(self.n, np.float32)
        call("brtm_vec_download", self.h, ptr(out))
        return out

    @property
    def data_ptr(self) -> int:
        return _lib.load().brtm_vec_data(self.h)

    def zero(self):
        call("brtm_vec_zero", self.h)
        return self

    def clone(self):
        v = DeviceVector(self.dev, self.n)
        call("brtm_vec_copy", v.h, self.h)
        return v

    def copy_from(self, other):
        call("brtm_vec_copy", self.h, other.h)
        return self

    def scale(self, s):
        call("brtm_vec_scale", self.h, float(s))
        return self

    def scale_add(self, a, x, b):
        """self = a*self + b*x (my_vector::scale_add)."""
        call("brtm_vec_scale_add", self.h, float(a), x.h, float(b))
        return self

    def add(self, x):
        call("brtm_vec_add", self.h, x.h)
        return self

    def mult(self, x):
        call("brtm_vec_mult", self.h, x.h)
        return self

    def dot(self, x) -> float:
        out = _D()
        call("brtm_vec_dot", self.h, x.h, C.byref(out))
        return out.value

    def random(self, seed: int):
        call("brtm_vec_random", self.h, C.c_ulonglong(seed))
        return self


class RtmZeroOp:
    """The multi-shot Born operator (rtm_zero_op) on one device context.

    ``workload`` is a ``configs.Workload``; the model grid is the compute
    grid interior from ``nbound`` (the model origin, -d*nbound in physical
    units) unless ``model_box`` = (m1, m2, m3, o1, o2, o3) is given."""

    def __init__(self, dev, workload, store=None, ckpt=None, model_box=None, mute=None,
                 nt_data=None):
        _ensure()
        w = workload
        self.dev = dev
        self.w = w
        n1, n2, n3 = w.n
        if model_box is None:
            model_box = (n1, n2, n3, 0, 0, 0)
        self.model_box = tuple(int(x) for x in model_box)
        m1, m2, m3, o1, o2, o3 = self.model_box
        ns = len(w.shots)
        nsrc = len(w.shots[0].src_locs)
        self._src_locs = i32(np.stack([s.src_locs for s in w.shots]))
        self._src_tr = f32(np.stack([s.src_traces for s in w.shots]))
        self._rec_locs = i32(np.stack([s.rec_locs for s in w.shots]))
        self._mute = None if mute is None else f32(mute)
        nt_data = int(nt_data or w.nt)
        self.nt_data = nt_data
        self.nrx, self.nry = w.shots[0].nrx, w.shots[0].nry
        store = store or w.store
        self.desc = OpDesc(m1, m2, m3, o1, o2, o3, ns, nsrc, ptr(self._src_locs),
                           ptr(self._src_tr), self._src_tr.shape[-1], self.nrx, self.nry,
                           ptr(self._rec_locs), nt_data, float(w.dt * w.jt), w.jt, w.jimg,
                           _lib.STORE_NAMES[store], int(ckpt or w.ckpt), ptr(self._mute))
        h = _P()
        call("brtm_op_create", dev.h, C.byref(self.desc), C.byref(h))
        self.h = h
        dev.adopt(self)
        self.model_shape = (m3, m2, m1)
        self.data_shape = (ns, self.nry * self.nrx, nt_data)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            if not self.dev.alive():
                self.h = None
                return
            call("brtm_op_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def model_size(self):
        return _lib.load().brtm_op_model_size(self.h)

    @property
    def data_size(self):
        return _lib.load().brtm_op_data_size(self.h)

    def prepare(self):
        """create_source_fields: propagate and store every shot's source field."""
        call("brtm_op_prepare", self.h)

    def model_vector(self, a=None):
        v = DeviceVector(self.dev, self.model_size)
        if a is not None:
            v.upload(a)
        return v

    def data_vector(self, a=None):
        v = DeviceVector(self.dev, self.data_size)
        if a is not None:
            v.upload(a)
        return v

    def forward(self, add, model, data, iter=0):
        call("brtm_op_forward", self.h, int(bool(add)), model.h, data.h)
        return True

    def adjoint(self, add, model, data, iter=0):
        call("brtm_op_adjoint", self.h, int(bool(add)), model.h, data.h)
        return True

    def illumination(self):
        """Source illumination sum_shots sum_k s_k^2 on the model grid (muted)."""
        v = DeviceVector(self.dev, self.model_size)
        call("brtm_op_illumination", self.h, v.h)
        return v

    def dot_test(self, verb=False, seed=12345):
        err = (_D * 2)()
        call("brtm_op_dot_test", self.h, C.c_ulonglong(seed), err)
        if verb:
            print(f"dot test: add=false {err[0]:.3e}, add=true {err[1]:.3e}")
        return err[0], err[1]


class cgstep:  # noqa: N801  (the reference's class name)
    """Marker for Claerbout's CG step (Born/cgstep.cpp); the step itself runs
    inside brtm_cg_solve."""


@dataclasses.dataclass
class SolveResult:
    iterations: int
    fallbacks: int
    converged: bool
    breakdown: bool
    seconds: float
    rel_resid: float
    resid_hist: np.ndarray
    time_hist: np.ndarray


class SimpleSolver:
    """simple_solver(step, data, op) with solve(niter) (Born/lin_solver.cpp)."""

    def __init__(self, step, data: DeviceVector, op: RtmZeroOp, precond=None, verbose=False,
                 history=0):
        """history: 0/1 = cgstep's plane search; h >= 2 = flexible CG over the
        new direction and the last h steps (for preconditioners that change
        between iterations)."""
        assert isinstance(step, cgstep)
        self.op = op
        self.history = int(history)
        self.data = data
        self.model = op.model_vector()
        self.verbose = verbose
        self._precond_cb = None
        self._precond_user = None
        self._precond = precond
        if isinstance(precond, tuple) and precond[0] == "native":
            # a C preconditioner exported by the library (e.g. brtm_unet_precond)
            self._precond_cb = PRECOND_FN(precond[1])
            self._precond_user = precond[2]
        elif precond is not None:
            def cb(user, g_h, z_h):
                try:
                    return int(precond(g_h, z_h) or 0)
                except Exception as e:  # report through the C status
                    print(f"preconditioner raised: {e}")
                    return 1

            self._precond_cb = PRECOND_FN(cb)

    def solve(self, niter, rel_tol=0.0, offset: DeviceVector | None = None) -> SolveResult:
        """offset: data-space vector e; the judged residual is |rr + e|/|d - e|
        (with e the synthetic noise: the misfit to the noise-free data)."""
        p = CgParams(int(niter), float(rel_tol),
                     self._precond_cb if self._precond_cb is not None else PRECOND_FN(),
                     self._precond_user, int(self.verbose),
                     offset.h if offset is not None else None, self.history)
        rh = np.zeros(niter + 1)
        th = np.zeros(niter + 1)
        st = CgStats()
        call("brtm_cg_solve", self.op.h, self.data.h, self.model.h, C.byref(p), ptr(rh), ptr(th),
             C.byref(st))
        k = st.iterations + 1
        return SolveResult(st.iterations, st.fallbacks, bool(st.converged), bool(st.breakdown),
                           st.seconds, st.rel_resid, rh[:k], th[:k])

    def return_model(self):
        return self.model
