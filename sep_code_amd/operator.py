"""Python mirror of the reference's operator / solver API over the C ABI of
include/born_op.h.

* ``DeviceVector``  -- my_vector (Born/my_vector.h:10-41): HBM storage,
  FP64 dot products reduced on the GPU;
* ``RtmZeroOp``     -- rtm_zero_op (Born/rtm_zero_op_3d.h:9-38):
  ``forward(add, model, data)`` / ``adjoint(add, model, data)`` /
  ``dot_test()``;
* ``LinOp`` and its kinds (include/born_solver.h) -- my_operator:
  ``BornOp`` (rtm_zero_op), ``IOp`` (i_op), ``WeightOp`` (pointwise
  weight), ``TMute`` (tmute), ``CallbackOp`` (a Python operator),
  ``RowOp`` / ``ColOp`` / ``DiagOp`` (row_op / col_op / diag_op,
  Born/combo_oper.cpp);
* ``cgstep`` / ``SimpleSolver`` / ``RegSolver`` / ``PrecSolver`` --
  Born/cgstep.h, Born/lin_solver.h: ``SimpleSolver(cgstep(), data,
  op).solve(niter)`` runs the CG of Born/lin_solver.cpp:115-143 on the
  device, with an optional relative residual stop and an optional
  preconditioner of the gradient; the regularised and preconditioned forms
  stack their rows as the reference's col_op does.
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
APPLY_FN = C.CFUNCTYPE(_I, _P, _I, _P, _P)


class CgParams(C.Structure):
    _fields_ = [("niter", _I), ("rel_tol", _D), ("precond", PRECOND_FN), ("precond_user", _P),
                ("verbose", _I), ("offset", _P), ("history", _I), ("data_resid_hist", _P)]


class CgStats(C.Structure):
    _fields_ = [("iterations", _I), ("fallbacks", _I), ("converged", _I), ("breakdown", _I),
                ("seconds", _D), ("rel_resid", _D), ("data_rel_resid", _D)]


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
    # born_solver.h
    "brtm_linop_born": [_P, C.POINTER(_P)],
    "brtm_linop_identity": [_P, _L, C.POINTER(_P)],
    "brtm_linop_diagonal": [_P, C.POINTER(_P)],
    "brtm_linop_tmute": [_P, _I, _F, _F, _I, _F, _F, _L, _F, _F, C.POINTER(_P)],
    "brtm_linop_callback": [_P, _L, _L, APPLY_FN, APPLY_FN, _P, C.POINTER(_P)],
    "brtm_linop_chain": [_I, _P, C.POINTER(_P)],
    "brtm_linop_column": [_I, _P, C.POINTER(_P)],
    "brtm_linop_blockdiag": [_I, _P, C.POINTER(_P)],
    "brtm_linop_set_scale": [_P, _D],
    "brtm_linop_domain_size": [_P],
    "brtm_linop_range_size": [_P],
    "brtm_linop_forward": [_P, _I, _P, _P],
    "brtm_linop_adjoint": [_P, _I, _P, _P],
    "brtm_linop_dot_test": [_P, C.c_ulonglong, C.POINTER(_D)],
    "brtm_linop_destroy": [_P],
    "brtm_solver_create": [_I, _P, _P, _P, _D, _P, _P, C.POINTER(_P)],
    "brtm_solver_solve": [_P, C.POINTER(CgParams), _P, _P, C.POINTER(CgStats)],
    "brtm_solver_return_model": [_P, _P],
    "brtm_solver_update_model": [_P, _P],
    "brtm_solver_iterate": [_P],
    "brtm_solver_residual": [_P],
    "brtm_solver_destroy": [_P],
}
_RES = {"brtm_vec_size": _L, "brtm_op_model_size": _L, "brtm_op_data_size": _L,
        "brtm_vec_data": _P, "brtm_linop_domain_size": _L, "brtm_linop_range_size": _L,
        "brtm_solver_iterate": _P, "brtm_solver_residual": _P}
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

    def linop(self) -> "LinOp":
        """This operator as a composable LinOp (my_operator)."""
        if getattr(self, "_linop", None) is None:
            self._linop = BornOp(self)
        return self._linop

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
    # |L m - d| / |d| of the data as given (the judged residual without an
    # offset; the noisy-data residual with one)
    data_rel_resid: float = float("nan")
    data_resid_hist: np.ndarray | None = None


def cg_solve(op: RtmZeroOp, data: DeviceVector, model: DeviceVector, niter: int,
             rel_tol=0.0, verbose=False):
    """brtm_cg_solve: plain cgstep on the Born operator from m = 0 in one
    call (the C entry point for callers without the solver objects)."""
    dh = np.zeros(niter + 1)
    p = CgParams(int(niter), float(rel_tol), PRECOND_FN(), None, int(verbose), None, 0,
                 ptr(dh))
    rh = np.zeros(niter + 1)
    th = np.zeros(niter + 1)
    st = CgStats()
    call("brtm_cg_solve", op.h, data.h, model.h, C.byref(p), ptr(rh), ptr(th), C.byref(st))
    k = st.iterations + 1
    return SolveResult(st.iterations, st.fallbacks, bool(st.converged), bool(st.breakdown),
                       st.seconds, st.rel_resid, rh[:k], th[:k], st.data_rel_resid, dh[:k])


# --- operators (my_operator and its kinds) -----------------------------------


class LinOp:
    """my_operator (Born/my_operator.h:7-36) over the library's brtm_linop:
    forward(add, model, data), adjoint(add, model, data), set_scale(s),
    dot_test().  Combinations keep references to their parts."""

    def __init__(self, dev, h, parts=()):
        self.dev = dev
        self.h = h
        self._parts = list(parts)
        dev.adopt(self)

    @property
    def domain_size(self) -> int:
        return _lib.load().brtm_linop_domain_size(self.h)

    @property
    def range_size(self) -> int:
        return _lib.load().brtm_linop_range_size(self.h)

    def domain_vector(self, a=None):
        v = DeviceVector(self.dev, self.domain_size)
        if a is not None:
            v.upload(a)
        return v

    def range_vector(self, a=None):
        v = DeviceVector(self.dev, self.range_size)
        if a is not None:
            v.upload(a)
        return v

    def forward(self, add, model, data, iter=0):
        call("brtm_linop_forward", self.h, int(bool(add)), model.h, data.h)
        return True

    def adjoint(self, add, model, data, iter=0):
        call("brtm_linop_adjoint", self.h, int(bool(add)), model.h, data.h)
        return True

    def set_scale(self, s):
        call("brtm_linop_set_scale", self.h, float(s))
        return self

    def dot_test(self, verb=False, seed=12345):
        err = (_D * 2)()
        call("brtm_linop_dot_test", self.h, C.c_ulonglong(seed), err)
        if verb:
            print(f"dot test: add=false {err[0]:.3e}, add=true {err[1]:.3e}")
        return err[0], err[1]

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            call("brtm_linop_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _new(fn, *args):
    h = _P()
    call(fn, *args, C.byref(h))
    return h


def as_linop(op) -> LinOp:
    """An RtmZeroOp as a LinOp (rtm_zero_op is a my_operator)."""
    if isinstance(op, LinOp):
        return op
    if isinstance(op, RtmZeroOp):
        return op.linop()
    raise TypeError(f"not an operator: {type(op).__name__}")


class BornOp(LinOp):
    """rtm_zero_op as a composable operator."""

    def __init__(self, op: RtmZeroOp):
        _ensure()
        super().__init__(op.dev, _new("brtm_linop_born", op.h), parts=[op])


class IOp(LinOp):
    """i_op (Born/i_op.cpp:4-28): out = scale * in."""

    def __init__(self, dev, n: int):
        _ensure()
        super().__init__(dev, _new("brtm_linop_identity", dev.h, int(n)))


class WeightOp(LinOp):
    """Pointwise weight w . in (self-adjoint); keeps w alive."""

    def __init__(self, w: DeviceVector):
        _ensure()
        super().__init__(w.dev, _new("brtm_linop_diagonal", w.h), parts=[w])


class TMute(LinOp):
    """tmute (Born/tmute.cpp:4-76) on data [nouter][nrx][nt]: zero the
    samples before t0 + x_rx / vmute, x_rx = orx + drx * rx."""

    def __init__(self, dev, t0, vmute, nt, ot, dt, nrx, orx, drx, nouter):
        _ensure()
        super().__init__(dev, _new("brtm_linop_tmute", dev.h, int(nt), float(ot), float(dt),
                                   int(nrx), float(orx), float(drx), int(nouter), float(t0),
                                   float(vmute)))


class CallbackOp(LinOp):
    """A caller-defined operator: forward(add, x, y) computes y (+)= A x and
    adjoint(add, y, x) computes x (+)= A' y on DeviceVector views."""

    def __init__(self, dev, ndomain, nrange, forward, adjoint):
        _ensure()
        nd, nr = int(ndomain), int(nrange)

        def wrap(fn, nin, nout):
            def cb(user, add, hin, hout):
                try:
                    fn(bool(add), DeviceVector.view(dev, hin, nin),
                       DeviceVector.view(dev, hout, nout))
                    return 0
                except Exception as e:  # reported through the C status
                    print(f"operator callback raised: {e}")
                    return 1
            return APPLY_FN(cb)

        self._cbs = (wrap(forward, nd, nr), wrap(adjoint, nr, nd))
        super().__init__(dev, _new("brtm_linop_callback", dev.h, nd, nr, self._cbs[0],
                                   self._cbs[1], None))


def _combine(fn, ops):
    ops = [as_linop(o) for o in ops]
    arr = (_P * len(ops))(*[o.h.value for o in ops])
    return ops, _new(fn, len(ops), arr)


class RowOp(LinOp):
    """row_op (Born/combo_oper.cpp:176-235): RowOp(A, B) applies A then B."""

    def __init__(self, *ops):
        _ensure()
        ops, h = _combine("brtm_linop_chain", ops)
        super().__init__(ops[0].dev, h, parts=ops)


class ColOp(LinOp):
    """col_op (Born/combo_oper.cpp:1-99): [A m; B m; ...], ranges stacked."""

    def __init__(self, *ops):
        _ensure()
        ops, h = _combine("brtm_linop_column", ops)
        super().__init__(ops[0].dev, h, parts=ops)


class DiagOp(LinOp):
    """diag_op (Born/combo_oper.cpp:103-175): block diagonal."""

    def __init__(self, *ops):
        _ensure()
        ops, h = _combine("brtm_linop_blockdiag", ops)
        super().__init__(ops[0].dev, h, parts=ops)


# --- solvers (lin_solver) ----------------------------------------------------
SOLVER_SIMPLE, SOLVER_REG, SOLVER_PREC = 0, 1, 2


class _Solver:
    """lin_solver (Born/lin_solver.cpp): solve(niter) continues from the
    current state; return_model() gives the model (the update when m0 was
    given, as in the reference); update_model(v) adds it to v."""

    def __init__(self, kind, step, data, op, aux, eps, wt, m0, precond, verbose, history):
        assert isinstance(step, cgstep)
        _ensure()
        self.op = as_linop(op)
        self.aux = None if aux is None else as_linop(aux)
        self.wt = None if wt is None else as_linop(wt)
        self.kind = kind
        self.dev = self.op.dev
        self.history = int(history)
        self.verbose = verbose
        self.data = data
        self._m0 = m0
        self.h = _new("brtm_solver_create", kind, self.op.h, data.h,
                      self.aux.h if self.aux is not None else None, float(eps),
                      self.wt.h if self.wt is not None else None,
                      m0.h if m0 is not None else None)
        self.dev.adopt(self)
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

    @property
    def model(self) -> DeviceVector:
        """The current iterate in the solver's domain (a view, updated in place)."""
        lib = _lib.load()
        return DeviceVector.view(self.dev, lib.brtm_solver_iterate(self.h),
                                 (self.aux if self.kind == SOLVER_PREC else self.op).domain_size)

    @property
    def residual(self) -> DeviceVector:
        """The stacked residual rr (a view)."""
        lib = _lib.load()
        h = lib.brtm_solver_residual(self.h)
        return DeviceVector.view(self.dev, h, lib.brtm_vec_size(h))

    def solve(self, niter, rel_tol=0.0, offset: DeviceVector | None = None) -> SolveResult:
        """offset: range vector e; the judged residual is |rr + e|/|rr0 + e|
        (with rr0 = -d and e the synthetic noise: the misfit to the
        noise-free data)."""
        p = CgParams(int(niter), float(rel_tol),
                     self._precond_cb if self._precond_cb is not None else PRECOND_FN(),
                     self._precond_user, int(self.verbose),
                     offset.h if offset is not None else None, self.history, None)
        rh = np.zeros(niter + 1)
        th = np.zeros(niter + 1)
        dh = np.zeros(niter + 1)
        p.data_resid_hist = ptr(dh)
        st = CgStats()
        call("brtm_solver_solve", self.h, C.byref(p), ptr(rh), ptr(th), C.byref(st))
        k = st.iterations + 1
        return SolveResult(st.iterations, st.fallbacks, bool(st.converged), bool(st.breakdown),
                           st.seconds, st.rel_resid, rh[:k], th[:k], st.data_rel_resid, dh[:k])

    def return_model(self) -> DeviceVector:
        v = DeviceVector(self.dev, self.op.domain_size)
        call("brtm_solver_return_model", self.h, v.h)
        return v

    def update_model(self, model: DeviceVector):
        call("brtm_solver_update_model", self.h, model.h)
        return model

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            call("brtm_solver_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SimpleSolver(_Solver):
    """simple_solver(step, data, op, wt, m0): min |W(L m - d)|
    (Born/lin_solver.cpp:7-17)."""

    def __init__(self, step, data: DeviceVector, op, precond=None, verbose=False, history=0,
                 wt=None, m0=None):
        """history: 0/1 = cgstep's plane search; h >= 2 = flexible CG over the
        new direction and the last h steps (for preconditioners that change
        between iterations)."""
        super().__init__(SOLVER_SIMPLE, step, data, op, None, 0.0, wt, m0, precond, verbose,
                         history)


class RegSolver(_Solver):
    """reg_solver(step, data, op, reg, eps, wt, m0): min |W(L m - d)|^2 +
    |eps R m|^2 (Born/lin_solver.cpp:39-53; sets reg's scale to eps)."""

    def __init__(self, step, data: DeviceVector, op, reg, eps, wt=None, m0=None, precond=None,
                 verbose=False, history=0):
        super().__init__(SOLVER_REG, step, data, op, reg, eps, wt, m0, precond, verbose, history)


class PrecSolver(_Solver):
    """prec_solver(step, data, op, prec, eps, wt, m0): m = P p, min
    |W(L P p - d)|^2 + |eps p|^2 (Born/lin_solver.cpp:21-37); return_model
    gives P p."""

    def __init__(self, step, data: DeviceVector, op, prec, eps, wt=None, m0=None, precond=None,
                 verbose=False, history=0):
        super().__init__(SOLVER_PREC, step, data, op, prec, eps, wt, m0, precond, verbose,
                         history)
