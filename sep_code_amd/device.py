"""Python face of the device layer (the C ABI of include/born_rtm.h).

``BornDevice`` mirrors the call surface of the reference's
Born/gpu_funcs_3d.h:5-27 (setup_cuda, create_gpu_space, transfer_*,
source_prop, rtm_forward, rtm_adjoint) over one explicit context, with numpy
arrays in and out.  Volumes are numpy arrays of shape (n3, n2, n1).
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import _lib
from ._lib import call, f32, i32, ptr

FMT_NONE, FMT_F32, FMT_BF16, FMT_F16 = 0, 1, 2, 3


class BornDevice:
    """One device context: replaces the reference's global device state."""

    def __init__(self, device: int = 0):
        self._lib = _lib.load()
        h = C.c_void_p()
        call("brtm_create", device, C.byref(h))
        self.h = h
        self.n = None
        self._children = weakref.WeakSet()

    # -- lifetime -----------------------------------------------------------
    def adopt(self, obj):
        """Register an object (vector, operator, network) created on this
        context; it is closed before the context is destroyed."""
        self._children.add(obj)

    def alive(self) -> bool:
        """False once the context is destroyed.  Children may still be
        destroyed after that, in any order (the garbage collector finalises
        a context and its children in arbitrary order): each child holds a
        reference to the context's host record in the library, and its
        destroy call frees its own device memory."""
        return getattr(self, "h", None) is not None and bool(self.h.value)

    def close(self):
        if self.alive():
            for ch in list(self._children):
                ch.close()
            call("brtm_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- set-up (create_gpu_space, transfer_*) ------------------------------
    def create_space(self, d, bc_a, bc_b, n):
        """d = (d1, d2, d3), n = (n1, n2, n3)."""
        call("brtm_create_space", self.h, float(d[0]), float(d[1]), float(d[2]), float(bc_a),
             float(bc_b), int(n[0]), int(n[1]), int(n[2]))
        self.n = tuple(int(x) for x in n)

    @property
    def shape(self):
        return (self.n[2], self.n[1], self.n[0])

    def set_velocity(self, V):
        V = f32(V)
        assert V.size == np.prod(self.shape)
        call("brtm_set_velocity", self.h, ptr(V))

    def set_sinc_table(self, which, table):
        t = f32(table)
        assert t.ndim == 2 and t.shape[1] == 8
        call("brtm_set_sinc_table", self.h, int(which), 8, t.shape[0], ptr(t))

    def set_source(self, locs, traces):
        locs = i32(locs)
        tr = f32(traces)
        assert tr.shape[0] == len(locs)
        call("brtm_set_source", self.h, len(locs), tr.shape[1], ptr(locs), ptr(tr))

    def set_receivers(self, locs, traces=None, ntr=None, nx=None, ny=None):
        locs = i32(locs).reshape(-1)
        nrec = len(locs)
        if nx is None:
            nx, ny = nrec, 1
        assert nx * ny == nrec
        if traces is not None:
            tr = f32(traces).reshape(nrec, -1)
            ntr = tr.shape[1]
        else:
            tr = None
        self.ntr_d, self.nrec = int(ntr), nrec
        call("brtm_set_receivers", self.h, int(nx), int(ny), int(ntr), ptr(locs), ptr(tr))

    def get_receivers(self):
        out = np.zeros((self.nrec, self.ntr_d), np.float32)
        call("brtm_get_receivers", self.h, ptr(out))
        return out

    def set_model(self, m):
        m = f32(m)
        assert m.size == np.prod(self.shape)
        call("brtm_set_model", self.h, ptr(m))

    def get_image(self):
        out = np.zeros(self.shape, np.float32)
        call("brtm_get_image", self.h, ptr(out))
        return out

    def zero_image(self):
        call("brtm_zero_image", self.h)

    def set_imaging_scale(self, c):
        call("brtm_set_imaging_scale", self.h, float(c))

    def select_shot(self, i):
        call("brtm_select_shot", self.h, int(i))

    def release_store(self, i):
        call("brtm_release_store", self.h, int(i))

    # -- propagation (source_prop, rtm_forward, rtm_adjoint) ----------------
    def source_prop(self, jt, nt, jimg, store="none", ckpt=0, get_last=False):
        kind = _lib.STORE_NAMES[store] if isinstance(store, str) else int(store)
        p0 = p1 = None
        if get_last:
            p0 = np.zeros(self.shape, np.float32)
            p1 = np.zeros(self.shape, np.float32)
        call("brtm_source_prop", self.h, int(jt), int(nt), int(jimg), kind, int(ckpt), ptr(p0),
             ptr(p1))
        return (p0, p1) if get_last else None

    def rtm_forward(self, jt, nt, jimg, accumulate=False):
        call("brtm_rtm_forward", self.h, int(jt), int(nt), int(jimg), int(bool(accumulate)))

    def rtm_adjoint(self, jt, nt, jimg):
        call("brtm_rtm_adjoint", self.h, int(jt), int(nt), int(jimg))

    def get_snapshot(self, k):
        out = np.zeros(self.shape, np.float32)
        call("brtm_get_snapshot", self.h, int(k), ptr(out))
        return out

    # -- tests and timing ---------------------------------------------------
    def propagate(self, form, nsteps, p_cur, p_old):
        a = np.array(p_cur, np.float32, copy=True, order="C")
        b = np.array(p_old, np.float32, copy=True, order="C")
        call("brtm_propagate", self.h, int(form), int(nsteps), ptr(a), ptr(b))
        return a, b

    def time_step(self, form, act, fmt, nrep):
        ms = C.c_double()
        call("brtm_time_step", self.h, int(form), int(act), int(fmt), int(nrep), C.byref(ms))
        return ms.value

    def copy_bench(self, nbytes, nrep=10):
        ms = C.c_double()
        call("brtm_copy_bench", self.h, float(nbytes), int(nrep), C.byref(ms))
        return ms.value

    def stream_bench(self, nin, nbytes, nt_store=True, nrep=10):
        """Mean ms of out = sum of nin input streams of nbytes each."""
        ms = C.c_double()
        call("brtm_stream_bench", self.h, int(nin), float(nbytes), int(bool(nt_store)), int(nrep),
             C.byref(ms))
        return ms.value

    def set_born_path(self, direct_lv):
        call("brtm_set_born_path", self.h, int(bool(direct_lv)))

    def kernel_timing(self, enable, form=0, act=0, fmt=0):
        call("brtm_kernel_timing", self.h, int(bool(enable)), int(form), int(act), int(fmt))

    def kernel_timing_result(self):
        t = C.c_double()
        n = C.c_int()
        call("brtm_kernel_timing_result", self.h, C.byref(t), C.byref(n))
        return t.value, n.value

    def set_graphs(self, mode=-1):
        """Propagation loops as graphs: 1 on, 0 off, -1 automatic (small grids)."""
        call("brtm_set_graphs", self.h, int(mode))

    def set_two_step(self, mode=1):
        """Source-field steps two per pass (temporal blocking): 1 on, 0 off
        (the default), -1 automatic (large grids)."""
        call("brtm_set_two_step", self.h, int(mode))

    def set_tuning(self, ty_tile=0, target_blocks=2048):
        call("brtm_set_tuning", self.h, int(ty_tile), int(target_blocks))

    def set_halo_ring(self, on):
        """VL steps on the six-plane LDS ring (halos loaded with the z-queue):
        1 on, 0 off, -1 automatic (the default: on for four cells per lane)."""
        call("brtm_set_halo_ring", self.h, -1 if on is None or on == -1 else int(bool(on)))

    def set_prefetch(self, slices):
        """Load lookahead of the plain four-cell ring step: 1 or 2 z slices,
        0 automatic (the default)."""
        call("brtm_set_prefetch", self.h, int(slices))

    def set_stream_loads(self, mode):
        """Nontemporal loads of p_old, V, 16-bit snapshots and the model in
        the four-cell ring step: 1 on, 0 off, -1 automatic (the default: on
        when a step moves more than 512 MB).  The image and FP32 snapshots
        stream at every size."""
        call("brtm_set_stream_loads", self.h, int(mode))

    def set_receiver_fusion(self, on):
        """Receivers injected / extracted inside the four-cell ring step when
        they share one plane (1, the default) or by separate kernels (0)."""
        call("brtm_set_receiver_fusion", self.h, int(bool(on)))

    def set_x_per_lane(self, cells):
        """1: 64-wide tiles; 2: 128-wide tiles with 8-byte accesses (VL, n1
        even); 4: 256-wide tiles with 16-byte accesses (VL, n1 % 4 == 0)."""
        call("brtm_set_x_per_lane", self.h, int(cells))

    def sync(self):
        call("brtm_sync", self.h)

    def device_bytes(self):
        b = C.c_double()
        call("brtm_device_bytes", self.h, C.byref(b))
        return b.value
