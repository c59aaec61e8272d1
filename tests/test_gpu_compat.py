"""The reference's device ABI (include/gpu_funcs_3d.h), called exactly as
Born/rtm_zero_op_3d.cpp calls it, against the FP64 oracle."""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import born_oracle as O
from sep_code_amd import _lib
from sep_code_amd import geometry as G

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_reference_abi_forward_adjoint(dev):
    print("compat ABI", file=sys.stderr, flush=True)
    lib = _lib.load()
    P, I, F = C.c_void_p, C.c_int, C.c_float
    lib.create_gpu_space.argtypes = [F, F, F, F, F, F, I, I, I]
    lib.transfer_vel_func1.argtypes = [I, I, I, P]
    lib.transfer_sinc_table_s.argtypes = [I, I, P]
    lib.transfer_sinc_table_d.argtypes = [I, I, P]
    lib.transfer_source_func.argtypes = [I, I, P, P]
    lib.transfer_receiver_func.argtypes = [I, I, I, P, P]
    lib.source_prop.argtypes = [I, I, I, C.c_bool, C.c_bool, P, P, I, I, I]
    lib.rtm_forward.argtypes = [I, I, I, I, P, P, I, I, I, I, I]
    lib.rtm_adjoint.argtypes = [I, I, I, I, P, P, P, I, I]
    lib.setup_cuda.argtypes = [I, I, P]
    for f in ("create_gpu_space", "transfer_vel_func1", "transfer_sinc_table_s",
              "transfer_sinc_table_d", "transfer_source_func", "transfer_receiver_func",
              "source_prop", "rtm_forward", "rtm_adjoint", "setup_cuda"):
        getattr(lib, f).restype = None

    r = np.random.default_rng(8)
    n = (40, 36, 44)
    n1, n2, n3 = n
    d = (10.0, 10.0, 10.0)
    V = ((r.uniform(1500, 3000, (n3, n2, n1)) * 0.001) ** 2).astype(np.float32)
    nt = 90
    ntb = nt + 8
    w = G.wavelet_derivative(G.ricker(25.0, 0.001, nt), 0.001)
    slocs, strc = G.source_points(n1, n2, 20.3, 17.6, 9.5, w, ntb)
    rlocs, nx, ny = G.surface_receivers(n, 4, 3.0, step=4, origin=2)
    lib.setup_cuda(0, 0, None)
    lib.create_gpu_space(*d, 8.0, 0.02, 0.02, n1, n2, n3)
    lib.transfer_vel_func1(n1, n2, n3, V.ctypes.data)
    row = np.zeros(8, np.float32)
    row[3] = 1
    tabs = (C.c_void_p * 1)(row.ctypes.data)
    lib.transfer_sinc_table_s(8, 1, tabs)
    lib.transfer_sinc_table_d(8, 1, tabs)
    lib.transfer_source_func(9, ntb, slocs.ctypes.data, strc.ctypes.data)
    zeros = np.zeros((len(rlocs), ntb), np.float32)
    lib.transfer_receiver_func(nx, ny, ntb, rlocs.ctypes.data, zeros.ctypes.data)
    m = np.zeros((n3, n2, n1), np.float32)
    m[4:-4, 4:-4, 4:-4] = r.standard_normal((n3 - 8, n2 - 8, n1 - 8))
    dat = np.zeros((len(rlocs), ntb), np.float32)
    lib.rtm_forward(n1, n2, n3, 1, m.ctypes.data, dat.ctypes.data, 9, nt, ntb, nx, ny)
    Pr = O.Problem(n, d, V.astype(np.float64), 8.0, 0.02, src_locs=slocs,
                   src_tr=strc.astype(np.float64), rec_locs=O.mask_inactive(rlocs, n), ntr_d=ntb,
                   nt=nt, jt=1, jimg=1, cimg=1e-6)
    Do = Pr.born_forward(m.ravel().astype(np.float64), None)
    assert rel(dat, Do) < 1e-5
    # adjoint: data in the receiver traces, image out
    dd = r.standard_normal(dat.shape).astype(np.float32)
    lib.transfer_receiver_func(nx, ny, ntb, rlocs.ctypes.data, dd.ctypes.data)
    img = np.zeros_like(m)
    lib.rtm_adjoint(n1, n2, n3, 1, None, None, img.ctypes.data, 9, nt)
    io = Pr.rtm_adjoint(dd.astype(np.float64), None, ckpt=10)
    assert rel(img, io) < 1e-4
    # source_prop returns the last two fields of nt + 1 steps
    p0 = np.zeros_like(m)
    p1 = np.zeros_like(m)
    lib.source_prop(n1, n2, n3, False, True, p0.ctypes.data, p1.ctypes.data, 1, 9, nt - 1)
    _, last2 = Pr.source_prop(want_snaps=False)
    assert rel(p0, last2[0]) < 1e-5 and rel(p1, last2[1]) < 1e-5


def test_reference_abi_source_prop_closed_form(dev):
    """The reference-signature source_prop runs it = 0 .. nt, nt + 1 steps
    (Born/gpu_funcs_3d.cu:227), and returns (s^{nt}, s^{nt+1}) in p0, p1
    (the pointer swap after the last step, :287-289).  With V = 1e-20 (its
    coupling is far below FP32 rounding) and no sponge, each source cell
    follows p' = 2p - p_old + f, so s^n = scale * sum_{t<=n-2} (n-1-t) w[t]
    at every source point: the +4 trace offset read through the zero-shift
    sinc row delays the wavelet by one sample."""
    print("compat ABI source_prop closed form", file=sys.stderr, flush=True)
    lib = _lib.load()
    P, I, F = C.c_void_p, C.c_int, C.c_float
    lib.create_gpu_space.argtypes = [F, F, F, F, F, F, I, I, I]
    lib.transfer_vel_func1.argtypes = [I, I, I, P]
    lib.transfer_sinc_table_s.argtypes = [I, I, P]
    lib.transfer_source_func.argtypes = [I, I, P, P]
    lib.source_prop.argtypes = [I, I, I, C.c_bool, C.c_bool, P, P, I, I, I]
    lib.setup_cuda.argtypes = [I, I, P]
    for f in ("create_gpu_space", "transfer_vel_func1", "transfer_sinc_table_s",
              "transfer_source_func", "source_prop", "setup_cuda"):
        getattr(lib, f).restype = None
    r = np.random.default_rng(21)
    n = (40, 36, 44)
    n1, n2, n3 = n
    nt = 50                      # the reference's nt argument
    ntb = nt + 10
    w = r.standard_normal(nt + 1)
    fx, fy, fz = 20.3, 17.6, 9.5
    slocs, strc = G.source_points(n1, n2, fx, fy, fz, w, ntb)
    lib.setup_cuda(0, 0, None)
    lib.create_gpu_space(10.0, 10.0, 10.0, 0.0, 0.0, 0.0, n1, n2, n3)
    V = np.full((n3, n2, n1), 1e-20, np.float32)
    lib.transfer_vel_func1(n1, n2, n3, V.ctypes.data)
    row = np.zeros(8, np.float32)
    row[3] = 1
    lib.transfer_sinc_table_s(8, 1, (C.c_void_p * 1)(row.ctypes.data))
    strc = np.ascontiguousarray(strc, np.float32)  # kept alive across the call
    lib.transfer_source_func(9, ntb, slocs.ctypes.data, strc.ctypes.data)
    p0 = np.zeros((n3, n2, n1), np.float32)
    p1 = np.zeros_like(p0)
    lib.source_prop(n1, n2, n3, False, True, p0.ctypes.data, p1.ctypes.data, 1, 9, nt)
    cw = np.cumsum(np.cumsum(w))
    iz, iy, ix = int(fz) - 1, int(fy), int(fx)
    j = 0
    for i2 in (iy - 1, iy, iy + 1):
        for i1 in (ix - 1, ix, ix + 1):
            scale = np.exp(-(fz - iz) ** 2 - (fx - i1) ** 2 - (fy - i2) ** 2)
            assert slocs[j] == i1 + i2 * n1 + iz * n1 * n2
            # s^{nt} in p0 and s^{nt+1} in p1
            for field, nn in ((p0, nt), (p1, nt + 1)):
                want = scale * cw[nn - 2]
                got = field.ravel()[slocs[j]]
                # FP32 rounding accumulates through the double sum: a scalar
                # FP32 run of this recursion stays within 5e-6 of the largest
                # |s| over seeds 0-4, so 3e-5 of it leaves a margin
                assert abs(got - want) <= 3e-5 * scale * np.abs(cw[:nt]).max(), (j, nn, got, want)
            j += 1
    other = np.ones(p1.size, bool)
    other[slocs] = False
    assert np.abs(p1.ravel()[other]).max() <= 1e-12 * np.abs(p1).max()
