"""CPU tests of the host side: the C-ABI library loads and exports every
function include/*.h declares (no compute calls without a GPU), the host
geometry mirrors the reference conventions, and the seeded generators are
deterministic."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import born_oracle as O
from sep_code_amd import _lib, configs
from sep_code_amd import geometry as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    inc = os.path.join(ROOT, "include")
    for fn in sorted(os.listdir(inc)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"typedef[^;]*;", "", text)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;",
                             text, flags=re.M):
            if m.group(1) not in ("if", "while", "for", "return"):
                names.append((fn, m.group(1)))
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    decl = declared_functions()
    assert len(decl) > 20
    missing = [f"{h}:{n}" for h, n in decl if not hasattr(lib, n)]
    assert not missing, missing


CALLER = r"""
#include "gpu_funcs_3d.h"
#include "born_rtm.h"
#include "born_op.h"
#include "born_solver.h"
#include "born_unet.h"

/* Every declared function, by address: the link fails on a missing one. */
static void (*const every_symbol[])(void) = {
%s
};

/* A host driver in the order the reference's rtm_zero_op calls its device
   layer (set-up, shot transfers, source propagation, imaging); it is only
   compiled and linked, never run. */
int drive(int n1, int n2, int n3, int nt, float *vel, float **sinc, int *slocs, float *src,
          int nx, int ny, int *rlocs, float *data, float *img, float *p0, float *p1) {
  setup_cuda(1, 0, 0);
  create_gpu_space(10.f, 10.f, 10.f, 30.f, 0.001f, 0.001f, n1, n2, n3);
  transfer_vel_func1(n1, n2, n3, vel);
  transfer_sinc_table_s(8, 1, sinc);
  transfer_sinc_table_d(8, 1, sinc);
  transfer_source_func(9, nt + 8, slocs, src);
  source_prop(n1, n2, n3, 0, 1, p0, p1, 1, 9, nt);
  transfer_receiver_func(nx, ny, nt + 8, rlocs, data);
  rtm_adjoint(n1, n2, n3, 1, p0, p1, img, 9, nt);
  rtm_forward(n1, n2, n3, 1, img, data, 9, nt, nt + 8, nx, ny);
  return (int)(sizeof every_symbol / sizeof every_symbol[0]);
}
"""


@pytest.mark.parametrize("lang", ["c", "c++"])
def test_headers_compile_and_link_as_c_and_cpp(tmp_path, lang):
    """Every header compiles as C and as C++ in one translation unit, and a
    caller that takes the address of every declared function and drives
    the reference-signature layer links against the library."""
    import shutil
    import subprocess
    cc = {"c": "gcc", "c++": "g++"}[lang]
    if shutil.which(cc) is None:
        pytest.skip(f"{cc} not installed")
    _lib.load()  # builds the library when it is missing
    names = sorted({n for _, n in declared_functions()})
    refs = ",\n".join(f"  (void (*)(void)){n}" for n in names)
    src = tmp_path / ("caller.c" if lang == "c" else "caller.cpp")
    src.write_text(CALLER % refs + "int main(void) { return 0; }\n")
    libdir = os.path.join(ROOT, "sep_code_amd", "lib")
    cmd = [cc, "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
           "-o", str(tmp_path / "caller"), "-L", libdir, "-lborn_rtm",
           "-Wl,--allow-shlib-undefined", "-Wl,-rpath," + libdir]
    if lang == "c":
        cmd.insert(1, "-std=c11")
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


def test_python_binding_covers_born_rtm_header():
    decl = {n for h, n in declared_functions() if h == "born_rtm.h"}
    assert decl <= set(_lib.SIGNATURES), decl - set(_lib.SIGNATURES)


def test_error_path_without_gpu_is_an_error_code():
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.brtm_create(-1, ctypes.byref(h))
    assert rc != 0 and lib.brtm_last_error()
    assert lib.brtm_create_space(None, 1.0, 1.0, 1.0, 0.0, 0.0, 16, 16, 16) == -1


def test_splitmix64_known_answer():
    """SplitMix64 of Steele, Lea & Flood from state 0: published first outputs."""
    r = configs.SplitMix64(0)
    assert [r.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                                  0x06C45D188009454F]
    a = configs.SplitMix64(99)
    b = configs.SplitMix64(99)
    u = a.uniform_array(5)
    ref = [(b.next_u64() >> 11) / 2.0 ** 53 for _ in range(5)]
    np.testing.assert_array_equal(u, ref)
    assert a.next_u64() == b.next_u64()


def test_geometry_matches_oracle_restatement():
    w = G.wavelet_derivative(G.ricker(20.0, 0.001, 50), 0.001)
    l1, t1 = G.source_points(40, 44, 20.3, 21.7, 9.2, w, 60)
    l2, t2 = O.source_geometry(40, 44, 9.2, 20.3, 21.7, w, 60)
    np.testing.assert_array_equal(l1, l2)
    np.testing.assert_allclose(t1, t2, rtol=1e-6, atol=1e-6)
    sx, sy, sz = np.array([3.2, 7.9]), np.array([4.4, 1.0]), np.array([6.5, 6.5])
    np.testing.assert_array_equal(G.receiver_locs(40, 44, sx, sy, sz),
                                  O.receiver_index(40, 44, sx, sy, sz))
    np.testing.assert_allclose(G.wavelet_derivative(w, 0.001), O.wavelet_derivative(w, 0.001))


def test_compute_size_matches_reference_padding():
    """Born/source_func_3d.cpp:37-63: pad 2 (nb + 4) per axis, axis 3 gets
    nb + nb + 8 more, each rounded up to a multiple of 16."""
    n, nbound = G.compute_size((100, 90, 80), 50, extra_z=True)
    assert nbound == 54
    assert n == (208, 208, 304)  # 208 = ceil16(208), 198 -> 208, 80+108+108=296 -> 304


def test_workloads_are_deterministic_and_stable():
    a = configs.rtm128(seed=3, nt=50)
    b = configs.rtm128(seed=3, nt=50)
    np.testing.assert_array_equal(a.vel, b.vel)
    np.testing.assert_array_equal(a.refl, b.refl)
    assert G.cfl_number(a.vel.max(), a.dt, a.d) < 0.4529
    s = a.shots[0]
    assert len(s.rec_locs) == 128 * 128 and len(s.src_locs) == 9
    c = configs.rtm128(seed=4, nt=50)
    assert not np.array_equal(a.vel, c.vel)


def test_library_sinc_table_matches_oracle():
    """The library's host-side table (brtm_sinc_table) against the oracle's
    FP64 restatement of Born/sinc_bob.cpp."""
    from sep_code_amd.operator import sinc_table
    for ntab in (1, 3, 4):
        np.testing.assert_allclose(sinc_table(ntab), O.sinc_table(ntab), atol=1e-6)
    # and against the reference's own sinc_bob output (tests/golden/ref_outputs.json)
    import json
    with open(os.path.join(ROOT, "tests", "golden", "ref_outputs.json")) as f:
        ref = json.load(f)["sinc"]
    for ntab, t in ref.items():
        np.testing.assert_allclose(sinc_table(int(ntab)), np.array(t), atol=1e-4)


def test_lsrtm_workload_shape():
    w = configs.lsrtm(n_side=64, nt=100)
    assert len(w.shots) == 8 and w.vel.shape == (64, 64, 64)
    assert G.cfl_number(w.vel.max(), w.dt, w.d) < 0.4529
    assert np.count_nonzero(w.refl) > 0


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU; --gpus N > 1 exits with status 2 and says why,
    before touching any device."""
    import json
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert msg["n_gpus"] == 2 and "one GPU" in msg["error"]


def test_bench_names_the_kernel_the_library_runs():
    """bench.py's roofline label mirrors the library's automatic kernel
    choice: four cells per lane with 16-row tiles at 512^3 and 256^3, the
    1-cell kernel with 8-row tiles where the automatic shape would leave
    MI355X's 256 compute units without a workgroup (128^3), and the forced
    cells per lane where n1 allows them."""
    import sys
    sys.path.insert(0, ROOT)
    import bench
    assert bench.plain_step_kernel((512, 512, 512), 0) == (4, 16)
    assert bench.plain_step_kernel((256, 256, 256), 0) == (4, 16)
    assert bench.plain_step_kernel((128, 128, 128), 0) == (1, 8)
    assert bench.plain_step_kernel((296, 296, 300), 0) == (2, 16)
    assert bench.plain_step_kernel((512, 512, 512), 2) == (2, 16)
    assert bench.cells_per_lane(70, 4) == 2 and bench.cells_per_lane(71, 4) == 1


def test_bench_cpu_full_shot_matches_fp64_oracle():
    """The whole-shot CPU baseline (bench.cpu_full_shot, the FP32 OpenMP
    restatement with denormals flushed) images the same shot as the FP64
    oracle to 1e-5, for both of the workload's source stores."""
    import sys
    sys.path.insert(0, ROOT)
    import bench
    w = configs.rtm128(nt=40)
    sh = w.shots[0]
    D = np.random.default_rng(5).standard_normal((len(sh.rec_locs), w.ntr)).astype(np.float32)
    P = O.Problem(w.n, w.d, w.V, w.bc_a, w.bc_b, src_locs=sh.src_locs, src_tr=sh.src_traces,
                  rec_locs=sh.rec_locs, ntr_d=w.ntr, jt=w.jt, nt=w.nt, jimg=w.jimg, cimg=w.cimg)
    ref = P.rtm_adjoint(D)
    assert np.abs(ref).max() > 0
    for store in ("f32", "ckpt"):
        w.store, w.ckpt = store, 16
        r = bench.cpu_full_shot(w, D, ref, runs=1)
        assert r["image_rel_diff_vs_gpu"] < 1e-5, (store, r)
        assert r["value"] > 0 and r["kind"] == "port"


def test_drop_in_header_has_the_reference_signatures():
    """include/gpu_funcs_3d.h against the reference's own header
    (Born/gpu_funcs_3d.h:5-27): oracle/Makefile's abi target compiles both in
    one C++ translation unit, where a differing parameter type is a
    conflicting extern "C" declaration."""
    import subprocess
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "abi"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    if "reference absent" in r.stdout:
        pytest.skip("the reference's sources are not in this container")
    assert "drop-in signatures agree" in r.stdout


def test_workloads_match_their_fingerprints():
    """The synthetic workloads the measurements are quoted on (128^3 and
    512^3 RTM shots, the 256^3 LSRTM problem for seeds 77 and 78) are the
    ones tests/golden/workloads.json recorded: sizes and source/receiver
    cells exactly, velocity, reflectivity and source traces to 1e-6."""
    import importlib.util
    import json
    path = os.path.join(ROOT, "tests", "golden", "make_workload_fixtures.py")
    spec = importlib.util.spec_from_file_location("make_workload_fixtures", path)
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    with open(os.path.join(ROOT, "tests", "golden", "workloads.json")) as f:
        want = json.load(f)
    assert set(want) == set(gen.WORKLOADS)

    def same(a, b, where):
        if isinstance(a, dict):
            assert set(a) == set(b), where
            for k in a:
                same(a[k], b[k], f"{where}.{k}")
        elif isinstance(a, list):
            assert len(a) == len(b), where
            for i, (x, y) in enumerate(zip(a, b)):
                same(x, y, f"{where}[{i}]")
        elif isinstance(a, float):
            assert abs(a - b) <= 1e-6 * max(abs(a), abs(b)) + 1e-30, (where, a, b)
        else:
            assert a == b, (where, a, b)

    for name, make in gen.WORKLOADS.items():
        same(json.loads(json.dumps(gen.fingerprint(make()))), want[name], name)


def test_gram_solve_drops_dependent_directions(tmp_path):
    """The flexible CG's host-side Gram solve (sep_code_amd/csrc/gram_solve.h,
    built here with g++ on its own): on a full-rank K x K Gram matrix
    (K up to BRTM_CG_MAX_HISTORY + 1) it solves M x = b; a direction that
    depends on earlier ones (exactly, or to a pivot below tol) gets x = 0 and
    the rest is the least-squares solution over the others; a direction of
    zero length is dropped the same way, and with none left nothing is kept."""
    import subprocess
    src = tmp_path / "gram.cpp"
    src.write_text('#include "gram_solve.h"\n'
                   'extern "C" int gram(int K, const double *M, const double *b, double *x,'
                   ' double tol) { return brtm::gram_solve(K, M, b, x, tol); }\n')
    so = tmp_path / "libgram.so"
    subprocess.run(["g++", "-std=c++17", "-O2", "-shared", "-fPIC",
                    "-I", os.path.join(ROOT, "sep_code_amd", "csrc"), "-o", str(so), str(src)],
                   check=True)
    lib = ctypes.CDLL(str(so))
    P = ctypes.POINTER(ctypes.c_double)
    lib.gram.argtypes = [ctypes.c_int, P, P, P, ctypes.c_double]

    def solve(B, t, tol=1e-9):
        M = np.ascontiguousarray(B.T @ B)
        b = np.ascontiguousarray(B.T @ t)
        x = np.zeros(B.shape[1])
        kept = lib.gram(B.shape[1], M.ctypes.data_as(P), b.ctypes.data_as(P),
                        x.ctypes.data_as(P), tol)
        return kept, x

    r = np.random.default_rng(40)
    for K in (1, 2, 5, 9, 17):
        B, t = r.standard_normal((60, K)), r.standard_normal(60)
        kept, x = solve(B, t)
        assert kept == K
        np.testing.assert_allclose(x, np.linalg.solve(B.T @ B, B.T @ t), rtol=1e-9, atol=1e-12)
    # column 2 = column 0 - 2 column 1: dropped, the others solve least squares
    B, t = r.standard_normal((40, 6)), r.standard_normal(40)
    B[:, 2] = B[:, 0] - 2 * B[:, 1]
    kept, x = solve(B, t)
    keep = [0, 1, 3, 4, 5]
    assert kept == 5 and x[2] == 0
    np.testing.assert_allclose(x[keep], np.linalg.lstsq(B[:, keep], t, rcond=None)[0],
                               rtol=1e-8, atol=1e-10)
    # nearly dependent (relative pivot 1e-12 < tol): dropped; with tol below it: kept
    B[:, 2] += 1e-6 * r.standard_normal(40) * np.linalg.norm(B[:, 2]) / np.sqrt(40)
    assert solve(B, t)[0] == 5
    assert solve(B, t, tol=1e-14)[0] == 6
    # a direction of zero length is dropped; with every one of zero length
    # nothing is kept
    B = r.standard_normal((40, 4))
    B[:, 0] = 0
    kept, x = solve(B, t)
    assert kept == 3 and x[0] == 0
    np.testing.assert_allclose(x[1:], np.linalg.lstsq(B[:, 1:], t, rcond=None)[0],
                               rtol=1e-8, atol=1e-10)
    assert solve(np.zeros((40, 3)), t)[0] == 0


def test_tmute_starts_follow_the_reference_float_arithmetic(tmp_path):
    """The library's first kept sample of each receiver column of tmute
    (sep_code_amd/csrc/tmute_start.h, built here with g++ on its own) equals
    the oracle's restatement of Born/tmute.cpp:28-29,63-64 on random
    geometries and on a case where FP32 and FP64 truncate to different
    samples (30 m at 1500 m/s with 4 ms samples: 4, not 5)."""
    import subprocess
    src = tmp_path / "tm.cpp"
    src.write_text('#include "tmute_start.h"\n'
                   'extern "C" void starts(int nt, float ot, float dt, int nrx, float orx,'
                   ' float drx, float t0, float vmute, int *s) {'
                   ' brtm::tmute_starts(nt, ot, dt, nrx, orx, drx, t0, vmute, s); }\n')
    so = tmp_path / "libtm.so"
    subprocess.run(["g++", "-std=c++17", "-O2", "-shared", "-fPIC",
                    "-I", os.path.join(ROOT, "sep_code_amd", "csrc"), "-o", str(so), str(src)],
                   check=True)
    lib = ctypes.CDLL(str(so))
    F, I = ctypes.c_float, ctypes.c_int
    lib.starts.argtypes = [I, F, F, I, F, F, F, F, ctypes.POINTER(I)]

    def starts(nt, ot, dt, nrx, orx, drx, t0, vmute):
        s = (I * nrx)()
        lib.starts(nt, ot, dt, nrx, orx, drx, t0, vmute, s)
        return list(s)

    cases = [(100, 0.0, 0.004, 4, 0.0, 10.0, 0.0, 1500.0),
             (100, 0.0, 0.004, 3, 100.0, 50.0, 0.1, 2000.0),
             (10, 0.5, 0.004, 2, 0.0, 10.0, 0.1, 2000.0),
             (10, 0.0, 0.004, 1, 0.0, 10.0, 5.0, 2000.0)]
    r = np.random.default_rng(12)
    for _ in range(300):
        cases.append((int(r.integers(1, 3000)), float(r.uniform(-0.1, 0.2)),
                      float(r.choice([0.001, 0.002, 0.004, 0.008])), int(r.integers(1, 40)),
                      float(r.uniform(-500, 500)), float(r.choice([5.0, 10.0, 12.5, 25.0, 50.0])),
                      float(r.uniform(0, 0.5)), float(r.uniform(1400, 5000))))
    assert starts(*cases[0]) == [0, 1, 3, 4]
    for c in cases:
        assert starts(*c) == [int(v) for v in O.tmute_start(*c)], c


def test_step_tables_equal_the_oracle_bit_for_bit(tmp_path):
    """The library's stencil weights and sponge weights
    (sep_code_amd/csrc/step_tables.h, built here with g++ on its own) equal
    the oracle's FP32 tables (Born/wave_fkernel.3d8o.cu:26-27, 201-202) bit
    for bit, over random spacings and sponge parameters."""
    import subprocess
    src = tmp_path / "tab.cpp"
    src.write_text('#include "step_tables.h"\n'
                   'extern "C" void coeffs(float d1, float d2, float d3, float *c) {'
                   ' brtm::stencil_coeffs(d1, d2, d3, c); }\n'
                   'extern "C" int sponge(float a, float b, float *w, int cap) {'
                   ' int nw; std::vector<float> t; brtm::sponge_weights(a, b, &nw, &t);'
                   ' for (int i = 0; i < (int)t.size() && i < cap; ++i) w[i] = t[i];'
                   ' return (int)t.size() == std::max(nw, 1) ? nw : -1; }\n')
    so = tmp_path / "libtab.so"
    subprocess.run(["g++", "-std=c++17", "-O2", "-shared", "-fPIC",
                    "-I", os.path.join(ROOT, "sep_code_amd", "csrc"), "-o", str(so), str(src)],
                   check=True)
    lib = ctypes.CDLL(str(so))
    F, P = ctypes.c_float, ctypes.c_void_p
    lib.coeffs.argtypes = [F, F, F, P]
    lib.sponge.argtypes = [F, F, P, ctypes.c_int]
    lib.sponge.restype = ctypes.c_int
    r = np.random.default_rng(21)
    for _ in range(200):
        d = [float(np.float32(x)) for x in r.uniform(1.0, 50.0, 3)]
        c = np.zeros(13, np.float32)
        lib.coeffs(*d, c.ctypes.data)
        np.testing.assert_array_equal(c, O.coeffs(*d, prec=32))
        a = float(np.float32(r.choice([r.uniform(0, 256), float(r.integers(0, 257))])))
        b = float(np.float32(r.uniform(0, 0.1)))
        w = np.zeros(258, np.float32)
        nw = lib.sponge(a, b, w.ctypes.data, len(w))
        ref = O.sponge_table(a, b, prec=32)
        assert nw == len(ref), (a, nw, len(ref))
        np.testing.assert_array_equal(w[:nw], ref)
        assert (w[nw:max(nw, 1)] == 1).all()
