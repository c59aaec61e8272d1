"""The oracle's pins (tests/test_oracle.py) catch plausible mistakes: each
mistake is planted in a copy of the oracle, the pinning tests run against
the copy must fail, and they pass on the oracle itself.  CPU only.  Four
groups:

* time indexing (closed-form recursions) and
* stencil and sponge (Fornberg weights, polynomial and plane-wave
  exactness, the sponge formulas, the VL/LV transpose, the stability limit,
  energy conservation): planted in copies of oracle/born_oracle.c, each built
  into its own FP64 library in a temporary directory;
* the numpy operators (D_t, D_z, the data window, the wavelet derivative,
  the receiver mask, tmute's first kept sample): mistaken versions swapped
  into the module, against their closed forms and hand examples;
* the solver forms (lin_solve) wrapped with assembly mistakes, against their
  closed forms.
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import oracle.born_oracle as O
import test_oracle as T

SRC_PATH = os.path.join(O.HERE, "born_oracle.c")

# (what the mistake is, the oracle's text, the text with the mistake), after
# Born/gpu_funcs_3d.cu:228-229 and 247-250 (sample index, step then inject),
# :398-462 (Born source from s^{it+1}, traces extracted at the step's own
# sample) and :600-667 (data injected before imaging)
MUTANTS = [
    ("source sample index one step late",
     "static void source_step(const orc_prob *P, int it, const real *cur, real *old_new) {\n"
     "  int id = it / P->jt, ii = it % P->jt;",
     "static void source_step(const orc_prob *P, int it, const real *cur, real *old_new) {\n"
     "  int id = (it > 0 ? it - 1 : 0) / P->jt, ii = (it > 0 ? it - 1 : 0) % P->jt;"),
    ("source injected before the step",
     "  FN(orc_step_vl)(P, cur, old_new, old_new);\n"
     "  FN(orc_inject)(P->nsrc, P->slocs, P->strc, P->ntr_s, P->sinc_s + 8 * ii, id,\n"
     "                 (real)1, old_new);",
     "  FN(orc_inject)(P->nsrc, P->slocs, P->strc, P->ntr_s, P->sinc_s + 8 * ii, id,\n"
     "                 (real)1, old_new);\n"
     "  FN(orc_step_vl)(P, cur, old_new, old_new);"),
    ("Born source from s^it instead of s^{it+1}",
     "    if (!snaps) { source_step(P, it, sb, sa); s = sa; }",
     "    if (!snaps) { s = sb; source_step(P, it, sb, sa); }"),
    ("traces extracted at the next step's sample and row",
     "    FN(orc_extract)(P->nrec, P->rlocs, D, P->ntr_d, P->sinc_d + 8 * ii, id, qa);",
     "    FN(orc_extract)(P->nrec, P->rlocs, D, P->ntr_d, P->sinc_d + 8 * ((it + 1) % P->jt),"
     " (it + 1) / P->jt, qa);"),
    ("imaging before the data injection",
     "    FN(orc_inject)(P->nrec, P->rlocs, D, P->ntr_d, P->sinc_d + 8 * ii, id,\n"
     "                   (real)1, *la);\n"
     "    if (it % P->jimg == 0) FN(orc_xcorr)(n, (real)P->cimg, segp[it - k0], *la, img);",
     "    if (it % P->jimg == 0) FN(orc_xcorr)(n, (real)P->cimg, segp[it - k0], *la, img);\n"
     "    FN(orc_inject)(P->nrec, P->rlocs, D, P->ntr_d, P->sinc_d + 8 * ii, id,\n"
     "                   (real)1, *la);"),
]

# after Born/wave_fkernel.3d8o.cu:31-106 (the VL update and its taps),
# :183-210 (the sponge: w on the update, w^2 on p_old; the edge distance)
STENCIL_MUTANTS = [
    ("y taps with the x spacing",
     "    c[2 + 3 * k] = (real)(C2 * f[k]);",
     "    c[2 + 3 * k] = (real)(C1 * f[k]);"),
    ("fourth-distance weight with the wrong sign",
     "  static const double f[4] = {1.0, -1.0 / 8.0, 1.0 / 63.0, -1.0 / 896.0};",
     "  static const double f[4] = {1.0, -1.0 / 8.0, 1.0 / 63.0, 1.0 / 896.0};"),
    ("p_old damped by w instead of w^2",
     "        pn[i] = w * (P->V[i] * l + 2 * p[i]) - w * w * po[i];",
     "        pn[i] = w * (P->V[i] * l + 2 * p[i]) - w * po[i];"),
    ("symmetric edge distance",
     "  int e = imin(i1 - RAD, n1 - RAD - i1);",
     "  int e = imin(i1 - RAD, n1 - RAD - 1 - i1);"),
]

STENCIL_PINS = [(name, [()]) for name in (
    "test_coeffs_are_fornberg_weights", "test_1d_weights_exact_on_polynomials",
    "test_stencil_laplacian_of_polynomial_anisotropic", "test_stencil_symbol_on_plane_wave",
    "test_sponge_table_closed_form", "test_sponge_edge_distance_matches_reference_formula",
    "test_lv_step_is_transpose_of_vl_step", "test_undamped_leapfrog_is_time_reversible",
    "test_sponge_dissipates_energy", "test_leapfrog_stability_limit_from_the_weights",
    "test_undamped_leapfrog_conserves_discrete_energy")]

CLOSED_FORM = [
    ("test_source_prop_one_sample_delay_jt1", [()]),
    ("test_source_prop_closed_form_two_steps_per_sample", [(2, 2), (2, 3)]),
    ("test_born_forward_closed_form", [(1, 1, False), (1, 1, True), (2, 2, False), (2, 3, True)]),
    ("test_rtm_adjoint_closed_form", [(1, 1), (2, 2), (2, 3)]),
]


def _failing(tests):
    out = []
    for name, cases in tests:
        for args in cases:
            try:
                getattr(T, name)(*args)
            except AssertionError:
                out.append(f"{name}{args}")
    return out


@pytest.fixture(scope="module")
def builds():
    if shutil.which("gcc") is None:
        pytest.skip("gcc is needed to build the planted copies")
    d = tempfile.mkdtemp(prefix="orc_planted_")
    yield d
    shutil.rmtree(d, ignore_errors=True)


def _use(lib_dir):
    """Load lib_dir/liborc64.so through the oracle's own loader."""
    here = O.HERE
    O.HERE = lib_dir
    O._LIBS.pop(64, None)
    try:
        O.lib(64)
    finally:
        O.HERE = here


def test_pinning_tests_pass_on_the_oracle():
    assert _failing(CLOSED_FORM) == []
    assert _failing(STENCIL_PINS) == []


def _planted(builds, tag, label, orig, planted, tests):
    src = open(SRC_PATH).read()
    assert src.count(orig) == 1, f"the oracle text for '{label}' has changed"
    d = os.path.join(builds, tag)
    os.makedirs(d)
    c_path = os.path.join(d, "born_oracle.c")
    with open(c_path, "w") as f:
        f.write(src.replace(orig, planted))
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-fopenmp", "-fno-fast-math",
                    "-ffp-contract=off", "-DORACLE_REAL=double", "-DORACLE_SUFFIX=_f64", "-o",
                    os.path.join(d, "liborc64.so"), c_path, "-lm"], check=True)
    saved = O._LIBS.get(64)
    try:
        _use(d)
        return _failing(tests)
    finally:
        O._LIBS.pop(64, None)
        if saved is not None:
            O._LIBS[64] = saved


@pytest.mark.parametrize("label,orig,planted", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_closed_form_tests_catch_a_planted_mistake(builds, label, orig, planted):
    tag = f"t{[m[0] for m in MUTANTS].index(label)}"
    failing = _planted(builds, tag, label, orig, planted, CLOSED_FORM)
    assert failing, f"no closed-form test caught: {label}"


@pytest.mark.parametrize("label,orig,planted", STENCIL_MUTANTS,
                         ids=[m[0] for m in STENCIL_MUTANTS])
def test_stencil_pins_catch_a_planted_mistake(builds, label, orig, planted):
    tag = f"s{[m[0] for m in STENCIL_MUTANTS].index(label)}"
    failing = _planted(builds, tag, label, orig, planted, STENCIL_PINS)
    assert failing, f"no stencil or sponge pin caught: {label}"


# Mistakes in the oracle's numpy operators, each made consistently in the
# forward and the adjoint (so the adjoint-pair tests still pass): after
# Born/deriv_3d.cpp:4-46, Born/laplac_3d.cpp:4-65, Born/map_data_3d.cpp:44-76,
# Born/source_func_3d.cpp:9-13 and the update set [4, n-4).
def _dt_fwd_diff(m, dt):
    m = np.asarray(m, np.float64)
    d = np.zeros_like(m)
    d[..., :-1] = (m[..., :-1] - m[..., 1:]) / dt
    return d


def _dt_fwd_diff_adj(d, dt):
    d = np.asarray(d, np.float64)
    m = np.zeros_like(d)
    m[..., :-1] += d[..., :-1] / dt
    m[..., 1:] -= d[..., :-1] / dt
    return m


def _dz_down(m):
    m = np.asarray(m, np.float64)
    d = np.zeros_like(m)
    d[:-1] = m[:-1] - m[1:]
    return d


def _dz_down_adj(d):
    d = np.asarray(d, np.float64)
    m = np.zeros_like(d)
    m[:-1] += d[:-1]
    m[1:] -= d[:-1]
    return m


def _window3(m, nt):
    return np.array(np.asarray(m)[..., 3:3 + nt], np.float64)


def _window3_adj(d, ntb):
    d = np.asarray(d, np.float64)
    m = np.zeros(d.shape[:-1] + (ntb,), np.float64)
    m[..., 3:3 + d.shape[-1]] = d
    return m


def _wavelet_backward(w, dt):
    w = np.asarray(w, np.float64)
    out = np.zeros_like(w)
    out[1:] = (w[1:] - w[:-1]) / dt
    return out


def _mask(lo, hi):
    def f(locs, n):
        n1, n2, n3 = n
        rl = np.asarray(locs, np.int64).copy()
        i1, i2, i3 = rl % n1, (rl // n1) % n2, rl // (n1 * n2)
        ok = ((i1 >= lo) & (i1 < n1 - hi) & (i2 >= lo) & (i2 < n2 - hi) & (i3 >= lo)
              & (i3 < n3 - hi))
        rl[~ok] = -1
        return rl.astype(np.int32)
    return f


def _mask_xy_swapped(locs, n):
    n1, n2, n3 = n
    rl = np.asarray(locs, np.int64).copy()
    i1, i2, i3 = rl % n2, (rl // n2) % n1, rl // (n1 * n2)
    ok = (i1 >= 4) & (i1 < n1 - 4) & (i2 >= 4) & (i2 < n2 - 4) & (i3 >= 4) & (i3 < n3 - 4)
    rl[~ok] = -1
    return rl.astype(np.int32)


def _tmute(conv):
    def f(nt, ot, dt, nrx, orx, drx, t0, vmute):
        out = np.zeros(nrx, np.int64)
        for i in range(nrx):
            out[i] = conv(nt, ot, dt, orx, drx, t0, vmute, i)
        return out
    return f


def _tmute_f64(nt, ot, dt, orx, drx, t0, vmute, i):
    q = (t0 + (orx + drx * i) / vmute - ot) / dt
    return 0 if q <= 0 else (nt if q >= nt else int(q))


def _tmute_f32(nt, ot, dt, orx, drx, t0, vmute, i):
    f32 = np.float32
    return float((f32(t0) + (f32(orx) + f32(drx) * f32(i)) / f32(vmute) - f32(ot)) / f32(dt))


def _tmute_rounded(nt, ot, dt, orx, drx, t0, vmute, i):
    q = _tmute_f32(nt, ot, dt, orx, drx, t0, vmute, i)
    return 0 if q <= 0 else (nt if q >= nt else int(round(q)))


def _tmute_unclamped(nt, ot, dt, orx, drx, t0, vmute, i):
    q = _tmute_f32(nt, ot, dt, orx, drx, t0, vmute, i)
    return nt if q >= nt else int(q)


PY_MUTANTS = [
    ("D_t with the opposite sign",
     {"dt_forward": lambda m, dt: -_ORIG["dt_forward"](m, dt),
      "dt_adjoint": lambda d, dt: -_ORIG["dt_adjoint"](d, dt)}),
    ("D_t as the difference toward the next sample",
     {"dt_forward": _dt_fwd_diff, "dt_adjoint": _dt_fwd_diff_adj}),
    ("D_z as the difference toward the next slice",
     {"dz_forward": _dz_down, "dz_adjoint": _dz_down_adj}),
    ("data window dropping 3 samples", {"window_forward": _window3,
                                        "window_adjoint": _window3_adj}),
    ("wavelet derivative as a backward difference", {"wavelet_derivative": _wavelet_backward}),
    ("receivers on the last halo cell kept", {"mask_inactive": _mask(4, 3)}),
    ("receivers on the first updated cell dropped", {"mask_inactive": _mask(5, 4)}),
    ("receiver mask with x and y swapped", {"mask_inactive": _mask_xy_swapped}),
    ("tmute start computed in FP64", {"tmute_start": _tmute(_tmute_f64)}),
    ("tmute start rounded, not truncated", {"tmute_start": _tmute(_tmute_rounded)}),
    ("tmute start not clamped at 0", {"tmute_start": _tmute(_tmute_unclamped)}),
]
_ORIG = {name: getattr(O, name) for m in PY_MUTANTS for name in m[1]}

PY_PINS = [(name, [()]) for name in (
    "test_dz_dt_window_are_adjoint_pairs", "test_dz_and_dt_closed_forms",
    "test_window_hand_example", "test_wavelet_derivative_hand_example",
    "test_inactive_receivers_hand_example", "test_tmute_hand_example")]


def test_numpy_pins_pass_on_the_oracle():
    assert _failing(PY_PINS) == []


@pytest.mark.parametrize("label,patch", PY_MUTANTS, ids=[m[0] for m in PY_MUTANTS])
def test_numpy_pins_catch_a_planted_mistake(label, patch):
    try:
        for name, fn in patch.items():
            setattr(O, name, fn)
        failing = _failing(PY_PINS)
    finally:
        for name in patch:
            setattr(O, name, _ORIG[name])
    assert failing, f"no pin caught: {label}"
    # the adjoint pair still holds for every mutated pair, so it is the
    # closed forms and hand examples that catch them
    assert "test_dz_dt_window_are_adjoint_pairs()" not in failing


# Mistakes in how the three solver forms are assembled (Born/lin_solver.cpp:
# 7-53, 79-101), planted around the oracle's lin_solve by changing what it is
# given or what it returns.
_LIN = O.lin_solve


def _lin(kind, L, d, niter, aux=None, eps=0.0, W=None, m0=None, history=1):
    return _LIN(kind, L, d, niter, aux=aux, eps=eps, W=W, m0=m0, history=history)


def _lin_no_eps(kind, L, d, niter, aux=None, eps=0.0, W=None, m0=None, history=1):
    return _lin(kind, L, d, niter, aux, 1.0 if (kind == "reg" and eps) else eps, W, m0, history)


def _lin_r_transposed(kind, L, d, niter, aux=None, eps=0.0, W=None, m0=None, history=1):
    if kind == "reg" and aux is not None:
        aux = np.asarray(aux).T
    return _lin(kind, L, d, niter, aux, eps, W, m0, history)


def _lin_model_for_update(kind, L, d, niter, aux=None, eps=0.0, W=None, m0=None, history=1):
    m, res = _lin(kind, L, d, niter, aux, eps, W, m0, history)
    return (m + m0 if m0 is not None else m), res


def _lin_weight_twice(kind, L, d, niter, aux=None, eps=0.0, W=None, m0=None, history=1):
    W = np.asarray(W) @ np.asarray(W) if W is not None else None
    return _lin(kind, L, d, niter, aux, eps, W, m0, history)


def _lin_weight_transposed(kind, L, d, niter, aux=None, eps=0.0, W=None, m0=None, history=1):
    W = np.asarray(W).T if W is not None else None
    return _lin(kind, L, d, niter, aux, eps, W, m0, history)


def _lin_prec_eps_on_m(kind, L, d, niter, aux=None, eps=0.0, W=None, m0=None, history=1):
    if kind != "prec":
        return _lin(kind, L, d, niter, aux, eps, W, m0, history)
    P = np.asarray(aux)
    # min |W(L P p - d)|^2 + eps^2 |P p|^2: the damping on m instead of p
    p, res = _lin("reg", np.asarray(L) @ P, d, niter, P, eps, W, None, history)
    return P @ p, res


LIN_MUTANTS = [
    ("reg_solver without its eps", _lin_no_eps),
    ("reg_solver with the regulariser transposed", _lin_r_transposed),
    ("the model returned in place of the update from m0", _lin_model_for_update),
    ("the data weight applied twice", _lin_weight_twice),
    ("the data weight transposed", _lin_weight_transposed),
    ("prec_solver damping m = P p instead of p", _lin_prec_eps_on_m),
]
LIN_PINS = [("test_lin_solve_forms_reach_their_closed_forms", [()])]


def test_solver_pins_pass_on_the_oracle():
    O.lin_solve = _lin
    try:
        assert _failing(LIN_PINS) == []
    finally:
        O.lin_solve = _LIN


@pytest.mark.parametrize("label,fn", LIN_MUTANTS, ids=[m[0] for m in LIN_MUTANTS])
def test_solver_pins_catch_a_planted_mistake(label, fn):
    O.lin_solve = fn
    try:
        failing = _failing(LIN_PINS)
    finally:
        O.lin_solve = _LIN
    assert failing, f"no closed form caught: {label}"
