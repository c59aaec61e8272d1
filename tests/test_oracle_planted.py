"""The oracle's pins (tests/test_oracle.py) catch the plausible mistakes:
each is planted in a copy of the oracle's C source (oracle/born_oracle.c),
built into a separate FP64 library, and the pinning tests run against it
must fail, while they pass on the oracle itself.  Two groups: the time
indexing (closed-form recursions) and the stencil and sponge (Fornberg
weights, polynomial and plane-wave exactness, the sponge formulas, the VL/LV
transpose).  CPU only; the copies are built in a temporary directory."""
import os
import shutil
import subprocess
import tempfile

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
    "test_sponge_dissipates_energy")]

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
