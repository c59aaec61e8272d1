"""The closed-form time-recursion tests (tests/test_oracle.py) catch the
plausible indexing mistakes: each is planted in a copy of the oracle's C
source (oracle/born_oracle.c), built into a separate FP64 library, and the
closed-form tests run against it must fail, while they pass on the oracle
itself.  CPU only; the copies are built in a temporary directory."""
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

CLOSED_FORM = [
    ("test_source_prop_one_sample_delay_jt1", [()]),
    ("test_source_prop_closed_form_two_steps_per_sample", [(2, 2), (2, 3)]),
    ("test_born_forward_closed_form", [(1, 1, False), (1, 1, True), (2, 2, False), (2, 3, True)]),
    ("test_rtm_adjoint_closed_form", [(1, 1), (2, 2), (2, 3)]),
]


def _failing():
    out = []
    for name, cases in CLOSED_FORM:
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


def test_closed_form_tests_pass_on_the_oracle():
    assert _failing() == []


@pytest.mark.parametrize("label,orig,planted", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_closed_form_tests_catch_a_planted_mistake(builds, label, orig, planted):
    src = open(SRC_PATH).read()
    assert src.count(orig) == 1, f"the oracle text for '{label}' has changed"
    d = os.path.join(builds, str(MUTANTS.index((label, orig, planted))))
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
        failing = _failing()
    finally:
        O._LIBS.pop(64, None)
        if saved is not None:
            O._LIBS[64] = saved
    assert failing, f"no closed-form test caught: {label}"
