"""
The LM branch pins (test_lm_branches.py) bite: each plausible misreading of
the reference's state machine, planted in a copy of the oracle's C source
(compiled here with gcc, on the CPU), makes the pin written for it fail,
while the unmodified oracle passes them all.  The oracle under test/ is not
modified; the copies live in pytest's temporary directory.
"""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(os.path.dirname(HERE), "oracle", "dao_oracle.c")

# (name, original text, planted text, the pin that must fail)
MUTATIONS = [
    ("lambda_down_half", "#define LM_DOWN 0.75", "#define LM_DOWN 0.5",
     "test_accepted_decrease_then_lambda_three_quarters"),
    ("lambda_up_ten", "#define LM_UP 4.0", "#define LM_UP 10.0",
     "test_singular_system_climbs_lambda_until_lost"),
    ("lambda_max_1e19", "#define LM_MAX 1.0e+20", "#define LM_MAX 1.0e+19",
     "test_singular_system_climbs_lambda_until_lost"),
    ("reset_keeps_lambda", "  w->lambda = lam * LM_UP;\n}", "  w->lambda = lam;\n}",
     "test_rejected_increase_retries_with_lambda_four"),
    ("increase_converges_at_start_params",
     "        if (((w->error - start) / start) < fd->tolerance) {\n          w->status = ST_CONVERGED;",
     "        if (((w->error - start) / start) < fd->tolerance) {\n          peak_stamp(fd, -1);\n"
     "          peak_reset(fd, i);\n          peak_shape(fd);\n          peak_stamp(fd, +1);\n"
     "          w->status = ST_CONVERGED;",
     "test_accepted_increase_below_tol_keeps_the_new_parameters"),
    ("increase_tol_over_err", "((w->error - start) / start)", "((w->error - start) / w->error)",
     "test_increase_tolerance_is_relative_to_the_start_error"),
    ("decrease_tol_over_err", "((start - w->error) / start)", "((start - w->error) / w->error)",
     "test_decrease_below_tol_converges_relative_to_the_start_error"),
    ("margin_exclusive", "(p->xi >= (fd->image_size_x - fd->fit_size_x))",
     "(p->xi > (fd->image_size_x - fd->fit_size_x))", "test_step_onto_the_far_margin_is_rejected"),
    ("recenter_half_band", "if ((dx < -0.25) || (dx > 1.25) || (dy < -0.25) || (dy > 1.25)) {",
     "if ((dx < -0.5) || (dx > 1.5) || (dy < -0.5) || (dy > 1.5)) {",
     "test_recentering_after_the_sweep_uses_a_quarter_pixel_band"),
    ("recenter_ignores_margin", "if ((dx < -0.25) || (dx > 1.25) || (dy < -0.25) || (dy > 1.25)) {",
     "if (((dx < -0.25) || (dx > 1.25) || (dy < -0.25) || (dy > 1.25)) && floor(p->params[P_XC]) >= 0) {",
     "test_recentering_off_the_image_loses_the_peak"),
    ("cycle_moves_late", "if ((dx < -0.5) || (dx > 1.5)) p->xi", "if ((dx < -1.0) || (dx > 2.0)) p->xi",
     "test_roi_moves_inside_the_cycle_past_one_and_a_half"),
    ("background_sum", "fd->f_data[k] + fd->bg_data[k] / ((double)fd->bg_counts[k])", "fd->f_data[k] + fd->bg_data[k]",
     "test_overlapping_peaks_see_the_mean_background"),
    ("clamp_both_widths_2d", "  if (m->shape == SH_3D) {\n    if (p->params[P_YW] < m->width_min)",
     "  if (m->shape == SH_3D || m->shape == SH_2D) {\n    if (p->params[P_YW] < m->width_min)",
     "test_2d_width_clamp_leaves_the_y_width_unclamped"),
    ("reset_keeps_cycle_count", "  int added = w->added;\n  double lam = w->lambda;\n  peak_copy(fd, &fd->fit[i], w);",
     "  int added = w->added;\n  double lam = w->lambda;\n  int it = w->iterations;\n  peak_copy(fd, &fd->fit[i], w);\n"
     "  w->iterations = it;", "test_negative_height_is_rejected"),
    ("reverse_update_order", "void mFitIterateLM(fitData *fd)\n{", "void mFitIterateLM(fitData *fd)\n{",
     "test_overlapping_peaks_update_in_index_order"),
]


def _build(tmp, name, text):
    src = os.path.join(tmp, name + ".c")
    with open(src, "w") as fh:
        fh.write(text)
    so = os.path.join(tmp, "lib" + name + ".so")
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-w", "-o", so, src, "-lm"], check=True)
    return so


def _pins(so, select=None):
    """Exit status of the pins against the oracle build `so` (None: it hung -- a failure too)."""
    env = dict(os.environ, DAO_ORACLE_SO=so)
    cmd = [sys.executable, "-m", "pytest", os.path.join(HERE, "test_lm_branches.py"), "-q", "-m", "not gpu",
           "-p", "no:cacheprovider"]
    if select:
        cmd += ["-k", select]
    try:
        r = subprocess.run(cmd, env=env, cwd=HERE, capture_output=True, text=True, timeout=30)
    except subprocess.TimeoutExpired:
        return None, "timed out"
    return r.returncode, r.stdout[-1500:]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_unmodified_oracle_passes_the_pins(tmp_path):
    with open(SRC) as fh:
        text = fh.read()
    rc, out = _pins(_build(str(tmp_path), "unmodified", text))
    assert rc == 0, out


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("name,old,new,pin", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_planted_misreading_fails_its_pin(tmp_path, name, old, new, pin):
    with open(SRC) as fh:
        text = fh.read()
    assert text.count(old) == 1, name
    text = text.replace(old, new)
    if name == "reverse_update_order":
        # process the sweep's peaks last to first: run the loop over a
        # reversed copy of the list, then put the list back in order
        text = text.replace(
            "  for (int i = 0; i < fd->nfit; i++) {\n    if (fd->fit[i].status != ST_RUNNING) continue;\n"
            "    peak_copy(fd, &fd->fit[i], w);",
            "  for (int ii = 0; ii < fd->nfit; ii++) {\n    int i = fd->nfit - 1 - ii;\n"
            "    if (fd->fit[i].status != ST_RUNNING) continue;\n    peak_copy(fd, &fd->fit[i], w);")
        assert "int i = fd->nfit - 1 - ii;" in text
    rc, out = _pins(_build(str(tmp_path), name, text), select=pin)
    assert rc != 0, (name, out)   # failed, or hung
