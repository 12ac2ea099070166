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


# The multiplane pins (test_mp_pins.py) against copies of oracle/mp_oracle.c
MP_SRC = os.path.join(os.path.dirname(HERE), "oracle", "mp_oracle.c")
MP_MUTATIONS = [
    ("mp_heights_unweighted", "      sum += d * w * mp->heights[c];\n      norm += w * mp->heights[c];",
     "      sum += d * w;\n      norm += w;", "test_first_coupled_step"),
    ("mp_back_with_0toN", "const double *tr = axis == 0 ? mp->xt_Nto0 : mp->yt_Nto0;",
     "const double *tr = axis == 0 ? mp->xt_0toN : mp->yt_0toN;", "test_first_coupled_step"),
    ("mp_transposed_coefficient",
     "double d = tr[3 * c + 1] * mp->w_jacobian[c][1];\n      d += tr[3 * c + 2] * mp->w_jacobian[c][2];",
     "double d = tr[3 * c + 1] * mp->w_jacobian[c][2];\n      d += tr[3 * c + 2] * mp->w_jacobian[c][1];",
     "test_first_coupled_step"),
    ("mp_weights_channel_major", "  double w1 = wts[nc * zi + c];\n  double w2 = wts[nc * (zi + 1) + c];",
     "  double w1 = wts[2 * c + zi];\n  double w2 = wts[2 * c + zi + 1];", "test_first_coupled_step"),
    ("mp_y_width_follows", "    mp->fit_data[c]->working_peak->params[P_XW] -= mp->w_jacobian[c][3];\n",
     "    mp->fit_data[c]->working_peak->params[P_XW] -= mp->w_jacobian[c][3];\n"
     "    mp->fit_data[c]->working_peak->params[P_YW] = mp->fit_data[c]->working_peak->params[P_XW];\n",
     "test_first_coupled_step"),
    ("mp_map_without_offset", "    p->params[P_XC] = t - xoff;", "    p->params[P_XC] = t;",
     "test_first_coupled_step"),
    ("mp_step_offset_added", "double d = tr[3 * c + 1] * mp->w_jacobian[c][1];",
     "double d = tr[3 * c] + tr[3 * c + 1] * mp->w_jacobian[c][1];", "test_a_whole_pixel_shifted_copy_changes_nothing"),
    ("mp_mean_unnormalised", "    p0->params[axis == 0 ? P_XC : P_YC] -= sum / norm;",
     "    p0->params[axis == 0 ? P_XC : P_YC] -= sum;", "test_a_whole_pixel_shifted_copy_changes_nothing"),
    ("mp_no_height_floor", "    if (p->params[P_HEIGHT] < 0.01) p->params[P_HEIGHT] = 0.01;\n", "",
     "test_heights_floor_at_a_hundredth"),
    ("mp_decide_on_channel_0", "        else\n          cur += mp->fit_data[c]->working_peak->error;",
     "        else if (c == 0)\n          cur += mp->fit_data[c]->working_peak->error;",
     "test_the_summed_deviance_decides"),
    ("mp_start_error_channel_0", "      start += fd->working_peak->error;",
     "      if (c == 0) start += fd->working_peak->error;", "test_a_whole_pixel_shifted_copy_changes_nothing"),
    ("mp_non_decr_channel_0", "          mp->fit_data[c]->n_non_decr++;",
     "          if (c == 0) mp->fit_data[c]->n_non_decr++;", "test_the_summed_deviance_decides"),
    ("mp_solve_every_channel", "      for (int c = 0; c < nc && !bad; c++) {", "      for (int c = 0; c < nc; c++) {",
     "test_failed_solve_stops_the_channel_loop"),
    ("mp_no_sync_at_new_peaks",
     "  for (int i = (stop - n_peaks > 0 ? stop - n_peaks : 0); i < stop; i++) check_group(mp, i);",
     "  (void)stop;", "test_a_group_off_one_channel_is_an_error_everywhere"),
    ("mp_sync_old_range",
     "  for (int i = (stop - n_peaks > 0 ? stop - n_peaks : 0); i < stop; i++) check_group(mp, i);",
     "  for (int i = start; i < start + n_peaks && i < stop; i++) check_group(mp, i);",
     "test_new_error_groups_after_storage_growth"),
]


def _build(tmp, name, text):
    src = os.path.join(tmp, name + ".c")
    with open(src, "w") as fh:
        fh.write(text)
    so = os.path.join(tmp, "lib" + name + ".so")
    # -I: the multiplane oracle includes dao_oracle.c from its own directory
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-w", "-I", os.path.dirname(SRC), "-o", so, src, "-lm"],
                   check=True)
    return so


def _pins(so, select=None, tests="test_lm_branches.py", var="DAO_ORACLE_SO"):
    """Exit status of the pins against the oracle build `so` (None: it hung -- a failure too)."""
    env = dict(os.environ)
    env[var] = so
    cmd = [sys.executable, "-m", "pytest", os.path.join(HERE, tests), "-q", "-m", "not gpu",
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


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_unmodified_multiplane_oracle_passes_the_pins(tmp_path):
    with open(MP_SRC) as fh:
        text = fh.read()
    rc, out = _pins(_build(str(tmp_path), "mp_unmodified", text), tests="test_mp_pins.py", var="DAO_MP_ORACLE_SO")
    assert rc == 0, out


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("name,old,new,pin", MP_MUTATIONS, ids=[m[0] for m in MP_MUTATIONS])
def test_planted_multiplane_misreading_fails_its_pin(tmp_path, name, old, new, pin):
    with open(MP_SRC) as fh:
        text = fh.read()
    assert text.count(old) == 1, name
    rc, out = _pins(_build(str(tmp_path), name, text.replace(old, new)), select=pin, tests="test_mp_pins.py",
                    var="DAO_MP_ORACLE_SO")
    assert rc != 0, (name, out)   # failed, or hung
