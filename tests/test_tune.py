"""
The autotuner's host logic, on the CPU with a stand-in fitter whose kernel
time is a known function of the knobs: tile sides map to reserved-box
margins (a tile below the ROI is no candidate), the fastest candidate is
picked and left set, trials fit `trial_sweeps` sweeps, and a pick stored in
the cache is reused for the same workload signature without trials (and
not for another workload).
"""
import pytest

from storm_analysis_amd.sa_library import dao_tune


class FakeFitter(object):
    """BatchFitter's tuning surface; the 'kernel' is fastest at waves 4, margin 2, team 0."""
    def __init__(self, shapes=((128, 128), (2048, 2048)), counts=(40, 9000), roi_size=18):
        self.shapes = list(shapes)
        self.counts = list(counts)
        self.model, self.error_model, self.roi_size, self.precision = "3d", "MLE", roi_size, "mixed"
        self.tuning = None
        self.ms = 0.0
        self.fits = []

    def getNFit(self, f):
        return self.counts[f]

    def setTuning(self, waves=0, margin=-1, team=-1):
        self.tuning = (waves, margin, team)

    def setTiming(self, on=True):
        pass

    def stats(self):
        return {"lm_ms": self.ms}

    def doFit(self, max_iterations=2000):
        w, m, t = self.tuning
        self.fits.append(max_iterations)
        self.ms += 100.0 + abs(w - 4) * 10.0 + abs(m - 2) * 7.0 + t * 0.01


def test_tiles_map_to_margins():
    assert dao_tune.margins_for((16, 18, 20, 22, 24, 25), 18) == (0, 1, 2, 3)
    assert dao_tune.margins_for((16,), 18) == ()


def test_picks_the_fastest_and_sets_it():
    bf = FakeFitter()
    calls = []
    best, trials = dao_tune.autotune(bf, lambda: calls.append(1), waves=(2, 4, 8), tiles=(16, 18, 20, 22, 24),
                                     teams=(0, 500))
    assert best == (4, 2, 0) and bf.tuning == best
    assert len(trials) == 3 * 4 * 2 and len(calls) == len(trials)
    assert set(t["tile"] for t in trials) == {18, 20, 22, 24}
    assert all(n == 2000 for n in bf.fits)


class TailFitter(FakeFitter):
    """Short fits favour team 0, whole fits (the long tail) team 500."""
    def doFit(self, max_iterations=2000):
        w, m, t = self.tuning
        self.fits.append(max_iterations)
        per_sweep = (1.0 if t == 0 else 1.6) if max_iterations <= 20 else (4.0 if t == 0 else 2.5)
        self.ms += max_iterations * per_sweep + abs(w - 4) * 10.0 + abs(m - 2) * 7.0


def test_two_stages_choose_the_team_on_whole_fits():
    bf = TailFitter()
    best, trials = dao_tune.autotune(bf, lambda: None, waves=(2, 4, 8), tiles=(18, 20, 22), teams=(0, 500),
                                     trial_sweeps=10)
    assert best == (4, 2, 500) and bf.tuning == best
    stage1 = [t for t in trials if t["stage"] == 1]
    stage2 = [t for t in trials if t["stage"] == 2]
    assert len(stage1) == 9 and all(t["sweeps"] == 10 and t["team"] == 0 for t in stage1)
    assert len(stage2) == 2 and all(t["sweeps"] == 2000 and (t["waves"], t["margin"]) == (4, 2) for t in stage2)


def test_no_tile_holds_the_roi():
    with pytest.raises(ValueError):
        dao_tune.autotune(FakeFitter(), lambda: None, tiles=(12, 16))


def test_cache_reuses_a_pick_for_the_same_workload(tmp_path):
    path = str(tmp_path / "tune.json")
    bf = FakeFitter()
    best, trials = dao_tune.autotune(bf, lambda: None, waves=(2, 4), margins=(1, 2), cache=path, trial_sweeps=5)
    assert trials and best == (4, 2, 0)
    # same signature (peak counts within the same power of two): no trials
    bf2 = FakeFitter(counts=(50, 10000))
    best2, trials2 = dao_tune.autotune(bf2, lambda: None, waves=(2, 4), margins=(1, 2), cache=path)
    assert trials2 == [] and best2 == best and bf2.tuning == best and bf2.fits == []
    # another workload: trials again
    bf3 = FakeFitter(shapes=((512, 512),), counts=(1000,))
    best3, trials3 = dao_tune.autotune(bf3, lambda: None, waves=(2, 4), margins=(1, 2), cache=path, trial_sweeps=5)
    assert trials3 and len(dao_tune._load_cache(path)) == 2
