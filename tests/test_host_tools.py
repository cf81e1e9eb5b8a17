"""Host-side tools without a GPU: the launch autotuner's search and cache,
and bench.py's one-GPU contract."""
import json
import os
import subprocess
import sys
import types

import pytest

from hipft_amd import tune

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# The tuner's clock in these tests: the fake model advances it by its cost,
# so the search does not depend on how long sleeps take on a busy host.
_CLOCK = [0.0]


@pytest.fixture(autouse=True)
def _fake_clock(monkeypatch):
    monkeypatch.setattr(tune, "time", types.SimpleNamespace(perf_counter=lambda: _CLOCK[0]))


class FakeModel:
    """Stands in for FluxTransport: records tunings, times by a cost table."""

    def __init__(self, nr=4, cost=None):
        self.nr, self.ntm, self.npm1 = nr, 33, 65
        self.cfg = {}
        self.cost = cost or (lambda c: 1.0)
        self.calls = 0

    def set_tuning(self, **kw):
        self.cfg.update(kw)

    def get_state(self, layout):
        return "state"

    def set_state(self, s, layout):
        assert s == "state"

    def synchronize(self):
        pass

    def has_flow(self):
        return True

    def diffusion_step(self, dt, subcycles):
        self.calls += 1
        _CLOCK[0] += 1e-4 * self.cost(self.cfg)

    def advection_step(self, dt):
        _CLOCK[0] += 1e-4 * (2.0 if self.cfg["advection_band"] == 0 else 1.0)

    def flux_transport_step(self, dt, strang, subcycles):
        self.steps = getattr(self, "steps", 0) + 1
        _CLOCK[0] += 1e-4 * self.step_cost(self.cfg)

    step_cost = staticmethod(lambda c: 1.0)


def test_candidates_cover_knobs():
    c1 = tune.candidates(1)
    c64 = tune.candidates(64)
    assert any(c["fused"] == 0 for c in c1)
    assert all(c["fused_min_realizations"] <= 1 for c in c1 if c["fused"])
    assert {c["fused_kernel"] for c in c64 if c["fused"]} == {1, 3}
    assert {c["fused_bands"] for c in c64 if c["fused"]} == {0, 1}
    assert all(c["fused_kernel"] == 3 for c in c64 if c["fused_bands"])
    assert all(set(tune.DIFF_KEYS) <= set(c) for c in c64)
    assert all(c["fused_kernel"] == 3 for c in c64 if c["fused"] and c["fused_stages"] == 5)
    assert {c["fused_band_share"] for c in c64 if c["fused"] and c["fused_bands"]} == {16, 32, 64}


def test_apply_sets_every_key_with_defaults():
    m = FakeModel()
    tune.apply(m, {"fused": 0})
    assert set(m.cfg) == set(tune.KEYS)
    assert m.cfg["fused"] == 0 and m.cfg["band_rows"] == tune.DEFAULTS["band_rows"]


def test_autotune_picks_fastest_and_caches(tmp_path, monkeypatch):
    monkeypatch.setattr(tune, "CACHE", str(tmp_path / "cache.json"))
    best_cfg = dict(fused=1, fused_stages=3, band_rows=64, fused_min_realizations=4,
                    fused_target_blocks=0, fused_kernel=1)
    cost = lambda c: 0.2 if all(c.get(k) == v for k, v in best_cfg.items()) else 3.0
    m = FakeModel(nr=4, cost=cost)
    got = tune.autotune(m, reps=1, use_cache=False)
    assert all(got[k] == v for k, v in best_cfg.items())
    assert got["advection_band"] == 1
    cache = json.load(open(tmp_path / "cache.json"))
    assert tune.cache_key(m) in cache
    m2 = FakeModel(nr=4, cost=cost)
    again = tune.autotune(m2, use_cache=True)
    assert again == cache[tune.cache_key(m)] and m2.calls == 0


def test_autotune_tail_stage_times_whole_steps(tmp_path, monkeypatch):
    """With the band-count kernel chosen, a last stage times whole steps for
    the tail knobs (one-stage threshold, thinnest band) and keeps the
    diffusion and advection choices."""
    monkeypatch.setattr(tune, "CACHE", str(tmp_path / "cache.json"))
    best_cfg = dict(fused=1, fused_stages=4, fused_kernel=3, fused_bands=1,
                    fused_min_realizations=4, fused_band_share=16)
    cost = lambda c: 0.2 if all(c.get(k) == v for k, v in best_cfg.items()) else 3.0
    m = FakeModel(nr=64, cost=cost)
    m.step_cost = lambda c: 0.3 if (c["fused_min_realizations"], c["fused_band_min_rows"]) == \
        (2, 4) else 2.0
    got = tune.autotune(m, reps=1, use_cache=False, step_runs=2)
    assert (got["fused_min_realizations"], got["fused_band_min_rows"]) == (2, 4)
    assert all(got[k] == v for k, v in best_cfg.items() if k != "fused_min_realizations")
    assert got["advection_band"] == 1
    assert m.steps == 2 * 2 * len(tune.tail_candidates())  # warm-up run + one timed run each
    assert "ms_per_flux_transport_step" in got and "ms_per_diffusion_step" in got


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def _counter_csv(d, counter, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "run_counter_collection.csv"), "w") as fh:
        fh.write("Kernel_Name,Counter_Name,Counter_Value\n")
        for name, value in rows:
            fh.write(f'"{name}",{counter},{value}\n')


def test_pmc_summary_scales_counters_and_keeps_probe(tmp_path):
    """The counter summary scales FETCH_SIZE/WRITE_SIZE so the calibration
    copy reads and writes its own size, averages per dispatch, and keeps the
    probe's algorithmic bytes per fused launch."""
    from hipft_amd import pmc_summary
    copy = 256 << 20
    kib = copy // 1024
    fdir, wdir = str(tmp_path / "f"), str(tmp_path / "w")
    # the copy's FETCH_SIZE reads half its size (a 2x scale), WRITE_SIZE its size
    _counter_csv(fdir, "FETCH_SIZE", [("__amd_rocclr_copyBuffer", kib // 2),
                                      ("k_fused", 1000), ("k_fused", 3000)])
    _counter_csv(wdir, "WRITE_SIZE", [("__amd_rocclr_copyBuffer", kib),
                                      ("k_fused", 500), ("k_fused", 500)])
    log = tmp_path / "probe.log"
    log.write_text('noise\n{"pmc_probe": "done", "copy_bytes": 268435456, '
                   '"fused_launches": 2, "fused_alg_bytes_per_launch": 3000000}\n')
    out_json, out_md = str(tmp_path / "s.json"), str(tmp_path / "s.md")
    pmc_summary.main([fdir, wdir, out_json, out_md, "--probe", str(log)])
    s = json.load(open(out_json))
    k = s["kernels"]["k_fused"]
    assert s["fetch_scale"] == 2.0 and s["write_scale"] == 1.0
    assert k["dispatches"] == 2
    assert k["read_bytes_avg"] == 2 * 1024 * 2000
    assert k["write_bytes_avg"] == 1024 * 500
    assert s["probe"]["fused_alg_bytes_per_launch"] == 3000000
    assert "k_fused" in open(out_md).read()
