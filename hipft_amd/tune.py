"""Launch autotuner for the batched RKG2 diffusion.

The knobs (``hft_set_tuning``) are the stages per fused launch (temporal
depth), the theta rows per block (tile height), the number of still-active
realizations below which stages go one per launch, the block count below
which a fused launch thins its bands (the tail of a batch whose
realizations have different stage counts), and whether each launch instead
picks its band count against whole rounds of co-resident blocks.  The WENO3 stages have their own
knobs (banded or one thread per point, rows per band), timed on advection
steps.  Results are identical
for every setting, so the tuner only measures: it times diffusion super-steps
of the live batch for each candidate and keeps the fastest.  Choices are
cached in ``tuning_cache.json`` beside this file, keyed by device name, grid
and batch size.
"""
import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tuning_cache.json")

DIFF_KEYS = ("fused", "fused_stages", "band_rows", "fused_min_realizations",
             "fused_target_blocks", "fused_kernel", "fused_bands", "fused_band_share",
             "fused_body", "fused_band_min_rows")
ADV_KEYS = ("advection_band", "advection_band_rows")
KEYS = DIFF_KEYS + ADV_KEYS
DEFAULTS = dict(fused=1, fused_stages=4, band_rows=32, fused_min_realizations=8,
                fused_target_blocks=0, fused_kernel=3, fused_bands=0, fused_band_share=32,
                fused_body=1, fused_band_min_rows=8, advection_band=1, advection_band_rows=16)


def adv_candidates():
    return [dict(advection_band=0, advection_band_rows=16)] + \
        [dict(advection_band=1, advection_band_rows=h) for h in (8, 16, 32, 64, 128)]


def candidates(nr):
    out = [dict(fused=0, fused_stages=4, band_rows=32, fused_min_realizations=1,
                fused_target_blocks=0, fused_kernel=3, fused_bands=0)]
    for K in (1, 3):
        for T in (3, 4):
            for H in (32, 64, 128):
                for m in sorted({1, 4, 8, 16, 32} & set(range(1, max(nr, 1) + 1))):
                    for B in (0, 256):
                        out.append(dict(fused=1, fused_stages=T, band_rows=H,
                                        fused_min_realizations=m, fused_target_blocks=B,
                                        fused_kernel=K, fused_bands=0))
    # five stages per launch: the two-column kernel only
    for H in (32, 64, 128):
        for m in sorted({1, 4, 8, 16, 32} & set(range(1, max(nr, 1) + 1))):
            for B in (0, 256):
                out.append(dict(fused=1, fused_stages=5, band_rows=H, fused_min_realizations=m,
                                fused_target_blocks=B, fused_kernel=3, fused_bands=0))
    # band count chosen per launch (two-column kernel only; band_rows unused),
    # with the band-count model's memory-sharing slowdown in percent
    for T in (3, 4, 5):
        for m in sorted({1, 4, 8, 16, 32} & set(range(1, max(nr, 1) + 1))):
            for share in (16, 32, 64):
                out.append(dict(fused=1, fused_stages=T, band_rows=32, fused_min_realizations=m,
                                fused_target_blocks=0, fused_kernel=3, fused_bands=1,
                                fused_band_share=share))
    # interior march steps without row tests: same results, never slower
    return [dict({"fused_band_share": DEFAULTS["fused_band_share"], "fused_body": 1,
                  "fused_band_min_rows": DEFAULTS["fused_band_min_rows"]}, **c)
            for c in out]


def tail_candidates():
    """Switch to one-stage launches below m realizations, bands down to h
    rows: the launches late in a super-step and in extra PTL rounds, which
    only whole steps of an evolving state exercise."""
    return [dict(fused_min_realizations=m, fused_band_min_rows=h)
            for m in (1, 2, 4, 8) for h in (2, 4, 8, 16)]


def cache_key(model):
    # libhft is built for gfx950 only; the marketing name that HIP reports
    # varies with how the runtime was initialised, so it is not part of the key
    return f"gfx950|{model.ntm}x{model.npm1}|nr={model.nr}"


def load_cache():
    if os.path.exists(CACHE):
        with open(CACHE) as fh:
            return json.load(fh)
    return {}


def apply(model, cfg):
    model.set_tuning(**{k: cfg.get(k, DEFAULTS[k]) for k in KEYS})


def nearest_cached(model, cache):
    """The cached choice for the same grid whose batch size is nearest (in
    ratio) to this one's, or None."""
    grid = f"gfx950|{model.ntm}x{model.npm1}|nr="
    best, best_d = None, None
    for k, v in cache.items():
        if k.startswith(grid):
            d = abs(np.log(int(k[len(grid):]) / model.nr))
            if best_d is None or d < best_d:
                best, best_d = v, d
    return best


def autotune(model, dt=1.0, reps=3, use_cache=True, write_cache=True, cands=None, log=None,
             borrow=False, step_runs=8):
    """Pick and apply the fastest launch configuration for this batch.
    borrow: with no cached choice for this batch size, apply the one cached
    for the nearest batch size on the same grid instead of measuring
    (results do not depend on the choice, only the speed does).
    step_runs: Strang steps per measurement of the last stage (the tail
    knobs, timed on whole steps from the batch's state); 0 skips it."""
    key = cache_key(model)
    cache = load_cache()
    if use_cache and key in cache:
        apply(model, cache[key])
        return cache[key]
    if use_cache and borrow:
        near = nearest_cached(model, cache)
        if near is not None:
            apply(model, near)
            return dict(near, borrowed=True)
    state = model.get_state("tp")

    def measure(cands, run, what):
        best, best_t = None, float("inf")
        for cfg in cands:
            apply(model, cfg)
            model.set_state(state, layout="tp")
            run()  # warm
            model.set_state(state, layout="tp")
            model.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                run()
            model.synchronize()
            el = (time.perf_counter() - t0) / reps
            if log:
                log(f"tune {what} {cfg} -> {1e3 * el:.2f} ms")
            if el < best_t:
                best, best_t = dict(cfg), el
        return best, best_t

    best, best_t = measure([dict(DEFAULTS, **c) for c in (cands or candidates(model.nr))],
                           lambda: model.diffusion_step(dt, subcycles=0), "diffusion")
    if model.has_flow():
        adv, adv_t = measure([dict(best, **c) for c in adv_candidates()],
                             lambda: model.advection_step(0.5 * dt), "advection")
        best = adv
        best["ms_per_advection_step"] = round(1e3 * adv_t, 3)
    best["ms_per_diffusion_step"] = round(1e3 * best_t, 3)
    if step_runs > 0 and best.get("fused") and best.get("fused_kernel") == 3 and \
            best.get("fused_bands"):
        def steps():
            for _ in range(step_runs):
                model.flux_transport_step(dt, strang=True, subcycles=0)
        keep = {k: v for k, v in best.items() if k.startswith("ms_per_")}
        tail, tail_t = measure([dict(best, **c) for c in tail_candidates()], steps, "steps")
        best = dict(tail, **keep)
        best["ms_per_flux_transport_step"] = round(1e3 * tail_t / step_runs, 3)
    model.set_state(state, layout="tp")
    apply(model, best)
    if write_cache:
        cache = load_cache()
        cache[key] = best
        with open(CACHE, "w") as fh:
            json.dump(cache, fh, indent=1, sort_keys=True)
    return best
