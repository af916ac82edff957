"""Autotuner for the batched stepper.

Three choices are timed with HIP events on the batch's own instances and
states (no state change) and installed:
  * per size class, the stress solve's R2C / C2R row tile (0: the default
    rows per block, 1: half of them; sides >= 256; cdd_batch_time_spectral,
    cdd_batch_set_spectral);
  * per size class, the flux tile of the per-size launch (0: 32x8, 1: 64x4,
    2: 16x16 cells of the two fastest axes; cdd_batch_time_flux,
    cdd_batch_set_tile) -- the flux of sizes outside the multi-instance
    launch, and of the per-size mode;
  * the persistent multi-instance flux launch (one launch per stage over
    every instance of every size): the tile walk (tiles 0-2, 1 or 2 blocks
    per CU) or the row march (tile 3, sides that are multiples of 128, 1-3
    blocks per CU), against the per-size launches as a whole
    (cdd_batch_time_flux_multi, cdd_batch_set_flux; per_cu 0 selects the
    per-size launches).
Choices are cached per (device name, sizes, dynamics) in a JSON file so a
later run can reuse them without timing.
"""
import json
import os

VARIANTS = {0: "32x8", 1: "64x4", 2: "16x16"}
TILE_SHAPE = {0: (32, 8), 1: (64, 4), 2: (16, 16)}
MULTI = {0: "32x8 tiles", 1: "64x4 tiles", 2: "16x16 tiles", 3: "row march"}
PER_CU = {0: (1, 2), 1: (1, 2), 2: (1, 2), 3: (1, 2, 3)}


def _valid(n, v):
    return {0: n >= 32, 1: n >= 64, 2: n >= 16}[v]


def _multi_ok(n, v):
    """Whether a group of side n joins the multi-instance launch of tile v."""
    if v == 3:
        return n >= 128 and n % 128 == 0
    tc, tr = TILE_SHAPE[v]
    return n >= tc and n >= tr and n % tc == 0 and n % tr == 0


def autotune_batch(batch, reps=5, cache_path=None, dynamics=None, verbose=False):
    """Returns (choices, timings): choices[n] is size n's per-size flux tile,
    choices["spectral"][n] its stress-solve row tile and choices["multi"] =
    (tile, per_cu) the multi-instance launch's configuration (per_cu 0:
    per-size launches); timings holds the measured ms per launch of every
    candidate."""
    import torch
    name = torch.cuda.get_device_name(batch.dev)
    cache = {}
    if cache_path and os.path.exists(cache_path):
        try:
            cache = json.load(open(cache_path))
        except ValueError:
            cache = {}
    choices, timings = {}, {}
    sides = sorted(set(batch.sizes))
    best_ms = {}
    for n in sides:
        key = "%s|%d|%s" % (name, n, dynamics)
        if key in cache:
            best = int(cache[key])
        else:
            t = {v: batch.time_flux(n, v, reps) for v in VARIANTS if _valid(n, v)}
            timings[n] = t
            best = min(t, key=t.get)
            best_ms[n] = t[best]
            cache[key] = best
        batch.set_tile(n, best)
        choices[n] = best
        if verbose:
            print("autotune n=%d -> %s %s" % (n, VARIANTS[best], timings.get(n, "(cached)")))
    choices["spectral"] = {}
    for n in sides:
        if n < 256:
            continue
        key = "%s|%d|%s|spectral" % (name, n, dynamics)
        if key in cache:
            best = int(cache[key])
        else:
            t = {v: batch.time_spectral(n, v, reps) for v in (0, 1)}
            timings["spectral %d" % n] = t
            best = min(t, key=t.get)
            cache[key] = best
        batch.set_spectral(n, best)
        choices["spectral"][n] = best
    counts = {n: batch.sizes.count(n) for n in sides}
    mkey = "%s|multi|%s|%s" % (name, ",".join("%dx%d" % (n, counts[n]) for n in sides), dynamics)
    if mkey in cache:
        tile, per_cu = (int(x) for x in cache[mkey])
    else:
        cand = {}
        for v in MULTI:
            if not any(_multi_ok(n, v) for n in sides):
                continue
            # sizes outside this variant's multi launch keep their per-size launch
            rest = sum(best_ms.get(n, 0.0) for n in sides if not _multi_ok(n, v))
            for p in PER_CU[v]:
                cand[(v, p)] = batch.time_flux_multi(v, p, reps) + rest
        if len(best_ms) == len(sides):
            cand[(0, 0)] = sum(best_ms.values())  # the per-size launches one after another
        timings["multi"] = {"%d/%d" % k: ms for k, ms in cand.items()}
        tile, per_cu = min(cand, key=cand.get) if cand else (0, 0)
        cache[mkey] = [tile, per_cu]
    batch.set_flux(tile, per_cu)
    choices["multi"] = (tile, per_cu)
    if verbose:
        print("autotune multi -> %s, %d blocks/CU %s" % (MULTI[tile], per_cu, timings.get("multi", "(cached)")))
    if cache_path:
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return choices, timings
