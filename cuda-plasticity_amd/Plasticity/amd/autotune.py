"""Autotuner for the batched stepper: picks the flux tile shape (and hence
the halo-to-interior ratio of the LDS tile) per grid-size class.

Candidates are the tile variants compiled into the library (0: 32x8,
1: 64x4, 2: 16x16 cells of the two fastest axes).  Each candidate is timed
on the batch's actual instances of that size with HIP events
(cdd_batch_time_flux); the fastest is installed with cdd_batch_set_tile.
Choices are cached per (device name, side, dynamics) in a JSON file so a
later run can reuse them without timing.
"""
import json
import os

VARIANTS = {0: "32x8", 1: "64x4", 2: "16x16"}


def _valid(n, v):
    return {0: n >= 32, 1: n >= 64, 2: n >= 16}[v]


def autotune_batch(batch, reps=5, cache_path=None, dynamics=None, verbose=False):
    import torch
    name = torch.cuda.get_device_name(batch.dev)
    cache = {}
    if cache_path and os.path.exists(cache_path):
        try:
            cache = json.load(open(cache_path))
        except ValueError:
            cache = {}
    choices, timings = {}, {}
    for n in sorted(set(batch.sizes)):
        key = "%s|%d|%s" % (name, n, dynamics)
        if key in cache:
            best = int(cache[key])
        else:
            t = {v: batch.time_flux(n, v, reps) for v in VARIANTS if _valid(n, v)}
            timings[n] = t
            best = min(t, key=t.get)
            cache[key] = best
        batch.set_tile(n, best)
        choices[n] = best
        if verbose:
            print("autotune n=%d -> %s %s" % (n, VARIANTS[best], timings.get(n, "(cached)")))
    if cache_path:
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return choices, timings
