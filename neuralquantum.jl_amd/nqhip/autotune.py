"""Launch-shape autotuner.

Four knobs decide how an SR iteration's work is spread over the device
without changing what is computed (nqh_set_launch_shape): estimator samples
per workgroup, estimator waves per workgroup, BF16 Gram split-K chunks and
sampler waves per workgroup.  `tune` runs coordinate descent over them: from
the problem-derived defaults, each knob in turn takes every candidate value
with the others held, and a value is kept when its time per iteration beats
the best so far by more than `gain`; two rounds.  A trial is `iters` timed
graph-replayed iterations after `warmup` untimed ones (the first replay after
a change captures the graph again).

The chosen shapes of the benchmark workloads are stored in
launch_shapes.json next to this file (written by `save`, read by `load`).
No reference counterpart: the reference has no device launches to shape.
"""
from __future__ import annotations

import json
import os
import time

KNOBS = ("est_samples_per_block", "est_waves", "gram_splits", "sampler_waves")
CANDIDATES = {
    "est_samples_per_block": (8, 16, 32, 64),
    "est_waves": (4, 8),
    "gram_splits": (1, 2, 4, 8, 16),
    "sampler_waves": (1, 2, 4),
}
PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "launch_shapes.json")


def _measure(apply, step, sync, shape, iters, warmup):
    apply(shape)
    for _ in range(warmup):
        step()
    sync()
    t0 = time.perf_counter()
    for _ in range(iters):
        step()
    sync()
    return (time.perf_counter() - t0) / iters * 1e3


def tune(apply, step, sync, iters=10, warmup=3, rounds=2, gain=0.01, candidates=CANDIDATES):
    """Coordinate descent over the four knobs.  apply(shape) sets a shape
    (list of four ints, 0 = default), step() runs one iteration, sync() waits
    for the device.  Returns {"shape", "ms", "default_ms", "trials"}."""
    shape = [0, 0, 0, 0]
    best = _measure(apply, step, sync, shape, iters, warmup)
    default_ms = best
    trials = [{"shape": list(shape), "ms": round(best, 4)}]
    for _ in range(rounds):
        changed = False
        for k, name in enumerate(KNOBS):
            for v in candidates[name]:
                if v == shape[k]:
                    continue
                trial = list(shape)
                trial[k] = v
                t = _measure(apply, step, sync, trial, iters, warmup)
                trials.append({"shape": trial, "ms": round(t, 4)})
                if t < best * (1.0 - gain):
                    best, shape, changed = t, trial, True
        if not changed:
            break
    apply(shape)
    return {"shape": dict(zip(KNOBS, shape)), "ms": round(best, 4), "default_ms": round(default_ms, 4),
            "trials": trials}


def tune_context(ctx, eps, lr, **kw):
    """Tune one context's SR iteration (graph replay)."""
    ctx.set_graph(True)
    return tune(lambda s: ctx.set_launch_shape(*s), lambda: ctx.sr_iteration(eps, lr), lambda: None, **kw)


def tune_batch(batch, members, eps, lr, **kw):
    """Tune a batch: one shape for every member (the packed form launches
    all members' workgroups together)."""
    def apply(s):
        for c in members:
            c.set_launch_shape(*s)
    return tune(apply, lambda: batch.sr_iteration(eps, lr), lambda: None, **kw)


def load(workload, path=PATH):
    """The stored shape of a workload as a list of four ints, or None."""
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    e = d.get(workload)
    if not e:
        return None
    return [int(e["shape"][k]) for k in KNOBS]


def save(workload, result, note="", path=PATH):
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        d = {}
    d[workload] = {"shape": result["shape"], "ms_per_iteration": result["ms"],
                   "default_ms_per_iteration": result["default_ms"],
                   "trials": [[t["shape"], t["ms"]] for t in result["trials"]], "note": note}
    with open(path, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
        f.write("\n")
