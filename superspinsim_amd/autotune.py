"""Autotuner for the persistent protocol launches.

Chooses, by measurement on the device, the launch shape (wavefronts per
workgroup) and the Taylor plan of a ``ProtocolBatch``:

* every candidate plan meets the same truncation bound (the propagator's
  measured plan at the batch tolerance), so the choice changes speed, not
  accuracy: degree m with s = 1, 2, 4 substeps, i.e. the exponent norm above
  which an exponential is split into 2^k parts;
* wavefronts per workgroup 1, 2, 4 or 8 (generator values are staged once
  per workgroup in LDS; more wavefronts share one copy).

Each candidate runs on a trial copy of the batch: its longest instances (one
grid's worth), truncated to ``trial_steps`` steps, timed with HIP events on
the launch stream.  Results are cached per device and problem shape in
``$XDG_CACHE_HOME/superspinsim_amd/autotune.json`` (default ``~/.cache``).
"""

import json
import os

import numpy as np

from superspinsim_amd import _lib
from superspinsim_amd.propagator import _torch, taylor_plan_measured

WAVES_PER_BLOCK = (1, 2, 4, 8)
SUBSTEPS = (1, 2, 4)


def cache_path():
    root = os.environ.get("XDG_CACHE_HOME") or os.path.join(
        os.path.expanduser("~"), ".cache")
    return os.path.join(root, "superspinsim_amd", "autotune.json")


def _load_cache():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(cache):
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass                            # a read-only home only loses caching


def plan_candidates(batch, samples):
    """(m, s) plans meeting the batch tolerance for s in SUBSTEPS."""
    prop = batch.prop
    exps = prop.sample_exponents(samples, batch.h)
    out = []
    for s in SUBSTEPS:
        try:
            # the measured plan restricted to exactly s substeps
            scaled = [a / s for a in exps]
            m, _ = taylor_plan_measured(scaled, tol=batch.tolerance / s,
                                        max_substeps=1)
        except ValueError:
            continue
        if (m, s) not in out:
            out.append((m, s))
    return out


def problem_key(batch):
    torch = _torch()
    prop = batch.prop
    name = torch.cuda.get_device_properties(0).name
    return "|".join(str(v) for v in (
        name, prop.n, prop.ell.slots, prop.K, prop.J, prop.S, prop.ell.nnz,
        batch.tolerance, batch.n_obs))


def tune(batch, trial_steps=20000, repeats=2, use_cache=True, verbose=False):
    """Measure every candidate on a trial copy; apply and return the best.

    Returns a dict with the chosen ``waves_per_block``, ``taylor``,
    ``substeps``, the per-candidate times (ms) and whether it came from the
    cache.
    """
    torch = _torch()
    cache = _load_cache() if use_cache else {}
    key = problem_key(batch)
    samples = batch.sim.coefficient_box(batch.waves_host,
                                        batch.program)
    plans = plan_candidates(batch, samples)
    if use_cache and key in cache:
        best = cache[key]
        if (best["taylor"], best["substeps"]) in [tuple(p) for p in plans]:
            apply(batch, best)
            return dict(best, cached=True)

    # trial copy: the longest instances, one grid's worth, truncated
    saved = (batch.steps.copy(), batch.read_from.copy(), batch.n_waves,
             batch.waves_per_block, batch.taylor, batch.substeps)
    count = min(batch.n_inst, batch.n_waves)
    trial_steps_arr = np.minimum(batch.steps, trial_steps)
    trial_read = np.minimum(batch.read_from, trial_steps_arr)
    order = batch.order[:count]
    mask = np.zeros(batch.n_inst, dtype=bool)
    mask[order] = True
    trial_steps_arr = np.where(mask, trial_steps_arr, 0)
    batch.d_steps.copy_(torch.from_numpy(trial_steps_arr))
    batch.d_read_from.copy_(torch.from_numpy(trial_read))
    results = []
    try:
        for m, s in plans:
            for wpb in WAVES_PER_BLOCK:
                batch.taylor, batch.substeps = m, s
                batch.waves_per_block = wpb
                batch.launch()              # warm
                stream = torch.cuda.current_stream()
                ev = [torch.cuda.Event(enable_timing=True)
                      for _ in range(2)]
                ev[0].record(stream)
                for _ in range(repeats):
                    batch.launch()
                ev[1].record(stream)
                torch.cuda.synchronize()
                ms = ev[0].elapsed_time(ev[1]) / repeats
                results.append(dict(taylor=m, substeps=s,
                                    waves_per_block=wpb, ms=ms))
                if verbose:
                    print(f"tune: m={m} s={s} wpb={wpb}: {ms:.3f} ms")
    finally:
        steps, read_from, n_waves, wpb, m, s = saved
        batch.d_steps.copy_(torch.from_numpy(steps))
        batch.d_read_from.copy_(torch.from_numpy(read_from))
        batch.n_waves, batch.waves_per_block = n_waves, wpb
        batch.taylor, batch.substeps = m, s
        torch.cuda.synchronize()
    best = min(results, key=lambda r: r["ms"])
    choice = dict(waves_per_block=best["waves_per_block"],
                  taylor=best["taylor"], substeps=best["substeps"],
                  trial_steps=int(trial_steps), trials=results)
    if use_cache:
        cache[key] = choice
        _store_cache(cache)
    apply(batch, choice)
    return dict(choice, cached=False)


def apply(batch, choice):
    batch.waves_per_block = int(choice["waves_per_block"])
    batch.taylor = int(choice["taylor"])
    batch.substeps = int(choice["substeps"])
    if not 1 <= batch.taylor <= _lib.MAX_TAYLOR:
        raise ValueError("tuned Taylor degree out of range")


__all__ = ["tune", "plan_candidates", "cache_path", "apply"]
