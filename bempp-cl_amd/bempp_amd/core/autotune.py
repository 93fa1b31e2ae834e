"""Launch-shape autotuner for the frequency-batched assembler.

The persistent kbatch kernel has two free launch parameters: the number of test elements
staged in LDS per tile (8, 16 or 32) and the number of resident workgroups per CU
(the persistent grid is CUs x that number). `tune` times every candidate on the actual
problem (same spaces, same wavenumbers, same output buffer) with HIP events and keeps the
fastest. Results are cached in memory and in a JSON file keyed by
(operator, shapesets, rows, cols, number of wavenumbers, precision, device name), so a
configuration is tuned once per machine. The cache file is $BEMPP_AMD_TUNE_CACHE, else
$XDG_CACHE_HOME/bempp_amd/tune.json, else ~/.cache/bempp_amd/tune.json.
"""

import json
import os

CANDIDATE_TILES = (8, 16, 32)
CANDIDATE_WG_PER_CU = (1, 2, 4, 8)
DEFAULT_CONFIG = (16, 4)

_MEMORY = {}


def cache_path():
    p = os.environ.get("BEMPP_AMD_TUNE_CACHE")
    if p:
        return p
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "bempp_amd", "tune.json")


def _load():
    try:
        with open(cache_path()) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def _store(key, value):
    data = _load()
    data[key] = value
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp%d" % os.getpid()
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only loses persistence, not correctness


def make_key(op, test_shapeset, trial_shapeset, rows, cols, nk, precision, device_name):
    return "|".join(str(x) for x in (op, test_shapeset, trial_shapeset, rows, cols, nk, precision, device_name))


def _valid_entry(entry):
    """The launch shape of a persisted entry, or None if the entry is malformed or outside the
    candidate set (a stale or hand-edited cache file then only costs a re-tune)."""
    try:
        cfg = entry["config"]
        if len(cfg) != 2:
            return None
        tile, wg = cfg
        if isinstance(tile, bool) or isinstance(wg, bool) or int(tile) != tile or int(wg) != wg:
            return None
        tile, wg = int(tile), int(wg)
    except (TypeError, KeyError, ValueError, IndexError):
        return None
    if tile not in CANDIDATE_TILES or wg not in CANDIDATE_WG_PER_CU:
        return None
    return (tile, wg)


def lookup(key):
    if key in _MEMORY:
        return _MEMORY[key]
    entry = _load().get(key)
    if entry is not None:
        cfg = _valid_entry(entry)
        if cfg is None:
            return None
        _MEMORY[key] = cfg
        return cfg
    return None


def tune(key, run, candidates=None, repeats=2, passes=2):
    """Time run(config) for every candidate config on the current stream; cache and return the best.

    run: callable enqueuing one complete launch with the given (test_tile, workgroups_per_cu).
    Each candidate is timed over `passes` sweeps through the candidate list (so a clock ramp
    early in the first sweep does not decide) and ranked by the median of its launches (so one
    fast or slow launch does not either).
    """
    import torch

    if candidates is None:
        candidates = [(t, w) for t in CANDIDATE_TILES for w in CANDIDATE_WG_PER_CU]
    for cfg in candidates:
        run(cfg)  # warm-up (code object load, caches)
    samples = {cfg: [] for cfg in candidates}
    for _ in range(passes):
        for cfg in candidates:
            for _ in range(repeats):
                start = torch.cuda.Event(enable_timing=True)
                end = torch.cuda.Event(enable_timing=True)
                start.record()
                run(cfg)
                end.record()
                end.synchronize()
                samples[cfg].append(start.elapsed_time(end))
    timings = {cfg: sorted(v)[len(v) // 2] for cfg, v in samples.items()}
    best = min(timings, key=timings.get)
    _MEMORY[key] = best
    _store(key, {"config": list(best), "ms": timings[best],
                 "all": {"%d,%d" % c: round(t, 4) for c, t in timings.items()}})
    return best


def clear_memory():
    _MEMORY.clear()
