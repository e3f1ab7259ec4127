"""Autotuner for the launch shapes of the summation kernels.

Two families of choices are tuned by timing them on the device:

* matrix-core grids (``gsx_summate_grid_mfma`` / ``_incompr_grid_mfma``):
  the K-split fill target (grids with fewer blocks than the target split
  their mode steps over more blocks: 2^8 .. 2^11 resident blocks) and
  whether a 3D grid with a short axis 0 may swap axes 0 and 1.  Candidates
  whose plans coincide (``gsx_summate_grid_mfma_plan``) are timed once;
* the persistent FP32 batch plan (``gsx_batch_plan_*``): points per work
  tile and modes per LDS chunk (:func:`tune_batch`).

Results are cached in a JSON file (``$GSX_AUTOTUNE_CACHE``, default
``~/.cache/gstools_amd/autotune.json``) keyed by device architecture, path,
dimension, mode-count bucket and point-count bucket (powers of two), plus
the per-axis size buckets for grids.  Every choice gives the same sums up
to rounding (the K split adds its partial sums in a fixed order).
"""

import ctypes
import json
import os
import threading

import numpy as np
import torch

from . import _native

__all__ = ["cache_path", "cache_key", "tune_mfma_grid", "mfma_variant", "tune_batch", "MFMA_FILL_LOG2"]

MFMA_FILL_LOG2 = (8, 9, 10, 11)
_lock = threading.Lock()
_memo = {}


def cache_path():
    return os.environ.get("GSX_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "gstools_amd", "autotune.json"))


def _bucket(n):
    return 0 if n <= 0 else 1 << int(np.floor(np.log2(n)))


def _arch():
    props = torch.cuda.get_device_properties(torch.cuda.current_device())
    return getattr(props, "gcnArchName", None) or props.name


def cache_key(path, d, N, P, extra=""):
    """``arch|path|d|N-bucket|P-bucket[|extra]``."""
    key = f"{_arch()}|{path}|d{d}|N{_bucket(N)}|P{_bucket(P)}"
    return f"{key}|{extra}" if extra else key


def _load():
    try:
        with open(cache_path()) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store(key, entry):
    with _lock:
        cache = _load()
        cache[key] = entry
        try:
            os.makedirs(os.path.dirname(cache_path()), exist_ok=True)
            tmp = cache_path() + f".{os.getpid()}.tmp"
            with open(tmp, "w") as fh:
                json.dump(cache, fh, indent=1, sort_keys=True)
            os.replace(tmp, cache_path())
        except OSError:
            pass
        _memo[key] = entry


def _lookup(key):
    if key in _memo:
        return _memo[key]
    entry = _load().get(key)
    if entry is not None:
        _memo[key] = entry
    return entry


def _time_ms(fn, reps):
    """Average milliseconds of fn() on the current stream (after one warm-up call)."""
    fn()
    stream = torch.cuda.current_stream()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        fn()
    end.record(stream)
    end.synchronize()
    return start.elapsed_time(end) / reps


# --------------------------------------------------------- matrix cores ----
def _mfma_plan(d, N, shape, comps, variant):
    info = (ctypes.c_int64 * 8)()
    shp = (ctypes.c_int64 * 3)(*(list(shape) + [1] * (3 - len(shape))))
    _native.check(_native.lib().gsx_summate_grid_mfma_plan(d, N, shp, comps, variant, info), "mfma_plan")
    return tuple(info)


def _mfma_candidates(d, N, shape, comps, base):
    """Tuning bits of the distinct plans among fill targets x swap choices."""
    seen, out = set(), []
    for f in MFMA_FILL_LOG2:
        for extra in (0, _native.MFMA_NO_SWAP):
            bits = (f << _native.MFMA_FILL_SHIFT) | extra
            plan = _mfma_plan(d, N, shape, comps, base | bits)
            if plan not in seen:
                seen.add(plan)
                out.append(bits)
    return out


def _mfma_key(d, N, shape, comps, base):
    shape = [int(n) for n in shape]
    prec = "fp16x3" if base & _native.MFMA_FP16X3 else "bf16x3"
    uni = "u" if base & _native.MFMA_UNIFORM_LAST else ""
    path = f"mfma-{prec}" + ("-incompr" if comps > 1 else "")
    return cache_key(path, d, N, int(np.prod(shape)), "x".join(str(_bucket(n)) for n in shape) + uni)


def tune_mfma_grid(d, N, shape, precision="bf16x3", comps=1, uniform_last=False, reps=3, force=False):
    """Time the distinct matrix-core plans of a grid; cache and return the entry
    ``{"bits": tuning bits, "ms": {candidate: ms}, "default_ms": ms}``."""
    from .summate import _stream

    base = _native.MFMA_FP16X3 if precision == "fp16x3" else _native.MFMA_BF16X3
    if uniform_last:
        base |= _native.MFMA_UNIFORM_LAST
    key = _mfma_key(d, N, shape, comps, base)
    if not force:
        entry = _lookup(key)
        if entry is not None:
            return entry
    shape = [int(n) for n in shape]
    dev = torch.device("cuda", torch.cuda.current_device())
    rs = np.random.RandomState(0)
    cov = torch.from_numpy(rs.normal(size=(d, N)) * 0.3).to(dev)
    z1 = torch.from_numpy(rs.normal(size=N)).to(dev)
    z2 = torch.from_numpy(rs.normal(size=N)).to(dev)
    axes = torch.cat([torch.arange(n, dtype=torch.float64, device=dev) * 0.5 for n in shape])
    P = int(np.prod(shape))
    out = torch.empty(comps * P, dtype=torch.float32, device=dev)
    lib = _native.lib()
    shp = (ctypes.c_int64 * 3)(*(shape + [1] * (3 - d)))
    q = lib.gsx_summate_incompr_grid_mfma_workspace if comps > 1 else lib.gsx_summate_grid_mfma_workspace
    ws_bytes = int(q(d, N, shp))
    if ws_bytes < 0:
        _native.check(ws_bytes, "mfma workspace")
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    fn = lib.gsx_summate_incompr_grid_mfma if comps > 1 else lib.gsx_summate_grid_mfma

    def run(bits):
        _native.check(fn(d, N, shp, ctypes.c_void_p(axes.data_ptr()), ctypes.c_void_p(cov.data_ptr()),
                         ctypes.c_void_p(z1.data_ptr()), ctypes.c_void_p(z2.data_ptr()), base | bits, 1.0,
                         ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ws.data_ptr()), ws_bytes, _stream()),
                      "mfma tune")

    cands = _mfma_candidates(d, N, shape, comps, base)
    ms = {}
    for bits in cands:
        ms[_bits_name(bits)] = _time_ms(lambda: run(bits), reps)
    best = min(cands, key=lambda b: ms[_bits_name(b)])
    default = _mfma_plan(d, N, shape, comps, base)
    default_bits = next(b for b in cands if _mfma_plan(d, N, shape, comps, base | b) == default)
    entry = {"bits": int(best), "choice": _bits_name(best), "ms": ms, "default_ms": ms[_bits_name(default_bits)],
             "candidates": len(cands)}
    _store(key, entry)
    return entry


def _bits_name(bits):
    f = (bits >> _native.MFMA_FILL_SHIFT) & 0xF
    return f"fill{1 << f}" + ("-noswap" if bits & _native.MFMA_NO_SWAP else "")


def mfma_variant(d, N, shape, base, comps=1, tune=True):
    """``base`` (precision and flag bits) or-ed with the tuned choice for this
    grid: from the cache, timed now when absent and ``tune``, else none."""
    key = _mfma_key(d, N, shape, comps, base)
    entry = _lookup(key)
    if entry is None and tune:
        prec = "fp16x3" if base & _native.MFMA_FP16X3 else "bf16x3"
        entry = tune_mfma_grid(d, N, shape, prec, comps, bool(base & _native.MFMA_UNIFORM_LAST))
    return base | (entry["bits"] if entry else 0)


# ---------------------------------------------------------- FP32 batch ----
def tune_batch(items, variants, reps=3, graph=True, force=False):
    """Time the batch plan variants ((tile points, mode chunk) pairs) on ``items``;
    cache and return ``(best, {name: ms})``."""
    from .batch import BatchPlan

    work = sum(it.P * it.N * (it.d if it.incompr else 1) for it in items)
    n_mean = int(np.mean([it.N for it in items])) if items else 0
    dims = "".join(sorted({f"{it.d}{'i' if it.incompr else ''}{'g' if it.axes is not None else 'p'}"
                           for it in items}))
    key = cache_key("fp32-batch", dims, n_mean, work, f"items{_bucket(len(items))}")
    if not force:
        entry = _lookup(key)
        if entry is not None:
            return tuple(entry["best"]), entry["ms"]
    timings = {}
    for tp, mc in variants:
        plan = BatchPlan(items, tp, mc)
        timings[f"{tp}x{mc}"] = _time_ms(lambda: plan.launch(graph), reps)
        plan.close()
    best = min(variants, key=lambda v: timings[f"{v[0]}x{v[1]}"])
    _store(key, {"best": list(best), "ms": timings})
    return best, timings
