"""Many realizations in one persistent launch (optionally a replayed HIP graph).

Each item is one generator call (``RandMeth.__call__`` /
``IncomprRandMeth.__call__``, ``src/gstools/field/generator.py:243-270``,
``:529-567``): its own modes, points (array or grid axes) and FP32 output
(unscaled sum).  :class:`BatchPlan` wraps ``gsx_batch_plan_*`` from the C ABI;
:func:`autotune` times the kernel variants (points per tile, modes per LDS
chunk) on a workload and caches the winner per device and workload shape.
"""

import ctypes

import numpy as np
import torch

from . import _native
from .summate import uniform_axis

__all__ = ["BatchItem", "BatchPlan", "MixedBatch", "sample_and_sum", "autotune", "TILE_VARIANTS"]

TILE_VARIANTS = [(512, 128), (512, 256), (1024, 128), (1024, 256), (2048, 128), (2048, 256)]


class BatchItem:
    """One realization: device tensors for modes, points or axes, and output.

    Axes given as host arrays are uploaded in one copy and kept on the host
    too (the uniform-axis check of the matrix-core plan reads them there);
    axes already on the device may come with ``axes_host``, host copies of
    the same values, for that check without a device round trip."""

    def __init__(self, cov, z1, z2, pos=None, axes=None, incompr=False, out=None, axes_host=None):
        dev = torch.device("cuda", torch.cuda.current_device())
        as_dev = lambda a: (a if isinstance(a, torch.Tensor) else torch.as_tensor(np.asarray(a, dtype=np.float64))
                            ).to(dev, torch.float64).contiguous()
        self.cov = as_dev(cov)
        if self.cov.dim() == 1:
            self.cov = self.cov.reshape(1, -1)
        self.d, self.N = int(self.cov.shape[0]), int(self.cov.shape[1])
        self.z1, self.z2 = as_dev(z1).reshape(-1), as_dev(z2).reshape(-1)
        if self.z1.numel() != self.N or self.z2.numel() != self.N:
            raise ValueError("z_1 / z_2 must have N entries")
        if (pos is None) == (axes is None):
            raise ValueError("give exactly one of pos and axes")
        self.pos = self.axes = None
        if pos is not None:
            self.pos = as_dev(pos).reshape(self.d, -1)
            self.shape = (int(self.pos.shape[1]), 1, 1)
            self.P = int(self.pos.shape[1])
        else:
            if any(isinstance(a, torch.Tensor) for a in axes):
                ax = [as_dev(a).reshape(-1) for a in axes]
                self.axes = torch.cat(ax)
                dims = [int(a.numel()) for a in ax]
                self.axes_host = None
                if axes_host is not None:
                    self.axes_host = [np.asarray(a, dtype=np.float64).reshape(-1) for a in axes_host]
                    if [a.size for a in self.axes_host] != dims:
                        raise ValueError("axes_host must match the axes' lengths")
            else:
                self.axes_host = [np.asarray(a, dtype=np.float64).reshape(-1) for a in axes]
                dims = [int(a.size) for a in self.axes_host]
                self.axes = torch.from_numpy(np.concatenate(self.axes_host)).to(dev)
            if len(dims) != self.d:
                raise ValueError("one axis per dimension")
            self.shape = tuple(dims + [1] * (3 - self.d))
            self.P = int(np.prod(dims))
        self.incompr = bool(incompr)
        if self.incompr and self.d < 2:
            raise ValueError(f"BatchItem: incompressible items need d = 2 or 3, got d = {self.d}")
        nout = self.d if self.incompr else 1
        self.out = out if out is not None else torch.empty((nout, self.P), dtype=torch.float32, device=dev)

    def to_c(self):
        it = _native.BatchItem()
        it.d, it.kind, it.N, it.P = self.d, int(self.incompr), self.N, self.P
        for i in range(3):
            it.shape[i] = self.shape[i]
        ptr = lambda t: t.data_ptr() if t is not None and t.numel() else None
        it.cov, it.z1, it.z2 = ptr(self.cov), ptr(self.z1), ptr(self.z2)
        it.pos, it.axes, it.out = ptr(self.pos), ptr(self.axes), ptr(self.out)
        return it


class BatchPlan:
    """A persistent-launch plan over ``items`` (kept alive by the plan).

    The plan's work queue is one device counter that every launch resets:
    launch a plan on one stream at a time (never an eager launch overlapping
    a replay of its graph); build one plan per stream for concurrent work.
    """

    def __init__(self, items, tile_points=0, mode_chunk=0):
        self.items = list(items)
        arr = (_native.BatchItem * max(1, len(self.items)))(*[it.to_c() for it in self.items])
        handle = ctypes.c_void_p()
        _native.check(_native.lib().gsx_batch_plan_create(arr, len(self.items), int(tile_points), int(mode_chunk),
                                                          ctypes.byref(handle)), "batch_plan_create")
        self._h = handle
        self.tile_points, self.mode_chunk = tile_points, mode_chunk

    @property
    def tiles(self):
        return int(_native.lib().gsx_batch_plan_tiles(self._h))

    def launch(self, graph=False):
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        fn = _native.lib().gsx_batch_plan_launch_graph if graph else _native.lib().gsx_batch_plan_launch
        _native.check(fn(self._h, s), "batch_plan_launch")

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _native.lib().gsx_batch_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MixedBatch:
    """Grid items (d = 2, 3; scalar or incompressible) on the matrix cores, the other
    items in one persistent FP32 launch; optionally captured once and replayed as a
    HIP graph.

    Grid items run ``gsx_summate_grid_mfma`` / ``gsx_summate_incompr_grid_mfma`` in
    split bf16 ("bf16x3") or fp16 ("fp16x3"), one workspace per branch (the launches
    of a branch are ordered on its stream); outputs are the unscaled FP32 sums, as in
    :class:`BatchPlan`.  The graph holds kernel launches only (no memset nodes).

    ``branches`` > 1 (default 4): the grid items are split over that many side streams
    (longest work first, each branch with its own workspace), forked from the
    caller's stream with an event and joined back, while the persistent launch
    runs on the caller's stream; captured, the graph keeps the branches
    parallel.  Each item's output is the same as with one branch, bit for bit.
    """

    def __init__(self, items, precision="bf16x3", tile_points=0, mode_chunk=0, tune=True, branches=4):
        """``tune``: each grid item runs the matrix-core plan the autotuner chose
        for its class (timed on first use, then cached; see :mod:`.autotune`)."""
        if precision not in ("bf16x3", "fp16x3"):
            raise ValueError("precision must be 'bf16x3' or 'fp16x3'")
        self.items = list(items)
        self.variant = _native.MFMA_BF16X3 if precision == "bf16x3" else _native.MFMA_FP16X3
        self.grid_items = [it for it in self.items if it.axes is not None and it.d >= 2]
        rest = [it for it in self.items if not (it.axes is not None and it.d >= 2)]
        self.plan = BatchPlan(rest, tile_points, mode_chunk) if rest else None
        lib = _native.lib()
        self._calls = []
        if int(branches) < 1:
            raise ValueError("branches must be >= 1")
        self.branches = int(branches)
        # grid items over the branches: longest work first, each to the least loaded
        load = [0.0] * self.branches
        need = [1] * self.branches
        order = sorted(self.grid_items, key=lambda it: -float(it.P) * it.N * (it.d if it.incompr else 1))
        branch_of = {}
        for it in order:
            b = min(range(self.branches), key=lambda i: load[i])
            load[b] += float(it.P) * it.N * (it.d if it.incompr else 1)
            branch_of[id(it)] = b
        for it in self.grid_items:
            shape = (ctypes.c_int64 * 3)(*it.shape)
            q = lib.gsx_summate_incompr_grid_mfma_workspace if it.incompr else lib.gsx_summate_grid_mfma_workspace
            nbytes = int(q(it.d, it.N, shape))
            if nbytes < 0:
                _native.check(nbytes, "mixed batch workspace")
            b = branch_of[id(it)]
            need[b] = max(need[b], nbytes)
            variant = self.variant
            last = it.axes_host[1] if getattr(it, "axes_host", None) else it.axes[it.shape[0]:it.shape[0] + it.shape[1]]
            if it.d == 2 and not it.incompr and uniform_axis(last):
                variant |= _native.MFMA_UNIFORM_LAST  # factorised last axis
            if tune:
                from .autotune import mfma_variant

                variant = mfma_variant(it.d, it.N, it.shape[:it.d], variant, comps=it.d if it.incompr else 1)
            self._calls.append((it, shape, variant, branch_of[id(it)]))
        self.tuned = bool(tune)
        dev = self.items[0].out.device if self.items else "cuda"
        self.ws = [torch.empty(n, dtype=torch.uint8, device=dev) for n in need]
        # side streams at the default priority, one per branch when branching
        self.streams = [torch.cuda.Stream(device=dev) for _ in range(self.branches)] if self.branches > 1 else []
        self.graph = None

    def tuning_summary(self):
        """Per grid class: the tuned choice, its time and the candidates' spread (ms)."""
        from .autotune import _lookup, _mfma_key

        out = {}
        for it, _, variant, _ in self._calls:
            base = variant & (_native.MFMA_FP16X3 | _native.MFMA_UNIFORM_LAST)
            key = _mfma_key(it.d, it.N, it.shape[:it.d], it.d if it.incompr else 1, base)
            entry = _lookup(key)
            if entry is None or key in out:
                continue
            ms = entry["ms"]
            out[key.split("|", 1)[1]] = {"choice": entry["choice"], "ms": round(min(ms.values()), 4),
                                         "default_ms": round(entry["default_ms"], 4),
                                         "spread_ms": [round(min(ms.values()), 4), round(max(ms.values()), 4)],
                                         "candidates": entry["candidates"], "items": 0}
        for it, _, variant, _ in self._calls:
            base = variant & (_native.MFMA_FP16X3 | _native.MFMA_UNIFORM_LAST)
            k = _mfma_key(it.d, it.N, it.shape[:it.d], it.d if it.incompr else 1, base).split("|", 1)[1]
            if k in out:
                out[k]["items"] += 1
        return out

    def _grid_calls(self, branch, stream):
        lib = _native.lib()
        s = ctypes.c_void_p(stream.cuda_stream)
        ws = self.ws[branch]
        for it, shape, variant, b in self._calls:
            if b != branch:
                continue
            fn = lib.gsx_summate_incompr_grid_mfma if it.incompr else lib.gsx_summate_grid_mfma
            _native.check(fn(it.d, it.N, shape, it.axes.data_ptr(), it.cov.data_ptr(), it.z1.data_ptr(),
                             it.z2.data_ptr(), variant, 1.0, it.out.data_ptr(), ctypes.c_void_p(ws.data_ptr()),
                             ws.numel(), s),
                          "mixed batch grid item")

    def _run(self):
        main = torch.cuda.current_stream()
        if self.branches == 1:
            self._grid_calls(0, main)
            if self.plan is not None:
                self.plan.launch(graph=False)
            return
        # fork: every branch waits for the work already on the caller's stream
        fork = torch.cuda.Event()
        fork.record(main)
        joins = []
        for b, st in enumerate(self.streams):
            st.wait_event(fork)
            self._grid_calls(b, st)
            ev = torch.cuda.Event()
            ev.record(st)
            joins.append(ev)
        if self.plan is not None:
            self.plan.launch(graph=False)  # on the caller's stream, beside the branches
        for ev in joins:
            main.wait_event(ev)  # join

    def launch(self, graph=False):
        if not graph:
            self._run()
            return
        if self.graph is None:
            self._run()  # eager warm-up before the capture
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._run()
            self.graph = g
        self.graph.replay()


def sample_and_sum(groups, items_of, precision="bf16x3", sampling="auto", nthreads=None, device=None,
                   tune=True):
    """Mode sampling and the batched sums of several groups of realizations,
    overlapped.  ``groups`` are ``(model, mode_nos, seeds)`` as for
    :func:`~gstools_amd.sampling.sample_modes_many`; ``items_of(i, modes)``
    returns the :class:`BatchItem` list of group i for its sampled modes.  Each
    group's sums (grids on the matrix cores, the rest in one persistent FP32
    launch) are launched on a stream of their own as soon as the group's modes
    are final (the order in which ``sample_modes_many`` hands the groups over:
    the device groups first once their chains have run, else the groups drawn
    on the host while the chains still run).  An ``items_of`` that uploads
    modes should stage them through pinned memory: a pageable copy waits for
    the kernels of the groups already launched.  Returns ``(modes, items, batches)``:
    modes and items per group, and the :class:`MixedBatch` objects whose
    launches fill the items' outputs (keep them until the outputs are read).
    The outputs are complete once the caller's stream has synchronized (the
    group streams are joined back into it)."""
    from .sampling import sample_modes_many

    main = torch.cuda.current_stream()
    start = torch.cuda.Event()
    start.record(main)  # the work before this call; the device chains come after it on `main`
    items = [None] * len(groups)
    streams, keep = [], []

    def ready(i, modes):
        st = torch.cuda.Stream()
        st.wait_event(start)
        with torch.cuda.stream(st):
            its = items_of(i, modes)
            if its:
                mb = MixedBatch(its, precision, tune=tune)
                mb.launch(graph=False)
                keep.append(mb)
        items[i] = its
        streams.append(st)

    modes = sample_modes_many(groups, sampling=sampling, nthreads=nthreads, device=device, on_ready=ready)
    for st in streams:
        main.wait_stream(st)  # join
    return modes, items, keep


def autotune(items, reps=3, graph=True, use_cache=True, verbose=False):
    """Time every (tile_points, mode_chunk) variant of the persistent plan on
    ``items``; return (best, timings_ms).  Cached by :mod:`.autotune` under the
    device architecture, the item kinds, the mean mode count and the work."""
    from .autotune import tune_batch

    best, timings = tune_batch(items, TILE_VARIANTS, reps=reps, graph=graph, force=not use_cache)
    if verbose:
        for name, ms in timings.items():
            print(f"autotune {name}: {ms:.3f} ms", flush=True)
    return tuple(best), timings
