"""GPU tests of the batched persistent launch and its HIP-graph replay.

The batch tiles run the same FP32 instruction sequence, on the same 64-mode
sub-blocks, as the single-realization FP32 kernels, so every item must equal
its own single launch bit for bit; each is also checked against the oracle
(<= 1e-5 relative).  Graph replays are interleaved with eager work and every
replay is compared with the eager result.
"""

import sys

import numpy as np
import pytest
import torch

import oracle
from gstools_amd import summate_device, summate_grid_device, summate_incompr_device
from gstools_amd.batch import TILE_VARIANTS, BatchItem, BatchPlan, MixedBatch, autotune
from gstools_amd.geometry import generate_grid

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mixed_items(seed=0):
    rs = np.random.RandomState(seed)
    specs = []
    # (d, N, kind, points or axes)
    specs.append((2, 1000, False, None, [np.arange(64.0), np.arange(70.0)]))
    specs.append((3, 300, False, rs.uniform(0, 1000, size=(3, 5003)) / 7.0, None))
    specs.append((3, 257, True, None, [np.arange(9.0), np.arange(10.0), np.arange(11.0)]))
    specs.append((2, 700, True, rs.uniform(0, 50, size=(2, 3000)), None))
    specs.append((1, 129, False, rs.uniform(-5, 5, size=(1, 999)), None))
    specs.append((2, 64, False, rs.uniform(0, 10, size=(2, 1)), None))
    # the second 2048-point segment reaches |x| ~ 3e6: there the large modes of
    # every 128-mode group take the fract reduction, the small ones the magic
    far = rs.uniform(0, 1000, size=(3, 6000))
    far[:, 2048:4096] *= 3000.0
    specs.append((3, 300, False, far, None))
    items, refs = [], []
    for d, n, incompr, pos, axes in specs:
        cov = rs.normal(size=(d, n)) * 0.3
        z1, z2 = rs.normal(size=n), rs.normal(size=n)
        items.append(BatchItem(cov, z1, z2, pos=pos, axes=axes, incompr=incompr))
        p = pos if pos is not None else generate_grid(axes)
        refs.append((cov, z1, z2, p, axes, incompr))
    return items, refs


def _single(cov, z1, z2, p, axes, incompr):
    if incompr:
        return summate_incompr_device(cov, z1, z2, p, "fp32").cpu().numpy()
    if axes is not None:
        return summate_grid_device(cov, z1, z2, axes, "fp32").cpu().numpy().reshape(1, -1)
    return summate_device(cov, z1, z2, p, "fp32").cpu().numpy().reshape(1, -1)


@pytest.mark.parametrize("variant", TILE_VARIANTS)
def test_batch_matches_single_launches(variant):
    _say(f"batch variant {variant}")
    items, refs = _mixed_items()
    plan = BatchPlan(items, *variant)
    assert plan.tiles >= len(items)
    plan.launch()
    torch.cuda.synchronize()
    for it, (cov, z1, z2, p, axes, incompr) in zip(items, refs):
        got = it.out.cpu().numpy()
        single = _single(cov, z1, z2, p, axes, incompr)
        assert np.array_equal(got, single)
        ref = oracle.summate_incompr_c(cov, z1, z2, p) if incompr else oracle.summate_c(cov, z1, z2, p)[None]
        assert oracle.rel_errors(got.astype(np.float64), ref)[0] <= 1e-5
    plan.close()


def test_batch_graph_replays():
    _say("batch graph replays")
    items, refs = _mixed_items(seed=1)
    plan = BatchPlan(items, 1024, 256)
    plan.launch()
    torch.cuda.synchronize()
    eager = [it.out.clone() for it in items]
    for rep in range(4):
        for it in items:
            it.out.fill_(float("nan"))
        # unrelated eager work between replays
        _ = summate_device(refs[1][0], refs[1][1], refs[1][2], refs[1][3], "fp32")
        torch.cuda.synchronize()
        plan.launch(graph=True)
        torch.cuda.synchronize()
        for it, e in zip(items, eager):
            assert torch.equal(it.out, e), f"replay {rep} differs"
    plan.close()


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3"])
def test_mixed_batch_grids_on_matrix_cores(precision):
    """MixedBatch: the 2D / 3D grid items (scalar and incompressible) on the matrix
    cores within 1e-5 (rms) of FP64, the others bitwise equal to their single FP32
    launches; graph replays (with eager work between) equal the eager run."""
    _say(f"mixed batch {precision}")
    items, refs = _mixed_items(seed=2)
    mb = MixedBatch(items, precision)
    assert len(mb.grid_items) == 2
    mb.launch()
    torch.cuda.synchronize()
    eager = [it.out.clone() for it in items]
    for it, (cov, z1, z2, p, axes, incompr) in zip(items, refs):
        got = it.out.cpu().numpy().astype(np.float64)
        ref = oracle.summate_incompr_c(cov, z1, z2, p) if incompr else oracle.summate_c(cov, z1, z2, p)[None]
        if it in mb.grid_items:
            for c in range(got.shape[0]):
                assert oracle.rel_errors(got[c], ref[c])[1] <= 1e-5, (precision, c)
        else:
            assert np.array_equal(it.out.cpu().numpy(), _single(cov, z1, z2, p, axes, incompr))
    for rep in range(3):
        for it in items:
            it.out.fill_(float("nan"))
        _ = summate_device(refs[1][0], refs[1][1], refs[1][2], refs[1][3], "fp32")
        torch.cuda.synchronize()
        mb.launch(graph=True)
        torch.cuda.synchronize()
        for it, e in zip(items, eager):
            assert torch.equal(it.out, e), f"replay {rep} differs"


def test_mixed_batch_branches_equal_one_branch():
    """Grid items split over side streams (forked and joined with events) give
    each item the same output as one branch, bit for bit, eager and in graph
    replays (with eager work and synchronization between the replays)."""
    _say("mixed batch branches")
    rs = np.random.RandomState(7)
    items = []
    for k, shape in enumerate([(64, 70), (200, 256), (333, 17), (48, 40, 36), (128, 130), (20, 21, 22)]):
        d, n = len(shape), int(rs.choice([300, 1000, 2048]))
        cov = rs.normal(size=(d, n)) * 0.3
        axes = [np.arange(m, dtype=np.float64) * 0.7 for m in shape]
        items.append(BatchItem(cov, rs.normal(size=n), rs.normal(size=n), axes=axes, incompr=(k == 5)))
    for k in range(3):
        cov = rs.normal(size=(3, 500)) * 0.3
        items.append(BatchItem(cov, rs.normal(size=500), rs.normal(size=500),
                               pos=rs.uniform(0, 100, size=(3, 4000 + 777 * k))))
    one = MixedBatch(items, "bf16x3", tune=False, branches=1)
    one.launch()
    torch.cuda.synchronize()
    ref = [it.out.clone() for it in items]
    mb = MixedBatch(items, "bf16x3", tune=False, branches=3)
    assert len({b for *_, b in mb._calls}) == 3
    for it in items:
        it.out.fill_(float("nan"))
    mb.launch()
    torch.cuda.synchronize()
    for it, r in zip(items, ref):
        assert torch.equal(it.out, r)
    for rep in range(3):
        for it in items:
            it.out.fill_(float("nan"))
        _ = summate_device(items[6].cov, items[6].z1, items[6].z2, items[6].pos, "fp32")
        torch.cuda.synchronize()
        mb.launch(graph=True)
        torch.cuda.synchronize()
        for it, r in zip(items, ref):
            assert torch.equal(it.out, r), f"replay {rep} differs"
    with pytest.raises(ValueError):
        MixedBatch(items, "bf16x3", tune=False, branches=0)


def test_batch_item_axes_forms_equal():
    """Grid items with host axes, device axes, and device axes with their host
    copies (``axes_host``) give the same outputs bit for bit (uniform and
    non-uniform last axes, 2D and 3D)."""
    _say("batch item axes forms")
    rs = np.random.RandomState(11)
    dev = torch.device("cuda", 0)
    specs = []
    for shape, uniform in [((300, 512), True), ((64, 300), False), ((40, 36, 300), True)]:
        n = int(rs.choice([500, 1000]))
        cov = rs.normal(size=(len(shape), n)) * 0.3
        axes = [np.arange(m, dtype=np.float64) * 0.7 for m in shape]
        if not uniform:
            axes[-1] = np.sort(rs.uniform(0, 200, shape[-1]))
        specs.append((cov, rs.normal(size=n), rs.normal(size=n), axes))
    outs = []
    for form in ("host", "device", "device+host"):
        items = []
        for cov, z1, z2, axes in specs:
            if form == "host":
                items.append(BatchItem(cov, z1, z2, axes=axes))
            else:
                dax = [torch.from_numpy(a).to(dev) for a in axes]
                items.append(BatchItem(cov, z1, z2, axes=dax, axes_host=axes if form == "device+host" else None))
        MixedBatch(items, "bf16x3", tune=False).launch(graph=False)
        torch.cuda.synchronize()
        outs.append([it.out.clone() for it in items])
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert torch.equal(a, b)
    cov, z1, z2, axes = specs[0]
    with pytest.raises(ValueError):
        BatchItem(cov, z1, z2, axes=[torch.from_numpy(a).to(dev) for a in axes], axes_host=[axes[0], axes[1][:-1]])


def test_batch_empty_and_invalid():
    plan = BatchPlan([], 1024, 256)
    assert plan.tiles == 0
    plan.launch()
    plan.launch(graph=True)
    torch.cuda.synchronize()
    plan.close()
    items, _ = _mixed_items()
    with pytest.raises(ValueError):
        BatchPlan(items, 777, 256)


def test_autotune_returns_a_variant(tmp_path, monkeypatch):
    _say("autotune")
    monkeypatch.setenv("GSX_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    items, _ = _mixed_items(seed=2)
    best, ms = autotune(items, reps=1)
    assert best in TILE_VARIANTS and len(ms) == len(TILE_VARIANTS)
    best2, _ = autotune(items, reps=1)  # from the cache
    assert best2 == best
