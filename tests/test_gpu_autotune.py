"""Autotuned matrix-core plans (K-split fill target, axis swap) on the GPU.

Every candidate plan of a grid computes the same sum up to rounding: each is
held to the split-precision bar against the FP64 oracle (rel_rms 1e-5).
Tuning times the distinct candidates, writes the cache entry keyed by
architecture, path, dimension, mode and point buckets, and is reused; a
``MixedBatch`` on tuned plans matches one on the default plans to rounding.
"""

import ctypes
import importlib
import sys

import numpy as np
import pytest
import torch

import oracle
from gstools_amd import _native, summate_grid_device, summate_incompr_grid_device

pytestmark = pytest.mark.gpu

autotune = importlib.import_module("gstools_amd.autotune")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture
def fresh_cache(tmp_path, monkeypatch):
    monkeypatch.setenv("GSX_AUTOTUNE_CACHE", str(tmp_path / "autotune.json"))
    autotune._memo.clear()
    yield tmp_path / "autotune.json"
    autotune._memo.clear()


def _modes(d, n, seed):
    rs = np.random.RandomState(seed)
    return rs.normal(size=(d, n)) * 0.4, rs.normal(size=n), rs.normal(size=n)


GRIDS = [(64, 64), (300, 200), (32, 96, 80), (40, 48, 48)]


@pytest.mark.parametrize("shape", GRIDS)
def test_every_candidate_plan_meets_the_bar(shape):
    d, n = len(shape), 1024
    cov, z1, z2 = _modes(d, n, 3)
    axes = [np.arange(k) * 0.7 for k in shape]
    grid = np.asarray(np.meshgrid(*axes, indexing="ij")).reshape(d, -1)
    sub = np.random.RandomState(1).choice(grid.shape[1], min(3000, grid.shape[1]), replace=False)
    ref = oracle.summate_c(cov, z1, z2, np.ascontiguousarray(grid[:, sub]))
    cands = autotune._mfma_candidates(d, n, shape, 1, _native.MFMA_BF16X3)
    _say(f"autotune candidates {shape}: {[autotune._bits_name(b) for b in cands]}")
    dev = torch.device("cuda", 0)
    cov_t, z1_t, z2_t = (torch.from_numpy(a).to(dev) for a in (cov, z1, z2))
    flat = torch.from_numpy(np.concatenate(axes)).to(dev)
    shp = (ctypes.c_int64 * 3)(*(list(shape) + [1] * (3 - d)))
    lib = _native.lib()
    ws_bytes = int(lib.gsx_summate_grid_mfma_workspace(d, n, shp))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    for bits in cands:
        out = torch.full((grid.shape[1],), float("nan"), dtype=torch.float32, device=dev)
        st = lib.gsx_summate_grid_mfma(d, n, shp, flat.data_ptr(), cov_t.data_ptr(), z1_t.data_ptr(),
                                       z2_t.data_ptr(), _native.MFMA_BF16X3 | bits, 1.0, out.data_ptr(),
                                       ws.data_ptr(), ws_bytes, torch.cuda.current_stream().cuda_stream)
        assert st == 0
        got = out.cpu().numpy().astype(np.float64)
        assert np.all(np.isfinite(got))
        rmax, rrms = oracle.rel_errors(got[sub], ref)
        assert rrms <= 1e-5 and rmax <= 5e-5, (autotune._bits_name(bits), rmax, rrms)


def test_tuning_caches_and_is_reused(fresh_cache):
    _say("autotune: tune, cache, reuse")
    e = autotune.tune_mfma_grid(3, 2048, (32, 96, 80), "bf16x3")
    assert e["choice"] in e["ms"] and e["candidates"] == len(e["ms"]) >= 2
    assert fresh_cache.exists()
    key = autotune._mfma_key(3, 2048, (32, 96, 80), 1, _native.MFMA_BF16X3)
    assert key.startswith(autotune._arch() + "|mfma-bf16x3|d3|N2048|P131072|")
    autotune._memo.clear()  # read back from the file
    assert autotune.mfma_variant(3, 2048, (32, 96, 80), _native.MFMA_BF16X3, tune=False) == e["bits"]
    # tuned grid functions agree with the default plans to rounding
    cov, z1, z2 = _modes(3, 2048, 5)
    axes = [np.arange(k) * 0.5 for k in (32, 96, 80)]
    a = summate_grid_device(cov, z1, z2, axes, "bf16x3").cpu().numpy()
    b = summate_grid_device(cov, z1, z2, axes, "bf16x3", tune=True).cpu().numpy()
    assert oracle.rel_errors(b.astype(np.float64), a.astype(np.float64))[1] <= 1e-5
    c = summate_incompr_grid_device(cov, z1, z2, [x[:40] for x in axes], "bf16x3").cpu().numpy()
    dd = summate_incompr_grid_device(cov, z1, z2, [x[:40] for x in axes], "bf16x3", tune=True).cpu().numpy()
    assert oracle.rel_errors(dd.ravel().astype(np.float64), c.ravel().astype(np.float64))[1] <= 1e-5


def test_mixed_batch_tuned_matches_default(fresh_cache):
    from gstools_amd.batch import BatchItem, MixedBatch

    _say("autotune: MixedBatch tuned vs default plans")
    dev = torch.device("cuda", 0)
    items = []
    for k, (shape, n) in enumerate([((64, 64), 1000), ((200, 256), 2048), ((48, 48, 48), 1000)]):
        cov, z1, z2 = _modes(len(shape), n, 10 + k)
        axes = [torch.arange(m, dtype=torch.float64, device=dev) for m in shape]
        items.append(BatchItem(cov, z1, z2, axes=axes, incompr=k == 2))
    pos = np.random.RandomState(2).uniform(0, 100, size=(3, 5000))
    cov, z1, z2 = _modes(3, 512, 20)
    items.append(BatchItem(cov, z1, z2, pos=pos))
    MixedBatch(items, "bf16x3", tune=False).launch(graph=True)
    torch.cuda.synchronize()
    ref = [it.out.clone() for it in items]
    mb = MixedBatch(items, "bf16x3", tune=True)
    mb.launch(graph=True)
    torch.cuda.synchronize()
    summary = mb.tuning_summary()
    assert len(summary) == 3 and all(v["items"] == 1 for v in summary.values())
    for it, r in zip(items, ref):
        got, want = it.out.cpu().numpy().astype(np.float64), r.cpu().numpy().astype(np.float64)
        assert oracle.rel_errors(got.ravel(), want.ravel())[1] <= 1e-5
