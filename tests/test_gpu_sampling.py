"""Mode sampling on the GPU (``gsx_stretch_run_device``) against the host sampler.

The host lockstep sampler is pinned bit for bit to one ``RandMeth`` per
realization and to numpy's own draws (``tests/test_sampling_batch.py``).
Here the device sampler (one wavefront per chain, MT19937 in LDS, the
closed-form log-pdf on the device) must give the same modes bit for bit:
for every closed-form model family, forced MCMC in 1D / 2D, odd and 64
walker counts, the reference's 3D MCMC known values
(``tests/test_randmeth.py:43-46``, ``tests/test_incomprrandmeth.py:40-44``)
and the 128 MCMC realizations of the 256-realization batch.
"""

import sys
import time

import numpy as np
import pytest
import torch

import oracle
from gstools_amd import Exponential, Gaussian, IncomprRandMeth, Matern, RandMeth, Stable
from gstools_amd.rng import RNG
from gstools_amd.sampling import device_model, sample_ln_pdf_batch, sample_modes_batch

pytestmark = pytest.mark.gpu

SEED = 19031977


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        for u, v in zip(x, y):
            assert np.array_equal(u, v)


CASES = [
    ("exponential3d", lambda: Exponential(dim=3, var=2.0, len_scale=[50, 20, 5], angles=[0.5, 0.2, 0.1]),
     [1000, 300, 64, 8192], [11, 22, 44, 45]),
    ("gaussian3d", lambda: Gaussian(dim=3, var=1.0, len_scale=10.0), [500, 100, 7], [5, 6, 7]),
    ("matern3d", lambda: Matern(dim=3, var=1.0, len_scale=8.0, nu=1.5), [256, 4096], [7, 8]),
    ("matern2d", lambda: Matern(dim=2, var=1.0, len_scale=3.0, nu=0.7), [300], [9]),
    ("exponential1d", lambda: Exponential(dim=1, var=1.0, len_scale=2.0), [200], [3]),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_device_modes_equal_host(case):
    name, make, mode_nos, seeds = CASES[case]
    _say(f"device sampling {name} N={mode_nos}")
    model = make()
    sampling = "mcmc" if model.dim == 1 else "auto"
    host = sample_modes_batch(model, mode_nos, seeds, sampling=sampling, device=False)
    dev = sample_modes_batch(model, mode_nos, seeds, sampling=sampling, device=True)
    _same(host, dev)
    for (cov, _, _), n, s in zip(dev[:2], mode_nos, seeds):
        assert np.array_equal(cov, RandMeth(model, mode_no=n, seed=s, sampling=sampling)._cov_sample)


@pytest.mark.parametrize("nw", [7, 64, 2])
def test_device_walker_counts(nw):
    """Odd, full-wave and minimal walker counts through ``sample_ln_pdf_batch``."""
    _say(f"device sampling nwalkers={nw}")
    model = Gaussian(dim=2, var=1.0, len_scale=4.0)
    md = device_model(model)
    sizes = [40, 25, 333]
    host = sample_ln_pdf_batch([RNG(21), RNG(22), RNG(23)], model.ln_spectral_rad_pdf, sizes, sample_around=0.25,
                               nwalkers=nw, burn_in=5)
    dev = sample_ln_pdf_batch([RNG(21), RNG(22), RNG(23)], model.ln_spectral_rad_pdf, sizes, sample_around=0.25,
                              nwalkers=nw, burn_in=5, device_model=md)
    for a, b in zip(host, dev):
        assert np.array_equal(a, b)


def test_device_reference_known_values():
    """3D MCMC values of ``tests/test_randmeth.py:43-46`` and ``tests/test_incomprrandmeth.py:40-44``."""
    _say("device sampling: reference known values")
    pos = np.array([np.linspace(0.0, 10.0, 10), np.linspace(-5.0, 5.0, 10), np.linspace(-6.0, 8.0, 10)])
    model = Gaussian(dim=3, var=1.5, len_scale=3.5)
    (cov, z1, z2), _ = sample_modes_batch(model, [100, 200], [SEED, SEED + 1], device=True)
    f = np.sqrt(1.5 / 100) * oracle.summate_c(cov, z1, z2, pos)
    assert f[0] == pytest.approx(1.3240234883187239, abs=1e-12)
    assert f[1] == pytest.approx(1.6367244277732766, abs=1e-12)
    m3 = Gaussian(dim=3, var=1.5, len_scale=2.5)
    (cov, z1, z2), = sample_modes_batch(m3, [100], [SEED], device=True)
    assert np.array_equal(cov, IncomprRandMeth(m3, mode_no=100, seed=SEED)._cov_sample)
    v = oracle.summate_incompr_c(cov, z1, z2, pos) * np.sqrt(1.5 / 100)
    v[0] += 1.0
    assert [v[0, 0], v[0, 1], v[1, 0]] == pytest.approx([0.7924546333550331, 1.660747056686244,
                                                         -0.28049855754819514], abs=1e-12)


def test_device_sampler_errors_and_fallback():
    """Models without a device log-pdf keep the host sampler; asking for the
    device with one is an error; a nu > 20 Matern has no device form."""
    assert device_model(Stable(dim=2, len_scale=5.0)) is None
    assert device_model(Matern(dim=3, nu=25.0)) is None
    with pytest.raises(ValueError):
        sample_modes_batch(Stable(dim=2, len_scale=5.0, density="series"), [50], [1], device=True)
    got = sample_modes_batch(Stable(dim=2, len_scale=5.0, density="series"), [50], [1])  # host path
    assert np.array_equal(got[0][0], RandMeth(Stable(dim=2, len_scale=5.0, density="series"), mode_no=50,
                                              seed=1)._cov_sample)


def test_device_batch_workload():
    """The 128 MCMC realizations of the 256-realization batch: device = host, bit for bit."""
    from bench import _batch_specs, _model_of

    specs = _batch_specs()
    total_dev = total_host = 0.0
    for kind in ("points3d", "incompr3d"):
        idx = [i for i, sp in enumerate(specs) if sp[0] == kind]
        ns, ss = [specs[i][1] for i in idx], [specs[i][2] for i in idx]
        _say(f"device sampling: batch workload {kind}, {len(idx)} realizations")
        model = _model_of(kind)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dev = sample_modes_batch(model, ns, ss, device=True)
        total_dev += time.perf_counter() - t0
        t0 = time.perf_counter()
        host = sample_modes_batch(model, ns, ss, device=False, nthreads=16)
        total_host += time.perf_counter() - t0
        _same(host, dev)
    _say(f"  device {total_dev:.3f} s, host lockstep {total_host:.3f} s")


def test_device_many_models_one_launch():
    """Groups of different models in one chains launch: each group's modes
    equal its own host sampling bit for bit (the batch workload's three kinds)."""
    from bench import _batch_specs, _model_of
    from gstools_amd.sampling import sample_modes_many

    specs = _batch_specs()
    kinds = ("grid2d", "points3d", "incompr3d")
    groups = []
    for kind in kinds:
        idx = [i for i, sp in enumerate(specs) if sp[0] == kind][:20]
        groups.append((_model_of(kind), [specs[i][1] for i in idx], [specs[i][2] for i in idx]))
    groups.append((Matern(dim=3, var=1.0, len_scale=8.0, nu=1.5), [300, 50], [1, 2]))
    _say("device sampling: four groups, one launch")
    dev = sample_modes_many(groups, device=None)
    for (model, ns, ss), got in zip(groups, dev):
        _same(sample_modes_batch(model, ns, ss, device=False), got)


def test_sample_and_sum_overlapped_equals_sequential():
    """``batch.sample_and_sum`` (each group's sums launched as soon as its modes
    are final, beside the device chains): the same modes as ``sample_modes_many``
    and the same outputs, bit for bit, as one MixedBatch after the sampling."""
    from bench import _batch_specs, _model_of
    from gstools_amd.batch import BatchItem, MixedBatch, sample_and_sum
    from gstools_amd.sampling import sample_modes_many

    _say("sample_and_sum: overlapped vs sequential")
    dev = torch.device("cuda", 0)
    specs = _batch_specs()
    kinds = ("grid2d", "points3d", "incompr3d")
    idx = {k: [i for i, sp in enumerate(specs) if sp[0] == k][:12] for k in kinds}
    groups = [(_model_of(k), [specs[i][1] for i in idx[k]], [specs[i][2] for i in idx[k]]) for k in kinds]
    rs = np.random.RandomState(3)
    geo = {}
    for k in kinds:
        for i in idx[k]:
            size = min(specs[i][3], 40 if k == "incompr3d" else (300 if k == "grid2d" else 20000))
            if k == "points3d":
                geo[i] = torch.from_numpy(rs.uniform(0, 100, size=(3, size))).to(dev)
            else:
                geo[i] = [torch.arange(size, dtype=torch.float64, device=dev)] * (2 if k == "grid2d" else 3)

    def items_of(g, modes):
        k = kinds[g]
        if k == "points3d":
            return [BatchItem(*m, pos=geo[i]) for i, m in zip(idx[k], modes)]
        return [BatchItem(*m, axes=geo[i], incompr=k == "incompr3d") for i, m in zip(idx[k], modes)]

    seen = []
    ref_modes = sample_modes_many(groups, on_ready=lambda i, m: seen.append(i))
    assert sorted(seen) == [0, 1, 2]
    seq = [it for g in range(3) for it in items_of(g, ref_modes[g])]
    MixedBatch(seq, "bf16x3", tune=False).launch()
    torch.cuda.synchronize()
    modes, items, keep = sample_and_sum(groups, items_of, tune=False)
    torch.cuda.synchronize()
    _same([m for g in modes for m in g], [m for g in ref_modes for m in g])
    got = [it for g in items for it in g]
    assert len(got) == len(seq)
    for a, b in zip(got, seq):
        assert torch.equal(a.out, b.out)


@pytest.mark.parametrize("chains_first", [True, False])
def test_many_groups_ready_order(monkeypatch, chains_first):
    """Both orders of ``sample_modes_many``'s finish (device groups read back
    before the host groups when their chains have run, or after the last) give
    the same modes, and ``on_ready`` sees every group once, in that order."""
    from gstools_amd import sampling as smp

    _say(f"ready order chains_first={chains_first}")
    groups = [(Gaussian(dim=2, var=1.0, len_scale=10.0), [300, 50], [1, 2]),          # host (inversion)
              (Exponential(dim=3, var=2.0, len_scale=5.0), [64, 200], [5, 6]),        # device chains
              (Gaussian(dim=3, var=1.0, len_scale=10.0), [], []),                     # empty
              (Matern(dim=3, var=1.0, len_scale=8.0, nu=1.5), [90], [7])]             # device chains
    ref = smp.sample_modes_many(groups, device=False)
    monkeypatch.setattr(smp._DeviceChains, "finished", lambda self: chains_first)
    seen = []
    got = smp.sample_modes_many(groups, on_ready=lambda i, m: seen.append(i))
    assert sorted(seen) == [0, 1, 2, 3]
    assert seen == ([1, 3, 0, 2] if chains_first else [0, 2, 1, 3])
    for a, b in zip(ref, got):
        _same(a, b)
