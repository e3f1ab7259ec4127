"""The differentiation and batching configurations at their own sizes, on the GPU.

* 2D Stable(alpha 1.5, len 15), 2048 MCMC modes (seed 20170519), 1e6 points
  uniform in [0, 500]^2, FP32: forward against the oracle; the FP32 adjoint
  gradients of the mean squared misfit at 2000 observation points (targets:
  the same model under seed 20170520) w.r.t. z_1, z_2, the wavenumbers and
  len_scale against torch float64 autograd, and the FP64 adjoint against
  central finite differences (h = 1e-4, FP64); the dense adjoint over all
  1e6 points in FP32 against FP64 and the oracle.
* 256 realizations (128 2D Gaussian grids, 96 3D Exponential point sets, 32
  incompressible 3D Gaussian 128^3 grids; mode_no in {1000, 2048, 4096,
  8192}; seeds from MasterRNG(20170519)) in one persistent plan replayed as
  a HIP graph: every item bit for bit equal to its own launch, spot rows of
  every item against the oracle, and the matrix-core MixedBatch within
  rel_rms 1e-5 of the oracle.

The forward formula is ``src/gstools/field/generator.py:193-199``; the modes
come from the restated sampler (``generator.py:346-387``, ``rng.py:38-104``).
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from gstools_amd import RandMeth, Stable, summate_adjoint_device, summate_device  # noqa: E402
from gstools_amd.autograd import randmeth_field  # noqa: E402

pytestmark = pytest.mark.gpu

N_C4, P_C4, OBS_C4 = 2048, 1_000_000, 2000


def _rel_max(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def c4():
    # the series density: the reference's truncated quadrature (the default)
    # draws radii up to ~1e10 for this model, phases no precision resolves
    model = Stable(dim=2, var=1.0, len_scale=15.0, alpha=1.5, density="series")
    rm = RandMeth(model, mode_no=N_C4, seed=20170519)
    target_rm = RandMeth(model, mode_no=N_C4, seed=20170520)
    pos = np.random.RandomState(19970222).uniform(0.0, 500.0, size=(2, P_C4))
    obs = np.random.RandomState(19970223).uniform(0.0, 500.0, size=(2, OBS_C4))
    target = target_rm(obs)
    return model, rm, pos, obs, target


def test_c4_forward_fp32_against_oracle(c4):
    model, rm, pos, _, _ = c4
    print(f"C4 forward: N={N_C4} P={P_C4}", flush=True)
    dev = torch.device("cuda", 0)
    cov, z1, z2 = (torch.from_numpy(a).to(dev) for a in (rm._cov_sample, rm._z_1, rm._z_2))
    post = torch.from_numpy(pos).to(dev)
    u32 = summate_device(cov, z1, z2, post, "fp32").cpu().numpy().astype(np.float64)
    u64 = summate_device(cov, z1, z2, post, "fp64").cpu().numpy()
    sub = np.random.RandomState(1).choice(P_C4, 4096, replace=False)
    ref = oracle.summate_c(rm._cov_sample, rm._z_1, rm._z_2, np.ascontiguousarray(pos[:, sub]))
    assert oracle.rel_errors(u64[sub], ref)[0] <= 1e-12
    rmax, rrms = oracle.rel_errors(u32[sub], ref)
    print(f"C4 fp32 vs oracle: rel_max {rmax:.2e} rel_rms {rrms:.2e}", flush=True)
    assert rmax <= 1e-5 and rrms <= 1e-5
    assert oracle.rel_errors(u32, u64)[0] <= 1e-5


def _torch_field(z1, z2, k_unit, pos, var, len_scale):
    """float64 torch restatement of sqrt(var / N) sum z1 cos(k.x) + z2 sin(k.x)."""
    ph = (k_unit / len_scale).transpose(0, 1).contiguous() @ pos  # (N, M)
    u = (z1[:, None] * torch.cos(ph) + z2[:, None] * torch.sin(ph)).sum(0)
    return torch.sqrt(var / z1.shape[0]) * u


def _misfit_grads(field_fn, z1, z2, k_unit, length, pos, target):
    leaves = [t.clone().requires_grad_(True) for t in (z1, z2, k_unit, length)]
    u = field_fn(*leaves, pos)
    loss = torch.mean((u.to(torch.float64) - target) ** 2)
    loss.backward()
    return [t.grad.detach().cpu().numpy() for t in leaves]


def test_c4_adjoint_gradients(c4):
    """FP32 adjoint vs float64 autograd; FP64 adjoint vs central FD (h = 1e-4)."""
    model, rm, _, obs, target = c4
    print(f"C4 misfit gradients: N={N_C4} obs={OBS_C4}", flush=True)
    dev = torch.device("cuda", 0)
    f64 = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    length = float(model.len_rescaled)
    z1, z2 = f64(rm._z_1), f64(rm._z_2)
    k_unit = f64(rm._cov_sample * length)  # k = k_unit / len_scale
    ell = torch.tensor(length, dtype=torch.float64, device=dev)
    pobs, tgt = f64(obs), f64(target)
    var = torch.as_tensor(model.var, dtype=torch.float64, device=dev)

    def kernel_field(precision):
        return lambda a, b, k, L, p: randmeth_field(a, b, k, p, var=model.var, len_scale=L, precision=precision)

    g_ref = _misfit_grads(lambda a, b, k, L, p: _torch_field(a, b, k, p, var, L), z1, z2, k_unit, ell, pobs, tgt)
    g32 = _misfit_grads(kernel_field("fp32"), z1, z2, k_unit, ell, pobs, tgt)
    g64 = _misfit_grads(kernel_field("fp64"), z1, z2, k_unit, ell, pobs, tgt)
    for name, r, a, b in zip(("z1", "z2", "k", "len"), g_ref, g32, g64):
        e32, e64 = _rel_max(a, r), _rel_max(b, r)
        print(f"C4 grad {name}: fp32 vs f64 autograd {e32:.2e}, fp64 vs autograd {e64:.2e}", flush=True)
        assert e64 <= 1e-10, name
        assert e32 <= 1e-5, name

    # central differences of the FP64 kernel forward at h = 1e-4, bounded by
    # their own truncation estimate (from h / 2: the wavenumbers reach |k| ~ 5,
    # so d^3 L / d len^3 is large and the len difference is off by ~1e-5),
    # and the Richardson value (4 D(h/2) - D(h)) / 3, whose h^2 term cancels
    def loss(a, b, k, L):
        u = randmeth_field(a, b, k, pobs, var=model.var, len_scale=L, precision="fp64")
        return float(torch.mean((u - tgt) ** 2))

    def central(which, idx, step):
        args_p = [z1.clone(), z2.clone(), k_unit.clone(), ell.clone()]
        args_m = [z1.clone(), z2.clone(), k_unit.clone(), ell.clone()]
        if idx is None:
            args_p[which] += step
            args_m[which] -= step
        else:
            args_p[which][idx] += step
            args_m[which][idx] -= step
        return (loss(*args_p) - loss(*args_m)) / (2 * step)

    h = 1e-4
    rs = np.random.RandomState(2)
    checks = [(0, j) for j in rs.choice(N_C4, 4, replace=False)] + [(1, j) for j in rs.choice(N_C4, 4, replace=False)]
    checks += [(2, (d, j)) for d in (0, 1) for j in rs.choice(N_C4, 3, replace=False)] + [(3, None)]
    for which, idx in checks:
        fd, fd2 = central(which, idx, h), central(which, idx, h / 2)
        trunc = abs(fd - fd2) * 4.0 / 3.0
        rich = (4.0 * fd2 - fd) / 3.0
        ad64 = float(g64[which] if idx is None else g64[which][idx])
        ad32 = float(g32[which] if idx is None else g32[which][idx])
        scale = float(np.max(np.abs(g_ref[which])))
        print(f"C4 FD {which} {idx}: ad64 {ad64:.10e} fd {fd:.10e} (trunc ~{trunc:.1e}) richardson {rich:.10e}",
              flush=True)
        assert abs(ad64 - fd) <= 1.5 * trunc + 1e-8 * scale, (which, idx, ad64, fd, trunc)
        assert abs(ad64 - rich) <= 1e-7 * scale + 1e-12, (which, idx, ad64, rich)
        assert abs(ad32 - rich) <= 1e-5 * scale + 1e-12, (which, idx, ad32, rich)


def test_c4_dense_adjoint_full_size(c4):
    """Adjoint over all 1e6 points with a dense upstream gradient (the timed
    form): FP32 against FP64 on the GPU, FP64 against the oracle on a subset."""
    _, rm, pos, _, _ = c4
    print(f"C4 dense adjoint: N={N_C4} P={P_C4}", flush=True)
    dev = torch.device("cuda", 0)
    cov, z1, z2 = (torch.from_numpy(a).to(dev) for a in (rm._cov_sample, rm._z_1, rm._z_2))
    post = torch.from_numpy(pos).to(dev)
    g = np.random.RandomState(3).normal(size=P_C4).astype(np.float32)
    a32 = [t.cpu().numpy() for t in summate_adjoint_device(cov, z1, z2, post, torch.from_numpy(g).to(dev), "fp32")]
    a64 = [t.cpu().numpy() for t in summate_adjoint_device(cov, z1, z2, post, g.astype(np.float64), "fp64")]
    for name, a, b in zip(("gz1", "gz2", "gk"), a32, a64):
        e = _rel_max(a, b)
        print(f"C4 dense adjoint {name}: fp32 vs fp64 {e:.2e}", flush=True)
        assert e <= 1e-5, name
    sub = np.arange(0, P_C4, 97)
    gs = g[sub].astype(np.float64)
    s64 = [t.cpu().numpy() for t in summate_adjoint_device(cov, z1, z2, post[:, sub].contiguous(), gs, "fp64")]
    ref = oracle.summate_adjoint_c(rm._cov_sample, rm._z_1, rm._z_2, np.ascontiguousarray(pos[:, sub]), gs)
    for a, b in zip(s64, ref):
        assert _rel_max(a, b) <= 1e-12


# --------------------------------------------------------------------- C5 --

@pytest.fixture(scope="module")
def c5():
    """The 256 items with their modes, built as the benchmark builds them."""
    from bench import _batch_specs, _model_of
    from gstools_amd.batch import BatchItem
    from gstools_amd.sampling import sample_modes_batch

    specs = _batch_specs()
    modes = [None] * len(specs)
    for kind in ("grid2d", "points3d", "incompr3d"):
        idx = [i for i, sp in enumerate(specs) if sp[0] == kind]
        got = sample_modes_batch(_model_of(kind), [specs[i][1] for i in idx], [specs[i][2] for i in idx])
        for i, m in zip(idx, got):
            modes[i] = m
    dev = torch.device("cuda", 0)
    expo = _model_of("points3d")
    items = []
    for (kind, n, seed, size), (cov, z1, z2) in zip(specs, modes):
        covt, z1t, z2t = (torch.from_numpy(a).to(dev) for a in (cov, z1, z2))
        if kind == "grid2d":
            items.append(BatchItem(covt, z1t, z2t, axes=[torch.arange(size, dtype=torch.float64, device=dev)] * 2))
        elif kind == "points3d":
            pos = np.random.RandomState(seed).uniform(0.0, 1000.0, size=(3, size))
            items.append(BatchItem(covt, z1t, z2t, pos=np.ascontiguousarray(expo.isometrize(pos))))
        else:
            items.append(BatchItem(covt, z1t, z2t, axes=[torch.arange(size, dtype=torch.float64, device=dev)] * 3,
                                   incompr=True))
    return specs, modes, items


def _item_points(it, idx):
    """(d, len(idx)) FP64 host coordinates of an item's points ``idx`` (numpy)."""
    if it.axes is None:
        return it.pos[:, torch.as_tensor(idx, device=it.pos.device)].cpu().numpy()
    ax = [a.cpu().numpy() for a in torch.split(it.axes, list(it.shape[:it.d]))]
    sub = np.unravel_index(idx, tuple(it.shape[:it.d]))
    return np.stack([ax[i][sub[i]] for i in range(it.d)])


def _spot_rows(it, count, seed):
    rows = np.random.RandomState(seed).choice(it.P, min(count, it.P), replace=False)
    return np.unique(np.concatenate([rows, [0, it.P - 1]]))


def _oracle_rows(it, cov, z1, z2, rows):
    p = np.ascontiguousarray(_item_points(it, rows))
    return oracle.summate_incompr_c(cov, z1, z2, p) if it.incompr else oracle.summate_c(cov, z1, z2, p)[None]


def test_c5_modes_match_one_realization_at_a_time(c5):
    specs, modes, _ = c5
    from bench import _model_of

    for r in (0, 130, 250):
        kind, n, seed, _ = specs[r]
        rm = RandMeth(_model_of(kind), mode_no=n, seed=seed)
        for a, b in zip(modes[r], (rm._cov_sample, rm._z_1, rm._z_2)):
            assert np.array_equal(a, b)


def test_c5_persistent_graph_equals_single_launches(c5):
    from gstools_amd import summate_grid_device, summate_incompr_device
    from gstools_amd.batch import BatchPlan

    specs, modes, items = c5
    work = sum(it.P * it.N for it in items)
    print(f"C5 batch: {len(items)} items, {work:.3e} mode.points", flush=True)
    plan = BatchPlan(items, 1024, 256)
    plan.launch(graph=True)
    torch.cuda.synchronize()
    batched = [it.out.clone() for it in items]
    plan.launch(graph=True)  # a second replay gives the same bits
    torch.cuda.synchronize()
    assert all(torch.equal(it.out, b) for it, b in zip(items, batched))
    plan.close()
    meshes = {}  # incompressible items run on the materialised meshgrid, one per grid shape
    for r, it in enumerate(items):
        if it.axes is not None:
            axes = list(torch.split(it.axes, list(it.shape[:it.d])))
            if it.incompr:
                key = tuple(it.shape[:it.d])
                if key not in meshes:
                    g = torch.meshgrid(*axes, indexing="ij")
                    meshes[key] = torch.stack([a.reshape(-1) for a in g]).contiguous()
                single = summate_incompr_device(it.cov, it.z1, it.z2, meshes[key], "fp32")
            else:
                single = summate_grid_device(it.cov, it.z1, it.z2, axes, "fp32").reshape(batched[r].shape)
        else:
            single = summate_device(it.cov, it.z1, it.z2, it.pos, "fp32").reshape(batched[r].shape)
        assert torch.equal(single, batched[r]), f"item {r} ({specs[r][0]}) differs from its single launch"
    worst = 0.0
    for r, (it, (cov, z1, z2)) in enumerate(zip(items, modes)):
        rows = _spot_rows(it, 24, r)
        ref = _oracle_rows(it, cov, z1, z2, rows)
        got = batched[r][:, rows].cpu().numpy().astype(np.float64)
        for c in range(got.shape[0]):
            err = float(np.max(np.abs(got[c] - ref[c])) / max(np.max(np.abs(ref[c])), 1e-300))
            worst = max(worst, err)
            assert err <= 1e-5, f"item {r} ({specs[r][0]}) component {c}: {err:.2e}"
    print(f"C5 fp32 batch: worst spot rel_max vs oracle {worst:.2e}", flush=True)


def test_c5_mixed_batch_matrix_cores(c5):
    from gstools_amd.batch import MixedBatch

    specs, modes, items = c5
    print(f"C5 mixed batch: {len(items)} items, grids on the matrix cores (bf16x3)", flush=True)
    mb = MixedBatch(items, "bf16x3", 1024, 256)
    mb.launch(graph=True)
    torch.cuda.synchronize()
    first = [it.out.clone() for it in items]
    mb.launch(graph=True)
    torch.cuda.synchronize()
    assert all(torch.equal(it.out, b) for it, b in zip(items, first))
    worst = 0.0
    for r, (it, (cov, z1, z2)) in enumerate(zip(items, modes)):
        rows = _spot_rows(it, 256, 1000 + r)
        ref = _oracle_rows(it, cov, z1, z2, rows)
        got = first[r][:, rows].cpu().numpy().astype(np.float64)
        for c in range(got.shape[0]):
            rrms = oracle.rel_errors(got[c], ref[c])[1]
            worst = max(worst, rrms)
            assert rrms <= 1e-5, f"item {r} ({specs[r][0]}) component {c}: rel_rms {rrms:.2e}"
    print(f"C5 mixed batch: worst spot rel_rms vs oracle {worst:.2e}", flush=True)
