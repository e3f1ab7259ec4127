"""Spectral mode sampling for many realizations at once.

``RandMeth.reset_seed`` (``src/gstools/field/generator.py:346-387``) draws
``z_1``, ``z_2``, unit directions and -- for models without a closed-form
inverse cdf -- the radii by ensemble MCMC (``RNG.sample_ln_pdf``,
``src/gstools/random/rng.py:38-104``, emcee's stretch move).  One MCMC run
is ~``size / 5`` steps of two half-moves with eight small draws each, so
sampling 8192 modes spends ~0.5 s in interpreter overhead per realization.

:func:`sample_modes_batch` gives, for a list of ``(mode_no, seed)`` pairs
sharing one model, exactly the arrays ``RandMeth(model, mode_no=n,
seed=s)`` holds (``_cov_sample``, ``_z_1``, ``_z_2``), bit for bit:

* every draw of every MCMC step comes from ``gsx_stretch_draws`` (a native
  restatement of the MT19937 stream and numpy's legacy draw routines, run
  for all chains in one call per block of steps);
* the chains then advance in lockstep: per half-move one native call forms
  the proposals of all chains (``gsx_stretch_propose``), the model's own
  ``ln_spectral_rad_pdf`` is evaluated once on all of them, and one native
  call applies the acceptance tests (``gsx_stretch_accept``) -- the
  arithmetic of the one-chain sampler (``mcmc.py``), rounded the same way,
  with the log of the acceptance uniforms taken by numpy as there;
* the per-realization draws that are not MCMC (the master seeds, normals,
  directions, initial walkers, the final ``choice``) stay on numpy.

For the closed-form models (Gaussian, Exponential, Matern with nu <= 20)
the MCMC runs on the GPU instead (``device=True``, the default when a HIP
device is present): ``gsx_stretch_run_device`` advances every chain in one
launch, one block of two wavefronts per chain (the stream wave draws the
next step while the walker wave applies the current one), drawing from the
chain's own MT19937 streams in the same order and evaluating the
closed-form log-pdf on the device; the final ``choice`` indices come from
numpy and are gathered on the device (``gsx_stretch_pick_device``).  Accept decisions compare the same
quantities up to the last bits of exp / log, so the modes come out bit for
bit equal to the host sampler's unless a test lands within an ulp.
"""

import ctypes
import os

import numpy as np
import torch

from . import _native
from .rng import RNG

__all__ = ["sample_ln_pdf_batch", "sample_modes_batch", "sample_modes_many", "device_model"]

_MT_WORDS = 625  # 624 key words + position


def _mt_state(random_state):
    """MT19937 state of a legacy ``RandomState`` as the (625,) uint32 row the draws use."""
    name, key, pos = random_state.get_state()[:3]
    if name != "MT19937":
        raise ValueError(f"sample_ln_pdf_batch: unsupported bit generator {name}")
    row = np.empty(_MT_WORDS, dtype=np.uint32)
    row[:624] = key
    row[624] = pos
    return row


def _vp(a):
    return ctypes.c_void_p(a.ctypes.data)


class _Draws:
    """Blocks of pre-drawn step randomness for R chains."""

    def __init__(self, states, steps, nwalkers, block, nthreads):
        r = states.shape[0]
        self.states = states
        self.steps = np.asarray(steps, dtype=np.int64)
        self.nw = int(nwalkers)
        self.block = int(block)
        self.nthreads = int(nthreads)
        shape = (r, self.block, self.nw)
        self.labels = np.zeros(shape, dtype=np.int8)
        self.uz = np.zeros(shape)
        self.partner = np.zeros(shape, dtype=np.int64)
        self.uacc = np.zeros(shape)
        self.log_uacc = np.zeros(shape)
        self.start = -1

    def load(self, s0):
        todo = np.clip(self.steps - s0, 0, self.block).astype(np.int64)
        st = _native.lib().gsx_stretch_draws(self.states.shape[0], _vp(self.states), _vp(todo), self.block,
                                               self.nw, _vp(self.labels), _vp(self.uz), _vp(self.partner),
                                               _vp(self.uacc), self.nthreads)
        _native.check(st, "stretch_draws")
        # the acceptance test compares with numpy's log of the uniforms
        with np.errstate(divide="ignore"):
            np.log(self.uacc, out=self.log_uacc)
        self.start = s0


def _log_prob(ln_pdf, q):
    """Vectorised ``ln_pdf`` on proposals (R, ns), with the sampler's input checks."""
    if np.any(np.isinf(q)):
        raise ValueError("At least one parameter value was infinite")
    if np.any(np.isnan(q)):
        raise ValueError("At least one parameter value was NaN")
    flat = np.ascontiguousarray(q).reshape(-1, 1)
    res = np.asarray(ln_pdf(flat), dtype=np.double)
    lp = res.reshape(flat.shape[0], -1)[:, 0].reshape(q.shape).copy()
    if np.any(np.isnan(lp)):
        raise ValueError("Probability function returned NaN")
    return lp


def _lockstep(ln_pdf, coords, log_prob, steps, states, a=2.0, record=False, block=64, nthreads=1):
    """Advance R chains of ``nwalkers`` one-dimensional walkers; chain r runs steps[r] steps.

    Per chain this is ``StretchEnsemble.run`` (``mcmc.py``) step for step.
    Returns (coords, log_prob, chain) with chain (R, max_steps, nwalkers) when
    ``record`` (rows past steps[r] unused).
    """
    lib = _native.lib()
    nr, nw = coords.shape
    states = np.ascontiguousarray(states, dtype=np.uint32)
    coords = np.ascontiguousarray(coords, dtype=np.double)
    log_prob = np.ascontiguousarray(log_prob, dtype=np.double)
    steps = np.ascontiguousarray(steps, dtype=np.int64)
    max_steps = int(steps.max(initial=0))
    chain = np.empty((nr, max_steps, nw)) if record else None
    draws = _Draws(states, steps, nw, block, nthreads)
    n0 = (nw + 1) // 2
    q = np.empty(nr * n0)
    new_lp = np.empty(nr * n0)
    n_act = ctypes.c_int64(0)
    p_steps, p_coords, p_lp, p_q, p_new = _vp(steps), _vp(coords), _vp(log_prob), _vp(q), _vp(new_lp)
    p_lab, p_uz, p_par, p_logu = _vp(draws.labels), _vp(draws.uz), _vp(draws.partner), _vp(draws.log_uacc)
    chain_p = (_vp(chain) if record else None, None)
    propose, accept = lib.gsx_stretch_propose, lib.gsx_stretch_accept
    a = float(a)
    for s in range(max_steps):
        if s % block == 0:
            draws.load(s)
        j = s - draws.start
        for split in range(2):
            ns = n0 if split == 0 else nw - n0
            st = propose(nr, nw, p_steps, s, j, block, split, a, p_coords, p_lab, p_uz, p_par, p_q,
                         ctypes.byref(n_act))
            if st:
                _native.check(st, "stretch_propose")
            k = int(n_act.value)
            new_lp[:k * ns] = _log_prob(ln_pdf, q[:k * ns].reshape(k, ns)).reshape(-1)
            st = accept(nr, nw, p_steps, s, j, block, split, p_lab, p_logu, p_q, p_new, p_coords, p_lp,
                        chain_p[split - 1], max_steps)
            if st:
                _native.check(st, "stretch_accept")
    return coords, log_prob, chain


def device_model(model):
    """``gsx_stretch_model`` for this package's Gaussian, Exponential and Matern
    (nu <= 20) models, with the scalar factors computed exactly as their
    ``spectral_density`` computes them; None for any other model."""
    import scipy.special as sps

    from . import covmodel as cm

    kind = type(model)
    if kind not in (cm.Gaussian, cm.Exponential, cm.Matern) or not 1 <= model.dim <= 3:
        return None
    md = _native.StretchModel()
    md.dim, md.len_rescaled = model.dim, float(model.len_rescaled)
    if kind is cm.Gaussian:
        md.kind = _native.STRETCH_GAUSSIAN
        md.c[0] = (model.len_rescaled / 2.0 / np.sqrt(np.pi)) ** model.dim
    elif kind is cm.Exponential:
        md.kind = _native.STRETCH_EXPONENTIAL
        md.c[0] = model.len_rescaled**model.dim * sps.gamma((model.dim + 1) / 2.0)
    else:
        if model.nu > 20.0:
            return None
        md.kind = _native.STRETCH_MATERN
        md.c[0] = (model.len_rescaled / np.sqrt(np.pi)) ** model.dim
        md.c[1] = -(model.nu + model.dim / 2.0)
        md.c[2] = model.nu
        md.c[3] = sps.loggamma(model.nu + model.dim / 2.0)
        md.c[4] = sps.loggamma(model.nu)
        md.c[5] = model.dim * np.log(np.sqrt(model.nu))
    return md


_STRETCH_ERRORS = {1: "At least one parameter value was infinite or NaN",
                   2: "Probability function returned NaN"}


class _DeviceChains:
    """The chains of one ``gsx_stretch_run_device`` launch (running on the
    current stream; :meth:`pick` waits for it)."""

    def __init__(self, models, init, burn, prod, burn_in, prod_steps, a=2.0):
        """``models``: one ``gsx_stretch_model`` for all chains or one per chain."""
        dev = torch.device("cuda", torch.cuda.current_device())
        nr, self.nw = init.shape
        if isinstance(models, _native.StretchModel):
            models = [models] * nr
        table = (_native.StretchModel * nr)(*models)
        self.max_steps = max(int(prod_steps.max(initial=0)), 1)
        states = np.ascontiguousarray(np.stack([burn, prod], axis=1), dtype=np.uint32)
        self._keep = [torch.from_numpy(states.view(np.int32)).to(dev),
                      torch.from_numpy(np.ascontiguousarray(prod_steps, dtype=np.int64)).to(dev),
                      torch.from_numpy(np.ascontiguousarray(init, dtype=np.double)).to(dev),
                      torch.frombuffer(bytearray(bytes(table)), dtype=torch.uint8).to(dev)]
        self.chain = torch.empty((nr, self.max_steps, self.nw), dtype=torch.float64, device=dev)
        self.status = torch.zeros(1, dtype=torch.int32, device=dev)
        st, steps_t, init_t, models_t = self._keep
        _native.check(_native.lib().gsx_stretch_run_device(nr, self.nw, _tp(st), int(burn_in), _tp(steps_t),
                                                           self.max_steps, float(a), _tp(models_t), _tp(init_t),
                                                           _tp(self.chain), None, _tp(self.status), _stream()),
                      "stretch_run_device")
        self._done = torch.cuda.Event()
        self._done.record()

    def finished(self):
        """True once the chains have run (does not wait)."""
        return self._done.query()

    def pick(self, picks):
        """Chain r's recorded walkers at the flat indices picks[r] -> list of arrays."""
        dev = self.chain.device
        counts = [len(p) for p in picks]
        owner = np.repeat(np.arange(len(picks), dtype=np.int64), counts)
        flat = np.concatenate(picks).astype(np.int64) if picks else np.zeros(0, dtype=np.int64)
        out = torch.empty(max(flat.size, 1), dtype=torch.float64, device=dev)
        owner_t, flat_t = torch.from_numpy(owner).to(dev), torch.from_numpy(flat).to(dev)
        _native.check(_native.lib().gsx_stretch_pick_device(flat.size, _tp(owner_t), _tp(flat_t), self.max_steps,
                                                            self.nw, _tp(self.chain), _tp(out), _stream()),
                      "stretch_pick_device")
        code = int(self.status.item())
        if code:
            raise ValueError(_STRETCH_ERRORS.get(code, f"device sampler status {code}"))
        rad = out[:flat.size].cpu().numpy()
        return np.split(rad, np.cumsum(counts)[:-1])


def _tp(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _chains_host(ln_pdf, init, burn, prod, burn_in, prod_steps, picks, block=64, nthreads=1):
    """The host lockstep sampler: chain r's recorded walkers at picks[r]."""
    nr = init.shape[0]
    log_prob = _log_prob(ln_pdf, init)
    if np.any(np.isnan(log_prob)):
        raise ValueError("The initial log_prob was NaN")
    coords, log_prob, _ = _lockstep(ln_pdf, init, log_prob, np.full(nr, burn_in, dtype=np.int64), burn,
                                    block=block, nthreads=nthreads)
    _, _, chain = _lockstep(ln_pdf, coords, log_prob, prod_steps, prod, record=True, block=block,
                            nthreads=nthreads)
    return [chain[r, :prod_steps[r]].reshape(-1)[picks[r]] for r in range(nr)]


def _prod_steps(sizes, nwalkers, burn_in, oversampling_factor):
    """Production steps of each run (``rng.py:88-101``)."""
    return np.array([int(burn_in if sz is None else max(burn_in, (sz / nwalkers) * oversampling_factor))
                     for sz in sizes], dtype=np.int64)


def sample_ln_pdf_batch(rngs, ln_pdf, sizes, sample_around=1.0, nwalkers=50, burn_in=20, oversampling_factor=10,
                        block=64, nthreads=1, device_model=None):
    """``RNG.sample_ln_pdf`` for many generators sharing one ``ln_pdf``.

    ``rngs`` are :class:`gstools_amd.rng.RNG` objects, each consumed exactly
    as ``rng.sample_ln_pdf(ln_pdf, size, sample_around, ...)`` would consume
    it (four master draws: initial walkers, burn-in stream, production stream,
    final choice).  Returns a list of radius arrays.  With ``device_model``
    (a ``gsx_stretch_model`` of the same distribution as ``ln_pdf``, see
    :func:`device_model`) and at most 64 walkers the chains run on the GPU.
    """
    nr = len(rngs)
    if nr == 0:
        return []
    sizes = list(sizes)
    if len(sizes) != nr:
        raise ValueError("sample_ln_pdf_batch: one size per generator")
    prod_steps = _prod_steps(sizes, nwalkers, burn_in, oversampling_factor)
    init = np.empty((nr, nwalkers))
    burn = np.empty((nr, _MT_WORDS), dtype=np.uint32)
    prod = np.empty((nr, _MT_WORDS), dtype=np.uint32)
    picks = []
    for r, rng in enumerate(rngs):
        init[r] = rng.random.rand(nwalkers) * sample_around
        burn[r] = _mt_state(rng.random)
        prod[r] = _mt_state(rng.random)
        # choice(chain, size) draws the indices choice(len(chain), size) does
        picks.append(rng.random.choice(int(prod_steps[r]) * nwalkers, sizes[r]))
    _check_init(init)
    if device_model is not None and nwalkers <= 64:
        return _DeviceChains(device_model, init, burn, prod, burn_in, prod_steps).pick(picks)
    return _chains_host(ln_pdf, init, burn, prod, burn_in, prod_steps, picks, block, nthreads)


def _check_init(init):
    for r in range(init.shape[0]):
        if np.all(init[r] == init[r, 0]):
            raise ValueError("Initial state has walkers that are not linearly independent")


def _sphere_chunk(dim, ang1, ang2, out):
    """``RNG.sample_sphere`` (``rng.py:142-193``) from its uniform draws, into out (dim, n)."""
    if dim == 1:
        out[0] = ang1
    elif dim == 2:
        np.cos(ang1, out=out[0])
        np.sin(ang1, out=out[1])
    else:
        rho = np.sqrt(1.0 - ang2**2)
        np.multiply(rho, np.cos(ang1, out=out[0]), out=out[0])
        np.multiply(rho, np.sin(ang1, out=out[1]), out=out[1])
        out[2] = ang2


_CHUNK = 1 << 15
_POOL = None


def _chunked(fn, n, threads):
    """Run fn(lo, hi) over [0, n) in chunks on a thread pool (numpy's elementwise
    loops release the GIL and give every element the same result in any chunk)."""
    global _POOL
    bounds = [(lo, min(n, lo + _CHUNK)) for lo in range(0, n, _CHUNK)]
    if threads <= 1 or len(bounds) <= 1:
        for lo, hi in bounds:
            fn(lo, hi)
        return
    if _POOL is None or _POOL._max_workers < threads:
        from concurrent.futures import ThreadPoolExecutor

        _POOL = ThreadPoolExecutor(max_workers=threads)
    list(_POOL.map(lambda b: fn(*b), bounds))


def _native_seed(seed):
    return isinstance(seed, (int, np.integer)) and not isinstance(seed, bool) and 0 <= int(seed) < 2**32


def sample_modes_batch(model, mode_nos, seeds, sampling="auto", nthreads=None, device=None, nwalkers=50,
                       burn_in=20, oversampling_factor=10):
    """``(cov_sample, z_1, z_2)`` of ``RandMeth(model, mode_no=n, seed=s)`` for every pair.

    Same draws, in the same order, as ``RandMeth.reset_seed``.  For integer
    seeds every numpy draw is made natively for all realizations at once
    (``gsx_mode_draws``, over ``nthreads`` host threads, default all); the
    MCMC runs of all realizations advance together, on the GPU when
    ``device`` is true (default: when a HIP device is present and the model
    has a device log-pdf, :func:`device_model`), while the host finishes the
    directions.  Other seeds take the one-generator-at-a-time draws.
    """
    return sample_modes_many([(model, mode_nos, seeds)], sampling=sampling, nthreads=nthreads, device=device,
                             nwalkers=nwalkers, burn_in=burn_in, oversampling_factor=oversampling_factor)[0]


def sample_modes_many(groups, sampling="auto", nthreads=None, device=None, nwalkers=50, burn_in=20,
                      oversampling_factor=10, on_ready=None):
    """:func:`sample_modes_batch` for several ``(model, mode_nos, seeds)`` groups:
    the MCMC chains of every group on the device run in one launch while the
    host draws and finishes all groups.  ``on_ready(i, modes)``, if given, is
    called for each group once its modes are final: the device groups'
    directions and then the host groups' modes are computed while the chains
    run; the host groups are then handed over in order, and the device
    groups are read back and handed over before the first host group still
    waiting once their chains have run (else after the last), so the
    caller's GPU work on any group starts as early as the chains and the host
    allow."""
    if sampling not in ("auto", "inversion", "mcmc"):
        raise ValueError("sample_modes_batch: sampling not in ['auto', 'inversion', 'mcmc']")
    threads = len(os.sched_getaffinity(0)) if nthreads is None else int(nthreads)
    # groups whose chains run on the device are drawn first, so that the
    # chains launch before the host draws the others
    on_device = [_kind(model, sampling) == 2 and _pick_device_model(model, 2, device) is not None
                 for model, _, _ in groups]
    preps = [None] * len(groups)
    for first in (True, False):
        for i, (model, mode_nos, seeds) in enumerate(groups):
            if on_device[i] == first:
                preps[i] = _prepare(model, mode_nos, seeds, sampling, device, nwalkers, burn_in,
                                    oversampling_factor, threads)
        if first:
            chains, dev_groups = _launch(preps, burn_in)
    out = [None] * len(preps)
    on_dev = [p is not None and p.get("md") is not None and p["n"].size > 0 for p in preps]
    # the directions of the device groups first, while the chains run
    for i, ((model, _, _), p) in enumerate(zip(groups, preps)):
        if on_dev[i]:
            out[i] = _finish(model, p)

    def device_groups():
        rads = chains.pick([pk for p in dev_groups for pk in p["picks"]])
        k = 0
        for i, p in enumerate(preps):
            if on_dev[i]:
                nr = p["init"].shape[0]
                out[i] = [(r * s, a, b) for r, (s, a, b) in zip(rads[k:k + nr], out[i])]
                k += nr
                if on_ready is not None:
                    on_ready(i, out[i])

    host = [i for i, p in enumerate(preps) if p is not None and not on_dev[i]]
    for i in host:
        model, p = groups[i][0], preps[i]
        if not p["n"].size:
            out[i] = []
            continue
        if p["kind"] == 2 and p["md"] is None:
            p["rads"] = _chains_host(model.ln_spectral_rad_pdf, p["init"], p["burn"], p["prod"], burn_in,
                                     p["prod_steps"], p["picks"], nthreads=threads)
        out[i] = _finish(model, p)
    pending = chains is not None
    for i in host:
        if pending and chains.finished():  # chains done: their groups before the next host group
            device_groups()
            pending = False
        if on_ready is not None:
            on_ready(i, out[i])
    if pending:
        device_groups()
    for i, ((model, mode_nos, seeds), p) in enumerate(zip(groups, preps)):
        if p is None:
            out[i] = _modes_python(model, [int(n) for n in mode_nos], list(seeds), _kind(model, sampling),
                                   _pick_device_model(model, _kind(model, sampling), device), nwalkers, burn_in,
                                   oversampling_factor, threads)
            if on_ready is not None:
                on_ready(i, out[i])
    return out


def _launch(preps, burn_in):
    """One chains launch for every prepared device group -> (chains, groups)."""
    dev_groups = [p for p in preps if p is not None and p.get("md") is not None and p["n"].size]
    if not dev_groups:
        return None, []
    init = np.concatenate([p["init"] for p in dev_groups])
    burn = np.concatenate([p["burn"] for p in dev_groups])
    prod = np.concatenate([p["prod"] for p in dev_groups])
    steps = np.concatenate([p["prod_steps"] for p in dev_groups])
    models = [p["md"] for p in dev_groups for _ in range(p["init"].shape[0])]
    return _DeviceChains(models, init, burn, prod, burn_in, steps), dev_groups


def _kind(model, sampling):
    if np.isclose(model.var, 0):
        return 0
    return 1 if sampling == "inversion" or (sampling == "auto" and model.has_ppf) else 2


def _pick_device_model(model, kind, device):
    if kind != 2 or device is False:
        return None
    md = device_model(model)
    if md is None and device:
        raise ValueError(f"sample_modes_batch: no device log-pdf for {type(model).__name__}")
    if md is not None and device is None and not torch.cuda.is_available():
        return None
    return md


def _prepare(model, mode_nos, seeds, sampling, device, nwalkers, burn_in, oversampling_factor, threads):
    """Native draws of one group (None: seeds the native draws do not take)."""
    mode_nos = [int(n) for n in mode_nos]
    seeds = list(seeds)
    if len(mode_nos) != len(seeds):
        raise ValueError("sample_modes_batch: one seed per mode_no")
    dim = model.dim
    kind = _kind(model, sampling)
    md = _pick_device_model(model, kind, device)
    if not all(_native_seed(x) for x in seeds) or dim > 3:
        return None
    nr = len(mode_nos)
    n = np.asarray(mode_nos, dtype=np.int64)
    total = int(n.sum())
    p = {"kind": kind, "dim": dim, "n": n, "md": md, "cuts": np.cumsum(n)[:-1], "threads": threads}
    p["z1"], p["z2"] = np.empty(total), np.empty(total)
    p["ang1"] = np.empty(total) if kind else None
    p["ang2"] = np.empty(total) if kind and dim == 3 else None
    p["u_inv"] = np.empty(total) if kind == 1 else None
    mc = kind == 2
    if mc:
        p["prod_steps"] = _prod_steps(mode_nos, nwalkers, burn_in, oversampling_factor)
        n_flat = p["prod_steps"] * nwalkers
        p["init"] = np.empty((nr, nwalkers))
        states = np.empty((nr, 2, _MT_WORDS), dtype=np.uint32)
        picks = np.empty(total, dtype=np.int64)
    else:
        n_flat = states = picks = None
    around = 1.0 / model.len_rescaled if mc else 0.0
    seeds_a = np.asarray([int(x) for x in seeds], dtype=np.int64)
    opt = lambda a: None if a is None else _vp(a)
    if nr:
        _native.check(_native.lib().gsx_mode_draws(nr, _vp(seeds_a), _vp(n), dim, kind, nwalkers, float(around),
                                                   opt(n_flat), _vp(p["z1"]), _vp(p["z2"]), opt(p["ang1"]),
                                                   opt(p["ang2"]), opt(p["u_inv"]), opt(p.get("init")), opt(states),
                                                   opt(picks), threads), "mode_draws")
    if mc:
        _check_init(p["init"])
        p["burn"], p["prod"] = np.ascontiguousarray(states[:, 0]), np.ascontiguousarray(states[:, 1])
        p["picks"] = np.split(picks, p["cuts"]) if nr else []
    return p


def _finish(model, p):
    """Modes of a prepared group; MCMC groups on the device still lack their radii
    (returned as (directions, z1, z2) for the caller to scale).

    The directions, the radii and their product are computed in one pass over
    chunks of all the group's modes on the thread pool (the same numpy
    operations per element as ``rad * sample_sphere(...)``); each realization
    then gets its (dim, k) block of one buffer."""
    kind, dim, cuts = p["kind"], p["dim"], p["cuts"]
    if not p["n"].size:
        return []
    z1s, z2s = np.split(p["z1"], cuts), np.split(p["z2"], cuts)
    if kind == 0:
        return [(np.full((dim, k), 0.0), a, b) for k, a, b in zip(p["n"], z1s, z2s)]
    total = int(p["n"].sum())
    ang1, ang2 = p["ang1"], p["ang2"]
    if kind == 1:
        ppf = model.dist_func[2]
        radius = lambda lo, hi: np.asarray(ppf(p["u_inv"][lo:hi]), dtype=np.double) * 1.0 + 0.0
    elif p["md"] is None:
        rad = np.concatenate(p["rads"]) if p["rads"] else np.empty(0)
        radius = lambda lo, hi: rad[lo:hi]
    else:
        radius = None
    sph = np.empty((dim, total))

    def chunk(lo, hi):
        s = sph[:, lo:hi]
        _sphere_chunk(dim, ang1[lo:hi], None if ang2 is None else ang2[lo:hi], s)
        if radius is not None:
            np.multiply(radius(lo, hi), s, out=s)

    _chunked(chunk, total, p["threads"])
    if radius is None:
        return list(zip(np.split(sph, cuts, axis=1), z1s, z2s))
    flat = np.empty(dim * total)
    covs, o = [], 0
    for k in p["n"].tolist():
        c = flat[dim * o:dim * (o + k)].reshape(dim, k)
        c[...] = sph[:, o:o + k]
        covs.append(c)
        o += k
    return list(zip(covs, z1s, z2s))


def _modes_python(model, mode_nos, seeds, kind, md, nwalkers, burn_in, oversampling_factor, threads):
    """Per-realization numpy draws (any seed ``RandomState`` accepts)."""
    dim = model.dim
    out, pending = [], []
    for n, seed in zip(mode_nos, seeds):
        rng = RNG(seed)
        z1 = rng.random.normal(size=n)
        z2 = rng.random.normal(size=n)
        if kind == 0:
            out.append([np.full((dim, n), 0.0), z1, z2])
            continue
        sph = rng.sample_sphere(dim, n)
        if kind == 1:
            pdf, cdf, ppf = model.dist_func
            rad = rng.sample_dist(size=n, pdf=pdf, cdf=cdf, ppf=ppf, a=0)
            out.append([rad * sph, z1, z2])
        else:
            out.append([sph, z1, z2])
            pending.append((len(out) - 1, rng, n))
    if pending:
        rads = sample_ln_pdf_batch([p[1] for p in pending], model.ln_spectral_rad_pdf, [p[2] for p in pending],
                                   sample_around=1.0 / model.len_rescaled, nwalkers=nwalkers, burn_in=burn_in,
                                   oversampling_factor=oversampling_factor, nthreads=threads, device_model=md)
        for (i, _, _), rad in zip(pending, rads):
            out[i][0] = rad * out[i][0]
    return [tuple(o) for o in out]
