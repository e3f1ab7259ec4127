"""ORACLE -- test infrastructure, never part of the shipped product.

CPU restatements of the randomisation-method summation that GSTools calls at
``src/gstools/field/generator.py:266-268`` (``summate``) and ``:554-560``
(``summate_incompr``).  The reference's arithmetic lives in the un-vendored
third-party package ``gstools-cython >=1,<2`` (``pyproject.toml:45``); these
functions restate its published formula (``generator.py:193-199`` and
``:479-494``) and its loop order (parallel over points, serial over modes,
phase accumulated in dimension order without fused multiply-add).

Three independent restatements are provided so that they check each other:

* ``summate_c`` / ``summate_incompr_c`` / ``summate_adjoint_c`` -- the C/OpenMP
  loop in ``summate_ref.c`` (``liboracle_summate.so``).  This is also the CPU
  baseline that the benchmark times (``cpu_baseline.kind == "port"``).
* ``summate_np`` / ``summate_incompr_np`` / ``summate_adjoint_np`` -- numpy,
  vectorised over points, serial over modes (same operation order).
* ``summate_mp`` / ``summate_incompr_mp`` -- mpmath brute force at 40 digits
  for tiny inputs.

Pinning (each function is checked by the tests in ``tests/test_oracle.py``):

* ``summate``: mpmath brute force; linearity in (z1, z2); N = 0 gives 0;
  k = 0 gives sum(z1); the (k, z1, z2) -> (-k, z1, -z2) symmetry; translation
  by t equals rotating (z1, z2) by k.t; the tensor-grid GEMM identity; the
  reference's own known-answer values (``tests/test_randmeth.py:33-71``,
  ``tests/test_pgs.py:71-75`` through the host sampler).
* ``summate_fourier``: equals ``summate`` with amplitudes scaled by the
  factor; the reference's Fourier known values and periodicity
  (``tests/test_fouriergen.py:47-81``) through the host mode grid.
* ``summate_incompr``: mpmath brute force; analytic divergence-free check
  (central differences); ``tests/test_incomprrandmeth.py:34-44`` and
  ``tests/test_srf.py:259-275`` known values.
* ``summate_adjoint``: central finite differences of ``summate`` in float64
  and the torch float64 autograd of the numpy formula.
* ``krige`` (``krige_ref.c``, ``calc_field_krige(_and_variance)`` of
  ``src/gstools/krige/base.py:42-61``): matrix identities computed by numpy
  (``cond @ M @ V``, column sums of ``V * (M @ V)``) and simple kriging's
  interpolation property (at a conditioning point the field is the
  conditioning value and the variance term is the covariance).
* variogram kernels (``variogram_ref.c``, ``variogram.py:44-113`` of the
  reference): the reference's known values (``tests/test_variogram_*``),
  brute force over explicit pair lists in numpy, the structured kernel
  against the unstructured one on grids (axis lags), direction tests
  against explicit angle computations.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package.
"""

from .oracle import (  # noqa: F401
    krige_c,
    lib_path,
    load_lib,
    max_threads,
    rel_errors,
    summate_adjoint_c,
    summate_adjoint_np,
    summate_c,
    summate_fourier_c,
    summate_incompr_c,
    summate_incompr_mp,
    summate_incompr_np,
    summate_mp,
    summate_np,
    vario_directional_c,
    vario_structured_c,
    vario_unstructured_c,
)
