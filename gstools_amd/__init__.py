"""gstools_amd -- GSTools' randomisation-method field summation on AMD MI355X (gfx950).

A drop-in for ``RandMeth.__call__`` / ``Fourier.__call__`` and the
``summate`` / ``summate_incompr`` / ``summate_fourier`` kernels they call
(``src/gstools/field/generator.py:42-75``, ``:243-270``, ``:529-567``,
``:659-691`` of GSTools), with the sum computed by hand-written HIP kernels
in ``libgsx.so`` (C ABI: ``include/gsx.h``).

Public surface:

* ``summate``, ``summate_incompr``, ``summate_fourier`` -- the backend
  functions (numpy in/out);
* ``summate_device``, ``summate_incompr_device``, ``summate_grid_device``,
  ``summate_adjoint_device``, ``summate_fourier_device``,
  ``summate_fourier_grid_device`` -- torch-tensor forms with FP32,
  structured-grid, matrix-core and adjoint paths;
* ``RandMeth``, ``IncomprRandMeth``, ``Fourier`` -- ``Generator`` plugins;
* ``SRF`` -- a thin caller (positions -> isometrize -> generator -> mean);
* ``Gaussian``, ``Exponential``, ``Matern``, ``Stable`` -- the spectral
  members of the covariance models the configurations use;
* ``krige`` (``calc_field_krige``, ``calc_field_krige_and_variance`` and the
  ``Krige`` / ``Simple`` / ``Ordinary`` / ``Universal`` / ``ExtDrift`` /
  ``Detrended`` front end) and ``CondSRF`` -- kriging sums on the GPU;
* ``variogram`` (the ``unstructured`` / ``directional`` / ``structured`` /
  ``ma_structured`` kernels) and ``vario_estimate``, ``vario_estimate_axis``,
  ``standard_bins`` -- pair binning on the GPU;
* ``RandMethSum`` (autograd), ``batch`` (many realizations per launch),
  ``sample_modes_batch`` (the modes of many realizations, sampled in
  lockstep, identical to one ``RandMeth`` each).
"""

from . import geometry, krige, variogram  # noqa: F401
from .condsrf import CondSRF  # noqa: F401
from .covmodel import HANKEL_DEFAULT, CovModel, Exponential, Gaussian, Matern, Spherical, Stable  # noqa: F401
from .generator import Fourier, Generator, IncomprRandMeth, RandMeth  # noqa: F401
from .geometry import DEGREE_SCALE, EARTH_RADIUS, KM_SCALE, RADIAN_SCALE  # noqa: F401
from .rng import RNG, MasterRNG  # noqa: F401
from .sampling import sample_ln_pdf_batch, sample_modes_batch  # noqa: F401
from .srf import SRF  # noqa: F401
from .summate import (  # noqa: F401
    summate,
    summate_adjoint_device,
    summate_device,
    summate_fourier,
    summate_fourier_device,
    summate_fourier_grid_device,
    summate_fourier_tensor_device,
    summate_grid_device,
    summate_incompr,
    summate_incompr_device,
    summate_incompr_grid_device,
)

from .variogram import standard_bins, vario_estimate, vario_estimate_axis  # noqa: F401

__version__ = "0.1.0"
