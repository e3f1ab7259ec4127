"""Thin spatial-random-field caller around the GPU generators.

Mirrors the parts of ``SRF.__call__`` (``src/gstools/field/srf.py:109-163``)
and ``Field.pre_pos`` (``src/gstools/field/base.py:254-297``) that sit on the
summation path: generator update, meshgrid of structured positions,
isometrize, generator call, reshape, and a constant (scalar or per-component)
mean as ``apply_mean_norm_trend`` adds it (``normalizer/tools.py:35-110``).
Normalizers, trends, upscaling and field storage are not restated.

Structured grids never build the meshgrid when the isometrize matrix M is
the identity, nor for the precisions other than "fp64" (``"fp32"``,
``"bf16x3"`` / ``"fp16x3"`` on the matrix cores, ``"fp64gemm"``): a rotation
or anisotropy is folded into the wavenumbers, k' = M^T k, since k.(M x) =
k'.x, so the grid stays a tensor grid.  "fp64" on a rotated or anisotropic
grid keeps the reference's route (meshgrid, ``np.dot(M, pos)``, per-point
sum), whose phases k.(M x) it reproduces bit for bit; so do vector fields,
whose projector needs the unfolded k.
"""

import numpy as np
import torch

from .generator import Fourier, IncomprRandMeth, RandMeth
from .summate import host_array
from .geometry import generate_grid, matrix_isometrize

__all__ = ["SRF", "GENERATOR"]

GENERATOR = {
    "RandMeth": RandMeth,
    "IncomprRandMeth": IncomprRandMeth,
    "VectorField": IncomprRandMeth,
    "Fourier": Fourier,
}


class SRF:
    """Spatial random field on the GPU generators."""

    def __init__(self, model, mean=0.0, generator="RandMeth", **generator_kwargs):
        self.model = model
        self.mean = mean
        gen = GENERATOR.get(generator, generator) if isinstance(generator, str) else generator
        if not isinstance(gen, type):
            raise ValueError(f"gstools.SRF: Unknown or wrong generator: {generator}")
        self.generator = gen(model, **generator_kwargs)
        self.value_type = self.generator.value_type
        self.field = None

    def __call__(self, pos, seed=np.nan, mesh_type="unstructured", precision=None):
        self.generator.update(self.model, seed)
        dim = self.model.dim
        if mesh_type != "unstructured":
            axes = [np.asarray(a, dtype=np.double).reshape(-1) for a in pos]
            if len(axes) != dim:
                raise ValueError("SRF: position tuple does not match the model dimension")
            shape = tuple(a.size for a in axes)
            iso = matrix_isometrize(dim, self.model.angles, self.model.anis)
            gen = self.generator
            if hasattr(gen, "structured"):
                prec = precision or gen.precision
                if np.array_equal(iso, np.eye(dim)):  # k.(I x) == k.x exactly
                    field = host_array(gen.structured(axes, precision=precision))
                    return self._post(np.asarray(field, dtype=np.double))
                if prec != "fp64" and self.value_type == "scalar":
                    field = host_array(gen.structured(axes, precision=precision, iso_matrix=iso))
                    return self._post(np.asarray(field, dtype=np.double))
            pos = generate_grid(axes)
        else:
            pos = np.asarray(pos, dtype=np.double).reshape(dim, -1)
            shape = (pos.shape[1],)
        iso = self.model.isometrize(pos)
        field = self._point_call(iso, precision)
        if self.value_type == "vector":
            field = np.reshape(field, (dim,) + shape)
        else:
            field = np.reshape(field, shape)
        return self._post(field)

    def _point_call(self, iso, precision):
        """Generator call on isometrized points, in ``precision`` when given."""
        gen = self.generator
        if precision is None or precision == getattr(gen, "precision", None):
            return gen(iso)
        if precision not in ("fp64", "fp32"):
            raise ValueError(f"SRF: precision {precision!r} needs a structured grid whose field is scalar "
                             "(point sets and vector fields on rotated grids run 'fp64' or 'fp32')")
        prev = gen.precision
        gen.precision = precision
        try:
            return gen(iso)
        finally:
            gen.precision = prev

    def _post(self, field):
        mean = np.asarray(self.mean, dtype=np.double)
        if mean.size == 1:
            field = field + mean.item()
        else:
            field = field + mean.reshape((-1,) + (1,) * (field.ndim - 1))
        self.field = field
        return field
