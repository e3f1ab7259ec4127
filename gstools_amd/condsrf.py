"""Conditioned spatial random fields: kriging plus a scaled unconditioned field.

The conditioned field at a target x (``src/gstools/field/cond_srf.py:27-311``)
is built from three parts:

* the kriged estimate k(x) of the conditioning data and its kriging
  variance s2(x) (:class:`gstools_amd.krige.Krige`, sums on the GPU);
* an unconditioned field u(x) of the same model without nugget
  (:class:`gstools_amd.RandMeth`, summation on the GPU);
* nugget noise where the model has a nugget n.

The kriging variance is shared between the random field and the noise: the
field part carries max(s2 - n, 0), the noise the rest (min(s2, n)), so

    field(x) = k(x) + sqrt(max(s2 - n, 0) / var) u(x) + sqrt(min(s2, n)) e(x)

with e standard normal noise drawn by the generator (its ``get_nugget``,
the same draw as the reference); without nugget the field part carries all
of s2.  Mean and trend are added afterwards.  A call on the positions of the
previous call reuses its kriging result.
"""

import numpy as np

from .generator import RandMeth
from .krige import Krige, _positions, eval_func

__all__ = ["CondSRF"]

GENERATOR = {"RandMeth": RandMeth}


def _same_positions(a, b):
    """Whether two position arguments hold the same coordinates."""
    if isinstance(a, (tuple, list)) != isinstance(b, (tuple, list)):
        return False
    if isinstance(a, (tuple, list)):
        return len(a) == len(b) and all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(a, b))
    try:
        return np.array_equal(np.asarray(a), np.asarray(b))
    except ValueError:
        return False


class CondSRF:
    """Conditioned random field of a kriging setup (``cond_srf.py:27-311``)."""

    default_field_names = ["field", "raw_field", "raw_krige"]

    def __init__(self, krige, generator="RandMeth", **generator_kwargs):
        if not isinstance(krige, Krige):
            raise ValueError(f"CondSRF: needs a Krige instance, got {type(krige).__name__}")
        self._krige = krige
        self.fields = {}
        self.pos = None
        self.mesh_type = "unstructured"
        self._generator = None
        self.set_generator(generator, **generator_kwargs)

    def set_generator(self, generator, **generator_kwargs):
        """Build the unconditioned generator (a name or a generator class)."""
        cls = GENERATOR.get(generator, generator) if isinstance(generator, str) else generator
        if not isinstance(cls, type):
            raise ValueError(f"CondSRF: unknown generator {generator!r}")
        gen = cls(self.model, **generator_kwargs)
        if gen.value_type != "scalar":
            raise ValueError(f"CondSRF: conditioning needs a scalar field, {cls.__name__} gives "
                             f"{gen.value_type} values")
        self._generator = gen
        self.value_type = gen.value_type

    @property
    def krige(self):
        return self._krige

    @property
    def generator(self):
        return self._generator

    @property
    def model(self):
        return self._krige.model

    @property
    def field_names(self):
        return list(self.fields)

    def __getitem__(self, name):
        return self.fields[name]

    def __call__(self, pos=None, seed=np.nan, mesh_type="unstructured", post_process=True, store=True,
                 krige_store=True, **kwargs):
        """The conditioned field at ``pos`` (``cond_srf.py:63-153``)."""
        if pos is None:
            if self.pos is None:
                raise ValueError("CondSRF: no positions given and none stored")
            pos, mesh_type = self.pos, self.mesh_type
        repeat = self.pos is not None and self.mesh_type == mesh_type and _same_positions(self.pos, pos)
        if not repeat:
            self.fields = {}
        self.pos, self.mesh_type = pos, mesh_type
        orig, shape = _positions(self.model.dim, pos, mesh_type)
        self.generator.update(self.model, seed)
        raw_field = np.reshape(self.generator(self.model.isometrize(orig), add_nugget=False), shape)
        krige_est, krige_var = self._kriging(pos, mesh_type, orig, shape, repeat, krige_store, kwargs)
        scale, noise = self.get_scaling(krige_var, shape)
        field = krige_est + scale * raw_field + noise
        if post_process:
            field = field + (eval_func(self.krige.mean, orig, self.model.dim, broadcast=True)
                             + eval_func(self.krige.trend, orig, self.model.dim, broadcast=True)).reshape(shape)
        if store:
            self.fields.update(raw_krige=krige_est, raw_field=raw_field, field=field)
        return field

    def _kriging(self, pos, mesh_type, orig, shape, repeat, krige_store, kwargs):
        """Kriged estimate (no mean / trend) and variance, reused for repeated positions."""
        if repeat and "raw_krige" in self.fields and "krige_var" in self.krige.fields:
            return self.fields["raw_krige"], self.krige.fields["krige_var"]
        est, var = self.krige(pos, mesh_type=mesh_type, only_mean=False, return_var=True, post_process=False,
                              store=bool(krige_store), **kwargs)
        if krige_store:
            self.krige.fields["field"] = self.krige._post(est.copy(), orig, shape)
        return est, var

    def get_scaling(self, krige_var, shape):
        """(field scale, nugget noise) for a kriging variance (``cond_srf.py:155-183``)."""
        nug, var = self.model.nugget, self.model.var
        if nug <= 0:
            return np.sqrt(krige_var / var), 0
        field_part = np.maximum(krige_var - nug, 0)
        noise_std = np.sqrt(krige_var - field_part)  # min(s2, n)
        # the generator's nugget draw is sqrt(n) e: scale it to sqrt(min(s2, n)) e
        return np.sqrt(field_part / var), noise_std / np.sqrt(nug) * self.generator.get_nugget(shape)

    def structured(self, *args, **kwargs):
        return self(*args, mesh_type="structured", **kwargs)

    def unstructured(self, *args, **kwargs):
        return self(*args, mesh_type="unstructured", **kwargs)

    def __repr__(self):
        return f"CondSRF(krige={self.krige!r}, generator={self.generator.name})"
