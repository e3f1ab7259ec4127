"""Conditioned spatial random fields: kriging plus a scaled unconditioned field.

Restates ``CondSRF`` (``src/gstools/field/cond_srf.py:27-311``): the kriged
field of the conditioning data plus an unconditioned RandMeth field (no
nugget) scaled by the kriging standard deviation,

    field = krige + sqrt(krige_var / var) * raw_field   (+ nugget noise),

with the nugget split of ``get_scaling`` (``cond_srf.py:155-183``).  The
random field comes from :class:`gstools_amd.RandMeth` (summation on the GPU)
and the kriging sums from :mod:`gstools_amd.krige` (GPU).  When the positions
are unchanged the kriging result of the previous call is reused, as in the
reference.
"""

import numpy as np

from .generator import RandMeth
from .krige import Krige, _positions, eval_func

__all__ = ["CondSRF"]

GENERATOR = {"RandMeth": RandMeth}


class CondSRF:
    """Conditioned SRF (``cond_srf.py:27-311``)."""

    default_field_names = ["field", "raw_field", "raw_krige"]

    def __init__(self, krige, generator="RandMeth", **generator_kwargs):
        if not isinstance(krige, Krige):
            raise ValueError("CondSRF: krige should be an instance of Krige.")
        self._krige = krige
        self.fields = {}
        self.pos = None
        self.mesh_type = "unstructured"
        self._generator = None
        self.set_generator(generator, **generator_kwargs)

    def set_generator(self, generator, **generator_kwargs):
        gen = GENERATOR[generator] if generator in GENERATOR else generator
        if not isinstance(gen, type):
            raise ValueError(f"gstools.CondSRF: Unknown or wrong generator: {generator}")
        self._generator = gen(self.model, **generator_kwargs)
        self.value_type = self._generator.value_type
        if self.value_type != "scalar":
            raise ValueError("CondSRF: only scalar generators can be conditioned")

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
        """The conditioned field at ``pos`` (``cond_srf.py:68-153``)."""
        if pos is None:
            if self.pos is None:
                raise ValueError("CondSRF: no positions given")
            pos, mesh_type = self.pos, self.mesh_type
        same = (self.pos is not None and self.mesh_type == mesh_type
                and _same_pos(self.pos, pos))
        if not same:
            self.fields = {}
        self.pos, self.mesh_type = pos, mesh_type
        self.generator.update(self.model, seed)
        xyz, shape = _positions(self.model.dim, pos, mesh_type)
        iso = self.model.isometrize(xyz)
        rawfield = np.reshape(self.generator(iso, add_nugget=False), shape)
        if same and "raw_krige" in self.fields and "krige_var" in self.krige.fields:
            rawkrige, krige_var = self.fields["raw_krige"], self.krige.fields["krige_var"]
        else:
            rawkrige, krige_var = self.krige(pos, mesh_type=mesh_type, only_mean=False, return_var=True,
                                             post_process=False, store=bool(krige_store), **kwargs)
            if krige_store:
                self.krige.fields["field"] = self.krige._post(rawkrige.copy(), xyz, shape)
        var_scale, nugget = self.get_scaling(krige_var, shape)
        if store:
            self.fields["raw_krige"] = rawkrige
            self.fields["raw_field"] = rawfield
        field = rawkrige + var_scale * rawfield + nugget
        if post_process:
            field = field + eval_func(self.krige.mean, xyz, self.model.dim, broadcast=True).reshape(shape)
            field = field + eval_func(self.krige.trend, xyz, self.model.dim, broadcast=True).reshape(shape)
        if store:
            self.fields["field"] = field
        return field

    def get_scaling(self, krige_var, shape):
        """Scale of the random field and nugget noise (``cond_srf.py:155-183``)."""
        if self.model.nugget > 0:
            var_scale = np.maximum(krige_var - self.model.nugget, 0)
            nug_scale = np.sqrt((krige_var - var_scale) / self.model.nugget)
            var_scale = np.sqrt(var_scale / self.model.var)
            nugget = nug_scale * self.generator.get_nugget(shape)
        else:
            var_scale = np.sqrt(krige_var / self.model.var)
            nugget = 0
        return var_scale, nugget

    def structured(self, *args, **kwargs):
        return self(*args, mesh_type="structured", **kwargs)

    def unstructured(self, *args, **kwargs):
        return self(*args, mesh_type="unstructured", **kwargs)

    def __repr__(self):
        return f"CondSRF(krige={self.krige!r}, generator={self.generator.name})"


def _same_pos(a, b):
    try:
        if isinstance(a, (tuple, list)) and isinstance(b, (tuple, list)):
            return len(a) == len(b) and all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(a, b))
        return np.array_equal(np.asarray(a), np.asarray(b))
    except ValueError:
        return False
