"""Differentiable randomisation-method field (forward + hand-written adjoint).

:class:`RandMethSum` is a ``torch.autograd.Function`` whose forward is the GPU
summation (``summate_device``) and whose backward is the adjoint kernel
(``summate_adjoint_device``): a reduction over points per mode giving

    dL/dz1_j = sum_i g_i cos(k_j.x_i),     dL/dz2_j = sum_i g_i sin(k_j.x_i),
    dL/dk_jd = sum_i g_i x_id (z2_j cos(k_j.x_i) - z1_j sin(k_j.x_i)).

The forward formula is ``src/gstools/field/generator.py:193-199``; the
reference has no gradient.  :func:`randmeth_field` adds the RandMeth scaling
``sqrt(var / N)`` (``generator.py:270``) and a differentiable length scale:
every mode sampler draws wavenumbers proportional to ``1 / len_rescaled``
(inverse cdf ``covmodel/models.py:168-180``; the stretch move is affine
invariant and starts at ``1 / len_rescaled``, ``generator.py:380-385``), so
``k = k_unit / len_scale`` and ``dL/dlen = sum_jd dL/dk_jd (-k_jd / len)``
follows from torch's chain rule.
"""

import torch

from .summate import summate_adjoint_device, summate_device

__all__ = ["RandMethSum", "randmeth_field"]


class RandMethSum(torch.autograd.Function):
    """u = sum_j z1 cos(k.x) + z2 sin(k.x); differentiable in z1, z2 and k."""

    @staticmethod
    def forward(ctx, z1, z2, cov, pos, precision="fp64"):
        u = summate_device(cov.detach(), z1.detach(), z2.detach(), pos.detach(), precision=precision)
        ctx.save_for_backward(z1, z2, cov, pos)
        ctx.precision = precision
        return u.to(torch.float64) if precision == "fp64" else u

    @staticmethod
    def backward(ctx, gu):
        z1, z2, cov, pos = ctx.saved_tensors
        g = gu.contiguous().to(torch.float64 if ctx.precision == "fp64" else torch.float32)
        gz1, gz2, gk = summate_adjoint_device(cov.detach(), z1.detach(), z2.detach(), pos.detach(), g,
                                              precision=ctx.precision)
        return (gz1.to(z1.dtype).reshape(z1.shape) if ctx.needs_input_grad[0] else None,
                gz2.to(z2.dtype).reshape(z2.shape) if ctx.needs_input_grad[1] else None,
                gk.to(cov.dtype).reshape(cov.shape) if ctx.needs_input_grad[2] else None,
                None, None)


def randmeth_field(z1, z2, k_unit, pos, var=1.0, len_scale=1.0, precision="fp64"):
    """Field ``sqrt(var / N) * sum(...)`` with wavenumbers ``k_unit / len_scale``.

    All of ``z1``, ``z2`` (N,), ``k_unit`` (d, N), ``var`` and ``len_scale``
    may require grad; ``pos`` (d, P) is data.  Tensors live on the GPU in
    float64.
    """
    cov = k_unit / len_scale
    n = z1.shape[0]
    u = RandMethSum.apply(z1, z2, cov.contiguous(), pos, precision)
    return torch.sqrt(torch.as_tensor(var, dtype=torch.float64, device=u.device) / n) * u
