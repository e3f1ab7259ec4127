"""Write the golden fixtures under tests/golden/ (inputs + expected outputs).

Every expected value comes from an independent high-precision computation
(mpmath at 30-40 digits) or from the oracle (the C restatement of the
reference loop), never from the code under test:

* ``tiny_mpmath.npz``         -- random tiny (cov, z1, z2, pos) cases for
  d = 1, 2, 3 with the scalar and incompressible sums by mpmath brute force;
* ``stable_density.npz``      -- the isotropic alpha-stable density p_d(y)
  (spectral density of exp(-r^alpha) at unit length scale) by mpmath.quadosc,
  alpha in {0.7, 1.5, 1.9}, d in {2, 3};
* ``config1.npz``             -- the 2D Gaussian configuration (var 1, len 10,
  mode_no 1000, seed 20170519, 100 x 100 grid): the sampled modes and the
  oracle's field (summed, unscaled);
* ``adjoint_case.npz``        -- a small adjoint case with gradients by mpmath.

Run: ``python tests/golden/make_golden.py`` (a few minutes, CPU only).
"""

import os
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def tiny_cases():
    out = {}
    for d in (1, 2, 3):
        rs = np.random.RandomState(500 + d)
        n, p = 13, 9
        cov = rs.normal(size=(d, n)) * 0.8
        z1, z2 = rs.normal(size=n), rs.normal(size=n)
        pos = rs.uniform(-20, 20, size=(d, p))
        out[f"d{d}_cov"], out[f"d{d}_z1"], out[f"d{d}_z2"], out[f"d{d}_pos"] = cov, z1, z2, pos
        out[f"d{d}_sum"] = oracle.summate_mp(cov, z1, z2, pos)
        if d > 1:
            out[f"d{d}_incompr"] = oracle.summate_incompr_mp(cov, z1, z2, pos)
    np.savez(os.path.join(HERE, "tiny_mpmath.npz"), **out)


def stable_density():
    rows = []
    with mp.workdps(30):
        for d in (2, 3):
            nu = mp.mpf(d) / 2 - 1
            for a in (0.7, 1.5, 1.9):
                for y in (0.5, 3.0, 8.0, 20.0, 60.0, 500.0, 1e4):
                    yy = mp.mpf(y)
                    fn = lambda t: t ** (mp.mpf(d) / 2) * mp.besselj(nu, yy * t) * mp.exp(-t ** mp.mpf(a))
                    integral = mp.quadosc(fn, [0, mp.inf], zeros=lambda n: mp.besseljzero(nu, n) / yy)
                    val = (2 * mp.pi) ** (-d) * (2 * mp.pi) ** (mp.mpf(d) / 2) * yy ** (1 - mp.mpf(d) / 2) * integral
                    rows.append((d, a, y, float(val)))
                    print(f"stable d={d} alpha={a} y={y}: {float(val):.6e}", flush=True)
    arr = np.array(rows)
    np.savez(os.path.join(HERE, "stable_density.npz"), d=arr[:, 0], alpha=arr[:, 1], y=arr[:, 2], p=arr[:, 3])


def config1():
    from gstools_amd import Gaussian, RandMeth
    from gstools_amd.geometry import generate_grid

    rm = RandMeth(Gaussian(dim=2, var=1.0, len_scale=10), mode_no=1000, seed=20170519)
    pos = generate_grid([np.arange(100.0), np.arange(100.0)])
    field = oracle.summate_c(rm._cov_sample, rm._z_1, rm._z_2, pos)
    np.savez(os.path.join(HERE, "config1.npz"), cov=rm._cov_sample, z1=rm._z_1, z2=rm._z_2, summed=field)


def adjoint_case():
    rs = np.random.RandomState(77)
    d, n, p = 2, 6, 11
    cov = rs.normal(size=(d, n)) * 0.3
    z1, z2 = rs.normal(size=n), rs.normal(size=n)
    pos = rs.uniform(0, 30, size=(d, p))
    g = rs.normal(size=p)
    gz1, gz2, gk = np.zeros(n), np.zeros(n), np.zeros((d, n))
    with mp.workdps(40):
        for j in range(n):
            c_acc = s_acc = mp.mpf(0)
            kx = [mp.mpf(0)] * d
            for i in range(p):
                ph = mp.fsum(mp.mpf(cov[k, j]) * mp.mpf(pos[k, i]) for k in range(d))
                c, s = mp.cos(ph), mp.sin(ph)
                c_acc += mp.mpf(g[i]) * c
                s_acc += mp.mpf(g[i]) * s
                w = mp.mpf(g[i]) * (mp.mpf(z2[j]) * c - mp.mpf(z1[j]) * s)
                for k in range(d):
                    kx[k] += w * mp.mpf(pos[k, i])
            gz1[j], gz2[j] = float(c_acc), float(s_acc)
            for k in range(d):
                gk[k, j] = float(kx[k])
    np.savez(os.path.join(HERE, "adjoint_case.npz"), cov=cov, z1=z1, z2=z2, pos=pos, g=g, gz1=gz1, gz2=gz2, gk=gk)


if __name__ == "__main__":
    tiny_cases()
    adjoint_case()
    config1()
    stable_density()
    print("golden fixtures written", flush=True)
