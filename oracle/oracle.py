"""ORACLE -- test infrastructure only (see the package docstring).

Restates ``summate`` / ``summate_incompr`` / ``summate_fourier`` of the
reference's backend (formulas at ``src/gstools/field/generator.py:193-199``,
``:479-494`` and ``:619-624``, call sites ``:266-268``, ``:554-560`` and
``:681-688``) and its kriging sums (``src/gstools/krige/base.py:42-61``,
called at ``:307-317``).
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)


def lib_path():
    return os.path.join(_HERE, "liboracle_summate.so")


def build():
    """Compile ``summate_ref.c`` with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load_lib():
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        i64 = ctypes.c_int64
        lib.oracle_summate.argtypes = [i64, i64, i64, _dp, _dp, _dp, _dp, _dp, ctypes.c_int]
        lib.oracle_summate.restype = None
        lib.oracle_summate_incompr.argtypes = lib.oracle_summate.argtypes
        lib.oracle_summate_incompr.restype = None
        lib.oracle_summate_adjoint.argtypes = [
            i64, i64, i64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int]
        lib.oracle_summate_adjoint.restype = None
        lib.oracle_summate_fourier.argtypes = [i64, i64, i64, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int]
        lib.oracle_summate_fourier.restype = None
        lib.oracle_krige.argtypes = [i64, i64, _dp, _dp, _dp, _dp, _dp]
        lib.oracle_krige.restype = None
        _lp = ctypes.POINTER(ctypes.c_int64)
        lib.oracle_vario_unstructured.argtypes = [i64, i64, i64, i64, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp,
                                                  _lp]
        lib.oracle_vario_unstructured.restype = None
        lib.oracle_vario_directional.argtypes = [i64, i64, i64, i64, i64, _dp, _dp, _dp, _dp, ctypes.c_double,
                                                 ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _lp]
        lib.oracle_vario_directional.restype = None
        lib.oracle_vario_structured.argtypes = [i64, i64, _dp, ctypes.c_void_p, ctypes.c_int, _dp, _lp]
        lib.oracle_vario_structured.restype = None
        lib.oracle_max_threads.restype = ctypes.c_int
        _LIB = lib
    return _LIB


def max_threads():
    return int(load_lib().oracle_max_threads())


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _check(cov, z1, z2, pos):
    cov, z1, z2, pos = _c(cov), _c(z1), _c(z2), _c(pos)
    if pos.ndim == 1:
        pos = pos.reshape(1, -1)
    if cov.ndim == 1:
        cov = cov.reshape(1, -1)
    d, n = cov.shape
    if pos.shape[0] != d or z1.shape != (n,) or z2.shape != (n,):
        raise ValueError("oracle: shape mismatch")
    return cov, z1, z2, pos


def summate_c(cov, z1, z2, pos, nthreads=0):
    """C/OpenMP restatement of ``summate`` -> (P,) float64."""
    cov, z1, z2, pos = _check(cov, z1, z2, pos)
    d, n = cov.shape
    p = pos.shape[1]
    out = np.empty(p, dtype=np.float64)
    load_lib().oracle_summate(d, n, p, _ptr(cov), _ptr(z1), _ptr(z2), _ptr(pos), _ptr(out), int(nthreads))
    return out


def summate_fourier_c(spectrum_factor, modes, z1, z2, pos, nthreads=0):
    """C/OpenMP restatement of ``summate_fourier`` -> (P,) float64."""
    cov, z1, z2, pos = _check(modes, z1, z2, pos)
    f = _c(spectrum_factor)
    d, n = cov.shape
    if f.shape != (n,):
        raise ValueError("oracle: spectrum factor shape mismatch")
    p = pos.shape[1]
    out = np.empty(p, dtype=np.float64)
    load_lib().oracle_summate_fourier(d, n, p, _ptr(f), _ptr(cov), _ptr(z1), _ptr(z2), _ptr(pos), _ptr(out),
                                      int(nthreads))
    return out


def krige_c(krig_mat, krig_vecs, cond, variance=True):
    """C/OpenMP restatement of ``calc_field_krige(_and_variance)`` -> field (P,) [, error (P,)]."""
    m, v, c = _c(krig_mat), _c(krig_vecs), _c(cond)
    if v.ndim != 2 or m.shape != (v.shape[0], v.shape[0]) or c.shape != (v.shape[0],):
        raise ValueError("oracle: kriging shape mismatch")
    k, p = v.shape
    field = np.empty(p)
    error = np.empty(p) if variance else None
    load_lib().oracle_krige(k, p, _ptr(m), _ptr(v), _ptr(c), _ptr(field),
                            _ptr(error) if variance else None)
    return (field, error) if variance else field


def _lptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def vario_unstructured_c(f, bin_edges, pos, estimator_type="m", distance_type="e"):
    """C restatement of the ``unstructured`` variogram kernel -> (gamma (B,), counts (B,)).

    ``distance_type`` "h": ``pos`` is (lat, lon) in degrees and distances are
    great-circle angles in radians (haversine).
    """
    f = np.atleast_2d(_c(f))
    pos = _c(pos)
    if pos.ndim == 1:
        pos = pos.reshape(1, -1)
    e = _c(bin_edges)
    nf, npnt = f.shape
    if pos.shape[1] != npnt:
        raise ValueError("oracle: pos / field mismatch")
    nb = max(len(e) - 1, 0)
    gamma, counts = np.empty(nb), np.empty(nb, dtype=np.int64)
    if distance_type not in ("e", "h"):
        raise ValueError(f"oracle: unknown distance_type {distance_type!r}")
    if distance_type == "h" and pos.shape[0] != 2:
        raise ValueError("oracle: haversine distances need (lat, lon) positions")
    load_lib().oracle_vario_unstructured(nf, npnt, pos.shape[0], nb, _ptr(f), _ptr(e), _ptr(pos),
                                         int(estimator_type == "c"), int(distance_type == "h"), _ptr(gamma),
                                         _lptr(counts))
    return gamma, counts


def vario_directional_c(f, bin_edges, pos, direction, angles_tol=np.pi / 8, bandwidth=-1.0, separate_dirs=False,
                        estimator_type="m"):
    """C restatement of the ``directional`` kernel -> (gamma (D, B), counts (D, B))."""
    f = np.atleast_2d(_c(f))
    pos = _c(pos)
    e = _c(bin_edges)
    dirs = np.atleast_2d(_c(direction))
    nf, npnt = f.shape
    nb, nd = max(len(e) - 1, 0), dirs.shape[0]
    gamma, counts = np.empty((nd, nb)), np.empty((nd, nb), dtype=np.int64)
    load_lib().oracle_vario_directional(nf, npnt, pos.shape[0], nb, nd, _ptr(f), _ptr(e), _ptr(pos), _ptr(dirs),
                                        float(angles_tol), float(bandwidth), int(bool(separate_dirs)),
                                        int(estimator_type == "c"), _ptr(gamma), _lptr(counts))
    return gamma, counts


def vario_structured_c(f, mask=None, estimator_type="m"):
    """C restatement of ``structured`` / ``ma_structured`` on an (n0, n1) array -> (gamma, counts)."""
    f = _c(f)
    if f.ndim == 1:
        f = f.reshape(-1, 1)
    n0, n1 = f.shape
    m = None if mask is None else np.ascontiguousarray(np.asarray(mask, dtype=bool).reshape(n0, n1),
                                                       dtype=np.uint8)
    gamma, counts = np.empty(n0), np.empty(n0, dtype=np.int64)
    load_lib().oracle_vario_structured(n0, n1, _ptr(f), None if m is None else m.ctypes.data, int(estimator_type == "c"),
                                       _ptr(gamma), _lptr(counts))
    return gamma, counts


def summate_incompr_c(cov, z1, z2, pos, nthreads=0):
    """C/OpenMP restatement of ``summate_incompr`` -> (d, P) float64."""
    cov, z1, z2, pos = _check(cov, z1, z2, pos)
    d, n = cov.shape
    p = pos.shape[1]
    out = np.empty((d, p), dtype=np.float64)
    load_lib().oracle_summate_incompr(d, n, p, _ptr(cov), _ptr(z1), _ptr(z2), _ptr(pos), _ptr(out), int(nthreads))
    return out


def summate_adjoint_c(cov, z1, z2, pos, g, nthreads=0):
    """Unscaled adjoint of ``summate``: (gz1 (N,), gz2 (N,), gk (d, N))."""
    cov, z1, z2, pos = _check(cov, z1, z2, pos)
    g = _c(g)
    d, n = cov.shape
    p = pos.shape[1]
    if g.shape != (p,):
        raise ValueError("oracle: g shape mismatch")
    gz1 = np.empty(n)
    gz2 = np.empty(n)
    gk = np.empty((d, n))
    load_lib().oracle_summate_adjoint(d, n, p, _ptr(cov), _ptr(z1), _ptr(z2), _ptr(pos), _ptr(g),
                                      _ptr(gz1), _ptr(gz2), _ptr(gk), int(nthreads))
    return gz1, gz2, gk


def summate_np(cov, z1, z2, pos):
    """numpy restatement: vectorised over points, serial over modes."""
    cov, z1, z2, pos = _check(cov, z1, z2, pos)
    d, n = cov.shape
    out = np.zeros(pos.shape[1])
    for j in range(n):
        phase = np.zeros(pos.shape[1])
        for k in range(d):
            phase += cov[k, j] * pos[k]
        out += z1[j] * np.cos(phase) + z2[j] * np.sin(phase)
    return out


def summate_incompr_np(cov, z1, z2, pos):
    cov, z1, z2, pos = _check(cov, z1, z2, pos)
    d, n = cov.shape
    out = np.zeros((d, pos.shape[1]))
    for j in range(n):
        phase = np.zeros(pos.shape[1])
        for k in range(d):
            phase += cov[k, j] * pos[k]
        w = z1[j] * np.cos(phase) + z2[j] * np.sin(phase)
        k2 = float(np.sum(cov[:, j] ** 2))
        for c in range(d):
            proj = (1.0 if c == 0 else 0.0) - cov[c, j] * cov[0, j] / k2
            out[c] += proj * w
    return out


def summate_adjoint_np(cov, z1, z2, pos, g):
    cov, z1, z2, pos = _check(cov, z1, z2, pos)
    g = _c(g)
    phase = cov.T @ pos  # (N, P)
    c = np.cos(phase)
    s = np.sin(phase)
    gz1 = c @ g
    gz2 = s @ g
    w = (z2[:, None] * c - z1[:, None] * s) * g[None, :]
    gk = (w @ pos.T).T
    return gz1, gz2, gk


def summate_mp(cov, z1, z2, pos, dps=40):
    """mpmath brute force (tiny inputs only)."""
    import mpmath as mp

    cov, z1, z2, pos = _check(cov, z1, z2, pos)
    d, n = cov.shape
    with mp.workdps(dps):
        out = []
        for i in range(pos.shape[1]):
            acc = mp.mpf(0)
            for j in range(n):
                ph = mp.fsum(mp.mpf(cov[k, j]) * mp.mpf(pos[k, i]) for k in range(d))
                acc += mp.mpf(z1[j]) * mp.cos(ph) + mp.mpf(z2[j]) * mp.sin(ph)
            out.append(acc)
        return np.array([float(v) for v in out])


def summate_incompr_mp(cov, z1, z2, pos, dps=40):
    import mpmath as mp

    cov, z1, z2, pos = _check(cov, z1, z2, pos)
    d, n = cov.shape
    with mp.workdps(dps):
        out = np.zeros((d, pos.shape[1]))
        for i in range(pos.shape[1]):
            acc = [mp.mpf(0)] * d
            for j in range(n):
                kk = [mp.mpf(cov[k, j]) for k in range(d)]
                ph = mp.fsum(kk[k] * mp.mpf(pos[k, i]) for k in range(d))
                w = mp.mpf(z1[j]) * mp.cos(ph) + mp.mpf(z2[j]) * mp.sin(ph)
                k2 = mp.fsum(v * v for v in kk)
                for c in range(d):
                    acc[c] += ((1 if c == 0 else 0) - kk[c] * kk[0] / k2) * w
            for c in range(d):
                out[c, i] = float(acc[c])
        return out


def rel_errors(u, u_ref):
    """(rel_max, rel_rms) = (max|u-u_ref| / max|u_ref|, ||u-u_ref|| / ||u_ref||)."""
    u = np.asarray(u, dtype=np.float64).ravel()
    u_ref = np.asarray(u_ref, dtype=np.float64).ravel()
    diff = u - u_ref
    den_max = float(np.max(np.abs(u_ref))) if u_ref.size else 0.0
    den_rms = float(np.linalg.norm(u_ref))
    rel_max = float(np.max(np.abs(diff))) / den_max if den_max > 0 else float(np.max(np.abs(diff), initial=0.0))
    rel_rms = float(np.linalg.norm(diff)) / den_rms if den_rms > 0 else float(np.linalg.norm(diff))
    return rel_max, rel_rms
