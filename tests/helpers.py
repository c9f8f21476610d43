"""Shared helpers for the tests: independent numpy forms of the operator and
small utilities.  The numpy Dslash below works on global coordinates
(Lattice.shifted) instead of the oracle's neighbour tables and slot storage."""
import numpy as np

from pyqphix import fields as F


def half(v_full_raw, lat, p):
    """Raw full array (6 reals/site) -> half of parity p."""
    return np.ascontiguousarray(v_full_raw[6 * p * lat.vh:6 * (p + 1) * lat.vh])


def full_from_half(h, lat, p, dtype=np.float64):
    out = np.zeros(6 * lat.V, dtype=dtype)
    out[6 * p * lat.vh:6 * (p + 1) * lat.vh] = h
    return out


def np_dslash(lat, fat, lng, fatback, lngback, psi):
    """D psi on the full lattice: sum_mu fat psi(x+mu) + fatback psi(x-mu)
    + lng psi(x+3mu) + lngback psi(x-3mu).  Complex (V,4,3,3) links, (V,3) psi."""
    out = np.zeros_like(psi)
    for mu in range(4):
        out += np.einsum("nij,nj->ni", fat[:, mu], psi[lat.shifted(mu, 1)])
        out += np.einsum("nij,nj->ni", fatback[:, mu], psi[lat.shifted(mu, -1)])
        out += np.einsum("nij,nj->ni", lng[:, mu], psi[lat.shifted(mu, 3)])
        out += np.einsum("nij,nj->ni", lngback[:, mu], psi[lat.shifted(mu, -3)])
    return out


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if not (np.iscomplexobj(a) or np.iscomplexobj(b)):
        a = a.astype(np.float64)
        b = b.astype(np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cdot(a, b):
    return np.vdot(a.ravel(), b.ravel())


def allowed_momenta(lat, n):
    """n momenta allowed by periodic x,y,z and antiperiodic t."""
    nx, ny, nz, nt = lat.dims
    rng = np.random.default_rng(7)
    ps = []
    for _ in range(n):
        k = [rng.integers(0, d) for d in lat.dims]
        ps.append((2 * np.pi * k[0] / nx, 2 * np.pi * k[1] / ny, 2 * np.pi * k[2] / nz,
                   (2 * k[3] + 1) * np.pi / nt))
    return ps


def random_links(lat, kind="haar", seed=11, eps=0.3):
    fat, lng = F.make_links(lat, kind, seed, eps=eps)
    fb = F.back_links(lat, fat, 1)
    lb = F.back_links(lat, lng, 3)
    return fat, lng, fb, lb
