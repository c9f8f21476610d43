"""Pins the oracle's Symanzik gauge action and force.

* Loop enumeration: the oracle's action equals a brute-force Python sum over
  every distinct geometric loop, found by generating all closed paths of each
  shape from every site and de-duplicating them as sets of links.
* Force = derivative of the action: central finite differences (h = 1e-5)
  of S(exp(tX) U) against -Re Tr(X W) on random links.
* Invariants: W transforms covariantly, the force vanishes at U = 1, the
  momentum stays traceless anti-Hermitian.
"""
import itertools

import numpy as np
import pytest

import oracle as O
from pyqphix import fields as F

COEF = (0.7, -0.11, -0.05)


def path_loops(lat, shape):
    """All closed paths of a shape family from every site, de-duplicated as
    link sets; returns list of (start site coords, steps)."""
    dirs = range(4)
    fams = []
    for mu in dirs:
        for nu in dirs:
            if nu == mu:
                continue
            for s in (1, -1):
                n = s * (nu + 1)
                m = mu + 1
                if shape == "plaq":
                    fams.append((m, n, -m, -n))
                elif shape == "rect":
                    fams.append((m, m, n, -m, -m, -n))
                    fams.append((m, n, n, -m, -n, -n))
            if shape == "pgm":
                for rho in dirs:
                    if rho in (mu, nu):
                        continue
                    for sa, sb in itertools.product((1, -1), (1, -1)):
                        fams.append((mu + 1, sa * (nu + 1), sb * (rho + 1), -(mu + 1),
                                     -sa * (nu + 1), -sb * (rho + 1)))
    seen = {}
    L = lat.dims
    for x in itertools.product(*[range(n) for n in L]):
        for steps in fams:
            y = list(x)
            links = []
            for s in steps:
                d = abs(s) - 1
                if s > 0:
                    links.append((tuple(v % L[k] for k, v in enumerate(y)), d))
                    y[d] += 1
                else:
                    y[d] -= 1
                    links.append((tuple(v % L[k] for k, v in enumerate(y)), d))
            key = frozenset(links)
            if key not in seen:
                seen[key] = (x, steps)
    return list(seen.values())


def loop_trace(lat, U, x, steps):
    y = list(x)
    P = np.eye(3, dtype=complex)
    for s in steps:
        d = abs(s) - 1
        if s > 0:
            P = P @ U[lat.index(*y), d]
            y[d] += 1
        else:
            y[d] -= 1
            P = P @ F.dagger(U[lat.index(*y), d])
    return np.trace(P).real


@pytest.fixture(scope="module")
def small():
    lat = F.Lattice((4, 4, 4, 4))
    U = F.gauge_field(lat, "weak", 5, eps=0.5)
    return lat, U


def test_loop_counts_per_site(small):
    lat, _ = small
    counts = {k: len(path_loops(lat, k)) / lat.V for k in ("plaq", "rect", "pgm")}
    assert counts == {"plaq": 6, "rect": 12, "pgm": 16}


def test_action_matches_bruteforce_loops(small):
    lat, U = small
    exp = 0.0
    for k, c in zip(("plaq", "rect", "pgm"), COEF):
        exp -= c * sum(loop_trace(lat, U, x, st) for x, st in path_loops(lat, k))
    got = O.gauge_action(lat.dims, F.to_raw(U), COEF)
    assert abs(got - exp) < 1e-10 * abs(exp)


def su3_exp(X):
    w, v = np.linalg.eigh(-1j * X)
    return (v * np.exp(1j * w)[None, :]) @ F.dagger(v)


@pytest.mark.parametrize("coef", [COEF, (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)])
def test_force_is_action_derivative(coef):
    lat = F.Lattice((4, 4, 4, 8))
    U = F.gauge_field(lat, "weak", 7, eps=0.4)
    raw = F.to_raw(U)
    mom0 = np.zeros(72 * lat.V)
    _, W = O.gauge_force(lat.dims, raw, coef, 0.0, mom0)
    W = W.view(np.complex128).reshape(lat.V, 4, 3, 3)
    rng = np.random.default_rng(3)
    h = 1e-5
    for _ in range(8):
        s = rng.integers(lat.V)
        mu = rng.integers(4)
        X = F.antihermitian_traceless(rng, ())
        Sp, Sm = [], []
        for sign, out in ((1, Sp), (-1, Sm)):
            U2 = U.copy()
            U2[s, mu] = su3_exp(sign * h * X) @ U[s, mu]
            box = [int(lat.x[s]), int(lat.y[s]), int(lat.z[s]), int(lat.t[s])]
            out.append(O.gauge_action(lat.dims, F.to_raw(U2), coef, box, 3, int(mu)))
        fd = (Sp[0] - Sm[0]) / (2 * h)
        an = -np.trace(X @ W[s, mu]).real
        assert abs(fd - an) < 1e-8 * max(1.0, abs(an))


def test_link_restricted_action_differences_are_exact():
    """Changing one link changes S exactly by the change of the loops through it."""
    lat = F.Lattice((4, 4, 4, 8))
    U = F.gauge_field(lat, "weak", 12, eps=0.4)
    rng = np.random.default_rng(5)
    s, mu = 77, 2
    U2 = U.copy()
    U2[s, mu] = su3_exp(0.3 * F.antihermitian_traceless(rng, ())) @ U[s, mu]
    box = [int(lat.x[s]), int(lat.y[s]), int(lat.z[s]), int(lat.t[s])]
    full = O.gauge_action(lat.dims, F.to_raw(U2), COEF) - O.gauge_action(lat.dims, F.to_raw(U), COEF)
    loc = (O.gauge_action(lat.dims, F.to_raw(U2), COEF, box, 2, mu)
           - O.gauge_action(lat.dims, F.to_raw(U), COEF, box, 2, mu))
    assert abs(full - loc) < 1e-11 * abs(full)


def test_force_covariance_and_identity():
    lat = F.Lattice((4, 4, 4, 4))
    U = F.gauge_field(lat, "haar", 9)
    g = F.haar_su3(np.random.default_rng(2), lat.V)
    U2 = np.empty_like(U)
    for mu in range(4):
        U2[:, mu] = g @ U[:, mu] @ F.dagger(g[lat.shifted(mu, 1)])
    z = np.zeros(72 * lat.V)
    _, W1 = O.gauge_force(lat.dims, F.to_raw(U), COEF, 0.0, z)
    _, W2 = O.gauge_force(lat.dims, F.to_raw(U2), COEF, 0.0, z)
    W1 = W1.view(np.complex128).reshape(lat.V, 4, 3, 3)
    W2 = W2.view(np.complex128).reshape(lat.V, 4, 3, 3)
    exp = np.einsum("nij,nmjk,nkl->nmil", g, W1, F.dagger(g))
    assert np.max(np.abs(W2 - exp)) < 1e-12
    one = np.broadcast_to(np.eye(3, dtype=complex), (lat.V, 4, 3, 3)).copy()
    mom, _ = O.gauge_force(lat.dims, F.to_raw(one), COEF, 0.3, z)
    assert np.max(np.abs(mom)) < 1e-14


def test_momentum_stays_traceless_antihermitian():
    lat = F.Lattice((4, 4, 4, 4))
    U = F.gauge_field(lat, "weak", 11)
    m0 = F.antihermitian_traceless(np.random.default_rng(1), (lat.V, 4))
    mom, _ = O.gauge_force(lat.dims, F.to_raw(U), COEF, 0.05, F.to_raw(m0))
    A = mom.view(np.complex128).reshape(lat.V, 4, 3, 3)
    assert np.max(np.abs(A + F.dagger(A))) < 1e-14
    assert np.max(np.abs(np.trace(A, axis1=-2, axis2=-1))) < 1e-14
