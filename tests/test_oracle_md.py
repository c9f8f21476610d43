"""Pins the oracle's gauge-field update and molecular-dynamics integrators.

* The matrix exponential (Taylor series with scaling and squaring) equals
  scipy's expm and, for anti-Hermitian arguments, the eigen-decomposition
  exp(X) = V exp(i w) V^dag.
* U <- exp(eps A) U per link equals the per-link product with scipy's expm.
* The MD step conserves H = -1/2 sum Tr A^2 + S(U) to O(eps^2): halving eps at
  fixed trajectory length divides |dH| by ~4 (both schemes), Omelyan's dH is
  smaller than leapfrog's; the step is reversible (flip the momentum, run
  again, flip: the start comes back); links stay SU(3); the free field with
  zero momentum does not move.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O
from pyqphix import fields as F

COEF = (0.7, -0.11, -0.05)


def su3_exp(X):
    w, v = np.linalg.eigh(-1j * X)
    return (v * np.exp(1j * w)[None, :]) @ F.dagger(v)


def test_matrix_exp():
    rng = np.random.default_rng(2)
    for scale in (1e-8, 0.01, 0.3, 1.0, 4.0, 20.0):
        Z = scale * (rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3)))
        assert np.abs(O.matrix_exp(Z) - sl.expm(Z)).max() < 1e-13 * max(1.0, np.abs(sl.expm(Z)).max())
        X = F.antihermitian_traceless(rng, ()) * scale
        E = O.matrix_exp(X)
        assert np.abs(E - su3_exp(X)).max() < 1e-13 * max(1, scale)
        assert np.abs(E @ F.dagger(E) - np.eye(3)).max() < 1e-13 * max(1, scale)


def kinetic(mom_raw, V):
    A = mom_raw.view(np.complex128).reshape(V, 4, 3, 3)
    return 0.5 * float(np.sum(np.abs(A) ** 2))


def hamiltonian(lat, U, mom):
    return kinetic(mom, lat.V) + O.gauge_action(lat.dims, U, COEF)


@pytest.fixture(scope="module")
def fields():
    lat = F.Lattice((4, 4, 4, 4))
    U = F.to_raw(F.gauge_field(lat, "weak", 21, eps=0.5))
    mom = F.to_raw(F.antihermitian_traceless(np.random.default_rng(22), (lat.V, 4)))
    return lat, U, mom


def test_gauge_update_matches_expm(fields):
    lat, U, mom = fields
    eps = 0.37
    got = O.gauge_update(lat.dims, U, mom, eps).view(np.complex128).reshape(lat.V, 4, 3, 3)
    u = U.view(np.complex128).reshape(lat.V, 4, 3, 3)
    A = mom.view(np.complex128).reshape(lat.V, 4, 3, 3)
    for s in (0, 7, 100, lat.V - 1):
        for mu in range(4):
            exp = sl.expm(eps * A[s, mu]) @ u[s, mu]
            assert np.abs(got[s, mu] - exp).max() < 1e-13


@pytest.mark.parametrize("scheme", [0, 1])
def test_md_energy_error_is_second_order(fields, scheme):
    lat, U, mom = fields
    H0 = hamiltonian(lat, U, mom)
    dh = []
    # trajectory length 0.5; |dH| ~ 4e-3 (leapfrog) at eps = 0.0125
    for eps, n in ((0.025, 20), (0.0125, 40), (0.00625, 80)):
        U1, m1 = O.gauge_md(lat.dims, U, mom, COEF, eps, n, scheme)
        dh.append(abs(hamiltonian(lat, U1, m1) - H0))
    r1, r2 = dh[0] / dh[1], dh[1] / dh[2]
    assert 3.6 < r1 < 4.4 and 3.8 < r2 < 4.2, dh
    if scheme == 1:
        U1, m1 = O.gauge_md(lat.dims, U, mom, COEF, 0.025, 20, 0)
        assert 5 * dh[0] < abs(hamiltonian(lat, U1, m1) - H0)


@pytest.mark.parametrize("scheme", [0, 1])
def test_md_reversible_and_unitary(fields, scheme):
    lat, U, mom = fields
    U1, m1 = O.gauge_md(lat.dims, U, mom, COEF, 0.07, 6, scheme)
    U2, m2 = O.gauge_md(lat.dims, U1, -m1, COEF, 0.07, 6, scheme)
    assert np.abs(U2 - U).max() < 1e-12 and np.abs(-m2 - mom).max() < 1e-12
    u = U1.view(np.complex128).reshape(lat.V, 4, 3, 3)
    assert np.abs(u @ F.dagger(u) - np.eye(3)).max() < 1e-13
    assert np.abs(np.linalg.det(u) - 1).max() < 1e-13
    # the trajectory moved the fields
    assert np.abs(U1 - U).max() > 1e-2


def test_md_free_field_at_rest():
    lat = F.Lattice((4, 4, 4, 4))
    one = F.to_raw(np.broadcast_to(np.eye(3, dtype=complex), (lat.V, 4, 3, 3)).copy())
    U1, m1 = O.gauge_md(lat.dims, one, np.zeros_like(one), COEF, 0.1, 3, 0)
    assert np.abs(U1 - one).max() < 1e-15 and np.abs(m1).max() < 1e-15
