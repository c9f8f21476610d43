"""Pins the oracle's fermion-force outer products (oracle_fermion_oprod_d).

The reference has no fermion force, so the oracle's outer products are pinned
by the action they differentiate,
    S(U) = sum_k c_k Re phi^dag X_k,   X_k = (sigma_k - D_eo D_oe)^{-1} phi,
through central finite differences in random complex directions E on single
forward links (fat and Naik), with MILC's tied back links -U^dag(x-mu):
    dS/dh = 2 Re Tr[E O(y,mu)].
S itself is a sum over the whole lattice, so S(U+hE) - S(U-hE) is formed
without cancellation from the exact identity
    S+ - S- = sum_k c_k Re X+_k^dag (D+_eo D+_oe - D-_eo D-_oe) X-_k,
whose middle factor vanishes identically away from the perturbed link.
A transposed or conjugated outer product, a dropped parity sign or a wrong
neighbour fails these checks (the planted-mistake test shows one).
"""
import numpy as np
import pytest

import oracle as O
from helpers import full_from_half, half
from pyqphix import fields as F

DIMS = (8, 8, 8, 8)
MASSES = [0.1, 0.25]
COEF = [0.6, 1.4]


def build(lat, fat, lng):
    raw = [F.to_raw(fat), F.to_raw(lng), F.to_raw(F.back_links(lat, fat, 1)),
           F.to_raw(F.back_links(lat, lng, 3))]
    return O.Oracle(lat.dims, *raw)


def solutions(orc, lat, phi):
    xs = []
    for m in MASSES:
        x, info = orc.invert(phi, m, O.EVEN, 1e-26, niter=3000, nrestart=10)
        assert info["final_rsq"] < 1e-26
        xs.append(x)
    return xs


def ddx(orc, lat, x_full):
    """D_eo D_oe on the even half of x."""
    return orc.dslash(orc.dslash(half(x_full, lat, 0), 1), 0)


def fd_directional(lat, fat, lng, phi, which, y, mu, E, h=1e-5):
    """[S(U+hE) - S(U-hE)] / 2h for forward link (y, mu) of `which` ('fat'/'long')."""
    def perturbed(sign):
        f, l = fat.copy(), lng.copy()
        (f if which == "fat" else l)[y, mu] += sign * h * E
        return build(lat, f, l)
    op, om = perturbed(+1), perturbed(-1)
    xp, xm = solutions(op, lat, phi), solutions(om, lat, phi)
    ds = 0.0
    for c, a, b in zip(COEF, xp, xm):
        w = ddx(op, lat, b) - ddx(om, lat, b)
        ds += c * float(np.real(np.vdot(half(a, lat, 0).view(np.complex128), w.view(np.complex128))))
    op.close()
    om.close()
    return ds / (2 * h)


@pytest.fixture(scope="module")
def setup():
    lat = F.Lattice(DIMS)
    fat, lng = F.make_links(lat, "weak", 2024, eps=0.3)
    phi = F.to_raw(F.gaussian_source(lat, 77, parity=0))
    orc = build(lat, fat, lng)
    xs = solutions(orc, lat, phi)
    o1, o3 = orc.fermion_oprod(COEF, xs)
    orc.close()
    O1 = o1.view(np.complex128).reshape(lat.V, 4, 3, 3)
    O3 = o3.view(np.complex128).reshape(lat.V, 4, 3, 3)
    return lat, fat, lng, phi, xs, O1, O3


def _cases(lat, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        y = int(rng.integers(0, lat.V))
        mu = int(rng.integers(0, 4))
        E = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
        out.append((y, mu, E))
    return out


@pytest.mark.parametrize("which", ["fat", "long"])
def test_oprod_is_the_link_derivative_of_the_action(setup, which):
    lat, fat, lng, phi, xs, O1, O3 = setup
    Om = O1 if which == "fat" else O3
    # two links of each parity and several directions
    cases = _cases(lat, 4, 5 if which == "fat" else 6)
    cases[0] = (3, cases[0][1], cases[0][2])                 # an even site
    cases[1] = (lat.vh + 5, cases[1][1], cases[1][2])        # an odd site
    for y, mu, E in cases:
        pred = 2.0 * float(np.real(np.trace(E @ Om[y, mu])))
        fd = fd_directional(lat, fat, lng, phi, which, y, mu, E)
        scale = 2.0 * np.linalg.norm(E) * np.linalg.norm(Om[y, mu])
        assert abs(fd - pred) < 1e-9 * scale, (which, y, mu, fd, pred)


def test_planted_transpose_fails(setup):
    """The FD check discriminates: O^T (a transposed operand) disagrees."""
    lat, fat, lng, phi, xs, O1, O3 = setup
    y, mu, E = _cases(lat, 1, 9)[0]
    fd = fd_directional(lat, fat, lng, phi, "fat", y, mu, E)
    wrong = 2.0 * float(np.real(np.trace(E @ O1[y, mu].T)))
    scale = 2.0 * np.linalg.norm(E) * np.linalg.norm(O1[y, mu])
    assert abs(fd - wrong) > 1e-3 * scale


def test_oprod_linear_in_coefficients(setup):
    """Outer products of a sum of terms are the sum of the per-term ones, and
    parity signs: O(y) on odd sites is minus the unsigned product."""
    lat, fat, lng, phi, xs, O1, O3 = setup
    orc = build(lat, fat, lng)
    a1, a3 = orc.fermion_oprod([COEF[0]], [xs[0]])
    b1, b3 = orc.fermion_oprod([COEF[1]], [xs[1]])
    np.testing.assert_allclose(a1 + b1, O1.view(np.float64).ravel(), rtol=0, atol=1e-12 * np.abs(a1).max())
    np.testing.assert_allclose(a3 + b3, O3.view(np.float64).ravel(), rtol=0, atol=1e-12 * np.abs(a3).max())
    # direct numpy form of one term at one odd site (global coordinates)
    x = xs[0]
    y_odd = orc.dslash(half(x, lat, 0), 1)
    psi = np.concatenate([half(x, lat, 0), y_odd]).view(np.complex128).reshape(lat.V, 3)
    s = lat.vh + 17
    for mu in range(4):
        ref1 = -COEF[0] * np.outer(psi[lat.shifted(mu, 1)[s]], np.conj(psi[s]))
        ref3 = -COEF[0] * np.outer(psi[lat.shifted(mu, 3)[s]], np.conj(psi[s]))
        got1 = a1.view(np.complex128).reshape(lat.V, 4, 3, 3)[s, mu]
        got3 = a3.view(np.complex128).reshape(lat.V, 4, 3, 3)[s, mu]
        np.testing.assert_allclose(got1, ref1, rtol=1e-13, atol=1e-13 * np.abs(ref1).max())
        np.testing.assert_allclose(got3, ref3, rtol=1e-13, atol=1e-13 * np.abs(ref3).max())
    orc.close()
