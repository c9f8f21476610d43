"""Pins the CPU oracle (oracle/qphix_oracle.cpp) before it is trusted as the
checker of the HIP library.

The reference ships no tests and no golden data, and its Dslash arithmetic
lives in the absent milc-qphix-codegen repository, so each oracle function is
pinned by what mathematics fixes:
  * neighbour tables   <- global-coordinate arithmetic (Lattice.shifted)
  * Dslash             <- free-field spectrum m^2 + sum(2c1 sin p + 2c3 sin 3p)^2,
                          anti-Hermiticity, gauge covariance, a coordinate-based
                          numpy operator; a planted back-link sign error must fail
  * CG                 <- plane-wave known answer (x = b/lambda, 3 iterations by the
                          reference's counting), dense solves at 4^4, restarts
  * multi-shift CG     <- dense solves per shift, plane-wave known answer
"""
import numpy as np
import pytest

import oracle as O
from helpers import (allowed_momenta, cdot, full_from_half, half, np_dslash, random_links,
                     rel_l2)
from pyqphix import fields as F


def make_oracle(lat, links, dtype=np.float64):
    raw = [F.to_raw(a) for a in links]
    return O.Oracle(lat.dims, *raw, dtype=dtype)


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (8, 4, 4, 8), (4, 8, 12, 4)])
def test_neighbour_table_matches_coordinates(dims):
    lat = F.Lattice(dims)
    orc = make_oracle(lat, random_links(lat, seed=1))
    for cb in (0, 1):
        tab = orc.neighbours(cb)
        sites = np.arange(lat.vh) + cb * lat.vh
        for mu in range(4):
            for fb, hop in ((0, -1), (1, 1)):
                exp1 = lat.shifted(mu, hop)[sites] - (1 - cb) * lat.vh
                exp3 = lat.shifted(mu, 3 * hop)[sites] - (1 - cb) * lat.vh
                np.testing.assert_array_equal(tab[:, 2 * mu + fb], exp1)
                np.testing.assert_array_equal(tab[:, 8 + 2 * mu + fb], exp3)


def apply_ddag(orc, lat, chi_even_raw_half):
    """D_eo D_oe on an even half vector."""
    t = orc.dslash(chi_even_raw_half, 1)
    return orc.dslash(t, 0)


@pytest.mark.parametrize("sign", [-1, +1])
def test_free_field_spectrum(sign):
    """(4m^2 - D_eo D_oe) chi_p = [4m^2 + sum (2c1 sin p + 2c3 sin 3p)^2] chi_p.
    sign=-1 is MILC's double-store back link; sign=+1 is a planted error that
    turns sin into cos and must break the identity."""
    lat = F.Lattice((8, 8, 4, 8))
    c1, c3, m = 0.9, -0.07, 0.05
    fat, lng = F.free_links(lat, c1, c3)
    fb = -sign * F.back_links(lat, fat, 1)
    lb = -sign * F.back_links(lat, lng, 3)
    orc = make_oracle(lat, (fat, lng, fb, lb))
    worst = 0.0
    for p in allowed_momenta(lat, 6):
        chi = F.plane_wave(lat, p, colour=1, parity=0)
        ce = half(F.to_raw(chi), lat, 0)
        mchi = 4 * m * m * ce - apply_ddag(orc, lat, ce)
        lam = F.free_eigenvalue(p, c1, c3, 4 * m * m)
        worst = max(worst, rel_l2(mchi, lam * ce))
    if sign == -1:
        assert worst < 1e-12
    else:
        assert worst > 1e-2


def test_anti_hermiticity():
    lat = F.Lattice((8, 4, 4, 8))
    orc = make_oracle(lat, random_links(lat, "haar", seed=3))
    rng = np.random.default_rng(5)
    chi = rng.standard_normal(6 * lat.vh)
    psi = rng.standard_normal(6 * lat.vh)
    chic = chi.view(np.complex128)
    psic = psi.view(np.complex128)
    lhs = cdot(chic, orc.dslash(psi, 0).view(np.complex128))      # <chi_e, D_eo psi_o>
    rhs = -cdot(orc.dslash(chi, 1).view(np.complex128), psic)     # -<D_oe chi_e, psi_o>
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_gauge_covariance():
    lat = F.Lattice((4, 4, 8, 4))
    fat, lng, _, _ = random_links(lat, "weak", seed=8)
    g = F.haar_su3(np.random.default_rng(9), lat.V)
    fat2, lng2 = F.gauge_transform(lat, fat, lng, g)
    orc1 = make_oracle(lat, (fat, lng, F.back_links(lat, fat, 1), F.back_links(lat, lng, 3)))
    psi = F.gaussian_source(lat, 10, parity=1)
    out1 = orc1.dslash(half(F.to_raw(psi), lat, 1), 0).view(np.complex128).reshape(-1, 3)
    orc1.close()
    orc2 = make_oracle(lat, (fat2, lng2, F.back_links(lat, fat2, 1), F.back_links(lat, lng2, 3)))
    gpsi = np.einsum("nij,nj->ni", g, psi)
    out2 = orc2.dslash(half(F.to_raw(gpsi), lat, 1), 0).view(np.complex128).reshape(-1, 3)
    expect = np.einsum("nij,nj->ni", g[:lat.vh], out1)
    assert rel_l2(out2, expect) < 1e-13


@pytest.mark.parametrize("kind", ["haar", "weak"])
def test_oracle_matches_coordinate_operator(kind):
    lat = F.Lattice((4, 8, 4, 8))
    links = random_links(lat, kind, seed=21)
    orc = make_oracle(lat, links)
    psi = F.gaussian_source(lat, 22)
    full = np_dslash(lat, *links, psi)
    for cb in (0, 1):
        got = orc.dslash(half(F.to_raw(psi), lat, 1 - cb), cb)
        exp = half(F.to_raw(full), lat, cb)
        assert rel_l2(got, exp) < 1e-14


def test_float_oracle_tracks_double():
    lat = F.Lattice((4, 4, 4, 8))
    links = random_links(lat, "weak", seed=4)
    od = make_oracle(lat, links)
    of = make_oracle(lat, links, dtype=np.float32)
    psi = half(F.to_raw(F.gaussian_source(lat, 1)), lat, 1)
    assert rel_l2(of.dslash(psi.astype(np.float32), 0), od.dslash(psi, 0)) < 1e-6


def test_cg_plane_wave_known_answer():
    """Free field + plane-wave source: CG is exact after one step.  By the
    reference's counting (C9) the first true-residual evaluation, the step and
    the final true-residual evaluation make 3 iterations."""
    lat = F.Lattice((8, 4, 4, 8))
    c1, c3, m = 0.9, -0.07, 0.1
    fat, lng = F.free_links(lat, c1, c3)
    orc = make_oracle(lat, (fat, lng, F.back_links(lat, fat, 1), F.back_links(lat, lng, 3)))
    p = allowed_momenta(lat, 1)[0]
    b = F.to_raw(F.plane_wave(lat, p, 2, parity=0))
    x, info = orc.invert(b, m, O.EVEN, resid=1e-20)
    lam = F.free_eigenvalue(p, c1, c3, 4 * m * m)
    assert info["ret"] == 3 and info["final_iter"] == 3
    assert rel_l2(half(x, lat, 0), half(b, lat, 0) / lam) < 1e-12


def dense_ddag(orc, lat, cb):
    """Dense complex matrix of D_{cb,1-cb} D_{1-cb,cb} on the half lattice."""
    n = 3 * lat.vh
    Mx = np.zeros((n, n), dtype=np.complex128)
    for k in range(n):
        e = np.zeros(n, dtype=np.complex128)
        e[k] = 1
        t = orc.dslash(e.view(np.float64).copy(), 1 - cb)
        Mx[:, k] = orc.dslash(t, cb).view(np.complex128)
    return Mx


@pytest.mark.parametrize("parity", [O.EVEN, O.ODD])
def test_cg_matches_dense_solve(parity):
    lat = F.Lattice((4, 4, 4, 4))
    orc = make_oracle(lat, random_links(lat, "weak", seed=31))
    cb = 0 if parity == O.EVEN else 1
    m = 0.1
    b = F.to_raw(F.gaussian_source(lat, 32, parity=cb))
    x, info = orc.invert(b, m, parity, resid=1e-24, niter=1000, nrestart=5)
    A = 4 * m * m * np.eye(3 * lat.vh) - dense_ddag(orc, lat, cb)
    xs = np.linalg.solve(A, half(b, lat, cb).view(np.complex128))
    assert rel_l2(half(x, lat, cb).view(np.complex128), xs) < 1e-10
    assert info["final_rsq"] < 1e-24
    # untouched parity stays zero
    assert not np.any(half(x, lat, 1 - cb))


def test_cg_restarts_and_counting():
    """Small `max` forces true-residual restarts every `max` iterations; each
    restart counts as an iteration (C9) and the answer is unchanged."""
    lat = F.Lattice((4, 4, 4, 4))
    orc = make_oracle(lat, random_links(lat, "weak", seed=41))
    b = F.to_raw(F.gaussian_source(lat, 42, parity=0))
    x1, i1 = orc.invert(b, 0.1, O.EVEN, resid=1e-20, niter=1000, nrestart=5)
    x2, i2 = orc.invert(b, 0.1, O.EVEN, resid=1e-20, niter=50, nrestart=100)
    assert i2["final_restart"] >= 2 and i1["final_restart"] >= 1
    assert i2["final_iter"] > i1["final_iter"]
    assert rel_l2(x2, x1) < 1e-8
    # the cap max*nrestart: restarts at iterations 1, 6 and 11; the one at 11
    # finds iteration >= max_cg = 10 and stops (ks_d_congrad_fn.cpp:424-427)
    x3, i3 = orc.invert(b, 0.1, O.EVEN, resid=1e-30, niter=5, nrestart=2)
    assert i3["final_iter"] == 11 and i3["final_rsq"] > 1e-30


def test_cg_initial_guess_is_used():
    lat = F.Lattice((4, 4, 4, 4))
    orc = make_oracle(lat, random_links(lat, "weak", seed=43))
    b = F.to_raw(F.gaussian_source(lat, 44, parity=0))
    x, _ = orc.invert(b, 0.1, O.EVEN, resid=1e-24)
    x2, info = orc.invert(b, 0.1, O.EVEN, resid=1e-20, x0=x)
    assert info["final_iter"] == 1  # the restart evaluation finds it converged
    assert rel_l2(x2, x) == 0.0


def test_cg_evenodd_solves_each_parity():
    lat = F.Lattice((4, 4, 4, 4))
    orc = make_oracle(lat, random_links(lat, "weak", seed=45))
    b = F.to_raw(F.gaussian_source(lat, 46))
    x, info = orc.invert(b, 0.1, O.EVENODD, resid=1e-22)
    for cb in (0, 1):
        xe, _ = orc.invert(full_from_half(half(b, lat, cb), lat, cb), 0.1,
                           O.EVEN if cb == 0 else O.ODD, resid=1e-22)
        assert rel_l2(half(x, lat, cb), half(xe, lat, cb)) < 1e-9


def test_zero_source():
    lat = F.Lattice((4, 4, 4, 4))
    orc = make_oracle(lat, random_links(lat, "weak", seed=47))
    x, info = orc.invert(np.zeros(6 * lat.V), 0.1, O.EVEN, resid=1e-10,
                         x0=np.ones(6 * lat.V))
    assert info["ret"] == 0 and not np.any(half(x, lat, 0))


def test_multishift_matches_dense_solves():
    lat = F.Lattice((4, 4, 4, 4))
    orc = make_oracle(lat, random_links(lat, "weak", seed=51))
    masses = np.array([0.2, 0.05, 0.1, 0.4])
    b = F.to_raw(F.gaussian_source(lat, 52, parity=0))
    xs, info = orc.invert_multi(b, masses, O.EVEN, resid=1e-22)
    DD = dense_ddag(orc, lat, 0)
    bh = half(b, lat, 0).view(np.complex128)
    for m, x in zip(masses, xs):
        ref = np.linalg.solve(4 * m * m * np.eye(3 * lat.vh) - DD, bh)
        assert rel_l2(half(x, lat, 0).view(np.complex128), ref) < 1e-8
    assert info["final_rsq"] <= 1e-22


def test_multishift_plane_wave():
    lat = F.Lattice((8, 4, 4, 8))
    c1, c3 = 0.9, -0.07
    fat, lng = F.free_links(lat, c1, c3)
    orc = make_oracle(lat, (fat, lng, F.back_links(lat, fat, 1), F.back_links(lat, lng, 3)))
    p = allowed_momenta(lat, 2)[1]
    b = F.to_raw(F.plane_wave(lat, p, 0, parity=1))
    masses = np.array([0.05, 0.3, 0.1])
    xs, info = orc.invert_multi(b, masses, O.ODD, resid=1e-24)
    assert info["final_iter"] == 1
    for m, x in zip(masses, xs):
        lam = F.free_eigenvalue(p, c1, c3, 4 * m * m)
        assert rel_l2(half(x, lat, 1), half(b, lat, 1) / lam) < 1e-12


def test_relative_residue_definition():
    """sqrt(2 sum_x |r(x)|^2/|x(x)|^2 / V), sites with r=0 skipped, x=0 counts 1
    (ks_blas_utils_c.cpp:110-149)."""
    lat = F.Lattice((4, 4, 4, 4))
    make_oracle(lat, random_links(lat, seed=1))
    rng = np.random.default_rng(3)
    r = rng.standard_normal(6 * lat.vh)
    x = rng.standard_normal(6 * lat.vh)
    r[:6] = 0
    x[6:12] = 0
    num = (r.reshape(-1, 6) ** 2).sum(1)
    den = (x.reshape(-1, 6) ** 2).sum(1)
    terms = np.where(num == 0, 0.0, np.where(den == 0, 1.0, num / np.where(den == 0, 1, den)))
    exp = np.sqrt(2 * terms.sum() / lat.V)
    got = O.lib().oracle_relative_residue_d(O._p(r), O._p(x))
    assert abs(got - exp) < 1e-14 * exp
