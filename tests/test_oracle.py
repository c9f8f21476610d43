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


def _rr(orc, r, x):
    return orc.relative_residue(r, x)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_relative_residue_closed_forms(dtype):
    """relative_residue = sqrt(2 sum_x |r(x)|^2/|x(x)|^2 / V) over one parity
    (ks_blas_utils_c.cpp:110-149), pinned by values fixed without its formula:
    r = c x at every site gives |c|; r = 0 gives 0; x = 0 with r != 0 at every
    site gives sqrt(2 V_h / V) = 1; a mixture of the three kinds of site gives
    the hand-computed 1.5."""
    lat = F.Lattice((4, 4, 4, 8))
    orc = make_oracle(lat, random_links(lat, seed=1), dtype=dtype)
    tol = 1e-13 if dtype == np.float64 else 2e-6
    rng = np.random.default_rng(3)
    x = rng.standard_normal(6 * lat.vh).astype(dtype)
    for c in (0.5, 3.0, 1e-4, 2.0 ** -20):
        got = _rr(orc, (c * x.astype(np.float64)).astype(dtype), x)
        assert abs(got - c) < tol * c, (c, got)
    # a complex factor c acts per colour as a rotation: |c r| = |c||r|
    xc = x.astype(np.float64).view(np.complex128)
    rc = ((0.6 - 0.8j) * 2.5 * xc).view(np.float64).astype(dtype)
    assert abs(_rr(orc, rc, x) - 2.5) < tol * 2.5
    assert _rr(orc, np.zeros_like(x), x) == 0.0
    assert abs(_rr(orc, x, np.zeros_like(x)) - 1.0) < tol
    # sites: first half r = 2x (term 4), next quarter r = 0 (skipped),
    # last quarter x = 0 (term 1): sqrt(2 (vh/2 * 4 + vh/4) / (2 vh)) = 1.5
    r = (2 * x.astype(np.float64)).astype(dtype).reshape(lat.vh, 6)
    xm = x.copy().reshape(lat.vh, 6)
    h, q = lat.vh // 2, lat.vh // 4
    r[h:h + q] = 0
    xm[h + q:] = 0
    assert abs(_rr(orc, r.ravel(), xm.ravel()) - 1.5) < tol * 1.5
    # a planted swap of r and x changes the mixed case (x = 0 sites would be
    # skipped, r = 0 sites count 1): the function is not symmetric
    assert abs(_rr(orc, xm.ravel(), r.ravel()) - 1.5) > 0.1


@pytest.mark.parametrize("parity", [O.EVEN, O.ODD])
def test_cg_relative_residue_plane_wave(parity):
    """CG with the relative-residue stop alone (resid = 0) on the free field,
    where everything is known in closed form: with initial guess x0 = chi and
    source b = (lambda + kappa) chi for a plane wave chi of eigenvalue lambda,
    the first true residual is r = kappa chi = kappa x0, so final_rel = |kappa|
    exactly.  relresid above |kappa| stops there (1 iteration by the
    reference's counting); below it, CG converges in one step (3 iterations)
    to x = b / lambda with a relative residue at rounding level."""
    lat = F.Lattice((8, 4, 4, 8))
    c1, c3 = 0.9, -0.07
    fat, lng = F.free_links(lat, c1, c3)
    orc = make_oracle(lat, (fat, lng, F.back_links(lat, fat, 1), F.back_links(lat, lng, 3)))
    cb = 0 if parity == O.EVEN else 1
    p = allowed_momenta(lat, 3)[2]
    m = 0.1
    lam = F.free_eigenvalue(p, c1, c3, 4 * m * m)
    chi = F.to_raw(F.plane_wave(lat, p, 1, parity=cb))
    for kappa in (0.03, -0.2):
        b = (lam + kappa) * chi
        x, info = orc.invert(b, m, parity, resid=0.0, relresid=0.5, x0=chi)
        assert info["final_iter"] == 1 and info["ret"] == 1
        assert abs(info["final_rel"] - abs(kappa)) < 1e-12 * abs(kappa)
        assert rel_l2(x, chi) == 0.0
        x, info = orc.invert(b, m, parity, resid=0.0, relresid=1e-3, x0=chi)
        assert info["final_iter"] == 3
        assert info["final_rel"] < 1e-12 and info["size_relr"] < 1e-12
        assert rel_l2(half(x, lat, cb), half(b, lat, cb) / lam) < 1e-13
    # with both stops, the absolute one can hold the solver after the relative
    # one is met: resid 1e-40 is never met in one step, relresid 0.5 is
    x, info = orc.invert((lam + 0.03) * chi, m, parity, resid=1e-40, relresid=0.5, x0=chi, niter=4,
                         nrestart=1)
    assert info["final_iter"] > 3


@pytest.mark.parametrize("parity", [O.EVEN, O.ODD])
def test_cg_float_matches_dense_solve(parity):
    """The F build's CG (scalars rounded to float, reductions in float as the
    reference's fptype accumulators) against a dense FP64 solve at 4^4."""
    lat = F.Lattice((4, 4, 4, 4))
    links = random_links(lat, "weak", seed=33)
    orc = make_oracle(lat, links, dtype=np.float32)
    orc64 = make_oracle(lat, links)
    cb = 0 if parity == O.EVEN else 1
    m = 0.1
    b = F.to_raw(F.gaussian_source(lat, 34, parity=cb))
    x, info = orc.invert(b.astype(np.float32), m, parity, resid=1e-10)
    A = 4 * m * m * np.eye(3 * lat.vh) - dense_ddag(orc64, lat, cb)
    xs = np.linalg.solve(A, half(b, lat, cb).view(np.complex128))
    assert x.dtype == np.float32
    assert rel_l2(half(x, lat, cb).astype(np.float64).view(np.complex128), xs) < 1e-5
    assert info["final_rsq"] < 1e-10


def test_multishift_float_matches_dense_solves():
    lat = F.Lattice((4, 4, 4, 4))
    links = random_links(lat, "weak", seed=53)
    orc = make_oracle(lat, links, dtype=np.float32)
    orc64 = make_oracle(lat, links)
    masses = np.array([0.2, 0.05, 0.1, 0.4])
    b = F.to_raw(F.gaussian_source(lat, 54, parity=0))
    xs, info = orc.invert_multi(b.astype(np.float32), masses.astype(np.float32), O.EVEN, resid=1e-10)
    DD = dense_ddag(orc64, lat, 0)
    bh = half(b, lat, 0).view(np.complex128)
    for m, x in zip(masses, xs):
        ref = np.linalg.solve(4 * m * m * np.eye(3 * lat.vh) - DD, bh)
        assert rel_l2(half(x, lat, 0).astype(np.float64).view(np.complex128), ref) < 2e-5


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("parity", [O.EVEN, O.ODD])
def test_cg_fewsums_matches_dense_solve(dtype, parity):
    """The FEWSUMS variant of the CG (single reduction per step, |r|^2 from
    the recurrence |r|^2 + 2a Re<Ap|r> + a^2 |Ap|^2) solves the same system:
    against a dense FP64 solve at 4^4, with restarts every 25 iterations."""
    lat = F.Lattice((4, 4, 4, 4))
    links = random_links(lat, "weak", seed=35)
    orc = make_oracle(lat, links, dtype=dtype)
    orc64 = orc if dtype == np.float64 else make_oracle(lat, links)
    cb = 0 if parity == O.EVEN else 1
    m = 0.1
    b = F.to_raw(F.gaussian_source(lat, 36, parity=cb))
    resid = 1e-24 if dtype == np.float64 else 1e-10
    A = 4 * m * m * np.eye(3 * lat.vh) - dense_ddag(orc64, lat, cb)
    xs = np.linalg.solve(A, half(b, lat, cb).view(np.complex128))
    for niter in (1000, 25):
        x, info = orc.invert(b.astype(dtype), m, parity, resid=resid, niter=niter, nrestart=50,
                             fewsums=True)
        # FP32: the stop is tested on the recurrence |r|^2, ~1e-10 from the true
        # one, so the error is a few times the plain CG's 1e-5
        tol = 1e-10 if dtype == np.float64 else 5e-5
        assert rel_l2(half(x, lat, cb).astype(np.float64).view(np.complex128), xs) < tol
        assert info["final_rsq"] < resid


def test_cg_fewsums_plane_wave_and_counts():
    """Free-field plane wave: one step, 3 iterations, x = b / lambda; on the
    8^4 random-link system the iteration count stays within 5% of the plain
    CG's, and the final (true, restart-evaluated) residual meets the target.
    The recurrence |r|^2 + 2a Re<Ap|r> + a^2 |Ap|^2 cancels to rounding level,
    ~1e-16 |r|^2 in one step, so the one-step known answer takes a target
    above that (1e-12); with 1e-24 the recurrence cannot register the exact
    drop and the solve runs to its restart interval, where the true residual
    ends it."""
    lat = F.Lattice((8, 4, 4, 8))
    c1, c3 = 0.9, -0.07
    fat, lng = F.free_links(lat, c1, c3)
    orc = make_oracle(lat, (fat, lng, F.back_links(lat, fat, 1), F.back_links(lat, lng, 3)))
    p = allowed_momenta(lat, 2)[0]
    b = F.to_raw(F.plane_wave(lat, p, 2, parity=0))
    x, info = orc.invert(b, 0.1, O.EVEN, resid=1e-12, fewsums=True)
    assert info["final_iter"] == 3
    _, info24 = orc.invert(b, 0.1, O.EVEN, resid=1e-24, fewsums=True, niter=50)
    assert info24["final_iter"] > 51 and info24["final_rsq"] < 1e-24
    lam = F.free_eigenvalue(p, c1, c3, 0.04)
    assert rel_l2(half(x, lat, 0), half(b, lat, 0) / lam) < 1e-12
    lat = F.Lattice((8, 8, 8, 8))
    orc = make_oracle(lat, random_links(lat, "haar", 1001))
    b = F.to_raw(F.gaussian_source(lat, 77, parity=0))
    _, i0 = orc.invert(b, 0.025, O.EVEN, resid=1e-16)
    x, i1 = orc.invert(b, 0.025, O.EVEN, resid=1e-16, fewsums=True)
    assert abs(i1["final_iter"] - i0["final_iter"]) <= 0.05 * i0["final_iter"]
    assert i1["final_rsq"] < 1e-16
