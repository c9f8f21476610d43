"""Symanzik gauge force on the GPU (C ABI) vs the oracle, plus finite
differences of the action at the 24^4 size."""
import numpy as np
import pytest

import oracle as O
from conftest import lattice_session
from pyqphix import api as A
from pyqphix import fields as F

pytestmark = pytest.mark.gpu

MILC_B6 = F.symanzik_coeffs(6.0, 0.86)   # beta = 6.0, u0 = 0.86


# (zero coefficients drop walks, groups and their closing steps from the step tables)
@pytest.mark.parametrize("coef", [MILC_B6, (1.0, 0.0, 0.0), (0.3, -0.2, 0.1), (0.0, -0.2, 0.0), (0.0, 0.0, 0.1)])
def test_gauge_force_matches_oracle(coef):
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    U = F.gauge_field(lat, "weak", 21, eps=0.3)
    m0 = F.antihermitian_traceless(np.random.default_rng(4), (lat.V, 4))
    raw_u, raw_m = F.to_raw(U), F.to_raw(m0)
    G = A.Gauge(s, raw_u)
    Fm = A.Force(s, raw_m)
    # the device layout is complex-major; extract returns the raw array exactly
    np.testing.assert_array_equal(G.extract(), raw_u)
    odd = np.zeros_like(raw_u)
    s.fn("D", "extract_G_to_raw")(A.ptr(odd), G.h, A.ODD)
    np.testing.assert_array_equal(odd[36 * lat.V:], raw_u[36 * lat.V:])
    assert not odd[:36 * lat.V].any()
    # C13 round trip; Im A22 is rebuilt as -Im A00 - Im A11 (rounding only)
    np.testing.assert_allclose(Fm.extract(), raw_m, rtol=0, atol=1e-14)
    info = A.gauge_force(G, Fm, coef, 0.02)
    assert info.status == 0
    got = Fm.extract()
    exp, _ = O.gauge_force(dims, raw_u, coef, 0.02, raw_m)
    d = got - raw_m
    e = exp - raw_m
    assert np.linalg.norm(d - e) < 1e-12 * np.linalg.norm(e)
    G.free()
    Fm.free()


def test_gauge_force_two_row_links():
    """SU(3) links: the force reads rows 0-1 and rebuilds row 2; the result
    agrees with the all-rows kernel to rounding, and a field that is not SU(3)
    keeps all rows and still matches the oracle."""
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw_u = F.to_raw(F.gauge_field(lat, "weak", 24, eps=0.3))
    z = np.zeros(72 * lat.V)
    got = {}
    try:
        for mode in (0, 18):
            s.lib.QPHIX_MI355X_set_gauge_recon(mode)
            G = A.Gauge(s, raw_u)
            assert s.fn("D", "gauge_links_recon")(G.h) == (12 if mode == 0 else 18)
            Fm = A.Force(s, z)
            A.gauge_force(G, Fm, MILC_B6, 0.02)
            got[mode] = Fm.extract()
            np.testing.assert_array_equal(G.extract(), raw_u)
            G.free()
            Fm.free()
    finally:
        s.lib.QPHIX_MI355X_set_gauge_recon(0)
    assert np.linalg.norm(got[0] - got[18]) < 1e-13 * np.linalg.norm(got[18])
    # a general complex field (not unitary): all rows are kept
    rng = np.random.default_rng(25)
    bad = raw_u + 0.05 * rng.standard_normal(raw_u.shape)
    G = A.Gauge(s, bad)
    assert s.fn("D", "gauge_links_recon")(G.h) == 18
    Fm = A.Force(s, z)
    A.gauge_force(G, Fm, MILC_B6, 0.02)
    exp, _ = O.gauge_force(dims, bad, MILC_B6, 0.02, z)
    assert np.linalg.norm(Fm.extract() - exp) < 1e-12 * np.linalg.norm(exp)
    G.free()
    Fm.free()


@pytest.mark.parametrize("dims", [(8, 8, 8, 8), (8, 12, 4, 16)])
@pytest.mark.parametrize("coef", [MILC_B6, (0.3, -0.2, 0.1), (0.0, 0.0, 0.1), (0.0, -0.2, 0.0)])
def test_gauge_force_staples_match_walks(dims, coef):
    """SU(3) links: the staple-product kernel (rectangles and parallelograms
    formed from stored plaquette staples) against the loop-walk kernel on the
    same two-row field -- every link of both parities, to rounding."""
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw_u = F.to_raw(F.gauge_field(lat, "weak", 26, eps=0.4))
    m0 = F.to_raw(F.antihermitian_traceless(np.random.default_rng(27), (lat.V, 4)))
    G = A.Gauge(s, raw_u)
    assert s.fn("D", "gauge_links_recon")(G.h) == 12
    got = {}
    try:
        for on in (0, 1):
            s.lib.QPHIX_MI355X_set_gauge_force_staples(on)
            Fm = A.Force(s, m0)
            print(f"[case] gauge force staples={on} dims={dims} coef={coef}", flush=True)
            A.gauge_force(G, Fm, coef, 0.05)
            got[on] = Fm.extract() - m0
            Fm.free()
    finally:
        s.lib.QPHIX_MI355X_set_gauge_force_staples(1)
    assert np.linalg.norm(got[1] - got[0]) < 1e-13 * np.linalg.norm(got[0])
    G.free()


def test_gauge_force_single_precision():
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    U = F.gauge_field(lat, "weak", 22, eps=0.3)
    raw_u = F.to_raw(U)
    z = np.zeros(72 * lat.V)
    G = A.Gauge(s, raw_u.astype(np.float32), prec="F")
    Fm = A.Force(s, z.astype(np.float32), prec="F")
    A.gauge_force(G, Fm, MILC_B6, 0.02)
    exp, _ = O.gauge_force(dims, raw_u, MILC_B6, 0.02, z)
    assert np.linalg.norm(Fm.extract() - exp) < 1e-5 * np.linalg.norm(exp)
    G.free()
    Fm.free()


def su3_exp(X):
    w, v = np.linalg.eigh(-1j * X)
    return (v * np.exp(1j * w)[None, :]) @ F.dagger(v)


def test_gauge_force_fd_24_4():
    """24^4 weak-field SU(3), beta = 6.0, u0 = 0.86: the momentum change equals
    -eps times the derivative of the action (central differences, h = 1e-5)
    on 64 random links."""
    dims = (24, 24, 24, 24)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    U = F.gauge_field(lat, "weak", 1004, eps=0.3)
    raw_u = F.to_raw(U)
    z = np.zeros(72 * lat.V)
    G = A.Gauge(s, raw_u)
    Fm = A.Force(s, z)
    eps = 1.0
    A.gauge_force(G, Fm, MILC_B6, eps)
    dA = Fm.extract().view(np.complex128).reshape(lat.V, 4, 3, 3)   # = -eps TA(W)
    rng = np.random.default_rng(8)
    h = 1e-5
    worst = 0.0
    U = np.ascontiguousarray(U)
    Uview = U.view(np.float64).reshape(-1)   # raw layout, shares memory with U
    for _ in range(64):
        site = int(rng.integers(lat.V))
        mu = int(rng.integers(4))
        X = F.antihermitian_traceless(rng, ())
        box = [int(lat.x[site]), int(lat.y[site]), int(lat.z[site]), int(lat.t[site])]
        S = []
        u2 = U[site, mu].copy()
        for sg in (1, -1):
            U[site, mu] = su3_exp(sg * h * X) @ u2
            S.append(O.gauge_action(dims, Uview, MILC_B6, box, 2, mu))
        U[site, mu] = u2
        fd = (S[0] - S[1]) / (2 * h)
        # dS/dt = -Re Tr(X W) = -Re Tr(X TA(W)) = Re Tr(X dA) / eps
        an = np.trace(X @ dA[site, mu]).real / eps
        worst = max(worst, abs(fd - an) / max(1.0, abs(an)))
    assert worst < 1e-8
    G.free()
    Fm.free()


def test_parity_mismatch_is_an_error():
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw_u = F.to_raw(F.gauge_field(lat, "weak", 23))
    s.lib.QPHIX_MI355X_set_abort_on_error(0)
    s.lib.QPHIX_MI355X_clear_error()
    try:
        G = A.Gauge(s, raw_u, parity=A.ODD)
        Fm = A.Force(s, np.zeros(72 * lat.V), parity=A.EVENODD)
        info = A.gauge_force(G, Fm, MILC_B6, 0.1)
        assert info.status == 1 and s.lib.last_error()[0] == 1
        G.free()
        Fm.free()
    finally:
        s.lib.QPHIX_MI355X_clear_error()
        s.lib.QPHIX_MI355X_set_abort_on_error(1)
