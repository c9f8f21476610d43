"""Pins the oracle's asqtad / HISQ link construction (oracle_links_from_gauge).

* U = 1: fat = c_tot eta b 1 with c_tot = c1 + 6 c3 + 24 c5 + 48 c7 + 6 cL
  (9/8 for asqtad and HISQ), Naik = cN eta b3 1; the free Dslash with them
  has the spectrum 4m^2 + sum (2 c_tot sin p + 2 cN sin 3p)^2.
* Pure gauge U_mu(x) = g(x) g(x+mu)^dag: every path from x to x+mu gives
  g(x) g(x+mu)^dag, so fat = c_tot eta b g(x) g(x+mu)^dag (HISQ: the fat7
  level is c_tot' = 1 times a unitary matrix, so the projection keeps it).
* Gauge covariance on random fields, asqtad and HISQ.
* A uniform abelian magnetic field in the x-y plane, U_x(x) = exp(i B y):
  each path's phase is its enclosed flux, so fat_x = exp(iBy) [c1 + c3 (4 +
  2 cos B) + c5 (8 + 16 cos B) + 48 c7 cos B + cL (4 + 2 cos 2B)], the same for
  y links, c_tot for z and t links; the Naik x link is cN exp(3iBy).
"""
import numpy as np
import pytest

import oracle as O
from helpers import allowed_momenta, half, rel_l2
from pyqphix import fields as F

ASQ = O.ASQTAD


def ctot(c):
    return c[0] + 6 * c[2] + 24 * c[3] + 48 * c[4] + 6 * c[5]


def as_links(raw, lat):
    return raw.view(np.complex128).reshape(lat.V, 4, 3, 3)


def phases(lat, hop, apbc=True):
    return np.stack([lat.eta(mu) * lat.bc(mu, hop, apbc) for mu in range(4)], axis=1)


@pytest.mark.parametrize("hisq", [False, True])
@pytest.mark.parametrize("apbc", [True, False])
def test_unit_field_closed_form(hisq, apbc):
    lat = F.Lattice((4, 4, 4, 8))
    one = np.broadcast_to(np.eye(3, dtype=complex), (lat.V, 4, 3, 3)).copy()
    fat, lng = O.links_from_gauge(lat.dims, F.to_raw(one), hisq=hisq, naik_eps=0.1 if hisq else 0.0, apbc=apbc)
    cn = -(1.1) / 24 if hisq else ASQ[1]
    ef = 9.0 / 8 * phases(lat, 1, apbc)[..., None, None] * np.eye(3)
    en = cn * phases(lat, 3, apbc)[..., None, None] * np.eye(3)
    assert np.abs(as_links(fat, lat) - ef).max() < 1e-14
    assert np.abs(as_links(lng, lat) - en).max() < 1e-14


def test_unit_field_free_spectrum():
    """The oracle Dslash (pinned separately) with the built links shows the
    free spectrum of c1 = 9/8, c3 = -1/24."""
    lat = F.Lattice((8, 4, 4, 8))
    one = np.broadcast_to(np.eye(3, dtype=complex), (lat.V, 4, 3, 3)).copy()
    fat, lng = O.links_from_gauge(lat.dims, F.to_raw(one))
    fl, ll = as_links(fat, lat), as_links(lng, lat)
    orc = O.Oracle(lat.dims, fat, lng, F.to_raw(F.back_links(lat, fl, 1)), F.to_raw(F.back_links(lat, ll, 3)))
    m2 = 0.04
    for p in allowed_momenta(lat, 3):
        chi = F.to_raw(F.plane_wave(lat, p, 0, parity=0))
        h = half(chi, lat, 0)
        got = m2 * h - orc.dslash(orc.dslash(h, 1), 0)
        lam = F.free_eigenvalue(p, 9.0 / 8, -1.0 / 24, m2)
        assert rel_l2(got, lam * h) < 1e-13


def gauge_rot(lat, seed):
    return F.haar_su3(np.random.default_rng(seed), lat.V)


@pytest.mark.parametrize("hisq", [False, True])
def test_pure_gauge(hisq):
    lat = F.Lattice((4, 4, 4, 4))
    g = gauge_rot(lat, 5)
    U = np.empty((lat.V, 4, 3, 3), dtype=complex)
    for mu in range(4):
        U[:, mu] = g @ F.dagger(g[lat.shifted(mu, 1)])
    fat, lng = O.links_from_gauge(lat.dims, F.to_raw(U), hisq=hisq)
    for mu in range(4):
        ef = 9.0 / 8 * (lat.eta(mu) * lat.bc(mu, 1))[:, None, None] * (g @ F.dagger(g[lat.shifted(mu, 1)]))
        en = -1.0 / 24 * (lat.eta(mu) * lat.bc(mu, 3))[:, None, None] * (g @ F.dagger(g[lat.shifted(mu, 3)]))
        assert np.abs(as_links(fat, lat)[:, mu] - ef).max() < 1e-13
        assert np.abs(as_links(lng, lat)[:, mu] - en).max() < 1e-13


@pytest.mark.parametrize("hisq", [False, True])
def test_gauge_covariance(hisq):
    lat = F.Lattice((4, 4, 4, 4))
    U = F.gauge_field(lat, "weak", 7, eps=0.5)
    g = gauge_rot(lat, 8)
    Ug = np.empty_like(U)
    for mu in range(4):
        Ug[:, mu] = g @ U[:, mu] @ F.dagger(g[lat.shifted(mu, 1)])
    f1, n1 = O.links_from_gauge(lat.dims, F.to_raw(U), hisq=hisq)
    f2, n2 = O.links_from_gauge(lat.dims, F.to_raw(Ug), hisq=hisq)
    for mu in range(4):
        ef = g @ as_links(f1, lat)[:, mu] @ F.dagger(g[lat.shifted(mu, 1)])
        en = g @ as_links(n1, lat)[:, mu] @ F.dagger(g[lat.shifted(mu, 3)])
        assert np.abs(as_links(f2, lat)[:, mu] - ef).max() < 1e-13
        assert np.abs(as_links(n2, lat)[:, mu] - en).max() < 1e-13
    if hisq:  # the projected field is unitary: its fat links differ from asqtad's
        fa, _ = O.links_from_gauge(lat.dims, F.to_raw(U))
        assert np.abs(fa - f1).max() > 1e-3


def test_abelian_magnetic_field():
    lat = F.Lattice((4, 12, 4, 4))
    B = 2 * np.pi / 12
    U = np.broadcast_to(np.eye(3, dtype=complex), (lat.V, 4, 3, 3)).copy()
    U[:, 0] *= np.exp(1j * B * lat.y)[:, None, None]
    c = ASQ
    fat, lng = O.links_from_gauge(lat.dims, F.to_raw(U), coef=c, apbc=False)
    inplane = c[0] + c[2] * (4 + 2 * np.cos(B)) + c[3] * (8 + 16 * np.cos(B)) + 48 * c[4] * np.cos(B) \
        + c[5] * (4 + 2 * np.cos(2 * B))
    fl, nl = as_links(fat, lat), as_links(lng, lat)
    for mu in range(4):
        eta = lat.eta(mu)[:, None, None]
        ph = np.exp(1j * B * lat.y)[:, None, None] if mu == 0 else 1.0
        amp = inplane if mu < 2 else ctot(c)
        assert np.abs(fl[:, mu] - eta * amp * ph * np.eye(3)).max() < 1e-13, mu
        nph = np.exp(3j * B * lat.y)[:, None, None] if mu == 0 else 1.0
        assert np.abs(nl[:, mu] - eta * c[1] * nph * np.eye(3)).max() < 1e-13
