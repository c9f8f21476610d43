"""Asqtad / HISQ fermion links built on the GPU from the thin gauge field
(QPHIX_*3_asqtad_create_L_from_G, QPHIX_*3_hisq_create_L_from_G) against the
oracle's path-by-path construction (pinned in tests/test_oracle_links.py).
Bars: FP64 1e-13, FP32 fields 1e-6 relative; the handle's Dslash equals the
oracle Dslash on the oracle's links."""
import numpy as np
import pytest

import oracle as O
from conftest import lattice_session
from helpers import half, rel_l2
from pyqphix import api as A
from pyqphix import fields as F

pytestmark = pytest.mark.gpu
DIMS = (8, 8, 4, 8)
NP = {"D": np.float64, "F": np.float32}


def fields(seed=11):
    lat = F.Lattice(DIMS)
    U = F.to_raw(F.gauge_field(lat, "weak", seed, eps=0.5))
    return lat, U


@pytest.mark.parametrize("prec", ["D", "F"])
@pytest.mark.parametrize("kind", ["asqtad", "hisq", "fat7_naik1"])
@pytest.mark.parametrize("periodic", [False, True])
def test_links_from_gauge_match_oracle(prec, kind, periodic):
    s = lattice_session(DIMS)
    lat, U = fields()
    G = A.Gauge(s, U.astype(NP[prec]), prec=prec)
    print(f"[case] link prep {kind} prec={prec} periodic={periodic}", flush=True)
    if kind == "hisq":
        L = A.LinksFromGauge(G, hisq=True, naik_eps=0.05, periodic_t=periodic)
        fo, lo = O.links_from_gauge(DIMS, U, hisq=True, naik_eps=0.05, apbc=not periodic)
    else:
        # fat7 with Naik -1 * U U U: SU(3) times a sign, so stored as 12 reals
        c = O.ASQTAD if kind == "asqtad" else (1 / 8, -1.0, 1 / 16, 1 / 64, 1 / 384, 0.0)
        L = A.LinksFromGauge(G, coeffs=c, periodic_t=periodic)
        fo, lo = O.links_from_gauge(DIMS, U, coef=c, apbc=not periodic)
    assert L.info.status == 0
    assert L.recon == (12 if kind == "fat7_naik1" else 18)
    fg, lg = L.extract()
    tol = 1e-13 if prec == "D" else 1e-6
    assert rel_l2(fg, fo) < tol and rel_l2(lg, lo) < tol
    # the operator: Dslash with the device-built links = oracle Dslash with the oracle's
    fl, ll = fo.view(np.complex128).reshape(lat.V, 4, 3, 3), lo.view(np.complex128).reshape(lat.V, 4, 3, 3)
    orc = O.Oracle(DIMS, fo, lo, F.to_raw(F.back_links(lat, fl, 1)), F.to_raw(F.back_links(lat, ll, 3)))
    psi = F.to_raw(F.gaussian_source(lat, 5))
    vin = A.Vector(s, psi.astype(NP[prec]), A.EVENODD, prec)
    vout = A.Vector(s, None, A.EVENODD, prec)
    for cb, par in ((0, A.EVEN), (1, A.ODD)):
        A.dslash(L, vout, vin, par)
        ref = orc.dslash(half(psi, lat, 1 - cb), cb)
        assert rel_l2(half(vout.extract(), lat, cb), ref) < (1e-13 if prec == "D" else 2e-6)
    for o in (vin, vout, L, G):
        o.free()


def test_links_from_gauge_solve(qlib):
    """A CG on the device-built HISQ links reaches the FP64 true residual of
    the oracle operator on the oracle's links."""
    s = lattice_session(DIMS)
    lat, U = fields(21)
    G = A.Gauge(s, U)
    L = A.LinksFromGauge(G, hisq=True)
    fo, lo = O.links_from_gauge(DIMS, U, hisq=True)
    fl, ll = fo.view(np.complex128).reshape(lat.V, 4, 3, 3), lo.view(np.complex128).reshape(lat.V, 4, 3, 3)
    orc = O.Oracle(DIMS, fo, lo, F.to_raw(F.back_links(lat, fl, 1)), F.to_raw(F.back_links(lat, ll, 3)))
    b = F.to_raw(F.gaussian_source(lat, 22, parity=0))
    vb, vx = A.Vector(s, b, A.EVEN), A.Vector(s, None, A.EVEN)
    ret, info, res = A.invert(L, vx, vb, 0.05, A.EVEN, 1e-20)
    xo, io = orc.invert(b, 0.05, O.EVEN, 1e-20)
    assert info.status == 0 and res.final_rsq < 1e-20
    assert abs(res.final_iter - io["final_iter"]) <= max(2, 0.05 * io["final_iter"])
    assert rel_l2(half(vx.extract(), lat, 0), half(xo, lat, 0)) < 1e-9
    for o in (vb, vx, L, G):
        o.free()
