"""Gauge-field update U <- exp(eps A) U and pure-gauge MD on the GPU, against
the oracle (tests/test_oracle_md.py pins it: expm, O(eps^2) energy error,
reversibility).  Bars: FP64 1e-13 per update, 1e-11 after an MD trajectory;
FP32 fields 2e-6."""
import numpy as np
import pytest

import oracle as O
from conftest import lattice_session
from pyqphix import api as A
from pyqphix import fields as F

pytestmark = pytest.mark.gpu
COEF = (0.7, -0.11, -0.05)
DIMS = (8, 8, 4, 8)
NP = {"D": np.float64, "F": np.float32}


def make(prec, seed=21):
    s = lattice_session(DIMS)
    lat = F.Lattice(DIMS)
    U = F.to_raw(F.gauge_field(lat, "weak", seed, eps=0.5))
    mom = F.to_raw(F.antihermitian_traceless(np.random.default_rng(seed + 1), (lat.V, 4)))
    G = A.Gauge(s, U.astype(NP[prec]), prec=prec)
    Fm = A.Force(s, mom.astype(NP[prec]), prec=prec)
    return s, lat, U, mom, G, Fm


@pytest.mark.parametrize("prec", ["D", "F"])
def test_update_gauge_matches_oracle(qlib, prec):
    s, lat, U, mom, G, Fm = make(prec)
    tol = 1e-13 if prec == "D" else 2e-6
    Uref = U
    for eps in (0.3, -0.05, 2.5):
        print(f"[case] update_gauge prec={prec} eps={eps}", flush=True)
        info = A.update_gauge(G, Fm, eps)
        assert info.status == 0
        Uref = O.gauge_update(lat.dims, Uref, mom, eps)
        got = G.extract()
        assert np.abs(got - Uref).max() < tol * 10 ** (eps == 2.5)
    # links stay SU(3)
    u = G.extract().astype(np.float64).view(np.complex128).reshape(lat.V, 4, 3, 3)
    assert np.abs(u @ F.dagger(u) - np.eye(3)).max() < (1e-13 if prec == "D" else 2e-6)
    assert np.abs(np.linalg.det(u) - 1).max() < (1e-13 if prec == "D" else 2e-6)
    G.free()
    Fm.free()


def test_update_keeps_two_row_copy(qlib):
    """The force kernel reads the two-row copy: after updates it must agree
    with a field created (all rows) from the updated links."""
    s, lat, U, mom, G, Fm = make("D", 31)
    assert s.lib.QPHIX_D3_gauge_links_recon(G.h) == 12
    for _ in range(3):
        A.update_gauge(G, Fm, 0.2)
    Uu = G.extract()
    qlib.QPHIX_MI355X_set_gauge_recon(18)
    try:
        G18 = A.Gauge(s, Uu)
    finally:
        qlib.QPHIX_MI355X_set_gauge_recon(0)
    assert s.lib.QPHIX_D3_gauge_links_recon(G18.h) == 18
    z = np.zeros(72 * lat.V)
    f1, f2 = A.Force(s, z), A.Force(s, z)
    A.gauge_force(G, f1, COEF, 0.1)
    A.gauge_force(G18, f2, COEF, 0.1)
    a, b = f1.extract(), f2.extract()
    assert np.abs(a - b).max() < 1e-13 * np.abs(b).max()
    for o in (G, G18, Fm, f1, f2):
        o.free()


@pytest.mark.parametrize("scheme", [0, 1])
def test_gauge_md_matches_oracle(scheme):
    s, lat, U, mom, G, Fm = make("D", 41)
    print(f"[case] gauge_md scheme={scheme}", flush=True)
    info = A.gauge_md(G, Fm, COEF, 0.05, 4, scheme)
    assert info.status == 0
    Uo, mo = O.gauge_md(lat.dims, U, mom, COEF, 0.05, 4, scheme)
    assert np.abs(G.extract() - Uo).max() < 1e-11
    assert np.abs(Fm.extract() - mo).max() < 1e-11
    G.free()
    Fm.free()


@pytest.mark.parametrize("scheme", [0, 1])
def test_gauge_md_reversible_and_second_order(scheme):
    s, lat, U, mom, G, Fm = make("D", 51)
    A.gauge_md(G, Fm, COEF, 0.05, 6, scheme)
    U1, m1 = G.extract(), Fm.extract()
    Fm.free()
    Fm = A.Force(s, -m1)
    A.gauge_md(G, Fm, COEF, 0.05, 6, scheme)
    assert np.abs(G.extract() - U).max() < 1e-12
    assert np.abs(-Fm.extract() - mom).max() < 1e-12
    G.free()
    Fm.free()

    def ham(Ur, mr):
        return 0.5 * float(np.sum(np.abs(mr.view(np.complex128)) ** 2)) + O.gauge_action(lat.dims, Ur, COEF)

    H0 = ham(U, mom)
    dh = []
    for eps, n in ((0.025, 20), (0.0125, 40)):
        G = A.Gauge(s, U)
        Fm = A.Force(s, mom)
        A.gauge_md(G, Fm, COEF, eps, n, scheme)
        dh.append(abs(ham(G.extract(), Fm.extract()) - H0))
        G.free()
        Fm.free()
    assert 3.5 < dh[0] / dh[1] < 4.5, dh
