"""Single-reduction (FEWSUMS) CG steps on the GPU (ks_d_congrad_fn.cpp:450-541).

QPHIX_MI355X_set_fewsums(mode): bit 0 the one-precision CG, bit 1 the sloppy
steps of the mixed-precision CG.  Each FEWSUMS step runs two Dslashes (the
second reducing p.Ap, Re<Ap|r>, |Ap|^2 and |r|^2 at once) and one fused
x/r/p update.  Compared with the oracle's restatement of the same variant
(pinned by dense solves in tests/test_oracle.py) and with the plain CG:
iteration counts within 5%, the same true residual, solutions to the plain
CG's bars.
"""
import numpy as np
import pytest

import oracle as O
from conftest import lattice_session
from helpers import half, random_links, rel_l2
from pyqphix import api as A
from pyqphix import fields as F

pytestmark = pytest.mark.gpu
NP = {"D": np.float64, "F": np.float32}


@pytest.fixture
def fewsums(qlib):
    def set_mode(m):
        qlib.QPHIX_MI355X_set_fewsums(m)
    yield set_mode
    qlib.QPHIX_MI355X_set_fewsums(2)


def true_rsq(orc64, lat, x, b, mass, cb):
    xh = half(x, lat, cb).astype(np.float64)
    bh = half(b, lat, cb).astype(np.float64)
    r = bh - (4 * mass * mass * xh - orc64.dslash(orc64.dslash(xh, 1 - cb), cb))
    return float(r @ r / (bh @ bh))


@pytest.mark.parametrize("prec", ["D", "F"])
@pytest.mark.parametrize("parity", [A.EVEN, A.ODD])
def test_cg_fewsums_matches_oracle(fewsums, prec, parity):
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw = [F.to_raw(a) for a in random_links(lat, "haar", 1001)]
    L = A.Links(s, *[r.astype(NP[prec]) for r in raw], prec=prec)
    orc = O.Oracle(dims, *raw, dtype=NP[prec])
    orc64 = O.Oracle(dims, *raw)
    cb = 0 if parity == A.EVEN else 1
    mass = 0.025
    resid = 1e-16 if prec == "D" else 1e-10
    b = F.to_raw(F.gaussian_source(lat, 77, parity=cb), NP[prec])
    vb = A.Vector(s, b, parity, prec)
    vx = A.Vector(s, None, parity, prec)
    fewsums(3)
    print(f"[case] fewsums cg prec={prec} parity={parity}", flush=True)
    ret, info, res = A.invert(L, vx, vb, mass, parity, resid)
    x = vx.extract()
    xf, iof = orc.invert(b, mass, parity, resid, fewsums=True)
    xp, iop = orc.invert(b, mass, parity, resid)
    assert info.status == 0
    assert abs(res.final_iter - iof["final_iter"]) <= max(2, 0.05 * iof["final_iter"])
    assert abs(res.final_iter - iop["final_iter"]) <= max(2, 0.05 * iop["final_iter"])
    tr = true_rsq(orc64, lat, x, b, mass, cb)
    if prec == "D":
        assert res.final_rsq < resid and tr < 1.5 * resid
        assert rel_l2(half(x, lat, cb), half(xf, lat, cb)) < 1e-9
        assert rel_l2(half(x, lat, cb), half(xp, lat, cb)) < 1e-9
    else:
        assert tr < 1e-8
        assert rel_l2(half(x, lat, cb), half(xf, lat, cb)) < 2e-4
    for o in (vb, vx, L):
        o.free()


def test_cg_fewsums_restarts_evenodd(fewsums):
    dims = (4, 4, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw = [F.to_raw(a) for a in random_links(lat, "weak", 1003)]
    L = A.Links(s, *raw)
    orc = O.Oracle(dims, *raw)
    b = F.to_raw(F.gaussian_source(lat, 78))
    vb = A.Vector(s, b, A.EVENODD)
    vx = A.Vector(s, None, A.EVENODD)
    fewsums(1)
    ret, info, res = A.invert(L, vx, vb, 0.1, A.EVENODD, 1e-20, niter=40, nrestart=50)
    xo, io = orc.invert(b, 0.1, O.EVENODD, 1e-20, niter=40, nrestart=50, fewsums=True)
    assert res.final_restart == io["final_restart"] and res.final_restart >= 1
    assert abs(res.final_iter - io["final_iter"]) <= max(2, 0.05 * io["final_iter"])
    assert rel_l2(vx.extract(), xo) < 1e-9
    # relresid > 0 takes the two-reduction path (same answer as without FEWSUMS)
    vx2 = A.Vector(s, None, A.EVENODD)
    fewsums(0)
    r0 = A.invert(L, vx2, vb, 0.1, A.EVENODD, 1e-20, relresid=1e-9, niter=40, nrestart=50)
    x0 = vx2.extract()
    vx2.free()
    vx2 = A.Vector(s, None, A.EVENODD)
    fewsums(1)
    r1 = A.invert(L, vx2, vb, 0.1, A.EVENODD, 1e-20, relresid=1e-9, niter=40, nrestart=50)
    assert r0[2].final_iter == r1[2].final_iter
    np.testing.assert_array_equal(vx2.extract(), x0)
    for o in (vb, vx, vx2, L):
        o.free()


@pytest.mark.parametrize("bits", [32, 16])
@pytest.mark.parametrize("lbits", [32, 16])
def test_mixed_fewsums_matches_two_reduction_steps(fewsums, qlib, bits, lbits):
    """The mixed solver with single-reduction sloppy steps reaches the same
    FP64 true residual as with the two-reduction steps, in about as many
    iterations, and its solution agrees with the oracle's FP64 CG."""
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw = [F.to_raw(a) for a in random_links(lat, "weak", 1001)]
    L = A.Links(s, *raw)
    orc = O.Oracle(dims, *raw)
    mass, resid = 0.01, 1e-20
    b = F.to_raw(F.gaussian_source(lat, 93, parity=0))
    vb = A.Vector(s, b, A.EVEN)
    qlib.QPHIX_MI355X_set_sloppy_link_bits(lbits)
    got = {}
    try:
        for mode in (0, 2):
            fewsums(mode)
            vx = A.Vector(s, None, A.EVEN)
            print(f"[case] mixed fewsums={mode} bits={bits} lbits={lbits}", flush=True)
            ret, info, res = A.invert_mixed(L, vx, vb, mass, A.EVEN, resid, sloppy_bits=bits,
                                            niter=2000, nrestart=10)
            assert info.status == 0 and res.final_rsq < resid
            got[mode] = (res.final_iter, vx.extract())
            assert true_rsq(orc, lat, got[mode][1], b, mass, 0) < 2 * resid
            vx.free()
    finally:
        qlib.QPHIX_MI355X_set_sloppy_link_bits(32)
    assert abs(got[2][0] - got[0][0]) <= 0.05 * got[0][0]
    xo, io = orc.invert(b, mass, O.EVEN, resid)
    for mode in (0, 2):
        assert rel_l2(half(got[mode][1], lat, 0), half(xo, lat, 0)) < 1e-7
    for o in (vb, L):
        o.free()


@pytest.mark.parametrize("prec", ["D", "F"])
def test_cg_fewsums_graph_replay(fewsums, prec):
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw = [F.to_raw(a) for a in random_links(lat, "weak", 77)]
    L = A.Links(s, *[r.astype(NP[prec]) for r in raw], prec=prec)
    b = F.to_raw(F.gaussian_source(lat, 12), NP[prec])
    fewsums(1)
    got = {}
    for graphs in (1, 0):
        s.lib.QPHIX_MI355X_set_graphs(graphs)
        vx = A.Vector(s, None, A.EVENODD, prec)
        vb = A.Vector(s, b, A.EVENODD, prec)
        ret, info, res = A.invert(L, vx, vb, 0.02, A.EVENODD, 1e-20 if prec == "D" else 1e-10, niter=50,
                                  nrestart=40)
        got[graphs] = (ret, res.final_iter, res.final_restart, vx.extract())
        vx.free()
        vb.free()
    s.lib.QPHIX_MI355X_set_graphs(1)
    assert got[0][:3] == got[1][:3]
    np.testing.assert_array_equal(got[0][3], got[1][3])
    L.free()
