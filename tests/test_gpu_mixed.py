"""Mixed-precision CG (FP32 / FP16-stored sloppy, FP64 reliable updates):
reaches the requested true residual of the FP64 system; its solution matches
the oracle's FP64 CG solution to the accuracy the residual allows."""
import numpy as np
import pytest

import oracle as O
from conftest import lattice_session
from helpers import half, random_links, rel_l2
from pyqphix import api as A
from pyqphix import fields as F

pytestmark = pytest.mark.gpu


def true_rsq(orc, lat, x, b, mass, cb):
    xh = half(x, lat, cb)
    bh = half(b, lat, cb)
    r = bh - (4 * mass * mass * xh - orc.dslash(orc.dslash(xh, 1 - cb), cb))
    return float(r @ r / (bh @ bh))


@pytest.mark.parametrize("bits", [32, 16])
@pytest.mark.parametrize("kind,mass", [("haar", 0.025), ("weak", 0.01)])
def test_mixed_reaches_fp64_residual(bits, kind, mass):
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    links = random_links(lat, kind, 1001)
    raw = [F.to_raw(a) for a in links]
    orc = O.Oracle(dims, *raw)
    L = A.Links(s, *raw)
    b = F.to_raw(F.gaussian_source(lat, 91, parity=0))
    vb = A.Vector(s, b, A.EVEN)
    vx = A.Vector(s, None, A.EVEN)
    print(f"[case] mixed bits={bits} kind={kind}", flush=True)
    resid = 1e-20
    ret, info, res = A.invert_mixed(L, vx, vb, mass, A.EVEN, resid, sloppy_bits=bits,
                                    niter=2000, nrestart=10)
    x = vx.extract()
    assert info.status == 0
    assert res.final_rsq < resid
    assert true_rsq(orc, lat, x, b, mass, 0) < 2 * resid
    xo, io = orc.invert(b, mass, O.EVEN, resid)
    assert rel_l2(half(x, lat, 0), half(xo, lat, 0)) < 1e-7
    assert res.final_restart >= 1
    # sloppy arithmetic costs iterations, but not many more than FP64 CG
    assert res.final_iter < (1.6 if bits == 32 else 2.5) * io["final_iter"]
    for o in (L, vb, vx):
        o.free()


def test_mixed_evenodd_and_guess():
    dims = (8, 4, 4, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw = [F.to_raw(a) for a in random_links(lat, "weak", 1005)]
    orc = O.Oracle(dims, *raw)
    L = A.Links(s, *raw)
    b = F.to_raw(F.gaussian_source(lat, 92))
    vb = A.Vector(s, b, A.EVENODD)
    vx = A.Vector(s, None, A.EVENODD)
    ret, info, res = A.invert_mixed(L, vx, vb, 0.05, A.EVENODD, 1e-22)
    x = vx.extract()
    for cb in (0, 1):
        assert true_rsq(orc, lat, x, b, 0.05, cb) < 2e-22
    ret2, info2, res2 = A.invert_mixed(L, vx, vb, 0.05, A.EVENODD, 1e-20)
    assert res2.final_iter == 2  # the initial guess already solves both halves
    for o in (L, vb, vx):
        o.free()
