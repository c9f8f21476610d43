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


@pytest.mark.parametrize("recon", [8, 12])
@pytest.mark.parametrize("bits", [32, 16])
@pytest.mark.parametrize("kind,mass", [("haar", 0.025), ("weak", 0.01)])
def test_mixed_reaches_fp64_residual(bits, kind, mass, recon):
    """recon: long links of the FP32 sloppy operator (8 reals, the default, or 12)."""
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
    print(f"[case] mixed bits={bits} kind={kind} long recon={recon}", flush=True)
    resid = 1e-20
    s.lib.QPHIX_MI355X_set_sloppy_long_recon(recon)
    try:
        ret, info, res = A.invert_mixed(L, vx, vb, mass, A.EVEN, resid, sloppy_bits=bits,
                                        niter=2000, nrestart=10)
    finally:
        s.lib.QPHIX_MI355X_set_sloppy_long_recon(8)
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


def gpu_true_rsq(L, vx, b, mass, lat):
    """|b - (4 m^2 - D_eo D_oe) x|^2 / |b|^2 on the even sites, with the
    library's FP64 Dslash (x and b are even-parity vectors)."""
    s = L.s
    t = A.Vector(s, None, A.EVENODD)
    u = A.Vector(s, None, A.EVENODD)
    A.dslash(L, t, vx, A.ODD)
    A.dslash(L, u, t, A.EVEN)
    n = 6 * lat.vh
    x = vx.extract()[:n]
    r = b[:n] - (4 * mass * mass * x - u.extract()[:n])
    t.free()
    u.free()
    return float(r @ r / (b[:n] @ b[:n]))


def test_mixed_32x4_full_size():
    """At the 32^4 benchmark size (weak links, m = 0.02, 1e-10): FP64 CG, mixed
    CG with FP32 links and vectors, and with 16-bit links and FP16-stored
    vectors all reach the requested true residual (computed with the FP64
    Dslash) and agree with each other to what that residual allows."""
    dims = (32, 32, 32, 32)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    fat, lng = F.make_links(lat, "weak", 1002, eps=0.3)
    L = A.Links(s, *F.raw_link_set(lat, fat, lng))
    del fat, lng
    b = F.to_raw(F.gaussian_source(lat, 5))
    b[6 * lat.vh:] = 0
    vb = A.Vector(s, b, A.EVEN)
    mass, resid = 0.01, 1e-20
    sols = {}
    for name, bits, lbits in (("f64", None, None), ("f32", 32, 32), ("f16_l16", 16, 16)):
        print(f"[case] 32^4 {name}", flush=True)
        vx = A.Vector(s, None, A.EVEN)
        if bits is None:
            ret, info, res = A.invert(L, vx, vb, mass, A.EVEN, resid, niter=5000, nrestart=10)
        else:
            s.lib.QPHIX_MI355X_set_sloppy_link_bits(lbits)
            try:
                ret, info, res = A.invert_mixed(L, vx, vb, mass, A.EVEN, resid, sloppy_bits=bits,
                                                niter=5000, nrestart=10)
            finally:
                s.lib.QPHIX_MI355X_set_sloppy_link_bits(32)
        assert info.status == 0
        assert gpu_true_rsq(L, vx, b, mass, lat) < 2 * resid
        sols[name] = (vx.extract()[:6 * lat.vh], res.final_iter)
        vx.free()
    x64, it64 = sols["f64"]
    for name in ("f32", "f16_l16"):
        x, it = sols[name]
        assert np.linalg.norm(x - x64) < 1e-8 * np.linalg.norm(x64)
        assert it < 1.1 * it64
    L.free()
    vb.free()
