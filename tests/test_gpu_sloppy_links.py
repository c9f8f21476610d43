"""16-bit fixed-point sloppy links: the Dslash they give against the FP64
oracle (the quantisation bound: per element at most s / 65534 of its link's
largest element s), and mixed-precision CG with them reaching the FP64
system's true residual."""
import numpy as np
import pytest

import oracle as O
from conftest import lattice_session
from helpers import half, random_links, rel_l2
from pyqphix import api as A
from pyqphix import fields as F

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("recon", [0, 18])
def test_sloppy_dslash_against_oracle(recon):
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw = [F.to_raw(a) for a in random_links(lat, "weak", 1013)]
    orc = O.Oracle(dims, *raw)
    s.lib.QPHIX_MI355X_set_long_recon(recon)
    try:
        L = A.Links(s, *raw)
    finally:
        s.lib.QPHIX_MI355X_set_long_recon(0)
    assert s.fn("D", "asqtad_links_recon")(L.h) == (12 if recon == 0 else 18)
    src = F.to_raw(F.gaussian_source(lat, 7), np.float32)
    vin = A.Vector(s, src, A.EVENODD, "F")
    vout = A.Vector(s, None, A.EVENODD, "F")
    for parity, cb in ((A.EVEN, 0), (A.ODD, 1)):
        ref = orc.dslash(half(src.astype(np.float64), lat, 1 - cb), cb)
        err = {}
        for bits in (32, 16):
            print(f"[case] sloppy dslash recon={recon} parity={parity} bits={bits}", flush=True)
            assert A.dslash_sloppy(L, vout, vin, parity, bits) > 0
            err[bits] = rel_l2(half(vout.extract(), lat, cb), ref)
        assert err[32] < 1e-6
        # uniform rounding of 16-bit elements: ~1e-5 relative; the 16-bit path was taken
        assert 1e-7 < err[16] < 6e-5
    for o in (L, vin, vout):
        o.free()


@pytest.mark.parametrize("kind", ["weak", "haar"])
def test_sloppy_dslash_8real_long_links(kind):
    """The FP32 sloppy operator with 8-real long links (U01, U02, U10 and two
    half-angle phase tangents; the rest from unitarity and det = +-1) against
    the FP64 oracle, beside the 12-real form: both within the FP32 Dslash bar,
    and the two outputs differ, so the 8-real kernels ran."""
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw = [F.to_raw(a) for a in random_links(lat, kind, 1017)]
    orc = O.Oracle(dims, *raw)
    L = A.Links(s, *raw)
    assert s.fn("D", "asqtad_links_recon")(L.h) == 12
    src = F.to_raw(F.gaussian_source(lat, 8), np.float32)
    vin = A.Vector(s, src, A.EVENODD, "F")
    vout = A.Vector(s, None, A.EVENODD, "F")
    try:
        for parity, cb in ((A.EVEN, 0), (A.ODD, 1)):
            ref = orc.dslash(half(src.astype(np.float64), lat, 1 - cb), cb)
            out = {}
            for recon in (12, 8):
                print(f"[case] sloppy dslash kind={kind} parity={parity} long recon={recon}", flush=True)
                s.lib.QPHIX_MI355X_set_sloppy_long_recon(recon)
                assert A.dslash_sloppy(L, vout, vin, parity, 32) > 0
                out[recon] = half(vout.extract(), lat, cb)
            e12, e8 = rel_l2(out[12], ref), rel_l2(out[8], ref)
            print(f"[case] rel err 12-real {e12:.2e}, 8-real {e8:.2e}", flush=True)
            assert e12 < 1e-6 and e8 < 1e-6
            assert not np.array_equal(out[8], out[12])
    finally:
        s.lib.QPHIX_MI355X_set_sloppy_long_recon(8)
    for o in (L, vin, vout):
        o.free()


def test_sloppy_8real_falls_back_without_the_form():
    """Unit long links (times phases) store 12 reals but have |U01|^2 +
    |U02|^2 = 0, which the 8-real form cannot represent: the sloppy operator
    keeps the 12-real form, so both settings give the same output, exact
    against the oracle to FP32 rounding."""
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    fat, lng = F.free_links(lat, 0.75, 1.0)
    raw = list(F.raw_link_set(lat, fat, lng))
    orc = O.Oracle(dims, *raw)
    L = A.Links(s, *raw)
    assert s.fn("D", "asqtad_links_recon")(L.h) == 12
    src = F.to_raw(F.gaussian_source(lat, 9), np.float32)
    vin = A.Vector(s, src, A.EVENODD, "F")
    vout = A.Vector(s, None, A.EVENODD, "F")
    ref = orc.dslash(half(src.astype(np.float64), lat, 1), 0)
    out = {}
    try:
        for recon in (8, 12):
            print(f"[case] sloppy dslash unit long links, long recon={recon}", flush=True)
            s.lib.QPHIX_MI355X_set_sloppy_long_recon(recon)
            assert A.dslash_sloppy(L, vout, vin, A.EVEN, 32) > 0
            out[recon] = half(vout.extract(), lat, 0)
    finally:
        s.lib.QPHIX_MI355X_set_sloppy_long_recon(8)
    assert np.array_equal(out[8], out[12])
    assert rel_l2(out[8], ref) < 1e-6
    for o in (L, vin, vout):
        o.free()


def test_sloppy_8real_falls_back_on_non_unitary_rows():
    """Long links (D W) + 1/det W, with W a random unitary 2x2 block and D =
    diag(1.01, 1/1.01), times random signs: row 2 equals s conj(row 0 x row 1)
    for each link and its adjoint, so they store 12 reals, but rows 0 and 1
    are not unit vectors, which the 8-real rebuild assumes (|U00| =
    sqrt(1 - |U01|^2 - |U02|^2) would be wrong, or NaN).  The sloppy operator
    keeps the 12-real form: both settings give the same output, within the
    FP32 bar of the oracle."""
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    fat = random_links(lat, "haar", 1019)[0]
    rng = np.random.default_rng(1021)
    z = rng.standard_normal((lat.V, 4, 2, 2)) + 1j * rng.standard_normal((lat.V, 4, 2, 2))
    w, _ = np.linalg.qr(z)
    lng = np.zeros((lat.V, 4, 3, 3), np.complex128)
    lng[..., :2, :2] = np.array([1.01, 1 / 1.01])[:, None] * w
    lng[..., 2, 2] = 1 / np.linalg.det(w)
    lng *= rng.choice([-1.0, 1.0], size=(lat.V, 4))[..., None, None]
    raw = list(F.raw_link_set(lat, fat, lng))
    orc = O.Oracle(dims, *raw)
    L = A.Links(s, *raw)
    assert s.fn("D", "asqtad_links_recon")(L.h) == 12
    src = F.to_raw(F.gaussian_source(lat, 10), np.float32)
    vin = A.Vector(s, src, A.EVENODD, "F")
    vout = A.Vector(s, None, A.EVENODD, "F")
    ref = orc.dslash(half(src.astype(np.float64), lat, 0), 1)
    out = {}
    try:
        for recon in (8, 12):
            print(f"[case] sloppy dslash non-unitary long links, long recon={recon}", flush=True)
            s.lib.QPHIX_MI355X_set_sloppy_long_recon(recon)
            assert A.dslash_sloppy(L, vout, vin, A.ODD, 32) > 0
            out[recon] = half(vout.extract(), lat, 1)
    finally:
        s.lib.QPHIX_MI355X_set_sloppy_long_recon(8)
    assert np.array_equal(out[8], out[12])
    assert rel_l2(out[8], ref) < 1e-6
    for o in (L, vin, vout):
        o.free()


def true_rsq(orc, lat, x, b, mass, cb):
    xh = half(x, lat, cb)
    bh = half(b, lat, cb)
    r = bh - (4 * mass * mass * xh - orc.dslash(orc.dslash(xh, 1 - cb), cb))
    return float(r @ r / (bh @ bh))


@pytest.mark.parametrize("bits", [32, 16])
@pytest.mark.parametrize("kind,mass", [("haar", 0.025), ("weak", 0.01)])
def test_mixed_with_16bit_links(bits, kind, mass):
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw = [F.to_raw(a) for a in random_links(lat, kind, 1001)]
    orc = O.Oracle(dims, *raw)
    L = A.Links(s, *raw)
    b = F.to_raw(F.gaussian_source(lat, 93, parity=0))
    vb = A.Vector(s, b, A.EVEN)
    vx = A.Vector(s, None, A.EVEN)
    resid = 1e-20
    print(f"[case] mixed 16-bit links bits={bits} kind={kind}", flush=True)
    s.lib.QPHIX_MI355X_set_sloppy_link_bits(16)
    try:
        ret, info, res = A.invert_mixed(L, vx, vb, mass, A.EVEN, resid, sloppy_bits=bits,
                                        niter=3000, nrestart=10)
    finally:
        s.lib.QPHIX_MI355X_set_sloppy_link_bits(32)
    x = vx.extract()
    assert info.status == 0
    assert res.final_rsq < resid
    assert true_rsq(orc, lat, x, b, mass, 0) < 2 * resid
    xo, io = orc.invert(b, mass, O.EVEN, resid)
    assert rel_l2(half(x, lat, 0), half(xo, lat, 0)) < 1e-7
    print(f"[case] iterations {res.final_iter} (FP64 oracle {io['final_iter']}), updates {res.final_restart}",
          flush=True)
    assert res.final_iter < 2.5 * io["final_iter"]
    for o in (L, vb, vx):
        o.free()
