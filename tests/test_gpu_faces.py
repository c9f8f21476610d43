"""Domain-decomposed Dslash with self-loopback faces (QPHIX_*3_asqtad_dslash_faces,
the one-GPU part of the reference's halo exchange, ks_boundary.cpp:440-690):
packing the 3-deep faces, computing interior and boundary sites separately and
reading every neighbour across a cut from the packed faces gives exactly the
periodic Dslash (bit-identical), for every cut pattern, both parities, both
precisions and both Naik storages."""
import numpy as np
import pytest

from conftest import lattice_session
from helpers import random_links
from pyqphix import api as A
from pyqphix import fields as F

pytestmark = pytest.mark.gpu
NP = {"D": np.float64, "F": np.float32}


@pytest.mark.parametrize("dims", [(8, 8, 8, 8), (16, 8, 12, 8)])
@pytest.mark.parametrize("prec", ["D", "F"])
@pytest.mark.parametrize("kind", ["haar", "weak"])
def test_faces_bit_identical(dims, prec, kind):
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw = [F.to_raw(a, NP[prec]) for a in random_links(lat, kind, 61)]
    if kind == "weak":  # non-unitary Naik links: 18-real storage
        raw[1] = (1.3 * raw[1]).astype(NP[prec])
        raw[3] = (1.3 * raw[3]).astype(NP[prec])
    L = A.Links(s, *raw, prec=prec)
    assert L.recon == (12 if kind == "haar" else 18)
    psi = F.to_raw(F.gaussian_source(lat, 62), NP[prec])
    vin = A.Vector(s, psi, A.EVENODD, prec)
    ref = A.Vector(s, None, A.EVENODD, prec)
    got = A.Vector(s, None, A.EVENODD, prec)
    for parity in (A.EVEN, A.ODD):
        A.dslash(L, ref, vin, parity)
        r = ref.extract()
        for cut in (1, 2, 4, 8, 15, 9, 6, 0):
            print(f"[case] faces dims={dims} prec={prec} {kind} parity={parity} cut={cut}", flush=True)
            A.dslash_faces(L, got, vin, parity, cut)
            np.testing.assert_array_equal(got.extract(), r)
    for o in (vin, ref, got, L):
        o.free()


def test_faces_large_lattice():
    """32^4, every direction cut: the same result as the periodic kernel."""
    dims = (32, 32, 32, 32)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    fat, lng = F.make_links(lat, "weak", 1002, eps=0.3)
    L = A.Links(s, *F.raw_link_set(lat, fat, lng))
    del fat, lng
    vin = A.Vector(s, F.to_raw(F.gaussian_source(lat, 63)), A.EVENODD)
    ref = A.Vector(s, None, A.EVENODD)
    got = A.Vector(s, None, A.EVENODD)
    A.dslash(L, ref, vin, A.EVEN)
    for cut in (8, 15):
        A.dslash_faces(L, got, vin, A.EVEN, cut)
        np.testing.assert_array_equal(got.extract(), ref.extract())
    for o in (vin, ref, got, L):
        o.free()


def test_faces_reject_short_cut(qlib):
    dims = (8, 4, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    L = A.Links(s, *[F.to_raw(a) for a in random_links(lat, "haar", 64)])
    vin = A.Vector(s, F.to_raw(F.gaussian_source(lat, 65)), A.EVENODD)
    vout = A.Vector(s, None, A.EVENODD)
    qlib.QPHIX_MI355X_set_abort_on_error(0)
    qlib.QPHIX_MI355X_clear_error()
    try:
        assert A.dslash_faces(L, vout, vin, A.EVEN, 2) < 0
        n, msg = qlib.last_error()
        assert n == 1 and "extent" in msg
        assert A.dslash_faces(L, vout, vin, A.EVEN, 13) == 0.0  # y not cut: fine
    finally:
        qlib.QPHIX_MI355X_clear_error()
        qlib.QPHIX_MI355X_set_abort_on_error(1)
    for o in (vin, vout, L):
        o.free()
