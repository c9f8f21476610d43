"""QPHIX_*3_asqtad_dslash_faces_noalloc: the face-exchange Dslash with its face
buffer in the library's fixed device scratch instead of allocated memory.  Its
output must equal asqtad_dslash_faces' (bit for bit: the same kernels on the
same inputs) on consecutive calls on one GPU, the later ones with smaller
faces (fewer cut directions, then a smaller lattice), so that a stale face
left in the shared scratch by the previous call would show.  Faces beyond
the scratch are rejected before any launch."""
import numpy as np
import pytest

from conftest import lattice_session
from helpers import random_links
from pyqphix import api as A
from pyqphix import fields as F

pytestmark = pytest.mark.gpu
NP = {"D": np.float64, "F": np.float32}


def _fields(s, dims, prec, seed):
    lat = F.Lattice(dims)
    raw = [F.to_raw(a, NP[prec]) for a in random_links(lat, "haar", seed)]
    L = A.Links(s, *raw, prec=prec)
    vin = A.Vector(s, F.to_raw(F.gaussian_source(lat, seed + 1), NP[prec]), A.EVENODD, prec)
    return L, vin


@pytest.mark.parametrize("prec", ["D", "F"])
def test_noalloc_matches_faces(prec):
    calls = [((16, 16, 16, 16), 15, A.EVEN), ((16, 16, 16, 16), 8, A.ODD), ((8, 8, 8, 8), 15, A.EVEN)]
    for k, (dims, cut, parity) in enumerate(calls):
        print(f"[case] noalloc prec={prec} call={k} dims={dims} cut={cut} parity={parity}", flush=True)
        s = lattice_session(dims)
        L, vin = _fields(s, dims, prec, 70 + 3 * k)
        ref = A.Vector(s, None, A.EVENODD, prec)
        got = A.Vector(s, None, A.EVENODD, prec)
        A.dslash_faces(L, ref, vin, parity, cut)
        assert A.dslash_faces_noalloc(L, got, vin, parity, cut) == 0
        np.testing.assert_array_equal(got.extract(), ref.extract())
        for o in (vin, ref, got, L):
            o.free()


def test_noalloc_rejects_oversized_faces(qlib):
    """64^4 with every direction cut needs 4 x 6 x 64^3/2 x 48 B = 151 MB of FP64
    faces (more than the 64 MiB scratch): rejected with a recorded error, before
    any argument is used (the handles here are null)."""
    dims = (64, 64, 64, 64)
    s = lattice_session(dims)
    qlib.QPHIX_MI355X_set_abort_on_error(0)
    qlib.QPHIX_MI355X_clear_error()
    try:
        st = s.lib.QPHIX_D3_asqtad_dslash_faces_noalloc(None, None, None, A.EVEN, 15)
        n, msg = qlib.last_error()
        assert st != 0 and n == 1 and "scratch" in msg
        qlib.QPHIX_MI355X_clear_error()
        # one direction (t): 6 x 64^3/2 x 48 B = 37.7 MB fits; then the null
        # handles are what is rejected
        st = s.lib.QPHIX_D3_asqtad_dslash_faces_noalloc(None, None, None, A.EVEN, 8)
        n, msg = qlib.last_error()
        assert st != 0 and n == 1 and "null" in msg
    finally:
        qlib.QPHIX_MI355X_clear_error()
        qlib.QPHIX_MI355X_set_abort_on_error(1)
