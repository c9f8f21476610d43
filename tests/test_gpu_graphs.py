"""hipGraph replay of the solvers' iteration batches: each solver gives the same
iteration counts and bit-identical solutions with replay on (the default) and
off.  One solve replays its graph many times, with host reads of the device
state and true-residual (restart / reliable-update) work between replays."""
import numpy as np
import pytest

from conftest import lattice_session
from helpers import random_links
from pyqphix import api as A
from pyqphix import fields as F

pytestmark = pytest.mark.gpu


def _solve(kind, s, lat, L, L32, b):
    if kind == "multi64":
        xs = [A.Vector(s, None, A.EVEN) for _ in range(3)]
        ret, info, res = A.invert_multi(L, xs, A.Vector(s, b, A.EVEN), [0.02, 0.1, 0.3], A.EVEN,
                                        1e-20, niter=3000)
        out = [x.extract() for x in xs]
        iters = [r.final_iter for r in res]
        for x in xs:
            x.free()
        return ret, iters, out
    if kind == "cg32":
        x = A.Vector(s, None, A.EVENODD, "F")
        ret, info, res = A.invert(L32, x, A.Vector(s, b.astype(np.float32), A.EVENODD, "F"), 0.02,
                                  A.EVENODD, 1e-10, niter=50, nrestart=40)
    elif kind == "cg64":
        x = A.Vector(s, None, A.EVENODD)
        ret, info, res = A.invert(L, x, A.Vector(s, b, A.EVENODD), 0.02, A.EVENODD, 1e-20,
                                  niter=50, nrestart=40)
    else:
        x = A.Vector(s, None, A.EVEN)
        bits = 32 if kind == "mixed32" else 16
        ret, info, res = A.invert_mixed(L, x, A.Vector(s, b, A.EVEN), 0.02, A.EVEN, 1e-20,
                                        sloppy_bits=bits, niter=3000, nrestart=2)
    out = x.extract()
    x.free()
    return ret, [res.final_iter, res.final_restart], [out]


@pytest.mark.parametrize("kind", ["cg64", "cg32", "multi64", "mixed32", "mixed16"])
def test_graph_replay_matches_plain_launches(kind):
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw = [F.to_raw(a) for a in random_links(lat, "weak", 77)]
    L = A.Links(s, *raw)
    L32 = A.Links(s, *[r.astype(np.float32) for r in raw], prec="F")
    b = F.to_raw(F.gaussian_source(lat, 12))
    if kind in ("multi64", "mixed32", "mixed16"):
        b[6 * lat.vh:] = 0
    got = {}
    for graphs in (1, 0):
        s.lib.QPHIX_MI355X_set_graphs(graphs)
        print(f"[case] {kind} graphs={graphs}", flush=True)
        got[graphs] = _solve(kind, s, lat, L, L32, b)
    s.lib.QPHIX_MI355X_set_graphs(1)
    assert got[0][0] == got[1][0]
    assert got[0][1] == got[1][1]
    for a, c in zip(got[0][2], got[1][2]):
        np.testing.assert_array_equal(a, c)
    L.free()
    L32.free()
