"""Out-of-device-memory paths.  With abort-on-error off, an allocation that
fails must leave a recorded error and an early return (null handle, or
info->status = QPHIX_MEM_ERROR) with nothing launched on the missing buffer.

The failures are injected by QPHIX_MI355X_debug_fail_alloc(k): the (k+1)-th
device allocation from then on fails as hipMalloc would.  For each entry point
every allocation it makes is failed in turn; the call after the hook is reset
must work normally (checked against the oracle).
"""
import numpy as np
import pytest

import oracle as O
from conftest import lattice_session
from helpers import half, random_links, rel_l2
from pyqphix import api as A
from pyqphix import fields as F

pytestmark = pytest.mark.gpu
DIMS = (4, 4, 8, 8)  # a lattice no other test uses: fresh solver workspaces


@pytest.fixture
def no_abort(qlib):
    qlib.QPHIX_MI355X_set_abort_on_error(0)
    qlib.QPHIX_MI355X_clear_error()
    try:
        yield qlib
    finally:
        qlib.QPHIX_MI355X_debug_fail_alloc(-1)
        qlib.QPHIX_MI355X_clear_error()
        qlib.QPHIX_MI355X_set_abort_on_error(1)


def injected(qlib, fn, nmax=16):
    """Calls fn() with the k-th allocation failing, k = 0, 1, ... until a call
    makes no failing allocation.  fn returns True when it reported failure.
    Returns the number of allocations the call makes."""
    for k in range(nmax):
        qlib.QPHIX_MI355X_clear_error()
        qlib.QPHIX_MI355X_debug_fail_alloc(k)
        print(f"[case] fail allocation {k}", flush=True)
        failed = fn()
        qlib.QPHIX_MI355X_debug_fail_alloc(-1)
        n, msg = qlib.last_error()
        if n == 0:
            assert not failed
            return k
        assert failed and "injected" in msg, (k, msg)
    raise AssertionError("the call never completed without a failing allocation")


def test_field_creation_out_of_memory(no_abort):
    qlib = no_abort
    s = lattice_session(DIMS)
    lat = F.Lattice(DIMS)
    raw = [F.to_raw(a) for a in random_links(lat, "haar", 5)]
    psi = F.to_raw(F.gaussian_source(lat, 6))
    for prec, dt in (("D", np.float64), ("F", np.float32)):
        keep = []

        def mk(cls, *args, **kw):
            try:
                keep.append(cls(s, *args, prec=prec, **kw))
                return False
            except RuntimeError:
                return True

        assert injected(qlib, lambda: mk(A.Vector, psi.astype(dt), A.EVENODD)) == 3
        assert injected(qlib, lambda: mk(A.Links, *[r.astype(dt) for r in raw])) == 8
        U = F.to_raw(F.gauge_field(lat, "weak", 7))
        # stage + 2 parities; the optional SU(3) scratch and two-row copy fall
        # back silently when they cannot be allocated
        assert injected(qlib, lambda: mk(A.Gauge, U.astype(dt))) == 3
        mom = np.zeros(72 * lat.V)
        assert injected(qlib, lambda: mk(A.Force, mom.astype(dt))) == 3
        # the handles made by the last (successful) calls work
        L, vin = keep[-3], keep[0]
        vout = A.Vector(s, None, A.EVENODD, prec)
        A.dslash(L, vout, vin, A.EVEN)
        orc = O.Oracle(DIMS, *raw, dtype=dt)
        ref = orc.dslash(half(psi, lat, 1).astype(dt), 0)
        assert rel_l2(half(vout.extract(), lat, 0), ref) < (1e-13 if prec == "D" else 1e-6)
        for o in keep + [vout]:
            o.free()


def test_solvers_out_of_memory(no_abort):
    qlib = no_abort
    s = lattice_session(DIMS)
    lat = F.Lattice(DIMS)
    raw = [F.to_raw(a) for a in random_links(lat, "haar", 8)]
    b = F.to_raw(F.gaussian_source(lat, 9, parity=0))
    L = A.Links(s, *raw)
    vb = A.Vector(s, b, A.EVEN)
    # outputs exist before the hook is armed: only the solvers' own
    # allocations (workspaces, the sloppy link copies) are failed
    vx = A.Vector(s, None, A.EVEN)
    outs = [A.Vector(s, None, A.EVEN) for _ in range(3)]

    def cg():
        ret, info, res = A.invert(L, vx, vb, 0.1, A.EVEN, 1e-16)
        return info.status == 2  # QPHIX_MEM_ERROR

    def multi():
        ret, info, res = A.invert_multi(L, outs, vb, [0.1, 0.2, 0.3], A.EVEN, 1e-16)
        return info.status == 2

    def mixed():
        ret, info, res = A.invert_mixed(L, vx, vb, 0.1, A.EVEN, 1e-16)
        return info.status == 2

    for fn in (cg, multi, mixed):
        print(f"[case] solver {fn.__name__}", flush=True)
        assert injected(qlib, fn) >= 1
    # the last calls completed: check the mixed solution and a shifted one
    orc = O.Oracle(DIMS, *raw)
    xo, io = orc.invert(b, 0.1, O.EVEN, 1e-18)
    assert rel_l2(half(vx.extract(), lat, 0), half(xo, lat, 0)) < 1e-7
    xo, io = orc.invert(b, 0.3, O.EVEN, 1e-18)
    assert rel_l2(half(outs[2].extract(), lat, 0), half(xo, lat, 0)) < 1e-7
    for o in outs + [vx, vb, L]:
        o.free()
