"""The relative-residue stopping rule on the GPU (res_arg->relresid > 0).

Reference: ks_d_congrad_fn.cpp:404-435 (restart check), 541-560 (inner loop)
and relative_residue, ks_blas_utils_c.cpp:110-149.  The library evaluates the
per-site |r|^2/|x|^2 sum in the Dslash restart epilogue and in the x/r update
kernel; these tests compare final_rel, size_relr, iteration counts (within 5%)
and solutions with the oracle, whose relative_residue is pinned by closed
forms in tests/test_oracle.py.
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
CFG1 = dict(dims=(8, 8, 8, 8), kind="haar", seed=1001, mass=0.025)


def setup(prec, dims=CFG1["dims"], kind=CFG1["kind"], seed=CFG1["seed"]):
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    links = random_links(lat, kind, seed)
    raw = [F.to_raw(a) for a in links]
    L = A.Links(s, *[r.astype(NP[prec]) for r in raw], prec=prec)
    orc = O.Oracle(lat.dims, *raw, dtype=NP[prec])
    return s, lat, raw, L, orc


def true_rel(orc64, lat, x_full, b_full, mass, cb):
    """The relative residue of the FP64 true residual b - (4m^2 - D D) x."""
    xh = half(x_full, lat, cb).astype(np.float64)
    bh = half(b_full, lat, cb).astype(np.float64)
    r = bh - (4 * mass * mass * xh - orc64.dslash(orc64.dslash(xh, 1 - cb), cb))
    return orc64.relative_residue(r, xh)


# (resid, relresid): relative stop alone; both, the relative one binding;
# both, the absolute one binding
CASES = {"D": [(0.0, 1e-7), (1e-10, 1e-7), (1e-16, 1e-4)],
         "F": [(0.0, 1e-5), (1e-6, 1e-5), (1e-7, 1e-3)]}


@pytest.mark.parametrize("prec", ["D", "F"])
@pytest.mark.parametrize("parity", [A.EVEN, A.ODD])
@pytest.mark.parametrize("case", [0, 1, 2])
def test_cg_relresid_matches_oracle(prec, parity, case):
    resid, relresid = CASES[prec][case]
    s, lat, raw, L, orc = setup(prec)
    cb = 0 if parity == A.EVEN else 1
    b = F.to_raw(F.gaussian_source(lat, 77 + case, parity=cb), NP[prec])
    vb = A.Vector(s, b, parity, prec)
    vx = A.Vector(s, None, parity, prec)
    print(f"[case] relresid prec={prec} parity={parity} resid={resid} relresid={relresid}", flush=True)
    ret, info, res = A.invert(L, vx, vb, CFG1["mass"], parity, resid, relresid=relresid)
    x = vx.extract()
    xo, io = orc.invert(b, CFG1["mass"], parity, resid, relresid=relresid)
    assert info.status == 0
    assert abs(res.final_iter - io["final_iter"]) <= max(2, 0.05 * io["final_iter"])
    assert res.final_restart == io["final_restart"] or prec == "F"
    # the relative residue met where the oracle met it
    if prec == "D":
        assert res.final_rel < relresid and io["final_rel"] < relresid
        assert res.size_relr < relresid
        assert abs(np.log10(res.final_rel / io["final_rel"])) < 0.3
        assert abs(np.log10(res.size_relr / io["size_relr"])) < 0.3
        assert rel_l2(half(x, lat, cb), half(xo, lat, cb)) < 1e-7
        # final_rel is the relative residue of the FP64 true residual
        orc64 = O.Oracle(lat.dims, *raw)
        tr = true_rel(orc64, lat, x, b, CFG1["mass"], cb)
        assert abs(res.final_rel - tr) < 1e-6 * tr
        if resid > 0:
            assert res.final_rsq < resid
    else:
        assert res.final_rel > 0 and io["final_rel"] > 0
        assert abs(np.log10(res.final_rel / io["final_rel"])) < 0.5
        assert rel_l2(half(x, lat, cb), half(xo, lat, cb)) < 2e-3
    for o in (vb, vx, L):
        o.free()


@pytest.mark.parametrize("prec", ["D", "F"])
def test_cg_relresid_restarts_match_oracle(prec):
    """relresid alone with a restart every 60 iterations, capped by nrestart."""
    s, lat, raw, L, orc = setup(prec)
    b = F.to_raw(F.gaussian_source(lat, 81, parity=0), NP[prec])
    vb = A.Vector(s, b, A.EVEN, prec)
    vx = A.Vector(s, None, A.EVEN, prec)
    ret, info, res = A.invert(L, vx, vb, CFG1["mass"], A.EVEN, 0.0, relresid=1e-6, niter=60, nrestart=20)
    xo, io = orc.invert(b, CFG1["mass"], O.EVEN, 0.0, relresid=1e-6, niter=60, nrestart=20)
    assert res.final_restart == io["final_restart"] == 20
    assert res.final_iter == io["final_iter"]
    assert abs(np.log10(res.final_rel / io["final_rel"])) < 0.05
    assert rel_l2(half(vx.extract(), lat, 0), half(xo, lat, 0)) < (1e-9 if prec == "D" else 1e-4)
    for o in (vb, vx, L):
        o.free()


@pytest.mark.parametrize("prec", ["D", "F"])
@pytest.mark.parametrize("parity", [A.EVEN, A.ODD])
def test_cg_relresid_plane_wave_closed_form(prec, parity):
    """Free field, x0 = chi, b = (lambda + kappa) chi: the first true residual is
    kappa x0, so final_rel = |kappa| exactly (the closed form of the oracle
    test, here through the device reduction)."""
    dims = (8, 4, 4, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    c1, c3 = 0.9, -0.07
    fat, lng = F.free_links(lat, c1, c3)
    raw = F.raw_link_set(lat, fat, lng, NP[prec])
    L = A.Links(s, *raw, prec=prec)
    cb = 0 if parity == A.EVEN else 1
    p = np.array([2 * np.pi * 1 / 8, 2 * np.pi * 1 / 4, 0.0, 3 * np.pi / 8])
    m = 0.1
    lam = F.free_eigenvalue(p, c1, c3, 4 * m * m)
    chi = F.to_raw(F.plane_wave(lat, p, 1, parity=cb))
    tol = 1e-12 if prec == "D" else 2e-6
    for kappa in (0.03, -0.2):
        vb = A.Vector(s, (lam + kappa) * chi, parity, prec)
        vx = A.Vector(s, chi, parity, prec)
        ret, info, res = A.invert(L, vx, vb, m, parity, 0.0, relresid=0.5)
        assert res.final_iter == 1 and ret == 1
        # FP32: r = b - M x0 carries the rounding of b and M x0 (~1e-7 lambda)
        assert abs(res.final_rel - abs(kappa)) < tol * abs(kappa) + (0 if prec == "D" else 3e-6)
        vx.free()
        vx = A.Vector(s, chi, parity, prec)
        ret, info, res = A.invert(L, vx, vb, m, parity, 0.0, relresid=1e-3)
        assert res.final_iter == 3
        assert res.final_rel < (1e-12 if prec == "D" else 1e-5)
        for o in (vb, vx):
            o.free()
    L.free()


def test_mixed_relresid():
    """Mixed-precision CG with the relative-residue stop alone: it stops at a
    reliable update whose FP64 true residual meets relresid, and final_rel is
    that residual's relative residue."""
    s, lat, raw, L, orc = setup("D")
    orc64 = orc
    b = F.to_raw(F.gaussian_source(lat, 83, parity=0))
    for sloppy in (32, 16):
        vb = A.Vector(s, b, A.EVEN)
        vx = A.Vector(s, None, A.EVEN)
        ret, info, res = A.invert_mixed(L, vx, vb, CFG1["mass"], A.EVEN, 0.0, relresid=1e-7,
                                        sloppy_bits=sloppy)
        x = vx.extract()
        assert info.status == 0
        assert 0 < res.final_rel < 1e-7 and res.size_relr == res.final_rel
        tr = true_rel(orc64, lat, x, b, CFG1["mass"], 0)
        assert abs(res.final_rel - tr) < 1e-6 * tr
        # the FP64 CG with the same stop takes about as many iterations
        _, io = orc.invert(b, CFG1["mass"], O.EVEN, 0.0, relresid=1e-7)
        assert res.final_iter < 1.3 * io["final_iter"]
        for o in (vb, vx):
            o.free()
    L.free()


def test_mixed_reports_failure_at_iteration_cap():
    s, lat, raw, L, orc = setup("D")
    b = F.to_raw(F.gaussian_source(lat, 84, parity=0))
    vb = A.Vector(s, b, A.EVEN)
    vx = A.Vector(s, None, A.EVEN)
    ret, info, res = A.invert_mixed(L, vx, vb, CFG1["mass"], A.EVEN, 1e-20, niter=50, nrestart=2)
    assert info.status == 1  # QPHIX_FAIL
    assert res.final_iter >= 100 and res.final_rsq > 1e-20
    vx.free()
    vx = A.Vector(s, None, A.EVEN)
    ret, info, res = A.invert_mixed(L, vx, vb, CFG1["mass"], A.EVEN, 1e-16)
    assert info.status == 0 and res.final_rsq < 1e-16
    for o in (vb, vx, L):
        o.free()


def test_multi_rhs_rejects_missing_target(qlib):
    """The batched multi-shift has no relative-residue stop (as the reference's
    multi-shift, ks_multicg_offset.cpp:282); a solution without a positive
    resid target is a recorded error, before any work is queued."""
    s, lat, raw, L, orc = setup("D")
    b = F.to_raw(F.gaussian_source(lat, 85, parity=0))
    vb = A.Vector(s, b, A.EVEN)
    outs = [[A.Vector(s, None, A.EVEN) for _ in range(2)]]
    qlib.QPHIX_MI355X_set_abort_on_error(0)
    qlib.QPHIX_MI355X_clear_error()
    try:
        ret, info, res = A.invert_multi_rhs(L, outs, [vb], [0.1, 0.2], A.EVEN, 0.0)
        n, msg = qlib.last_error()
        assert ret == 0 and info.status == 1 and n == 1 and "resid" in msg
    finally:
        qlib.QPHIX_MI355X_clear_error()
        qlib.QPHIX_MI355X_set_abort_on_error(1)
    for o in outs[0] + [vb, L]:
        o.free()
