"""Multi-right-hand-side Dslash and batched mixed-precision multi-shift CG,
against the oracle."""
import numpy as np
import pytest

import oracle as O
from conftest import lattice_session
from helpers import half, random_links, rel_l2
from pyqphix import api as A
from pyqphix import fields as F

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[None, "1", "2", "4"], ids=["tunedK", "K1", "K2", "K4"])
def sources_per_item(request, monkeypatch):
    """Runs a test with the autotuned sources-per-work-item, then each K forced."""
    if request.param is None:
        monkeypatch.delenv("QPHIX_MRHS_K", raising=False)
    else:
        monkeypatch.setenv("QPHIX_MRHS_K", request.param)
    return request.param


@pytest.mark.parametrize("n,dims", [(1, (8, 8, 8, 8)), (2, (8, 8, 8, 8)), (3, (8, 8, 8, 8)), (5, (8, 8, 8, 8)),
                                    (8, (8, 8, 8, 8)), (5, (4, 8, 4, 12)), (8, (4, 4, 4, 12))])
def test_dslash_multi_matches_oracle(n, dims, sources_per_item):
    """The multi-RHS Dslash against the oracle, per source; 4x8x4x12 and
    4x4x4x12 (768 and 96 sites per parity: the last logical block partly
    empty, nz not a multiple of 8, so the XCDs walk t-slabs instead of
    z-regions)."""
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw = [F.to_raw(a) for a in random_links(lat, "weak", 31)]
    orc = O.Oracle(dims, *raw)
    L = A.Links(s, *[r.astype(np.float32) for r in raw], prec="F")
    srcs = [F.to_raw(F.gaussian_source(lat, 100 + k), np.float32) for k in range(n)]
    vin = [A.Vector(s, x, A.EVENODD, "F") for x in srcs]
    vout = [A.Vector(s, None, A.EVENODD, "F") for _ in srcs]
    for parity, cb in ((A.EVEN, 0), (A.ODD, 1)):
        A.dslash_multi(L, vout, vin, parity)
        for k in range(n):
            ref = orc.dslash(half(srcs[k].astype(np.float64), lat, 1 - cb), cb)
            assert rel_l2(half(vout[k].extract(), lat, cb), ref) < 1e-6
    for o in vin + vout + [L]:
        o.free()


def true_rsq(orc, lat, x, b, mass, cb):
    xh = half(x, lat, cb)
    bh = half(b, lat, cb)
    r = bh - (4 * mass * mass * xh - orc.dslash(orc.dslash(xh, 1 - cb), cb))
    return float(r @ r / (bh @ bh))


@pytest.mark.parametrize("nrhs,block,dims", [(3, 0, (8, 8, 8, 8)), (4, 4, (8, 8, 8, 8)), (8, 8, (8, 8, 8, 8)),
                                             (2, 1, (8, 8, 8, 8)), (3, 0, (4, 8, 4, 12)), (8, 8, (4, 4, 4, 12))])
def test_multi_rhs_multishift(nrhs, block, dims, sources_per_item):
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw = [F.to_raw(a) for a in random_links(lat, "weak", 1003)]
    orc = O.Oracle(dims, *raw)
    L = A.Links(s, *raw)
    masses = [0.05, 0.02, 0.1, 0.3]       # API masses: shifts 4 m^2
    resid = 1e-20
    bs = [F.to_raw(F.gaussian_source(lat, 200 + r, parity=0)) for r in range(nrhs)]
    vb = [A.Vector(s, b, A.EVEN) for b in bs]
    vx = [[A.Vector(s, None, A.EVEN) for _ in masses] for _ in bs]
    print(f"[case] multi_rhs nrhs={nrhs} block={block} dims={dims} K={sources_per_item}", flush=True)
    ret, info, res = A.invert_multi_rhs(L, vx, vb, masses, A.EVEN, resid, niter=1000,
                                        nrestart=2, rhs_block=block)
    print(f"[case] ret={ret} iters={[[res[r][j].final_iter for j in range(len(masses))] for r in range(nrhs)]}",
          flush=True)
    assert info.status == 0 and ret > 0
    for r in range(nrhs):
        xo, io = orc.invert_multi(bs[r], np.array(masses), O.EVEN, resid)
        # the batched base system converges like the oracle's (within 10%)
        assert ret <= 1.1 * io["ret"] + 5
        for j, m in enumerate(masses):
            x = vx[r][j].extract()
            assert res[r][j].final_rsq < resid
            assert true_rsq(orc, lat, x, bs[r], m, 0) < 2 * resid
            # two solutions at a 1e-10 relative residual differ by up to cond(M) * 1e-10
            assert rel_l2(half(x, lat, 0), half(xo[j], lat, 0)) < 1e-5
    for o in [L] + vb + [v for row in vx for v in row]:
        o.free()


def test_multi_rhs_graph_replay_matches_plain_launches():
    """The batched solver's hipGraph replays (many per solve, with host-side
    reliable updates between them) give the same results as plain launches of
    the same kernels, bit for bit."""
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw = [F.to_raw(a) for a in random_links(lat, "weak", 1007)]
    L = A.Links(s, *raw)
    masses = [0.03, 0.1, 0.3]
    bs = [F.to_raw(F.gaussian_source(lat, 300 + r, parity=0)) for r in range(4)]
    vb = [A.Vector(s, b, A.EVEN) for b in bs]
    outs = {}
    for graphs in (1, 0):
        s.lib.QPHIX_MI355X_set_graphs(graphs)
        vx = [[A.Vector(s, None, A.EVEN) for _ in masses] for _ in bs]
        print(f"[case] multi_rhs graphs={graphs}", flush=True)
        ret, info, res = A.invert_multi_rhs(L, vx, vb, masses, A.EVEN, 1e-18, niter=1000,
                                            nrestart=2, rhs_block=4)
        assert info.status == 0
        outs[graphs] = (ret, [[x.extract() for x in row] for row in vx],
                        [[(r.final_iter, r.final_rsq) for r in row] for row in res])
        for row in vx:
            for x in row:
                x.free()
    s.lib.QPHIX_MI355X_set_graphs(1)
    assert outs[0][0] == outs[1][0]
    assert outs[0][2] == outs[1][2]
    for r0, r1 in zip(outs[0][1], outs[1][1]):
        for a, b in zip(r0, r1):
            np.testing.assert_array_equal(a, b)
    for o in [L] + vb:
        o.free()


@pytest.mark.parametrize("k", [1, 2, 4])
def test_lds_staged_links_match_per_lane_loads(k, monkeypatch):
    """The multi-RHS Dslash with links staged through LDS and the one that
    loads them per work-item give bit-identical results: plain Dslash on 8 and
    4 FP32 vectors, and a batched multi-shift solve (FP64 and FP32 phases, the
    p.Ap epilogue, refinement).  k: sources per work-item (2, 4: the LDS
    kernel that applies each staged link to k sources; the FP64 phase uses at
    most 2)."""
    monkeypatch.setenv("QPHIX_MRHS_K", str(k))
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw = [F.to_raw(a) for a in random_links(lat, "weak", 1011)]
    L32 = A.Links(s, *[r.astype(np.float32) for r in raw], prec="F")
    L = A.Links(s, *raw)
    srcs = [F.to_raw(F.gaussian_source(lat, 500 + k), np.float32) for k in range(8)]
    vin = [A.Vector(s, x, A.EVENODD, "F") for x in srcs]
    masses = [0.03, 0.2]
    bs = [F.to_raw(F.gaussian_source(lat, 520 + r, parity=0)) for r in range(8)]
    vb = [A.Vector(s, b, A.EVEN) for b in bs]
    got = {}
    try:
        for on in (1, 0):
            s.lib.QPHIX_MI355X_set_mrhs_lds(on)
            print(f"[case] mrhs lds={on} k={k}", flush=True)
            outs = []
            for n in (8, 4):
                vout = [A.Vector(s, None, A.EVENODD, "F") for _ in range(n)]
                A.dslash_multi(L32, vout, vin[:n], A.EVEN)
                outs.append([v.extract() for v in vout])
                for v in vout:
                    v.free()
            vx = [[A.Vector(s, None, A.EVEN) for _ in masses] for _ in bs]
            ret, info, res = A.invert_multi_rhs(L, vx, vb, masses, A.EVEN, 1e-18, niter=1000,
                                                nrestart=2, rhs_block=8)
            assert info.status == 0
            got[on] = (outs, ret, [[x.extract() for x in row] for row in vx],
                       [[(r.final_iter, r.final_rsq) for r in row] for row in res])
            for row in vx:
                for x in row:
                    x.free()
    finally:
        s.lib.QPHIX_MI355X_set_mrhs_lds(1)
    for a, b in zip(got[1][0], got[0][0]):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    assert got[1][1] == got[0][1] and got[1][3] == got[0][3]
    for r1, r0 in zip(got[1][2], got[0][2]):
        for x, y in zip(r1, r0):
            np.testing.assert_array_equal(x, y)
    for o in [L, L32] + vin + vb:
        o.free()


@pytest.mark.parametrize("block,steps", [(8, 8), (4, 8), (8, 6), (8, 4), (4, 4), (8, 3), (8, 2)])
def test_deferred_shifted_updates_match_single_steps(block, steps):
    """The batched solver's shifted systems updated `steps` steps at a time
    (one read and write of x_j, p_j per group of steps) give the same solutions,
    bit for bit, as one step at a time: through shift retirements,
    reliable-update pauses, the FP64-to-FP32 switch and the per-shift
    refinement."""
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw = [F.to_raw(a) for a in random_links(lat, "weak", 1015)]
    L = A.Links(s, *raw)
    masses = [0.02, 0.05, 0.12, 0.3, 0.6]
    bs = [F.to_raw(F.gaussian_source(lat, 540 + r, parity=0)) for r in range(6)]
    bs[3][:] = 0.0
    vb = [A.Vector(s, b, A.EVEN) for b in bs]
    got = {}
    try:
        for fuse in (steps, 1):
            s.lib.QPHIX_MI355X_set_msx_fuse(fuse)
            print(f"[case] shifted updates every {fuse} steps, block={block}", flush=True)
            vx = [[A.Vector(s, None, A.EVEN) for _ in masses] for _ in bs]
            ret, info, res = A.invert_multi_rhs(L, vx, vb, masses, A.EVEN, 1e-18, niter=1000,
                                                nrestart=2, rhs_block=block)
            assert info.status == 0
            got[fuse] = (ret, [[x.extract() for x in row] for row in vx],
                         [[(r.final_iter, r.final_rsq) for r in row] for row in res])
            for row in vx:
                for x in row:
                    x.free()
    finally:
        s.lib.QPHIX_MI355X_set_msx_fuse(8)  # the default
    assert got[steps][0] == got[1][0]
    assert got[steps][2] == got[1][2]
    for r1, r0 in zip(got[steps][1], got[1][1]):
        for x, y in zip(r1, r0):
            np.testing.assert_array_equal(x, y)
    for o in [L] + vb:
        o.free()


def test_multi_rhs_ragged_and_zero_source():
    """5 sources in a block of 8 (three empty slots), one of them zero: the
    zero source returns zero solutions, the others meet their targets."""
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw = [F.to_raw(a) for a in random_links(lat, "weak", 1009)]
    orc = O.Oracle(dims, *raw)
    L = A.Links(s, *raw)
    masses = [0.04, 0.15]
    resid = 1e-18
    bs = [F.to_raw(F.gaussian_source(lat, 400 + r, parity=0)) for r in range(5)]
    bs[2][:] = 0.0
    vb = [A.Vector(s, b, A.EVEN) for b in bs]
    vx = [[A.Vector(s, None, A.EVEN) for _ in masses] for _ in bs]
    print("[case] multi_rhs ragged 5 in 8, zero source", flush=True)
    ret, info, res = A.invert_multi_rhs(L, vx, vb, masses, A.EVEN, resid, niter=1000, nrestart=2,
                                        rhs_block=8)
    assert info.status == 0
    for r in range(5):
        for j, m in enumerate(masses):
            x = vx[r][j].extract()
            if r == 2:
                assert not x.any() and res[r][j].final_rsq == 0.0
            else:
                assert true_rsq(orc, lat, x, bs[r], m, 0) < 2 * resid
    for o in [L] + vb + [v for row in vx for v in row]:
        o.free()


def test_dslash_multi_equals_single_32x32x32x64():
    """At the batched workload's full size the 8-vector Dslash equals eight
    single-vector FP32 Dslashes (same operator, same per-pair order)."""
    dims = (32, 32, 32, 64)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    rng = np.random.default_rng(5)
    fat = F.near_identity(rng, lat.V * 4, 0.3).reshape(lat.V, 4, 3, 3)
    lng = F.near_identity(rng, lat.V * 4, 0.3).reshape(lat.V, 4, 3, 3)   # 18-real Naik storage
    raw = [r.astype(np.float32) for r in F.raw_link_set(lat, F.fold_phases(lat, fat, 1),
                                                        F.fold_phases(lat, lng, 3))]
    del fat, lng
    L = A.Links(s, *raw, prec="F")
    del raw
    vin = [A.Vector(s, (rng.standard_normal(6 * lat.V)).astype(np.float32), A.EVENODD, "F")
           for _ in range(8)]
    vm = [A.Vector(s, None, A.EVENODD, "F") for _ in range(8)]
    v1 = A.Vector(s, None, A.EVENODD, "F")
    print("[case] dslash_multi vs single at 32^3x64", flush=True)
    A.dslash_multi(L, vm, vin, A.ODD)
    for k in range(8):
        A.dslash(L, v1, vin[k], A.ODD)
        a = vm[k].extract()[6 * lat.vh:]
        b = v1.extract()[6 * lat.vh:]
        assert rel_l2(a, b) < 1e-6
    for o in [L, v1] + vin + vm:
        o.free()


def test_batched_multishift_32x32x32x64_full_size():
    """At the 32^3x64 benchmark size (weak links, 12 shifts in [1e-4, 1e-1],
    8 sources, 1e-10 per shift): every one of the 96 solutions reaches its true
    residual, computed with the FP64 Dslash, and source 0's solutions agree with
    the FP64 single-source multi-shift solver's."""
    dims = (32, 32, 32, 64)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    fat, lng = F.make_links(lat, "weak", 1003, eps=0.3)
    L = A.Links(s, *F.raw_link_set(lat, fat, lng))
    del fat, lng
    sig = np.logspace(-4, -1, 12)
    masses = list(np.sqrt(sig) / 2)          # shifts 4 m^2 = sigma
    resid = 1e-20
    n = 6 * lat.vh
    bs = [F.to_raw(F.gaussian_source(lat, 600 + r, parity=0)) for r in range(8)]
    vb = [A.Vector(s, b, A.EVEN) for b in bs]
    vx = [[A.Vector(s, None, A.EVEN) for _ in masses] for _ in bs]
    print("[case] batched multi-shift 32^3x64, 8 x 12", flush=True)
    ret, info, res = A.invert_multi_rhs(L, vx, vb, masses, A.EVEN, resid, niter=20000, nrestart=2)
    assert info.status == 0
    t = A.Vector(s, None, A.EVENODD)
    u = A.Vector(s, None, A.EVENODD)
    worst = 0.0
    for r in range(8):
        bb = bs[r][:n]
        for j, m in enumerate(masses):
            A.dslash(L, t, vx[r][j], A.ODD)
            A.dslash(L, u, t, A.EVEN)
            x = vx[r][j].extract()[:n]
            rr = bb - (4 * m * m * x - u.extract()[:n])
            worst = max(worst, float(rr @ rr / (bb @ bb)))
    print(f"[case] worst true residual {worst:.3e}", flush=True)
    assert worst < 2 * resid
    xo = [A.Vector(s, None, A.EVEN) for _ in masses]
    ret64, info64, res64 = A.invert_multi(L, xo, vb[0], masses, A.EVEN, resid, niter=20000)
    for j in range(len(masses)):
        a = vx[0][j].extract()[:n]
        c = xo[j].extract()[:n]
        assert np.linalg.norm(a - c) < 1e-8 * np.linalg.norm(c)
    for o in [L, t, u] + vb + xo + [v for row in vx for v in row]:
        o.free()


def test_site_orders_agree(monkeypatch):
    """The multi-RHS kernels' traversal orders (z-regions per XCD, the
    default; t-slabs; (y, z) patches) only permute which lane computes which
    site: the plain multi-RHS Dslash is bit-identical under all three, and the
    batched multi-shift (whose per-block partial sums follow the order) gives
    the same solutions to the solver's accuracy.  On 8 x 8 x 16 x 16 (nz a
    multiple of 8, so the z-region order applies)."""
    dims = (8, 8, 16, 16)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw = [F.to_raw(a) for a in random_links(lat, "weak", 71)]
    L64 = A.Links(s, *raw)
    L32 = A.Links(s, *[r.astype(np.float32) for r in raw], prec="F")
    rng = np.random.default_rng(72)
    vin = [A.Vector(s, rng.standard_normal(6 * lat.V).astype(np.float32), A.EVENODD, "F") for _ in range(8)]
    vout = [A.Vector(s, None, A.EVENODD, "F") for _ in range(8)]
    masses = [0.02, 0.05, 0.1]
    bs = [F.to_raw(F.gaussian_source(lat, 73 + r, parity=0)) for r in range(4)]
    vb = [A.Vector(s, b, A.EVEN) for b in bs]
    got_d, got_x = {}, {}
    for order in ("z", "t", "2,4"):
        monkeypatch.setenv("QPHIX_SITE_ORDER", order)
        print(f"[case] site order {order}", flush=True)
        A.dslash_multi(L32, vout, vin, A.ODD)
        got_d[order] = [v.extract() for v in vout]
        xs = [[A.Vector(s, None, A.EVEN) for _ in masses] for _ in bs]
        ret, info, res = A.invert_multi_rhs(L64, xs, vb, masses, A.EVEN, 1e-20, niter=4000, nrestart=2)
        assert info.status == 0
        got_x[order] = [[x.extract()[:6 * lat.vh] for x in row] for row in xs]
        for row in xs:
            for x in row:
                x.free()
    for order in ("t", "2,4"):
        for a, b in zip(got_d[order], got_d["z"]):
            np.testing.assert_array_equal(a, b)
        for ra, rb in zip(got_x[order], got_x["z"]):
            for a, b in zip(ra, rb):
                assert rel_l2(a, b) < 1e-8
    for o in [L64, L32] + vin + vout + vb:
        o.free()
