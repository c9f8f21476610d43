"""One GPU's share of a production lattice, 48^3x96, at full size.

The workload of the combined multi-shift + force case: weak-field links
(fat 1 + 0.3 Gaussian, Naik exp(i 0.3 H), stored as 12 reals), m = 0.0102,
16 shifts sigma_j log-spaced in [4 m^2, 1], 4 Gaussian sources, 1e-10 true
relative residual per shift (resid = 1e-20).  The oracle would need hours at
this size, so the checks are properties that fix the answer:
  - every one of the 64 batched mixed-precision solutions satisfies
    |b - (sigma_j - D_eo D_oe) x|^2 / |b|^2 < 2e-20 with the FP64 Dslash;
  - source 0's solutions agree with the FP64 single-source multi-shift
    solver (the reference's algorithm, ks_multicg_offset.cpp:205-689) to 1e-8;
  - the fermion-force outer products of source 0's 16 solutions agree with
    those of the FP64 solutions to 1e-8 (the force is linear in each
    solution pair, so its error is bounded by the solutions' errors).
About 2 minutes and ~110 GB of HBM; host memory ~60 GB.
"""
import numpy as np
import pytest

from conftest import lattice_session
from pyqphix import api as A
from pyqphix import fields as F

pytestmark = pytest.mark.gpu

DIMS = (48, 48, 48, 96)
MASS = 0.0102
NSHIFT = 16
NRHS = 4
RESID = 1e-20


@pytest.mark.timeout(900)
def test_combined_48x96_full_size():
    s = lattice_session(DIMS)
    lat = F.Lattice(DIMS)
    print("[case] 48^3x96: generating links", flush=True)
    fat, lng = A.weak_links_on_device(s, lat, 51, 0.3)
    raw = F.raw_link_set(lat, fat, lng)
    del fat, lng
    L = A.Links(s, *raw)
    del raw
    assert L.recon == 12
    sig = np.logspace(np.log10(4 * MASS * MASS), 0.0, NSHIFT)
    masses = list(np.sqrt(sig) / 2)
    n = 6 * lat.vh
    bs = [F.to_raw(F.gaussian_source(lat, 5000 + r, parity=0)) for r in range(NRHS)]
    vb = [A.Vector(s, b, A.EVEN) for b in bs]
    vx = [[A.Vector(s, None, A.EVEN) for _ in masses] for _ in bs]
    print(f"[case] 48^3x96: batched mixed multi-shift {NRHS} x {NSHIFT}", flush=True)
    ret, info, res = A.invert_multi_rhs(L, vx, vb, masses, A.EVEN, RESID, niter=20000, nrestart=2)
    assert info.status == 0, s.lib.last_error()
    assert max(r.final_rsq for row in res for r in row) < RESID

    print("[case] 48^3x96: true residuals with the FP64 Dslash", flush=True)
    t = A.Vector(s, None, A.EVENODD)
    u = A.Vector(s, None, A.EVENODD)
    worst = 0.0
    for r in range(NRHS):
        bb = bs[r][:n]
        b2 = float(bb @ bb)
        for j, m in enumerate(masses):
            A.dslash(L, t, vx[r][j], A.ODD)
            A.dslash(L, u, t, A.EVEN)
            x = vx[r][j].extract()[:n]
            rr = bb - (4 * m * m * x - u.extract()[:n])
            worst = max(worst, float(rr @ rr) / b2)
    print(f"[case] 48^3x96: worst true residual {worst:.3e}", flush=True)
    assert worst < 2 * RESID
    t.free()
    u.free()

    print("[case] 48^3x96: FP64 multi-shift on source 0", flush=True)
    xo = [A.Vector(s, None, A.EVEN) for _ in masses]
    ret64, info64, res64 = A.invert_multi(L, xo, vb[0], masses, A.EVEN, RESID, niter=20000, nrestart=2)
    assert info64.status == 0
    worst_err = 0.0
    for j in range(NSHIFT):
        a = vx[0][j].extract()[:n]
        c = xo[j].extract()[:n]
        worst_err = max(worst_err, float(np.linalg.norm(a - c) / np.linalg.norm(c)))
    print(f"[case] 48^3x96: source 0 vs FP64, worst rel err {worst_err:.2e}", flush=True)
    assert worst_err < 1e-8

    print("[case] 48^3x96: fermion force, mixed vs FP64 solutions", flush=True)
    coef = list(1.0 / (1.0 + np.arange(NSHIFT)))
    o1m, o3m, _ = A.fermion_force_oprod(L, coef, vx[0])
    o1d, o3d, _ = A.fermion_force_oprod(L, coef, xo)
    e1 = float(np.linalg.norm(o1m - o1d) / np.linalg.norm(o1d))
    e3 = float(np.linalg.norm(o3m - o3d) / np.linalg.norm(o3d))
    print(f"[case] 48^3x96: force rel err fat {e1:.2e} long {e3:.2e}", flush=True)
    assert e1 < 1e-8 and e3 < 1e-8
    del o1m, o3m, o1d, o3d
    for o in [L] + vb + xo + [v for row in vx for v in row]:
        o.free()
