"""Fermion-force outer products (QPHIX_*3_asqtad_fermion_force_oprod) against
the oracle, and the torch.autograd harness against central finite differences
and the oracle's CPU FP64 force at 24^4 with 4 shifts.

The finite differences use the exact identity
    S(U+hE) - S(U-hE) = sum_k c_k Re X+_k^dag (D+_eo D+_oe - D-_eo D-_oe) X-_k
(X+-_k solved on the perturbed links), whose middle factor is identically
zero away from the perturbed links, so the difference carries no
cancellation from the lattice-wide sum S; its truncation error is O(h^2)."""
import numpy as np
import pytest
import torch

import oracle as O
from conftest import lattice_session
from helpers import rel_l2
from pyqphix import api as A
from pyqphix import fields as F
from pyqphix.autograd import FermionAction, solve

pytestmark = pytest.mark.gpu


def _solutions(s, lat, L, masses, seeds):
    xs = []
    for seed in seeds:
        b = F.to_raw(F.gaussian_source(lat, seed, parity=0))
        vb = A.Vector(s, b, A.EVEN)
        out = [A.Vector(s, None, A.EVEN) for _ in masses]
        ret, info, res = A.invert_multi(L, out, vb, masses, A.EVEN, 1e-20, niter=2000)
        assert info.status == 0
        xs += [o.extract() for o in out]
        for o in out + [vb]:
            o.free()
    return xs


@pytest.mark.parametrize("prec", ["D", "F"])
def test_oprod_matches_oracle(prec):
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    fat, lng = F.make_links(lat, "weak", 4242, eps=0.3)
    raw = F.raw_link_set(lat, fat, lng)
    L = A.Links(s, *raw)
    xs = _solutions(s, lat, L, [0.05, 0.2, 0.5], (1, 2, 3))   # 9 terms: one full group + 1
    coef = [0.3, -0.7, 0.5, 1.1, 0.2, 0.9, -0.4, 0.6, 1.3]
    if prec == "F":
        xs = [x.astype(np.float32).astype(np.float64) for x in xs]
        Lp = A.Links(s, *[r.astype(np.float32) for r in raw], prec="F")
    else:
        Lp = L
    vx = [A.Vector(s, x, A.EVEN, prec) for x in xs]
    print(f"[case] oprod {prec} 8^4 9 terms", flush=True)
    o1, o3, info = A.fermion_force_oprod(Lp, coef, vx)
    orc = O.Oracle(dims, *raw)
    r1, r3 = orc.fermion_oprod(coef, xs)
    orc.close()
    r1 = r1.view(np.complex128).reshape(o1.shape)
    r3 = r3.view(np.complex128).reshape(o3.shape)
    tol = 1e-12 if prec == "D" else 2e-6
    assert rel_l2(o1, r1) < tol
    assert rel_l2(o3, r3) < tol
    assert info.final_flop > 0
    for o in vx + [L] + ([Lp] if Lp is not L else []):
        o.free()


@pytest.mark.parametrize("nterms", [2, 3, 5])
def test_oprod_term_groups(nterms):
    """Groups of 2..8 terms get their D_oe X_k from one multi-RHS Dslash with
    2, 4 or 8 interleaved sources (3 and 5 terms leave padding sources
    unused): the outer products against the oracle."""
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    fat, lng = F.make_links(lat, "weak", 4243, eps=0.3)
    raw = F.raw_link_set(lat, fat, lng)
    L = A.Links(s, *raw)
    xs = _solutions(s, lat, L, [0.07, 0.3], (4, 5, 6))[:nterms]
    coef = [0.4, -0.9, 0.7, 1.2, -0.3][:nterms]
    vx = [A.Vector(s, x, A.EVEN) for x in xs]
    print(f"[case] oprod D 8^4 {nterms} terms", flush=True)
    o1, o3, info = A.fermion_force_oprod(L, coef, vx)
    orc = O.Oracle(dims, *raw)
    r1, r3 = orc.fermion_oprod(coef, xs)
    orc.close()
    assert rel_l2(o1, r1.view(np.complex128).reshape(o1.shape)) < 1e-12
    assert rel_l2(o3, r3.view(np.complex128).reshape(o3.shape)) < 1e-12
    for o in vx + [L]:
        o.free()


def test_oprod_edge_cases():
    dims = (8, 8, 8, 8)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    raw = F.raw_link_set(lat, *F.make_links(lat, "weak", 5, eps=0.3))
    L = A.Links(s, *raw)
    o1, o3, _ = A.fermion_force_oprod(L, [], [])
    assert not o1.any() and not o3.any()
    xs = _solutions(s, lat, L, [0.1], (9,))
    vx = [A.Vector(s, x, A.EVEN) for x in xs]
    full1, full3, _ = A.fermion_force_oprod(L, [1.0], vx)
    # NULL outputs are skipped; the other is unchanged
    import ctypes as C
    from pyqphix.lib import Info, ptr
    o = np.zeros(72 * lat.V)
    info = Info()
    s.lib.QPHIX_D3_asqtad_fermion_force_oprod(C.byref(info), L.h, 1, (C.c_double * 1)(1.0),
                                              (C.c_void_p * 1)(vx[0].h), None, ptr(o))
    assert info.status == 0
    np.testing.assert_array_equal(o.view(np.complex128).reshape(full3.shape), full3)
    for v in vx + [L]:
        v.free()


def _exact_central_difference(s, lat, fat, lng, phi, masses, coef, Ef, El, h, resid):
    sp = solve(s, fat + h * Ef, lng + h * El, phi, masses, resid)
    sm = solve(s, fat - h * Ef, lng - h * El, phi, masses, resid)
    tmp = A.Vector(s, None, A.EVENODD)
    wp = A.Vector(s, None, A.EVENODD)
    wm = A.Vector(s, None, A.EVENODD)
    vh = lat.vh
    ds = 0.0
    for c, xp, xm in zip(coef, sp.xs, sm.xs):
        A.dslash(sp.L, tmp, xm, A.ODD)
        A.dslash(sp.L, wp, tmp, A.EVEN)
        A.dslash(sm.L, tmp, xm, A.ODD)
        A.dslash(sm.L, wm, tmp, A.EVEN)
        w = (wp.extract()[:6 * vh] - wm.extract()[:6 * vh]).view(np.complex128)
        ds += c * float(np.real(np.vdot(xp.extract()[:6 * vh].view(np.complex128), w)))
    for o in (tmp, wp, wm):
        o.free()
    sp.free()
    sm.free()
    return ds / (2 * h)


def test_autograd_force_24_four_shifts():
    """24^4, FP64, 4 shifts: torch.autograd through the adjoint kernel against
    the oracle's CPU force and central differences (h = 1e-5) on 64 random links."""
    dims = (24, 24, 24, 24)
    s = lattice_session(dims)
    lat = F.Lattice(dims)
    fat, lng = F.make_links(lat, "weak", 99, eps=0.3)
    phi = F.to_raw(F.gaussian_source(lat, 5, parity=0))
    masses = [0.05, 0.1, 0.2, 0.4]
    coef = [0.5, 0.25, 0.125, 0.0625]
    resid = 1e-22
    ft = torch.tensor(fat, requires_grad=True)
    lt = torch.tensor(lng, requires_grad=True)
    print("[case] autograd forward 24^4", flush=True)
    S = FermionAction.apply(ft, lt, s, phi, masses, coef, resid)
    print("[case] autograd backward 24^4", flush=True)
    S.backward()
    gf, gl = ft.grad.numpy(), lt.grad.numpy()

    # the oracle's CPU FP64 force from the same solutions
    st = solve(s, fat, lng, phi, masses, resid)
    xs = [x.extract() for x in st.xs]
    st.free()
    orc = O.Oracle(dims, *F.raw_link_set(lat, fat, lng))
    r1, r3 = orc.fermion_oprod(coef, xs)
    orc.close()
    r1 = r1.view(np.complex128).reshape(lat.V, 4, 3, 3)
    r3 = r3.view(np.complex128).reshape(lat.V, 4, 3, 3)
    dag = lambda a: np.conj(np.swapaxes(a, -1, -2))
    assert rel_l2(gf, 2 * dag(r1)) < 1e-11
    assert rel_l2(gl, 2 * dag(r3)) < 1e-11

    # central differences along random directions on 64 random links
    rng = np.random.default_rng(3)
    sites = rng.choice(lat.V, 64, replace=False)
    mus = rng.integers(0, 4, 64)
    E = rng.standard_normal((64, 3, 3)) + 1j * rng.standard_normal((64, 3, 3))
    for which, idx in (("fat", range(0, 32)), ("long", range(32, 64))):
        Ef = np.zeros_like(fat)
        El = np.zeros_like(lng)
        Et = Ef if which == "fat" else El
        g = gf if which == "fat" else gl
        for j in idx:
            Et[sites[j], mus[j]] = E[j]
        pred = float(np.real(np.vdot(g, Et)))   # Re sum conj(grad) E
        print(f"[case] central difference, 32 {which} links", flush=True)
        fd = _exact_central_difference(s, lat, fat, lng, phi, masses, coef, Ef, El, 1e-5, resid)
        scale = float(np.linalg.norm(Et) * np.linalg.norm(g[sites[list(idx)], mus[list(idx)]]))
        print(f"[case]   fd {fd:.12e} autograd {pred:.12e} rel {abs(fd - pred) / scale:.2e}", flush=True)
        assert abs(fd - pred) < 1e-8 * scale
