"""Dslash and field-layout parity: HIP library (through the C ABI) vs the CPU
oracle on identical seeded fields.

Tolerances (relative L2): FP64 1e-13, FP32 1e-6 against the FP64 oracle.
Full-size (32^4) cases also check size-independent properties.
"""
import numpy as np
import pytest

import oracle as O
from conftest import lattice_session
from helpers import allowed_momenta, cdot, half, random_links, rel_l2
from pyqphix import api as A
from pyqphix import fields as F

pytestmark = pytest.mark.gpu

TOL = {"D": 1e-13, "F": 1e-6}
NP = {"D": np.float64, "F": np.float32}

_CACHE = {}


def links_for(dims, kind, seed=11):
    key = (dims, kind, seed)
    if key not in _CACHE:
        lat = F.Lattice(dims)
        links = random_links(lat, kind, seed)
        raw = [F.to_raw(a) for a in links]
        _CACHE.clear()
        _CACHE[key] = (lat, links, raw)
    return _CACHE[key]


def oracle_for(lat, raw):
    return O.Oracle(lat.dims, *raw)


@pytest.mark.parametrize("prec", ["D", "F"])
@pytest.mark.parametrize("parity", [A.EVEN, A.ODD, A.EVENODD])
def test_vector_roundtrip(prec, parity):
    s = lattice_session((8, 8, 8, 8))
    lat = F.Lattice(s.dims)
    raw = F.to_raw(F.gaussian_source(lat, 3), NP[prec])
    v = A.Vector(s, raw, parity, prec)
    out = v.extract()
    for cb, bit in ((0, A.EVEN), (1, A.ODD)):
        sl = slice(6 * cb * lat.vh, 6 * (cb + 1) * lat.vh)
        if parity & bit:
            np.testing.assert_array_equal(out[sl], raw[sl])
        else:
            assert not np.any(out[sl])
    v.free()


@pytest.mark.parametrize("prec", ["D", "F"])
@pytest.mark.parametrize("kind,recon", [("haar", 12), ("weak", 12)])
def test_links_roundtrip_and_recon(prec, kind, recon):
    s = lattice_session((8, 8, 8, 8))
    lat, links, raw = links_for(s.dims, kind)
    L = A.Links(s, *[r.astype(NP[prec]) for r in raw], prec=prec)
    assert L.recon == recon
    fat, lng = L.extract()
    np.testing.assert_array_equal(fat, raw[0].astype(NP[prec]))
    tol = 1e-14 if prec == "D" else 2e-6
    assert np.max(np.abs(lng - raw[1].astype(NP[prec]))) < tol
    L.free()
    s.lib.QPHIX_MI355X_set_long_recon(18)
    L = A.Links(s, *[r.astype(NP[prec]) for r in raw], prec=prec)
    assert L.recon == 18
    fat, lng = L.extract()
    np.testing.assert_array_equal(lng, raw[1].astype(NP[prec]))
    L.free()
    s.lib.QPHIX_MI355X_set_long_recon(0)


@pytest.mark.parametrize("prec", ["D", "F"])
@pytest.mark.parametrize("recon", [0, 18])
@pytest.mark.parametrize("parity", [A.EVEN, A.ODD], ids=["even", "odd"])
def test_single_parity_links(qlib, prec, recon, parity):
    """asqtad_create_L_from_raw / extract_L_to_raw with QPHIX_EVEN or QPHIX_ODD
    (qphix_d3.cpp:970-1206): the raw arrays are parity-blocked (even sites,
    then odd), and only the given parity's block is read and written.  The
    one-parity set stores the same links as the EVENODD set (same Naik form,
    same decoded links) and gives a bit-identical Dslash onto its parity; the
    Dslash onto the other parity and the solvers, which need both parities,
    leave a recorded error and launch nothing."""
    s = lattice_session((8, 8, 8, 8))
    lat, links, raw = links_for(s.dims, "haar")
    rawp = [r.astype(NP[prec]) for r in raw]
    cb = 0 if parity == A.EVEN else 1
    n = 72 * lat.V // 2
    mine, other = slice(cb * n, (cb + 1) * n), slice((1 - cb) * n, (2 - cb) * n)
    # the other parity's blocks hold garbage: the import must not read them
    junk = [r.copy() for r in rawp]
    for r in junk:
        r[other] = np.nan
    qlib.QPHIX_MI355X_set_long_recon(recon)
    try:
        L1 = A.Links(s, *junk, parity=parity, prec=prec)
        Lb = A.Links(s, *rawp, prec=prec)
    finally:
        qlib.QPHIX_MI355X_set_long_recon(0)
    print(f"[case] single-parity links prec={prec} recon={recon} parity={parity}", flush=True)
    assert L1.recon == Lb.recon == (12 if recon == 0 else 18)
    fat, lng = L1.extract(parity)
    fat_b, lng_b = Lb.extract(A.EVENODD)
    np.testing.assert_array_equal(fat[mine], rawp[0][mine])
    np.testing.assert_array_equal(lng[mine], lng_b[mine])
    assert not fat[other].any() and not lng[other].any()
    if recon == 18:
        np.testing.assert_array_equal(lng[mine], rawp[1][mine])
    else:
        assert np.max(np.abs(lng[mine] - rawp[1][mine])) < (1e-14 if prec == "D" else 2e-6)
    src = F.to_raw(F.gaussian_source(lat, 17), NP[prec])
    vin = A.Vector(s, src, A.EVENODD, prec)
    v1 = A.Vector(s, None, A.EVENODD, prec)
    vb = A.Vector(s, None, A.EVENODD, prec)
    A.dslash(L1, v1, vin, parity)
    A.dslash(Lb, vb, vin, parity)
    np.testing.assert_array_equal(half(v1.extract(), lat, cb), half(vb.extract(), lat, cb))
    qlib.QPHIX_MI355X_set_abort_on_error(0)
    qlib.QPHIX_MI355X_clear_error()
    try:
        before = v1.extract()
        A.dslash(L1, v1, vin, A.ODD if parity == A.EVEN else A.EVEN)
        nerr, msg = qlib.last_error()
        assert nerr > 0 and "parity missing" in msg, msg
        np.testing.assert_array_equal(v1.extract(), before)
        qlib.QPHIX_MI355X_clear_error()
        b = F.to_raw(F.gaussian_source(lat, 18, parity=cb), NP[prec])
        vbs = A.Vector(s, b, parity, prec)
        vx = A.Vector(s, None, parity, prec)
        ret, info, _ = A.invert(L1, vx, vbs, 0.05, parity, 1e-12)
        nerr, msg = qlib.last_error()
        assert ret == 0 and info.status != 0 and nerr > 0 and "both link parities" in msg, msg
        qlib.QPHIX_MI355X_clear_error()
        outs = [A.Vector(s, None, parity, prec) for _ in range(2)]
        ret, info, _ = A.invert_multi(L1, outs, vbs, [0.05, 0.1], parity, 1e-12)
        nerr, msg = qlib.last_error()
        assert ret == 0 and nerr > 0 and "both link parities" in msg, msg
        for o in outs + [vbs, vx]:
            o.free()
    finally:
        qlib.QPHIX_MI355X_clear_error()
        qlib.QPHIX_MI355X_set_abort_on_error(1)
    for o in (L1, Lb, vin, v1, vb):
        o.free()


def test_non_unitary_long_links_stay_18():
    s = lattice_session((8, 8, 8, 8))
    lat = F.Lattice(s.dims)
    fat, lng = F.free_links(lat, 0.9, -0.07)   # scaled identity: not +-SU(3)
    raw = F.raw_link_set(lat, fat, lng)
    L = A.Links(s, *raw)
    assert L.recon == 18
    L.free()


def test_backlink_convention_probe():
    s = lattice_session((8, 8, 8, 8))
    lat, links, raw = links_for(s.dims, "haar")
    L = A.Links(s, *raw)
    fm, fp, lm, lp = L.backcheck()
    assert fm == 0.0 and lm < 1e-13       # back = -U^dagger (MILC double store)
    assert fp > 0.1 and lp > 0.1
    L.free()


def run_dslash(s, raw_links, psi_raw, parity, prec, recon_mode=0):
    s.lib.QPHIX_MI355X_set_long_recon(recon_mode)
    L = A.Links(s, *[r.astype(NP[prec]) for r in raw_links], prec=prec)
    vin = A.Vector(s, psi_raw.astype(NP[prec]), A.EVENODD, prec)
    vout = A.Vector(s, None, A.EVENODD, prec)
    A.dslash(L, vout, vin, parity)
    out = vout.extract()
    recon = L.recon
    for o in (L, vin, vout):
        o.free()
    s.lib.QPHIX_MI355X_set_long_recon(0)
    return out, recon


@pytest.mark.parametrize("dims", [(8, 8, 8, 8), (4, 8, 4, 12)])
@pytest.mark.parametrize("prec", ["D", "F"])
@pytest.mark.parametrize("recon_mode", [0, 18])
@pytest.mark.parametrize("kind", ["haar", "weak"])
def test_dslash_matches_oracle(dims, prec, recon_mode, kind):
    s = lattice_session(dims)
    lat, links, raw = links_for(dims, kind)
    orc = oracle_for(lat, raw)
    psi = F.to_raw(F.gaussian_source(lat, 5))
    for parity, cb in ((A.EVEN, 0), (A.ODD, 1)):
        print(f"[case] dims={dims} prec={prec} recon={recon_mode} kind={kind} cb={cb}", flush=True)
        out, recon = run_dslash(s, raw, psi, parity, prec, recon_mode)
        assert recon == (18 if recon_mode == 18 else 12)
        ref = orc.dslash(half(psi, lat, 1 - cb), cb)
        assert rel_l2(half(out, lat, cb), ref) < TOL[prec]
        assert not np.any(half(out, lat, 1 - cb))   # the other half is untouched


@pytest.mark.parametrize("prec", ["D", "F"])
def test_dslash_free_field_spectrum(prec):
    s = lattice_session((8, 8, 8, 8))
    lat = F.Lattice(s.dims)
    c1, c3, m = 0.9, -0.07, 0.05
    fat, lng = F.free_links(lat, c1, c3)
    raw = F.raw_link_set(lat, fat, lng)
    L = A.Links(s, *[r.astype(NP[prec]) for r in raw], prec=prec)
    tmp = A.Vector(s, None, A.ODD, prec)
    out = A.Vector(s, None, A.EVEN, prec)
    for p in allowed_momenta(lat, 4):
        chi = F.to_raw(F.plane_wave(lat, p, 1, parity=0), NP[prec])
        vin = A.Vector(s, chi, A.EVEN, prec)
        A.dslash(L, tmp, vin, A.ODD)
        A.dslash(L, out, tmp, A.EVEN)
        ce = half(chi, lat, 0).astype(np.float64)
        mchi = 4 * m * m * ce - half(out.extract(), lat, 0)
        lam = F.free_eigenvalue(p, c1, c3, 4 * m * m)
        assert rel_l2(mchi, lam * ce) < (1e-13 if prec == "D" else 2e-6)
        vin.free()
    for o in (L, tmp, out):
        o.free()


def test_dslash_errors_are_reported_not_launched():
    s = lattice_session((8, 8, 8, 8))
    lat, links, raw = links_for(s.dims, "weak")
    s.lib.QPHIX_MI355X_set_abort_on_error(0)
    s.lib.QPHIX_MI355X_clear_error()
    try:
        L = A.Links(s, *raw)
        ve = A.Vector(s, None, A.EVEN)
        ve2 = A.Vector(s, None, A.EVEN)
        A.dslash(L, ve2, ve, A.EVEN)          # in lacks the odd half
        n, msg = s.lib.last_error()
        assert n == 1 and "opposite" in msg
        A.dslash(L, ve2, ve, A.EVENODD)       # parity must be EVEN or ODD
        assert s.lib.last_error()[0] == 2
        for o in (L, ve, ve2):
            o.free()
    finally:
        s.lib.QPHIX_MI355X_clear_error()
        s.lib.QPHIX_MI355X_set_abort_on_error(1)


@pytest.mark.parametrize("prec", ["D", "F"])
def test_dslash_full_size_32_4(prec):
    """32^4 (the benchmark lattice): parity against the oracle plus
    anti-Hermiticity, which holds at any size."""
    dims = (32, 32, 32, 32)
    s = lattice_session(dims)
    lat, links, raw = links_for(dims, "weak", seed=1002)
    orc = oracle_for(lat, raw)
    psi = F.to_raw(F.gaussian_source(lat, 7))
    chi = F.to_raw(F.gaussian_source(lat, 8))
    out_e, recon = run_dslash(s, raw, psi, A.EVEN, prec)
    out_o, _ = run_dslash(s, raw, chi, A.ODD, prec)
    assert recon == 12
    assert rel_l2(half(out_e, lat, 0), orc.dslash(half(psi, lat, 1), 0)) < TOL[prec]
    # <chi_e, D_eo psi_o> = -<D_oe chi_e, psi_o>
    lhs = cdot(half(chi, lat, 0).view(np.complex128), half(out_e, lat, 0).astype(np.float64).view(np.complex128))
    rhs = -cdot(half(out_o, lat, 1).astype(np.float64).view(np.complex128), half(psi, lat, 1).view(np.complex128))
    assert abs(lhs - rhs) < (1e-11 if prec == "D" else 1e-4) * abs(lhs)
