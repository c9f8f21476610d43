"""Object wrapper over the C interface (ctypes), for tests and benchmarks.

Mirrors the reference's API one to one: Vector = QPHIX_*3_ColorVector,
Links = QPHIX_*3_FermionLinksAsqtad, Gauge = QPHIX_*3_GaugeField,
Force = QPHIX_*3_Force; dslash / invert / invert_multi /
symanzik_1loop_gauge_force call the functions of the same name.
"""
import ctypes as C

import numpy as np

from . import lib as _L
from .lib import EVEN, ODD, EVENODD, AsqtadCoeffs, Info, InvertArg, ResidArg, GaugeCoeffs, ptr

_NP = {"D": np.float64, "F": np.float32}
_CR = {"D": C.c_double, "F": C.c_float}


class Session:
    """QPHIX_init .. QPHIX_finalize for one lattice."""

    def __init__(self, dims, lib=None):
        self.lib = lib or _L.load()
        self.dims = tuple(dims)
        self.V = int(np.prod(dims))
        self.vh = self.V // 2
        st = self.lib.init(self.dims)
        if st != 0:
            n, msg = self.lib.last_error()
            raise RuntimeError(f"QPHIX_init failed: {msg}")

    def close(self):
        self.lib.QPHIX_finalize()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def fn(self, prec, name):
        return getattr(self.lib, f"QPHIX_{prec}3_{name}")


class Vector:
    def __init__(self, s, raw, parity, prec="D"):
        self.s, self.prec, self.parity = s, prec, parity
        if raw is None:
            self._src = None
            self.h = s.fn(prec, "create_V_from_raw")(None, parity)
        else:
            self._src = np.ascontiguousarray(raw, dtype=_NP[prec])
            assert self._src.size == 6 * s.V
            self.h = s.fn(prec, "create_V_from_raw")(ptr(self._src), parity)
        if not self.h:
            raise RuntimeError("create_V failed: %s" % (s.lib.last_error(),))

    def extract(self):
        out = np.zeros(6 * self.s.V, dtype=_NP[self.prec])
        self.s.fn(self.prec, "extract_V_to_raw")(ptr(out), self.h, self.parity)
        return out

    def free(self):
        if self.h:
            self.s.fn(self.prec, "destroy_V")(self.h)
            self.h = None


class Links:
    def __init__(self, s, fat, lng, fatback, lngback, parity=EVENODD, prec="D"):
        self.s, self.prec, self.parity = s, prec, parity
        arrs = [np.ascontiguousarray(a, dtype=_NP[prec]) for a in (fat, lng, fatback, lngback)]
        self.h = s.fn(prec, "asqtad_create_L_from_raw")(*[ptr(a) for a in arrs], parity)
        if not self.h:
            raise RuntimeError("create_L failed: %s" % (s.lib.last_error(),))

    @property
    def recon(self):
        return self.s.fn(self.prec, "asqtad_links_recon")(self.h)

    def extract(self, parity=EVENODD):
        fat = np.zeros(72 * self.s.V, dtype=_NP[self.prec])
        lng = np.zeros_like(fat)
        self.s.fn(self.prec, "asqtad_extract_L_to_raw")(ptr(fat), ptr(lng), self.h, parity)
        return fat, lng

    def backcheck(self):
        out = (C.c_double * 4)()
        self.s.lib.QPHIX_D3_asqtad_links_backcheck(self.h, out)
        return list(out)

    def free(self):
        if self.h:
            self.s.fn(self.prec, "asqtad_destroy_L")(self.h)
            self.h = None


class LinksFromGauge(Links):
    """Fermion links built on the device from a Gauge field:
    QPHIX_*3_asqtad_create_L_from_G (coeffs = (one_link, naik, three_staple,
    five_staple, seven_staple, lepage)) or QPHIX_*3_hisq_create_L_from_G."""

    def __init__(self, G, coeffs=None, hisq=False, naik_eps=0.0, periodic_t=False):
        self.s, self.prec, self.parity = G.s, G.prec, EVENODD
        self.info = Info()
        flags = 1 if periodic_t else 0
        if hisq:
            self.h = G.s.fn(G.prec, "hisq_create_L_from_G")(C.byref(self.info), G.h, naik_eps, flags)
        else:
            c = AsqtadCoeffs(*coeffs)
            self.h = G.s.fn(G.prec, "asqtad_create_L_from_G")(C.byref(self.info), C.byref(c), G.h, flags)
        if not self.h:
            raise RuntimeError("create_L_from_G failed: %s" % (G.s.lib.last_error(),))


class Gauge:
    """QPHIX_*3_GaugeField from a raw 72-real/site array (all links)."""

    def __init__(self, s, raw, parity=EVENODD, prec="D"):
        self.s, self.prec = s, prec
        self._src = np.ascontiguousarray(raw, dtype=_NP[prec])
        self.h = s.fn(prec, "create_G_from_raw")(ptr(self._src), parity)
        if not self.h:
            raise RuntimeError("create_G failed: %s" % (s.lib.last_error(),))

    def extract(self, parity=EVENODD):
        out = np.zeros(72 * self.s.V, dtype=_NP[self.prec])
        self.s.fn(self.prec, "extract_G_to_raw")(ptr(out), self.h, parity)
        return out

    def free(self):
        if self.h:
            self.s.fn(self.prec, "destroy_G")(self.h)
            self.h = None


class Force:
    """QPHIX_*3_Force (momentum) from a raw anti-Hermitian 72-real/site array."""

    def __init__(self, s, raw, parity=EVENODD, prec="D"):
        self.s, self.prec = s, prec
        self._src = np.ascontiguousarray(raw, dtype=_NP[prec])
        self.h = s.fn(prec, "create_F_from_raw")(ptr(self._src), parity)
        if not self.h:
            raise RuntimeError("create_F failed: %s" % (s.lib.last_error(),))

    def extract(self, parity=EVENODD):
        out = np.zeros(72 * self.s.V, dtype=_NP[self.prec])
        self.s.fn(self.prec, "extract_F_to_raw")(ptr(out), self.h, parity)
        return out

    def free(self):
        if self.h:
            self.s.fn(self.prec, "destroy_F")(self.h)
            self.h = None


def weak_links_on_device(s, lat, seed, eps=0.3, fat_sigma=0.3, chunk=1 << 20):
    """fields.make_links(kind="weak") for lattices too large for the host
    eigensolver: fat = 1 + fat_sigma Gaussian (host RNG); long = exp(eps A)
    with A = (z - z^dag)/2 traceless, z complex Gaussian -- the distribution of
    exp(i eps H), H Gaussian Hermitian traceless -- formed by the library's own
    gauge update (QPHIX_D3_update_gauge on unit links).  Same distribution as
    make_links, not the same numbers.  Synthetic input only.  Returns the
    forward links (complex (V, 4, 3, 3)) with the staggered phases and the
    antiperiodic time boundary folded in."""
    from . import fields as F
    rng = np.random.default_rng(seed)
    fat = F.near_identity(rng, lat.V * 4, fat_sigma).reshape(lat.V, 4, 3, 3)
    mom = np.empty((lat.V, 4, 3, 3), dtype=np.complex128)
    for a in range(0, lat.V, chunk):
        b = min(lat.V, a + chunk)
        mom[a:b] = F.antihermitian_traceless(rng, (b - a, 4))
    Fm = Force(s, mom.view(np.float64).ravel())
    del mom
    one = np.zeros((lat.V, 4, 3, 3), dtype=np.complex128)
    for k in range(3):
        one[..., k, k] = 1.0
    G = Gauge(s, one.view(np.float64).ravel())
    del one
    info = update_gauge(G, Fm, eps)
    if info.status != 0:
        raise RuntimeError("update_gauge failed: %s" % (s.lib.last_error(),))
    Fm.free()
    del Fm
    lng = G.extract().view(np.complex128).reshape(lat.V, 4, 3, 3)
    G.free()
    del G
    return F.fold_phases(lat, fat, 1), F.fold_phases(lat, lng, 3)


def gauge_force(G, Fm, coeffs, eps):
    info = Info()
    c = GaugeCoeffs(coeffs[0], coeffs[1], coeffs[2], coeffs[3] if len(coeffs) > 3 else 0.0)
    G.s.fn(G.prec, "symanzik_1loop_gauge_force")(C.byref(info), G.h, Fm.h, C.byref(c), eps)
    return info


def update_gauge(G, Fm, eps):
    """QPHIX_*3_update_gauge: U <- exp(eps A) U with A the momentum of Fm."""
    info = Info()
    G.s.fn(G.prec, "update_gauge")(C.byref(info), G.h, Fm.h, eps)
    return info


def gauge_md(G, Fm, coeffs, eps, nsteps, scheme=0):
    """QPHIX_*3_gauge_md: pure-gauge MD steps (0 leapfrog, 1 Omelyan)."""
    info = Info()
    c = GaugeCoeffs(coeffs[0], coeffs[1], coeffs[2], coeffs[3] if len(coeffs) > 3 else 0.0)
    G.s.fn(G.prec, "gauge_md")(C.byref(info), G.h, Fm.h, C.byref(c), eps, nsteps, scheme)
    return info


def dslash(L, out, inp, parity):
    L.s.fn(L.prec, "asqtad_dslash")(L.h, out.h, inp.h, parity)


def dslash_faces(L, out, inp, parity, cut_mask, reps=0):
    """QPHIX_*3_asqtad_dslash_faces: the Dslash with the lattice cut along the
    directions of cut_mask, faces exchanged by self-loopback (average ms when
    reps > 0)."""
    return L.s.fn(L.prec, "asqtad_dslash_faces")(L.h, out.h, inp.h, parity, cut_mask, reps)


def dslash_faces_noalloc(L, out, inp, parity, cut_mask):
    """QPHIX_*3_asqtad_dslash_faces_noalloc: dslash_faces with the face buffer
    in the library's fixed device scratch (no allocation); returns the status."""
    return L.s.fn(L.prec, "asqtad_dslash_faces_noalloc")(L.h, out.h, inp.h, parity, cut_mask)


def dslash_multi(L, outs, inps, parity, reps=0):
    """QPHIX_F3_asqtad_dslash_multi (reps=0) or its timing variant (average ms)."""
    assert L.prec == "F" and len(outs) == len(inps)
    n = len(inps)
    ho = (C.c_void_p * n)(*[v.h for v in outs])
    hi = (C.c_void_p * n)(*[v.h for v in inps])
    if reps:
        return L.s.lib.QPHIX_F3_asqtad_dslash_multi_time(L.h, ho, hi, n, parity, reps)
    L.s.lib.QPHIX_F3_asqtad_dslash_multi(L.h, ho, hi, n, parity)
    return None


def mrhs_tune(L):
    """QPHIX_F3_asqtad_mrhs_tune: (best block size NR, its best sources per
    work-item K, {(NR, K): microseconds per source})."""
    assert L.prec == "F"
    us = (C.c_double * 12)()
    k = C.c_int(0)
    nr = L.s.lib.QPHIX_F3_asqtad_mrhs_tune(L.h, us, C.byref(k))
    if nr == 0:
        raise RuntimeError("QPHIX_F3_asqtad_mrhs_tune failed")
    table = {(n, kk): us[3 * c + q] for c, n in enumerate((1, 2, 4, 8))
             for q, kk in enumerate((1, 2, 4)) if us[3 * c + q] > 0}
    return nr, k.value, table


def dslash_time(L, out, inp, parity, reps):
    return L.s.fn(L.prec, "asqtad_dslash_time")(L.h, out.h, inp.h, parity, reps)


def dslash_sloppy(L, out, inp, parity, link_bits, reps=1):
    """out = D inp with the FP64 links' sloppy copy (link_bits 32: FP32,
    16: 16-bit fixed point) on FP32 vectors; milliseconds per application."""
    return L.s.fn("D", "asqtad_dslash_sloppy")(L.h, out.h, inp.h, parity, link_bits, reps)


def invert(L, out, inp, mass, parity, resid, relresid=0.0, niter=1000, nrestart=5):
    info = Info()
    inv = InvertArg(parity, niter, 0, nrestart, parity)
    res = ResidArg(resid, 0, relresid, 0, 0, 0, 0, 0)
    ret = L.s.fn(L.prec, "asqtad_invert")(C.byref(info), L.h, C.byref(inv), C.byref(res),
                                          mass, out.h, inp.h)
    return ret, info, res


def invert_mixed(L, out, inp, mass, parity, resid, sloppy_bits=32, niter=1000, nrestart=5,
                 relresid=0.0):
    """QPHIX_D3_asqtad_invert_mixed: FP32 or FP16-stored sloppy CG with FP64
    reliable updates (D3 links and vectors)."""
    assert L.prec == "D"
    info = Info()
    inv = InvertArg(parity, niter, 0, nrestart, parity)
    res = ResidArg(resid, 0, relresid, 0, 0, 0, 0, 0)
    ret = L.s.lib.QPHIX_D3_asqtad_invert_mixed(C.byref(info), L.h, C.byref(inv), C.byref(res),
                                               mass, out.h, inp.h, sloppy_bits)
    return ret, info, res


def invert_multi_rhs(L, outs, inps, masses, parity, resid, niter=1000, nrestart=5, rhs_block=0):
    """QPHIX_D3_asqtad_invert_multi_rhs: outs[r][j] for sources inps[r] and
    masses[j]; returns (ret, info, res[r][j])."""
    assert L.prec == "D"
    nr, ns = len(inps), len(masses)
    info = Info()
    inv = InvertArg(parity, niter, 0, nrestart, parity)
    res = [[ResidArg(resid, 0, 0, 0, 0, 0, 0, 0) for _ in masses] for _ in inps]
    resp = (C.POINTER(ResidArg) * (nr * ns))(*[C.pointer(res[r][j]) for r in range(nr) for j in range(ns)])
    m = (C.c_double * ns)(*masses)
    ho = (C.c_void_p * (nr * ns))(*[outs[r][j].h for r in range(nr) for j in range(ns)])
    hi = (C.c_void_p * nr)(*[v.h for v in inps])
    ret = L.s.lib.QPHIX_D3_asqtad_invert_multi_rhs(C.byref(info), L.h, C.byref(inv), resp, m, ns,
                                                   nr, ho, hi, rhs_block)
    return ret, info, res


def invert_multi(L, outs, inp, masses, parity, resid, niter=1000, nrestart=5):
    info = Info()
    inv = InvertArg(parity, niter, 0, nrestart, parity)
    res = [ResidArg(resid, 0, 0, 0, 0, 0, 0, 0) for _ in masses]
    resp = (C.POINTER(ResidArg) * len(masses))(*[C.pointer(r) for r in res])
    m = (_CR[L.prec] * len(masses))(*masses)
    hs = (C.c_void_p * len(masses))(*[o.h for o in outs])
    ret = L.s.fn(L.prec, "asqtad_invert_multi")(C.byref(info), L.h, C.byref(inv), resp, m,
                                                len(masses), hs, inp.h)
    return ret, info, res


def fermion_force_oprod(L, coef, xs):
    """QPHIX_*3_asqtad_fermion_force_oprod: outer products (O_fat, O_long) as
    complex (V, 4, 3, 3) arrays in raw site order, for
    S = sum_k coef[k] Re phi_k^dag X_k given the solutions xs[k] (Vectors
    holding their even halves)."""
    n = len(xs)
    V = L.s.V
    o1 = np.zeros(72 * V, dtype=_NP[L.prec])
    o3 = np.zeros(72 * V, dtype=_NP[L.prec])
    info = Info()
    c = (C.c_double * max(n, 1))(*coef)
    hs = (C.c_void_p * max(n, 1))(*[x.h for x in xs])
    L.s.fn(L.prec, "asqtad_fermion_force_oprod")(C.byref(info), L.h, n, c, hs, ptr(o1), ptr(o3))
    if info.status != 0:
        raise RuntimeError("fermion_force_oprod failed: %s" % (L.s.lib.last_error(),))
    cv = lambda a: a.astype(np.float64).view(np.complex128).reshape(V, 4, 3, 3)
    return cv(o1), cv(o3), info


def fermion_force_time(L, coef, xs, reps=10):
    """(ms of the device part, ms of the outer-product kernel alone)."""
    n = len(xs)
    c = (C.c_double * n)(*coef)
    hs = (C.c_void_p * n)(*[x.h for x in xs])
    k = C.c_double(0.0)
    t = L.s.fn(L.prec, "asqtad_fermion_force_time")(L.h, n, c, hs, reps, C.byref(k))
    return t, k.value
