"""ctypes wrapper of the CPU oracle (liboracle.so, built from qphix_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py, always as the checker or the baseline, never as a
compute path of the library.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")

EVEN, ODD, EVENODD = 2, 1, 3


def build(force=False):
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp = C.c_void_p
        L.oracle_setup.restype = C.c_int
        L.oracle_setup.argtypes = [C.POINTER(C.c_int)]
        L.oracle_neighbours.argtypes = [C.c_int, vp]
        L.oracle_num_threads.restype = C.c_int
        for p in ("d", "f"):
            getattr(L, f"oracle_links_create_{p}").restype = vp
            getattr(L, f"oracle_links_create_{p}").argtypes = [vp, vp, vp, vp]
            getattr(L, f"oracle_links_destroy_{p}").argtypes = [vp]
            getattr(L, f"oracle_dslash_{p}").argtypes = [vp, vp, vp, C.c_int]
            R = C.c_double if p == "d" else C.c_float
            getattr(L, f"oracle_invert_{p}").restype = C.c_int
            getattr(L, f"oracle_invert_{p}").argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double,
                                                         C.c_double, R, vp, vp, vp, vp]
            getattr(L, f"oracle_invert_multi_{p}").restype = C.c_int
            getattr(L, f"oracle_invert_multi_{p}").argtypes = [vp, C.c_int, C.c_int, C.c_int,
                                                               C.c_double, vp, C.c_int, vp, vp, vp, vp]
        for p in ("d", "f"):
            getattr(L, f"oracle_relative_residue_{p}").restype = C.c_double
            getattr(L, f"oracle_relative_residue_{p}").argtypes = [vp, vp]
        for p, R in (("d", C.c_double), ("f", C.c_float)):
            getattr(L, f"oracle_invert_fewsums_{p}").restype = C.c_int
            getattr(L, f"oracle_invert_fewsums_{p}").argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double,
                                                                 R, vp, vp, vp, vp]
        L.oracle_gauge_action.restype = C.c_double
        L.oracle_gauge_action.argtypes = [vp, vp, vp, vp, C.c_int, C.c_int]
        L.oracle_gauge_force.argtypes = [vp, vp, vp, C.c_double, vp, vp]
        L.oracle_gauge_update.argtypes = [vp, vp, vp, C.c_double]
        L.oracle_matrix_exp.argtypes = [vp, vp]
        L.oracle_gauge_md.argtypes = [vp, vp, vp, vp, C.c_double, C.c_int, C.c_int]
        L.oracle_links_from_gauge.argtypes = [vp, vp, vp, C.c_int, C.c_double, C.c_int, vp, vp]
        L.oracle_fermion_oprod_d.argtypes = [vp, C.c_int, vp, vp, vp, vp]
        _lib = L
    return _lib


def gauge_action(dims, U_raw, coef, box=None, box_r=0, link_mu=-1):
    """S(U) = -sum over geometric loops of c Re Tr U_loop; `box` restricts the
    sum to loops starting within Chebyshev distance box_r of site box, and
    link_mu >= 0 further to the loops that contain link (box, link_mu)."""
    d = np.ascontiguousarray(dims, dtype=np.int32)
    U = np.ascontiguousarray(U_raw, dtype=np.float64)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    b = None if box is None else np.ascontiguousarray(box, dtype=np.int32)
    return lib().oracle_gauge_action(_p(d), _p(U), _p(c), None if b is None else _p(b), box_r,
                                     link_mu)


def gauge_force(dims, U_raw, coef, eps, mom_raw):
    """Returns (updated momentum, W) as raw 72-real/site arrays."""
    d = np.ascontiguousarray(dims, dtype=np.int32)
    U = np.ascontiguousarray(U_raw, dtype=np.float64)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    mom = np.ascontiguousarray(mom_raw, dtype=np.float64).copy()
    W = np.zeros_like(mom)
    lib().oracle_gauge_force(_p(d), _p(U), _p(c), eps, _p(mom), _p(W))
    return mom, W


def gauge_update(dims, U_raw, mom_raw, eps):
    """U <- exp(eps A) U (raw 72-real/site arrays); returns the new U."""
    d = np.ascontiguousarray(dims, dtype=np.int32)
    U = np.ascontiguousarray(U_raw, dtype=np.float64).copy()
    mom = np.ascontiguousarray(mom_raw, dtype=np.float64)
    lib().oracle_gauge_update(_p(d), _p(U), _p(mom), eps)
    return U


def matrix_exp(X):
    """exp of one complex 3x3 matrix (the oracle's Taylor / scaling-and-squaring)."""
    x = np.ascontiguousarray(np.asarray(X, dtype=np.complex128).reshape(3, 3)).view(np.float64).ravel()
    out = np.empty(18)
    lib().oracle_matrix_exp(_p(x), _p(out))
    return out.view(np.complex128).reshape(3, 3)


def gauge_md(dims, U_raw, mom_raw, coef, eps, nsteps, scheme=0):
    """Pure-gauge MD (0 leapfrog, 1 Omelyan); returns (U, mom)."""
    d = np.ascontiguousarray(dims, dtype=np.int32)
    U = np.ascontiguousarray(U_raw, dtype=np.float64).copy()
    mom = np.ascontiguousarray(mom_raw, dtype=np.float64).copy()
    c = np.ascontiguousarray(coef, dtype=np.float64)
    lib().oracle_gauge_md(_p(d), _p(U), _p(mom), _p(c), eps, nsteps, scheme)
    return U, mom


ASQTAD = (5.0 / 8, -1.0 / 24, 1.0 / 16, 1.0 / 64, 1.0 / 384, -1.0 / 16)
FAT7 = (1.0 / 8, 0.0, 1.0 / 16, 1.0 / 64, 1.0 / 384, 0.0)


def links_from_gauge(dims, U_raw, coef=ASQTAD, hisq=False, naik_eps=0.0, apbc=True):
    """Phased forward fat and Naik links (raw 72-real/site arrays) of the thin
    field U: coef = (one_link, naik, three_staple, five_staple, seven_staple,
    lepage), or HISQ."""
    d = np.ascontiguousarray(dims, dtype=np.int32)
    U = np.ascontiguousarray(U_raw, dtype=np.float64)
    c = np.ascontiguousarray(coef, dtype=np.float64)
    fat = np.zeros_like(U)
    lng = np.zeros_like(U)
    lib().oracle_links_from_gauge(_p(d), _p(U), _p(c), int(hisq), naik_eps, int(apbc), _p(fat), _p(lng))
    return fat, lng


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


class Oracle:
    """One lattice's geometry plus one link set, in one precision."""

    def __init__(self, dims, fat, lng, fatback, lngback, dtype=np.float64):
        self.L = lib()
        self.dims = tuple(dims)
        d = (C.c_int * 4)(*dims)
        if self.L.oracle_setup(d) != 0:
            raise ValueError("bad lattice")
        self.vh = int(np.prod(dims)) // 2
        self.dtype = np.dtype(dtype)
        self.p = "d" if self.dtype == np.float64 else "f"
        self._arrs = [np.ascontiguousarray(a, dtype=self.dtype) for a in (fat, lng, fatback, lngback)]
        self.h = getattr(self.L, f"oracle_links_create_{self.p}")(*[_p(a) for a in self._arrs])

    def close(self):
        if self.h:
            getattr(self.L, f"oracle_links_destroy_{self.p}")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def neighbours(self, cb):
        out = np.empty(16 * self.vh, dtype=np.uint32)
        self.L.oracle_neighbours(cb, _p(out))
        return out.reshape(self.vh, 16)

    def dslash(self, inp_half, cb):
        """out(half of parity cb) = D in(half of parity 1-cb)."""
        inp = np.ascontiguousarray(inp_half, dtype=self.dtype)
        out = np.empty(6 * self.vh, dtype=self.dtype)
        getattr(self.L, f"oracle_dslash_{self.p}")(self.h, _p(out), _p(inp), cb)
        return out

    def invert(self, b_full, mass, parity, resid, relresid=0.0, niter=1000, nrestart=5, x0=None,
               fewsums=False):
        """CG (ks_d_congrad_fn.cpp:221-619); fewsums=True: its FEWSUMS variant
        (relresid must then be 0)."""
        b = np.ascontiguousarray(b_full, dtype=self.dtype)
        x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, dtype=self.dtype).copy()
        od = np.zeros(5)
        oi = np.zeros(2, dtype=np.int32)
        if fewsums:
            assert relresid == 0.0
            ret = getattr(self.L, f"oracle_invert_fewsums_{self.p}")(self.h, parity, niter, nrestart, resid,
                                                                     mass, _p(x), _p(b), _p(od), _p(oi))
        else:
            ret = getattr(self.L, f"oracle_invert_{self.p}")(self.h, parity, niter, nrestart, resid,
                                                             relresid, mass, _p(x), _p(b), _p(od), _p(oi))
        return x, dict(ret=ret, final_rsq=od[0], final_rel=od[1], size_r=od[2], size_relr=od[3],
                       final_flop=od[4], final_iter=int(oi[0]), final_restart=int(oi[1]))

    def relative_residue(self, r_half, x_half):
        """sqrt(2 sum_x |r(x)|^2 / |x(x)|^2 / V) over one parity's sites
        (ks_blas_utils_c.cpp:110-149), in this oracle's precision."""
        r = np.ascontiguousarray(r_half, dtype=self.dtype)
        x = np.ascontiguousarray(x_half, dtype=self.dtype)
        assert r.size == x.size == 6 * self.vh
        return getattr(self.L, f"oracle_relative_residue_{self.p}")(_p(r), _p(x))

    def invert_multi(self, b_full, masses, parity, resid, niter=1000, nrestart=5):
        b = np.ascontiguousarray(b_full, dtype=self.dtype)
        xs = [np.zeros_like(b) for _ in masses]
        ptrs = (C.c_void_p * len(masses))(*[x.ctypes.data for x in xs])
        m = np.ascontiguousarray(masses, dtype=self.dtype)
        od = np.zeros(2)
        oi = np.zeros(1, dtype=np.int32)
        ret = getattr(self.L, f"oracle_invert_multi_{self.p}")(
            self.h, parity, niter, nrestart, resid, _p(m), len(masses), C.cast(ptrs, C.c_void_p),
            _p(b), _p(od), _p(oi))
        return xs, dict(ret=ret, final_rsq=od[0], final_flop=od[1], final_iter=int(oi[0]))

    def fermion_oprod(self, coef, xs):
        """Fermion-force outer products (raw 72-real/site arrays, both parities)
        for S = sum_k coef[k] Re phi^dag X_k given the solutions X_k (full raw
        vectors, even half used): returns (O_fat, O_long)."""
        assert self.p == "d"
        xs = [np.ascontiguousarray(x, dtype=np.float64) for x in xs]
        ptrs = (C.c_void_p * len(xs))(*[x.ctypes.data for x in xs])
        c = np.ascontiguousarray(coef, dtype=np.float64)
        o1 = np.zeros(72 * 2 * self.vh)
        o3 = np.zeros(72 * 2 * self.vh)
        self.L.oracle_fermion_oprod_d(self.h, len(xs), _p(c), C.cast(ptrs, C.c_void_p), _p(o1), _p(o3))
        return o1, o3
