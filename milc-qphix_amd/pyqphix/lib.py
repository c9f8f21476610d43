"""ctypes binding of libqphixmilc_hip.so -- the MILC-facing C interface.

Every function here is a thin call through the C ABI declared in
include/qphix*.h; there is no other compute path.  Loading fails loudly when
the shared library has not been built.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "lib", "libqphixmilc_hip.so"))

QPHIX_SUCCESS, QPHIX_FAIL, QPHIX_MEM_ERROR = 0, 1, 2
QPHIX_EVEN, QPHIX_ODD, QPHIX_EVENODD = 0x02, 0x01, 0x03
EVEN, ODD, EVENODD = QPHIX_EVEN, QPHIX_ODD, QPHIX_EVENODD


class Layout(C.Structure):
    _fields_ = [("node_number", C.c_void_p), ("node_index", C.c_void_p),
                ("latdim", C.c_int), ("latsize", C.POINTER(C.c_int)),
                ("machdim", C.c_int), ("machsize", C.POINTER(C.c_int)),
                ("this_node", C.c_int), ("sites_on_node", C.c_size_t),
                ("even_sites_on_node", C.c_size_t), ("mpi_comm", C.c_void_p)]


class Info(C.Structure):
    _fields_ = [("final_sec", C.c_double), ("final_flop", C.c_double),
                ("status", C.c_int), ("count1", C.c_int), ("count2", C.c_int)]


class GaugeCoeffs(C.Structure):
    _fields_ = [("plaquette", C.c_double), ("rectangle", C.c_double),
                ("parallelogram", C.c_double), ("adjoint_plaquette", C.c_double)]


class AsqtadCoeffs(C.Structure):
    _fields_ = [("one_link", C.c_double), ("naik", C.c_double), ("three_staple", C.c_double),
                ("five_staple", C.c_double), ("seven_staple", C.c_double), ("lepage", C.c_double)]


class InvertArg(C.Structure):
    _fields_ = [("parity", C.c_int), ("max", C.c_int), ("restart", C.c_int),
                ("nrestart", C.c_int), ("evenodd", C.c_int)]


class ResidArg(C.Structure):
    _fields_ = [("resid", C.c_double), ("final_rsq", C.c_double),
                ("relresid", C.c_double), ("final_rel", C.c_double),
                ("size_r", C.c_double), ("size_relr", C.c_double),
                ("final_iter", C.c_int), ("final_restart", C.c_int)]


_dp = C.c_void_p  # raw array / handle pointers


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


def declared_symbols():
    """Every function name declared in include/*.h (parsed from the headers)."""
    import re
    inc = os.path.normpath(os.path.join(_HERE, "..", "..", "include"))
    names = set()
    for fn in os.listdir(inc):
        if not fn.endswith(".h") or fn.endswith("_generic.h"):
            continue
        txt = open(os.path.join(inc, fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"//[^\n]*", "", txt)
        for m in re.finditer(r"\b(QPHIX_[A-Za-z0-9_]+)\s*\(", txt):
            if not txt[:m.start()].rstrip().endswith("define"):
                names.add(m.group(1))
    names.discard("QPHIX_OPP_PAR")
    names.discard("QPHIX_OPP_DIR")
    return sorted(names)


class Lib:
    """Loaded library with argument types set for every entry point."""

    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        self.path = path
        self.lib = lib = C.CDLL(path)
        for P, R in (("D", C.c_double), ("F", C.c_float)):
            _sig(lib, f"QPHIX_{P}3_create_V_from_raw", _dp, [_dp, C.c_int])
            _sig(lib, f"QPHIX_{P}3_extract_V_to_raw", None, [_dp, _dp, C.c_int])
            _sig(lib, f"QPHIX_{P}3_destroy_V", None, [_dp])
            _sig(lib, f"QPHIX_{P}3_asqtad_create_L_from_raw", _dp, [_dp, _dp, _dp, _dp, C.c_int])
            _sig(lib, f"QPHIX_{P}3_asqtad_extract_L_to_raw", None, [_dp, _dp, _dp, C.c_int])
            _sig(lib, f"QPHIX_{P}3_asqtad_destroy_L", None, [_dp])
            _sig(lib, f"QPHIX_{P}3_asqtad_links_recon", C.c_int, [_dp])
            _sig(lib, f"QPHIX_{P}3_gauge_links_recon", C.c_int, [_dp])
            _sig(lib, f"QPHIX_{P}3_create_G_from_raw", _dp, [_dp, C.c_int])
            _sig(lib, f"QPHIX_{P}3_extract_G_to_raw", None, [_dp, _dp, C.c_int])
            _sig(lib, f"QPHIX_{P}3_destroy_G", None, [_dp])
            _sig(lib, f"QPHIX_{P}3_create_F_from_raw", _dp, [_dp, C.c_int])
            _sig(lib, f"QPHIX_{P}3_extract_F_to_raw", None, [_dp, _dp, C.c_int])
            _sig(lib, f"QPHIX_{P}3_destroy_F", None, [_dp])
            _sig(lib, f"QPHIX_{P}3_symanzik_1loop_gauge_force", None,
                 [C.POINTER(Info), _dp, _dp, C.POINTER(GaugeCoeffs), R])
            _sig(lib, f"QPHIX_{P}3_asqtad_dslash", None, [_dp, _dp, _dp, C.c_int])
            _sig(lib, f"QPHIX_{P}3_update_gauge", None, [C.POINTER(Info), _dp, _dp, R])
            _sig(lib, f"QPHIX_{P}3_asqtad_dslash_faces", C.c_double, [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int])
            _sig(lib, f"QPHIX_{P}3_asqtad_dslash_faces_noalloc", C.c_int, [_dp, _dp, _dp, C.c_int, C.c_int])
            _sig(lib, f"QPHIX_{P}3_asqtad_create_L_from_G", _dp,
                 [C.POINTER(Info), C.POINTER(AsqtadCoeffs), _dp, C.c_int])
            _sig(lib, f"QPHIX_{P}3_hisq_create_L_from_G", _dp, [C.POINTER(Info), _dp, C.c_double, C.c_int])
            _sig(lib, f"QPHIX_{P}3_gauge_md", None,
                 [C.POINTER(Info), _dp, _dp, C.POINTER(GaugeCoeffs), R, C.c_int, C.c_int])
            _sig(lib, f"QPHIX_{P}3_asqtad_dslash_time", C.c_double, [_dp, _dp, _dp, C.c_int, C.c_int])
            _sig(lib, f"QPHIX_{P}3_asqtad_invert", C.c_int,
                 [C.POINTER(Info), _dp, C.POINTER(InvertArg), C.POINTER(ResidArg), R, _dp, _dp])
            _sig(lib, f"QPHIX_{P}3_asqtad_invert_multi", C.c_int,
                 [C.POINTER(Info), _dp, C.POINTER(InvertArg), C.POINTER(C.POINTER(ResidArg)),
                  C.POINTER(R), C.c_int, C.POINTER(_dp), _dp])
            _sig(lib, f"QPHIX_{P}3_asqtad_fermion_force_oprod", None,
                 [C.POINTER(Info), _dp, C.c_int, C.POINTER(C.c_double), C.POINTER(_dp), _dp, _dp])
            _sig(lib, f"QPHIX_{P}3_asqtad_fermion_force_time", C.c_double,
                 [_dp, C.c_int, C.POINTER(C.c_double), C.POINTER(_dp), C.c_int,
                  C.POINTER(C.c_double)])
        for n in ("QPHIX_init", "QPHIX_init_F", "QPHIX_init_D"):
            _sig(lib, n, C.c_int, [C.POINTER(Layout)])
        for n in ("QPHIX_finalize", "QPHIX_finalize_F", "QPHIX_finalize_D"):
            _sig(lib, n, C.c_int, [])
        _sig(lib, "QPHIX_MI355X_set_abort_on_error", None, [C.c_int])
        _sig(lib, "QPHIX_MI355X_last_error", C.c_int, [C.c_char_p, C.c_int])
        _sig(lib, "QPHIX_MI355X_clear_error", None, [])
        _sig(lib, "QPHIX_MI355X_set_sync", None, [C.c_int])
        _sig(lib, "QPHIX_MI355X_sync", None, [])
        _sig(lib, "QPHIX_MI355X_stream", _dp, [])
        _sig(lib, "QPHIX_MI355X_device_info", C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_double)])
        _sig(lib, "QPHIX_MI355X_set_long_recon", None, [C.c_int])
        _sig(lib, "QPHIX_MI355X_set_gauge_recon", None, [C.c_int])
        _sig(lib, "QPHIX_MI355X_set_gauge_force_staples", None, [C.c_int])
        _sig(lib, "QPHIX_MI355X_set_mrhs_lds", None, [C.c_int])
        _sig(lib, "QPHIX_MI355X_set_msx_fuse", None, [C.c_int])
        _sig(lib, "QPHIX_MI355X_set_sloppy_link_bits", None, [C.c_int])
        _sig(lib, "QPHIX_MI355X_set_sloppy_long_recon", None, [C.c_int])
        _sig(lib, "QPHIX_D3_asqtad_dslash_sloppy", C.c_double, [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int])
        _sig(lib, "QPHIX_MI355X_set_graphs", None, [C.c_int])
        _sig(lib, "QPHIX_MI355X_debug_fail_alloc", None, [C.c_int])
        _sig(lib, "QPHIX_MI355X_set_fewsums", None, [C.c_int])
        _sig(lib, "QPHIX_D3_asqtad_links_backcheck", None, [_dp, C.POINTER(C.c_double)])
        _sig(lib, "QPHIX_MI355X_copy_time", C.c_double, [C.c_size_t, C.c_int])
        _sig(lib, "QPHIX_D3_asqtad_invert_mixed", C.c_int,
             [C.POINTER(Info), _dp, C.POINTER(InvertArg), C.POINTER(ResidArg), C.c_double, _dp,
              _dp, C.c_int])
        _sig(lib, "QPHIX_F3_asqtad_dslash_multi", None,
             [_dp, C.POINTER(_dp), C.POINTER(_dp), C.c_int, C.c_int])
        _sig(lib, "QPHIX_F3_asqtad_dslash_multi_time", C.c_double,
             [_dp, C.POINTER(_dp), C.POINTER(_dp), C.c_int, C.c_int, C.c_int])
        _sig(lib, "QPHIX_F3_asqtad_mrhs_tune", C.c_int, [_dp, C.POINTER(C.c_double), C.POINTER(C.c_int)])
        _sig(lib, "QPHIX_D3_asqtad_invert_multi_rhs", C.c_int,
             [C.POINTER(Info), _dp, C.POINTER(InvertArg), C.POINTER(C.POINTER(ResidArg)),
              C.POINTER(C.c_double), C.c_int, C.c_int, C.POINTER(_dp), C.POINTER(_dp), C.c_int])

    def __getattr__(self, name):
        return getattr(self.lib, name)

    def exported(self, name):
        try:
            getattr(self.lib, name)
            return True
        except AttributeError:
            return False

    # -- convenience ---------------------------------------------------------
    def last_error(self):
        buf = C.create_string_buffer(512)
        n = self.lib.QPHIX_MI355X_last_error(buf, 512)
        return n, buf.value.decode()

    def init(self, L):
        lat = (C.c_int * 4)(*L)
        mach = (C.c_int * 4)(1, 1, 1, 1)
        V = int(np.prod(L))
        lay = Layout(None, None, 4, lat, 4, mach, 0, V, V // 2, None)
        self._keep = (lat, mach, lay)
        return self.lib.QPHIX_init(C.byref(lay))


def ptr(a):
    """Data pointer of a C-contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


_LIB = None


def load():
    global _LIB
    if _LIB is None:
        _LIB = Lib()
    return _LIB
