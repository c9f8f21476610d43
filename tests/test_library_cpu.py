"""CPU-only checks of the C-ABI library: it builds, loads, exports every symbol
include/*.h declares, its headers compile as C and C++ in both precisions, its
public structs have the reference's memory layout, and it fails cleanly
(without touching a GPU) on bad input."""
import ctypes as C
import os
import subprocess
import textwrap

import pytest

from conftest import ROOT

INC = os.path.join(ROOT, "include")
REF = "/root/reference"


def _lib():
    from pyqphix import lib as L
    if not os.path.exists(L.LIB_PATH):
        subprocess.run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "milc-qphix_amd")], check=True)
    return L


def test_every_declared_symbol_is_exported():
    L = _lib()
    lib = L.Lib()
    names = L.declared_symbols()
    assert len(names) >= 50
    missing = [n for n in names if not lib.exported(n)]
    assert not missing, missing


@pytest.mark.parametrize("lang,prec", [("c", "F"), ("c", "D"), ("c++", "F"), ("c++", "D")])
def test_headers_compile_and_map_generic_names(tmp_path, lang, prec):
    src = tmp_path / ("t.c" if lang == "c" else "t.cpp")
    src.write_text(textwrap.dedent(f"""
        #define QPHIX_Precision '{prec}'
        #include <qphix.h>
        #include <qphix_mi355x.h>
        int main(void) {{
          QPHIX_ColorVector *v = QPHIX_create_V_from_raw(0, QPHIX_EVEN);
          QPHIX_{prec}3_ColorVector *w = v;
          (void)w;
        #ifndef __cplusplus
          /* C99 compound-literal defaults (C only, as in the reference) */
          QPHIX_invert_arg_t a = QPHIX_INVERT_ARG_DEFAULT;
          QPHIX_resid_arg_t r = QPHIX_RESID_ARG_DEFAULT;
          (void)a; (void)r;
        #endif
          return QPHIX_OPP_DIR(3) == 2 ? 0 : 1;
        }}
        """))
    cc = "gcc" if lang == "c" else "g++"
    subprocess.run([cc, "-Wall", "-Werror", "-I", INC, "-c", str(src), "-o", str(tmp_path / "t.o")],
                   check=True)


LAYOUT_PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include <qphix_int.h>
#define P(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))
int main(void) {
  printf("layout %zu info %zu coeffs %zu inv %zu resid %zu\n", sizeof(QPHIX_layout_t),
         sizeof(QPHIX_info_t), sizeof(QPHIX_gauge_coeffs_t), sizeof(QPHIX_invert_arg_t),
         sizeof(QPHIX_resid_arg_t));
  P(QPHIX_layout_t, latsize); P(QPHIX_layout_t, machsize); P(QPHIX_layout_t, sites_on_node);
  P(QPHIX_layout_t, mpi_comm); P(QPHIX_info_t, status); P(QPHIX_invert_arg_t, nrestart);
  P(QPHIX_resid_arg_t, relresid); P(QPHIX_resid_arg_t, final_iter);
  printf("enum %d %d %d %d\n", QPHIX_EVEN, QPHIX_ODD, QPHIX_EVENODD, QPHIX_MEM_ERROR);
  return 0;
}
"""


def _probe(tmp_path, incdir):
    src = tmp_path / "probe.c"
    src.write_text(LAYOUT_PROBE)
    exe = tmp_path / ("probe_" + os.path.basename(incdir))
    subprocess.run(["gcc", "-I", incdir, str(src), "-o", str(exe)], check=True)
    return subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout


def test_struct_layout_matches_reference(tmp_path):
    """The public structs must be layout-identical to the reference's
    qphix_int.h (read as text from the reference tree when present)."""
    mine = _probe(tmp_path, INC)
    ref_hdr = os.path.join(REF, "qphix_int.h")
    if not os.path.exists(ref_hdr):
        pytest.skip("reference tree not present")
    assert _probe(tmp_path, REF) == mine


def test_ctypes_structs_match_c_layout(tmp_path):
    from pyqphix import lib as L
    out = _probe(tmp_path, INC).split("\n")[0].split()
    sizes = dict(zip(out[0::2], map(int, out[1::2])))
    assert C.sizeof(L.Layout) == sizes["layout"]
    assert C.sizeof(L.Info) == sizes["info"]
    assert C.sizeof(L.GaugeCoeffs) == sizes["coeffs"]
    assert C.sizeof(L.InvertArg) == sizes["inv"]
    assert C.sizeof(L.ResidArg) == sizes["resid"]


def test_bad_layout_and_uninitialised_calls_fail_cleanly():
    """Runs in a child process: checks happen before any device query."""
    code = textwrap.dedent(f"""
        import sys; sys.path.insert(0, {os.path.join(ROOT, 'milc-qphix_amd')!r})
        from pyqphix import lib as L
        lib = L.Lib()
        lib.QPHIX_MI355X_set_abort_on_error(0)
        assert lib.init((6, 4, 4, 4)) == L.QPHIX_FAIL
        n, msg = lib.last_error()
        assert n == 1 and "multiple of 4" in msg, msg
        assert not lib.QPHIX_D3_create_V_from_raw(None, L.QPHIX_EVEN)
        assert lib.last_error()[0] == 2
        print("ok")
    """)
    r = subprocess.run(["python", "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ok" in r.stdout, r.stderr


def test_abort_on_error_default_exits_like_myassert():
    code = textwrap.dedent(f"""
        import sys; sys.path.insert(0, {os.path.join(ROOT, 'milc-qphix_amd')!r})
        from pyqphix import lib as L
        lib = L.Lib()
        lib.init((6, 4, 4, 4))
        print("not reached")
    """)
    r = subprocess.run(["python", "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "not reached" not in r.stdout
    assert "QPHIX ERROR" in r.stderr
