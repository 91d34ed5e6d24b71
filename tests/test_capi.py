"""The C-ABI library loads and exports every entry point include/lddmm.h
declares (no compute call: these run without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "lddmm.h")
LIB = os.path.join(ROOT, "lddmm_amd", "liblddmm_amd.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lddmm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_reference_boundary():
    names = declared()
    for n in ("lddmm_run_2d", "lddmm_run_3d", "lddmm_run_dev", "lddmm_objgrd", "lddmm_multiply_kernel",
              "lddmm_multiply_kernel_cross", "lddmm_dq_kernel", "lddmm_varifold", "lddmm_get_elm_ifo",
              "lddmm_compute_node_position"):
        assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_exports_are_plain_c():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l and "lddmm_" in l}
    assert set(declared()) <= exported


def test_header_compiles_as_c():
    r = subprocess.run(["gcc", "-std=c99", "-fsyntax-only", "-x", "c", HEADER], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_get_elm_ifo_host_only():
    """lddmm_get_elm_ifo is host code: callable without a GPU."""
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    import numpy as np
    import oracle
    from lddmm_amd import geometry as G
    import lddmm_amd
    q, tri = G.icosphere(2)
    ifo = lddmm_amd.getElmIfo(tri, q.shape[0], one_based=False)
    assert np.array_equal(ifo, oracle.get_elm_ifo(tri, q.shape[0]))
    q2, seg = G.circle_curve(50)
    assert np.array_equal(lddmm_amd.getElmIfo(seg + 1, 50), oracle.get_elm_ifo(seg, 50))


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_batch_start_entries_check_arguments_on_the_host():
    """lddmm_batch_objective_scan and lddmm_batch_run_scaled reject null
    arguments with an error status and message before any device work (so
    this runs without a GPU)."""
    L = ctypes.CDLL(LIB)
    L.lddmm_last_error.restype = ctypes.c_char_p
    v = ctypes.c_double(1.0)
    assert L.lddmm_batch_objective_scan(None, ctypes.byref(v), ctypes.byref(v), 1, ctypes.byref(v)) != 0
    assert b"null argument" in L.lddmm_last_error()
    assert L.lddmm_batch_run_scaled(None, ctypes.byref(v), ctypes.byref(v), ctypes.byref(v), 1, ctypes.c_double(0),
                                    ctypes.c_double(0), ctypes.c_double(0.5), 5, None, None, None, None) != 0
    assert b"null argument" in L.lddmm_last_error()


def test_option_structs_match_ctypes(tmp_path):
    """lddmm_options / lddmm_result offsets in C equal the ctypes mirror's."""
    import lddmm_amd
    fields = {"lddmm_options": lddmm_amd.Options, "lddmm_result": lddmm_amd.Result}
    src = ['#include "lddmm.h"', "#include <stdio.h>", "#include <stddef.h>", "int main(void){"]
    for cname, py in fields.items():
        src.append('printf("%%zu\\n", sizeof(%s));' % cname)
        for f, _ in py._fields_:
            src.append('printf("%%zu\\n", offsetof(%s, %s));' % (cname, f))
    src.append("return 0;}")
    c = tmp_path / "t.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "t"
    r = subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    want = []
    for py in fields.values():
        want.append(ctypes.sizeof(py))
        want += [getattr(py, f).offset for f, _ in py._fields_]
    assert got == want
