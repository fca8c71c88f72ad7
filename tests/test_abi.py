"""The C ABI: the library loads, exports every function include/*.h
declares, and the ctypes mirror has the C layout (no device calls)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from openvort_amd import tangle as T
from openvort_amd.gpu import LIB_PATH, SYMBOLS, load_library

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "openvort_gpu.h")
HOST_HEADER = os.path.join(ROOT, "include", "openvort_host.h")


def declared_functions(header=HEADER):
    src = open(header).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(ov_\w+)\s*\(", src, flags=re.M)))


def test_header_functions_are_exported():
    from openvort_amd.host import HOST_SYMBOLS
    lib = C.CDLL(LIB_PATH)
    decl = declared_functions()
    assert len(decl) >= 12
    host = declared_functions(HOST_HEADER)
    assert len(host) >= 8
    missing = [f for f in decl + host if not hasattr(lib, f)]
    assert not missing, missing
    assert set(decl) == set(SYMBOLS)
    assert set(host) == set(HOST_SYMBOLS)
    headers = sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h"))
    assert headers == ["openvort_gpu.h", "openvort_host.h"]


def test_binding_types_load():
    from openvort_amd import host
    lib = load_library()
    host._lib()
    for f in SYMBOLS + host.HOST_SYMBOLS:
        assert getattr(lib, f).argtypes is not None


def test_params_default_matches_python():
    lib = load_library()
    p = T.Params()
    lib.ov_params_default(C.byref(p))
    q = T.default_params()
    for name, _ in T.Params._fields_:
        a, b = getattr(p, name), getattr(q, name)
        if isinstance(a, T.Flow):
            assert a.type == b.type == T.FLOW_NONE
        else:
            assert a == b, name


LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "openvort_gpu.h"
#include "openvort_host.h"
#define F(T, m) printf(#T " " #m " %zu\n", offsetof(T, m));
int main(void) {
  printf("ov_tangle_state size %zu\n", sizeof(ov_tangle_state));
  printf("ov_params size %zu\n", sizeof(ov_params));
  printf("ov_flow size %zu\n", sizeof(ov_flow));
  F(ov_flow, v) F(ov_flow, a) F(ov_flow, w) F(ov_flow, q)
  F(ov_tangle_state, connections) F(ov_tangle_state, status) F(ov_tangle_state, box)
  F(ov_tangle_state, bimg) F(ov_tangle_state, N) F(ov_tangle_state, total_free)
  F(ov_params, max_curvature_scale) F(ov_params, BH_quadtree) F(ov_params, precision)
  F(ov_params, near_radius) F(ov_params, vn) F(ov_params, vs) F(ov_params, vb)
  printf("ov_config size %zu\n", sizeof(ov_config));
  printf("ov_v_conf size %zu\n", sizeof(ov_v_conf));
  F(ov_config, vn) F(ov_config, box) F(ov_config, images) F(ov_config, init_mode)
  F(ov_config, init_r_KW) F(ov_v_conf, v_params)
  return 0;
}
"""


def test_ctypes_layout_matches_c():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(src, "w").write(LAYOUT_C)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    vals = {}
    for line in out.strip().splitlines():
        a, b, c = line.split()
        vals[(a, b)] = int(c)
    assert vals[("ov_tangle_state", "size")] == C.sizeof(T.TangleState)
    assert vals[("ov_params", "size")] == C.sizeof(T.Params)
    assert vals[("ov_flow", "size")] == C.sizeof(T.Flow)
    from openvort_amd import host as H
    assert vals[("ov_config", "size")] == C.sizeof(H.Config)
    assert vals[("ov_v_conf", "size")] == C.sizeof(H.VConf)
    for (struct, field), off in vals.items():
        if field == "size":
            continue
        cls = {"ov_tangle_state": T.TangleState, "ov_params": T.Params, "ov_flow": T.Flow,
               "ov_config": H.Config, "ov_v_conf": H.VConf}[struct]
        assert getattr(cls, field).offset == off, (struct, field)


def test_oracle_is_not_imported_by_the_product():
    for root, _, files in os.walk(os.path.join(ROOT, "openvort_amd")):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".c", ".cpp")):
                txt = open(os.path.join(root, f)).read()
                assert "oracle" not in txt.replace("oracle/", "").lower() or f == "build.py", f
