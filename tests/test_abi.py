"""The C-ABI library loads and exports every symbol include/bit_amd.h declares;
no compute call is made here (no GPU needed)."""
import ctypes
import os
import re
import subprocess

import bit_amd
from conftest import ROOT


def _declared():
    txt = open(os.path.join(ROOT, "include", "bit_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bit_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_exported():
    lib = ctypes.CDLL(bit_amd.LIB_PATH)
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(bit_amd.EXPORTS) == names


def test_library_has_gfx950_code():
    out = subprocess.run(["/opt/rocm/bin/roc-obj-ls", bit_amd.LIB_PATH], capture_output=True,
                         text=True)
    if out.returncode != 0:
        return  # tool unavailable
    assert "gfx950" in out.stdout


def test_context_without_device_fails_cleanly():
    if os.environ.get("HIP_VISIBLE_DEVICES", None) is None:
        try:
            c = bit_amd.Context(0)
        except bit_amd.BitError as e:
            assert e.code in (bit_amd.BIT_E_HIP, bit_amd.BIT_E_INVALID)
            assert bit_amd.lib.bit_last_error()
            return
        c.close()  # a device is present: nothing to check here


def test_params_default_follow_problem_h():
    p = bit_amd.Params()
    bit_amd.lib.bit_params_default(ctypes.byref(p))
    assert p.acc == 0.5 and p.kappa == 0.0 and p.precision == bit_amd.BIT_FP64


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of bit_params / bit_stats have the header's size and
    field offsets (a C program compiled against include/bit_amd.h)."""
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stddef.h>\n#include <stdio.h>\n#include "bit_amd.h"\n'
        "int main(void) {\n"
        '  printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(bit_params), offsetof(bit_params, acc),\n'
        "         offsetof(bit_params, precision), offsetof(bit_params, kernel_order),\n"
        "         offsetof(bit_params, root), sizeof(bit_stats));\n"
        "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    P = bit_amd.Params
    assert got == [ctypes.sizeof(P), P.acc.offset, P.precision.offset, P.kernel_order.offset,
                   P.root.offset, ctypes.sizeof(bit_amd.Stats)]
    p = bit_amd.Params()
    p.root = 7
    bit_amd.lib.bit_params_default(ctypes.byref(p))
    assert p.kernel_order == 0 and p.mac_sub == 0.0 and p.root == 0


def test_oracle_params_layout_matches_header(tmp_path):
    from oracle.pyoracle import OraParams
    src = tmp_path / "ora.c"
    src.write_text(
        '#include <stddef.h>\n#include <stdio.h>\n#include "bit_oracle.h"\n'
        "int main(void) {\n"
        '  printf("%zu %zu %zu\\n", sizeof(ora_params), offsetof(ora_params, kernel_order),\n'
        "         offsetof(ora_params, root));\n"
        "  return 0;\n}\n")
    exe = tmp_path / "ora"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "oracle"), str(src), "-o", str(exe)],
                   check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    assert got == [ctypes.sizeof(OraParams), OraParams.kernel_order.offset,
                   OraParams.root.offset]
