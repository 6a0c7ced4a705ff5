"""The ctypes mirrors of the C ABI's structs (structured_amd/abi.py) have the
size and field offsets the C compiler gives the declarations in
include/structured_gpu.h: a C program compiled here from the header prints
sizeof and offsetof of every field, and the mirrors must agree field by field
(a field added on one side only, or in another order, fails)."""
import ctypes as C
import os
import subprocess

import pytest

from structured_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PAIRS = [("sg_freestream", abi.Freestream), ("sg_bc", abi.Bc), ("sg_opts", abi.Opts),
         ("sg_adjoint_opts", abi.AdjointOpts), ("sg_adjoint_report", abi.AdjointReport),
         ("sg_newton_opts", abi.NewtonOpts), ("sg_solve_report", abi.SolveReport),
         ("sg_explicit_opts", abi.ExplicitOpts), ("sg_inversion_opts", abi.InversionOpts),
         ("sg_inversion_report", abi.InversionReport)]


def _c_layout(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "structured_gpu.h"', "int main(void) {"]
    for cname, cls in PAIRS:
        lines.append('  printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in cls._fields_:
            lines.append('  printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return dict((k, int(v)) for k, v in (l.rsplit(" ", 1) for l in out.splitlines()))


def test_ctypes_mirrors_match_the_header(tmp_path):
    lay = _c_layout(tmp_path)
    for cname, cls in PAIRS:
        assert lay["%s size" % cname] == C.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert lay["%s.%s" % (cname, f)] == getattr(cls, f).offset, (cname, f)


@pytest.mark.parametrize("cname,cls", PAIRS[:1])
def test_a_misordered_mirror_is_caught(tmp_path, cname, cls):
    """The check is not vacuous: swapping two fields of a mirror changes an
    offset the header disagrees with."""
    lay = _c_layout(tmp_path)
    f = list(cls._fields_)
    f[0], f[1] = f[1], f[0]
    Bad = type("Bad", (C.Structure,), {"_fields_": f})
    assert any(lay["%s.%s" % (cname, n)] != getattr(Bad, n).offset for n, _ in f)
