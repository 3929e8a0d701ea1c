"""The C ABI: library loads, exports every declared symbol, and the ctypes
mirror of desalm_params has the C layout.  No compute call (no GPU here)."""
import ctypes
import os
import re
import subprocess
import tempfile

from desalm import lib as desalm_lib
from desalm.params import Params

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
HEADER = os.path.join(ROOT, "include", "desalm.h")


def declared():
    txt = open(HEADER).read()
    return re.findall(r"^(?:int|int64_t|double|const char\*)\s+\**(desalm_\w+)\(", txt, re.M)


def test_exports_every_declared_symbol():
    L = desalm_lib.lib()
    names = declared()
    assert len(names) >= 17
    for n in names:
        assert hasattr(L, n), n


def test_params_layout_matches_c():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "desalm.h"
int main(void){
 printf("%zu %zu %zu %zu %zu %zu\n", sizeof(desalm_params), offsetof(desalm_params, er_min),
        offsetof(desalm_params, ep), offsetof(desalm_params, btype), offsetof(desalm_params, ifix),
        offsetof(desalm_params, precision));
 return 0;}
'''
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "p.c")
        open(c, "w").write(src)
        exe = os.path.join(td, "p")
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        vals = [int(v) for v in subprocess.check_output([exe]).split()]
    assert vals[0] == ctypes.sizeof(Params)
    assert vals[1] == Params.er_min.offset
    assert vals[2] == Params.ep.offset
    assert vals[3] == Params.btype.offset
    assert vals[4] == Params.ifix.offset
    assert vals[5] == Params.precision.offset


def test_init_without_device_fails_loudly():
    import pytest
    p = Params()
    p.ng = 1
    p.img = p.jmg = p.kmg = 9
    p.ifix = p.jfix = p.kfix = 2
    s = None
    with pytest.raises(desalm_lib.DesalmError):
        s = desalm_lib.Solver(p, device=0)
    assert s is None
