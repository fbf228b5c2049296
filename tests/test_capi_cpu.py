"""The C-ABI library loads on a host without a GPU and exports every function
include/grhd_c.h declares; the ctypes mirror of grhd_config has the header's
layout.  No compute call is made here."""
import ctypes
import subprocess
import sys
from pathlib import Path

import pytest

from grhd_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _lib.declared_symbols()
    assert len(declared) >= 20
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    bound = set(_lib._SIGS)
    assert set(declared) <= bound, set(declared) - bound


def test_config_layout_matches_header(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "grhd_c.h"\n'
                   'int main(){printf("%zu %zu %zu %zu\\n", sizeof(grhd_config),'
                   ' offsetof(grhd_config, cfl), offsetof(grhd_config, fm_polytropic_constant),'
                   ' offsetof(grhd_config, allow_open_theta));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.check_call(["gcc", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)])
    got = [int(v) for v in subprocess.check_output([str(exe)]).split()]
    C = _lib.Config
    assert got == [ctypes.sizeof(C), C.cfl.offset, C.fm_polytropic_constant.offset,
                   C.allow_open_theta.offset]


def test_invalid_config_is_rejected_before_any_device_call():
    c = _lib.Config()
    c.problem = 99
    h = ctypes.c_void_p()
    rc = _lib.lib().grhd_create(ctypes.byref(c), ctypes.byref(h))
    assert rc == 1
    assert b"problem" in _lib.lib().grhd_last_error(None)
