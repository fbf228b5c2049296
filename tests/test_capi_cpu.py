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


def test_c_caller_links_and_validates(tmp_path):
    """A C program compiled against include/grhd_c.h and linked with
    libgrhd_hip.so validates a configuration and sizes its footprint without
    a device (the --check path of a C or Fortran caller)."""
    src = tmp_path / "caller.c"
    src.write_text(r'''
#include <stdio.h>
#include <string.h>
#include "grhd_c.h"
int main(void) {
    grhd_config c;
    memset(&c, 0, sizeof c);
    c.problem = 1; c.metric = GRHD_METRIC_MINKOWSKI; c.geometry = GRHD_GEOM_CARTESIAN;
    c.adiabatic_index = 5.0 / 3.0; c.nx = 400; c.ny = 1; c.nz = 1;
    c.r_min = 0.0; c.r_max = 1.0; c.y_min = 0.0; c.y_max = 1.0; c.z_min = 0.0; c.z_max = 1.0;
    c.reconstruction = GRHD_REC_WENO5; c.riemann_solver = GRHD_RS_HLLE; c.tvd_limiter = GRHD_LIM_MC;
    c.cfl = 0.4; c.final_time = 0.4;
    c.sod_interface = 0.5; c.sod_density_left = 10.0; c.sod_density_right = 1.0;
    c.sod_pressure_left = 13.333333333333334; c.sod_pressure_right = 6.666666666666667e-7;
    uint64_t bytes = 0;
    int ok = grhd_validate(&c);
    int fp = grhd_footprint(&c, &bytes);
    c.reconstruction = 9;
    int bad = grhd_validate(&c);
    printf("%d %d %llu %d %s\n", ok, fp, (unsigned long long)bytes, bad, grhd_last_error(NULL));
    return 0;
}
''')
    exe = tmp_path / "caller"
    subprocess.check_call(["gcc", f"-I{ROOT / 'include'}", str(src), "-o", str(exe),
                           f"-L{ROOT / 'grhd_amd'}", "-lgrhd_hip", f"-Wl,-rpath,{ROOT / 'grhd_amd'}"])
    out = subprocess.check_output([str(exe)]).decode().split(maxsplit=4)
    ok, fp, nbytes, bad, msg = int(out[0]), int(out[1]), int(out[2]), int(out[3]), out[4]
    assert ok == 0 and fp == 0 and bad == 1
    # p (407 x 5 doubles) + 2 U (400 x 5) + x faces (401 x 5) + coordinate
    # table (3 x 407 + 3 x 401 + 2 x 7 + 2 x 2 doubles) + the reduction
    # buffer of the CFL, mass and observer partials (4096 doubles) + small buffers
    base = 407 * 40 + 2 * 400 * 40 + 401 * 40 + (3 * 407 + 3 * 401 + 18) * 8 + 4096 * 8
    assert base <= nbytes < base + 1024
    assert "reconstruction" in msg
