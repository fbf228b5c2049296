"""VTK frames (output.f90:154-471) written by the Fortran driver against the
frames the reference build writes for the same problem.

CPU: the centre lapse and shift the frames carry (grhd_lapse_shift, host
evaluation) equal the reference's alpha_c / beta_c cache bit for bit on every
case of tests/cases.py.
GPU: the driver's frame 0 (the cold start's initial state) and, for Sod, the
frame at the final time are parsed section by section against the reference's
save_vtk_at_time of the same state: every header line, the points (structured
faces, or the full-sphere unstructured grid with its pole wedges), the cell
connectivity and types, the lapse and the shift exactly; the primitives
within round-off (the initial data are computed on the device: rtol 1e-12 of
the field's largest magnitude; the Sod frame after 0.2 time units to L1
1e-10, as for the driver's own Sod parity).
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from grhd_amd.params import config_from_par
from tests.cases import CASES

ROOT = Path(__file__).resolve().parents[1]
DRIVER = ROOT / "grhd_amd" / "bin" / "grhd2_hip"

pytestmark = pytest.mark.skipif(not oracle.reference_available(), reason="reference build absent")


def parse_vtk(path):
    """Sections of a legacy binary VTK file: the text lines in order, and the
    binary blocks by name (POINTS, CELLS, CELL_TYPES, each SCALARS field)."""
    b = Path(path).read_bytes()
    pos = 0
    lines, blocks = [], {}

    def line():
        nonlocal pos
        e = b.index(b"\n", pos)
        s = b[pos:e].decode()
        pos = e + 1
        return s

    def block(dtype, n):
        nonlocal pos
        nbytes = np.dtype(dtype).itemsize * n
        a = np.frombuffer(b[pos:pos + nbytes], dtype=dtype).copy()
        pos += nbytes
        assert b[pos:pos + 1] == b"\n"
        pos += 1
        return a

    for _ in range(3):
        lines.append(line())
    ncell = 0
    while pos < len(b):
        s = line()
        lines.append(s)
        w = s.split()
        if w[0] == "POINTS":
            blocks["POINTS"] = block(">f8", 3 * int(w[1]))
        elif w[0] == "CELLS":
            blocks["CELLS"] = block(">i4", int(w[2]))
        elif w[0] == "CELL_TYPES":
            blocks["CELL_TYPES"] = block(">i4", int(w[1]))
        elif w[0] == "CELL_DATA":
            ncell = int(w[1])
        elif w[0] == "SCALARS":
            lines.append(line())                     # LOOKUP_TABLE default
            blocks[w[1]] = block(">f8", ncell)
    return lines, blocks


def test_lapse_shift_matches_reference_cache():
    from grhd_amd.solver import lapse_shift
    for name, text in CASES.items():
        cfg = config_from_par(text)
        if cfg.nx * cfg.ny * cfg.nz > 300_000:
            continue
        a, b = lapse_shift(cfg)
        R = oracle.Reference(cfg)
        try:
            ra, rb = R.lapse_shift()
        finally:
            R.close()
        assert np.array_equal(a, ra), name
        assert np.array_equal(b, rb), name


PRIMS = ("Density", "Pressure", "Velocity_X", "Velocity_Y", "Velocity_Z")


def _compare(mine, ref, prim_check):
    lm, bm = parse_vtk(mine)
    lr, br = parse_vtk(ref)
    assert lm == lr
    assert sorted(bm) == sorted(br)
    for k in ("POINTS", "CELLS", "CELL_TYPES", "Lapse_alpha", "Shift_beta_r"):
        if k in br:
            assert np.array_equal(bm[k], br[k]), k
    for k in PRIMS:
        prim_check(k, bm[k], br[k])
    return lm


def _run_driver(par, cwd):
    r = subprocess.run([str(DRIVER), str(par)], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _scheme(cfg):
    rec = {1: "godunov", 3: "weno3", 4: "mp5", 5: "weno5"}.get(cfg.reconstruction)
    if rec is None:
        rec = {1: "minmod", 2: "superbee", 3: "mc"}[cfg.tvd_limiter]
    return rec + ("_hllc" if cfg.riemann_solver == 2 else "_hlle")


@pytest.mark.gpu
@pytest.mark.skipif(not DRIVER.exists(), reason="driver not built")
@pytest.mark.parametrize("name,mapping,dataset", [
    ("sod_c1", "physical", "STRUCTURED_GRID"),
    ("khi_weno3_sensor", "physical", "STRUCTURED_GRID"),
    ("michel_ef_log", "physical", "STRUCTURED_GRID"),
    ("fm_sagittal_ef_log", "physical", "STRUCTURED_GRID"),
    ("fm3d_ks_mp5_hllc_sensor", "physical", "UNSTRUCTURED_GRID"),
    ("fm3d_ks_mp5_hllc_sensor", "untwisted", "UNSTRUCTURED_GRID"),
])
def test_initial_frame_matches_reference(tmp_path, name, mapping, dataset):
    import re
    text = re.sub(r"(?m)^final_time\s*=.*$", "", CASES[name])
    extra = f"final_time = 0.0\nvtk_mapping = {mapping}\n"
    cfg = config_from_par(text + extra)
    print("frame 0 of", name, mapping, flush=True)
    par = tmp_path / "run.par"
    par.write_text(text + extra + f"output_folder = {tmp_path}/out\noutput_prefix = f\n")
    _run_driver(par, tmp_path)
    mine = tmp_path / "out" / _scheme(cfg) / f"f_{_scheme(cfg)}_0000.vtk"
    R = oracle.Reference(par_text=text + extra + f"output_folder = {tmp_path}/ref\noutput_prefix = f\n")
    try:
        ref = R.save_vtk(str(tmp_path / "ref"), _scheme(cfg), 0.0)
    finally:
        R.close()

    def close(k, a, b):
        scale = max(np.abs(b).max(), 1e-300)
        assert np.abs(a - b).max() <= 1e-12 * scale, (k, np.abs(a - b).max(), scale)

    lines = _compare(mine, ref, close)
    assert lines[3] == f"DATASET {dataset}"


@pytest.mark.gpu
@pytest.mark.skipif(not DRIVER.exists(), reason="driver not built")
def test_sod_frames_follow_the_save_calendar(tmp_path):
    """save_interval 0.1 to t = 0.2: frames 0, 1 and 2; frame 2 holds the
    state at the final time and matches the reference's."""
    cfg = config_from_par(CASES["sod_c1"])
    par = tmp_path / "run.par"
    par.write_text(CASES["sod_c1"].replace("final_time = 0.4", "") +
                   f"final_time = 0.2\noutput_folder = {tmp_path}/out\noutput_prefix = s\n")
    print("sod frames", flush=True)
    _run_driver(par, tmp_path)
    sch = _scheme(cfg)
    names = sorted(p.name for p in (tmp_path / "out" / sch).glob("*.vtk"))
    assert names == [f"s_{sch}_000{i}.vtk" for i in range(3)], names
    R = oracle.Reference(config_from_par(CASES["sod_c1"].replace("final_time = 0.4", "final_time = 0.2")))
    try:
        R.set_time(0.0, 0.2)
        R.step(10 ** 6)
        ref = R.save_vtk(str(tmp_path / "ref"), sch, 0.2)
    finally:
        R.close()

    def l1(k, a, b):
        den = np.abs(b).sum()
        err = np.abs(a - b).sum() / den if den > 0 else np.abs(a).sum()
        assert err <= 1e-10, (k, err)

    _compare(tmp_path / "out" / sch / f"s_{sch}_0002.vtk", ref, l1)
