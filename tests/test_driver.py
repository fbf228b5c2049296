"""The Fortran driver grhd2_hip (grhd_amd/fortran/grhd2_hip.f90).

CPU: ``--check`` validates a .par file through grhd_validate without touching
a device, like the reference's --check (grhd2.f90:42-55, parameters.f90).
GPU: the driver runs the Sod C1 problem to t = 0.4 and its final conserved
state equals the reference's (tests/golden/sod_c1_final.npz).
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
DRIVER = ROOT / "grhd_amd" / "bin" / "grhd2_hip"
GOLDEN = Path(__file__).resolve().parent / "golden"

pytestmark = pytest.mark.skipif(not DRIVER.exists(), reason="driver not built (build())")

SOD = """problem = sod
nx = 400
reconstruction = weno5
riemann_solver = hllc
shock_sensor = false
"""


def run(args, cwd, timeout=60):
    return subprocess.run([str(DRIVER), *args], cwd=cwd, capture_output=True, text=True,
                          timeout=timeout)


def test_check_valid(tmp_path):
    par = tmp_path / "sod.par"
    par.write_text(SOD)
    r = run(["--check", str(par)], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "parameter file valid" in r.stdout
    assert "grid nx ny nz: 400 1 1" in r.stdout


@pytest.mark.parametrize("bad, message", [
    ("reconstruction = ppm\n", "reconstruction"),
    ("nx = 0\n", "nx"),
    ("metric = kerr_schild\ngeometry = cartesian\n", "metric"),
    ("spin = 1.5\nmetric = kerr_schild\ngeometry = spheroidal\nr_min = 2\nr_max = 10\n", "spin"),
    ("unknown_key = 3\n", "unknown_key"),
    ("vtk_mapping = fancy\n", "vtk_mapping"),
])
def test_check_rejects(tmp_path, bad, message):
    par = tmp_path / "bad.par"
    par.write_text(SOD + bad)
    r = run(["--check", str(par)], tmp_path)
    assert r.returncode != 0
    assert "ERROR" in r.stdout + r.stderr
    assert message in (r.stdout + r.stderr).lower() or message in r.stdout + r.stderr


def test_usage(tmp_path):
    r = run([], tmp_path)
    assert r.returncode != 0 and "usage" in r.stdout


@pytest.mark.gpu
def test_driver_sod_c1_matches_reference(tmp_path):
    from tests.cases import CASES
    par = tmp_path / "sod.par"
    par.write_text(CASES["sod_c1"] + f"output_folder = {tmp_path}/out\noutput_prefix = sod\n")
    print("driver: sod_c1 to t = 0.4", flush=True)
    r = run([str(par)], tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    g = np.load(GOLDEN / "sod_c1_final.npz")
    assert f"time steps: {int(g['n_steps'])}" in r.stdout
    u = np.fromfile(tmp_path / "out" / "sod_final_u.bin", dtype=np.float64).reshape(g["u"].shape)
    ref = g["u"]
    for v in range(5):
        den = np.abs(ref[..., v]).sum()
        if den > 0:
            assert np.abs(u[..., v] - ref[..., v]).sum() / den <= 1e-10, v


@pytest.mark.gpu
def test_driver_perturbation_and_monitor_match_python(tmp_path):
    """Seeded density noise injected by the driver once t >= perturbation_time
    = 0.2 (grhd2.f90:186-227), then the run to t = 1: the final state equals
    the same sequence through the Python front-end bit for bit."""
    from grhd_amd.params import config_from_par
    from grhd_amd.solver import Solver
    from tests.cases import CASES
    text = CASES["michel_ef_log"].replace("final_time = 500", "") + "final_time = 1.0\n"
    par = tmp_path / "michel.par"
    par.write_text(text + "perturbation = density_noise\nperturbation_time = 0.2\n"
                   "perturbation_amplitude = 0.001\nperturbation_seed = 77\nmass_monitor_stride = 50\n"
                   f"output_folder = {tmp_path}/out\noutput_prefix = m\n")
    print("driver: michel with perturbation", flush=True)
    r = run([str(par)], tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "perturbation injected" in r.stdout and "total mass" in r.stdout
    cfg = config_from_par(text)
    with Solver(cfg) as S:
        S.initial_conditions()
        # the reference loop checks t >= perturbation_time before each step
        # and does not clamp to it: count the steps of dt that reach 0.2
        dt, t, n = S.compute_dt(), 0.0, 0
        while t < 0.2:
            t, n = t + dt, n + 1
        S.advance(t_final=1.0, max_steps=n)
        S.perturb(2, 1e-3, 77)
        S.advance(t_final=1.0)
        ref = S.get_conserved()
    u = np.fromfile(tmp_path / "out" / "m_final_u.bin", dtype=np.float64).reshape(ref.shape)
    assert np.array_equal(u, ref)


@pytest.mark.gpu
def test_driver_observers_match_reference(tmp_path):
    """ppi_diagnostics and extract_mdot in the driver: the records it appends
    every diagnostic_stride / extraction_stride steps and at the final time
    (grhd2.f90:139-145, 266-280, 344-354) follow the reference's formats, and
    the first ones (the initial state) equal the reference's own
    calc_global_diagnostics, calc_ppi_modes and calc_m_dot of its initial
    state."""
    import oracle
    from grhd_amd.params import config_from_par
    from tests.cases import CASES
    if not oracle.reference_available():
        pytest.skip("reference build absent")
    text = CASES["fm3d_ef"]
    par = tmp_path / "obs.par"
    par.write_text(text + "final_time = 0.05\nppi_diagnostics = true\nextract_mdot = true\n"
                   "diagnostic_stride = 2\nextraction_stride = 2\n"
                   f"output_folder = {tmp_path}/out\noutput_prefix = o\n")
    print("driver: observers on fm3d_ef", flush=True)
    r = run([str(par)], tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    steps = int(r.stdout.split("time steps: ")[1].split()[0])
    d = tmp_path / "out" / "weno5_hllc"
    glob = [np.array(l.split(), float) for l in (d / "global_diagnostics.dat").read_text().splitlines()]
    modes = [np.array(l.split(), float) for l in (d / "ppi_modes.dat").read_text().splitlines()]
    mdot = [np.array(l.split(), float) for l in (d / "m_dot.dat").read_text().splitlines()]
    # the initial sample, one every 2 steps, and one at the final time when the
    # stride missed it (grhd2.f90:344-354)
    n_expected = 1 + steps // 2 + (steps % 2 != 0)
    assert len(glob) == len(modes) == n_expected and len(mdot) == n_expected
    assert all(len(g) == 11 for g in glob) and all(len(m) == 18 for m in modes)
    R = oracle.Reference(config_from_par(text))
    try:
        ref = R.observers(str(tmp_path / "ref"), "weno5_hllc")
    finally:
        R.close()
    assert glob[0][0] == 0.0 and ref["global_diagnostics"][0] == 0.0
    # the two initial states agree to 1e-13 (initial-data parity), not bit for bit
    assert np.allclose(glob[0][1:9], ref["global_diagnostics"][1:9], rtol=1e-12)
    assert np.array_equal(glob[0][9:], ref["global_diagnostics"][9:])
    assert np.allclose(modes[0][1:2], ref["ppi_modes"][1:2], rtol=1e-12)
    assert mdot[0][1] == pytest.approx(ref["m_dot"][1], rel=1e-6)
