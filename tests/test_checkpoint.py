"""GRHDCP_V4 checkpoints (output.f90:474-609) between the reference and the
Fortran driver, both ways.

CPU: the reference build writes a checkpoint; the driver's --check reads its
metadata and state through a restart .par (restart = true, restart_file),
and rejects keys that belong to the checkpoint (is_restart_control,
parameters.f90:716-731).
GPU: (a) the reference runs 20 steps and writes a checkpoint; the driver
restarts from it and runs to the reference's time after 15 more steps: the
final state equals the reference's continued run.  (b) the driver writes a
checkpoint during its own run; the reference restarts from it through its
restart branch (grhd2.f90:63-117) and runs to the same final time: the final
states agree.  Tolerance: L1 <= 1e-10 per variable, as for the driver's Sod
run (the one-step parity is 1e-12, and these runs take tens of steps).
"""
import shutil
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

import oracle
from grhd_amd.params import config_from_par
from tests.cases import CASES

ROOT = Path(__file__).resolve().parents[1]
DRIVER = ROOT / "grhd_amd" / "bin" / "grhd2_hip"

pytestmark = pytest.mark.skipif(not DRIVER.exists() or not oracle.reference_available(),
                                reason="driver or reference build absent (build())")

CASE = "michel_ef_log"


def run(args, cwd, timeout=120):
    return subprocess.run([str(DRIVER), *args], cwd=cwd, capture_output=True, text=True, timeout=timeout)


def _reference_checkpoint(tmp_path, nsteps):
    cfg = config_from_par(CASES[CASE])
    R = oracle.Reference(cfg)
    R.set_time(0.0, 1.0e9)
    t, _, n = R.step(nsteps)
    path = R.save_checkpoint(str(tmp_path / "ref_ck"))
    return R, path, t, n


def _l1(a, b):
    out = []
    for v in range(5):
        den = np.abs(b[..., v]).sum()
        out.append(np.abs(a[..., v] - b[..., v]).sum() / den if den > 0 else np.abs(a[..., v]).sum())
    return np.array(out)


def test_check_reads_reference_checkpoint(tmp_path):
    R, path, t, n = _reference_checkpoint(tmp_path, 20)
    R.close()
    par = tmp_path / "restart.par"
    par.write_text(f"restart = true\nrestart_file = {path}\nfinal_time = 2.0\n")
    r = run(["--check", str(par)], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "problem: michel" in r.stdout
    assert "grid nx ny nz: 128 1 1" in r.stdout
    assert f"step {n}" in r.stdout and "checkpoint valid" in r.stdout
    # the Michel preset extracts the accretion rate (initialization.f90:656-682)
    assert "observers: extract_mdot T  ppi_diagnostics F  diagnostic_stride 100" in r.stdout
    line = next(l for l in r.stdout.splitlines() if l.startswith("checkpoint at t ="))
    assert abs(float(line.split("=")[1].split()[0]) - t) <= 1e-13 * t


@pytest.mark.parametrize("bad", ["nx = 64\n", "reconstruction = mp5\n", "michel_critical_radius = 7\n"])
def test_restart_rejects_checkpoint_keys(tmp_path, bad):
    R, path, _, _ = _reference_checkpoint(tmp_path, 2)
    R.close()
    par = tmp_path / "restart.par"
    par.write_text(f"restart = true\nrestart_file = {path}\n" + bad)
    r = run(["--check", str(par)], tmp_path)
    assert r.returncode != 0
    assert "cannot change on restart" in r.stdout + r.stderr


def test_restart_requires_restart_file(tmp_path):
    par = tmp_path / "restart.par"
    par.write_text("restart = true\n")
    r = run(["--check", str(par)], tmp_path)
    assert r.returncode != 0 and "restart_file" in r.stdout + r.stderr


@pytest.mark.gpu
def test_driver_restarts_from_reference_checkpoint(tmp_path):
    R, path, t20, n20 = _reference_checkpoint(tmp_path, 20)
    t35, _, n35 = R.step(15)
    u_ref = R.get_cons()
    R.close()
    par = tmp_path / "restart.par"
    par.write_text(f"restart = true\nrestart_file = {path}\nfinal_time = {float(t35)!r}\n"
                   f"output_folder = {tmp_path}/out\noutput_prefix = r\n")
    print("driver restart from the reference at step", n20, flush=True)
    r = run([str(par)], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"time steps: {n35}" in r.stdout, r.stdout
    u = np.fromfile(tmp_path / "out" / "r_final_u.bin", dtype=np.float64).reshape(u_ref.shape)
    l1 = _l1(u, u_ref)
    assert np.all(l1 <= 1e-10), l1


@pytest.mark.gpu
def test_reference_restarts_from_driver_checkpoint(tmp_path):
    text = CASES[CASE].replace("final_time = 500", "")
    par = tmp_path / "run.par"
    par.write_text(text + "final_time = 0.3\ncheckpoint_interval = 0.1\n"
                   f"output_folder = {tmp_path}/out\noutput_prefix = d\n")
    print("driver run writing checkpoints", flush=True)
    r = run([str(par)], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    cks = sorted((tmp_path / "out").glob("checkpoint_weno5_hlle_step_*.rst"))
    assert len(cks) >= 2, r.stdout
    # the reference holds restart_file in 100 characters (variables.f90:267):
    # hand it a copy under a short path
    short = Path(tempfile.mkdtemp(prefix="ck", dir="/tmp"))
    ck = short / "c.rst"
    shutil.copyfile(cks[0], ck)
    rpar = tmp_path / "ref_restart.par"
    rpar.write_text(f"restart = true\nrestart_file = {ck}\nfinal_time = 0.3\n")
    try:
        R = oracle.Reference(par_text=rpar.read_text(), restart=True)
    finally:
        shutil.rmtree(short, ignore_errors=True)
    try:
        dims = R.dims
        assert dims[:4] == [128, 1, 1, 3] and dims[4] == 7 and dims[5] == 5 and dims[6] == 1
        t0, _, n0 = R.step(0)
        assert n0 == int(cks[0].stem.split("_")[-1]) and 0.1 <= t0
        t1, _, n1 = R.step(10 ** 6)
        u_ref = R.get_cons()
    finally:
        R.close()
    assert f"time steps: {n1}" in r.stdout, (n1, r.stdout)
    u_drv = np.fromfile(tmp_path / "out" / "d_final_u.bin", dtype=np.float64).reshape(u_ref.shape)
    l1 = _l1(u_drv, u_ref)
    assert np.all(l1 <= 1e-10), l1


def test_reference_wrapper_rejects_long_restart_file():
    """The reference's parser ends the process on a restart_file longer than
    its 100 characters; the wrapper refuses such a .par before calling it."""
    with pytest.raises(ValueError):
        oracle.Reference(par_text="restart = true\nrestart_file = /" + "a" * 120 + ".rst\n", restart=True)
