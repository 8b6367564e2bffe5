"""The C++ host side (include/lumagpu_facade.hpp: GridObj, ObjectManager with LUMA's member
names) driving the GPU step, against the FP64 oracle bit for bit: a channel with a
parabolic inlet, outlet and an IB cylinder read from a geometry.config, run as
LBM_multi_run + LBM_multi_opt, then written as a restart file and read back."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from luma_amd import setups
from oracle.oracle import OracleIBM

from test_gpu_ibm import oracle_grid

pytestmark = pytest.mark.gpu

DRIVER = Path(__file__).resolve().parent / "cpp" / "_build" / "facade_run"


def write_case(case, path):
    D = case.dims
    hdr = np.array([D, case.Q, case.N, case.M, case.K, 0, int(case.regularised), int(case.forced),
                    int(case.fz_compat), case.info.get("level", 0)], np.int32)
    o = case.info.get("origin", (0.0, 0.0, 0.0))
    reals = np.array([case.omega, case.rho_in, 1.0, *case.edges, *case.wall_thk, case.info["dh"], *o,
                      case.info.get("dt", 1.0), case.info.get("dm", 1.0), 0.0, 0.0, 0.0], np.float64)
    with open(path, "wb") as fh:
        fh.write(hdr.tobytes())
        fh.write(reals.tobytes())
        fh.write(np.ascontiguousarray(case.lattyp, np.int32).tobytes())
        for a in (case.xpos, case.ypos, case.zpos, case.ux_in, case.uy_in, case.uz_in, case.rho, case.u):
            fh.write(np.ascontiguousarray(a, np.float64).tobytes())


def test_cpp_facade_channel_with_geometry_config_fp64_bitwise(tmp_path):
    assert DRIVER.exists(), "build the driver with __graft_entry__.build()"
    c = setups.channel_2d(80, 40, 10.0, 0.04, 100.0)
    dh = c.info["dh"]
    write_case(c, tmp_path / "case.bin")
    (tmp_path / "geometry.config").write_text("# one body\nCIRCLE_SPHERE IBM 0 0 2.05 2.03 0.0 0.5 RIGID\n")
    steps = 30
    print(f"[facade] channel {c.N}x{c.M} + cylinder, {steps} steps", flush=True)
    r = subprocess.run([str(DRIVER), str(tmp_path / "case.bin"), str(tmp_path / "geometry.config"), str(steps),
                        str(tmp_path / "out.bin"), str(tmp_path / "restart.out")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert f"t = {steps}" in r.stdout
    out = np.fromfile(tmp_path / "out.bin", np.float64).reshape(2, c.N, c.M, c.K, c.Q)

    og = oracle_grid(c)
    ib = OracleIBM(og, dh)
    ib.add_body(setups.circle_markers((2.05, 2.03, 0.0), 0.5, dh))
    ib.initialise()
    og.run(steps)
    ref = og.get_f()
    m = c.lattyp != setups.SOLID
    assert np.array_equal(out[0][m], ref[m])
    # the restart round trip reproduces the state to rounding
    assert np.abs(out[1][m] - ref[m]).max() <= 1e-13 * np.abs(ref[m]).max()


def test_cpp_facade_split_step_with_ib_stages_fp64_bitwise(tmp_path):
    # the same channel, the second half of the steps driven as LBM_stream_bc_macro,
    # ObjectManager::ibm_interpolate / ibm_computeForce / ibm_spread / ibm_updateMacroscopic and
    # LBM_force_collide_swap: still the oracle's LBM_multi_opt bit for bit
    c = setups.channel_2d(80, 40, 10.0, 0.04, 100.0)
    dh = c.info["dh"]
    write_case(c, tmp_path / "case.bin")
    (tmp_path / "geometry.config").write_text("CIRCLE_SPHERE IBM 0 0 2.05 2.03 0.0 0.5 RIGID\n")
    steps = 24
    r = subprocess.run([str(DRIVER), str(tmp_path / "case.bin"), str(tmp_path / "geometry.config"), str(steps),
                        str(tmp_path / "out.bin"), "-", "split"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert f"t = {steps}" in r.stdout
    out = np.fromfile(tmp_path / "out.bin", np.float64).reshape(c.N, c.M, c.K, c.Q)
    og = oracle_grid(c)
    ib = OracleIBM(og, dh)
    ib.add_body(setups.circle_markers((2.05, 2.03, 0.0), 0.5, dh))
    ib.initialise()
    og.run(steps)
    m = c.lattyp != setups.SOLID
    assert np.array_equal(out[m], og.get_f()[m])


def test_cpp_facade_reports_errors_through_the_handler(tmp_path):
    c = setups.channel_2d(40, 20, 10.0, 0.04, 100.0)
    write_case(c, tmp_path / "case.bin")
    (tmp_path / "geometry.config").write_text("CIRCLE_SPHERE IBM 0 0 1.0 1.0 0.0 0.3 FLEXIBLE\n")
    r = subprocess.run([str(DRIVER), str(tmp_path / "case.bin"), str(tmp_path / "geometry.config"), "4",
                        str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "Circle/sphere cannot be flexible" in r.stderr
