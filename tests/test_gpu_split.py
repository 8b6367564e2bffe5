"""LUMA's per-stage boundary: the step of one grid in the halves of LBM_multi_opt's two sweeps
(lumagpu_stream_bc_macro / lumagpu_force_collide_swap, GridObj_ops_lbm_optimised.cpp:53-172)
with ObjectManager's IB stages between them (ibm_interpolate, ibm_computeForce, ibm_spread,
ibm_updateMacroscopic, ObjectManager_ops_ibm.cpp:498-766, or ibm_apply, :33-54), stepped by the
caller, against lumagpu_step bit for bit; sweep 1's macroscopic values against the oracle's."""
import numpy as np
import pytest

from luma_amd import setups
from luma_amd._abi import ESTATE, LumaGpuError
from luma_amd.grid import GpuGrid
from oracle.oracle import OracleIBM

from test_gpu_ibm import gpu_grid, oracle_grid

pytestmark = pytest.mark.gpu


def split_step(g, subcycle=0, stages=True):
    g.stream_bc_macro(subcycle)
    if stages:
        g.ibm_interpolate()
        g.ibm_compute_force()
        g.ibm_spread()
        g.ibm_update_macro()
    else:
        g.ibm_apply()
    g.force_collide_swap()


def same_state(a, b, tag):
    assert np.array_equal(a.download_f(), b.download_f()), tag
    (ra, ua), (rb, ub) = a.download_macro(), b.download_macro()
    assert np.array_equal(ra, rb) and np.array_equal(ua, ub), tag


def same_markers(a, b, body=0):
    ma, mb = a.ibm_markers(body), b.ibm_markers(body)
    for key in ("pos", "vel", "force", "rho", "mom", "eps", "ds"):
        assert np.array_equal(ma[key], mb[key]), key


@pytest.mark.parametrize("arrangement", ["channel2d", "inflow3d"])
@pytest.mark.parametrize("stages", [True, False])
def test_split_step_with_ib_stages_equals_lumagpu_step(arrangement, stages):
    if arrangement == "channel2d":
        c = setups.channel_2d(64, 32, 8.0, 0.04, 80.0)
        pos = setups.circle_markers((2.05, 2.03, 0.0), 0.5, c.info["dh"])
    else:
        c = setups.uniform_inflow_3d(32, 16, 16, 0.03, 0.62)
        pos = setups.sphere_markers((10.3, 8.1, 7.9), 3.0, 1.0)
    print(f"[split] {arrangement}, stages={stages}, {len(pos)} markers", flush=True)
    with gpu_grid(c) as a, gpu_grid(c) as b:
        for g in (a, b):
            g.ibm_add_body(pos)
            g.ibm_initialise()
        for _ in range(25):
            split_step(a, 0, stages)
            b.LBM_multi_opt(0)
        same_state(a, b, arrangement)
        same_markers(a, b)


def test_split_step_flexible_filament_through_ibm_apply():
    c = setups.channel_2d(64, 32, 8.0, 0.04, 100.0)
    dh = c.info["dh"]
    density, E = setups.filament_material(1.0, 0.05, dh, mass_ratio=10.0, bending=1e-2)
    start = (1.5, 2.0 + 0.013, 0.0)
    print("[split] flexible filament", flush=True)
    with gpu_grid(c) as a, gpu_grid(c) as b:
        for g in (a, b):
            g.ibm_add_filament(start, 1.0, 0.05, 8.0, 8, True, density, E)
            g.ibm_initialise()
        for _ in range(12):
            split_step(a, 0, stages=False)
            b.LBM_multi_opt(0)
            assert a.ibm_subiterations() == b.ibm_subiterations()
        same_state(a, b, "filament")
        same_markers(a, b)


def test_split_steps_of_a_refined_level():
    # the caller drives LBM_multi_opt's recursion itself: the child's two sub-steps in halves with
    # the IB stages, then the parent's halves
    parent = setups.channel_2d(64, 32, 8.0, 0.04, 80.0)
    child = setups.add_refined_region(parent, (8, 31, 6, 25), parent.info["dh"], parent.info["dt"])
    pos = setups.circle_markers((2.5 + 0.01, 2.0 + 0.13, 0.0), 0.5, child.info["dh"])
    print(f"[split] refined level, child {child.N}x{child.M}", flush=True)
    with gpu_grid(parent) as pa, gpu_grid(child) as ca, gpu_grid(parent) as pb, gpu_grid(child) as cb:
        pa.add_child(ca)
        pb.add_child(cb)
        for g in (ca, cb):
            g.ibm_add_body(pos)
            g.ibm_initialise()
        for _ in range(10):
            split_step(ca, 0)
            split_step(ca, 1)
            split_step(pa, 0)
            pb.LBM_multi_opt(0)
        same_state(pa, pb, "parent")
        same_state(ca, cb, "child")
        same_markers(ca, cb)


def test_sweep1_macro_values_match_the_oracle():
    # after lumagpu_stream_bc_macro the macro arrays hold sweep 1's rho, u (with the IB reset's
    # zero force) at fluid and regularised velocity-face sites, as the oracle's arrays between
    # _LBM_resetForces + sweep 1 and the IB step
    c = setups.channel_2d(48, 24, 8.0, 0.04, 80.0)
    pos = setups.circle_markers((1.55, 1.53, 0.0), 0.4, c.info["dh"])
    og = oracle_grid(c)
    ib = OracleIBM(og, c.info["dh"])
    ib.add_body(pos)
    ib.initialise()
    with gpu_grid(c) as g:
        g.ibm_add_body(pos)
        g.ibm_initialise()
        for _ in range(17):
            og.step()
            g.LBM_multi_opt()
        og.lib.ora_reset_forces(og.h)
        og.lib.ora_sweep1(og.h, 0)
        g.stream_bc_macro(0)
        rg, ug = g.download_macro()
        ro, uo = og.get_macro()
        m = np.isin(c.lattyp, [setups.FLUID, setups.VELOCITY])   # the inlet sites are faces here
        assert np.array_equal(rg[m], ro[m]) and np.array_equal(ug[m], uo[m])
        g.ibm_apply()
        g.force_collide_swap()


def test_reset_forces_zeroes_the_body_force():
    # _LBM_resetForces on a grid carrying a body force: the next step equals an unforced step
    c = setups.body_force_channel(24, 12, dims=2, tau=0.8)
    F = np.zeros((c.N, c.M, c.K, 2))
    F[..., 0] = 1e-5
    with GpuGrid.from_case(c, "fp64", store_macro=True) as a, GpuGrid.from_case(c, "fp64", store_macro=True) as b:
        a.set_force(F)
        a.run(3)
        b.set_force(F)
        b.run(3)
        a.reset_forces()
        b.set_force(np.zeros_like(F))
        a.run(5)
        b.run(5)
        same_state(a, b, "reset")


def test_split_calls_out_of_order_are_refused():
    c = setups.channel_2d(32, 16, 8.0, 0.04, 80.0)
    with gpu_grid(c) as g:
        g.ibm_add_body(setups.circle_markers((1.05, 1.03, 0.0), 0.3, c.info["dh"]))
        g.ibm_initialise()
        for fn in (g.ibm_interpolate, g.ibm_spread, g.ibm_apply, g.force_collide_swap):
            with pytest.raises(LumaGpuError) as e:
                fn()
            assert e.value.code == ESTATE
        g.stream_bc_macro(0)
        for fn in (lambda: g.LBM_multi_opt(0), lambda: g.run(2), lambda: g.stream_bc_macro(0)):
            with pytest.raises(LumaGpuError) as e:
                fn()
            assert e.value.code == ESTATE
        g.ibm_apply()
        g.force_collide_swap()
        g.LBM_multi_opt(0)
