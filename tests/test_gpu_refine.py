"""Two-level refinement (explode / coalesce, GridObj_ops_lbm_optimised.cpp:591-663) on the GPU
against the FP64 oracle, bit for bit."""
import numpy as np
import pytest

from luma_amd import setups
from luma_amd.grid import GpuGrid
from oracle.oracle import OracleGrid

pytestmark = pytest.mark.gpu


def oracle_grid(case):
    ci = case.info.get("coarse_lims", (0,) * 6)
    g = OracleGrid(case.dims, case.Q, case.N, case.M, case.K, omega=case.omega, regularised=case.regularised,
                   forced=case.forced, fz_compat=case.fz_compat, rho_in=case.rho_in, edges=case.edges,
                   wall_thk=case.wall_thk, level=case.info.get("level", 0), coarse_lims=ci)
    if case.lattyp is not None:
        g.set_labels(case.lattyp)
    if case.xpos is not None:
        g.set_positions(case.xpos, case.ypos, case.zpos)
    if case.ux_in is not None:
        g.set_inlet(case.ux_in, case.uy_in, case.uz_in)
    g.init_feq(case.rho, case.u)
    return g


def run_pair(parent, child, nsteps, re_ramp=None):
    print(f"[refine] {parent.name} + {child.name} steps={nsteps}", flush=True)
    op, oc = oracle_grid(parent), oracle_grid(child)
    op.add_child(oc)
    with GpuGrid.from_case(parent, "fp64", store_macro=True) as gp, \
            GpuGrid.from_case(child, "fp64", store_macro=True) as gc:
        gp.add_child(gc)
        if re_ramp is not None:
            re, T = re_ramp
            for o, case in ((op, parent), (oc, child)):
                o.set_reynolds_ramp(re, T, case.info["dt"], case.info["dh"])
            gp.set_reynolds_ramp(re, T)
        for _ in range(nsteps):
            op.step()
            gp.LBM_multi_opt()
        assert gp.t == nsteps and gc.t == 2 * nsteps
        out = {"pf": gp.download_f(), "pm": gp.download_macro(), "cf": gc.download_f(), "cm": gc.download_macro()}
    ref = {"pf": op.get_f(), "pm": op.get_macro(), "cf": oc.get_f(), "cm": oc.get_macro()}
    mp = (parent.lattyp != setups.REFINED) & (parent.lattyp != setups.SOLID)
    mc = child.lattyp != setups.SOLID
    for key, m in (("p", mp), ("c", mc)):
        d = np.abs(out[key + "f"][m] - ref[key + "f"][m]).max()
        assert d == 0.0, f"{key}: f differs by {d}"
        for a, b in zip(out[key + "m"], ref[key + "m"]):
            assert np.array_equal(a[mc if key == "c" else (parent.lattyp != setups.REFINED)],
                                  b[mc if key == "c" else (parent.lattyp != setups.REFINED)])


def test_two_level_tgv2d_fp64_bitwise():
    parent = setups.taylor_green((24, 20, 1), u0=0.04, tau=0.6, dims=2)
    child = setups.add_refined_region(parent, (6, 13, 5, 12), 1.0, 1.0)
    setups.velocity_field_on(child, parent)
    run_pair(parent, child, 30)


def test_two_level_tgv3d_fp64_bitwise():
    parent = setups.taylor_green((14, 12, 12), u0=0.04, tau=0.6, dims=3)
    child = setups.add_refined_region(parent, (3, 8, 3, 8, 2, 7), 1.0, 1.0)
    setups.velocity_field_on(child, parent)
    run_pair(parent, child, 12)


def test_two_level_cavity_interior_patch_fp64_bitwise():
    parent = setups.lid_driven_cavity(32, re=1280.0 * 32 / 256)
    child = setups.add_refined_region(parent, (8, 23, 12, 27), parent.info["dh"], parent.info["dt"])
    run_pair(parent, child, 60)


def test_two_level_cavity_reynolds_ramp_fp64_bitwise():
    # L_REYNOLDS_RAMP on both levels: each resets its omega from its own dt and dh at the start
    # of each of its steps (the parent before its child's two sub-steps)
    parent = setups.lid_driven_cavity(32, re=1280.0 * 32 / 256)
    child = setups.add_refined_region(parent, (8, 23, 12, 27), parent.info["dh"], parent.info["dt"])
    run_pair(parent, child, 40, re_ramp=(parent.info["re"], 25 * parent.info["dt"]))


def test_refined_patch_touching_the_lid_fp64_bitwise():
    # The lid's velocity sites pull from refined sites, which LUMA never updates: they stream
    # the populations those sites were initialised with (f = fNew at initialisation,
    # GridObj_init_grids.cpp:333, and neither sweep visits eRefined, :91-95, :125-156)
    parent = setups.lid_driven_cavity(32, re=1280.0 * 32 / 256)
    child = setups.add_refined_region(parent, (10, 21, 22, 31), parent.info["dh"], parent.info["dt"])
    run_pair(parent, child, 60)


@pytest.mark.parametrize("where", ["inlet", "outlet"])
def test_refined_patch_touching_inlet_walls_or_outlet_fp64_bitwise(where):
    # A channel region spanning the full height (both solid walls): from the inlet column on, or
    # up to the column before the ExtrapolateRight outlet, whose sources two cells left are then
    # refined sites (the outlet's side buffer takes their initial populations as well)
    parent = setups.channel_2d(32, 20, 8.0, 0.04, 20.0, outlet=True)
    lims = (0, 13, 0, 19) if where == "inlet" else (14, 30, 0, 19)
    child = setups.add_refined_region(parent, lims, parent.info["dh"], parent.info["dt"])
    setups.velocity_field_on(child, parent)
    if where == "outlet":
        assert (parent.lattyp[29, 1:-1] == setups.REFINED).all()
    run_pair(parent, child, 40)


def test_region_over_the_outlet_column_is_refused():
    # its child would hold a two-column ExtrapolateRight layer; the outlet sources are built for
    # the last x column only
    from luma_amd._abi import LumaGpuError
    parent = setups.channel_2d(32, 20, 8.0, 0.04, 20.0, outlet=True)
    child = setups.add_refined_region(parent, (14, 31, 0, 19), parent.info["dh"], parent.info["dt"])
    with pytest.raises(LumaGpuError, match="last x column"):
        GpuGrid.from_case(child, "fp64").close()


@pytest.mark.parametrize("prec", ["fp32s", "fp16s"])
def test_two_level_shifted_forms_track_fp64(prec):
    parent = setups.taylor_green((32, 24, 1), u0=0.04, tau=0.6, dims=2)
    child = setups.add_refined_region(parent, (8, 23, 6, 17), 1.0, 1.0)
    setups.velocity_field_on(child, parent)
    res = {}
    for pr in ("fp64", prec):
        with GpuGrid.from_case(parent, pr, store_macro=True) as gp, \
                GpuGrid.from_case(child, pr, store_macro=True) as gc:
            gp.add_child(gc)
            gp.run(200)
            res[pr] = (gp.download_macro()[1], gc.download_macro()[1])
    for i in range(2):
        err = np.linalg.norm(res[prec][i] - res["fp64"][i]) / np.linalg.norm(res["fp64"][i])
        print(prec, "level", i, "rel L2 u error", err)
        assert err < (1e-5 if prec == "fp32s" else 5e-2)


@pytest.mark.parametrize("dims", [2, 3])
def test_three_levels_fp64_bitwise(dims):
    # GridObj_ops_lbm_optimised.cpp:45-49 recursively: the level-1 grid steps its own child twice
    # per sub-step, so level 2 takes four steps per coarse step, each level explodes from and
    # coalesces into the one above
    if dims == 2:
        l0 = setups.taylor_green((28, 24, 1), u0=0.04, tau=0.6, dims=2)
        l1 = setups.add_refined_region(l0, (6, 19, 5, 17), 1.0, 1.0)
        l2 = setups.add_refined_region(l1, (6, 19, 6, 17), 0.5, 0.5, parent_origin=l1.info["origin"], level=2)
    else:
        l0 = setups.taylor_green((16, 14, 14), u0=0.04, tau=0.6, dims=3)
        l1 = setups.add_refined_region(l0, (4, 11, 3, 10, 3, 10), 1.0, 1.0)
        l2 = setups.add_refined_region(l1, (4, 11, 4, 11, 4, 11), 0.5, 0.5, parent_origin=l1.info["origin"], level=2)
    setups.velocity_field_on(l1, l0)
    setups.velocity_field_on(l2, l1)
    print(f"[refine] three levels {dims}D: {l0.N}x{l0.M}x{l0.K}, {l1.N}x{l1.M}x{l1.K}, {l2.N}x{l2.M}x{l2.K}", flush=True)
    cases = (l0, l1, l2)
    ors = [oracle_grid(c) for c in cases]
    ors[0].add_child(ors[1])
    ors[1].add_child(ors[2])
    gs = [GpuGrid.from_case(c, "fp64", store_macro=True) for c in cases]
    try:
        gs[0].add_child(gs[1])
        gs[1].add_child(gs[2])
        for _ in range(12):
            ors[0].step()
            gs[0].LBM_multi_opt()
        assert (gs[0].t, gs[1].t, gs[2].t) == (12, 24, 48)
        for lev, (c, o, g) in enumerate(zip(cases, ors, gs)):
            m = ~np.isin(c.lattyp, [setups.REFINED, setups.SOLID])
            d = np.abs(g.download_f()[m] - o.get_f()[m]).max()
            assert d == 0.0, f"level {lev}: f differs by {d}"
            (rg, ug), (ro, uo) = g.download_macro(), o.get_macro()
            assert np.array_equal(rg[m], ro[m]) and np.array_equal(ug[m], uo[m]), lev
    finally:
        for g in reversed(gs):
            g.close()


def run_tree(cases, parent_of, nsteps):
    """Step a hierarchy (cases[i] a child of cases[parent_of[i]]; parent_of[0] = -1) on the GPU
    and in the oracle, then compare every level bit for bit."""
    ors = [oracle_grid(c) for c in cases]
    gs = [GpuGrid.from_case(c, "fp64", store_macro=True) for c in cases]
    try:
        for i, p in enumerate(parent_of):
            if p >= 0:
                ors[p].add_child(ors[i])
                gs[p].add_child(gs[i])
        for _ in range(nsteps):
            ors[0].step()
            gs[0].LBM_multi_opt()
        for lev, (c, o, g) in enumerate(zip(cases, ors, gs)):
            m = ~np.isin(c.lattyp, [setups.REFINED, setups.SOLID])
            d = np.abs(g.download_f()[m] - o.get_f()[m]).max()
            assert d == 0.0, f"grid {lev}: f differs by {d}"
            (rg, ug), (ro, uo) = g.download_macro(), o.get_macro()
            assert np.array_equal(rg[m], ro[m]) and np.array_equal(ug[m], uo[m]), lev
        return gs
    finally:
        for g in reversed(gs):
            g.close()


@pytest.mark.parametrize("dims", [2, 3])
def test_two_regions_on_one_level_fp64_bitwise(dims):
    # L_NUM_REGIONS > 1: the coarse grid holds two refined regions (GridObj::subGrid), each
    # stepped twice in turn before the coarse step; a TL2F site coalesces from the region that
    # holds it (GridUtils::getSubGrid) and copies its rho, u into that region.  In 2D the second
    # region carries a level-2 region of its own.
    if dims == 2:
        l0 = setups.taylor_green((44, 24, 1), u0=0.04, tau=0.6, dims=2)
        a = setups.add_refined_region(l0, (4, 15, 5, 16), 1.0, 1.0)
        b = setups.add_refined_region(l0, (22, 37, 3, 18), 1.0, 1.0)
        setups.velocity_field_on(a, l0)
        setups.velocity_field_on(b, l0)
        b2 = setups.add_refined_region(b, (8, 23, 8, 23), 0.5, 0.5, parent_origin=b.info["origin"], level=2)
        setups.velocity_field_on(b2, b)
        cases, parents, n = (l0, a, b, b2), (-1, 0, 0, 2), 10
    else:
        l0 = setups.taylor_green((24, 14, 12), u0=0.04, tau=0.6, dims=3)
        a = setups.add_refined_region(l0, (2, 9, 3, 10, 2, 9), 1.0, 1.0)
        b = setups.add_refined_region(l0, (13, 20, 2, 9, 3, 8), 1.0, 1.0)
        setups.velocity_field_on(a, l0)
        setups.velocity_field_on(b, l0)
        cases, parents, n = (l0, a, b), (-1, 0, 0), 8
    print(f"[refine] two regions {dims}D: " + ", ".join(f"{c.N}x{c.M}x{c.K}" for c in cases), flush=True)
    run_tree(cases, parents, n)


def test_overlapping_regions_are_refused():
    from luma_amd._abi import LumaGpuError
    l0 = setups.taylor_green((40, 24, 1), u0=0.04, tau=0.6, dims=2)
    a = setups.add_refined_region(l0, (4, 15, 5, 16), 1.0, 1.0)
    b = setups.add_refined_region(l0, (15, 27, 5, 16), 1.0, 1.0)
    with GpuGrid.from_case(l0, "fp64") as g0, GpuGrid.from_case(a, "fp64") as ga, GpuGrid.from_case(b, "fp64") as gb:
        g0.add_child(ga)
        with pytest.raises(LumaGpuError, match="overlap"):
            g0.add_child(gb)
