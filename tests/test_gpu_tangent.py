"""Tangent form (precision "fp64t"): the step on dual numbers, carrying the derivative of every
value along one seeded parameter direction.  Its values must equal the FP64 form bit for bit
(the FP64 form equals the oracle, see the other GPU tests); with nothing seeded every
derivative is exactly zero; with a filament's Young's modulus seeded, the derivatives of the
marker trajectory and of the flow fields must match central finite differences of FP64 runs.
Bound: 1e-4 relative (central differences with relative step 1e-4, the difference runs pinned
to the derivative run's FSI sub-iteration schedule; the delta kernel's second derivative jumps
at |r| = 1/2, so a much wider stencil that crosses it is only first-order accurate)."""
import numpy as np
import pytest

from luma_amd import setups
from luma_amd.grid import GpuGrid

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def grid(case, prec):
    g = GpuGrid.from_case(case, prec, store_macro=True)
    if case.lattyp is None and case.xpos is not None:
        g.set_positions(case.xpos, case.ypos, case.zpos)
    return g


def assert_same_values(a, b, tag):
    fa, fb = a.download_f(), b.download_f()
    assert np.array_equal(fa, fb), f"{tag}: f differs by {np.abs(fa - fb).max()}"
    (ra, ua), (rb, ub) = a.download_macro(), b.download_macro()
    assert np.array_equal(ra, rb) and np.array_equal(ua, ub), tag


def assert_zero_tangent(g, tag):
    df, dr, du = g.download_tangent()
    assert not df.any() and not dr.any() and not du.any(), tag


def test_values_equal_fp64_and_derivatives_vanish_unseeded():
    cases = [("cavity", setups.lid_driven_cavity(32, re=1280.0 * 32 / 256), 40),
             ("channel", setups.channel_2d(48, 24, 8.0, 0.04, 50.0), 40),
             ("tgv3d", setups.taylor_green((12, 10, 8), dims=3, tau=0.55, u0=0.05), 20)]
    for name, case, steps in cases:
        print(f"[tangent] {name}", flush=True)
        with grid(case, "fp64") as a, grid(case, "fp64t") as b:
            a.run(steps)
            b.run(steps)
            assert_same_values(a, b, name)
            assert_zero_tangent(b, name)


def test_refined_rigid_body_values_equal_fp64():
    parent = setups.channel_2d(64, 32, 8.0, 0.04, 80.0)
    child = setups.add_refined_region(parent, (8, 31, 6, 25), parent.info["dh"], parent.info["dt"])
    pos = setups.circle_markers((2.5 + 0.01, 2.0 + 0.13, 0.0), 0.5, child.info["dh"])
    out = {}
    for prec in ("fp64", "fp64t"):
        print(f"[tangent] refined rigid {prec}", flush=True)
        with grid(parent, prec) as gp, grid(child, prec) as gc:
            gp.add_child(gc)
            gc.ibm_add_body(pos)
            gc.ibm_initialise()
            for _ in range(20):
                gp.LBM_multi_opt()
            out[prec] = (gp.download_f(), gc.download_f(), gc.ibm_markers(0), gc.ibm_body_force(0))
            if prec == "fp64t":
                assert_zero_tangent(gp, "parent")
                assert_zero_tangent(gc, "child")
                assert not gc.ibm_body_force_tangent(0)[1].any()
    a, b = out["fp64"], out["fp64t"]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for key in ("force", "rho", "mom", "eps", "ds"):
        assert np.array_equal(a[2][key], b[2][key]), key
    assert np.array_equal(a[3], b[3])


def test_sphere_3d_values_equal_fp64():
    c = setups.taylor_green((20, 16, 16), dims=3, tau=0.6, u0=0.0)
    c.u[...] = 0.0
    c.u[..., 0] = 0.03
    c.rho[...] = 1.0
    c.forced = True
    c.xpos = 0.5 + np.arange(c.N, dtype=float)
    c.ypos = 0.5 + np.arange(c.M, dtype=float)
    c.zpos = 0.5 + np.arange(c.K, dtype=float)
    c.edges = (float(c.N), float(c.M), float(c.K))
    c.info.update({"dh": 1.0, "origin": (0.0, 0.0, 0.0)})
    pos = setups.sphere_markers((8.3, 8.1, 7.9), 2.5, 1.0)
    out = {}
    for prec in ("fp64", "fp64t"):
        with grid(c, prec) as g:
            g.ibm_add_body(pos)
            g.ibm_initialise()
            g.run(12)
            out[prec] = (g.download_f(), g.ibm_markers(0)["force"])
    assert np.array_equal(out["fp64"][0], out["fp64t"][0])
    assert np.array_equal(out["fp64"][1], out["fp64t"][1])


# ---- flexible filament: d/dE against central differences ---------------------------------

def filament_setup(refined):
    if refined:
        parent = setups.channel_2d(64, 32, 8.0, 0.05, 100.0)
        child = setups.add_refined_region(parent, (10, 41, 8, 23), parent.info["dh"], parent.info["dt"])
        setups.velocity_field_on(child, parent)
        fine = child
        start, angle = (2.0, 2.0 + 0.013, 0.0), 2.0
    else:
        parent, child = setups.channel_2d(64, 32, 8.0, 0.04, 100.0), None
        fine = parent
        start, angle = (1.5, 2.0 + 0.013, 0.0), 8.0
    density, E = setups.filament_material(1.0, 0.05, fine.info["dh"], mass_ratio=10.0, bending=1e-2)
    return parent, child, dict(start=start, length=1.0, height=0.05, angle=angle, elements=8,
                               density=density, E=E)


def run_filament(parent, child, fil, prec, E, steps, seed=False, pin=None):
    """Tip trajectory (y of the last marker, per coarse step), FSI sub-iteration counts of
    every fine step, and the final fine-grid state; derivatives too in the tangent form.
    pin: a sub-iteration schedule to follow instead of the residual test."""
    grids = [grid(parent, prec)]
    if child is not None:
        grids.append(grid(child, prec))
        grids[0].add_child(grids[1])
    fg = grids[-1]
    try:
        b = fg.ibm_add_filament(fil["start"], fil["length"], fil["height"], fil["angle"], fil["elements"], True,
                                fil["density"], E)
        fg.ibm_initialise()
        if seed:
            fg.ibm_seed_youngs(b, 1.0)
        if pin is not None:
            fg.ibm_pin_subiterations(pin)
        tip, dtip = [], []
        for _ in range(steps):
            grids[0].LBM_multi_opt()
            tip.append(fg.ibm_markers(b)["pos"][-1, 1])
            if seed:
                dtip.append(fg.ibm_marker_tangents(b)["pos"][-1, 1])
        f = fg.download_f()
        rho, u = fg.download_macro()
        tan = fg.download_tangent() if seed else None
        force = fg.ibm_body_force_tangent(b) if seed else (fg.ibm_body_force(b), None)
        its = list(fg.ibm_subiteration_log())
        return dict(tip=np.array(tip), dtip=np.array(dtip), its=its, f=f, rho=rho, u=u, tan=tan, force=force)
    finally:
        for g in reversed(grids):
            g.close()


@pytest.mark.parametrize("refined,steps", [(False, 30), (True, 12)])
def test_filament_youngs_derivative_matches_central_differences(refined, steps):
    parent, child, fil = filament_setup(refined)
    E = fil["E"]
    print(f"[tangent] filament refined={refined} E={E:.5g} steps={steps}", flush=True)
    ref = run_filament(parent, child, fil, "fp64", E, steps)
    tan = run_filament(parent, child, fil, "fp64t", E, steps, seed=True)
    # the values of the tangent form are the FP64 run's, seed or not
    assert tan["its"] == ref["its"]
    assert np.array_equal(tan["tip"], ref["tip"])
    assert np.array_equal(tan["f"], ref["f"]) and np.array_equal(tan["u"], ref["u"])
    # The FSI loop stops on a residual threshold, so E +- h may take other sub-iteration
    # counts than E; the difference runs follow E's schedule (pinned), so that both sides
    # differentiate the same discrete program the tangent form does.
    h = 1e-4 * E
    up = run_filament(parent, child, fil, "fp64", E + h, steps, pin=ref["its"])
    dn = run_filament(parent, child, fil, "fp64", E - h, steps, pin=ref["its"])
    assert up["its"] == ref["its"] == dn["its"]
    fd_tip = (up["tip"] - dn["tip"]) / (2 * h)
    scale = np.abs(tan["dtip"]).max()
    assert scale > 0.0
    err = np.abs(tan["dtip"] - fd_tip).max() / scale
    print(f"d tip/dE: max {scale:.4g}, max rel diff vs central differences {err:.3g}", flush=True)
    assert err < RTOL
    # mean tip displacement over the run (the C5 objective's form) and a projection of the fields
    J_tan, J_fd = tan["dtip"].mean(), fd_tip.mean()
    assert abs(J_tan - J_fd) <= RTOL * abs(J_tan)
    rng = np.random.default_rng(5)
    for key, d_tan in (("u", tan["tan"][2]), ("rho", tan["tan"][1]), ("f", tan["tan"][0])):
        r = rng.standard_normal(d_tan.shape)
        pt = float((d_tan * r).sum())
        pf = float(((up[key] - dn[key]) / (2 * h) * r).sum())
        print(f"d{key}/dE . r: tangent {pt:.8g} fd {pf:.8g}", flush=True)
        assert abs(pt - pf) <= RTOL * abs(pt)
    # body force derivative
    (F, dF), (Fu, _), (Fd, _) = tan["force"], up["force"], dn["force"]
    fdF = (Fu - Fd) / (2 * h)
    assert np.abs(dF - fdF).max() <= RTOL * np.abs(dF).max()


def test_pinned_subiteration_schedule_replays_a_run():
    # a run pinned to another run's logged FSI schedule repeats it bit for bit; an unpinned
    # step after the schedule ends goes back to the residual test
    parent, child, fil = filament_setup(False)
    free = run_filament(parent, child, fil, "fp64", fil["E"], 10)
    pinned = run_filament(parent, child, fil, "fp64", fil["E"], 10, pin=free["its"][:6])
    assert pinned["its"] == free["its"]
    assert np.array_equal(pinned["tip"], free["tip"]) and np.array_equal(pinned["f"], free["f"])
    forced = run_filament(parent, child, fil, "fp64", fil["E"], 3, pin=[1, 1, 1])
    assert forced["its"] == [1, 1, 1]


def test_fp32_shifted_tangent_tracks_fp64_tangent():
    # the shifted tangent form: FP32-shifted storage and collision on float duals; with nothing
    # seeded its derivatives vanish, its values follow the FP32-shifted form, and seeded its
    # derivatives track the FP64 tangent form's within FP32 accuracy
    case = setups.channel_2d(48, 24, 8.0, 0.04, 50.0)
    with grid(case, "fp32s") as a, grid(case, "fp32t") as b:
        a.run(40)
        b.run(40)
        fa, fb = a.download_f(), b.download_f()
        print("fp32t vs fp32s values, max |df|", np.abs(fa - fb).max(), flush=True)
        assert np.abs(fa - fb).max() <= 1e-6
        assert_zero_tangent(b, "channel fp32t")
    parent, child, fil = filament_setup(True)
    E = fil["E"]
    r64 = run_filament(parent, child, fil, "fp64t", E, 12, seed=True)
    r32 = run_filament(parent, child, fil, "fp32t", E, 12, seed=True)
    s32 = run_filament(parent, child, fil, "fp32s", E, 12)
    # the same FP32 arithmetic up to the order of fused operations: far below a cell
    assert np.abs(r32["tip"] - s32["tip"]).max() <= 1e-6 * child.info["dh"]
    err = np.abs(r32["dtip"] - r64["dtip"]).max() / np.abs(r64["dtip"]).max()
    print(f"fp32t vs fp64t d tip/dE: max rel diff {err:.3g}", flush=True)
    assert err < 1e-3
    for k in (0, 1, 2):
        d64, d32 = r64["tan"][k], r32["tan"][k]
        rel = np.abs(d32 - d64).max() / np.abs(d64).max()
        print(f"fp32t vs fp64t tangent field {k}: max rel diff {rel:.3g}", flush=True)
        assert rel < 1e-2


def test_tangent_and_reverse_adjoint_agree_on_domega():
    # two independent derivative codes: the FP64 tangent form seeded in omega and the reverse
    # discrete adjoint (lumagpu_adj_run) give the same d(time-averaged drag)/d(omega)
    from test_gpu_adjoint import channel_case, make_grid
    case, pos = channel_case()
    T = 60
    with make_grid(case, pos) as g:
        adj = g.adj_run(T, 16)
    with make_grid(case, pos, prec="fp64t") as g:
        g.seed_omega(1.0)
        J, dJ = 0.0, 0.0
        for _ in range(T):
            g.LBM_multi_opt()
            F, dF = g.ibm_body_force_tangent(0)
            J += F[0] / T
            dJ += dF[0] / T
    print(f"J {J!r} adjoint {adj['J']!r}; dJ/domega tangent {dJ!r} adjoint {adj['dJ_domega']!r}", flush=True)
    assert abs(J - adj["J"]) <= 1e-12 * abs(adj["J"])   # same values, another summation order
    assert abs(dJ - adj["dJ_domega"]) <= 1e-9 * abs(adj["dJ_domega"])


def test_tangent_entry_points_refuse_other_forms():
    from luma_amd._abi import LumaGpuError
    case = setups.channel_2d(48, 24, 8.0, 0.04, 50.0)
    with grid(case, "fp64") as g:
        with pytest.raises(LumaGpuError, match="tangent-form"):
            g.download_tangent()
        with pytest.raises(LumaGpuError, match="omega seed"):
            g.seed_omega(1.0)
    with grid(case, "fp32t") as g:
        with pytest.raises(LumaGpuError, match="omega seed"):
            g.seed_omega(1.0)
