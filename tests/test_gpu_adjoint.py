"""Discrete adjoint of the step (lumagpu_adj_run) against central finite differences of the
forward step, whose FP64 form equals the oracle bit for bit (the other GPU tests): gradients
of the time-averaged IB drag with respect to the relaxation frequency, the inlet velocity
scale, the marker positions and the initial populations, in 2D (channel with a forced-
equilibrium or a regularised velocity inlet, ExtrapolateRight or regularised pressure outlet,
solid walls, IB cylinder;
a box of regularised velocity walls meeting at corner sites, IB cylinder) and 3D (periodic box,
IB sphere; uniform inflow); checkpointing with a partial last segment; FP32-shifted adjoint
against FP64.
Bound: 1e-4 relative (central differences, step sizes in each test)."""
import numpy as np
import pytest

from luma_amd import setups
from luma_amd.grid import GpuGrid

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def channel_case():
    c = setups.channel_2d(48, 24, 8.0, 0.04, 20.0, regularised=False)
    pos = setups.circle_markers((1.52, 1.49, 0.0), 0.3, c.info["dh"])
    return c, pos


def channel_reg_case():
    # the same channel with LUMA's regularised velocity inlet (Latt 2008 faces; the inlet's
    # ends meet the solid walls, so there are no edge sites)
    c = setups.channel_2d(48, 24, 8.0, 0.04, 20.0, regularised=True)
    pos = setups.circle_markers((1.52, 1.49, 0.0), 0.3, c.info["dh"])
    return c, pos


def channel_pressure_case():
    # regularised inlet and LUMA's default outlet, a regularised pressure face (its tangential
    # velocity extrapolated from the two inward neighbours)
    c = setups.channel_2d(48, 24, 8.0, 0.04, 20.0, regularised=True, outlet="pressure")
    pos = setups.circle_markers((1.52, 1.49, 0.0), 0.3, c.info["dh"])
    return c, pos


def corner_box_case():
    # regularised velocity walls left, right and on top, meeting at corner sites (the density
    # extrapolated from two inward neighbours), solid bottom; an IB cylinder inside
    import dataclasses
    c = dataclasses.replace(setups.velocity_corner_box(32, 24, u_in=0.02, u_lid=0.05, tau=0.6), forced=True)
    dh = c.info["dh"]
    pos = setups.circle_markers((16.3 * dh, 11.6 * dh, 0.0), 3.0 * dh, dh)
    return c, pos


def sphere_case():
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
    return c, pos


def inflow_case():
    c = setups.uniform_inflow_3d(24, 14, 14, 0.03, 0.6)
    pos = setups.sphere_markers((9.2, 7.1, 6.9), 2.5, 1.0)
    return c, pos


def make_grid(case, pos, prec="fp64", omega=None, inlet_scale=1.0, f0=None, ramp=None, re_ramp=None):
    c = case
    if omega is not None:
        import dataclasses
        c = dataclasses.replace(case, omega=omega)
    g = GpuGrid.from_case(c, prec, store_macro=True)
    if c.lattyp is None:
        g.set_positions(c.xpos, c.ypos, c.zpos)
    if c.ux_in is not None and inlet_scale != 1.0:
        g.set_inlet(c.ux_in * inlet_scale, c.uy_in * inlet_scale, c.uz_in * inlet_scale)
    if f0 is not None:
        g.upload_f(f0)
    if ramp is not None:
        g.set_velocity_ramp(ramp)
    if re_ramp is not None:
        g.set_reynolds_ramp(*re_ramp)
    g.ibm_add_body(pos)
    g.ibm_initialise()
    return g


def J_of(case, pos, T, **kw):
    with make_grid(case, pos, **kw) as g:
        return g.adj_run(T, T, backward=False)["J"]


def close(a, b, rtol=RTOL, scale=None):
    s = max(abs(b), scale or 0.0)
    return abs(a - b) <= rtol * s


@pytest.mark.parametrize("build,T,K", [(channel_case, 60, 16), (channel_reg_case, 60, 16),
                                       (channel_pressure_case, 60, 16), (corner_box_case, 40, 12),
                                       (sphere_case, 24, 10), (inflow_case, 30, 12)])
def test_adjoint_matches_central_differences(build, T, K):
    case, pos = build()
    print(f"[adjoint] {case.name} T={T} K={K} markers={len(pos)}", flush=True)
    with make_grid(case, pos) as g:
        f0 = g.download_f()
        adj = g.adj_run(T, K, positions=True, initial=True)
    J = adj["J"]
    assert J != 0.0

    # relaxation frequency
    h = 1e-5 * case.omega
    fd = (J_of(case, pos, T, omega=case.omega + h) - J_of(case, pos, T, omega=case.omega - h)) / (2 * h)
    print("dJ/domega adjoint", adj["dJ_domega"], "fd", fd, flush=True)
    assert close(adj["dJ_domega"], fd)

    # inlet velocity scale
    if case.ux_in is not None and np.any(case.lattyp == setups.VELOCITY):
        h = 1e-5
        fd = (J_of(case, pos, T, inlet_scale=1 + h) - J_of(case, pos, T, inlet_scale=1 - h)) / (2 * h)
        print("dJ/dinlet adjoint", adj["dJ_dinlet"], "fd", fd, flush=True)
        assert close(adj["dJ_dinlet"], fd)

    # initial populations along a random direction on the fluid and velocity sites (a
    # forced-equilibrium source's own populations are never read: both sides give 0 there)
    rng = np.random.default_rng(11)
    fluid = np.ones(case.N * case.M * case.K, bool) if case.lattyp is None else \
        np.isin(case.lattyp, (setups.FLUID, setups.VELOCITY, setups.PRESSURE)).reshape(-1)
    dirn = rng.standard_normal(f0.shape) * fluid.reshape(f0.shape[:3])[..., None]
    h = 1e-6
    fd = (J_of(case, pos, T, f0=f0 + h * dirn) - J_of(case, pos, T, f0=f0 - h * dirn)) / (2 * h)
    pred = float((adj["lambda0"] * dirn).sum())
    print("dJ/df0 . d adjoint", pred, "fd", fd, flush=True)
    assert close(pred, fd)

    # marker positions (through delta weights, epsilon and ds)
    gscale = np.abs(adj["dJ_dpos"]).max()
    dh = case.info.get("dh", 1.0)
    for m in (0, len(pos) // 3):
        for a in range(case.dims):
            h = 1e-5 * dh
            pp, pm = pos.copy(), pos.copy()
            pp[m, a] += h
            pm[m, a] -= h
            fd = (J_of(case, pp, T) - J_of(case, pm, T)) / (2 * h)
            print(f"dJ/dx[{m},{a}] adjoint", adj["dJ_dpos"][m, a], "fd", fd, flush=True)
            assert close(adj["dJ_dpos"][m, a], fd, rtol=RTOL, scale=1e-2 * gscale)


@pytest.mark.parametrize("regularised", [False, True])
def test_adjoint_with_velocity_ramp(regularised):
    # L_VELOCITY_RAMP over the first 30 of the 60 steps: every reversed step takes the ramp
    # factor of its own time (forced-equilibrium sources and regularised faces)
    case = channel_case()[0] if not regularised else channel_reg_case()[0]
    pos = channel_case()[1]
    ramp = 30 * case.info["dt"]
    T, K = 60, 16
    print(f"[adjoint] velocity ramp regularised={regularised}", flush=True)
    with make_grid(case, pos, ramp=ramp) as g:
        f0 = g.download_f()
        adj = g.adj_run(T, K, initial=True)
    h = 1e-5 * case.omega
    fd = (J_of(case, pos, T, omega=case.omega + h, ramp=ramp) - J_of(case, pos, T, omega=case.omega - h, ramp=ramp)) / (2 * h)
    print("dJ/domega adjoint", adj["dJ_domega"], "fd", fd, flush=True)
    assert close(adj["dJ_domega"], fd)
    h = 1e-5
    fd = (J_of(case, pos, T, inlet_scale=1 + h, ramp=ramp) - J_of(case, pos, T, inlet_scale=1 - h, ramp=ramp)) / (2 * h)
    print("dJ/dinlet adjoint", adj["dJ_dinlet"], "fd", fd, flush=True)
    assert close(adj["dJ_dinlet"], fd)
    rng = np.random.default_rng(5)
    m = np.isin(case.lattyp, (setups.FLUID, setups.VELOCITY)).reshape(f0.shape[:3])
    dirn = rng.standard_normal(f0.shape) * m[..., None]
    h = 1e-6
    fd = (J_of(case, pos, T, f0=f0 + h * dirn, ramp=ramp) - J_of(case, pos, T, f0=f0 - h * dirn, ramp=ramp)) / (2 * h)
    pred = float((adj["lambda0"] * dirn).sum())
    print("dJ/df0 . d adjoint", pred, "fd", fd, flush=True)
    assert close(pred, fd)


@pytest.mark.parametrize("regularised", [False, True])
def test_adjoint_with_reynolds_ramp(regularised):
    # L_REYNOLDS_RAMP over the first 40 of the 60 steps: every reversed step relaxes with its own
    # omega, and dJ_domega is the derivative with respect to the omega after the ramp, which
    # L_RE sets (omega = 1 / (3 nu + 1/2), nu = dt / (dh^2 L_RE)); with the regularised inlet
    # the faces' omega partials take the same per-step weights
    case = channel_case()[0] if not regularised else channel_reg_case()[0]
    pos = channel_case()[1]
    dt, dh = case.info["dt"], case.info["dh"]

    def re_of(omega):
        return dt / (dh * dh) / ((1.0 / omega - 0.5) / 3.0)

    ramp = 40 * dt
    T, K = 60, 16
    print(f"[adjoint] Reynolds ramp regularised={regularised}", flush=True)
    with make_grid(case, pos, re_ramp=(re_of(case.omega), ramp)) as g:
        adj = g.adj_run(T, K)
        om_adj = g.omega
    with make_grid(case, pos, re_ramp=(re_of(case.omega), ramp)) as g:
        g.adj_run(T, T, backward=False)
        om_fwd = g.omega
    # the grid is left at the end of the run with that step's omega
    assert om_adj == om_fwd
    h = 1e-5 * case.omega
    fd = (J_of(case, pos, T, re_ramp=(re_of(case.omega + h), ramp)) -
          J_of(case, pos, T, re_ramp=(re_of(case.omega - h), ramp))) / (2 * h)
    print("dJ/domega adjoint", adj["dJ_domega"], "fd", fd, flush=True)
    assert close(adj["dJ_domega"], fd)
    # against the unramped run: the weights change the answer
    with make_grid(case, pos) as g:
        flat = g.adj_run(T, K)["dJ_domega"]
    assert not close(adj["dJ_domega"], flat)
    h = 1e-5
    fd = (J_of(case, pos, T, inlet_scale=1 + h, re_ramp=(re_of(case.omega), ramp)) -
          J_of(case, pos, T, inlet_scale=1 - h, re_ramp=(re_of(case.omega), ramp))) / (2 * h)
    print("dJ/dinlet adjoint", adj["dJ_dinlet"], "fd", fd, flush=True)
    assert close(adj["dJ_dinlet"], fd)


def test_adjoint_universal_epsilon():
    # L_UNIVERSAL_EPSILON_CALC: two cylinders whose supports overlap share one epsilon system;
    # the position gradient goes through that joint system (and ds body by body)
    case, pa = channel_case()
    dh = case.info["dh"]
    pb = setups.circle_markers((2.1, 1.5, 0.0), 0.2, dh)
    T, K = 50, 16

    def grid(a, b):
        g = GpuGrid.from_case(case, "fp64", store_macro=True)
        g.ibm_set_universal_epsilon(True)
        g.ibm_add_body(a)
        g.ibm_add_body(b)
        g.ibm_initialise()
        return g

    print(f"[adjoint] universal epsilon, {len(pa)} + {len(pb)} markers", flush=True)
    with grid(pa, pb) as g:
        adj = g.adj_run(T, K, positions=True)
    grad = adj["dJ_dpos"]
    gscale = np.abs(grad).max()
    for body, m in ((0, 3), (1, 0), (1, len(pb) // 2)):
        idx = m if body == 0 else len(pa) + m
        for a in range(2):
            h = 1e-5 * dh
            pos = [pa.copy(), pb.copy()]
            pos[body][m, a] += h
            with grid(*pos) as g:
                jp = g.adj_run(T, T, backward=False)["J"]
            pos[body][m, a] -= 2 * h
            with grid(*pos) as g:
                jm = g.adj_run(T, T, backward=False)["J"]
            fd = (jp - jm) / (2 * h)
            print(f"dJ/dx[{body}:{m},{a}] adjoint", grad[idx, a], "fd", fd, flush=True)
            assert close(grad[idx, a], fd, rtol=RTOL, scale=1e-2 * gscale)


def test_checkpoint_interval_does_not_change_the_gradient():
    case, pos = channel_case()
    res = []
    for K in (7, 30, 60):
        with make_grid(case, pos) as g:
            res.append(g.adj_run(60, K))
    for r in res[1:]:
        assert r["J"] == res[0]["J"]
        assert r["dJ_domega"] == res[0]["dJ_domega"]
        assert np.array_equal(r["dJ_dpos"], res[0]["dJ_dpos"])


def test_fp32_shifted_adjoint_tracks_fp64():
    case, pos = channel_case()
    out = {}
    for prec in ("fp64", "fp32s"):
        with make_grid(case, pos, prec=prec) as g:
            out[prec] = g.adj_run(60, 16)
    a, b = out["fp32s"], out["fp64"]
    print("fp32s vs fp64 adjoint", a["dJ_domega"], b["dJ_domega"], a["dJ_dinlet"], b["dJ_dinlet"], flush=True)
    assert close(a["J"], b["J"], rtol=1e-4)
    assert close(a["dJ_domega"], b["dJ_domega"], rtol=1e-3)
    assert close(a["dJ_dinlet"], b["dJ_dinlet"], rtol=1e-3)
    assert np.abs(a["dJ_dpos"] - b["dJ_dpos"]).max() <= 1e-3 * np.abs(b["dJ_dpos"]).max()


def test_torch_autograd_binding():
    import torch

    from luma_amd.adjoint import DragObjective

    case, pos = channel_case()

    def build(omega, inlet_scale, p):
        return make_grid(case, p, omega=omega, inlet_scale=inlet_scale)

    om = torch.tensor(case.omega, dtype=torch.float64, requires_grad=True)
    s = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    p = torch.tensor(pos, dtype=torch.float64, requires_grad=True)
    J = DragObjective.apply(om, s, p, build, 40, 16)
    J.backward()
    with make_grid(case, pos) as g:
        ref = g.adj_run(40, 16)
    assert J.item() == ref["J"]
    assert om.grad.item() == ref["dJ_domega"] and s.grad.item() == ref["dJ_dinlet"]
    assert np.array_equal(p.grad.numpy(), ref["dJ_dpos"])
