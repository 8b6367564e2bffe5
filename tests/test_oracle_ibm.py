"""Pin the IB oracle (ObjectManager_ops_ibm.cpp restated) with closed forms and an
independent numpy assembly of the epsilon system."""
import math

import numpy as np
import pytest

from luma_amd import setups
from oracle.oracle import OracleGrid, OracleIBM, delta_kernel, lu_solve


@pytest.mark.parametrize("r0", [0.0, 0.17, 0.5, 0.73, 0.999])
def test_delta_kernel_moments(r0):
    # Roma et al. 3-point kernel: sum phi = 1, sum r phi = 0, sum phi^2 = 1/2, support 1.5
    js = np.arange(-3, 4)
    ph = np.array([delta_kernel(r0 + j) for j in js])
    assert abs(ph.sum() - 1.0) < 1e-15
    assert abs(((r0 + js) * ph).sum()) < 1e-15
    assert abs((ph ** 2).sum() - 0.5) < 1e-15
    assert delta_kernel(1.5) == 0.0 and delta_kernel(1.51) == 0.0 and delta_kernel(0.0) == pytest.approx(2 / 3)


def test_lu_solve_matches_numpy():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((40, 40)) + 3 * np.eye(40)
    b = rng.standard_normal(40)
    assert np.allclose(lu_solve(A, b), np.linalg.solve(A, b), rtol=1e-12, atol=1e-13)


def periodic(N, M, u, tau=0.8):
    c = setups.taylor_green((N, M, 1), dims=2, tau=tau, u0=0.0)
    c.u[...] = 0.0
    c.u[..., 0], c.u[..., 1] = u
    c.rho[...] = 1.0
    g = OracleGrid(2, 9, N, M, 1, omega=c.omega, forced=True, edges=(float(N), float(M), 0.0))
    g.set_positions(0.5 + np.arange(N), 0.5 + np.arange(M), np.array([0.0]))
    g.init_feq(c.rho, c.u)
    return g


def test_uniform_flow_interpolates_exactly_and_exerts_no_force():
    U = (0.03, -0.01)
    g = periodic(32, 28, U)
    pos = setups.circle_markers((16.2, 14.1, 0.0), 4.0, 1.0, luma_skip_first=False)
    ib = OracleIBM(g, 1.0)
    ib.add_body(pos)
    ib.initialise()
    ib.set_velocity(0, np.tile([U[0], U[1], 0.0], (len(pos), 1)))
    g.step()
    mk = ib.markers(0)
    assert np.allclose(mk["rho"], 1.0, rtol=0, atol=1e-14)
    assert np.allclose(mk["mom"][:, 0], U[0], atol=1e-15) and np.allclose(mk["mom"][:, 1], U[1], atol=1e-15)
    assert np.abs(mk["force"]).max() < 1e-15
    assert (mk["ns"] == 9).all()


def test_epsilon_against_independent_numpy_assembly_and_spread_balance():
    g = periodic(40, 36, (0.02, 0.0))
    pos = setups.circle_markers((20.3, 18.2, 0.0), 6.0, 1.0)        # LUMA's n-1 markers
    ib = OracleIBM(g, 1.0)
    ib.add_body(pos)
    ib.initialise()
    mk = ib.markers(0)
    n = len(pos)
    # ds: distance to the nearest other marker (regular polygon with one gap)
    d = np.linalg.norm(pos[:, None, :] - pos[None, :, :], axis=-1)
    np.fill_diagonal(d, np.inf)
    assert np.allclose(mk["ds"], d.min(1), rtol=1e-14)
    assert np.allclose(mk["ds"], 2 * 6.0 * math.sin(math.pi / (n + 1)), rtol=1e-12)
    # A_IJ = ds_J sum_s delta_I(s) delta(x_J - x_s), assembled here from the support lists
    A = np.zeros((n, n))
    for I in range(n):
        ijk, dI = ib.support(0, I)
        xs = 0.5 + ijk[:, 0]
        ys = 0.5 + ijk[:, 1]
        for J in range(n):
            dJ = np.array([delta_kernel(pos[J, 0] - x) * delta_kernel(pos[J, 1] - y) for x, y in zip(xs, ys)])
            A[I, J] = (dI * dJ).sum() * mk["ds"][J]
    eps = np.linalg.solve(A, np.ones(n))
    assert np.allclose(mk["eps"], eps, rtol=1e-11)
    # spread: momentum given to the grid equals minus the marker forces weighted by eps ds
    g.step()
    mk = ib.markers(0)
    F = g.get_force().sum((0, 1, 2))
    ref = -(mk["force"][:, :2] * (mk["eps"] * mk["ds"])[:, None]).sum(0)
    assert np.allclose(F, ref, rtol=1e-12, atol=1e-16)
    assert np.abs(mk["force"]).max() > 1e-6


def _joint_epsilon(ib, bodies, pos_all):
    """A eps = 1 over the markers of several bodies together, assembled from the support lists."""
    n = len(pos_all)
    ds = np.concatenate([ib.markers(b)["ds"] for b in bodies])
    sup = [ib.support(b, m) for b in bodies for m in range(ib.nm[b])]
    A = np.zeros((n, n))
    for I, (ijk, dI) in enumerate(sup):
        xs, ys = 0.5 + ijk[:, 0], 0.5 + ijk[:, 1]
        for J in range(n):
            dJ = np.array([delta_kernel(pos_all[J, 0] - x) * delta_kernel(pos_all[J, 1] - y) for x, y in zip(xs, ys)])
            A[I, J] = (dI * dJ).sum() * ds[J]
    return np.linalg.solve(A, np.ones(n))


def test_universal_epsilon_couples_overlapping_bodies_only():
    # L_UNIVERSAL_EPSILON_CALC (ObjectManager_ops_ibm.cpp:775-786, :905-909, :1067-1141): one
    # system over every marker in body order.  Bodies far apart give a block-diagonal system, so
    # the joint solve returns each body's own epsilon exactly; bodies whose supports overlap are
    # coupled, and the joint epsilon solves the joint numpy assembly
    a = setups.circle_markers((10.3, 12.2, 0.0), 4.0, 1.0)
    far = setups.circle_markers((30.1, 12.4, 0.0), 3.0, 1.0)
    near = setups.circle_markers((16.6, 12.9, 0.0), 3.0, 1.0)       # supports overlap a's
    eps = {}
    for name, b2, uni in (("sep", far, False), ("sep_u", far, True), ("ov", near, False), ("ov_u", near, True)):
        g = periodic(44, 26, (0.02, 0.0))
        ib = OracleIBM(g, 1.0)
        if uni:
            ib.set_universal_epsilon(True)
        ib.add_body(a)
        ib.add_body(b2)
        ib.initialise()
        eps[name] = np.concatenate([ib.markers(0)["eps"], ib.markers(1)["eps"]])
        if name == "ov_u":
            ref = _joint_epsilon(ib, (0, 1), np.concatenate([a, near]))
            assert np.allclose(eps[name], ref, rtol=1e-11)
    assert np.array_equal(eps["sep"], eps["sep_u"])
    assert np.abs(eps["ov"] - eps["ov_u"]).max() > 1e-3


def test_support_stencil_3d_has_27_sites_led_by_the_enclosing_voxel():
    N = 12
    g = OracleGrid(3, 19, N, N, N, omega=1.2, forced=True, edges=(float(N),) * 3)
    x = 0.5 + np.arange(N)
    g.set_positions(x, x, x)
    g.init_feq(np.ones((N, N, N)), np.zeros((N, N, N, 3)))
    pos = setups.sphere_markers((6.2, 5.9, 6.1), 2.0, 1.0)
    ib = OracleIBM(g, 1.0)
    ib.add_body(pos)
    ib.initialise()
    mk = ib.markers(0)
    assert (mk["ns"] == 27).all()
    for m in range(len(pos)):
        ijk, dl = ib.support(0, m)
        assert (ijk[0] == np.floor(pos[m]).astype(int)).all()
        assert abs(dl.sum() - 1.0) < 1e-14


def test_gravity_reset_momentum_balance():
    # L_GRAVITY_ON with an IB body on the level (GridObj_ops_lbm_optimised.cpp:53-57, :1327-1337).
    # A forced BGK step adds exactly the final force field to the total momentum, and streaming
    # in a periodic box keeps it, so per step:
    #   along gravity: dP = sum_sites rho_prev * g + (this step's spread),
    #   across gravity: dP = sum of every spread so far -- the reset leaves that component alone,
    # where a spread gives the grid -sum_m F_m eps_m ds_m (the balance pinned above).
    from oracle.oracle import lattice
    grav = 3e-6
    g = periodic(40, 36, (0.0, 0.0))
    g.set_gravity(grav, 0)
    pos = setups.circle_markers((20.3, 18.2, 0.0), 6.0, 1.0)
    ib = OracleIBM(g, 1.0)
    ib.add_body(pos)
    ib.initialise()
    c, _, _ = lattice(9)
    p_prev = (g.get_f() @ c[:, :2]).sum((0, 1, 2))
    rho_prev, _ = g.get_macro()
    spread_y = 0.0
    for step in range(6):
        g.step()
        mk = ib.markers(0)
        sp = -(mk["force"][:, :2] * (mk["eps"] * mk["ds"])[:, None]).sum(0)
        spread_y += sp[1]
        p = (g.get_f() @ c[:, :2]).sum((0, 1, 2))
        dp = p - p_prev
        # momentum sums over 1440 sites x 9 populations round at ~1e-13
        assert abs(dp[0] - (rho_prev.sum() * grav + sp[0])) <= 1e-10 * abs(dp[0]) + 1e-13, step
        assert abs(dp[1] - spread_y) <= 1e-10 * abs(dp[1]) + 1e-13, step
        p_prev = p
        rho_prev, _ = g.get_macro()
    assert abs(spread_y) > 0.0
