"""Pins for the oracle functions that no golden vector covers: the forced-equilibrium and
ExtrapolateRight stream sources, explode/coalesce across a refinement interface, the FSI
sub-iteration, the corotational beam at large rotation and the IB-marker <-> beam mapping.

Each check is fixed by mathematics or by a symmetry of the discrete scheme, and is chosen so that
a plausible mistake in the oracle (a wrong source cell, population, row or offset, a dropped
average, a small-angle beam force, a flipped element coordinate) makes it fail.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp
from scipy.optimize import brentq

from luma_amd import setups
from oracle.oracle import OracleFEM, OracleGrid, OracleIBM, lattice

FLUID, SOLID, VELOCITY, EXTRAP = setups.FLUID, setups.SOLID, setups.VELOCITY, setups.EXTRAP_RIGHT


def feq_numpy(Q, rho, u):
    c, w, _ = lattice(Q)
    cu = c[:, :len(u)] @ np.asarray(u, float)
    return w * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * float(np.dot(u, u)))


def inflow_grid(dims, N, M, K, ux, U=0.04, tau=0.7, forced=False):
    """Forced-equilibrium eVelocity inlet at x = 0 (L_REGULARISED_BOUNDARIES off), eExtrapolateRight
    outlet at x = N-1, periodic in y (and z), initial equilibrium of u = (ux(y), 0, 0), rho = 1."""
    Q = 9 if dims == 2 else 19
    g = OracleGrid(dims, Q, N, M, K, omega=1.0 / tau, regularised=False, forced=forced, fz_compat=False,
                   edges=(N, M, K if dims == 3 else 0.0))
    lt = np.full((N, M, K), FLUID, np.int32)
    lt[0] = VELOCITY
    lt[N - 1] = EXTRAP
    g.set_labels(lt)
    g.set_positions(np.arange(N) + 0.5, np.arange(M) + 0.5, (np.arange(K) + 0.5) if dims == 3 else np.zeros(1))
    g.set_inlet(ux, np.zeros(M), np.zeros(M))
    u = np.zeros((N, M, K, dims))
    u[..., 0] = ux[None, :, None]
    g.init_feq(np.ones((N, M, K)), u)
    return g, lt


# ---- stream sources: GridObj_ops_lbm_optimised.cpp:246-270 ----------------------------------

@pytest.mark.parametrize("dims", [2, 3])
def test_uniform_flow_is_a_fixed_point_through_inlet_and_outlet(dims):
    # Every site holds feq(1, U): the inlet supplies feq of its own (rho, u), the outlet the value
    # two cells left of it, so nothing changes.  A source taking the opposite population, another
    # density or velocity, or the post-stream value of the wrong cell type breaks it.
    N, M, K = 12, 6, (5 if dims == 3 else 1)
    g, lt = inflow_grid(dims, N, M, K, np.full(M, 0.04))
    f0 = g.get_f()
    for _ in range(7):
        g.step()
    f1 = g.get_f()
    live = lt != VELOCITY   # the inlet column is skipped by sweep 1 and only collided
    assert np.abs(f1[live] - f0[live]).max() < 1e-15
    rho, u = g.get_macro()
    assert np.abs(rho[lt == FLUID] - 1).max() < 1e-15 and np.abs(u[lt == FLUID][:, 0] - 0.04).max() < 1e-15


@pytest.mark.parametrize("dims", [2, 3])
def test_sheared_flow_near_the_outlet_follows_the_periodic_box(dims):
    # An x-invariant shear wave u_x(y) decays the same way in every column of a periodic box.  The
    # outlet hands each puller the value two cells left of the SOURCE cell (in y and z), which in an
    # x-invariant state is exactly what the periodic neighbour would give, and the inlet supplies
    # the initial equilibrium, whose disturbance moves at most one column per step.  So after n
    # steps the columns x >= n + 1 equal the periodic box bit for bit, and after one step all
    # columns 1 .. N-2 do.  Taking the pulling site's y, or another population, breaks it.
    N, M, K = 24, 8, (4 if dims == 3 else 1)
    Q = 9 if dims == 2 else 19
    y = np.arange(M) + 0.5
    ux = 0.03 + 0.01 * np.sin(2 * np.pi * y / M)
    g, lt = inflow_grid(dims, N, M, K, ux)
    per = OracleGrid(dims, Q, N, M, K, omega=1.0 / 0.7, regularised=False, fz_compat=False)
    u = np.zeros((N, M, K, dims))
    u[..., 0] = ux[None, :, None]
    per.init_feq(np.ones((N, M, K)), u)
    for n in range(1, 9):
        g.step()
        per.step()
        fg, fp = g.get_f(), per.get_f()
        lo = 1 if n == 1 else n + 1
        assert np.array_equal(fg[lo:N - 1], fp[lo:N - 1]), n
        if n > 1:
            assert not np.array_equal(fg[1], fp[1])   # the inlet's disturbance is real


def _sweep1(g):
    g.lib.ora_sweep1.argtypes = [C.c_void_p, C.c_int]
    g.lib.ora_sweep1(g.h, 0)
    fn = g.lib.ora_fnew
    fn.restype = C.POINTER(C.c_double)
    fn.argtypes = [C.c_void_p]
    N, M, K = g.shape
    return np.ctypeslib.as_array(fn(g.h), shape=(N * M * K * g.Q,)).reshape(N, M, K, g.Q).copy()


@pytest.mark.parametrize("dims", [2, 3])
def test_outlet_source_is_two_cells_left_of_the_source_cell(dims):
    # Known-answer test on tagged populations: f[v] at (i, j, k) holds a number that encodes
    # (i, j, k, v).  After the stream, a population pulled from an eExtrapolateRight cell
    # (i_s, j_s, k_s) must decode to (i_s - 2, j_s, k_s, v), for the pullers at x = N-2 and for
    # the outlet cells themselves (sources on the outlet, c_x = 0, rest population included).
    N, M, K = 9, 5, (4 if dims == 3 else 1)
    Q = 9 if dims == 2 else 19
    c, _, _ = lattice(Q)
    g, lt = inflow_grid(dims, N, M, K, np.full(M, 0.02))
    ii, jj, kk, vv = np.meshgrid(np.arange(N), np.arange(M), np.arange(K), np.arange(Q), indexing="ij")
    tag = 1.0 + vv + Q * (kk + K * (jj + M * ii))
    g.set_f(tag)
    fn = _sweep1(g)
    checked = 0
    for i in (N - 2, N - 1):
        for j in range(M):
            for k in range(K):
                for v in range(Q):
                    si, sj, sk = (i - c[v, 0]) % N, (j - c[v, 1]) % M, (k - c[v, 2]) % K
                    if lt[si, sj, sk] != EXTRAP:
                        continue
                    t = int(fn[i, j, k, v] - 1.0)
                    dv, rest = t % Q, t // Q
                    dk, rest = rest % K, rest // K
                    dj, di = rest % M, rest // M
                    assert (di, dj, dk, dv) == (si - 2, sj, sk, v), (i, j, k, v)
                    checked += 1
    assert checked == M * K * (int((c[:, 0] == -1).sum()) + int((c[:, 0] == 0).sum()))


def test_ramped_forced_equilibrium_source_uses_the_pulling_row():
    # L_VELOCITY_RAMP on a forced-equilibrium inlet (:258-266): the source's u is set to the
    # profile of the PULLING site's row j times the ramp coefficient at t dt, and the pulled value
    # is feq(rho_src, that u).  With a different profile value in every row, the diagonal pulls
    # tell the source row from the pulling row; checked against an independent numpy equilibrium.
    N, M = 8, 6
    ux = 0.01 * (1 + np.arange(M))
    g, lt = inflow_grid(2, N, M, 1, ux)
    T, dt = 10.0, 1.0
    g.set_ramp(T, dt)
    for _ in range(3):
        g.step()
    t = g.t
    fn = _sweep1(g)
    c, _, _ = lattice(9)
    coef = (1 - math.cos(math.pi * t * dt / T)) / 2
    for j in range(M):
        for v in range(9):
            if c[v, 0] != 1:
                continue
            want = feq_numpy(9, 1.0, [ux[j] * coef, 0.0])[v]
            assert fn[1, j, 0, v] == pytest.approx(want, rel=1e-14, abs=1e-17), (j, v)
            if c[v, 1] != 0:
                other = feq_numpy(9, 1.0, [ux[(j - c[v, 1]) % M] * coef, 0.0])[v]
                assert abs(other - want) > 1e-6


# ---- explode / coalesce: GridObj_ops_lbm_optimised.cpp:591-663 -----------------------------

def _two_level(dims, N, M, K, lims, rho, u, crho=None, cu=None, tau=0.7):
    parent = setups.taylor_green((N, M, K), u0=0.0, tau=tau, dims=dims)
    parent.xpos, parent.ypos = np.arange(N) + 0.5, np.arange(M) + 0.5
    parent.zpos = (np.arange(K) + 0.5) if dims == 3 else np.zeros(1)
    child = setups.add_refined_region(parent, lims, 1.0, 1.0)
    Q = parent.Q

    def mk(c, r, v):
        g = OracleGrid(dims, Q, c.N, c.M, c.K, omega=c.omega, regularised=False, fz_compat=False,
                       level=c.info.get("level", 0), coarse_lims=c.info.get("coarse_lims", (0,) * 6))
        g.set_labels(c.lattyp)
        g.set_positions(c.xpos, c.ypos, c.zpos)
        g.init_feq(r, v)
        return g
    op = mk(parent, rho, u)
    oc = mk(child, crho if crho is not None else np.ones((child.N, child.M, child.K)),
            cu if cu is not None else np.zeros((child.N, child.M, child.K, dims)))
    op.add_child(oc)
    return op, oc, parent, child


@pytest.mark.parametrize("dims", [2, 3])
def test_uniform_flow_is_a_fixed_point_across_a_refined_patch(dims):
    # In lattice units a uniform flow has the same (rho, u) on both levels (dt/dh is the same), so
    # exploded copies and coalesced averages of feq are feq again; a missing 1/2^D in the average,
    # a wrong population or a collided transition-to-coarser layer breaks the fixed point.
    N, M, K = 12, 10, (8 if dims == 3 else 1)
    lims = (3, 7, 2, 6) + ((2, 5) if dims == 3 else ())
    U = np.array([0.03, -0.02, 0.01][:dims])
    rho = np.full((N, M, K), 1.01)
    u = np.broadcast_to(U, (N, M, K, dims)).copy()
    op, oc, parent, child = _two_level(dims, N, M, K, lims, rho, u)
    oc.init_feq(np.full((child.N, child.M, child.K), 1.01), np.broadcast_to(U, (child.N, child.M, child.K, dims)).copy())
    fp0, fc0 = op.get_f(), oc.get_f()
    for _ in range(4):
        op.step()
    live_p = np.isin(parent.lattyp, [FLUID, setups.TL2F])
    assert np.abs(op.get_f()[live_p] - fp0[live_p]).max() < 1e-15
    assert np.abs(oc.get_f() - fc0).max() < 1e-15


def _random_two_level(dims, N, M, K, lims, seed, shift=(0, 0, 0)):
    rng = np.random.default_rng(seed)
    rho = 1 + 0.01 * rng.standard_normal((N, M, K))
    u = 0.03 * rng.standard_normal((N, M, K, dims))
    rho = np.roll(rho, shift, axis=(0, 1, 2))
    u = np.roll(u, shift, axis=(0, 1, 2))
    l = list(lims)
    for a in range(dims):
        l[2 * a] += shift[a]
        l[2 * a + 1] += shift[a]
    return _two_level(dims, N, M, K, tuple(l), rho, u)


@pytest.mark.parametrize("dims", [2, 3])
def test_two_level_step_is_translation_covariant(dims):
    # Moving the patch and the initial state together by whole coarse cells (periodic box) moves
    # the result on both levels, bit for bit.  An explode or coalesce that reads absolute instead
    # of patch-relative indices (CoarseLims offsets, getCoarseIndices/getFineIndices) does not.
    N, M, K = 12, 10, (8 if dims == 3 else 1)
    lims = (3, 6, 2, 5) + ((2, 4) if dims == 3 else ())
    shift = (2, 3, 1 if dims == 3 else 0)
    a = _random_two_level(dims, N, M, K, lims, 11)
    b = _random_two_level(dims, N, M, K, lims, 11, shift)
    for _ in range(3):
        a[0].step()
        b[0].step()
    pa, pb = a[0].get_f(), b[0].get_f()
    live = np.isin(a[2].lattyp, [FLUID, setups.TL2F])
    assert np.array_equal(np.roll(pa, shift, axis=(0, 1, 2))[np.roll(live, shift, axis=(0, 1, 2))],
                          pb[np.roll(live, shift, axis=(0, 1, 2))])
    assert np.array_equal(a[1].get_f(), b[1].get_f())


def test_two_level_step_is_mirror_symmetric():
    # Reflecting x -> -x (cells i -> N-1-i, populations with c_x negated, patch mirrored) maps the
    # two-level solution onto itself up to summation order.  An explode that rounds the fine
    # index the wrong way, or a coalesce that averages the cluster one cell off, is not symmetric.
    N, M = 14, 10
    lims = (4, 8, 3, 6)
    rng = np.random.default_rng(7)
    rho = 1 + 0.01 * rng.standard_normal((N, M, 1))
    u = 0.03 * rng.standard_normal((N, M, 1, 2))
    mirror_v = [1, 0, 2, 3, 7, 6, 5, 4, 8]     # D2Q9 in LUMA's order with c_x negated
    a = _two_level(2, N, M, 1, lims, rho, u)
    rho_m = rho[::-1].copy()
    u_m = u[::-1].copy()
    u_m[..., 0] *= -1
    b = _two_level(2, N, M, 1, (N - 1 - lims[1], N - 1 - lims[0], lims[2], lims[3]), rho_m, u_m)
    for _ in range(3):
        a[0].step()
        b[0].step()
    live = np.isin(a[2].lattyp, [FLUID, setups.TL2F])
    fa, fb = a[0].get_f(), b[0].get_f()[::-1][..., mirror_v]
    assert np.allclose(fa[live], fb[live], rtol=0, atol=1e-15)
    ca, cb = a[1].get_f(), b[1].get_f()[::-1][..., mirror_v]
    assert np.allclose(ca, cb, rtol=0, atol=1e-15)
    assert np.abs(ca - a[1].get_f()[::-1][..., mirror_v]).max() > 1e-4   # the state is not symmetric itself


@pytest.mark.parametrize("dims", [2, 3])
def test_two_regions_each_step_as_if_alone(dims):
    # Two refined regions of one grid (L_NUM_REGIONS = 2, GridUtils::getSubGrid) far enough apart
    # that in three coarse steps neither reaches the other: each region's fine state, and the
    # coarse state around it, equal those of a run with that region alone, bit for bit.  A
    # transition site that coalesced from, or copied into, the wrong region breaks this.
    N, M, K = (30, 12, 1) if dims == 2 else (26, 10, 8)
    la = (3, 7, 2, 6) + ((2, 5) if dims == 3 else ())
    lb = (N - 9, N - 5, 3, 7) + ((1, 4) if dims == 3 else ())
    rng = np.random.default_rng(5)
    rho = 1 + 0.01 * rng.standard_normal((N, M, K))
    u = 0.03 * rng.standard_normal((N, M, K, dims))

    def hierarchy(regions):
        parent = setups.taylor_green((N, M, K), u0=0.0, tau=0.7, dims=dims)
        parent.xpos, parent.ypos = np.arange(N) + 0.5, np.arange(M) + 0.5
        parent.zpos = (np.arange(K) + 0.5) if dims == 3 else np.zeros(1)
        kids = [setups.add_refined_region(parent, l, 1.0, 1.0) for l in regions]

        def mk(c, r, v):
            g = OracleGrid(dims, parent.Q, c.N, c.M, c.K, omega=c.omega, regularised=False, fz_compat=False,
                           level=c.info.get("level", 0), coarse_lims=c.info.get("coarse_lims", (0,) * 6))
            g.set_labels(c.lattyp)
            g.set_positions(c.xpos, c.ypos, c.zpos)
            g.init_feq(r, v)
            return g
        op = mk(parent, rho, u)
        oks = []
        for c in kids:
            crng = np.random.default_rng(c.info["coarse_lims"][0])
            oc = mk(c, 1 + 0.01 * crng.standard_normal((c.N, c.M, c.K)), 0.03 * crng.standard_normal((c.N, c.M, c.K, dims)))
            op.add_child(oc)
            oks.append(oc)
        for _ in range(3):
            op.step()
        return op, oks, parent

    both, kids2, parent2 = hierarchy([la, lb])
    for idx, lims in ((0, la), (1, lb)):
        alone, kids1, parent1 = hierarchy([lims])
        assert np.array_equal(kids2[idx].get_f(), kids1[0].get_f()), idx
        # coarse sites within two cells of the region, outside both regions' refined interiors
        near = np.zeros((N, M, K), bool)
        sl = [slice(max(lims[2 * a] - 2, 0), lims[2 * a + 1] + 3) for a in range(dims)] + ([slice(None)] if dims == 2 else [])
        near[tuple(sl)] = True
        live = near & np.isin(parent2.lattyp, [FLUID, setups.TL2F])
        assert live.sum() > 20
        assert np.array_equal(both.get_f()[live], alone.get_f()[live]), idx
    assert not np.array_equal(kids2[0].get_f()[:4, :4], kids2[1].get_f()[:4, :4])


# ---- corotational beam at large rotation: FEMElement.cpp:273-302, FEMBody.cpp:183-325 -------

# Bisshopp & Drucker (1945), Q. Appl. Math. 3:272, cantilever under a dead tip load P,
# alpha = P L^2 / EI: tip deflection delta_v / L and horizontal shortening delta_h / L
ELASTICA_TABLE = {1.0: (0.30172, 0.05643), 3.0: (0.60325, 0.25442), 10.0: (0.81061, 0.55500)}


def elastica(alpha):
    """Shooting solution of EI theta'' = -P cos(theta), theta(0) = 0, theta'(L) = 0 (L = EI = 1)."""
    def rhs(s, y):
        return [y[1], -alpha * np.cos(y[0]), np.cos(y[0]), np.sin(y[0])]

    def end(k0):
        return solve_ivp(rhs, (0, 1), [0, k0, 0, 0], rtol=1e-12, atol=1e-14).y[:, -1]
    k0 = brentq(lambda k: end(k)[1], 0.0, alpha * 1.0001, xtol=1e-15)
    th, _, x, y = end(k0)
    return y, 1 - x, th


@pytest.mark.parametrize("alpha", [1.0, 3.0, 10.0])
def test_cantilever_large_deflection_matches_the_elastica(alpha):
    # Dead tip load raised over 40 quasi-static steps (Newmark with dt = 1e6: inertia vanishes and
    # each step is a Newton solve of F_int(U) = R).  At alpha = 10 the tip turns 82 degrees, so a
    # small-rotation internal force, a wrong corotated angle or a stale element length misses the
    # elastica by far more than the discretisation error (< 7e-4 with 20 and 40 elements).
    yv, xh, th = elastica(alpha)
    tv, th_tab = ELASTICA_TABLE[alpha]
    assert yv == pytest.approx(tv, abs=1e-5) and xh == pytest.approx(th_tab, abs=1e-5)
    L, H, E = 1.0, 0.02, 1e5
    EI = E * H ** 3 / 12.0        # depth = dh = 1
    err = {}
    for nel in (20, 40):
        f = OracleFEM((0.0, 0.0, 0.0), L, H, 0.0, nel, True, 10.0, E, dh=1.0, dt=1e6, dm=1.0)
        for s in range(1, 41):
            R = np.zeros(f.n)
            R[-2] = alpha * EI / L ** 2 * s / 40
            assert 1 <= f.step(R) < 20
        U, _ = f.state()
        f.close()
        err[nel] = (U[-2] / yv - 1, -U[-3] / xh - 1, U[-1] / th - 1)
        assert max(abs(e) for e in err[nel]) < 1e-3, err[nel]
    assert abs(err[40][0]) < abs(err[20][0])


@pytest.mark.parametrize("phi", [0.3, 1.0])
def test_rigid_rotation_moves_markers_by_the_hermite_interpolant(phi):
    # A rigid rotation by phi about the clamped node: nodal displacements (R - I)(x - x0), nodal
    # rotations phi.  In the rotated element frame the axial displacement is linear and exact; the
    # transverse one is s sin(phi) at the nodes but the Hermite cubic takes the nodal slope phi, so
    # a marker at element coordinate h = (zeta + 1)/2 is off the rotated line by
    # (phi - sin phi) l h (1 - h)(1 - 2h) along the rotated normal.  Markers at nodes and element
    # midpoints (r = 1/2: 2 markers per element) therefore rotate exactly; C5's 5 markers per
    # element show the defect term.  A wrong parent element or a flipped zeta fails both.
    x0, angle = np.array([2.0, 3.0]), 20.0
    a = math.radians(angle)
    R = np.array([[math.cos(phi), -math.sin(phi)], [math.sin(phi), math.cos(phi)]])
    for nel, per_el in ((10, 2), (40, 5)):
        nib = nel * per_el + 1
        L = float(nib - 1)              # dh = 1: markers at s = i dh
        f = OracleFEM((x0[0], x0[1], 0.0), L, 0.05, angle, nel, True, 10.0, 1e3, dh=1.0, dt=1.0, dm=1.0)
        t0 = np.array([math.cos(a), math.sin(a)])
        nodes = x0 + np.outer(np.arange(nel + 1) * (L / nel), t0)
        U = np.zeros(f.n)
        d = (nodes - x0) @ R.T - (nodes - x0)
        U[0::3], U[1::3], U[2::3] = d[:, 0], d[:, 1], phi
        f.set_state(U)
        s = np.arange(nib, dtype=float)
        pos0 = np.zeros((nib, 3))
        pos0[:, :2] = x0 + np.outer(s, t0)
        pos, _ = f.markers(pos0)
        f.close()
        l = L / nel
        h = (s / l) - np.floor(s / l)
        nrm = R @ np.array([-math.sin(a), math.cos(a)])
        want = x0 + (pos0[:, :2] - x0) @ R.T + np.outer((phi - math.sin(phi)) * l * h * (1 - h) * (1 - 2 * h), nrm)
        assert np.abs(pos[:, :2] - want).max() < 1e-12 * L
        if per_el == 2:
            assert np.abs(pos[:, :2] - (x0 + (pos0[:, :2] - x0) @ R.T)).max() < 1e-12 * L


# ---- FSI sub-iteration: ObjectManager_ops_ibm.cpp:111-185, :988-1032 -----------------------

def _filament_case(E_scale, flexible=True, steps=6):
    c = setups.channel_2d(48, 24, 8.0, 0.04, 100.0)
    dh, dt, dm = c.info["dh"], c.info["dt"], c.info["dm"]
    length, height = 1.0, 0.05
    density, E = setups.filament_material(length, height, dh, mass_ratio=10.0, bending=1e-2)
    start = (1.2, 1.5 + 0.013, 0.0)
    g = OracleGrid(c.dims, c.Q, c.N, c.M, c.K, omega=c.omega, regularised=c.regularised, forced=c.forced,
                   rho_in=c.rho_in, edges=c.edges, wall_thk=c.wall_thk)
    g.set_labels(c.lattyp)
    g.set_positions(c.xpos, c.ypos, c.zpos)
    g.set_inlet(c.ux_in, c.uy_in, c.uz_in)
    g.init_feq(c.rho, c.u)
    ib = OracleIBM(g, dh)
    if flexible:
        ib.add_filament(start, length, height, dh, 8.0, 8, True, density, E * E_scale, dt, dm)
    else:
        ib.add_body(setups.filament_markers(start, length, 8.0, dh))
    ib.initialise()
    its, res = [], []
    for _ in range(steps):
        g.step()
        i, r = ib.subiterations()
        its.append(i)
        res.append(r)
    assert ib.error() == 0
    return g, ib, its, res, c


def test_stiff_filament_takes_one_subiteration_and_tends_to_the_rigid_body():
    # E x 1e9 and x 1e12: the beam barely moves, the relaxed marker velocities settle in the first
    # pass (residual below 1e-4), and the flow departs from the rigid body's (the same markers held
    # still) by O(1/E): a thousandfold stiffer beam is a thousandfold closer.
    gr, ibr, _, _, _ = _filament_case(1.0, flexible=False)
    _, ur = gr.get_macro()
    dev = []
    for scale in (1e9, 1e12):
        gf, ibf, its, res, c = _filament_case(scale)
        assert its == [1] * len(its) and max(res) < 1e-4
        _, uf = gf.get_macro()
        m = c.lattyp == FLUID
        dev.append(np.abs(uf[m] - ur[m]).max() / np.abs(ur[m]).max())
        assert np.abs(ibf.markers(0)["pos"] - ibr.markers(0)["pos"]).max() < 1e-7
    assert dev[1] < 1e-7
    assert dev[0] / dev[1] == pytest.approx(1e3, rel=0.05)


def test_subiteration_residual_is_the_max_marker_velocity_change():
    # The soft filament needs more than one pass; the residual the loop stopped on is
    # max_m |U_m - U_m^(k-1)| / (L_UX0 dt / dh) over the last pass (L_UX0 = 1), and the loop stops
    # at the first pass below 1e-4 or at 10.
    g, ib, its, res, c = _filament_case(1.0, steps=4)
    assert max(its) >= 2
    mk = ib.markers(0)
    dt, dh = c.info["dt"], c.info["dh"]
    want = np.max(np.linalg.norm(mk["vel"] - mk["vel_km1"], axis=1)) / (1.0 * dt / dh)
    assert res[-1] == pytest.approx(want, rel=1e-12)
    assert (res[-1] <= 1e-4) == (its[-1] < 10) or its[-1] == 10


# ---- threaded oracle (CPU baseline mode) ------------------------------------------------------

def test_threaded_run_matches_serial_on_a_labelled_grid():
    # ora_run steps labelled grids serially whatever the thread count, so the CPU baseline mode
    # cannot race on regularised edge/corner sites or ramped forced-equilibrium sources
    c = setups.lid_driven_cavity(24, re=100.0, u_lid_lbm=0.1)
    runs = []
    for nt in (1, 4):
        g = OracleGrid(c.dims, c.Q, c.N, c.M, c.K, omega=c.omega, regularised=c.regularised, forced=c.forced,
                       rho_in=c.rho_in, edges=c.edges, wall_thk=c.wall_thk)
        g.set_labels(c.lattyp)
        g.set_positions(c.xpos, c.ypos, c.zpos)
        g.set_inlet(c.ux_in, c.uy_in, c.uz_in)
        g.init_feq(c.rho, c.u)
        g.run(25, nthreads=nt)
        runs.append(g.get_f())
    assert np.array_equal(runs[0], runs[1])
    # and the all-fluid periodic box runs in parallel with the same arithmetic
    tg = setups.taylor_green((16, 12, 10), dims=3)
    out = []
    for nt in (1, 4):
        g = OracleGrid(3, 19, 16, 12, 10, omega=tg.omega)
        g.init_feq(tg.rho, tg.u)
        g.run(5, nthreads=nt)
        out.append(g.get_f())
    assert np.array_equal(out[0], out[1])
