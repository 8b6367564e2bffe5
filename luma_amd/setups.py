"""Initial states and labels of the benchmark cases, built the way LUMA builds them.

Everything here is host-side set-up in LUMA's layouts (arrays shaped (N, M, K[, D]) in C
order, i.e. id = k + j*K + i*K*M).  It restates the initialisation code the reference runs
before the time loop:
  positions       GridObj_init_grids.cpp:234-240 (GridUtils::linspace, GridUtils.h:289-309)
  wall labels     LBM_initBoundLab, GridObj_init_grids.cpp:983-1097, precedence :1372-1377
  inlet profile   _LBM_initSetInletProfile, :1322-1360
  velocity        LBM_initVelocity (L_NO_FLOW), :37-133; density LBM_initRho, :138-151
  omega           GridObj_init_grids.cpp:336-344 with GridUnits::nud2nulbm (GridUnits.h:128-131)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

SOLID, FLUID, REFINED, TL2C, TL2F, BFL, VELOCITY, PRESSURE, SLIP, EXTRAP_RIGHT = range(10)
CS = 1.0 / math.sqrt(3.0)

# velocity sets in LUMA's order (stdafx.cpp:81-153); rest particle last
_C9 = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (0, 0, 0)]
_W9 = [1 / 9] * 4 + [1 / 36] * 4 + [4 / 9]
_C19 = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, -1, 0), (1, -1, 0),
        (-1, 1, 0), (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1), (1, 0, 1), (-1, 0, -1), (-1, 0, 1), (1, 0, -1),
        (0, 0, 0)]
_W19 = [1 / 18] * 6 + [1 / 36] * 12 + [1 / 3]


def lattice_c(Q: int, v: int):
    return (_C9 if Q == 9 else _C19)[v]


def lattice_w(Q: int, v: int) -> float:
    return (_W9 if Q == 9 else _W19)[v]


@dataclass
class Case:
    name: str
    dims: int
    Q: int
    N: int
    M: int
    K: int
    omega: float
    rho: np.ndarray
    u: np.ndarray
    lattyp: Optional[np.ndarray] = None
    xpos: Optional[np.ndarray] = None
    ypos: Optional[np.ndarray] = None
    zpos: Optional[np.ndarray] = None
    ux_in: Optional[np.ndarray] = None
    uy_in: Optional[np.ndarray] = None
    uz_in: Optional[np.ndarray] = None
    force: Optional[np.ndarray] = None
    regularised: bool = True
    forced: bool = False
    fz_compat: bool = True
    rho_in: float = 1.0
    edges: tuple = (1.0, 1.0, 1.0)
    wall_thk: tuple = (0.0,) * 6
    info: dict = field(default_factory=dict)

    @property
    def ncell(self):
        return self.N * self.M * self.K


def linspace(lo: float, hi: float, n: int) -> np.ndarray:
    """GridUtils::linspace: lo + spacing*i with spacing = (hi - lo)/(n - 1)."""
    n = max(n, 2)
    spacing = (hi - lo) / float(n - 1)
    return lo + spacing * np.arange(n, dtype=np.float64)


def omega_from_re(re: float, dt: float, dh: float) -> float:
    nu = ((1.0 / re) * dt) / (dh * dh)
    return 1.0 / ((nu / (CS * CS)) + 0.5)


def bc_precedence(current: int, desired: int) -> int:
    if current == SOLID or desired == SOLID:
        return SOLID
    if current == VELOCITY:
        return VELOCITY
    return desired


def label_walls(N, M, K, dims, xpos, ypos, zpos, edges, thk, walls):
    """LBM_initBoundLab: walls = dict(left, right, bottom, top[, front, back]) of eType."""
    lt = np.full((N, M, K), FLUID, dtype=np.int32)
    prec = np.vectorize(bc_precedence, otypes=[np.int32])

    def apply(mask_axis, idx, wall):
        sl = [slice(None)] * 3
        sl[mask_axis] = idx
        lt[tuple(sl)] = prec(lt[tuple(sl)], wall)

    for i in range(N):
        if xpos[i] <= thk[0]:
            apply(0, i, walls["left"])
    for i in range(N):
        if xpos[i] >= edges[0] - thk[1]:
            apply(0, i, walls["right"])
    if dims == 3:
        for k in range(K):
            if zpos[k] <= thk[4]:
                apply(2, k, walls["front"])
        for k in range(K):
            if zpos[k] >= edges[2] - thk[5]:
                apply(2, k, walls["back"])
    for j in range(M):
        if ypos[j] <= thk[2]:
            apply(1, j, walls["bottom"])
    for j in range(M):
        if ypos[j] >= edges[1] - thk[3]:
            apply(1, j, walls["top"])
    return lt


def taylor_green(n: int | tuple, u0: float = 0.05, tau: float = 0.55, dims: int = 3,
                 rho0: float = 1.0) -> Case:
    """Triply (doubly) periodic Taylor-Green vortex, cell centres at x = i.

    u = u0 (sin kx cos ky cos kz, -cos kx sin ky cos kz, 0), k = 2 pi / n per axis;
    rho = rho0 + (p - p0)/cs^2, p - p0 = rho0 u0^2/16 (cos 2kx + cos 2ky)(cos 2kz + 2) in 3D
    and rho0 u0^2/4 (cos 2kx + cos 2ky) in 2D.
    """
    if isinstance(n, int):
        n = (n, n, n) if dims == 3 else (n, n, 1)
    N, M, K = n
    kx, ky = 2 * math.pi / N, 2 * math.pi / M
    x = np.arange(N, dtype=np.float64)[:, None, None]
    y = np.arange(M, dtype=np.float64)[None, :, None]
    if dims == 3:
        kz = 2 * math.pi / K
        z = np.arange(K, dtype=np.float64)[None, None, :]
        ux = u0 * np.sin(kx * x) * np.cos(ky * y) * np.cos(kz * z)
        uy = -u0 * np.cos(kx * x) * np.sin(ky * y) * np.cos(kz * z)
        uz = np.zeros((N, M, K))
        dp = rho0 * u0 * u0 / 16.0 * (np.cos(2 * kx * x) + np.cos(2 * ky * y)) * (np.cos(2 * kz * z) + 2.0)
        u = np.stack(np.broadcast_arrays(ux, uy, uz), axis=-1)
        Q = 19
    else:
        ux = u0 * np.sin(kx * x) * np.cos(ky * y)
        uy = -u0 * np.cos(kx * x) * np.sin(ky * y)
        dp = rho0 * u0 * u0 / 4.0 * (np.cos(2 * kx * x) + np.cos(2 * ky * y))
        u = np.stack(np.broadcast_arrays(ux, uy), axis=-1)
        Q = 9
    rho = np.broadcast_to(rho0 + dp / (CS * CS), (N, M, K)).copy()
    u = np.ascontiguousarray(u, dtype=np.float64)
    return Case(name=f"tgv{dims}d_{N}x{M}x{K}", dims=dims, Q=Q, N=N, M=M, K=K, omega=1.0 / tau,
                rho=rho, u=u, info={"u0": u0, "tau": tau, "k": (kx, ky)})


def lid_driven_cavity(n: int = 256, re: float = 1280.0, u_lid_lbm: float = 0.1,
                      regularised: bool = True) -> Case:
    """D2Q9 lid-driven cavity as LUMA sets it up.

    L_RESOLUTION = n, L_BX = L_BY = 1 (dh = 1/n), L_TIMESTEP = u_lid_lbm*dh with L_UX0 = 1,
    L_RE = re (n = 256, re = 1280 gives tau = 0.56).  Left, right and bottom walls are
    eSolid, the top wall eVelocity; wall thickness one cell; L_NO_FLOW initial state.
    """
    N = M = n
    K = 1
    dh = 1.0 / n
    dt = u_lid_lbm * dh
    edges = (1.0, 1.0, 0.0)
    thk = (dh, dh, dh, dh, dh, dh)
    xpos = linspace(dh / 2.0, edges[0] - dh / 2.0, N)
    ypos = linspace(dh / 2.0, edges[1] - dh / 2.0, M)
    zpos = np.array([0.0])
    lt = label_walls(N, M, K, 2, xpos, ypos, zpos, edges, thk,
                     dict(left=SOLID, right=SOLID, bottom=SOLID, top=VELOCITY))
    def inlet_fn(y, dt_, dh_):
        # uniform profile from L_UX0 = 1: GridUnits::ud2ulbm = (u * dt) / dh
        return np.full(len(y), (1.0 * dt_) / dh_), np.zeros(len(y)), np.zeros(len(y))

    ux_in, uy_in, uz_in = inlet_fn(ypos, dt, dh)
    u_lbm = ux_in[0]
    u = np.zeros((N, M, K, 2))
    vel = lt == VELOCITY
    u[..., 0][vel] = np.broadcast_to(ux_in[None, :, None], (N, M, K))[vel]
    rho = np.ones((N, M, K))
    return Case(name=f"cavity_{n}", dims=2, Q=9, N=N, M=M, K=K, omega=omega_from_re(re, dt, dh),
                rho=rho, u=u, lattyp=lt, xpos=xpos, ypos=ypos, zpos=zpos, ux_in=ux_in, uy_in=uy_in,
                uz_in=uz_in, regularised=regularised, edges=edges, wall_thk=thk,
                info={"re": re, "dh": dh, "dt": dt, "u_lid": u_lbm, "inlet_fn": inlet_fn})


def velocity_corner_box(N: int, M: int, K: int = 1, dims: int = 2, u_in: float = 0.02, u_lid: float = 0.05,
                        tau: float = 0.6) -> Case:
    """Box whose left, right and top walls (and front wall in 3D) are regularised velocity
    walls, so velocity walls meet at edges and corners (GridObj_ops_lbm_optimised.cpp:346-364);
    bottom (and back) walls are solid, which wins at the corners it shares (:1374).
    ux_in[j] = u_in below the lid row and u_lid on it; one-cell walls, cell-centred positions."""
    Q = 9 if dims == 2 else 19
    dh = 1.0 / M
    edges = (N * dh, M * dh, K * dh if dims == 3 else 0.0)
    thk = (dh,) * 6
    xpos = linspace(dh / 2.0, edges[0] - dh / 2.0, N)
    ypos = linspace(dh / 2.0, edges[1] - dh / 2.0, M)
    zpos = linspace(dh / 2.0, edges[2] - dh / 2.0, K) if dims == 3 else np.array([0.0])
    walls = dict(left=VELOCITY, right=VELOCITY, bottom=SOLID, top=VELOCITY, front=VELOCITY, back=SOLID)
    lt = label_walls(N, M, K, dims, xpos, ypos, zpos, edges, thk, walls)
    ux_in = np.full(M, u_in)
    ux_in[M - 1] = u_lid
    uy_in = np.zeros(M)
    uz_in = np.zeros(M)
    u = np.zeros((N, M, K, dims))
    vel = lt == VELOCITY
    u[..., 0][vel] = np.broadcast_to(ux_in[None, :, None], (N, M, K))[vel]
    rho = np.ones((N, M, K))
    return Case(name=f"velocity_corners{dims}d_{N}x{M}x{K}", dims=dims, Q=Q, N=N, M=M, K=K, omega=1.0 / tau,
                rho=rho, u=u, lattyp=lt, xpos=xpos, ypos=ypos, zpos=zpos, ux_in=ux_in, uy_in=uy_in,
                uz_in=uz_in, regularised=True, edges=edges, wall_thk=thk, info={"dh": dh})


def uniform_inflow_3d(N: int, M: int, K: int, U: float, tau: float, dh: float = 1.0) -> Case:
    """D3Q19 box with a uniform inflow: the x = 0 plane is an eVelocity inlet with forced
    equilibrium (L_REGULARISED_BOUNDARIES off, GridObj_ops_lbm_optimised.cpp:254-270) at
    u = (U, 0, 0), the x = N-1 plane eExtrapolateRight, periodic in y and z; uniform initial
    flow U, rho = 1; cell-centred positions with spacing dh; forced (IB) grid."""
    lt = np.full((N, M, K), FLUID, dtype=np.int32)
    lt[0] = VELOCITY
    lt[N - 1] = EXTRAP_RIGHT
    u = np.zeros((N, M, K, 3))
    u[..., 0] = U
    rho = np.ones((N, M, K))
    edges = (N * dh, M * dh, K * dh)
    return Case(name=f"inflow3d_{N}x{M}x{K}", dims=3, Q=19, N=N, M=M, K=K, omega=1.0 / tau, rho=rho, u=u,
                lattyp=lt, xpos=(np.arange(N) + 0.5) * dh, ypos=(np.arange(M) + 0.5) * dh,
                zpos=(np.arange(K) + 0.5) * dh, ux_in=np.full(M, U), uy_in=np.zeros(M), uz_in=np.zeros(M),
                regularised=False, forced=True, edges=edges, wall_thk=(0.0,) * 6,
                info={"dh": dh, "origin": (0.0, 0.0, 0.0), "U": U})


def pressure_outlet_duct(N: int, M: int, K: int = 1, dims: int = 2, U: float = 0.03, tau: float = 0.8,
                         rho_out: float = 1.0, dh: float = 1.0, forced: bool = False) -> Case:
    """Duct periodic in y (and z): a regularised eVelocity inlet at x = 0 with u = (U, 0, 0) and a
    regularised ePressure outlet at x = N-1 holding rho = rho_out (L_RHOIN, GridObj_ops_lbm_optimised.cpp
    :333-343, :389-402); one-cell wall thicknesses in x put both planes on faces (isWithinDomainWall).
    Uniform initial flow U, rho = 1."""
    Q = 9 if dims == 2 else 19
    lt = np.full((N, M, K), FLUID, dtype=np.int32)
    lt[0] = VELOCITY
    lt[N - 1] = PRESSURE
    u = np.zeros((N, M, K, dims))
    u[..., 0] = U
    rho = np.ones((N, M, K))
    edges = (N * dh, M * dh, K * dh if dims == 3 else 0.0)
    return Case(name=f"pressure{dims}d_{N}x{M}x{K}", dims=dims, Q=Q, N=N, M=M, K=K, omega=1.0 / tau, rho=rho, u=u,
                lattyp=lt, xpos=(np.arange(N) + 0.5) * dh, ypos=(np.arange(M) + 0.5) * dh,
                zpos=(np.arange(K) + 0.5) * dh if dims == 3 else np.array([0.0]), ux_in=np.full(M, U),
                uy_in=np.zeros(M), uz_in=np.zeros(M), regularised=True, forced=forced, rho_in=rho_out,
                edges=edges, wall_thk=(dh, dh, 0.0, 0.0, 0.0, 0.0),
                info={"dh": dh, "origin": (0.0, 0.0, 0.0), "U": U})


def body_force_channel(N: int, M: int, K: int = 1, dims: int = 2, tau: float = 0.8,
                       g: float = 1e-6) -> Case:
    """Channel with solid walls at j = 0 and j = M-1, periodic in x (and z), driven by a
    uniform body force g along x (Poiseuille flow)."""
    Q = 9 if dims == 2 else 19
    lt = np.full((N, M, K), FLUID, dtype=np.int32)
    lt[:, 0, :] = SOLID
    lt[:, M - 1, :] = SOLID
    F = np.zeros((N, M, K, dims))
    F[..., 0] = np.where(lt == FLUID, g, 0.0)
    rho = np.ones((N, M, K))
    u = np.zeros((N, M, K, dims))
    return Case(name=f"poiseuille{dims}d_{N}x{M}x{K}", dims=dims, Q=Q, N=N, M=M, K=K, omega=1.0 / tau,
                rho=rho, u=u, lattyp=lt, force=F, forced=True, regularised=True,
                info={"g": g, "tau": tau})


def position_vector(start: float, dh: float, n: int, end_lim: float, start_lim: float) -> np.ndarray:
    """GridObj::LBM_initPositionVector (GridObj_init_grids.cpp:902-965): sequential sums of dh,
    wrapping to start_lim + dh/2 past end_lim."""
    arr = [start]
    while len(arr) < n:
        nxt = arr[-1] + dh
        arr.append(start_lim + dh / 2.0 if nxt > end_lim else nxt)
    return np.array(arr, dtype=np.float64)


def add_refined_region(parent: Case, lims, dh_parent: float, dt_parent: float, parent_origin=(0.0, 0.0, 0.0),
                       level: int = 1) -> Case:
    """Attach a 2:1 refined region covering parent cells lims = (i0, i1, j0, j1[, k0, k1]).

    Restates LBM_initRefinedLab (GridObj_init_grids.cpp:1106-1224) with transition layers on
    every side: parent sites on the region's outer ring become eTransitionToFiner, the rest
    eRefined (fluid sites only); each child site takes eTransitionToCoarser under a TL2F parent,
    eFluid under a refined one, and the parent's label otherwise.  Sizes, positions, omega and
    the initial state follow LBM_initSubGrid (:390-536) and LBM_initGridToGridMappings (:545-).
    The parent's labels are modified in place.
    """
    D = parent.dims
    lims = list(lims) + ([0, 0] if D == 2 else [])
    i0, i1, j0, j1, k0, k1 = lims
    if parent.lattyp is None:
        parent.lattyp = np.full((parent.N, parent.M, parent.K), FLUID, np.int32)
    if parent.xpos is None:
        parent.xpos = parent_origin[0] + dh_parent / 2 + dh_parent * np.arange(parent.N)
        parent.ypos = parent_origin[1] + dh_parent / 2 + dh_parent * np.arange(parent.M)
        parent.zpos = (parent_origin[2] + dh_parent / 2 + dh_parent * np.arange(parent.K)) if D == 3 else np.array([0.0])
    px, py, pz = parent.xpos, parent.ypos, parent.zpos
    edges = [px[i0] - dh_parent / 2, px[i1] + dh_parent / 2, py[j0] - dh_parent / 2, py[j1] + dh_parent / 2,
             (pz[k0] - dh_parent / 2) if D == 3 else 0.0, (pz[k1] + dh_parent / 2) if D == 3 else 0.0]
    lt = parent.lattyp
    for i in range(parent.N):
        for j in range(parent.M):
            for k in range(parent.K):
                inside = edges[0] < px[i] < edges[1] and edges[2] < py[j] < edges[3]
                if D == 3:
                    inside = inside and edges[4] < pz[k] < edges[5]
                if not inside:
                    continue
                tl = (px[i] < edges[0] + dh_parent or px[i] > edges[1] - dh_parent or
                      py[j] < edges[2] + dh_parent or py[j] > edges[3] - dh_parent)
                if D == 3:
                    tl = tl or pz[k] < edges[4] + dh_parent or pz[k] > edges[5] - dh_parent
                if lt[i, j, k] == FLUID:
                    lt[i, j, k] = TL2F if tl else REFINED
    dh = dh_parent / 2.0
    dt = dt_parent / 2.0
    N, M = 2 * (i1 - i0 + 1), 2 * (j1 - j0 + 1)
    K = 2 * (k1 - k0 + 1) if D == 3 else 1
    xpos = position_vector(px[i0] - dh / 2.0, dh, N, edges[1], edges[0])
    ypos = position_vector(py[j0] - dh / 2.0, dh, M, edges[3], edges[2])
    zpos = position_vector(pz[k0] - dh / 2.0, dh, K, edges[5], edges[4]) if D == 3 else np.array([0.0])
    clt = np.empty((N, M, K), np.int32)
    for i in range(N):
        for j in range(M):
            for k in range(K):
                pl = lt[(i // 2) + i0, (j // 2) + j0, ((k // 2) + k0) if D == 3 else 0]
                clt[i, j, k] = TL2C if pl == TL2F else (FLUID if pl == REFINED else pl)
    omega = 1.0 / (((1.0 / parent.omega - 0.5) * 2.0) + 0.5)
    child = Case(name=parent.name + f"_L{level}", dims=D, Q=parent.Q, N=N, M=M, K=K, omega=omega,
                 rho=np.full((N, M, K), parent.rho_in), u=np.zeros((N, M, K, D)), lattyp=clt,
                 xpos=xpos, ypos=ypos, zpos=zpos, regularised=parent.regularised, forced=parent.forced,
                 fz_compat=parent.fz_compat, rho_in=parent.rho_in, edges=parent.edges, wall_thk=parent.wall_thk,
                 info={"level": level, "coarse_lims": (i0, i1, j0, j1, k0, k1), "dh": dh, "dt": dt,
                       "region_edges": edges, "origin": (edges[0], edges[2], edges[4]),
                       "dm": parent.info.get("dm", 1.0) / 8.0})
    if "inlet_fn" in parent.info:
        # _LBM_initSetInletProfile evaluated on the child's positions, dt and dh
        child.ux_in, child.uy_in, child.uz_in = parent.info["inlet_fn"](ypos, dt, dh)
        child.info["inlet_fn"] = parent.info["inlet_fn"]
        vel = clt == VELOCITY
        child.u[..., 0][vel] = np.broadcast_to(child.ux_in[None, :, None], (N, M, K))[vel]
    return child


def velocity_field_on(child: Case, parent: Case, init: str = "interp"):
    """Initial child velocity: copy the parent's u at the enclosing coarse cell."""
    i0, _, j0, _, k0, _ = child.info["coarse_lims"]
    for i in range(child.N):
        for j in range(child.M):
            for k in range(child.K):
                child.u[i, j, k] = parent.u[(i // 2) + i0, (j // 2) + j0, ((k // 2) + k0) if child.dims == 3 else 0]


# ---- immersed-boundary bodies ----------------------------------------------------------

def circle_markers(centre, radius, dh, luma_skip_first=True, n=None):
    """Body<IBMarker> circle (Body.h:230-243): n = floor(2 pi r / dh) angles from linspace(0,
    2 pi - 2 pi/n, n); LUMA keeps angles 1..n-1 (luma_skip_first), giving n - 1 markers."""
    if n is None:
        n = int(math.floor(2.0 * math.pi * radius / dh))
    theta = linspace(0.0, 2.0 * math.pi - (2.0 * math.pi / float(n)), n)
    if luma_skip_first:
        theta = theta[1:]
    return np.stack([centre[0] + radius * np.cos(theta), centre[1] + radius * np.sin(theta),
                     np.full(len(theta), centre[2] if len(centre) > 2 else 0.0)], axis=1)


def sphere_markers(centre, radius, dh, n=None):
    """Body<IBMarker> Fibonacci sphere (Body.h:215-228): n = floor(4 pi r^2 / dh^2) unless given."""
    if n is None:
        n = int(math.floor(4.0 * math.pi * radius ** 2 / dh ** 2))
    inc = math.pi * (3 - math.sqrt(5))
    off = 2.0 / float(np.float32(n))
    pts = []
    for k in range(n):
        y = k * off - 1 + (off / 2)
        r = math.sqrt(1 - y * y)
        phi = k * inc
        pts.append((centre[0] + (math.cos(phi) * r * radius), y * radius + centre[1],
                    centre[2] + (math.sin(phi) * r * radius)))
    return np.array(pts)


def filament_markers(start, length, angle_deg, dh):
    """Body<IBMarker> filament (Body.h:144-187), 2D: floor(L/dh) + 1 markers along angle_deg."""
    n = int(math.floor(length / dh)) + 1
    spacing = length / (n - 1)
    spacing_h = spacing * math.cos(angle_deg * math.pi / 180)
    return np.array([(start[0] + i * spacing_h * math.cos(0.0), start[1] + i * spacing * math.sin(angle_deg * math.pi / 180.0),
                      start[2] + i * spacing_h * math.sin(0.0)) for i in range(n)])


def channel_2d(N: int, M: int, resolution: float, umax_lbm: float, re: float, regularised: bool = True,
               outlet=True, no_flow: bool = False) -> Case:
    """D2Q9 channel as LUMA sets it up: parabolic inlet (eVelocity, L_PARABOLIC_INLET) on the left,
    eExtrapolateRight on the right (outlet=True; False: periodic in x; "pressure": an ePressure
    outlet at L_RHOIN, LUMA's default L_WALL_RIGHT), eSolid top and bottom, one-cell walls.

    resolution = L_RESOLUTION (cells per unit length); umax_lbm = 1.5 * L_UX0 in lattice units;
    re = L_RE based on L_UX0 (the mean velocity) and unit length."""
    dh = 1.0 / resolution
    u0_lbm = umax_lbm / 1.5
    dt = u0_lbm * dh                       # L_UX0 = 1: ud2ulbm(1) = dt/dh
    edges = (N * dh, M * dh, 0.0)
    thk = (dh, dh, dh, dh, dh, dh)
    xpos = linspace(dh / 2.0, edges[0] - dh / 2.0, N)
    ypos = linspace(dh / 2.0, edges[1] - dh / 2.0, M)
    zpos = np.array([0.0])
    lt = label_walls(N, M, 1, 2, xpos, ypos, zpos, edges, thk,
                     dict(left=VELOCITY,
                          right=PRESSURE if outlet == "pressure" else (EXTRAP_RIGHT if outlet else FLUID),
                          bottom=SOLID, top=SOLID))

    def inlet_fn(y, dt_, dh_):
        # _LBM_initSetInletProfile, L_PARABOLIC_INLET (GridObj_init_grids.cpp:1334-1347)
        b = edges[1] - thk[3]
        p = (b + thk[2]) / 2.0
        q = b - p
        ux = ((1.5 * 1.0) * dt_) / dh_ * (1.0 - np.power((y - p) / q, 2.0))
        return ux, np.zeros(len(y)), np.zeros(len(y))

    ux_in, uy_in, uz_in = inlet_fn(ypos, dt, dh)
    u = np.zeros((N, 1 * M, 1, 2))
    if no_flow:
        u[..., 0][lt == VELOCITY] = np.broadcast_to(ux_in[None, :, None], (N, M, 1))[lt == VELOCITY]
    else:
        u[..., 0] = np.broadcast_to(ux_in[None, :, None], (N, M, 1))
    u[lt == SOLID] = 0.0
    rho = np.ones((N, M, 1))
    return Case(name=f"channel_{N}x{M}", dims=2, Q=9, N=N, M=M, K=1, omega=omega_from_re(re, dt, dh),
                rho=rho, u=u, lattyp=lt, xpos=xpos, ypos=ypos, zpos=zpos, ux_in=ux_in, uy_in=uy_in,
                uz_in=uz_in, regularised=regularised, forced=True, edges=edges, wall_thk=thk,
                info={"re": re, "dh": dh, "dt": dt, "inlet_fn": inlet_fn, "origin": (0.0, 0.0, 0.0),
                      "dm": (1000.0 / 1.0) * dh * dh * dh})


def filament_material(length: float, height: float, dh: float, mass_ratio: float = 10.0,
                      bending: float = 1e-3, rho_f: float = 1000.0, u_ref: float = 1.0):
    """Density and Young's modulus of a 2D filament for LUMA's FEM from dimensionless groups:
    density = mass_ratio * rho_f (L_PHYSICAL_RHO), EI = bending * rho_f u_ref^2 L^3 with the
    element's I = depth h^3 / 12 and depth = dh (FEMElement.cpp:79, FEMBody.cpp:103-105)."""
    I = dh * height ** 3 / 12.0
    return mass_ratio * rho_f, bending * rho_f * u_ref ** 2 * length ** 3 / I
