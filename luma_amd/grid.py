"""Python handle over the C ABI: one GPU grid in LUMA's host layouts.

Method names follow GridObj (LUMA/inc/GridObj.h): LBM_multi_opt(subcycle) is the time step.
Arrays are numpy float64 shaped (N, M, K[, Q or D]) in C order, i.e. LUMA's f[v + Q*id].
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import GridDesc, call

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class GpuGrid:
    def __init__(self, dims, Q, N, M, K=1, *, precision="fp64", omega=1.0, regularised=True,
                 forced=False, fz_compat=True, rho_in=1.0, rho0=1.0, edges=(1.0, 1.0, 1.0),
                 wall_thk=(0.0,) * 6, level=0, coarse_lims=(0,) * 6, dh=1.0, origin=(0.0, 0.0, 0.0), dt=1.0, dm=1.0):
        self.lib = _abi.load()
        d = GridDesc()
        d.dims, d.Q, d.N, d.M, d.K = dims, Q, N, M, K
        d.precision = _abi.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        d.regularised, d.forced, d.fz_compat, d.level = int(regularised), int(forced), int(fz_compat), level
        d.omega, d.rho_in, d.rho0 = omega, rho_in, rho0
        d.xmax, d.ymax, d.zmax = edges
        for i in range(6):
            d.wall_thk[i] = wall_thk[i]
            d.coarse_lims[i] = coarse_lims[i]
        d.dh, d.dt, d.dm = dh, dt, dm
        for i in range(3):
            d.origin[i] = origin[i]
        self.desc = d
        self.dims, self.Q, self.shape = dims, Q, (N, M, K)
        h = C.c_void_p()
        call("lumagpu_grid_create", C.byref(d), C.byref(h))
        self.h = h

    @classmethod
    def from_case(cls, case, precision="fp64", store_macro=False):
        g = cls(case.dims, case.Q, case.N, case.M, case.K, precision=precision, omega=case.omega,
                regularised=case.regularised, forced=case.forced, fz_compat=case.fz_compat,
                rho_in=case.rho_in, edges=case.edges, wall_thk=case.wall_thk,
                level=case.info.get("level", 0), coarse_lims=case.info.get("coarse_lims", (0,) * 6),
                dh=case.info.get("dh", 1.0), origin=case.info.get("origin", (0.0, 0.0, 0.0)),
                dt=case.info.get("dt", 1.0), dm=case.info.get("dm", 1.0))
        if case.lattyp is not None:
            g.set_labels(case.lattyp, case.xpos, case.ypos, case.zpos)
        if case.ux_in is not None:
            g.set_inlet(case.ux_in, case.uy_in, case.uz_in)
        if case.force is not None:
            g.set_force(case.force)
        g.init_equilibrium(case.rho, case.u)
        g.set_store_macro(store_macro)
        return g

    # lifetime -------------------------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.lumagpu_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_child(self, child: "GpuGrid"):
        """GridObj::LBM_addSubGrid: the child is stepped twice inside every step of this grid."""
        call("lumagpu_grid_add_child", self.h, child.h)
        self._kids = getattr(self, "_kids", []) + [child]

    # state ----------------------------------------------------------------------------
    def set_labels(self, lattyp, xpos=None, ypos=None, zpos=None):
        lt = np.ascontiguousarray(lattyp, dtype=np.int32)
        x, y, z = _f64(xpos), _f64(ypos), _f64(zpos)
        call("lumagpu_set_labels", self.h, lt.ctypes.data_as(_ip), _dptr(x), _dptr(y), _dptr(z))

    def set_inlet(self, ux, uy, uz):
        a, b, c = _f64(ux), _f64(uy), _f64(uz)
        call("lumagpu_set_inlet", self.h, _dptr(a), _dptr(b), _dptr(c))

    def init_equilibrium(self, rho, u):
        r, uu = _f64(rho), _f64(u)
        call("lumagpu_init_equilibrium", self.h, _dptr(r), _dptr(uu))

    def upload_f(self, f):
        a = _f64(f)
        call("lumagpu_upload_f", self.h, _dptr(a))

    def upload_macro(self, rho, u):
        r, uu = _f64(rho), _f64(u)
        call("lumagpu_upload_macro", self.h, _dptr(r), _dptr(uu))

    def download_f(self):
        N, M, K = self.shape
        f = np.empty((N, M, K, self.Q))
        call("lumagpu_download_f", self.h, _dptr(f))
        return f

    def download_macro(self):
        N, M, K = self.shape
        rho = np.empty((N, M, K))
        u = np.empty((N, M, K, self.dims))
        call("lumagpu_download_macro", self.h, _dptr(rho), _dptr(u))
        return rho, u

    def set_force(self, F):
        a = _f64(F)
        call("lumagpu_set_force", self.h, _dptr(a))

    def set_store_macro(self, on=True):
        call("lumagpu_set_store_macro", self.h, int(bool(on)))

    # stepping -------------------------------------------------------------------------
    def adj_run(self, nsteps, checkpoint_every, direction=(1.0, 0.0, 0.0), backward=True,
                positions=True, initial=False):
        """Advance nsteps with checkpoints and, with backward, run the discrete adjoint of them.

        Returns a dict: J (time-averaged dir . IB body force), dJ_domega (under a Reynolds ramp:
        with respect to the omega after the ramp), dJ_dinlet (derivative
        with respect to a scale of the inlet profile), dJ_dpos (n_markers x 3, all bodies in
        marker order) and lambda0 (dJ/df of the initial populations, LUMA layout
        (N, M, K, Q)) when requested."""
        res = _abi.AdjResult()
        d = np.ascontiguousarray(direction, dtype=np.float64)
        nb = self.ibm_count()[0]
        nm = sum(self.ibm_count(b)[1] for b in range(nb))
        dpos = np.zeros((nm, 3)) if (backward and positions and nm) else None
        lam = np.zeros(self.shape + (self.Q,)) if (backward and initial) else None
        call("lumagpu_adj_run", self.h, int(nsteps), int(checkpoint_every), _dptr(d), int(bool(backward)),
             C.byref(res), _dptr(dpos) if dpos is not None else None, _dptr(lam) if lam is not None else None)
        out = {"J": res.J}
        if backward:
            out.update(dJ_domega=res.dJ_domega, dJ_dinlet=res.dJ_dinlet)
            if dpos is not None:
                out["dJ_dpos"] = dpos
            if lam is not None:
                out["lambda0"] = lam
        return out

    # device checkpoints ---------------------------------------------------------------------
    def checkpoint(self):
        """Copy of the populations, macro arrays and step counter in HBM; pass it to restore()
        (and release it with free_checkpoint())."""
        ck = C.c_void_p()
        call("lumagpu_checkpoint_save", self.h, C.byref(ck))
        return ck

    def restore(self, ck):
        call("lumagpu_checkpoint_restore", self.h, ck)

    @staticmethod
    def free_checkpoint(ck):
        call("lumagpu_checkpoint_free", ck)

    # LUMA's on-disk formats ---------------------------------------------------------------
    def write_restart(self, path, append=False, region=0):
        call("lumagpu_write_restart", self.h, str(path).encode(), int(append), int(region))

    def read_restart(self, path, region=0):
        call("lumagpu_read_restart", self.h, str(path).encode(), int(region))

    def io_lite(self, path, rank=0, precision=8):
        call("lumagpu_io_lite", self.h, str(path).encode(), int(rank), int(precision))

    def ibm_write_ld(self, path, body, t, time, precision=8):
        call("lumagpu_ibm_write_ld", self.h, str(path).encode(), int(body), int(t), float(time), int(precision))

    def ibm_write_tip(self, path, body, t, time, precision=10):
        call("lumagpu_ibm_write_tip", self.h, str(path).encode(), int(body), int(t), float(time), int(precision))

    def set_zslab(self, on=True):
        """z planes 0 and K-1 become halo planes exchanged around every odd step (one rank of a
        z-sharded periodic box, its own neighbour on both faces)."""
        call("lumagpu_set_zslab", self.h, int(on))

    def LBM_multi_opt(self, subcycle=0):
        call("lumagpu_step", self.h, int(subcycle))

    def run(self, nsteps):
        call("lumagpu_run", self.h, int(nsteps))

    # the step in the halves of LBM_multi_opt's two sweeps, with the IB stages between them
    def stream_bc_macro(self, subcycle=0):
        call("lumagpu_stream_bc_macro", self.h, int(subcycle))

    def force_collide_swap(self):
        call("lumagpu_force_collide_swap", self.h)

    def reset_forces(self):
        call("lumagpu_reset_forces", self.h)

    def ibm_apply(self):
        call("lumagpu_ibm_apply", self.h)

    def ibm_interpolate(self):
        call("lumagpu_ibm_interpolate", self.h)

    def ibm_compute_force(self):
        call("lumagpu_ibm_compute_force", self.h)

    def ibm_spread(self):
        call("lumagpu_ibm_spread", self.h)

    def ibm_update_macro(self):
        call("lumagpu_ibm_update_macro", self.h)

    def run_timed(self, nsteps):
        tot, ker = C.c_double(), C.c_double()
        call("lumagpu_run_timed", self.h, int(nsteps), C.byref(tot), C.byref(ker))
        return tot.value, ker.value

    def sync(self):
        call("lumagpu_sync", self.h)

    @property
    def t(self):
        t = C.c_int()
        call("lumagpu_get_time", self.h, C.byref(t))
        return t.value

    def moments(self):
        out = (C.c_double * _abi.NUM_MOMENTS)()
        call("lumagpu_moments", self.h, out)
        return {"mass": out[0], "momentum": (out[1], out[2], out[3]), "ke": out[4]}

    def footprint(self):
        a, b = C.c_size_t(), C.c_size_t()
        call("lumagpu_footprint", self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    # immersed boundary (ObjectManager) ------------------------------------------------
    def set_positions(self, x, y, z):
        a, b, c = _f64(x), _f64(y), _f64(z)
        call("lumagpu_set_positions", self.h, _dptr(a), _dptr(b), _dptr(c))

    def ibm_add_body(self, pos, flexible=False):
        p = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        body = C.c_int()
        call("lumagpu_ibm_add_body", self.h, len(p), _dptr(p), int(flexible), C.byref(body))
        return body.value

    def ibm_add_filament(self, start, length, height, angle_deg, n_elements, clamped, density, youngs, depth=0.0):
        st = np.ascontiguousarray(start, dtype=np.float64)
        body = C.c_int()
        call("lumagpu_ibm_add_filament", self.h, _dptr(st), length, height, depth, angle_deg, int(n_elements),
             int(clamped), density, youngs, C.byref(body))
        return body.value

    def ibm_subiterations(self):
        it, res = C.c_int(), C.c_double()
        call("lumagpu_ibm_subiterations", self.h, C.byref(it), C.byref(res))
        return it.value, res.value

    def ibm_initialise(self):
        call("lumagpu_ibm_initialise", self.h)

    def ibm_set_velocity(self, body, vel):
        v = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        call("lumagpu_ibm_set_velocity", self.h, int(body), _dptr(v))

    def ibm_set_epsilon(self, body, eps):
        e = _f64(eps)
        call("lumagpu_ibm_set_epsilon", self.h, int(body), _dptr(e))

    def ibm_count(self, body=None):
        """(number of bodies, markers of `body`) as the grid holds them."""
        nb, nm = C.c_int(), C.c_int()
        call("lumagpu_ibm_count", self.h, C.byref(nb), -1 if body is None else int(body),
             C.byref(nm) if body is not None else None)
        return nb.value, (nm.value if body is not None else None)

    def ibm_markers(self, body):
        rec = np.empty((self.ibm_count(body)[1], 16))
        call("lumagpu_ibm_get_markers", self.h, int(body), _dptr(rec))
        return {"pos": rec[:, 0:3], "vel": rec[:, 3:6], "force": rec[:, 6:9], "mom": rec[:, 9:12],
                "rho": rec[:, 12], "eps": rec[:, 13], "ds": rec[:, 14], "ns": rec[:, 15].astype(int)}

    def ibm_body_force(self, body):
        out = (C.c_double * 3)()
        call("lumagpu_ibm_body_force", self.h, int(body), out)
        return np.array(out[:])

    def ibm_subiteration_log(self):
        """FSI sub-iterations of every step since ibm_initialise."""
        n = C.c_int()
        call("lumagpu_ibm_subiteration_log", self.h, None, 0, C.byref(n))
        out = np.zeros(n.value, dtype=np.int32)
        call("lumagpu_ibm_subiteration_log", self.h, out.ctypes.data_as(_ip), n.value, C.byref(n))
        return out

    def ibm_pin_subiterations(self, counts):
        """The next len(counts) steps run exactly these FSI sub-iteration counts."""
        c = np.ascontiguousarray(counts, dtype=np.int32)
        call("lumagpu_ibm_pin_subiterations", self.h, c.ctypes.data_as(_ip), len(c))

    # --- tangent form (precision "fp64t"): derivatives along one seeded parameter direction --
    def download_tangent(self):
        """(df, drho, du) in the shapes of download_f / download_macro."""
        N, M, K = self.shape
        df = np.empty((N, M, K, self.Q))
        dr = np.empty((N, M, K))
        du = np.empty((N, M, K, self.dims))
        call("lumagpu_download_tangent", self.h, _dptr(df), _dptr(dr), _dptr(du))
        return df, dr, du

    def set_velocity_ramp(self, ramp_time):
        """L_VELOCITY_RAMP: ramp the inlet over ramp_time (dimensionless time); 0 turns it off."""
        call("lumagpu_set_velocity_ramp", self.h, float(ramp_time))

    def set_temporal_blocking(self, shape=0, zchunk=16):
        """Two steps per launch on a periodic D3Q19 box (shape < 0 turns it off)."""
        call("lumagpu_set_temporal_blocking", self.h, int(shape), int(zchunk))

    def set_reynolds_ramp(self, re, ramp_time):
        """L_REYNOLDS_RAMP over L_RE = re on this grid and the grids below it (0 turns it off)."""
        call("lumagpu_set_reynolds_ramp", self.h, float(re), float(ramp_time))

    @property
    def omega(self):
        o = C.c_double()
        call("lumagpu_get_omega", self.h, C.byref(o))
        return o.value

    def ibm_set_universal_epsilon(self, on=True):
        """L_UNIVERSAL_EPSILON_CALC: one epsilon system over all bodies (before ibm_initialise)."""
        call("lumagpu_ibm_set_universal_epsilon", self.h, int(on))

    def set_pressure_delta(self, p):
        """L_PRESSURE_DELTA: pressure sites hold L_RHOIN + pd2dlbm(p) (p dimensionless)."""
        call("lumagpu_set_pressure_delta", self.h, float(p))

    def set_gravity(self, gravity, direction):
        """L_GRAVITY_ON: the IB force reset writes rho * gravity * refinement_ratio each step."""
        call("lumagpu_set_gravity", self.h, float(gravity), int(direction))

    def read_initial_velocity(self, path, u=None):
        """GridObj::LBM_initVelocity from a velocity file: u (N, M, K, D) in lattice units,
        starting from `u` (zeros by default) where no row lands."""
        N, M, K = self.shape
        out = np.zeros((N, M, K, self.dims)) if u is None else np.array(u, dtype=np.float64, order="C")
        call("lumagpu_read_initial_velocity", self.h, str(path).encode(), _dptr(out))
        return out

    def read_inlet_profile(self, path):
        """GridObj::_LBM_initGetInletProfileFromFile: the inlet profile from an x y z ux uy uz file."""
        call("lumagpu_read_inlet_profile", self.h, str(path).encode())

    def seed_omega(self, d_omega=1.0):
        """Seed d(omega) = d_omega (FP64 tangent form): derivatives become d/domega."""
        call("lumagpu_tangent_seed_omega", self.h, float(d_omega))

    def ibm_seed_youngs(self, body, dE=1.0):
        """Seed d(Young's modulus) = dE for a flexible filament: derivatives become d/dE."""
        call("lumagpu_ibm_seed_youngs", self.h, int(body), float(dE))

    def ibm_marker_tangents(self, body):
        rec = np.empty((self.ibm_count(body)[1], 16))
        call("lumagpu_ibm_get_marker_tangents", self.h, int(body), _dptr(rec))
        return {"pos": rec[:, 0:3], "vel": rec[:, 3:6], "force": rec[:, 6:9], "mom": rec[:, 9:12],
                "rho": rec[:, 12], "eps": rec[:, 13], "ds": rec[:, 14]}

    def ibm_body_force_tangent(self, body):
        out, dout = (C.c_double * 3)(), (C.c_double * 3)()
        call("lumagpu_ibm_body_force_tangent", self.h, int(body), out, dout)
        return np.array(out[:]), np.array(dout[:])


def read_geometry(grids, path, shift=(0.0, 0.0, 0.0)):
    """ObjectManager::io_readInGeomConfig over GpuGrids: builds the file's IBM bodies on the grid
    of their level; returns the number of bodies built."""
    arr = (C.c_void_p * len(grids))(*[g.h for g in grids])
    sh = np.ascontiguousarray(shift, dtype=np.float64)
    n = C.c_int()
    call("lumagpu_ibm_read_geometry", arr, len(grids), str(path).encode(), _dptr(sh), C.byref(n))
    return n.value


def ibm_write_restart(grids, path):
    """ObjectManager::io_restart(eWrite): the bodies of the grids, in grid order, in LUMA's IB
    restart layout at full precision."""
    arr = (C.c_void_p * len(grids))(*[g.h for g in grids])
    call("lumagpu_ibm_write_restart", arr, len(grids), str(path).encode())


def ibm_read_restart(grids, path):
    """ObjectManager::io_restart(eRead): marker positions for the bodies already built on the
    grids; call ibm_initialise on each grid with bodies afterwards."""
    arr = (C.c_void_p * len(grids))(*[g.h for g in grids])
    call("lumagpu_ibm_read_restart", arr, len(grids), str(path).encode())
