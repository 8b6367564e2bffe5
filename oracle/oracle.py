"""ctypes binding of the FP64 CPU oracle (oracle/lbm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the luma_amd package.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "_build" / "liboracle.so"

SOLID, FLUID, REFINED, TL2C, TL2F, BFL, VELOCITY, PRESSURE, SLIP, EXTRAP_RIGHT = range(10)


class OraDesc(C.Structure):
    _fields_ = [
        ("dims", C.c_int), ("Q", C.c_int), ("N", C.c_int), ("M", C.c_int), ("K", C.c_int),
        ("regularised", C.c_int), ("forced", C.c_int), ("fz_compat", C.c_int), ("level", C.c_int),
        ("omega", C.c_double), ("rho_in", C.c_double),
        ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double),
        ("wall_thk", C.c_double * 6), ("coarse_lims", C.c_int * 6),
    ]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            from oracle.build import build_oracle
            build_oracle()
        lib = C.CDLL(str(LIB))
        protos = {
            "ora_create": ([C.POINTER(OraDesc)], _vp), "ora_destroy": ([_vp], None),
            "ora_add_child": ([_vp, _vp], C.c_int),
            "ora_set_labels": ([_vp, _ip], None), "ora_set_positions": ([_vp, _dp, _dp, _dp], None),
            "ora_set_inlet": ([_vp, _dp, _dp, _dp], None), "ora_set_macro": ([_vp, _dp, _dp], None),
            "ora_set_force": ([_vp, _dp], None), "ora_set_f": ([_vp, _dp], None),
            "ora_init_feq": ([_vp], None), "ora_get_f": ([_vp, _dp], None),
            "ora_get_macro": ([_vp, _dp, _dp], None), "ora_get_force": ([_vp, _dp], None),
            "ora_get_t": ([_vp], C.c_int), "ora_multi_opt": ([_vp, C.c_int], None),
            "ora_sweep1": ([_vp, C.c_int], None), "ora_sweep2": ([_vp], None),
            "ora_reset_forces": ([_vp], None), "ora_run": ([_vp, C.c_int, C.c_int], None),
            "ora_set_gravity": ([_vp, C.c_double, C.c_int], None),
            "ora_set_ramp": ([_vp, C.c_double, C.c_double], None),
            "ora_set_pressure_delta": ([_vp, C.c_double, C.c_double, C.c_double, C.c_double], None),
            "ora_set_reynolds_ramp": ([_vp, C.c_double, C.c_double, C.c_double, C.c_double], None),
            "ora_get_omega": ([_vp], C.c_double),
            "ora_equilibrium": ([_vp, C.c_int, C.c_int], C.c_double),
            "ora_lattice": ([C.c_int, _ip, _dp, _ip], None), "ora_cs": ([], C.c_double),
            "ora_ibm_create": ([_vp, C.c_double, _dp], _vp), "ora_ibm_destroy": ([_vp], None),
            "ora_ibm_add_body": ([_vp, C.c_int, _dp, C.c_int], C.c_int),
            "ora_ibm_initialise": ([_vp], C.c_int), "ora_ibm_attach": ([_vp], None),
            "ora_ibm_set_velocity": ([_vp, C.c_int, _dp], None),
            "ora_ibm_set_epsilon": ([_vp, C.c_int, _dp], None),
            "ora_ibm_set_universal_epsilon": ([_vp, C.c_int], None),
            "ora_ibm_get_markers": ([_vp, C.c_int, _dp], None),
            "ora_ibm_get_support": ([_vp, C.c_int, C.c_int, _ip, _dp], C.c_int),
            "ora_ibm_delta": ([C.c_double, C.c_double], C.c_double),
            "ora_lu_solve_t": ([C.c_int, _dp, _dp], C.c_int),
            "ora_ibm_add_filament": ([_vp, _dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                      C.c_double, C.c_double, C.c_double, C.c_double], C.c_int),
            "ora_ibm_error": ([_vp], C.c_int),
            "ora_fem_create": ([_dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double,
                                C.c_double, C.c_double, C.c_double, C.c_double], _vp),
            "ora_fem_destroy": ([_vp], None), "ora_fem_ndof": ([_vp], C.c_int),
            "ora_fem_assemble": ([_vp, _dp, _dp], None),
            "ora_fem_set_state": ([_vp, _dp, _dp, _dp], None),
            "ora_fem_step_load": ([_vp, _dp], C.c_int),
            "ora_fem_state": ([_vp, _dp, _dp, _ip, _dp], None),
            "ora_fem_markers": ([_vp, _dp, _dp, _dp, _dp], None), "ora_fem_nib": ([_vp], C.c_int),
            "ora_ibm_subiter": ([_vp, _ip, _dp], None),
            "ora_ibm_fem_state": ([_vp, C.c_int, _dp, _dp, _ip, _dp], C.c_int),
        }
        for name, (args, res) in protos.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def lattice(Q):
    lib = _load()
    c = np.zeros(Q * 3, np.int32)
    w = np.zeros(Q)
    o = np.zeros(Q, np.int32)
    lib.ora_lattice(Q, c.ctypes.data_as(_ip), w.ctypes.data_as(_dp), o.ctypes.data_as(_ip))
    return c.reshape(Q, 3), w, o


class OracleGrid:
    """LUMA-layout FP64 grid: arrays shaped (N, M, K[, Q or D]) in C order."""

    def __init__(self, dims, Q, N, M, K=1, omega=1.0, regularised=True, forced=False, fz_compat=True,
                 rho_in=1.0, edges=(1.0, 1.0, 1.0), wall_thk=(0.0,) * 6, level=0, coarse_lims=(0,) * 6):
        self.lib = _load()
        d = OraDesc()
        d.dims, d.Q, d.N, d.M, d.K = dims, Q, N, M, K
        d.regularised, d.forced, d.fz_compat, d.level = int(regularised), int(forced), int(fz_compat), level
        d.omega, d.rho_in = omega, rho_in
        d.xmax, d.ymax, d.zmax = edges
        for i in range(6):
            d.wall_thk[i] = wall_thk[i]
            d.coarse_lims[i] = coarse_lims[i]
        self.desc = d
        self.dims, self.Q, self.shape = dims, Q, (N, M, K)
        self.h = self.lib.ora_create(C.byref(d))
        self._owned = True

    def close(self):
        if self.h and self._owned:
            self.lib.ora_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_child(self, child: "OracleGrid"):
        assert self.lib.ora_add_child(self.h, child.h) == 0
        child._owned = False   # destroyed with the parent
        self._kids = getattr(self, "_kids", []) + [child]

    def set_labels(self, lattyp):
        a = np.ascontiguousarray(lattyp, dtype=np.int32)
        self.lib.ora_set_labels(self.h, a.ctypes.data_as(_ip))

    def set_positions(self, x, y, z):
        (x, px), (y, py), (z, pz) = _d(x), _d(y), _d(z)
        self.lib.ora_set_positions(self.h, px, py, pz)

    def set_inlet(self, ux, uy, uz):
        (a, pa), (b, pb), (c, pc) = _d(ux), _d(uy), _d(uz)
        self.lib.ora_set_inlet(self.h, pa, pb, pc)

    def set_macro(self, rho, u):
        (r, pr), (uu, pu) = _d(rho), _d(u)
        self.lib.ora_set_macro(self.h, pr, pu)

    def set_force(self, F):
        a, p = _d(F)
        self.lib.ora_set_force(self.h, p)

    def set_gravity(self, gravity, direction):
        self.lib.ora_set_gravity(self.h, float(gravity), int(direction))

    def set_ramp(self, ramp_time, dt):
        self.lib.ora_set_ramp(self.h, float(ramp_time), float(dt))

    def set_reynolds_ramp(self, re, ramp_time, dt, dh):
        self.lib.ora_set_reynolds_ramp(self.h, float(re), float(ramp_time), float(dt), float(dh))

    @property
    def omega(self):
        return self.lib.ora_get_omega(self.h)

    def set_pressure_delta(self, p, dh, dt, dm):
        self.lib.ora_set_pressure_delta(self.h, float(p), float(dh), float(dt), float(dm))

    def set_f(self, f):
        a, p = _d(f)
        self.lib.ora_set_f(self.h, p)

    def init_feq(self, rho, u):
        self.set_macro(rho, u)
        self.lib.ora_init_feq(self.h)

    def get_f(self):
        N, M, K = self.shape
        f = np.empty((N, M, K, self.Q))
        self.lib.ora_get_f(self.h, f.ctypes.data_as(_dp))
        return f

    def get_macro(self):
        N, M, K = self.shape
        rho = np.empty((N, M, K))
        u = np.empty((N, M, K, self.dims))
        self.lib.ora_get_macro(self.h, rho.ctypes.data_as(_dp), u.ctypes.data_as(_dp))
        return rho, u

    def get_force(self):
        N, M, K = self.shape
        F = np.empty((N, M, K, self.dims))
        self.lib.ora_get_force(self.h, F.ctypes.data_as(_dp))
        return F

    def step(self, subcycle=0):
        self.lib.ora_multi_opt(self.h, subcycle)

    def run(self, n, nthreads=1):
        self.lib.ora_run(self.h, n, nthreads)

    @property
    def t(self):
        return self.lib.ora_get_t(self.h)

    def equilibrium(self, id_, v):
        return self.lib.ora_equilibrium(self.h, id_, v)


class OracleIBM:
    """ObjectManager's IB bodies on one oracle grid (serial LUMA semantics)."""

    def __init__(self, grid: OracleGrid, dh: float, origin=(0.0, 0.0, 0.0)):
        self.lib = _load()
        self.grid = grid
        o = np.ascontiguousarray(origin, dtype=np.float64)
        self.h = self.lib.ora_ibm_create(grid.h, dh, o.ctypes.data_as(_dp))
        self.nm = []
        self.dh = dh

    def add_body(self, pos, flexible=False):
        p = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        self.nm.append(len(p))
        return self.lib.ora_ibm_add_body(self.h, len(p), p.ctypes.data_as(_dp), int(flexible))

    def add_filament(self, start, length, height, depth, angle_deg, n_el, clamped, density, E, dt, dm):
        st = np.ascontiguousarray(start, dtype=np.float64)
        ib = self.lib.ora_ibm_add_filament(self.h, st.ctypes.data_as(_dp), length, height, depth, angle_deg, n_el,
                                           int(clamped), density, E, dt, dm)
        self.nm.append(int(np.floor(length / self.dh)) + 1)
        return ib

    def error(self):
        return self.lib.ora_ibm_error(self.h)

    def subiterations(self):
        it, res = C.c_int(), C.c_double()
        self.lib.ora_ibm_subiter(self.h, C.byref(it), C.byref(res))
        return it.value, res.value

    def fem_state(self, ib, ndof):
        U = np.empty(ndof)
        Ud = np.empty(ndof)
        it, res = C.c_int(), C.c_double()
        self.lib.ora_ibm_fem_state(self.h, ib, U.ctypes.data_as(_dp), Ud.ctypes.data_as(_dp), C.byref(it), C.byref(res))
        return U, Ud, it.value, res.value

    def initialise(self):
        rc = self.lib.ora_ibm_initialise(self.h)
        if rc != 0:
            raise RuntimeError(f"oracle ibm_initialise failed ({rc})")
        self.lib.ora_ibm_attach(self.h)

    def set_velocity(self, ib, vel):
        v = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        self.lib.ora_ibm_set_velocity(self.h, ib, v.ctypes.data_as(_dp))

    def set_universal_epsilon(self, on=True):
        """L_UNIVERSAL_EPSILON_CALC: one epsilon system over every body (call before initialise)."""
        self.lib.ora_ibm_set_universal_epsilon(self.h, int(on))

    def set_epsilon(self, ib, eps):
        e = np.ascontiguousarray(eps, dtype=np.float64)
        self.lib.ora_ibm_set_epsilon(self.h, ib, e.ctypes.data_as(_dp))

    def markers(self, ib):
        out = np.empty((self.nm[ib], 19))
        self.lib.ora_ibm_get_markers(self.h, ib, out.ctypes.data_as(_dp))
        return {"pos": out[:, 0:3], "vel": out[:, 3:6], "force": out[:, 6:9], "mom": out[:, 9:12],
                "rho": out[:, 12], "eps": out[:, 13], "ds": out[:, 14], "ns": out[:, 15].astype(int),
                "vel_km1": out[:, 16:19]}

    def support(self, ib, m):
        ijk = np.empty(27 * 3, np.int32)
        d = np.empty(27)
        n = self.lib.ora_ibm_get_support(self.h, ib, m, ijk.ctypes.data_as(_ip), d.ctypes.data_as(_dp))
        return ijk[:3 * n].reshape(n, 3), d[:n]

    def close(self):
        if self.h:
            self.lib.ora_ibm_destroy(self.h)
            self.h = None


def delta_kernel(r, dilation=1.0):
    return _load().ora_ibm_delta(float(r), float(dilation))


def lu_solve(A, b):
    """A x = b through the oracle's LU (row-major A, as GridUtils::solveLinearSystem takes it)."""
    a = np.ascontiguousarray(A, dtype=np.float64).copy()
    x = np.ascontiguousarray(b, dtype=np.float64).copy()
    rc = _load().ora_lu_solve_t(len(x), a.ctypes.data_as(_dp), x.ctypes.data_as(_dp))
    if rc != 0:
        raise RuntimeError("singular matrix")
    return x


class OracleFEM:
    """Stand-alone FEMBody (2D corotational beam) for pinning tests."""

    def __init__(self, start, length, height, angle_deg, n_el, clamped, density, E, dh=1.0, dt=1.0, dm=1.0):
        self.lib = _load()
        st = np.ascontiguousarray(start, dtype=np.float64)
        self.h = self.lib.ora_fem_create(st.ctypes.data_as(_dp), length, height, dh, angle_deg, n_el, int(clamped),
                                         density, E, dh, dt, dm)
        self.n = self.lib.ora_fem_ndof(self.h)

    def assemble(self):
        K = np.empty((self.n, self.n))
        M = np.empty((self.n, self.n))
        self.lib.ora_fem_assemble(self.h, K.ctypes.data_as(_dp), M.ctypes.data_as(_dp))
        return K, M

    def set_state(self, U, Ud=None, Udd=None):
        z = np.zeros(self.n)
        a, b, c = [np.ascontiguousarray(x if x is not None else z, dtype=np.float64) for x in (U, Ud, Udd)]
        self.lib.ora_fem_set_state(self.h, a.ctypes.data_as(_dp), b.ctypes.data_as(_dp), c.ctypes.data_as(_dp))

    def step(self, R=None):
        r = np.ascontiguousarray(R if R is not None else np.zeros(self.n), dtype=np.float64)
        return self.lib.ora_fem_step_load(self.h, r.ctypes.data_as(_dp))

    def markers(self, pos0):
        """Marker positions and velocities of the current state (updateIBMarkers)."""
        p0 = np.ascontiguousarray(pos0, dtype=np.float64).reshape(-1, 3)
        assert len(p0) == self.lib.ora_fem_nib(self.h)
        pos, vel, vkm1 = np.zeros_like(p0), np.zeros_like(p0), np.zeros_like(p0)
        self.lib.ora_fem_markers(self.h, p0.ctypes.data_as(_dp), pos.ctypes.data_as(_dp), vel.ctypes.data_as(_dp),
                                 vkm1.ctypes.data_as(_dp))
        return pos, vel

    def state(self):
        U, Ud = np.empty(self.n), np.empty(self.n)
        it, res = C.c_int(), C.c_double()
        self.lib.ora_fem_state(self.h, U.ctypes.data_as(_dp), Ud.ctypes.data_as(_dp), C.byref(it), C.byref(res))
        return U, Ud

    def close(self):
        if self.h:
            self.lib.ora_fem_destroy(self.h)
            self.h = None
