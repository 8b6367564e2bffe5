"""Two steps per launch (temporal blocking of the periodic D3Q19 step) against one step per
launch and against the FP64 oracle: bit for bit in the FP64 and FP32-shifted forms, to within
the storage rounding in the FP16-shifted form."""
import numpy as np
import pytest

from luma_amd import setups
from luma_amd._abi import LumaGpuError
from luma_amd.grid import GpuGrid
from oracle.oracle import OracleGrid

pytestmark = pytest.mark.gpu


def _tgv(shape):
    return setups.taylor_green(shape, u0=0.05, tau=0.55, dims=3)


def _run(case, prec, nsteps, fuse=None, parity_first=False):
    with GpuGrid.from_case(case, prec, store_macro=True) as g:
        if parity_first:
            g.run(1)
        if fuse is not None:
            g.set_temporal_blocking(*fuse)
        g.run(nsteps)
        assert g.t == nsteps + (1 if parity_first else 0)
        return g.download_f(), g.download_macro()


CASES = [(p, s, z) for p in ("fp64", "fp32s", "fp16s") for s, z in ((0, 7), (1, 5), (2, 64), (0, 128))]
CASES += [(p, s, z) for p in ("fp32s", "fp16s") for s, z in ((8, 6), (9, 7), (10, 128), (10, 3))]
# loads one plane ahead (24, 25, 26; zc 7 and 5 leave an odd last plane pair), windowed LDS
# copies (40, 41) and both (57)
CASES += [(p, s, z) for p in ("fp32s", "fp16s")
          for s, z in ((24, 7), (24, 6), (25, 5), (25, 128), (26, 128), (40, 6), (41, 5), (57, 3), (57, 128))]


@pytest.mark.parametrize("prec,shape,zc", CASES)
def test_two_step_kernel_matches_single_steps(prec, shape, zc):
    # 128 x 32 x 20: two or four tiles in x, two to eight in y, z chunks that do not divide the
    # depth (zc = 7, 5, 6, 3) and one chunk over the whole depth (zc = 64, 128); shapes 8 and 10
    # alternate two LDS copies, 24 to 26 load a plane ahead, 40 and 41 keep only the tile's
    # windows in LDS, 57 both
    case = _tgv((128, 32, 20))
    print(f"[fused] {prec} shape={shape} zc={zc}", flush=True)
    f1, (r1, u1) = _run(case, prec, 6)
    f2, (r2, u2) = _run(case, prec, 6, fuse=(shape, zc))
    if prec == "fp16s":
        # the compiler may fold a collision's last FMA and the conversion to half into one
        # mixed-precision FMA (rounding once, straight to half) in one kernel and not in the
        # other: equal to within the half storage rounding (the shifted values are below 2^-4,
        # where one unit in the last place of a half is at most 2^-15 = 3.05e-5)
        assert np.abs(f1 - f2).max() <= 3.05e-5
        assert np.abs(u1 - u2).max() < 1e-5
        return
    assert np.array_equal(f1, f2)
    assert np.array_equal(r1, r2) and np.array_equal(u1, u2)


def test_two_step_kernel_odd_counts_and_parity():
    # an odd step count ends with a single step; a grid at odd parity takes one single step first
    case = _tgv((64, 16, 9))
    print("[fused] odd counts", flush=True)
    for parity_first in (False, True):
        f1, (r1, u1) = _run(case, "fp32s", 7, parity_first=parity_first)
        f2, (r2, u2) = _run(case, "fp32s", 7, fuse=(8, 4), parity_first=parity_first)
        assert np.array_equal(f1, f2) and np.array_equal(u1, u2)


def test_two_step_kernel_fp64_equals_oracle():
    case = _tgv((64, 16, 12))
    print("[fused] fp64 vs oracle", flush=True)
    ref = OracleGrid(case.dims, case.Q, case.N, case.M, case.K, omega=case.omega,
                     regularised=case.regularised, forced=case.forced, fz_compat=case.fz_compat,
                     rho_in=case.rho_in, edges=case.edges, wall_thk=case.wall_thk)
    ref.init_feq(case.rho, case.u)
    ref.run(10)
    f, (r, u) = _run(case, "fp64", 10, fuse=(0, 5))
    rr, ur = ref.get_macro()
    assert np.array_equal(f, ref.get_f())
    assert np.array_equal(r, rr) and np.array_equal(u, ur)


def test_two_step_kernel_run_timed_and_switch_off():
    case = _tgv((64, 16, 12))
    print("[fused] timed, off", flush=True)
    f1, _ = _run(case, "fp32s", 8)
    with GpuGrid.from_case(case, "fp32s") as g:
        g.set_temporal_blocking(8, 6)
        tot, ker = g.run_timed(4)
        assert tot > 0 and ker > 0
        g.set_temporal_blocking(-1, 0)
        g.run(4)
        assert np.array_equal(g.download_f(), f1)


def test_two_step_kernel_refuses_unsuitable_grids():
    with GpuGrid.from_case(_tgv((64, 12, 8)), "fp32s") as g:
        with pytest.raises(LumaGpuError, match="whole numbers"):
            g.set_temporal_blocking(0, 4)     # M = 12 is not a whole number of 8-row tiles
    with GpuGrid.from_case(_tgv((64, 16, 8)), "fp64") as g:
        with pytest.raises(LumaGpuError, match="at most 4 bytes"):
            g.set_temporal_blocking(8, 4)     # two FP64 LDS copies exceed a CU's LDS
    c2 = setups.taylor_green((32, 32, 1), u0=0.05, tau=0.56, dims=2)
    with GpuGrid.from_case(c2, "fp32s") as g:
        with pytest.raises(LumaGpuError, match="D3Q19"):
            g.set_temporal_blocking(1, 4)


def test_two_step_kernel_full_size_256():
    # the metric box, 256^3 FP32-shifted: the two-step kernel's state after six steps equals six
    # single steps exactly (FP64 moments of the populations and the last step's macro arrays),
    # and mass is conserved to the shifted storage's rounding
    case = _tgv(256)
    print("[fused] 256^3 fp32s", flush=True)
    out = []
    for fuse in (None, (8, 128)):
        with GpuGrid.from_case(case, "fp32s", store_macro=True) as g:
            m0 = g.moments()
            if fuse is not None:
                g.set_temporal_blocking(*fuse)
            g.run(6)
            m = g.moments()
            r, u = g.download_macro()
        out.append((m, r, u))
        assert abs(m["mass"] - m0["mass"]) < 1e-6 * m0["mass"]
    (ma, ra, ua), (mb, rb, ub) = out
    assert ma == mb
    assert np.array_equal(ra, rb) and np.array_equal(ua, ub)


@pytest.mark.parametrize("prec", ["fp64", "fp32s"])
def test_two_step_graph_replays_match_single_steps(prec):
    # lumagpu_run replays a captured pair of two-step launches in long runs: several runs in a
    # row, with downloads and reductions between them, each checked against single steps
    case = _tgv((64, 16, 12))
    print(f"[fused] graph replays {prec}", flush=True)
    segments = (1, 20, 3, 17, 12)
    with GpuGrid.from_case(case, prec, store_macro=True) as ref, \
            GpuGrid.from_case(case, prec, store_macro=True) as g:
        g.set_temporal_blocking(8 if prec == "fp32s" else 0, 5)   # two LDS copies need <= 4-byte storage
        for n in segments:
            ref.run(n)
            g.run(n)
            assert g.t == ref.t
            assert np.array_equal(g.download_f(), ref.download_f())
            (rg, ug), (rr, ur) = g.download_macro(), ref.download_macro()
            assert np.array_equal(rg, rr) and np.array_equal(ug, ur)
            assert g.moments() == ref.moments()
