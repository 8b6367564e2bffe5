"""Pin the oracle's lattice tables and equilibrium against LUMA's own t = 0 dumps and
against the closed-form moments every lattice must satisfy."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle.oracle import OracleGrid, lattice

GOLD = Path(__file__).resolve().parent / "golden" / "t0_feq_d2q9.json"


def test_t0_equilibrium_matches_luma_testsuite_dumps():
    # LUMA/cases/testsuite/base/case00{0,3,4}/io_lite.Lev0.Reg0.Rnk0.0.dat, 8 decimals
    vecs = json.loads(GOLD.read_text())["vectors"]
    n = len(vecs)
    assert n > 50
    g = OracleGrid(2, 9, n, 1, 1, omega=1.0)
    rho = np.array([v["rho"] for v in vecs]).reshape(n, 1, 1)
    u = np.array([[v["ux"], v["uy"]] for v in vecs]).reshape(n, 1, 1, 2)
    g.init_feq(rho, u)
    f = g.get_f().reshape(n, 9)
    ref = np.array([v["f"] for v in vecs])
    # printed to 8 decimals; u itself printed to 8 decimals (|df/du| <= 3 w (1 + 3|u|))
    assert np.abs(f - ref).max() < 1e-8


@pytest.mark.parametrize("Q,D", [(9, 2), (19, 3)])
def test_lattice_isotropy(Q, D):
    c, w, opp = lattice(Q)
    cs2 = 1.0 / 3.0
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.allclose((w[:, None] * c).sum(0), 0.0, atol=1e-16)
    second = np.einsum("v,va,vb->ab", w, c, c)
    assert np.allclose(second[:D, :D], cs2 * np.eye(D), atol=1e-15)
    fourth = np.einsum("v,va,vb,vc,vd->abcd", w, c, c, c, c)
    for a in range(D):
        for b in range(D):
            assert abs(fourth[a, a, b, b] - (cs2 * cs2 * (3.0 if a == b else 1.0))) < 1e-15
    assert (c[opp] == -c).all()
    assert (c[-1] == 0).all()     # rest population last (LUMA's c_opt order)


@pytest.mark.parametrize("Q,D", [(9, 2), (19, 3)])
def test_equilibrium_moments_closed_form(Q, D):
    rng = np.random.default_rng(7)
    n = 64
    rho = 1.0 + 0.05 * rng.standard_normal((n, 1, 1))
    u = 0.08 * rng.standard_normal((n, 1, 1, D))
    g = OracleGrid(D, Q, n, 1, 1, omega=1.0)
    g.init_feq(rho, u)
    f = g.get_f().reshape(n, Q)
    c, w, _ = lattice(Q)
    c = c[:, :D].astype(float)
    r = rho.reshape(n)
    uu = u.reshape(n, D)
    assert np.allclose(f.sum(1), r, rtol=1e-14, atol=0)
    assert np.allclose(f @ c, r[:, None] * uu, rtol=1e-12, atol=1e-15)
    pi = np.einsum("nv,va,vb->nab", f, c, c)
    ref = r[:, None, None] * (np.eye(D)[None] / 3.0 + uu[:, :, None] * uu[:, None, :])
    assert np.allclose(pi, ref, rtol=1e-12, atol=1e-15)


def test_setup_lattice_tables_match_the_oracle():
    from luma_amd import setups
    for Q in (9, 19):
        c, w, _ = lattice(Q)
        for v in range(Q):
            assert tuple(c[v]) == setups.lattice_c(Q, v)
            assert w[v] == setups.lattice_w(Q, v)


def test_c4_marker_fixture_is_reproduced():
    # the committed, seeded marker positions of the sphere-adjoint configuration
    from luma_amd.cases import c4_markers
    fix = np.load(Path(__file__).resolve().parent / "golden" / "c4_sphere_markers.npy")
    pos = c4_markers()
    assert fix.shape == (1810, 3) and np.array_equal(pos, fix)
    r = np.linalg.norm(pos - np.array([96.0, 64.0, 64.0]), axis=1)
    assert abs(r.mean() - 12.0) < 0.01 and 0.04 < (r - 12.0).std() < 0.06
