"""Pin the FEM oracle (FEMBody.cpp / FEMElement.cpp restated) with beam theory closed forms."""
import numpy as np
import pytest

from oracle.oracle import OracleFEM

L, H, E, RHO, NEL = 1.0, 0.02, 1e5, 10.0, 20
I = 1.0 * H ** 3 / 12.0      # depth = dh = 1 in 2D
A = H * 1.0


def beam(angle=0.0, clamped=True, nel=NEL):
    return OracleFEM((0.0, 0.0, 0.0), L, H, angle, nel, clamped, RHO, E, dh=1.0, dt=1.0, dm=1.0)


def test_static_tip_deflection_is_PL3_over_3EI():
    K, M = beam().assemble()
    P = 1e-6
    rhs = np.zeros(K.shape[0] - 3)
    rhs[-2] = P
    u = np.linalg.solve(K[3:, 3:], rhs)
    assert u[-2] == pytest.approx(P * L ** 3 / (3 * E * I), rel=1e-10)
    assert u[-1] == pytest.approx(P * L ** 2 / (2 * E * I), rel=1e-10)       # tip rotation


def test_first_cantilever_frequency_and_mass():
    K, M = beam().assemble()
    w2 = np.sort(np.real(np.linalg.eigvals(np.linalg.solve(M[3:, 3:], K[3:, 3:]))))
    assert np.sqrt(w2[0]) == pytest.approx(1.8751040687 ** 2 * np.sqrt(E * I / (RHO * A * L ** 4)), rel=1e-6)
    assert M[0::3, 0::3].sum() == pytest.approx(RHO * A * L, rel=1e-14)
    assert np.allclose(K, K.T, atol=1e-12 * np.abs(K).max())


def test_rotated_beam_matrices_are_congruent():
    # the element transformation T rotates the beam; stiffness eigenvalues are invariant
    K0, M0 = beam(0.0).assemble()
    K1, M1 = beam(30.0).assemble()
    e0 = np.sort(np.linalg.eigvalsh(K0))
    e1 = np.sort(np.linalg.eigvalsh(K1))
    assert np.allclose(e0, e1, rtol=1e-9, atol=1e-9 * e0.max())


def test_newmark_trapezoidal_conserves_energy_of_small_free_vibration():
    f = beam(nel=10)
    K, M = f.assemble()
    n = K.shape[0]
    x = np.linspace(0, L, 11)
    U = np.zeros(n)
    U[1::3] = 1e-6 * (x / L) ** 2           # small transverse deflection, zero velocity
    U[2::3] = 1e-6 * 2 * x / L ** 2
    U[:3] = 0.0
    # consistent start: M Udd + K U = 0 (trapezoidal Newmark conserves energy from there)
    Udd = np.zeros(n)
    Udd[3:] = -np.linalg.solve(M[3:, 3:], K[3:, 3:] @ U[3:])
    f.set_state(U, None, Udd)
    e0 = 0.5 * U[3:] @ K[3:, 3:] @ U[3:]
    for _ in range(200):
        assert f.step() >= 1
    U1, Ud1 = f.state()
    e1 = 0.5 * U1[3:] @ K[3:, 3:] @ U1[3:] + 0.5 * Ud1[3:] @ M[3:, 3:] @ Ud1[3:]
    assert abs(e1 / e0 - 1) < 1e-5
    assert np.abs(U1[:3]).max() == 0.0                  # clamped DOFs stay fixed
