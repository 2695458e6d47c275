"""Pins the FP64 spherical-harmonics basis the SH pre-pass is checked
against: the 16 functions are orthonormal on the unit sphere (Lebedev-free
check: Gauss-Legendre in cos(theta) x uniform phi, exact for degree <= 6
products), and degree-1 functions are proportional to the coordinates with
the 3DGS signs (-y, z, -x)."""
import numpy as np

from oracle import fp64


def test_sh_basis_orthonormal():
    ct, wt = np.polynomial.legendre.leggauss(16)
    phi = np.linspace(0, 2 * np.pi, 32, endpoint=False)
    st = np.sqrt(1 - ct**2)
    v = np.stack([st[:, None] * np.cos(phi)[None], st[:, None] * np.sin(phi)[None],
                  np.repeat(ct[:, None], 32, 1)], -1).reshape(-1, 3)
    w = (np.repeat(wt[:, None], 32, 1) * (2 * np.pi / 32)).reshape(-1)
    Y = fp64.sh_basis_fp64(v, 3)
    G = (Y * w[:, None]).T @ Y
    assert np.allclose(G, np.eye(16), atol=1e-12)


def test_sh_degree1_signs():
    Y = fp64.sh_basis_fp64(np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]), 1)
    c1 = fp64.SH_C1
    assert np.allclose(Y[0, 1:], [0, c1, 0]) and np.allclose(Y[1, 1:], [0, 0, -c1])
    assert np.allclose(Y[2, 1:], [-c1, 0, 0])
