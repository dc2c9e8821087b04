"""The matrix-free Hubbard FCI of tests/fci.py, pinned against the general determinant
FCI on small chains and against the U = 0 closed form
E = 2 sum_{k <= N/2} -2t cos(k pi / (L + 1)) of an open chain."""
import numpy as np
import pytest

from fci import fci_energy, hubbard, hubbard_fci_energy


@pytest.mark.parametrize("L,N,U", [(6, 6, 4.0), (6, 4, 2.0), (5, 5, 3.0)])
def test_against_general_fci(L, N, U):
    T, V = hubbard(L, U=U)
    assert abs(hubbard_fci_energy(L, N, U=U) - fci_energy(L, [0] * L, N, N % 2, 0, T, V)) < 1e-10


def test_u0_closed_form():
    L, N = 12, 12
    ref = 2 * sum(-2 * np.cos(k * np.pi / (L + 1)) for k in range(1, N // 2 + 1))
    assert abs(hubbard_fci_energy(L, N, U=0.0) - ref) < 1e-10
