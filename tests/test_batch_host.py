"""Closed-shell cluster choice of the batch helper (host only)."""
import pytest

from kagome_vmc.batch import pi_flux_cluster
from kagome_systems import system


@pytest.mark.parametrize("L", [2, 4, 6, 8, 10, 12])
def test_pi_flux_cluster_is_closed_shell_and_matches_known_bc(L):
    ham = pi_flux_cluster(L)
    assert ham.ns == 3 * L * L and ham.N_up == ham.N_down == ham.ns // 2
    w = ham.spectrum
    assert w[ham.N_up] - w[ham.N_up - 1] > 1e-8
    # agrees with the boundary conditions the parity tests use
    assert ham.nn == system(L).nn
    assert abs(w - system(L).spectrum).max() < 1e-12
