"""MI355X-native Gutzwiller-projected-determinant VMC sampler for the Kagome
Heisenberg model, with the host entry points of hz-xiaxz/KagomeDSL.jl.

Host setup (lattice, pi-flux ansatz, orbitals, bond list) runs in numpy; the
sampler state and every hot-path operation (Metropolis exchange step,
Sherman-Morrison W update, local energy, W refresh, log-derivatives) run in
libkagome_vmc.so on the GPU.
"""
from .lattice import (DoubleKagome, get_boundary_shifts, get_site_coord, ns, unitcell_coord,
                      unitcell_diff)
from .hamiltonian import Hamiltonian, Hmat, get_nn, orbitals, pi_link_in, pi_link_inter, shell_gap
from .montecarlo import (MC, MCContext, find_initial_configuration, init, init_conf_qr, measure,
                         read_checkpoint, register_evaluables, run, sweep, write_checkpoint)
from .gradient import (AnsatzFamily, GradientSampler, ansatz_family, energy, exhaustive_average,
                       orbital_derivatives)
from . import _native as native

__all__ = [
    "DoubleKagome", "ns", "unitcell_coord", "unitcell_diff", "get_site_coord",
    "get_boundary_shifts", "Hamiltonian", "Hmat", "get_nn", "orbitals", "pi_link_in",
    "pi_link_inter", "shell_gap", "MC", "MCContext", "init", "sweep", "measure", "run",
    "register_evaluables", "write_checkpoint", "read_checkpoint", "native", "AnsatzFamily",
    "init_conf_qr", "find_initial_configuration",
    "GradientSampler", "ansatz_family", "energy", "exhaustive_average", "orbital_derivatives",
]
