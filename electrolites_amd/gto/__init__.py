"""Molecules, basis sets and shell tables for the density-fitted path."""
from .mole import Mole, ShellTable, cart2sph, cart_powers, ncart, BOHR
from .cluster import water_cluster, water_cluster_atoms, cube_side

__all__ = ['Mole', 'ShellTable', 'cart2sph', 'cart_powers', 'ncart', 'BOHR',
           'water_cluster', 'water_cluster_atoms', 'cube_side']
