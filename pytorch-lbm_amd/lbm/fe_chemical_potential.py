"""compute_density (lbm/fe_chemical_potential.py:106-125 of the reference)."""


def compute_density(C, rho_l, rho_g):
    """rho = C*rho_l + (1 - C)*rho_g."""
    return C * rho_l + (1.0 - C) * rho_g
