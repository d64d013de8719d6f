"""Units of the reference (eV, Angstrom) -- restates constants.py:10-20.

Pinned by: CODATA 2018 values (scipy.constants) in tests/test_oracle_units.py.
"""

HARTREE_EV = 27.211386245988
RY_EV = HARTREE_EV / 2.0
BOHR_ANG = 0.529177210903
HBAR2_2M = RY_EV * BOHR_ANG**2  # eV*A^2, T(G) = HBAR2_2M |k+G|^2
E2 = HARTREE_EV * BOHR_ANG  # eV*A, Coulomb prefactor
MINUS_I_POW = (1.0 + 0.0j, -1.0j, -1.0 + 0.0j, 1.0j, 1.0 + 0.0j)
