"""observables:: mirror (src/observables/)."""
from .hamiltonian import ks_hamiltonian


def current(ham: ks_hamiltonian, phi, occ=2.0):
    """observables::current(ions, ham, el) (current.hpp:25-53) for equally
    occupied bands phi: the covariant total current (what the reference
    returns) and its Cartesian form (viewables.hpp:100-102)."""
    return ham.current(phi, occ)
