"""``qmb.hamiltonian`` import path (qmb/hamiltonian.py of the reference) served by qmb_hip."""

from qmb_hip.hamiltonian import Hamiltonian

__all__ = ["Hamiltonian"]
