"""
qmb_hip: MI355X (gfx950) implementation of qmb's Hamiltonian operators on bit-packed configurations.

    from qmb_hip import Hamiltonian          # drop-in for qmb.hamiltonian.Hamiltonian
    import qmb_hip.ops                        # torch.ops.qmb_hamiltonian_{nq}_{pc}.* registration
"""

from .hamiltonian import Hamiltonian

__all__ = ["Hamiltonian"]
