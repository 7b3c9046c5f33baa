"""
Import path of the reference package for the operators this build provides: ``qmb.hamiltonian``
and ``qmb.bitspack`` resolve to the MI355X implementation in ``qmb_hip``.
"""
