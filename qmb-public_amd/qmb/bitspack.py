"""``qmb.bitspack`` import path (qmb/bitspack.py of the reference) served by qmb_hip."""

from qmb_hip.bitspack import pack_int, unpack_int

__all__ = ["pack_int", "unpack_int"]
