"""MI355X-native block-sparse tensor engine and PEPS amplitude path.

``tnsp_amd.TAT``   drop-in for the reference ``TAT`` binding (device tensors)
``tnsp_amd.peps``  sampled-configuration single-layer PEPS amplitude, plain
                   per-configuration path and batched engine
"""
from ._lib import LIB_PATH, lib  # noqa: F401

__all__ = ["LIB_PATH", "lib"]
