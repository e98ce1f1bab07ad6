"""MI355X-native BioBloom Tools read-classification path.

libbbt_hip.so (gfx950 kernels behind the C ABI in include/bbt_hip.h) does the
work; this package binds it (lib), drives device batches (device) and makes
synthetic workloads (synth).
"""
from . import lib  # noqa: F401
