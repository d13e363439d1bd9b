"""hbtplus_amd -- MI355X-native HBT+ subhalo unbinding (gfx950 HIP behind a C-ABI)."""
from . import abi, synth  # noqa: F401

__version__ = "0.1.0"
