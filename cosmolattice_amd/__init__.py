"""MI355X-native lattice evolution for CosmoLattice models (HIP kernels behind the libclmi C ABI)."""
from . import clmi  # noqa: F401

__all__ = ["clmi"]
