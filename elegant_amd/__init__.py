"""elegant_amd -- MI355X (gfx950) implementation of elegant's per-particle
element tracking (KQUAD/KSEXT/KOCT/KQUSE, CSBEND, DRIF/EDRIFT, RFCA, MAXAMP sweeps).

The compute path is libelegant_hip.so (built from csrc/ by build.py); this
package only marshals lattices and beams to its C ABI.
"""
from . import structs
from .lattice import Element, Lattice

__all__ = ["structs", "Element", "Lattice", "tracking", "configs", "optics"]
