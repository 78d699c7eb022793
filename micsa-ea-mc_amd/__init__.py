"""micSA on MI355X: multispin-coded microcanonical simulated annealing of the
3D Edwards-Anderson +-J spin glass, one sample per bit lane, on gfx950.

The product is C/HIP (libmsa_hip.so, libmsa_host.so, bin/msa_anneal); this
package only exposes ctypes bindings of it for Python callers.  The directory
name contains hyphens, so load it by path (see load_package in bench.py).
"""
from .msa import *  # noqa: F401,F403
from .msa import Lattice, MsaError, footprint_bytes, hip  # noqa: F401
