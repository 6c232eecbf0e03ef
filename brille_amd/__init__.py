"""brille_amd: MI355X implementation of brille's batched Q-point path.

The Python surface mirrors the reference module `brille` for the
``BrillouinZone`` / ``BZTrellisQdc.ir_interpolate_at`` path.  Construction of
lattices, zones and trellises runs on the host (C++); the hot path runs as
HIP kernels on gfx950 behind the C ABI in ``include/brille_amd.h``.  There is
no CPU fallback: importing fails if the native library has not been built,
and interpolation fails if no GPU is present.
"""
import os

_HERE = os.path.dirname(os.path.abspath(__file__))

try:
    from . import _brille  # noqa: F401
except ImportError as exc:  # pragma: no cover - exercised only on a broken build
    raise ImportError("brille_amd native extension is not built; run `python brille_amd/build.py` "
                      f"({exc})") from exc

from ._brille import (  # noqa: E402
    ApproxConfig,
    Basis,
    BrillouinZone,
    BZMeshQdc,
    BZNestQdc,
    BZTrellisQdc,
    Direct,
    LengthUnit,
    Lattice,
    NodeType,
    PointSymmetry,
    Polyhedron,
    Reciprocal,
    RotatesLike,
    Symmetry,
    hall_symbol,
    spacegroup_number,
)
from . import capi  # noqa: E402,F401

__version__ = "0.1.0"
__all__ = ["ApproxConfig", "Basis", "BrillouinZone", "BZMeshQdc", "BZNestQdc", "BZTrellisQdc", "Direct", "LengthUnit", "Lattice",
           "NodeType", "PointSymmetry", "Polyhedron", "Reciprocal", "RotatesLike", "Symmetry", "hall_symbol",
           "spacegroup_number", "capi"]
