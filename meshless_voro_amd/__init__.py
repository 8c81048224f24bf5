"""meshless_voro_amd -- MI355X-native (gfx950 HIP) per-cell Voronoi
construction with the API of the meshless_voronoi crate's construction path
(src/lib.rs:108-111 re-exports: Voronoi, VoronoiCell, VoronoiFace,
Dimensionality)."""
from .voronoi import (Dimensionality, Voronoi, VoronoiCell, VoronoiFace, VoronoiError,
                      BuildOptions, BuildGraph, MODE_PLAIN, MODE_LEARNED)
from . import _native

__all__ = ["Dimensionality", "Voronoi", "VoronoiCell", "VoronoiFace", "VoronoiError",
           "BuildOptions", "BuildGraph", "MODE_PLAIN", "MODE_LEARNED"]
