"""Host side of the MI355X nonbonded-force path of SPONGE (Python binding).

The compute lives in libsponge_nb.so (HIP, gfx950); this package loads it
(lib), writes/reads SPONGE native inputs (native_io) and generates the
synthetic benchmark systems (systems).
"""
__all__ = ["lib", "native_io", "systems"]
