"""nfempy_amd: MI355X-native assembly-and-solve path for nfempy's mixed finite elements.

Host side mirrors nfempy's mesh / product-space / weak-form / assembler API;
compute runs through libnfx (HIP, gfx950), see include/nfx.h.
"""

__version__ = "0.1.0"
