"""genetic_amd: genetIC's field-generation path (white noise -> multi-level
delta -> modifications -> Zel'dovich/2LPT displacements -> particles) on
AMD MI355X, behind a C-ABI (include/gic.h, libgic_hip.so)."""
from . import gic  # noqa: F401
