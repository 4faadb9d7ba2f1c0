"""MI355X-native RRQ close evaluation for 3D Laplace single/double layers.

Host-side mirror of qotential's close-evaluation entry points (qol_rrq,
qol_getnearquad_lap_rrq, qol_lap3dsdlpmat_levels) on top of the C-ABI library
libqotential_hip.so, plus the stellarator workload generator.
"""
from ._lib import QolError, lib  # noqa: F401
