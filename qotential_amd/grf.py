"""Green's-representation (GRF) check on top of the engine.

For a density u harmonic inside the surface, S[du/dn] - D[u] equals u at
interior targets and 0 at exterior ones; on the surface the exterior limit is 0
(stellarator_grf_core_mod.f90:909-928 uses max|u|/max|du/dn| there).
"""
import numpy as np

from .engine import Engine


def layer_potential(geo, tx, ub, ubn, exterior=True, side_tol=1e-10, chart=None, engine=None):
    """S[ubn] - D[ub] at targets tx: far field by direct summation plus the
    close correction (RRQ / levels minus the naive block) on the near list."""
    eng = engine or Engine(geo, chart=chart)
    eng.set_targets(tx)
    eng.nearlist()
    eng.eval(exterior=exterior, side_tol=side_tol)
    far = eng.far(ub, ubn)
    u, val = eng.apply(ub, ubn, far)
    return u, eng
