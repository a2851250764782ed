"""sfb_amd -- MI355X-native cat2amln with the reference's Julia-level API.

Mirrors the pieces of hsgg/SphericalFourierBesselDecompositions.jl that the
catalogue -> a_nlm path uses:

    AnlmModes, getidx, getnlm, getlmsize, getnlmsize, estimate_nside   (src/modes.jl)
    ConfigurationSpaceModes, window_r, integrate_window, win_rhat_ln   (src/windows.jl)
    SeparableArray                                                     (src/SeparableArrays.jl)
    cat2amln / cat2nlm                                                  (src/cat2anlm.jl)

All numerical work runs in libsfb_mi355x.so (HIP for gfx950 plus native host
setup); this package only marshals arguments.
"""
from ._lib import SFBError, lib  # noqa: F401
from .cat2anlm import (Plan, anlm2field, ang2pix_ring, cat2amln, cat2amln_vjp, cat2nlm, field2anlm,  # noqa: F401
                       get_plan)  # noqa: F401
from .catalog import apply_window, winweights2galweights, xyz2rtp  # noqa: F401
from .clnn import ClnnModes, amln2clnn, anlm2clnn, getlnn, getlnnsize, healpix_pixwin, pixwin  # noqa: F401
from .modes import (AnlmModes, GNL, Spline1D, estimate_nside, getidx, getlmsize, getnlm,  # noqa: F401
                    getnlmsize, isvalidnlm)
from .windows import (ConfigurationSpaceModes, SeparableArray, get_rbounds, integrate_window,  # noqa: F401
                      win_rhat_ln, window_r)

__all__ = [
    "AnlmModes", "GNL", "Spline1D", "estimate_nside", "getidx", "getlmsize", "getnlm", "getnlmsize",
    "isvalidnlm", "ConfigurationSpaceModes", "SeparableArray", "get_rbounds", "integrate_window",
    "win_rhat_ln", "window_r", "xyz2rtp", "apply_window", "winweights2galweights", "ClnnModes",
    "amln2clnn", "anlm2clnn", "getlnn", "getlnnsize", "pixwin", "healpix_pixwin", "cat2amln", "cat2amln_vjp", "cat2nlm", "field2anlm", "anlm2field", "Plan", "get_plan", "ang2pix_ring", "SFBError",
]
