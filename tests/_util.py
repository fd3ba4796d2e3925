"""Shared helpers of the test suite."""
import numpy as np

from musica2_newgrid_amd import grid as gr


def relerr(A, B):
    """max over the interior of |A - B| divided by the magnitude of the
    variable: max|rho| for rho, max|rho*vec(u)| for the three momentum
    components (one of them may start at zero, e.g. rho*w in the TGV),
    max|rho*E| for the energy.  Infinite when A holds a non-finite value
    (a NaN would otherwise compare as no error)."""
    Bi = [gr.interior(b) for b in B]
    mom = float(np.max(np.sqrt(Bi[1] ** 2 + Bi[2] ** 2 + Bi[3] ** 2)))
    scale = [float(np.max(np.abs(Bi[0]))), mom, mom, mom, float(np.max(np.abs(Bi[4])))]
    out = 0.0
    for a, b, s in zip(A, Bi, scale):
        d = np.abs(gr.interior(a) - b)
        if not np.all(np.isfinite(d)):
            return float("inf")
        out = max(out, float(np.max(d)) / s)
    return out
