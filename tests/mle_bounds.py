"""How far two runs of the reference maximiser may end apart.

GSL Brent stops once its bracket is within 1% of its estimate
(src/PhyloCSFModel.ml:89-95, relative to the mlin/PhyloCSF tree), so a run
ends within |ln x - ln x*| <= -ln(0.99) < 0.0101 of the maximum x* of lpr,
where lpr is below its maximum by at most c d^2 / 2, c = |d2 lpr / d(ln
rho)^2| (lpr is flat to second order there).  Two runs whose likelihoods
differ only in rounding -- FP64 against an MFMA precision -- can therefore
end at lpr values that differ by that much per model, whatever the rounding
did to the path between.  A run that stops at a bound of [0.01, 10] (or at
the same rho as the other run) differs only by the rounding of one
evaluation.  c is taken from the FP64 adjoint's d lpr / d rho at
ln rho +- 0.005, with a factor 1.5 for its change across the bracket.
"""
import numpy as np

import phylocsf_amd as P

D_LN_RHO = 0.0101


def curvature(batch, rho):
    """|d2 lpr / d(ln rho)^2| per (region, model) from FP64 adjoint gradients."""
    d = 0.005
    hs = []
    for s in (np.exp(d), np.exp(-d)):
        batch.forward(rho * s, P.PCSF_FP64, per_region=True)
        g, _ = batch.backward(np.ones_like(rho))
        hs.append(rho * s * g)
    return np.abs(hs[0] - hs[1]) / (2 * d)


def brent_stop_bound_db(batch, rho_a, rho_b):
    """Per region: the largest |score_a - score_b| in decibans that the 1%
    stopping rule allows between two runs ending at rho_a and rho_b
    ([R, 2] each), before the rounding of the final evaluations."""
    c = curvature(batch, rho_a)
    same = rho_a == rho_b
    per_model = np.where(same, 0.0, 0.5 * 1.5 * c * D_LN_RHO ** 2)
    return P.DB * per_model.sum(axis=1)
