"""SCF driver over the C ABI (the oneComp/twoComp loops of fJmodels.cpp:38-144).

The per-iteration work (moments, projection, Poisson) runs on the GPU through
fj_compute_new_phi; this module only sequences the calls as the reference
driver does and sums component potentials (updatePhil, potLeg.cpp:21-29).
"""
import numpy as np

from .capi import Context, df_params, FJ_POISSON_QUADPACK


def log_grid(n, rmin, rmax):
    """SetGrid(rmax) (include/Utils.h:103-108) with an explicit rmin."""
    i = np.arange(n)
    return np.array([10.0 ** (np.log10(rmin) + (np.log10(rmax) - np.log10(rmin)) * k / (n - 1)) for k in i])


class Model:
    """One- or two-component model (fJmodels.cpp:38-144) on one context."""

    def __init__(self, ctx: Context, ar, ntheta, npoly, components, ngrid=30,
                 poisson_mode=FJ_POISSON_QUADPACK):
        self.ctx = ctx
        self.components = components   # list of dicts: model, mass, r0, q, df (DFParams)
        self.mode = poisson_mode
        ctx.set_grid(ar, ntheta, npoly, ngrid)

    def setup(self):
        """Guess potential of every component, total potential, Delta table
        (fJmodels.cpp:47-63 / 89-127)."""
        c = self.ctx
        self.philP = []
        for comp in self.components:
            rhl = c.guess_rhl(comp["model"], comp["mass"], comp["r0"], comp.get("q", 1.0))
            self.philP.append(c.poisson(rhl, self.mode))
        self._install_total()
        self.Egrid, self.Dgrid = c.tabulate_delta()

    def _install_total(self):
        tot = [sum(p[k] for p in self.philP) for k in range(3)]
        self.ctx.set_potential(*tot)
        self.total = tot

    def iterate(self):
        """One SCF iteration: computeNewPhi per component, then updatePhil."""
        out = []
        for k, comp in enumerate(self.components):
            self.ctx.set_df(comp["df"])
            res = self.ctx.compute_new_phi(self.mode)
            self.philP[k] = (res["phil"], res["Pr"], res["Pr2"])
            out.append(res)
        self._install_total()
        return out
