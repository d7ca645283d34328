"""Hamiltonian Monte Carlo over a batch of chains driven by the adjoint
gradient, the way LaplacesDemon's HMC drives ProFit's likelihood
(R/profitLikeModel.R:39 returns LP for the sampler; the gradient is the C-ABI's
profit_amd_engine_like_grad).

Parameters move in a unit-scaled space: theta = theta0 + scale * u, so that a
step size of 0.01 means 1% of each parameter's scale.  Every leapfrog step is
one batched like_grad call over all chains.
"""
from __future__ import annotations

import numpy as np

from . import GRAD_SLOTS, PROFILE_TYPES, SHAPE_NAMES, make_profile

SLOT = {n: i for i, n in enumerate(GRAD_SLOTS)}


def grad_slot(ptype, name):
    """Gradient slot of a parameter of a profile type (shape parameters map to s0..s4)."""
    if name in SLOT:
        return SLOT[name]
    return SLOT["s%d" % SHAPE_NAMES[PROFILE_TYPES[ptype]].index(name)]


# parameter ranges libprofit accepts (profile validate(): radial.cpp, sersic.cpp,
# moffat.cpp, ferrer.cpp, king.cpp); positions outside them have logL = -inf
BOUNDS = {"axrat": (0.0, 1.0), "re": (0.0, np.inf), "nser": (0.0, np.inf), "fwhm": (0.0, np.inf),
          "con": (1.0, np.inf), "rout": (0.0, np.inf), "rc": (0.0, np.inf), "rt": (0.0, np.inf)}


class Problem:
    """logL and its gradient in unit-scaled coordinates for `nchain` copies of
    one component list on one galaxy (the engine's batch is set here)."""

    def __init__(self, engine, galaxy_id, components, free, scale, nchain):
        self.eng = engine
        self.base = [dict(c) for c in components]
        self.free = list(free)  # (component index, parameter name)
        self.scale = np.asarray(scale, dtype=np.float64)
        self.theta0 = np.array([self.base[k][n] for k, n in self.free], dtype=np.float64)
        self.nchain = nchain
        self.ncomp = len(components)
        engine.set_batch([galaxy_id] * nchain, [self.ncomp] * nchain)

    def theta(self, u):
        return self.theta0[None, :] + self.scale[None, :] * u

    def valid(self, u):
        """Chains whose parameters lie inside the accepted ranges (axrat may equal 1)."""
        th = self.theta(u)
        ok = np.ones(u.shape[0], dtype=bool)
        for j, (_, n) in enumerate(self.free):
            lo, hi = BOUNDS.get(n, (-np.inf, np.inf))
            ok &= (th[:, j] > lo) & ((th[:, j] <= hi) if n == "axrat" else (th[:, j] < hi))
        return ok

    def _params(self, u):
        # chains out of range are evaluated at the centre and reported as -inf
        u = np.where(self.valid(u)[:, None], u, 0.0)
        th = self.theta(u)
        out = []
        for c in range(self.nchain):
            comps = [dict(x) for x in self.base]
            for j, (k, n) in enumerate(self.free):
                comps[k][n] = th[c, j]
            for comp in comps:
                comp = dict(comp)
                t = comp.pop("type")
                out.append(make_profile(t, convolve=True, **comp))
        return out

    def logp_grad(self, u):
        """logL per chain and d logL / d u, shape (nchain, nfree)."""
        ll, g = self.eng.like_grad(self._params(u))
        ll = np.asarray(ll, dtype=np.float64)
        g = g.reshape(self.nchain, self.ncomp, -1)
        du = np.empty_like(u)
        for j, (k, n) in enumerate(self.free):
            du[:, j] = g[:, k, grad_slot(self.base[k]["type"], n)] * self.scale[j]
        bad = ~self.valid(u)
        ll[bad] = -np.inf
        du[bad] = 0.0
        return ll, du

    def logp(self, u):
        ll = np.asarray(self.eng.like(self._params(u)), dtype=np.float64)
        ll[~self.valid(u)] = -np.inf
        return ll


def unit_scales(engine, galaxy_id, components, free, rel=1e-4):
    """Posterior widths 1/sqrt(-d2 logL / d theta^2) at `components`, from
    central differences of the adjoint gradient (one batched call)."""
    nfree = len(free)
    theta0 = np.array([components[k][n] for k, n in free], dtype=np.float64)
    h = rel * np.maximum(np.abs(theta0), 1e-3)
    prob = Problem(engine, galaxy_id, components, free, np.ones(nfree), 2 * nfree)
    u = np.zeros((2 * nfree, nfree))
    for j in range(nfree):
        u[2 * j, j] = h[j]
        u[2 * j + 1, j] = -h[j]
    _, g = prob.logp_grad(u)
    curv = np.array([(g[2 * j, j] - g[2 * j + 1, j]) / (2 * h[j]) for j in range(nfree)])
    if np.any(curv >= 0):
        raise ValueError("logL is not locally concave along every free parameter")
    return 1.0 / np.sqrt(-curv)


def hmc(problem, u0, n_leapfrog=1000, step=0.01, path=10, seed=0):
    """Runs n_leapfrog leapfrog steps in trajectories of `path` steps on every
    chain (identity mass matrix, Metropolis accept/reject per trajectory).
    Returns the final positions, the per-trajectory energy errors, the
    acceptance rate and the logL of the final positions."""
    rng = np.random.default_rng(seed)
    u = np.array(u0, dtype=np.float64)
    ll, g = problem.logp_grad(u)
    n_traj = max(1, n_leapfrog // path)
    dH = np.zeros((n_traj, u.shape[0]))
    accepted = 0
    for t in range(n_traj):
        p = rng.standard_normal(u.shape)
        h0 = -ll + 0.5 * np.sum(p * p, axis=1)
        un, pn, gn = u.copy(), p + 0.5 * step * g, None
        for s in range(path):
            un = un + step * pn
            lln, gn = problem.logp_grad(un)
            if s < path - 1:
                pn = pn + step * gn
        pn = pn + 0.5 * step * gn
        h1 = -lln + 0.5 * np.sum(pn * pn, axis=1)
        dH[t] = h1 - h0
        acc = np.log(rng.random(u.shape[0])) < -(h1 - h0)
        acc &= np.isfinite(h1)
        u[acc] = un[acc]
        ll[acc] = lln[acc]
        g[acc] = gn[acc]
        accepted += int(acc.sum())
    return dict(u=u, ll=ll, dH=dH, acceptance=accepted / (n_traj * u.shape[0]))
