"""Writes tests/golden/special_golden.json: independent golden scalars for the
oracle's special functions and luminosity integrals, computed with scipy
(scipy.special for the gamma family, scipy.integrate.quad -- which wraps the
same QUADPACK dqagse/dqagie routines libprofit calls through R -- for the
numerically integrated luminosities, with libprofit's epsabs = epsrel = 1e-4,
limit = 100)."""
import json
import os

import numpy as np
import scipy.integrate as si
import scipy.special as sp


def main():
    out = {"qgamma": [], "pgamma": [], "gamma": [], "beta": [], "quad_finite": [], "quad_infinite": []}
    for n in [0.3, 0.5, 0.8, 1.0, 1.7, 2.5, 4.0, 5.5, 8.0]:
        a = 2 * n
        for p in [0.5, 0.9999, 1 - n * n / 2000]:
            out["qgamma"].append([p, a, float(sp.gammaincinv(a, p))])
    for a in [0.6, 1.0, 3.0, 8.0, 16.0]:
        for x in [0.1, 1.0, 5.0, 20.0]:
            out["pgamma"].append([x, a, float(sp.gammainc(a, x))])
    for x in [0.2, 1.0, 2.5, 5.0, 8.0, 16.0]:
        out["gamma"].append([x, float(sp.gamma(x))])
    for a, b in [(0.5, 0.5), (1 / 2.5, 1 / 2.5), (1 / 3.0, 1 / 3.0), (2.0, 3.0), (1.5, 1 + 2 / (2 - 0.7))]:
        out["beta"].append([a, b, float(sp.beta(a, b))])

    # King r f(r) on [0, rt]
    for rc, rt, a in [(1.0, 3.0, 2.0), (2.5, 30.0, 1.5), (0.7, 12.0, 3.0)]:
        f = lambda r: r * (1 / (1 + (r / rc) ** 2) ** (1 / a) - 1 / (1 + (rt / rc) ** 2) ** (1 / a)) ** a if r < rt else 0.0
        v, _, info = si.quad(f, 0, rt, epsabs=1e-4, epsrel=1e-4, limit=100, full_output=1)[:3]
        out["quad_finite"].append({"kind": "king", "args": [rc, rt, a], "value": v, "neval": info["neval"],
                                   "last": info["last"]})

    # core-Sersic r f(r) on [0, inf)
    for re, rb, nser, a, b in [(5.0, 1.0, 4.0, 2.0, 0.5), (3.0, 0.5, 2.0, 1.0, 1.0)]:
        bn = float(sp.gammaincinv(2 * nser, 0.5))
        f = lambda r: r * (1 + (r / rb) ** (-a)) ** (b / a) * np.exp(-bn * ((r ** a + rb ** a) / re ** a) ** (1 / (nser * a)))
        v, _, info = si.quad(f, 0, np.inf, epsabs=1e-4, epsrel=1e-4, limit=100, full_output=1)[:3]
        out["quad_infinite"].append({"kind": "coresersic", "args": [re, rb, nser, a, b], "value": v,
                                     "neval": info["neval"], "last": info["last"]})
    # broken exponential r f(r) on [0, inf)
    for h1, h2, rb, a in [(2.0, 5.0, 6.0, 1.0), (4.0, 1.5, 10.0, 2.0)]:
        def f(r, h1=h1, h2=h2, rb=rb, a=a):
            base = r - rb
            if a * base < 40:
                base = np.log(1 + np.exp(a * base)) / a
            return r * np.exp(-r / h1 + (1 / h1 - 1 / h2) * base)
        v, _, info = si.quad(f, 0, np.inf, epsabs=1e-4, epsrel=1e-4, limit=100, full_output=1)[:3]
        out["quad_infinite"].append({"kind": "brokenexp", "args": [h1, h2, rb, a], "value": v,
                                     "neval": info["neval"], "last": info["last"]})
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "special_golden.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
