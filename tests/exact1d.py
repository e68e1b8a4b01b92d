"""Exact NPT averages of 1D nearest-neighbour systems (Takahashi 1942), and
exact fixed-length averages of two and three particles between walls.

With nearest-neighbour interactions only, the NPT ensemble factorises into
independent gaps with density w(g) ~ exp(-beta (phi(g) + P g)).

walls (reference geometry, N ln s or (N+1) ln s Jacobian with all N
coordinates scaled): N-1 interior gaps of law w plus two wall gaps of law
exp(-beta P g):
    <L> = 2/(beta P) + (N-1) E[g],   <U> = (N-1) E[phi]

periodic ring (the N gaps sum to L, and the Jacobian makes the gap law
weighted by L = sum g):
    <L> = N E[g] + Var[g]/E[g],      <U> = E[phi g]/E[g] + (N-1) E[phi]
"""
import math

import numpy as np
from scipy.integrate import quad


def gap_moments(phi, beta, P, gmin, gmax=math.inf, breaks=()):
    # exponent offset: the smallest beta (phi + P g) on a grid, so the
    # integrands stay below 1
    top = gmin + 50.0 / max(beta * P, 1e-3) if math.isinf(gmax) else gmax
    grid = np.linspace(max(gmin, 1e-9), top, 20001)
    ref = min(beta * (phi(g) + P * g) for g in grid)
    pts = [gmin, *breaks]

    def integ(f):
        tot = 0.0
        edges = pts + [gmax]
        for a, b in zip(edges[:-1], edges[1:]):
            if b <= a:
                continue
            v, _ = quad(lambda g: f(g) * math.exp(-beta * (phi(g) + P * g) + ref), a, b,
                        limit=400, epsabs=0, epsrel=1e-13)
            tot += v
        return tot

    Z = integ(lambda g: 1.0)
    m = integ(lambda g: g) / Z
    m2 = integ(lambda g: g * g) / Z
    ephi = integ(phi) / Z
    ephig = integ(lambda g: phi(g) * g) / Z
    ephi2 = integ(lambda g: phi(g) * phi(g)) / Z
    return dict(m=m, v=m2 - m * m, ephi=ephi, ephig=ephig, ephi2=ephi2,
                lnZ=math.log(Z) - ref)


def walls_npt(N, mom, beta, P):
    return 2.0 / (beta * P) + (N - 1) * mom["m"], (N - 1) * mom["ephi"]


def periodic_npt(N, mom):
    m, v = mom["m"], mom["v"]
    return N * m + v / m, mom["ephig"] / m + (N - 1) * mom["ephi"]


def walls_nvt_small(N, L, beta, phi, gmin, nbn=-1, breaks=()):
    """<U> of N = 2 or 3 particles between walls at +-L/2 at fixed L: the
    configurational integral over the interior gaps, the two wall gaps
    integrated out as the factor (L - sum of interior gaps).  With N = 3
    the outer pair (separation g1 + g2) interacts unless nbn is 1."""
    pts = sorted(b for b in (gmin, *breaks) if gmin <= b <= L)
    if N == 2:
        def f(d, k):
            e = math.exp(-beta * phi(d))
            return (L - d) * e * (phi(d) if k else 1.0)
        Z = sum(quad(lambda d: f(d, 0), a, b, limit=400, epsrel=1e-12)[0]
                for a, b in zip(pts, pts[1:] + [L]) if b > a)
        EU = sum(quad(lambda d: f(d, 1), a, b, limit=400, epsrel=1e-12)[0]
                 for a, b in zip(pts, pts[1:] + [L]) if b > a)
        return EU / Z
    assert N == 3
    outer = nbn < 0 or nbn >= 2

    def u3(g1, g2):
        return phi(g1) + phi(g2) + (phi(g1 + g2) if outer else 0.0)

    def inner(g1, k):
        hi = L - g1
        if hi <= gmin:
            return 0.0
        cuts = sorted({gmin, hi, *(b for b in breaks if gmin < b < hi),
                       *(b - g1 for b in breaks if gmin < b - g1 < hi)})
        tot = 0.0
        for a, b in zip(cuts[:-1], cuts[1:]):
            tot += quad(lambda g2: (L - g1 - g2) * math.exp(-beta * u3(g1, g2))
                        * (u3(g1, g2) if k else 1.0), a, b, limit=200, epsrel=1e-11)[0]
        return tot

    def outer_int(k):
        cuts = sorted({gmin, L - gmin, *(b for b in breaks if gmin < b < L - gmin)})
        return sum(quad(lambda g1: inner(g1, k), a, b, limit=200, epsrel=1e-10)[0]
                   for a, b in zip(cuts[:-1], cuts[1:]))

    return outer_int(1) / outer_int(0)


def yukawa_phi(A=1.0, kappa=0.2, rc=20.0):
    return lambda g: A * math.exp(-kappa * g) / g if g <= rc else 0.0


def hrsw_phi(rod=1.0, eps=1.0, edge=1.5):
    return lambda g: (-eps if g < edge else 0.0) if g >= rod else math.inf


def lj_phi(eps=1.0, sigma=1.0, rc=math.inf, shift=0.0):
    def f(g):
        if g > rc:
            return 0.0
        s6 = (sigma / g) ** 6
        return 4 * eps * (s6 * s6 - s6) - shift
    return f


def harmonic_phi():
    return lambda g: (g - 1.0) ** 2


def block_mean_sem(x, nblocks=20):
    x = np.asarray(x, dtype=float)
    n = len(x) // nblocks * nblocks
    b = x[:n].reshape(nblocks, -1).mean(axis=1)
    return float(b.mean()), float(b.std(ddof=1) / math.sqrt(nblocks))
