"""Synthetic workloads of the boiling path (nondimensional deck units).

Conventions mapping physical sweeps onto the deck (which is nondimensional):
  * Jakob number Ja = 9.0e-4 per kelvin of wall superheat, so 10 K gives the
    deck's JAKOB=0.009 (lica.in:27; c_v/h_fg of water at 1 atm);
  * reference length 0.2 mm / 0.0607, so the deck's interface height 0.0607
    (ASSIGN_LVS two_phase.f:108) is a 0.2 mm film and z0 = h / L_ref;
  * vapour temperature T = z / z0 below the interface (the deck's
    16.47 = 1 / 0.0607, MAKEFLD 1d.f:909), T = 1 in the liquid.
Random draws come from SplitMix64 so that every case is reproducible from
its seed.
"""
import math

from ._abi import CaseDesc

L_REF_MM = 0.2 / 0.0607
JA_PER_K = 9.0e-4
BATCH_SEED = 20261202
MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / (1 << 53))

    def randint(self, lo, hi):
        """uniform integer in [lo, hi]"""
        return lo + self.next() % (hi - lo + 1)


def deck_case(n=256, r=2):
    """The shipped lica.in/grid.in case (values as INPUT_SIMULATION reads them)."""
    return CaseDesc(n=n, r=r, zl=1.0, re=1.42e-2, visr=22.22, denr=1605.36, surf_j=7.36e-8, fr=9.7,
                    prm=1.0231, scr=2.077, tcr=27.16, tsat=1.0, ja=0.009, dt0=2.5e-6, cflmax=0.5,
                    idtopt=1, z0=0.0607, tslope=16.47)


def single_large_case(seed=1, n=8192, r=4, superheat_k=10.0, film_mm=0.2, amp=0.01):
    """Large single case: base grid n, refinement r, the film height perturbed
    by amp * sin(2 pi u) of its thickness with u drawn from the seed (the
    column has no lateral direction, so the perturbation acts on the only
    interface degree of freedom)."""
    u = SplitMix64(seed).uniform()
    z0 = film_mm / L_REF_MM * (1.0 + amp * math.sin(2.0 * math.pi * u))
    return deck_case(n, r).copy(ja=JA_PER_K * superheat_k, z0=z0, tslope=1.0 / z0)


def batched_cases(count=4096, seed=BATCH_SEED, nmin=128, nmax=4096, r=2):
    """Independent cases with N ~ U{nmin..nmax}, superheat ~ U(2,20) K and
    film thickness ~ U(0.05,0.5) mm."""
    g = SplitMix64(seed)
    out = []
    for _ in range(count):
        n = g.randint(nmin, nmax)
        dT = 2.0 + 18.0 * g.uniform()
        h = 0.05 + 0.45 * g.uniform()
        z0 = h / L_REF_MM
        out.append(deck_case(n, r).copy(ja=JA_PER_K * dT, z0=z0, tslope=1.0 / z0))
    return out
