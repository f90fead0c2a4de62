"""Poling design helpers that feed the per-step poling of the engine
(``polingProfile=`` / ``setPolingProfile``) and take an optimised profile back to
domains, after the reference's poling.py:8-29 and :439-470.

  periodicPoling(deltaBeta0, L)      domain lengths of first-order quasi-phase matching
  linearPoling(k0, kf, L, dL)        domains with a linearly swept phase-matching wavenumber
  domainsToSpace(poling, nZSteps)    domain lengths -> per-step orientation, the
                                     discretisation of _NonlinearMedium::setPoling
                                     (src/_NonlinearMedium.cpp:470-503)
  spaceToDomains(profile, length)    per-step profile -> domain lengths of its sign
                                     pattern and the first orientation (a
                                     manufacturable design from a continuous profile,
                                     e.g. one optimised with the adjoint)
"""
import math

import numpy as np


def periodicPoling(deltaBeta0, L):
    """Domains of half the coherence length 2 pi / |deltaBeta0| over a medium of
    length L; the last domain is the remaining fraction, and the lengths are
    rescaled to sum to L.  (When L is a whole number of domains the reference
    appends a zero-length domain, which setPoling rejects; here there is none.)"""
    half = math.pi / abs(deltaBeta0)
    n = L / half
    whole = int(n)
    frac = n - whole
    lengths = np.ones(whole + (1 if frac > 0 else 0))
    if frac > 0:
        lengths[-1] = frac
    return lengths * (L / lengths.sum())


def linearPoling(k0, kf, L, dL):
    """Domains whose quasi-phase-matching wavenumber sweeps linearly from k0 at z = 0
    to kf at z = L (poling.py:17-29 of the reference): the orientation is the sign of
    sin(phi(z)), phi(z) = (kf - k0) z^2 / (2 L) + k0 z (whose rate k0 + (kf - k0) z / L
    is the local wavenumber), sampled at round(L / dL) points from 0 to L inclusive
    (a zero counts as +1), and the domains are the runs of equal sign, each sample
    counting dL."""
    n = int(round(L / dL))
    z = np.linspace(0.0, L, n)
    s = np.sign(np.sin(0.5 * (kf - k0) * z * z / L + k0 * z))
    s[s == 0] = 1.0
    return _run_lengths(s) * dL


def _run_lengths(s):
    """Lengths (in samples) of the runs of equal values of s."""
    if s.size == 0:
        return np.zeros(0)
    edges = np.flatnonzero(s[1:] != s[:-1]) + 1
    return np.diff(np.concatenate([[0], edges, [s.size]])).astype(np.float64)


def domainsToSpace(poling, nZSteps):
    """Orientation (+1 / -1, alternating from +1) of each of nZSteps steps for
    domains of the given lengths, scaled to the step grid.  A step holding a
    domain boundary at fractional position f gets direction * (2 f - 1), the
    value setPoling assigns; fewer than two domains give a uniform medium."""
    d = np.asarray(poling, dtype=np.float64).ravel()
    out = np.ones(nZSteps)
    if d.size <= 1:
        return out
    if np.any(d <= 0):
        raise ValueError("Poling contains non-positive length domains")
    ends = np.cumsum(d)
    ends *= nZSteps / ends[-1]
    start = 0
    sign = 1.0
    for e in ends:
        k = int(e)
        if e < start:
            raise ValueError("Poling period too small for given resolution")
        out[start:k] = sign
        if k < nZSteps:
            out[k] = sign * (2 * abs(math.fmod(e, 1.0)) - 1)
        sign = -sign
        start = k + 1
    return out


def spaceToDomains(profile, length):
    """Domain lengths of the sign pattern of a per-step profile over a medium of
    the given length, and the orientation of the first domain (+1 or -1).  Zeros
    continue the previous orientation (+1 at the start).  For a +1-first pattern
    the lengths reproduce the profile's signs through domainsToSpace when the
    boundaries fall on step edges."""
    p = np.asarray(profile, dtype=np.float64).ravel()
    if p.size == 0:
        raise ValueError("empty profile")
    s = np.sign(p)
    for i in range(s.size):
        if s[i] == 0:
            s[i] = s[i - 1] if i > 0 else 1.0
    return _run_lengths(s) * (length / s.size), float(s[0])
