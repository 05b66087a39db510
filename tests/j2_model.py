"""Closed-form response of a homogeneous J2 voxel field under uniaxial stress,
used to pin the return map of the oracle and of the GPU path.

The reference's "ep" behaviour (FFT_analysis3D.cpp:640-691) yields at
sigma_eq = sy + H p^n with p the accumulated equivalent plastic strain, which
starts at EP0 in every voxel (:67-69), and returns radially.  Under uniaxial
stress the radial return puts the whole equivalent plastic increment on the
axial component, so after monotone loading to the axial strain eps the
stress sigma and the plastic strain p satisfy

    sigma = sy + H p^n,    eps = sigma / E + (p - EP0).

The root in p is found by bisection: a method independent of the reference's
Newton iteration, so a slip there (derivative, exponent, initial p) fails.
"""
from __future__ import annotations

import numpy as np

EP0 = 1.236547e-9

# the reference CLI's material table with the input deck's moduli
# (example/3d/main.cpp:107-108, example/3d/in_struct.txt): row 1 is
# power-law hardening with n = 0.2
MAT_REF_TABLE = np.array([[31e9, 0.30, 40e6, 100e6, 1.0, 8e-2, 8e2],
                          [30e9, 0.20, 250e6, 8006.0, 0.2, 8e-4, 8e3]])


def uniaxial_stress(eps, E, sy, H, n):
    s_el = E * eps
    if s_el <= sy + H * EP0 ** n:
        return s_el
    lo, hi = EP0, EP0 + eps

    def g(p):
        return (sy + H * p ** n) / E + (p - EP0) - eps

    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if g(mid) > 0:
            hi = mid
        else:
            lo = mid
    return sy + H * (0.5 * (lo + hi)) ** n


POWER_LAW_CASES = [  # (E, nu, sy, H, n, final axial strain)
    (70e9, 0.30, 100e6, 1e9, 0.2, 0.004),     # strong power-law hardening
    tuple(MAT_REF_TABLE[1, :5]) + (0.02,),    # the reference table's n = 0.2 row
]


def check_uniaxial_rows(P, E, sy, H, n, rtol=1e-8):
    """Every prop_homogenized row [eps_axial, sigma_axial, sigma_2..6, ...]
    against the closed form; returns the number of plastic rows."""
    plastic = 0
    for row in P:
        eps, s0 = row[0], row[1]
        ref = uniaxial_stress(eps, E, sy, H, n)
        plastic += s0 > sy
        assert abs(s0 - ref) < rtol * max(abs(ref), sy), (eps, s0, ref)
        assert np.abs(row[2:7]).max() < 1e-6 * max(abs(s0), 1.0)
    return plastic
