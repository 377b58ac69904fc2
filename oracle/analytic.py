"""Closed-form grounded-sphere solutions (Kelvin image), numpy FP64.

Test infrastructure only (see oracle/__init__.py).

Restates core/images.py:165-203 (image charge q' = -q R/d at (R^2/d^2) r0 for a
grounded sphere of radius R centred at the origin, d = |r0| > R) and the
textbook induced surface charge density on that sphere,
    sigma(x) = -q (d^2 - R^2) / (4 pi R |x - r0|^3),
whose total is -q R / d.  Units follow the reference: potentials are
K_E * charge / distance, so sigma is a charge per area.

Parity: pinned in tests by the total induced charge -qR/d (closed form) and by
zero potential on the sphere surface.
"""
import numpy as np

from .constants import K_E


def kelvin_image(R, r0, q):
    r0 = np.asarray(r0, dtype=float)
    d = float(np.linalg.norm(r0))
    return -q * R / d, (R * R / (d * d)) * r0


def kelvin_sigma(points, R, r0, q):
    p = np.asarray(points, dtype=float)
    r0 = np.asarray(r0, dtype=float)
    d2 = float(r0 @ r0)
    dist = np.linalg.norm(p - r0[None, :], axis=1)
    return -q * (d2 - R * R) / (4.0 * np.pi * R * dist ** 3)


def point_charge_potential(points, r0, q):
    p = np.asarray(points, dtype=float)
    return K_E * q / np.maximum(np.linalg.norm(p - np.asarray(r0)[None, :], axis=1), 1e-12)


def sphere_capacitance(R):
    """Q / V = R / K_E (= 4 pi eps0 R)."""
    return R / K_E
