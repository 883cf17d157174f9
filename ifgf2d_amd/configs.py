"""Workload definitions: the scattering problems the forward map is measured on.

Each returns the keyword arguments of ``ForwardMap`` plus the incidence
direction. Patches are equal-parameter, Np = 10 Fejer first-rule nodes each;
nlevels is chosen for ~30 points per leaf box; the near-singular distance is
0.75 x the shortest patch chord; the RP change of variables uses p = 6.
"""
import numpy as np

from . import circle, equal_parameter_lims, kite, star


def circle_k10(**kw):
    """Unit circle, k = 10, 40 patches (N = 400), nlevels 3, d = (1, 0)."""
    return dict(curve=circle(), patch_lims=equal_parameter_lims(40), k=10.0, nlevels=3, p=6.0,
                **kw), (1.0, 0.0)


def kite_k200(**kw):
    """Kite x = cos t + 0.65 cos 2t - 0.65, y = 1.5 sin t, k = 200, 6400 patches (N = 64,000),
    nlevels 10, d = (cos pi/4, sin pi/4)."""
    return dict(curve=kite(), patch_lims=equal_parameter_lims(6400), k=200.0, nlevels=10, p=6.0,
                **kw), (np.cos(np.pi / 4), np.sin(np.pi / 4))


def star_k500(**kw):
    """Star r = 1 + 0.3 cos 5t, k = 500, 25,600 patches (N = 256,000), nlevels 12, d = (1, 0)."""
    return dict(curve=star(a=[0, 0, 0, 0, 0, 0.3]), patch_lims=equal_parameter_lims(25600),
                k=500.0, nlevels=12, p=6.0, **kw), (1.0, 0.0)


def random_stars(count=16, k=100.0, seed=0, nodes_per_wavelength=10, **kw):
    """Seeded random star obstacles r(t) = R (1 + sum_{m=2..8} a_m cos mt + b_m sin mt),
    a_m, b_m ~ U(-0.05, 0.05), R ~ U(0.5, 4), at wavenumber k: Np = 10 Fejer nodes per
    patch and patches of about one wavelength (~10 nodes per wavelength), nlevels for
    ~30 points per leaf box. Returns a list of ForwardMap keyword dicts."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        R = rng.uniform(0.5, 4.0)
        a = [0.0, 0.0] + list(rng.uniform(-0.05, 0.05, 7))
        b = [0.0, 0.0] + list(rng.uniform(-0.05, 0.05, 7))
        npatch = int(np.ceil(1.1 * R * k * 10 / nodes_per_wavelength))  # ~ perimeter / wavelength
        nlev = max(3, int(np.round(np.log2(npatch * 10 / 100))) + 1)
        out.append(dict(curve=star(R=R, a=a, b=b), patch_lims=equal_parameter_lims(npatch), k=k,
                        nlevels=nlev, p=6.0, **kw))
    return out


def scaled(curve, npatches, k, nlevels, **kw):
    """Any curve at any size (used for reduced-size parity cases)."""
    return dict(curve=curve, patch_lims=equal_parameter_lims(npatches), k=k, nlevels=nlevels, p=6.0,
                **kw)


def plane_wave(x, y, k, d):
    """Incident field exp(i k d.x)."""
    return np.exp(1j * k * (d[0] * x + d[1] * y))
