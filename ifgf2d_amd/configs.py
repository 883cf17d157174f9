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


_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed, n, start=0):
    """Counter-based splitmix64 (Vigna): value i = mix(seed + (start + i + 1) * golden
    gamma), so any index range is reproducible on its own; the same generator runs in
    the oracle's C++ (oracle_splitmix)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.arange(start, start + n, dtype=np.uint64) + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def seeded_density(n, seed=0, column=0):
    """The benchmark's seeded complex density: entry i of column c takes the pair of
    splitmix64 values (2m, 2m + 1), m = c * n + i, as uniforms u = (v >> 11 + 1) / 2^53
    in (0, 1] and returns their Box-Muller normals r cos(2 pi u2) + i r sin(2 pi u2),
    r = sqrt(-2 ln u1): real and imaginary parts iid N(0, 1)."""
    v = splitmix64(seed, 2 * n, start=2 * column * n)
    u = ((v >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * (1.0 / 9007199254740992.0)
    r = np.sqrt(-2.0 * np.log(u[0::2]))
    a = 2.0 * np.pi * u[1::2]
    return r * np.cos(a) + 1j * (r * np.sin(a))
