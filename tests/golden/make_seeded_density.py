#!/usr/bin/env python3
"""Write tests/golden/seeded_density.json: the benchmark generator's first values.

An independent restatement in Python integers and mpmath (no numpy, no
oracle): splitmix64 value i of seed s is mix(s + (i + 1) * 0x9E3779B97F4A7C15)
mod 2^64; a density entry takes values (2i, 2i + 1) as uniforms
u = ((v >> 11) + 1) / 2^53 and returns the Box-Muller pair
sqrt(-2 ln u1) (cos 2 pi u2, sin 2 pi u2), here rounded from 30 digits.
Stored: the first 16 values of seeds 0 and 12345, and the first 8 complex
entries of seed 0 (real, imaginary).
"""
import json
import os

import mpmath as mp

mp.mp.dps = 30
MASK = (1 << 64) - 1


def splitmix(seed, i):
    z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def main():
    out = {"splitmix64": {str(s): [str(splitmix(s, i)) for i in range(16)] for s in (0, 12345)}}
    dens = []
    for i in range(8):
        u1 = mp.mpf((splitmix(0, 2 * i) >> 11) + 1) / 2**53
        u2 = mp.mpf((splitmix(0, 2 * i + 1) >> 11) + 1) / 2**53
        r = mp.sqrt(-2 * mp.log(u1))
        dens.append([float(r * mp.cos(2 * mp.pi * u2)), float(r * mp.sin(2 * mp.pi * u2))])
    out["density_seed0"] = dens
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "seeded_density.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
