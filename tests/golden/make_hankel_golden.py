#!/usr/bin/env python3
"""Write tests/golden/hankel_mpmath.npy: H0^(1)(x), H1^(1)(x) at 30 digits.

Columns: x, Re H0, Im H0, Re H1, Im H1. 2000 log-spaced x in [1e-6, 2e3] plus
200 points around the evaluator's region split at x = 8.
"""
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 30


def main():
    xs = np.concatenate([np.logspace(-6, np.log10(2e3), 2000), np.linspace(7.9, 8.1, 200)])
    rows = []
    for x in xs:
        h0 = mp.hankel1(0, mp.mpf(float(x)))
        h1 = mp.hankel1(1, mp.mpf(float(x)))
        rows.append([x, float(h0.real), float(h0.imag), float(h1.real), float(h1.imag)])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hankel_mpmath.npy")
    np.save(out, np.array(rows))


if __name__ == "__main__":
    main()
