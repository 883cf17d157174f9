#!/usr/bin/env python3
"""Generate the Chebyshev coefficient tables of the project's FP64 Bessel/Hankel
evaluator (``bessel_coeffs.h``) with mpmath at 40 significant digits.

The evaluator replaces the Amos ZBESH call that the reference makes for every
kernel evaluation (ForwardMap/GreenFunctions.hpp:49,56 ->
complex_bessel-master/include/complex_bessel_bits/besselFunctions.h:312-341).
Only orders 0 and 1 at real positive argument are needed.

Two regions, split at x = 8:

* 0 < x <= 8, variable t = x^2/32 - 1 in [-1, 1]:
    J0(x) = sum a_k T_k(t)
    J1(x) = x * sum b_k T_k(t)
    Y0(x) = (2/pi) ln(x/2) J0(x) + sum c_k T_k(t)
    Y1(x) = (2/pi) ln(x/2) J1(x) - 2/(pi x) + x * sum d_k T_k(t)
  Every fitted function is an entire function of x^2 (A&S 9.1.10-9.1.11), so
  the coefficients decay super-geometrically.

* x > 8, variable z = 64/x^2 in (0, 1), t = 2z - 1:
    H_n(x) = sqrt(2/(pi x)) (P_n + i (8/x) Qt_n) exp(i(x - (2n+1) pi/4))
  with P_n, Qt_n smooth functions of 1/x^2 (the amplitude functions of
  A&S 9.2.5-9.2.6, Qt_n = Q_n x/8).

Run:  python gen_bessel_coeffs.py > bessel_coeffs.h
"""
import sys

import mpmath as mp

mp.mp.dps = 40
NFIT = 48          # sample count of the discrete Chebyshev transform
TOL = mp.mpf("1e-19")


def cheb_coeffs(f, n=NFIT):
    nodes = [mp.cos(mp.pi * (j + mp.mpf(1) / 2) / n) for j in range(n)]
    vals = [f(t) for t in nodes]
    out = []
    for k in range(n):
        s = mp.fsum(vals[j] * mp.cos(mp.pi * k * (j + mp.mpf(1) / 2) / n) for j in range(n))
        out.append(s * (1 if k == 0 else 2) / n)
    # truncate the tail below TOL (relative to the leading magnitude 1)
    last = max(k for k in range(n) if abs(out[k]) > TOL)
    return out[: last + 1]


def x_small(t):
    return mp.sqrt(32 * (t + 1))


def j0_small(t):
    return mp.besselj(0, x_small(t))


def j1_small(t):
    x = x_small(t)
    return mp.besselj(1, x) / x if x != 0 else mp.mpf(1) / 2


def r0_small(t):
    x = x_small(t)
    if x == 0:
        return 2 / mp.pi * mp.euler
    return mp.bessely(0, x) - 2 / mp.pi * mp.log(x / 2) * mp.besselj(0, x)


def r1_small(t):
    x = x_small(t)
    if x == 0:
        x = mp.mpf("1e-30")
    return (mp.bessely(1, x) - 2 / mp.pi * mp.log(x / 2) * mp.besselj(1, x) + 2 / (mp.pi * x)) / x


def pq_large(n, t):
    z = (t + 1) / 2
    if z == 0:
        z = mp.mpf("1e-40")
    x = 8 / mp.sqrt(z)
    h = mp.hankel1(n, x) * mp.sqrt(mp.pi * x / 2) * mp.exp(-1j * (x - (2 * n + 1) * mp.pi / 4))
    return h.real, h.imag * x / 8


def emit(name, coeffs):
    lines = [f"static constexpr int {name}_N = {len(coeffs)};",
             f"IFGF_BESSEL_TABLE double {name}[{len(coeffs)}] = {{"]
    for c in coeffs:
        lines.append(f"    {float(c)!r},")
    lines.append("};")
    return "\n".join(lines)


def emit_interleaved(name, cols):
    n = max(len(c) for c in cols)
    lines = [f"static constexpr int {name}_N = {n};",
             f"IFGF_BESSEL_TABLE double {name}[{n}][{len(cols)}] = {{"]
    for k in range(n):
        vals = [repr(float(c[k])) if k < len(c) else "0.0" for c in cols]
        lines.append("    {" + ", ".join(vals) + "},")
    lines.append("};")
    return "\n".join(lines)


def main():
    tables = [
        ("kJ0Small", cheb_coeffs(j0_small)),
        ("kJ1Small", cheb_coeffs(j1_small)),
        ("kR0Small", cheb_coeffs(r0_small)),
        ("kR1Small", cheb_coeffs(r1_small)),
        ("kP0Large", cheb_coeffs(lambda t: pq_large(0, t)[0])),
        ("kQ0Large", cheb_coeffs(lambda t: pq_large(0, t)[1])),
        ("kP1Large", cheb_coeffs(lambda t: pq_large(1, t)[0])),
        ("kQ1Large", cheb_coeffs(lambda t: pq_large(1, t)[1])),
    ]
    out = sys.stdout
    out.write("// Generated by gen_bessel_coeffs.py (mpmath, 40 digits). Do not edit.\n")
    out.write("// Chebyshev coefficients of the FP64 J0/Y0/J1/Y1 evaluator; the functions\n")
    out.write("// they represent are listed in the generator's docstring.\n")
    out.write("#pragma once\n#ifndef IFGF_BESSEL_TABLE\n#define IFGF_BESSEL_TABLE static constexpr\n#endif\n\n")
    for name, coeffs in tables:
        out.write(emit(name, coeffs) + "\n\n")
    t = dict(tables)
    # interleaved tables for the multi-chain Clenshaw evaluation:
    # small x: (J0, R0) and (J0, J1/x, R0, R1/x); large x: (P0, Qt0) and (P0, Qt0, P1, Qt1)
    out.write(emit_interleaved("kSmall2", [t["kJ0Small"], t["kR0Small"]]) + "\n\n")
    out.write(emit_interleaved("kSmall4", [t["kJ0Small"], t["kJ1Small"], t["kR0Small"], t["kR1Small"]]) + "\n\n")
    out.write(emit_interleaved("kLarge2", [t["kP0Large"], t["kQ0Large"]]) + "\n\n")
    out.write(emit_interleaved("kLarge4", [t["kP0Large"], t["kQ0Large"], t["kP1Large"], t["kQ1Large"]]) + "\n\n")


if __name__ == "__main__":
    main()
