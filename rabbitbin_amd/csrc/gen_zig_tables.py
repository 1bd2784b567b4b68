#!/usr/bin/env python3
"""Generate the 256-layer ziggurat tables for the Exp(1) distribution.

Construction (Marsaglia & Tsang, "The Ziggurat Method for Generating Random
Variables", J. Stat. Softw. 5(8), 2000): layers of equal area v under
f(x) = exp(-x).  With r the right edge of the base layer,
    v      = exp(-r) * (r + 1)            (base strip + tail area)
    x[0]   = v / f(r) = r + 1             (pseudo-width of the base strip)
    x[1]   = r,  y[1] = exp(-r)
    y[i+1] = y[i] + v / x[i],  x[i+1] = -ln(y[i+1])
    y[0]   = 0, and r is the root of x[256] = 0 (y[256] = 1).
The sampler in ProbMinHash's RNG stream (reference src/probmh.cpp:319-338)
uses exactly this layout.  Values are solved with 80-digit decimal
arithmetic and written as C99 hex-float literals (correctly rounded).

Usage: python3 gen_zig_tables.py > zig_tables.h
"""
from decimal import Decimal, getcontext
import sys

getcontext().prec = 90
N = 256


def build(r):
    r = Decimal(r)
    v = (-r).exp() * (r + 1)
    x = [Decimal(0)] * (N + 1)
    y = [Decimal(0)] * (N + 1)
    x[0] = r + 1
    x[1] = r
    y[1] = (-r).exp()
    for i in range(1, N):
        y[i + 1] = y[i] + v / x[i]
        if y[i + 1] >= 1:
            return x, y, i + 1  # ran past the top: r too small
        x[i + 1] = -(y[i + 1].ln())
    return x, y, None


def solve():
    lo, hi = Decimal("7.6"), Decimal("7.8")
    for _ in range(300):
        mid = (lo + hi) / 2
        x, y, over = build(mid)
        if over is not None:
            lo = mid          # layers reached y=1 early -> r too small
        else:
            if y[N] + (Decimal(0) if x is None else 0) < 1 and x[N] > 0:
                hi = mid
            else:
                lo = mid
    return (lo + hi) / 2


def main():
    r = solve()
    x, y, _ = build(r)
    x[N] = Decimal(0)
    y[N] = Decimal(1)
    y[0] = Decimal(0)
    out = sys.stdout
    out.write("// Generated by gen_zig_tables.py (Marsaglia-Tsang 256-layer Exp(1) ziggurat).\n")
    out.write("// r = %s\n" % format(r, ".30f"))
    out.write("#pragma once\n")
    out.write("static const double kZigX[257] = {\n")
    for i in range(N + 1):
        out.write("  %s,\n" % float(x[i]).hex())
    out.write("};\n")
    out.write("static const double kZigY[257] = {\n")
    for i in range(N + 1):
        out.write("  %s,\n" % float(y[i]).hex())
    out.write("};\n")


if __name__ == "__main__":
    main()
