"""Generate erfcx_table.hpp: piecewise polynomials for erfcx(x) = exp(x^2) erfc(x) on [0, 8].

The range is cut into 32 intervals of width 1/4. On each, erfcx (an entire function) is interpolated at 64
Chebyshev nodes in 50-digit arithmetic (mpmath), the Chebyshev series is truncated at degree 11 and converted
exactly to monomials in t = 8 (x - c) (c the interval's center, t in [-1, 1]), and the coefficients are rounded
to double. Evaluated by Horner in double the pieces are within 2.5e-16 of erfcx (tests/test_erfcx_table.py).

    python ryoanji_amd/csrc/gen_erfcx_table.py > ryoanji_amd/csrc/erfcx_table.hpp
"""

import mpmath as mp

mp.mp.dps = 50
WIDTH = mp.mpf(1) / 4
INTERVALS = 32
DEGREE = 11
NODES = 64


def erfcx(x):
    return mp.exp(x * x) * mp.erfc(x)


def piece(j):
    c = (j + mp.mpf(1) / 2) * WIDTH
    h = WIDTH / 2
    th = [mp.pi * (k + mp.mpf(1) / 2) / NODES for k in range(NODES)]
    fv = [erfcx(c + h * mp.cos(a)) for a in th]
    cheb = [2 * mp.fsum(fv[k] * mp.cos(i * th[k]) for k in range(NODES)) / NODES for i in range(DEGREE + 1)]
    cheb[0] /= 2
    mono = [mp.mpf(0)] * (DEGREE + 1)
    t_prev, t_cur = [mp.mpf(1)], [mp.mpf(0), mp.mpf(1)]
    mono[0] += cheb[0]
    for i in range(1, DEGREE + 1):
        for k, v in enumerate(t_cur):
            mono[k] += cheb[i] * v
        t_next = [mp.mpf(0)] * (len(t_cur) + 1)
        for k, v in enumerate(t_cur):
            t_next[k + 1] += 2 * v
        for k, v in enumerate(t_prev):
            t_next[k] -= v
        t_prev, t_cur = t_cur, t_next
    return [float(v) for v in mono]


def main():
    print("// Generated by gen_erfcx_table.py: erfcx(x) = exp(x^2) erfc(x) on [0, 8] as 32 polynomials of degree 11")
    print("// in t = 8 (x - (j + 1/2) / 4) on interval j = [j/4, (j+1)/4] (monomial coefficients, constant first).")
    print("#pragma once")
    print()
    print(f"constexpr int kErfcxIntervals = {INTERVALS};")
    print(f"constexpr int kErfcxTerms     = {DEGREE + 1};")
    print()
    print("__constant__ double kErfcx[kErfcxIntervals][kErfcxTerms] = {")
    for j in range(INTERVALS):
        c = ", ".join(repr(v) for v in piece(j))
        print(f"    {{{c}}},")
    print("};")


if __name__ == "__main__":
    main()
