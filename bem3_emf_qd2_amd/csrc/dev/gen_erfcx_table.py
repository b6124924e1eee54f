"""Write bq_erfcx_table.inc: the coefficients of the piecewise-polynomial erfcx of
bq_faddeeva.hpp (erfcx_tab).

erfcx(x) = exp(x^2) erfc(x) for 0 <= x <= 50 is taken in the variable t = 4 / (4 + x) in
[4/54, 1], split into 64 equal intervals; on each, a degree-7 polynomial in the local
coordinate u in [-1, 1] interpolates 40-digit values at the 8 Chebyshev points (monomial
coefficients, lowest degree first).  Piecewise fits in 4 / (4 + x) are the approach of
S. G. Johnson's Faddeeva package (its erfcx uses 100 Chebyshev fits in the same variable);
this table is fitted here from mpmath values.  Run from this directory:

    python gen_erfcx_table.py > bq_erfcx_table.inc
"""
import mpmath as mp

N, D = 64, 7
mp.mp.dps = 40
T0 = mp.mpf(4) / 54


def erfcx(x):
    return mp.exp(x * x) * mp.erfc(x)


def main():
    nodes = [mp.cos(mp.pi * (j + mp.mpf(1) / 2) / (D + 1)) for j in range(D + 1)]
    V = mp.matrix([[u ** p for p in range(D + 1)] for u in nodes])
    print("// generated by gen_erfcx_table.py: 64 intervals of t = 4/(4+x) in [4/54, 1], degree 7 in u")
    for k in range(N):
        a = T0 + (1 - T0) * k / N
        b = T0 + (1 - T0) * (k + 1) / N
        vals = [erfcx(4 / ((a + b) / 2 + (b - a) / 2 * u) - 4) for u in nodes]
        c = mp.lu_solve(V, mp.matrix(vals))
        print("    " + ", ".join(mp.nstr(c[p], 17, min_fixed=1, max_fixed=0) for p in range(D + 1)) + ",")


if __name__ == "__main__":
    main()
