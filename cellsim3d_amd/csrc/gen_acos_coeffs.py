"""Coefficients of cs3d_acos.h.

The degree-12 Chebyshev interpolant of f(z) = (asin(sqrt z)/sqrt z - 1)/z on
[0, 1/4], computed in 60-digit decimal arithmetic from the series of asin,
converted to monomial coefficients in z and rounded to double.  Prints the
coefficients (constant term first) and the largest relative error of the
rounded polynomial against the series on a grid of 1001 points.

    python gen_acos_coeffs.py [degree]
"""
from decimal import Decimal as D, getcontext
import sys
getcontext().prec = 60
def f_exact(z):
    z = D(z); tot = D(0); t = D(1); zk = D(1)
    for k in range(1, 400):
        t = t * D(2 * k - 1) / D(2 * k)
        term = t / D(2 * k + 1) * zk
        tot += term; zk *= z
        if abs(term) < D(10) ** -58: break
    return tot
def pi_dec():
    # Machin: pi = 16 atan(1/5) - 4 atan(1/239)
    def atan_inv(n):
        x = D(1) / n; x2 = x * x; s = x; t = x; k = 1
        while True:
            t = -t * x2; k += 2; term = t / k
            if abs(term) < D(10) ** -59: break
            s += term
        return s
    return 16 * atan_inv(5) - 4 * atan_inv(239)
PI = pi_dec()
def cos_dec(x):
    s = D(1); t = D(1); k = 0
    while True:
        k += 2; t = -t * x * x / (k * (k - 1))
        if abs(t) < D(10) ** -59: break
        s += t
    return s
deg = int(sys.argv[1]) if len(sys.argv) > 1 else 12
n = deg + 1
# nodes in t in [-1, 1] -> z = (t + 1) / 8
nodes = [cos_dec(PI * (2 * j + 1) / (2 * n)) for j in range(n)]
fz = [f_exact((t + 1) / 8) for t in nodes]
# Chebyshev coefficients a_k = 2/n sum f(t_j) T_k(t_j)
def T(k, t):
    a, b = D(1), t
    if k == 0: return a
    for _ in range(k - 1): a, b = b, 2 * t * b - a
    return b
a = [sum(fz[j] * T(k, nodes[j]) for j in range(n)) * 2 / n for k in range(n)]
a[0] /= 2
# expand sum a_k T_k(t) with t = 8 z - 1 into monomials of z
def polymul(p, q):
    r = [D(0)] * (len(p) + len(q) - 1)
    for i, x in enumerate(p):
        for j, y in enumerate(q): r[i + j] += x * y
    return r
Tz = [[D(1)], [D(-1), D(8)]]
for k in range(2, n):
    t2 = polymul([D(-2), D(16)], Tz[-1])
    prev = Tz[-2] + [D(0)] * (len(t2) - len(Tz[-2]))
    Tz.append([x - y for x, y in zip(t2, prev)])
c = [D(0)] * n
for k in range(n):
    for i, x in enumerate(Tz[k]): c[i] += a[k] * x
for i, x in enumerate(c):
    print(f"    {float(x).hex()},  // {x:.20e}")
# check the interpolant on a dense grid (decimal, coefficients as rounded doubles)
cd = [D(float(x)) for x in c]
worst = D(0)
for j in range(1001):
    z = D(j) / 4000
    p = D(0)
    for x in reversed(cd): p = p * z + x
    e = abs(p / f_exact(z) - 1)
    worst = max(worst, e)
import math
print("max rel err", float(worst), math.log2(float(worst)))
