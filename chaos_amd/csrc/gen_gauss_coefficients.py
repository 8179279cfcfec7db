"""Print the s = 3 Gauss-Legendre coefficients of cl_gauss.hpp: nodes c (roots of the shifted Legendre
polynomial, ascending), weights b and A from the collocation conditions sum_j A_ij c_j^(k-1) = c_i^k / k,
solved in 320-bit arithmetic, each split into a double hi part and a double remainder lo (hi + lo is the
value to about twice double precision), as the reference builds its Gauss tables."""
import mpmath as mp

mp.mp.prec = 320
s = 3
r15 = mp.sqrt(15)
c = [mp.mpf(1) / 2 - r15 / 10, mp.mpf(1) / 2, mp.mpf(1) / 2 + r15 / 10]
V = mp.matrix([[c[j] ** k for j in range(s)] for k in range(s)])
b = mp.lu_solve(V, mp.matrix([mp.mpf(1) / (k + 1) for k in range(s)]))
A = mp.matrix(s, s)
for i in range(s):
    row = mp.lu_solve(V, mp.matrix([c[i] ** (k + 1) / (k + 1) for k in range(s)]))
    for j in range(s):
        A[i, j] = row[j]


def split(x):
    hi = float(x)
    lo = float(x - mp.mpf(hi))
    return hi, lo


if __name__ == "__main__":
    print("A hi/lo (row-major):")
    for i in range(s):
        print("  ", ", ".join(f"{split(A[i, j])[0].hex()}" for j in range(s)))
        print("  ", ", ".join(f"{split(A[i, j])[1].hex()}" for j in range(s)))
    print("b hi/lo:", [split(x)[0].hex() for x in b], [split(x)[1].hex() for x in b])
    print("c:", [float(x) for x in c])
    print("check b = 5/18, 8/18, 5/18:", [float(x * 18) for x in b])
