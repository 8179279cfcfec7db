"""Minimax polynomial coefficients for the kernels' portable sin, cos and log (cl_math.h).

Each polynomial is fitted by the Remez exchange algorithm (E. Remez, 1934) for the smallest maximum
relative error on its reduced interval, in 200-bit arithmetic (mpmath), and then rounded to double:

  sin   sin(r) = r + r^3 S(z),                z = r^2 in [0, 1],          S of degree 6 (S(0) ~ -1/6)
  cos   cos(r) = 1 - z/2 + z^2 C(z),          z = r^2 in [0, 1],          C of degree 6 (C(0) ~ 1/24)
  log   log((1+s)/(1-s)) = 2s + s^3 P(z),     z = s^2 in [0, (3 - 2 sqrt 2)^2], P of degree 6 (P(0) ~ 2/3)

The sin/cos interval is |r| <= 1, not pi/4: near |x| = 2^50 the quadrant chosen from the rounded product
x * 2/pi can be off by one, which leaves |r| up to about 0.9.

Run it to print the coefficients as C hex-float literals and the fitted and rounded relative errors.
"""
import mpmath as mp

mp.mp.prec = 200


def remez(f, a, b, deg, iters=40):
    """Coefficients c[0..deg] of the polynomial p minimising max |p(z) / f(z) - 1| on [a, b]."""
    n = deg + 2
    xs = [(a + b) / 2 - (b - a) / 2 * mp.cos(mp.pi * i / (n - 1)) for i in range(n)]
    coef = None
    for _ in range(iters):
        # p(x_i) - f(x_i) = (-1)^i E f(x_i)
        A = mp.matrix(n, n)
        rhs = mp.matrix(n, 1)
        for i, x in enumerate(xs):
            fx = f(x)
            for j in range(deg + 1):
                A[i, j] = x ** j
            A[i, deg + 1] = (-1) ** i * fx
            rhs[i] = fx
        sol = mp.lu_solve(A, rhs)
        coef = [sol[j] for j in range(deg + 1)]
        err = lambda x: mp.polyval(coef[::-1], x) / f(x) - 1
        # new reference: the extrema of err between sign changes on a dense grid, refined by golden section
        grid = [a + (b - a) * mp.mpf(k) / 4000 for k in range(4001)]
        vals = [err(x) for x in grid]
        ext = []
        k = 0
        while k < len(grid):
            s = mp.sign(vals[k])
            best = k
            while k < len(grid) and mp.sign(vals[k]) == s:
                if abs(vals[k]) > abs(vals[best]):
                    best = k
                k += 1
            ext.append(grid[best])
        if len(ext) < n:
            break
        # keep n alternating extrema with the largest errors
        while len(ext) > n:
            e0, e1 = abs(err(ext[0])), abs(err(ext[-1]))
            ext = ext[1:] if e0 < e1 else ext[:-1]
        xs = ext
    emax = max(abs(mp.polyval(coef[::-1], x) / f(x) - 1) for x in
               [a + (b - a) * mp.mpf(k) / 20000 for k in range(20001)])
    return coef, emax


def rounded_error(f, a, b, coef):
    c = [mp.mpf(float(v)) for v in coef]
    return max(abs(mp.polyval(c[::-1], x) / f(x) - 1) for x in [a + (b - a) * mp.mpf(k) / 20000 for k in range(20001)])


def hexf(v):
    return float(v).hex().replace("0x1.", "0x1.").rstrip()


def series_sin(z):
    if z == 0:
        return mp.mpf(-1) / 6
    r = mp.sqrt(z)
    return (mp.sin(r) - r) / (r * z)


def series_cos(z):
    if z == 0:
        return mp.mpf(1) / 24
    return (mp.cos(mp.sqrt(z)) - 1 + z / 2) / (z * z)


def series_log(z):
    # (log((1+s)/(1-s)) - 2s) / s^3 = sum_{j>=1} 2 z^(j-1) / (2j + 1)
    if z == 0:
        return mp.mpf(2) / 3
    s = mp.sqrt(z)
    return (mp.log((1 + s) / (1 - s)) - 2 * s) / (s * z)


if __name__ == "__main__":
    zt = mp.mpf("1.0001")
    zl = (3 - 2 * mp.sqrt(2)) ** 2 * mp.mpf("1.0001")
    for name, f, b, deg in (("S (sin)", series_sin, zt, 6), ("C (cos)", series_cos, zt, 6),
                            ("P (log)", series_log, zl, 6)):
        coef, e = remez(f, mp.mpf(0), b, deg)
        er = rounded_error(f, mp.mpf(0), b, coef)
        print(f"{name}: fitted max rel err {float(e):.3e} (2^{float(mp.log(e, 2)):.1f}), "
              f"rounded {float(er):.3e} (2^{float(mp.log(er, 2)):.1f})")
        print("  " + ", ".join(hexf(c) for c in coef))
