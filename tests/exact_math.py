"""Reference values of sin, cos, asin and log to 60 significant digits (Python's
decimal module), rounded once to the nearest double: the correctly rounded
results that the math library of the tracking maps and the C library are
measured against.  Test infrastructure only."""
import math
from decimal import Decimal, getcontext, localcontext

PREC = 60
# pi to 70 digits (a constant of mathematics, not of any library)
PI = Decimal("3.141592653589793238462643383279502884197169399375105820974944592307816")


def _sin_cos(x):
    """(sin x, cos x) of a Decimal x"""
    with localcontext() as ctx:
        ctx.prec = PREC + 10
        two_pi = 2 * PI
        k = (x / two_pi).to_integral_value()
        r = x - k * two_pi
        # Taylor series; |r| <= pi
        s, c = Decimal(0), Decimal(0)
        term = r
        n = 1
        while True:
            s += term
            term *= -r * r / ((n + 1) * (n + 2))
            n += 2
            if abs(term) < Decimal(10) ** -(PREC + 8):
                break
        term = Decimal(1)
        n = 0
        while True:
            c += term
            term *= -r * r / ((n + 1) * (n + 2))
            n += 2
            if abs(term) < Decimal(10) ** -(PREC + 8):
                break
        return +s, +c


def sin(x):
    return float(_sin_cos(Decimal(x))[0])


def cos(x):
    return float(_sin_cos(Decimal(x))[1])


def asin(x):
    """Newton's iteration on sin(y) = x from the double estimate"""
    if x == 0:
        return x
    xd = Decimal(x)
    with localcontext() as ctx:
        ctx.prec = PREC + 10
        y = Decimal(math.asin(x))
        for _ in range(4):
            s, c = _sin_cos(y)
            y -= (s - xd) / c
        return float(+y)


def log(x):
    with localcontext() as ctx:
        ctx.prec = PREC + 10
        return float(Decimal(x).ln())


FUNCS = {"sin": sin, "cos": cos, "asin": asin, "log": log}
