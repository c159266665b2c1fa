"""Operands on which correctly rounded division and square root are hardest.

A p-bit binary format (p = 53 for binary64, 24 for binary32) rounds a
quotient or root to the nearest p-bit significand; the hard operands are
those whose exact result lies a tiny distance from a midpoint between two
neighbours, where a refinement that is off in its last bits rounds the wrong
way.  Both constructions below are exact integer arithmetic:

* Division: for an odd p-bit B and a small odd r, A = r * 2^-p (mod B) + B
  gives 2^p A - M B = r with M odd, i.e. A / B = M / 2^p + r / (2^p B): within
  |r| / (2^p B) of the midpoint M / 2^p of two neighbours in [1, 2).
  Correct rounding goes up for r > 0 and down for r < 0.
* Square root: for t = 1 (mod 8) the congruence M^2 = t (mod 2^k) has four
  odd roots (Hensel lifting); with k = p + 1 + b and N = (M^2 - t) / 2^k the
  p-bit radicand X = N / 2^(p-1-b) (X in [1, 2) for b = 0, [2, 4) for b = 1)
  satisfies (M / 2^p)^2 = X + t / 2^(2p), so sqrt(X) lies within about
  t / 2^(p+2) ulp of the midpoint M / 2^p, below it for t > 0 and above it
  for t < 0.

Scaling an operand by a power of two (of four for a radicand) keeps a case
exactly as hard, so the cases are spread over the solver's magnitudes.
"""
import random

import numpy as np


def div_near_midpoint(n, p, seed, max_r=15):
    """n tuples (A, B, M, r): p-bit integers with 2^p A - M B = r, r odd."""
    rng = random.Random(seed)
    lo, hi = 1 << (p - 1), 1 << p
    out = []
    while len(out) < n:
        b = rng.randrange(lo, hi) | 1
        r = rng.randrange(-max_r, max_r + 1, 2)
        a = (r * pow(2, -p, b)) % b + b
        if a >= hi:
            continue
        m, rem = divmod((a << p) - r, b)
        assert rem == 0 and m & 1
        out.append((a, b, m, r))
    return out


def _sqrt_mod_pow2(t, k):
    """An odd x with x^2 = t (mod 2^k), for t = 1 (mod 8) and k >= 3."""
    x = 1
    for j in range(3, k):
        if (x * x - t) % (1 << (j + 1)):
            x += 1 << (j - 1)
    return x


def sqrt_near_midpoint(n, p, seed, max_t=1 << 20):
    """n tuples (N, b, M, t): the radicand N / 2^(p-1-b) has its root within
    about t / 2^(p+2) ulp of the midpoint M / 2^p."""
    rng = random.Random(seed)
    lo, hi = 1 << (p - 1), 1 << p
    out = []
    while len(out) < n:
        t = 8 * rng.randrange(-max_t // 8, max_t // 8) + 1
        b = rng.randrange(2)
        k = p + 1 + b
        mod = 1 << k
        x = _sqrt_mod_pow2(t % mod, k)
        for m in {x, mod - x, (x + (mod >> 1)) % mod, (mod - x + (mod >> 1)) % mod}:
            if not (1 << p) <= m < (1 << (p + 1)):
                continue
            big, rem = divmod(m * m - t, mod)
            if rem == 0 and lo <= big < hi:
                out.append((big, b, m, t))
    return out[:n]


def scaled_div_operands(cases, p, seed, lo_exp=-70, hi_exp=28):
    """(a, b) float64 arrays of the division cases, each operand scaled by its own
    power of two into [2^lo_exp, 2^hi_exp) and given a random sign (exact)."""
    rng = np.random.default_rng(seed)
    a = np.array([c[0] for c in cases], dtype=np.float64)
    b = np.array([c[1] for c in cases], dtype=np.float64)
    ea = rng.integers(lo_exp, hi_exp, len(cases)) - (p - 1)
    eb = rng.integers(lo_exp, hi_exp, len(cases)) - (p - 1)
    sa = rng.choice([-1.0, 1.0], len(cases))
    sb = rng.choice([-1.0, 1.0], len(cases))
    return sa * np.ldexp(a, ea), sb * np.ldexp(b, eb)


def scaled_radicands(cases, p, seed, lo_exp=-35, hi_exp=14):
    """float64 radicands of the square-root cases, scaled by 4^s, s in [lo_exp, hi_exp)."""
    rng = np.random.default_rng(seed)
    x = np.array([c[0] for c in cases], dtype=np.float64)
    shift = np.array([-(p - 1 - c[1]) for c in cases])
    s = rng.integers(lo_exp, hi_exp, len(cases))
    return np.ldexp(x, shift + 2 * s)


def div_by_root_operands(x, p, seed, max_r=15):
    """Numerators a that put a / RN(sqrt(x)) next to a midpoint, for the radicands x
    whose rounded root has an odd p-bit significand (the others are dropped).
    Returns (a, x) float64 arrays; the root is rounded in the p-bit format."""
    rng = random.Random(seed)
    dt = np.float64 if p == 53 else np.float32
    x = x.astype(dt)
    s = np.sqrt(x)
    mant, ex = np.frexp(s)
    sig = np.ldexp(mant.astype(np.float64), p).astype(np.int64)
    keep = (sig & 1) == 1
    a_out, x_out = [], []
    hi = 1 << p
    for big_s, e, xv in zip(sig[keep].tolist(), ex[keep].tolist(), x[keep].tolist()):
        r = rng.randrange(-max_r, max_r + 1, 2)
        a = (r * pow(2, -p, big_s)) % big_s + big_s
        if a >= hi:
            continue
        # a / 2^p * 2^e over s = big_s / 2^p * 2^e: the quotient A / S, scaled
        scale = rng.randrange(-40, 20)
        a_out.append(float(np.ldexp(float(a), e - p + scale)) * rng.choice([-1.0, 1.0]))
        x_out.append(float(xv))
    return np.array(a_out, dtype=np.float64), np.array(x_out, dtype=np.float64)


def quotients_next_to_midpoint(tol, n, seed, max_r=1 << 20, p=53):
    """(a, d, above): float64 pairs, p-bit numbers, whose exact quotient a / d
    lies next to the midpoint mu = tol + h between the p-bit tol and the next
    one up (h half an ulp of tol), above it where `above` is true.  With
    mu = M / 2^p * 2^E (M odd, p + 1 bits, E the exponent of tol) and an odd r,
    B = -r / M (mod 2^p) and A = (M B + r) / 2^p give A / B = M / 2^p + r / (2^p B);
    then a = A 2^(E+s), d = B 2^s for a random scale s keep that distance."""
    rng = random.Random(seed)
    m_tol, e_tol = np.frexp(tol)
    tol_int = int(np.ldexp(m_tol, p))           # tol = tol_int * 2^(e_tol - p)
    big_m = 2 * tol_int + 1                      # mu = big_m * 2^(e_tol - p - 1)
    exp = int(e_tol) - 1                         # mu = big_m / 2^p * 2^exp
    mod = 1 << p
    inv_m = pow(big_m, -1, mod)
    a_out, d_out, above = [], [], []
    while len(a_out) < n:
        r = 2 * rng.randrange(-max_r // 2, max_r // 2) + 1
        b = (-r * inv_m) % mod
        if b == 0:
            continue
        a, rem = divmod(big_m * b + r, mod)
        if rem or a >= mod or a == 0:
            continue
        s = rng.randrange(-20, 20) - b.bit_length()
        a_out.append(float(np.ldexp(float(a), exp + s)))
        d_out.append(float(np.ldexp(float(b), s)))
        above.append(r > 0)
    return np.array(a_out), np.array(d_out), np.array(above)
