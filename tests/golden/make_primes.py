"""Writes primes.json: the prime chains of the workloads, found independently with sympy.

For each ring degree N: the largest primes below 2^50 and the smallest primes above 2^50
with q = 1 mod 2N, each with its smallest primitive root (find_primes.py:25-33 convention).
Run: python tests/golden/make_primes.py
"""
import json
import os

from sympy import isprime, primitive_root


def below(n, count):
    step, q, out = 2 * n, ((1 << 50) - 1) // (2 * n) * (2 * n) + 1, []
    while len(out) < count:
        if q < 1 << 50 and isprime(q):
            out.append([q, int(primitive_root(q))])
        q -= step
    return out


def above(n, count):
    step, q, out = 2 * n, ((1 << 50) // (2 * n) + 1) * (2 * n) + 1, []
    while len(out) < count:
        if isprime(q):
            out.append([q, int(primitive_root(q))])
        q += step
    return out


data = {
    "below_2^50": {str(n): below(n, 28) for n in (1 << 12, 1 << 14, 1 << 15, 1 << 16)},
    "above_2^50": {str(n): above(n, 28) for n in (1 << 12, 1 << 14, 1 << 15, 1 << 16)},
}
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "primes.json"), "w") as f:
    json.dump(data, f, indent=1)
