"""Prime chains and parameter presets for the RNS-CKKS hot path.

Mirrors RnsPrime / RnsCkksParams (src/rns/basis.rs:7-22, src/rns/context.rs:15-21,
src/rns/context/params.rs:9-46).  Prime search follows find_primes.py:3-55 (q = 1 mod 2N,
smallest primitive root), with a deterministic Miller-Rabin test valid for all 64-bit q.
"""
from dataclasses import dataclass, field
from typing import List

_MR_BASES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)


def is_prime(n: int) -> bool:
    if n < 2:
        return False
    for p in _MR_BASES:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in _MR_BASES:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def _factor(n: int) -> List[int]:
    fs, p = [], 2
    while p * p <= n:
        if n % p == 0:
            fs.append(p)
            while n % p == 0:
                n //= p
        p += 1 if p == 2 else 2
        if p > 1 << 20:  # q-1 = 2^k * m with m small for NTT primes; fall back to trial by odd p
            break
    if n > 1:
        if is_prime(n):
            fs.append(n)
        else:
            fs.extend(_factor_rho(n))
    return sorted(set(fs))


def _factor_rho(n: int) -> List[int]:
    import math
    import random
    if is_prime(n):
        return [n]
    rnd = random.Random(n)
    while True:
        c = rnd.randrange(1, n)
        f = lambda x: (x * x + c) % n
        x = y = rnd.randrange(2, n)
        d = 1
        while d == 1:
            x = f(x)
            y = f(f(y))
            d = math.gcd(abs(x - y), n)
        if d != n:
            return sorted(set(_factor_rho(d) + _factor_rho(n // d)))


def smallest_primitive_root(q: int) -> int:
    """find_primes.py:25-33 convention: the smallest generator of (Z/qZ)^*."""
    fs = _factor(q - 1)
    g = 2
    while True:
        if all(pow(g, (q - 1) // f, q) != 1 for f in fs):
            return g
        g += 1


def primes_below(bits: int, n: int, count: int, skip: int = 0):
    """The largest primes q < 2^bits with q = 1 mod 2N, in descending order."""
    step = 2 * n
    q = ((1 << bits) - 1) // step * step + 1
    out = []
    while len(out) < count + skip:
        if q < (1 << bits) and is_prime(q):
            out.append(q)
        q -= step
    return out[skip:]


def primes_above(bits: int, n: int, count: int, skip: int = 0):
    """The smallest primes q > 2^bits with q = 1 mod 2N, in ascending order."""
    step = 2 * n
    q = ((1 << bits) // step + 1) * step + 1
    out = []
    while len(out) < count + skip:
        if is_prime(q):
            out.append(q)
        q += step
    return out[skip:]


@dataclass
class RnsCkksParams:
    """RnsCkksParams (src/rns/context.rs:15-21): primes and roots as parallel lists."""
    poly_degree: int
    primes: List[int]
    roots: List[int]
    aux_primes: List[int] = field(default_factory=list)
    aux_roots: List[int] = field(default_factory=list)
    scale_bits: int = 40

    @classmethod
    def with_primes(cls, n, primes, aux=(), scale_bits=40):
        return cls(n, list(primes), [smallest_primitive_root(q) for q in primes], list(aux),
                   [smallest_primitive_root(q) for q in aux], scale_bits)


def config_plain_cpu():
    """N=2^12, the 3 largest primes < 2^50 with q = 1 mod 2N, scale 2^40, Q-only keys."""
    return RnsCkksParams.with_primes(1 << 12, primes_below(50, 1 << 12, 3), (), 40)


def config_per_op_gpu(n=1 << 16, limbs=24, aux=2):
    """N=2^16, the 24 largest primes < 2^50 with q = 1 mod 2N, plus the next 2 as aux primes P."""
    ps = primes_below(50, n, limbs + aux)
    return RnsCkksParams.with_primes(n, ps[:limbs], ps[limbs:], 40)


def config_batched(n=1 << 15, limbs=20, aux=2):
    """Primes just above 2^50 (exact rescale bookkeeping), aux primes taken after them."""
    ps = primes_above(50, n, limbs + aux)
    return RnsCkksParams.with_primes(n, ps[:limbs], ps[limbs:], 40)


def realistic_8_level():
    """The reference preset realistic_8_level (src/rns/context/params.rs:29-46)."""
    return RnsCkksParams(
        1 << 15,
        [1125899908022273, 1125899908612097, 1125899909398529, 1125899910316033, 1125899911168001,
         1125899911364609, 1125899913527297, 1125899914510337],
        [3, 3, 3, 5, 3, 3, 3, 3],
        [562949954142209, 562949955125249], [3, 3], 45)
