"""Pins the oracle's Zq layer (math/src/zq/mod.rs, primes.rs) against Python
integer arithmetic and the reference's known-answer tests."""
import random

import numpy as np
import pytest

P62 = 4611686018326724609


def _rand_moduli():
    return [P62, 4611686018309947393, 1153, 2, 4, 1152921504606748673, (1 << 61) - 1, 65537, 12289]


def test_nfl_62bit_primes_kat(oracle):
    # math/src/zq/primes.rs:67-102
    expected = [4611686018326724609, 4611686018309947393, 4611686018282684417, 4611686018257518593,
                4611686018232352769, 4611686018171535361, 4611686018106523649, 4611686018058289153,
                4611686018051997697, 4611686017974403073, 4611686017812922369, 4611686017781465089,
                4611686017773076481, 4611686017678704641, 4611686017666121729, 4611686017647247361,
                4611686017590624257, 4611686017554972673, 4611686017529806849, 4611686017517223937]
    got, ub = [], (2**64 - 1) >> 2
    while len(got) != 20:
        p = oracle.generate_prime(62, 2 * 1048576, ub)
        assert p is not None
        ub = p
        got.append(p)
    assert got == expected


def test_generate_prime_none_cases(oracle):
    # primes.rs:112-122
    assert oracle.generate_prime(10, 2048, 1 << 10) is None
    assert oracle.generate_prime(11, 16, 1033) is None


def test_generate_moduli_kat(oracle):
    # fhe/src/bfv/parameters.rs:740-773
    assert oracle.generate_moduli([62, 62, 62, 61, 60, 11], 16) == [
        4611686018427387617, 4611686018427387329, 4611686018427387073, 2305843009213693921,
        1152921504606845473, 2017]


def test_config_moduli(oracle):
    # Moduli the benchmark configurations use (derived with generate_moduli).
    assert oracle.generate_moduli([62] * 3, 8192) == [4611686018427322369, 4611686018427289601,
                                                      4611686018426454017]
    assert oracle.generate_moduli([60] * 7, 16384) == [
        1152921504606748673, 1152921504606683137, 1152921504606584833, 1152921504605962241,
        1152921504604979201, 1152921504600260609, 1152921504599080961]


def _is_prime_py(n):
    if n < 2:
        return False
    for p in [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37]:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37]:
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


def test_is_prime_against_trial_division(oracle):
    for n in range(0, 3000):
        tp = n >= 2 and all(n % d for d in range(2, int(n ** 0.5) + 1))
        assert oracle.is_prime(n) == tp, n
    # Carmichael / strong pseudoprimes to small bases
    for n in [561, 1105, 3215031751, 2152302898747, 3474749660383, 341550071728321]:
        assert not oracle.is_prime(n)


def test_supports_opt_closed_form(oracle):
    rng = random.Random(1)
    vals = [P62, 4611686018427387617, 1153, 2, 65537, 1152921504606748673, (1 << 62) - 57]
    vals += [rng.randrange(2, 1 << 62) for _ in range(200)]
    for p in vals:
        s0 = 64 - p.bit_length()
        expect = s0 >= 1 and ((1 << (3 * s0)) + 1) << 64 < (1 << (3 * s0)) * ((1 << s0) + 1) * p
        assert oracle.supports_opt(p) == expect, p


@pytest.mark.parametrize("p", _rand_moduli())
def test_modulus_ops_vs_python(oracle, p):
    rng = random.Random(p)
    n = 400
    a = [rng.randrange(p) for _ in range(n)] + [0, p - 1]
    b = [rng.randrange(p) for _ in range(n)] + [p - 1, p - 1]
    A, B = np.array(a, np.uint64), np.array(b, np.uint64)
    assert [int(x) for x in oracle.modulus_op(p, "add", A, B)] == [(x + y) % p for x, y in zip(a, b)]
    assert [int(x) for x in oracle.modulus_op(p, "sub", A, B)] == [(x - y) % p for x, y in zip(a, b)]
    assert [int(x) for x in oracle.modulus_op(p, "mul", A, B)] == [(x * y) % p for x, y in zip(a, b)]
    assert [int(x) for x in oracle.modulus_op(p, "neg", A)] == [(-x) % p for x in a]
    assert [int(x) for x in oracle.modulus_op(p, "shoup", A)] == [(x << 64) // p for x in a]
    assert [int(x) for x in oracle.modulus_op(p, "mul_shoup", A, B)] == [(x * y) % p for x, y in zip(a, b)]
    lz = oracle.modulus_op(p, "lazy_mul_shoup", A, B)
    for x, y, r in zip(a, b, lz):
        assert int(r) < 2 * p and int(r) % p == x * y % p
    # arbitrary u64 inputs for the reductions
    w = [rng.randrange(1 << 64) for _ in range(n)] + [2**64 - 1, 0]
    W = np.array(w, np.uint64)
    assert [int(x) for x in oracle.modulus_op(p, "reduce", W)] == [x % p for x in w]
    for x, r in zip(w, oracle.modulus_op(p, "lazy_reduce", W)):
        assert int(r) < 2 * p and int(r) % p == x % p
    # u128 reduction (a = lo | hi << 64)
    hi = [rng.randrange(1 << 64) for _ in range(len(w))]
    H = np.array(hi, np.uint64)
    assert [int(x) for x in oracle.modulus_op(p, "reduce_u128", W, H)] == \
        [((h << 64) | x) % p for x, h in zip(w, hi)]
    assert [int(x) for x in oracle.modulus_op(p, "mul_shoup", W, B)] == [(x * y) % p for x, y in zip(w, b)]
    if oracle.supports_opt(p):
        assert [int(x) for x in oracle.modulus_op(p, "mul_opt", A, B)] == [(x * y) % p for x, y in zip(a, b)]
        for x, r in zip(w, oracle.modulus_op(p, "lazy_reduce_opt", W)):
            assert int(r) < 2 * p and int(r) % p == x % p
    e = [rng.randrange(p) for _ in range(50)]
    assert [int(x) for x in oracle.modulus_op(p, "pow", A[:50], np.array(e, np.uint64))] == \
        [pow(x, y, p) for x, y in zip(a[:50], e)]
    s = [rng.randrange(-(1 << 63), 1 << 63) for _ in range(n)]
    S = np.array(s, np.int64).view(np.uint64)
    assert [int(x) for x in oracle.modulus_op(p, "reduce_i64", S)] == [x % p for x in s]
    if _is_prime_py(p):
        nz = [x for x in a if x][:50]
        assert [int(x) for x in oracle.modulus_op(p, "inv", np.array(nz, np.uint64))] == [pow(x, -1, p) for x in nz]
    c = oracle.modulus_op(p, "center", A).view(np.int64)
    for x, r in zip(a, c):
        assert int(r) == (x - p if x >= p // 2 else x)
