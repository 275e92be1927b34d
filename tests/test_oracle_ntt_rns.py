"""Pins the oracle's NTT (math/src/ntt/native.rs), RnsContext (math/src/rns/mod.rs)
and RnsScaler (math/src/rns/scaler.rs) with closed forms and the reference KATs."""
import random

import numpy as np
import pytest

P62 = 4611686018326724609


def brv(i, logn):
    return int(format(i, f"0{logn}b")[::-1], 2) if logn else 0


def ntt_closed_form(a, p, psi):
    """A[i] = a(psi^(2 brv(i) + 1)) mod p — the negacyclic evaluation the
    forward operator produces, in bit-reversed order."""
    n = len(a)
    logn = n.bit_length() - 1
    out = []
    for i in range(n):
        x = pow(psi, 2 * brv(i, logn) + 1, p)
        acc, xp = 0, 1
        for c in a:
            acc = (acc + c * xp) % p
            xp = xp * x % p
        out.append(acc)
    return out


@pytest.mark.parametrize("p,n", [(P62, 8), (P62, 16), (P62, 64), (12289, 16), (65537, 32),
                                 (1152921504606748673, 32), (4611686018427322369, 128)])
def test_ntt_closed_form_and_roundtrip(oracle, p, n):
    rng = random.Random(n + p)
    psi = oracle.primitive_root(n, p)
    assert pow(psi, 2 * n, p) == 1 and pow(psi, n, p) == p - 1
    a = [rng.randrange(p) for _ in range(n)]
    A = oracle.ntt(p, n, np.array(a, np.uint64), "forward")
    assert [int(x) for x in A] == ntt_closed_form(a, p, psi)
    assert [int(x) for x in oracle.ntt(p, n, A, "forward_vt")] == [int(x) for x in oracle.ntt(p, n, A, "forward")]
    lazy = oracle.ntt(p, n, np.array(a, np.uint64), "forward_vt_lazy")
    assert all(int(x) < 4 * p for x in lazy)
    assert [int(x) % p for x in lazy] == [int(x) for x in A]
    back = oracle.ntt(p, n, A, "backward")
    assert [int(x) for x in back] == a


@pytest.mark.parametrize("n", [8, 32])
def test_ntt_negacyclic_convolution(oracle, n):
    p = P62
    rng = random.Random(n)
    a = [rng.randrange(p) for _ in range(n)]
    b = [rng.randrange(p) for _ in range(n)]
    c = [0] * n
    for i in range(n):
        for j in range(n):
            k = i + j
            if k < n:
                c[k] = (c[k] + a[i] * b[j]) % p
            else:
                c[k - n] = (c[k - n] - a[i] * b[j]) % p
    A = oracle.ntt(p, n, np.array(a, np.uint64))
    B = oracle.ntt(p, n, np.array(b, np.uint64))
    prod = np.array([(int(x) * int(y)) % p for x, y in zip(A, B)], np.uint64)
    assert [int(x) for x in oracle.ntt(p, n, prod, "backward")] == c


def test_ntt_tables_match_definition(oracle):
    p, n = P62, 64
    psi = oracle.primitive_root(n, p)
    tabs, ninv = oracle.ntt_tables(p, n)
    logn = 6
    inv = pow(psi, -1, p)
    for i in range(n):
        j = brv(i, logn)
        assert int(tabs[0][i]) == pow(psi, j, p)
        assert int(tabs[1][i]) == (int(tabs[0][i]) << 64) // p
        assert int(tabs[2][i]) == pow(inv, j + 1, p)
        assert int(tabs[3][i]) == (int(tabs[2][i]) << 64) // p
    assert ninv == pow(n, -1, p)


def test_rns_project_lift_kat(oracle):
    # math/src/rns/mod.rs:205-242
    m = [4, 15, 1153]
    prod = 4 * 15 * 1153
    for v, rests in [(0, [0, 0, 0]), (4, [0, 4, 4]), (15, [3, 0, 15]), (1153, [1, 13, 0]), (prod - 1, [3, 14, 1152])]:
        assert [v % x for x in m] == rests
        assert oracle.rns_lift(m, rests) == v
    g = oracle.rns_garner(m)
    for i, q in enumerate(m):
        qs = prod // q
        assert g[i] == qs * pow(qs, -1, q)
    rng = random.Random(3)
    for _ in range(100):
        b = rng.randrange(prod)
        assert oracle.rns_lift(m, [b % x for x in m]) == b


def _scaled_round(x_lift, Q, n, d, R):
    """Closed form of the reference's own scaler test (math/src/rns/scaler.rs:397-414)."""
    sign = x_lift >= (Q >> 1)
    if sign:
        x_lift = Q - x_lift
        if d % 2 == 0:
            return R - ((x_lift * n + ((d >> 1) - 1)) // d) % R
        return R - ((x_lift * n + (d >> 1)) // d) % R
    return (x_lift * n + (d >> 1)) // d


NUMS = [1, 2, 3, 100, 1000, 4611686018326724610]
DENS = [1, 2, 3, 4, 100, 101, 1000, 1001, 4611686018326724610]


@pytest.mark.parametrize("different", [False, True])
def test_rns_scaler_closed_form(oracle, different):
    q = [4, 4611686018326724609, 1153]
    r = q + [4611686018309947393, 4611686018282684417, 4611686018257518593, 4611686018232352769,
             4611686018171535361, 4611686018106523649, 4611686018058289153] if different else q
    Q, R = 4 * 4611686018326724609 * 1153, 1
    for x in r:
        R *= x
    rng = random.Random(5 + different)
    for n in NUMS:
        for d in DENS:
            sc = oracle.RnsScaler(q, r, n, d)
            xs = [[rng.randrange(m) for m in q] for _ in range(40)]
            out = sc.scale(np.array(xs, np.uint64))
            for x, y in zip(xs, out):
                xl = oracle.rns_lift(q, x)
                exp = _scaled_round(xl, Q, n, d, R)
                assert [int(v) for v in y] == [exp % m for m in r], (n, d, x)


def test_rns_scaler_config_shapes(oracle):
    """Base extension 6->13 and the t/Q down-scaling 13->6 at the cfg3 moduli,
    against the exact rounding closed form on random columns."""
    q = oracle.generate_moduli([60] * 6, 16384)
    ext = [4611686018427322369, 4611686018427289601, 4611686018425815041, 4611686018424733697,
           4611686018423881729, 4611686018423390209, 4611686018423062529]
    m = q + ext
    Q, M = 1, 1
    for x in q:
        Q *= x
    for x in m:
        M *= x
    rng = random.Random(9)
    up = oracle.RnsScaler(q, m, 1, 1)
    xs = [[rng.randrange(v) for v in q] for _ in range(200)]
    out = up.scale(np.array(xs, np.uint64), start=6, length=7)
    for x, y in zip(xs, out):
        xl = oracle.rns_lift(q, x)
        v = xl if xl < (Q >> 1) else xl - Q  # centered representative
        assert [int(a) for a in y] == [v % e for e in ext]
    t = 65537
    down = oracle.RnsScaler(m, q, t, Q)
    xs = [[rng.randrange(v) for v in m] for _ in range(200)]
    out = down.scale(np.array(xs, np.uint64))
    for x, y in zip(xs, out):
        xl = oracle.rns_lift(m, x)
        exp = _scaled_round(xl, M, t, Q, Q)
        assert [int(a) for a in y] == [exp % e for e in q]
