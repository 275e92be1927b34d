"""Pins the oracle's ring layer (switch_down, substitute, rq::Scaler, key switch)
and BFV layer (Multiplicator, relinearization, Galois keys) with closed forms,
Python big-integer recomputation and decrypt-correctness."""
import random

import numpy as np
import pytest

from helpers import mul_and_relin_2_basis

# math/src/rq/mod.rs test moduli
MODULI = [1153, 4611686018326724609, 4611686018309947393, 4611686018232352769, 4611686018171535361,
          4611686018106523649]


def lift_poly(oracle, moduli, a, n):
    a = np.asarray(a, np.uint64).reshape(len(moduli), n)
    return [oracle.rns_lift(moduli, [int(a[i, j]) for i in range(len(moduli))]) for j in range(n)]


def project(vals, moduli):
    return np.array([[v % m for v in vals] for m in moduli], np.uint64).reshape(-1)


def test_switch_down_closed_form(oracle):
    n = 16
    ctx = oracle.Context(MODULI, n)
    a = ctx.random(1, seed=3)
    ref = lift_poly(oracle, MODULI, a, n)
    moduli = list(MODULI)
    cur = a
    while len(moduli) > 1:
        Q = 1
        for m in moduli:
            Q *= m
        c = oracle.Context(moduli, n)
        cur = c.switch_down(cur)
        moduli = moduli[:-1]
        Qn = Q // MODULI[len(moduli)]
        exp = [((b * Qn) + (Q >> 1)) // Q % Qn for b in ref]
        assert lift_poly(oracle, moduli, cur, n) == exp
        ref = exp


def negacyclic_mul(a, b, q):
    n = len(a)
    c = [0] * n
    for i in range(n):
        if a[i] == 0:
            continue
        for j in range(n):
            k = i + j
            if k < n:
                c[k] += a[i] * b[j]
            else:
                c[k - n] -= a[i] * b[j]
    return [x % q for x in c]


def test_substitute_closed_form(oracle):
    n = 16
    ctx = oracle.Context(MODULI[:3], n)
    a = ctx.random(1, seed=5)
    A = a.reshape(3, n)
    for k in [1, 3, 5, 11, 31, 2 * n - 1]:
        out = ctx.substitute(a, k, ntt_repr=False).reshape(3, n)
        for li, q in enumerate(MODULI[:3]):
            exp = [0] * n
            for i in range(n):
                e = (k * i) % (2 * n)
                if e >= n:
                    exp[e - n] = (exp[e - n] - int(A[li, i])) % q
                else:
                    exp[e] = (exp[e] + int(A[li, i])) % q
            assert [int(x) for x in out[li]] == exp
        # NTT-domain branch: substitute commutes with the NTT
        ntt_sub = ctx.substitute(ctx.forward(a), k, ntt_repr=True)
        assert np.array_equal(ntt_sub, ctx.forward(out.reshape(-1)))
    # 11 = 3^-1 mod 32
    assert np.array_equal(ctx.substitute(ctx.substitute(a, 3, False), 11, False), a)


def test_poly_scaler_switcher_closed_form(oracle):
    # math/src/rq/mod.rs:1044-1067 (Switcher = scale by to.modulus / from.modulus)
    n = 16
    c1 = oracle.Context(MODULI[:2], n)
    c2 = oracle.Context(MODULI[3:], n)
    Q1 = MODULI[0] * MODULI[1]
    Q2 = 1
    for m in MODULI[3:]:
        Q2 *= m
    sw = oracle.PolyScaler(c1, c2, Q2, Q1)
    a = c1.random(1, seed=8)
    ref = lift_poly(oracle, MODULI[:2], a, n)
    out = sw.scale(a, ntt_repr=False)
    assert lift_poly(oracle, MODULI[3:], out, n) == [((b * Q2) + (Q1 >> 1)) // Q1 for b in ref]
    # NTT branch equals power-basis branch framed by transforms
    out_ntt = sw.scale(c1.forward(a), ntt_repr=True)
    assert np.array_equal(out_ntt, c2.forward(out))


def test_key_switch_bigint_recomputation(oracle):
    """key_switch(p) = sum_i p_i * k_i over the key moduli, where p_i is the i-th
    residue row of p read as an integer polynomial (key_switching_key.rs:214-241),
    recomputed with Python integers from the key in power basis."""
    n = 16
    bfv = oracle.Bfv(n, 1153, oracle.generate_moduli([62, 62, 62], n), sk_seed=2)
    ct_level, key_level = 1, 0
    rk = bfv.relin_key(ct_level=ct_level, key_level=key_level, seed=4)
    cm = bfv.level_moduli(ct_level)
    km = bfv.level_moduli(key_level)
    kctx = oracle.Context(km, n)
    c0, c1 = rk.export()
    p = oracle.random_polys(cm, n, 1, seed=77)
    o0, o1 = rk.key_switch(p)
    P = p.reshape(len(cm), n)
    for key, out in ((c0, o0), (c1, o1)):
        kpb = [kctx.backward(key[i].reshape(-1)).reshape(len(km), n) for i in range(len(cm))]
        out_pb = kctx.backward(out).reshape(len(km), n)
        for j, q in enumerate(km):
            acc = [0] * n
            for i in range(len(cm)):
                term = negacyclic_mul([int(x) for x in P[i]], [int(x) for x in kpb[i][j]], q)
                acc = [(x + y) % q for x, y in zip(acc, term)]
            assert [int(x) for x in out_pb[j]] == acc


def _slots(rng, n, t):
    return [rng.randrange(t) for _ in range(n)]


@pytest.mark.parametrize("L,ct_level,key_level", [(2, 0, 0), (3, 0, 0), (3, 1, 0), (4, 1, 0), (4, 2, 1)])
def test_mul_relin_decrypts_tiny(oracle, L, ct_level, key_level):
    n, t = 16, 1153
    bfv = oracle.Bfv(n, t, oracle.generate_moduli([62] * L, n), sk_seed=10 + L)
    rng = random.Random(L)
    v1, v2 = _slots(rng, n, t), _slots(rng, n, t)
    c1 = bfv.encrypt(v1, level=ct_level, seed=21)
    c2 = bfv.encrypt(v2, level=ct_level, seed=22)
    assert [int(x) for x in bfv.decrypt(c1, level=ct_level)] == v1
    prod = [(a * b) % t for a, b in zip(v1, v2)]
    # no relinearization: 3-part ciphertext decrypts to the product
    m0 = oracle.Multiplicator(bfv, None, level=ct_level)
    ct3 = m0.multiply(c1, c2)
    assert [int(x) for x in bfv.decrypt(ct3, level=ct_level)] == prod
    assert np.array_equal(ct3, bfv.ct_mul(c1, c2, level=ct_level))
    rk = bfv.relin_key(ct_level=ct_level, key_level=key_level, seed=5)
    m = oracle.Multiplicator(bfv, rk, level=ct_level)
    ct = m.multiply(c1, c2)
    assert [int(x) for x in bfv.decrypt(ct, level=ct_level)] == prod
    # mul.rs path == (&ct * &ct) then relinearizes (relinearization_key.rs:71-98)
    assert np.array_equal(ct, bfv.relinearize(rk, ct3))


def _rotate_columns(v, k):
    h = len(v) // 2
    return v[k:h] + v[:k] + v[h + k:] + v[h:h + k]


@pytest.mark.parametrize("key_level", [0, 1])
def test_galois_rotation_decrypts_tiny(oracle, key_level):
    n, t = 16, 1153
    bfv = oracle.Bfv(n, t, oracle.generate_moduli([62] * 3, n), sk_seed=31)
    rng = random.Random(key_level)
    v = _slots(rng, n, t)
    ct = bfv.encrypt(v, level=1, seed=3)
    for k in [1, 2, 3]:
        e = pow(3, k, 2 * n)
        gk = bfv.galois_key(e, ct_level=1, key_level=key_level, seed=9 + k)
        out = bfv.galois(gk, ct)
        assert [int(x) for x in bfv.decrypt(out, level=1)] == _rotate_columns(v, k)
    gk = bfv.galois_key(2 * n - 1, ct_level=1, key_level=key_level, seed=40)
    out = bfv.galois(gk, ct)
    h = n // 2
    assert [int(x) for x in bfv.decrypt(out, level=1)] == v[h:] + v[:h]


@pytest.mark.slow
def test_mul_relin_decrypts_cfg1(oracle):
    """N=8192, 3 x 62-bit, t=65537 (the single-instance configuration)."""
    n, t = 8192, 65537
    bfv = oracle.Bfv(n, t, oracle.generate_moduli([62] * 3, n), sk_seed=1)
    rng = random.Random(1)
    v1, v2 = _slots(rng, n, t), _slots(rng, n, t)
    c1, c2 = bfv.encrypt(v1, seed=2), bfv.encrypt(v2, seed=3)
    rk = bfv.relin_key(seed=4)
    m = oracle.Multiplicator(bfv, rk)
    assert m.ext_moduli()[3:] == [4611686018426257409, 4611686018425815041, 4611686018424881153,
                                  4611686018424733697]
    ct = m.multiply(c1, c2)
    assert [int(x) for x in bfv.decrypt(ct)] == [(a * b) % t for a, b in zip(v1, v2)]


@pytest.mark.parametrize("L", [2, 3, 4])
def test_ciphertext_switch_down_decrypts(oracle, L):
    """Ciphertext::switch_down (ciphertext.rs:86-97) composed from the oracle's
    pinned NTT and Poly::switch_down: the ciphertext decrypts to the same
    slots one level down, including after a multiplication with modulus
    switching (mul.rs:221-223), for every level it can reach."""
    n, t = 16, 1153
    bfv = oracle.Bfv(n, t, oracle.generate_moduli([62] * L, n), sk_seed=50 + L)
    rng = random.Random(L)
    v1, v2 = _slots(rng, n, t), _slots(rng, n, t)
    ct = bfv.encrypt(v1, level=0, seed=61)
    for level in range(L - 1):
        cf = oracle.Context(bfv.level_moduli(level), n)
        ctn = oracle.Context(bfv.level_moduli(level + 1), n)
        ct = oracle.ct_switch_down(cf, ctn, ct)
        assert [int(x) for x in bfv.decrypt(ct, level=level + 1)] == v1
    c1 = bfv.encrypt(v1, level=0, seed=62)
    c2 = bfv.encrypt(v2, level=0, seed=63)
    rk = bfv.relin_key(ct_level=0, key_level=0, seed=64)
    prod = oracle.Multiplicator(bfv, rk, level=0).multiply(c1, c2)
    low = oracle.ct_switch_down(oracle.Context(bfv.level_moduli(0), n), oracle.Context(bfv.level_moduli(1), n),
                                prod)
    assert [int(x) for x in bfv.decrypt(low, level=1)] == [(a * b) % t for a, b in zip(v1, v2)]


@pytest.mark.parametrize("n,sizes", [(16, [62, 62, 62]), (64, [50, 55, 55])])
def test_mul_and_relin_2_decrypts(oracle, n, sizes):
    """Multiplicator::new(1, P/Q, basis, t/P) + relinearization (the
    mul_and_relin_2 strategy) decrypts to the slot products."""
    t = 1153
    bfv = oracle.Bfv(n, t, oracle.generate_moduli(sizes, n), sk_seed=17)
    ext, P, Q = mul_and_relin_2_basis(oracle, bfv.level_moduli(0), n)
    rk = bfv.relin_key(seed=18)
    om = oracle.Multiplicator(bfv, rk, 0, extended_basis=ext, rhs=(P, Q), post=(t, P))
    assert om.ext_moduli() == ext
    rng = random.Random(n)
    v1, v2 = _slots(rng, n, t), _slots(rng, n, t)
    prod = om.multiply(bfv.encrypt(v1, seed=1), bfv.encrypt(v2, seed=2))
    assert [int(x) for x in bfv.decrypt(prod)] == [(a * b) % t for a, b in zip(v1, v2)]


def _small(rng, n, bound=16):
    return np.array([rng.randint(-bound, bound) for _ in range(n)], np.int64)


@pytest.mark.parametrize("L,level", [(3, 0), (3, 1)])
def test_encode_encrypt_with_decrypt(oracle, L, level):
    """Plaintext::try_encode + to_poly, SecretKey::encrypt_poly with explicit
    samples, try_decrypt: the round trip returns the values, SIMD and
    coefficient encodings."""
    n, t = 16, 1153
    bfv = oracle.Bfv(n, t, oracle.generate_moduli([62] * L, n), sk_seed=4)
    moduli = bfv.level_moduli(level)
    rng = random.Random(L + level)
    for poly in (False, True):
        v = _slots(rng, n, t)
        pt = bfv.encode(v, level=level, poly=poly)
        a = oracle.random_polys(moduli, n, 1, seed=3)
        ct = bfv.encrypt_with(pt, a, _small(rng, n), level=level)
        assert np.array_equal(ct[len(moduli) * n:], a)
        assert [int(x) for x in bfv.decrypt(ct, level=level, poly=poly)] == v


def test_plaintext_poly_ntt_is_the_product_operand(oracle):
    """Plaintext::poly_ntt (the unscaled encoding) multiplies a ciphertext
    slot-wise (Mul<&Plaintext> for &Ciphertext): both parts times it decrypt
    to the slot products."""
    n, t = 16, 1153
    bfv = oracle.Bfv(n, t, oracle.generate_moduli([62] * 3, n), sk_seed=8)
    moduli = bfv.level_moduli(0)
    rng = random.Random(5)
    v1, v2 = _slots(rng, n, t), _slots(rng, n, t)
    ct = bfv.encrypt(v1, seed=9)
    p = bfv.encode(v2, scaled=False)
    prod = ct.copy()
    for part in range(2):
        for i, q in enumerate(moduli):
            o = (part * len(moduli) + i) * n
            prod[o:o + n] = [int(x) * int(y) % q for x, y in zip(ct[o:o + n], p[i * n:(i + 1) * n])]
    assert [int(x) for x in bfv.decrypt(prod)] == [(x * y) % t for x, y in zip(v1, v2)]


def test_public_key_encryption_decrypts(oracle):
    """PublicKey::new (a secret-key encryption of zero, public_key.rs:25-37)
    and PublicKey::try_encrypt with explicit u, e1, e2: decrypts to the slots."""
    n, t = 16, 1153
    bfv = oracle.Bfv(n, t, oracle.generate_moduli([62] * 3, n), sk_seed=12)
    moduli = bfv.level_moduli(0)
    rng = random.Random(12)
    pk = bfv.encrypt_with(bfv.encode([0] * n, poly=True), oracle.random_polys(moduli, n, 1, seed=4), _small(rng, n))
    v = _slots(rng, n, t)
    ct = bfv.encrypt_pk_with(pk, bfv.encode(v), _small(rng, n, 2), _small(rng, n), _small(rng, n))
    assert [int(x) for x in bfv.decrypt(ct)] == v


@pytest.mark.parametrize("L,level", [(2, 0), (3, 1)])
def test_rgsw_external_product_decrypts(oracle, L, level):
    """The reference's external_product test (rgsw_ciphertext.rs:176-210):
    ct1 x RGSW(m2) decrypts like ct1 x ct2, to the slot products."""
    n, t = 16, 1153
    bfv = oracle.Bfv(n, t, oracle.generate_moduli([62] * L, n), sk_seed=14)
    rng = random.Random(L)
    v1, v2 = _slots(rng, n, t), _slots(rng, n, t)
    ct1 = bfv.encrypt(v1, level=level, seed=5)
    rgsw = bfv.rgsw_encrypt(v2, level=level)
    prod = bfv.rgsw_product(rgsw, ct1)
    assert [int(x) for x in bfv.decrypt(prod, level=level)] == [(a * b) % t for a, b in zip(v1, v2)]


@pytest.mark.parametrize("L,level", [(4, 0), (4, 1)])
def test_ct_mul_any_parts_decrypts(oracle, L, level):
    """&ct * &ct for any part counts (fhe/src/bfv/ops/mod.rs:259-342): a
    3-part product times a 2-part ciphertext (4 parts), a square, and a
    1-part trivial encryption times a ciphertext decrypt to the slot
    products; the 2 x 2 case equals the fixed-shape path."""
    import random
    n, t = 64, 257
    bfv = oracle.Bfv(n, t, oracle.generate_moduli([62] * L, n), sk_seed=L + level)
    rng = random.Random(level)
    v1, v2, v3 = ([rng.randrange(t) for _ in range(n)] for _ in range(3))
    c1, c2, c3 = (bfv.encrypt(v, level=level, seed=s) for v, s in ((v1, 1), (v2, 2), (v3, 3)))
    p12 = bfv.ct_mul_parts(c1, c2, level)
    assert np.array_equal(p12, bfv.ct_mul(c1, c2, level))
    p123 = bfv.ct_mul_parts(p12, c3, level)
    assert [int(x) for x in bfv.decrypt(p123, level=level)] == [(a * b * c) % t for a, b, c in zip(v1, v2, v3)]
    sq = bfv.ct_mul_parts(c1, c1, level)
    assert [int(x) for x in bfv.decrypt(sq, level=level)] == [(a * a) % t for a in v1]
    triv = bfv.encode(v3, level=level, scaled=True)   # the one-part ciphertext (delta m)
    assert [int(x) for x in bfv.decrypt(triv, level=level)] == v3
    p = bfv.ct_mul_parts(triv, c2, level)
    assert [int(x) for x in bfv.decrypt(p, level=level)] == [(a * b) % t for a, b in zip(v3, v2)]


def test_level_multiplication_basis(oracle):
    """MultiplicationParameters (fhe/src/bfv/parameters.rs:585-612): at level
    l the multiplication context is the level's moduli followed by the first
    ceil((bits of the level's moduli + 60) / 62) of L + 1 generated 62-bit
    primes (= 1 mod 2N, none of the moduli); the same list at every level."""
    n = 256
    moduli = oracle.generate_moduli([62, 50, 62, 55], n)
    bfv = oracle.Bfv(n, 65537, moduli, sk_seed=1)
    ext, ub = [], 1 << 62
    while len(ext) != len(moduli) + 1:
        ub = oracle.generate_prime(62, 2 * n, ub)
        if ub not in ext and ub not in moduli:
            ext.append(ub)
    for level in range(len(moduli)):
        lm = moduli[:len(moduli) - level]
        k = -(-(sum(int(q).bit_length() for q in lm) + 60) // 62)
        assert bfv.mul_moduli(level) == [int(q) for q in lm] + ext[:k]


@pytest.mark.parametrize("sizes", [[62, 62], [37, 40], [30, 20, 37]])
def test_galois_rotation_decrypts_single_modulus_decomposition(oracle, sizes):
    """GaloisKey at the last level, where the key context has one modulus and
    KeySwitchingKey uses the power-of-two decomposition
    (fhe/src/bfv/keys/key_switching_key.rs:70-86, 172-210, 292-329): digits
    of ceil(log2 q / 2) bits; rotations still decrypt to the rotated slots,
    and the key has ceil(log2 q / log_base) rows of one limb."""
    n, t = 16, 97
    bfv = oracle.Bfv(n, t, oracle.generate_moduli(sizes, n), sk_seed=41)
    last = len(sizes) - 1
    q = bfv.level_moduli(last)[0]
    log_modulus = (q - 1).bit_length()
    digits = -(-log_modulus // (log_modulus // 2))
    rng = random.Random(len(sizes))
    v = _slots(rng, n, t)
    ct = bfv.encrypt(v, level=last, seed=3)
    for k in [1, 3]:
        e = pow(3, k, 2 * n)
        gk = bfv.galois_key(e, ct_level=last, key_level=last, seed=9 + k)
        c0, c1 = gk.export()
        assert c0.shape == (digits, 1, n)
        out = bfv.galois(gk, ct)
        assert [int(x) for x in bfv.decrypt(out, level=last)] == _rotate_columns(v, k)
