"""GPU parity of key switching with keys any number of levels above the
ciphertext (RelinearizationKey::new_leveled, relinearization_key.rs:34-67;
GaloisKey::new, galois_key.rs:26-60): the switch is followed by
switch_down_to over every dropped modulus (math/src/rq/mod.rs:455-462), on
both the fused (N <= 16384) and the split (N = 32768) pipelines. Each result
is compared bit for bit with the oracle and decrypted."""
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

T = 65537


@pytest.fixture(scope="module")
def fh():
    import torch
    import fhehip
    assert torch.cuda.is_available()
    fhehip.lib()
    return fhehip


def _sync():
    import torch
    torch.cuda.synchronize()


def _gpu_key(fh, bfv, key, ct_level, key_level):
    n = bfv.degree
    cctx = fh.Context(bfv.level_moduli(ct_level), n)
    kctx = fh.Context(bfv.level_moduli(key_level), n)
    c0, c1 = key.export()
    return fh.KeySwitchingKey(cctx, kctx, c0, c1), cctx


def _rotate_columns(v, k):
    h = len(v) // 2
    return v[k:h] + v[:k] + v[h + k:] + v[h:h + k]


# (degree, moduli sizes, ciphertext level, key level): the key 2, 3 and 4
# moduli above the ciphertext at N = 16384 and N = 32768, and a 6-level gap.
DEEP = [
    (64, [60] * 8, 6, 0),
    (16384, [60] * 7, 4, 2),
    (16384, [60] * 7, 3, 0),
    (16384, [60] * 7, 4, 0),
    (32768, [60] * 6, 4, 2),
    (32768, [60] * 6, 3, 0),
    (32768, [60] * 6, 4, 0),
]


@pytest.mark.parametrize("n,sizes,ct_level,key_level", DEEP)
def test_deep_relinearization_and_multiply(fh, oracle, n, sizes, ct_level, key_level):
    print(f"deep key relin/multiply n={n} sizes={sizes} levels={ct_level},{key_level}", flush=True)
    bfv = oracle.Bfv(n, T, oracle.generate_moduli(sizes, n), sk_seed=ct_level + 3)
    rk = bfv.relin_key(ct_level=ct_level, key_level=key_level, seed=5)
    gk, cctx = _gpu_key(fh, bfv, rk, ct_level, key_level)
    L = cctx.L
    rng = random.Random(n + ct_level)
    batch = 2
    vs = [[rng.randrange(T) for _ in range(n)] for _ in range(2 * batch)]
    lhs = np.concatenate([bfv.encrypt(vs[2 * i], level=ct_level, seed=10 + i) for i in range(batch)])
    rhs = np.concatenate([bfv.encrypt(vs[2 * i + 1], level=ct_level, seed=20 + i) for i in range(batch)])
    # relinearization of &ct * &ct
    ct3 = np.concatenate([bfv.ct_mul(lhs[i * 2 * L * n:(i + 1) * 2 * L * n], rhs[i * 2 * L * n:(i + 1) * 2 * L * n],
                                     level=ct_level) for i in range(batch)])
    out = fh.empty(batch * 2 * L * n)
    gk.relinearize(fh.to_device(ct3), out, batch)
    _sync()
    got = fh.to_host(out)
    for i in range(batch):
        g = got[i * 2 * L * n:(i + 1) * 2 * L * n]
        assert np.array_equal(g, bfv.relinearize(rk, ct3[i * 3 * L * n:(i + 1) * 3 * L * n])), i
        assert [int(x) for x in bfv.decrypt(g, level=ct_level)] == \
            [(a * b) % T for a, b in zip(vs[2 * i], vs[2 * i + 1])]
    # Multiplicator::default with the deep key
    om = oracle.Multiplicator(bfv, rk, level=ct_level)
    gm = fh.Multiplicator(rk=gk, plaintext=T)
    out2 = fh.empty(batch * 2 * L * n)
    gm.multiply(fh.to_device(lhs), fh.to_device(rhs), out2, batch)
    _sync()
    assert np.array_equal(fh.to_host(out2), om.multiply(lhs, rhs))


@pytest.mark.parametrize("n,sizes,ct_level,key_level", DEEP)
def test_deep_galois(fh, oracle, n, sizes, ct_level, key_level):
    print(f"deep key galois n={n} sizes={sizes} levels={ct_level},{key_level}", flush=True)
    bfv = oracle.Bfv(n, T, oracle.generate_moduli(sizes, n), sk_seed=ct_level + 7)
    rng = random.Random(n)
    v = [rng.randrange(T) for _ in range(n)]
    ct = bfv.encrypt(v, level=ct_level, seed=4)
    k = 3
    e = pow(3, k, 2 * n)
    gkey = bfv.galois_key(e, ct_level=ct_level, key_level=key_level, seed=k)
    ksk, _ = _gpu_key(fh, bfv, gkey, ct_level, key_level)
    g = fh.GaloisKey(ksk, e)
    out = fh.empty(ct.size)
    g.relinearize(fh.to_device(ct), out, 1)
    _sync()
    got = fh.to_host(out)
    assert np.array_equal(got, bfv.galois(gkey, ct))
    assert [int(x) for x in bfv.decrypt(got, level=ct_level)] == _rotate_columns(v, k)


# One-modulus key contexts: the power-of-two decomposition of
# KeySwitchingKey (key_switching_key.rs:70-86, 172-210, 292-329), used by
# Galois keys at the last level.
@pytest.mark.parametrize("n,sizes", [(64, [62, 62]), (64, [37, 40]), (4096, [60, 60]), (16384, [62] * 3),
                                     (32768, [60] * 3)])
def test_single_modulus_galois_decomposition(fh, oracle, n, sizes):
    print(f"decomposition galois n={n} sizes={sizes}", flush=True)
    t = 257 if n == 64 else T
    bfv = oracle.Bfv(n, t, oracle.generate_moduli(sizes, n), sk_seed=n % 23)
    last = len(sizes) - 1
    ctx = fh.Context(bfv.level_moduli(last), n)
    rng = random.Random(n)
    v = [rng.randrange(t) for _ in range(n)]
    ct = bfv.encrypt(v, level=last, seed=4)
    for k in (1, 2):
        e = pow(3, k, 2 * n)
        gkey = bfv.galois_key(e, ct_level=last, key_level=last, seed=k)
        c0, c1 = gkey.export()
        ksk = fh.KeySwitchingKey(ctx, ctx, c0, c1)
        g = fh.GaloisKey(ksk, e)
        out = fh.empty(ct.size)
        g.relinearize(fh.to_device(ct), out, 1)
        # the key switch itself on a batch of two power-basis polynomials
        p = oracle.random_polys(bfv.level_moduli(last), n, 2, seed=k)
        ks = fh.empty(2 * 2 * n)
        ksk.key_switch(fh.to_device(p), ks, 2)
        _sync()
        got = fh.to_host(out)
        assert np.array_equal(got, bfv.galois(gkey, ct)), k
        assert [int(x) for x in bfv.decrypt(got, level=last)] == _rotate_columns(v, k)
        kh = fh.to_host(ks).reshape(2, 2, n)
        for i in range(2):
            e0, e1 = gkey.key_switch(p[i * n:(i + 1) * n])
            assert np.array_equal(kh[i, 0], e0) and np.array_equal(kh[i, 1], e1)
    # a key with the wrong number of decomposition rows is refused
    with pytest.raises(fh.FheHipError):
        fh.KeySwitchingKey(ctx, ctx, c0[:1], c1[:1])
