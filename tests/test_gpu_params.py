"""GPU parity of the BfvParameters handle (fhe/src/bfv/parameters.rs:431-641):
its level chain and per-level multiplication bases equal the oracle's, and
&ct * &ct (fhe/src/bfv/ops/mod.rs:259-342) at every level, for 2 x 2 parts,
a 3 x 2 product, a square and a one-part operand, equals the oracle bit for bit
and decrypts to the slot products."""
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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


@pytest.mark.parametrize("n,sizes,t", [(64, [62, 62, 62, 62], 257), (4096, [55, 60, 62, 50], 65537),
                                       (16384, [60] * 7, 65537), (32768, [60] * 4, 65537)])
def test_ct_mul_at_every_level(fh, oracle, n, sizes, t):
    print(f"parameters n={n} sizes={sizes}", flush=True)
    moduli = oracle.generate_moduli(sizes, n)
    bfv = oracle.Bfv(n, t, moduli, sk_seed=n % 89)
    par = fh.Parameters(n, t, moduli)
    assert (par.max_level, par.degree, par.plaintext) == (len(moduli) - 1, n, t)
    from helpers import plaintext_moduli
    assert par.plaintext_context().moduli == [int(q) for q in plaintext_moduli(moduli, t)]
    rng = random.Random(n)
    for level in range(len(moduli)):
        ctx = par.context(level)
        assert ctx.L == len(moduli) - level and ctx.degree == n
        assert par.mul_basis(level) == bfv.mul_moduli(level)
        vs = [[rng.randrange(t) for _ in range(n)] for _ in range(4)]
        c = [bfv.encrypt(v, level=level, seed=10 * level + k) for k, v in enumerate(vs)]
        L = len(bfv.level_moduli(level))
        w = L * n
        # 2 x 2 on a batch of two pairs
        lhs, rhs = np.concatenate([c[0], c[2]]), np.concatenate([c[1], c[3]])
        out = fh.empty(2 * 3 * w)
        par.ct_mul(level, fh.to_device(lhs), 2, fh.to_device(rhs), 2, out, 2)
        _sync()
        got = fh.to_host(out).reshape(2, 3 * w)
        assert np.array_equal(got[0], bfv.ct_mul_parts(c[0], c[1], level)), level
        assert np.array_equal(got[1], bfv.ct_mul_parts(c[2], c[3], level)), level
        decrypts = L >= 2   # one modulus leaves too little noise room for products
        if decrypts:
            assert [int(x) for x in bfv.decrypt(got[0], level=level)] == \
                [(a * b) % t for a, b in zip(vs[0], vs[1])]
        # 3 x 2, a square and a one-part operand
        d3 = fh.empty(4 * w)
        par.ct_mul(level, fh.to_device(got[0]), 3, fh.to_device(c[2]), 2, d3, 1)
        sq = fh.empty(3 * w)
        dc = fh.to_device(c[0])
        par.ct_mul(level, dc, 2, dc, 2, sq, 1)
        triv = bfv.encode(vs[3], level=level, scaled=True)
        d1 = fh.empty(2 * w)
        par.ct_mul(level, fh.to_device(triv), 1, fh.to_device(c[1]), 2, d1, 1)
        _sync()
        assert np.array_equal(fh.to_host(d3), bfv.ct_mul_parts(got[0], c[2], level)), level
        assert np.array_equal(fh.to_host(sq), bfv.ct_mul_parts(c[0], c[0], level)), level
        assert np.array_equal(fh.to_host(d1), bfv.ct_mul_parts(triv, c[1], level)), level
        if decrypts and n <= 4096:
            assert [int(x) for x in bfv.decrypt(fh.to_host(sq), level=level)] == [(a * a) % t for a in vs[0]]


def test_parameters_errors(fh, oracle):
    n = 1024
    moduli = oracle.generate_moduli([60, 60], n)
    with pytest.raises(fh.FheHipError) as e:
        fh.Parameters(1000, 65537, moduli)
    assert e.value.code == fh.ERR_INVALID_ARG
    with pytest.raises(fh.FheHipError) as e:
        fh.Parameters(n, 65537, [1 << 61])
    assert e.value.code == fh.ERR_UNSUPPORTED   # not an NTT-friendly prime (NttOperator::new returns None)
    with pytest.raises(fh.FheHipError):
        fh.Parameters(n, 1, moduli)
    par = fh.Parameters(n, 65537, moduli)
    with pytest.raises(fh.FheHipError) as e:
        par.context(2)
    assert e.value.code == fh.ERR_BAD_LEVEL
    d = fh.empty(4 * 2 * n)
    with pytest.raises(fh.FheHipError) as e:
        par.ct_mul(2, d, 2, d, 2, d, 1)
    assert e.value.code == fh.ERR_BAD_LEVEL
    with pytest.raises(fh.FheHipError):
        par.ct_mul(0, d, 9, d, 9, d, 1)   # 17 output parts
    with pytest.raises(fh.FheHipError):
        par.ct_mul(0, d, 0, d, 2, d, 1)
