"""Encoding, secret-key encryption and decryption on the GPU, bit for bit
against the oracle: Plaintext::try_encode / to_poly (fhe/src/bfv/
plaintext_vec.rs:37-101, plaintext.rs:71-98), SecretKey::encrypt_poly with
explicit samples (secret_key.rs:102-137) and try_decrypt with SIMD or
coefficient decoding (secret_key.rs:200-282)."""
import random

import numpy as np
import pytest

from helpers import plaintext_moduli

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fh():
    import torch
    import fhehip
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    fhehip.lib()
    return fhehip


def _sync():
    import torch
    torch.cuda.synchronize()


def _codec(fh, bfv, level, with_sk=True):
    n = bfv.degree
    ctx = fh.Context(bfv.level_moduli(level), n)
    plain = fh.Context(plaintext_moduli(bfv.moduli, bfv.t), n)
    return fh.Codec(ctx, plain, bfv.t, sk=bfv.secret() if with_sk else None)


CASES = [
    # n, moduli sizes, t, level, num
    (16, [62, 62, 62], 1153, 0, 3),
    (16, [62, 62, 62], 1153, 2, 2),
    (4096, [36, 36, 37], 65537, 1, 2),
    (8192, [50, 55, 55], (1 << 20) + (1 << 19) + (1 << 17) + (1 << 16) + (1 << 14) + 1, 1, 2),
    (16384, [60] * 7, 65537, 1, 2),
    (32768, [60] * 4, 65537, 0, 1),
    (65536, [60] * 3, 786433, 1, 1),  # 786433 = 3 * 2^18 + 1: SIMD slots at degree 65536
]


@pytest.mark.parametrize("n,sizes,t,level,num", CASES)
def test_codec_matches_oracle(fh, oracle, n, sizes, t, level, num):
    print(f"codec n={n} sizes={sizes} t={t} level={level} num={num}", flush=True)
    bfv = oracle.Bfv(n, t, oracle.generate_moduli(sizes, n), sk_seed=n + level)
    moduli = bfv.level_moduli(level)
    L = len(moduli)
    codec = _codec(fh, bfv, level)
    rng = random.Random(n)
    for simd in (True, False):
        vals = [[rng.randrange(t) for _ in range(n)] for _ in range(num)]
        dv = fh.to_device(np.array(vals, np.uint64).reshape(-1))
        for scaled in (True, False):
            out = fh.empty(num * L * n)
            codec.encode(dv, num, out, simd=simd, scaled=scaled)
            _sync()
            got = fh.to_host(out).reshape(num, L * n)
            for b in range(num):
                assert np.array_equal(got[b], bfv.encode(vals[b], level=level, poly=not simd, scaled=scaled)), \
                    (simd, scaled, b)
        pt = fh.empty(num * L * n)
        codec.encode(dv, num, pt, simd=simd)
        a = oracle.random_polys(moduli, n, num, seed=7)
        e = np.array([rng.randint(-20, 20) for _ in range(num * n)], np.int64)
        e[0] = -(1 << 62)   # a large error coefficient, reduced like any other
        ct = fh.empty(num * 2 * L * n)
        codec.encrypt(pt, fh.to_device(a), fh.to_device(e.view(np.uint64)), ct, num)
        _sync()
        gct = fh.to_host(ct).reshape(num, 2 * L * n)
        hpt = fh.to_host(pt).reshape(num, L * n)
        for b in range(num):
            want = bfv.encrypt_with(hpt[b], a[b * L * n:(b + 1) * L * n], e[b * n:(b + 1) * n], level=level)
            assert np.array_equal(gct[b], want), (simd, b)
        e[0] = 0
        codec.encrypt(pt, fh.to_device(a), fh.to_device(e.view(np.uint64)), ct, num)
        dec = fh.empty(num * n)
        codec.decrypt(ct, 2, num, dec, simd=simd)
        _sync()
        assert np.array_equal(fh.to_host(dec).reshape(num, n), np.array(vals, np.uint64)), simd


@pytest.mark.parametrize("n,sizes", [(16, [62, 62, 62]), (8192, [62] * 3)])
def test_decrypt_three_parts_matches_oracle(fh, oracle, n, sizes):
    """try_decrypt of c0 + c1 s + c2 s^2 (a product before relinearization)."""
    t = 65537 if n > 16 else 1153
    bfv = oracle.Bfv(n, t, oracle.generate_moduli(sizes, n), sk_seed=3)
    rng = random.Random(n)
    v1 = [rng.randrange(t) for _ in range(n)]
    v2 = [rng.randrange(t) for _ in range(n)]
    ct3 = bfv.ct_mul(bfv.encrypt(v1, seed=1), bfv.encrypt(v2, seed=2))
    codec = _codec(fh, bfv, 0)
    out = fh.empty(n)
    codec.decrypt(fh.to_device(ct3), 3, 1, out)
    _sync()
    assert [int(x) for x in fh.to_host(out)] == [int(x) for x in bfv.decrypt(ct3)]
    assert [int(x) for x in fh.to_host(out)] == [(a * b) % t for a, b in zip(v1, v2)]


def test_codec_errors(fh, oracle):
    n = 1024
    moduli = oracle.generate_moduli([62, 62], n)
    ctx = fh.Context(moduli, n)
    plain = fh.Context(moduli[:2], n)
    c = fh.Codec(ctx, plain, 1153)   # 1153 != 1 mod 2048: coefficient encoding only
    buf = fh.empty(2 * 2 * n)
    with pytest.raises(fh.FheHipError):
        c.encode(buf, 1, buf, simd=True)
    c.encode(buf[:n], 1, buf, simd=False)
    with pytest.raises(fh.FheHipError):
        c.encrypt(buf, buf, buf, buf, 1)      # no secret key
    with pytest.raises(fh.FheHipError):
        c.decrypt(buf, 2, 1, buf)
    with pytest.raises(fh.FheHipError):
        fh.Codec(ctx, fh.Context(moduli[1:], n), 65537)   # plaintext context not a prefix


@pytest.mark.parametrize("n,sizes,level", [(16, [62, 62, 62], 1), (16384, [60] * 7, 1)])
def test_public_key_encryption_matches_oracle(fh, oracle, n, sizes, level):
    print(f"encrypt_pk n={n} level={level}", flush=True)
    t = 65537 if n > 16 else 1153
    bfv = oracle.Bfv(n, t, oracle.generate_moduli(sizes, n), sk_seed=21)
    moduli = bfv.level_moduli(level)
    rng = random.Random(n)
    small = lambda b: np.array([rng.randint(-b, b) for _ in range(n)], np.int64)
    pk = bfv.encrypt_with(bfv.encode([0] * n, level=level, poly=True), oracle.random_polys(moduli, n, 1, seed=2),
                          small(16), level=level)
    codec = _codec(fh, bfv, level, with_sk=False)
    v = [rng.randrange(t) for _ in range(n)]
    pt = bfv.encode(v, level=level)
    u, e1, e2 = small(1), small(16), small(16)
    ct = fh.empty(2 * len(moduli) * n)
    codec.encrypt_pk(fh.to_device(pk), fh.to_device(pt), fh.to_device(u.view(np.uint64)),
                     fh.to_device(e1.view(np.uint64)), fh.to_device(e2.view(np.uint64)), ct, 1)
    _sync()
    got = fh.to_host(ct)
    assert np.array_equal(got, bfv.encrypt_pk_with(pk, pt, u, e1, e2, level=level))
    assert [int(x) for x in bfv.decrypt(got, level=level)] == v


def test_codec_explicit_root_modulo_t(fh, oracle):
    """psi_t (the root of the parameters' NttOperator modulo t) is honoured:
    another primitive 2N-th root changes the slot layout of the encoding,
    and encryption and decryption under it still return the values."""
    n, t = 1024, 65537
    bfv = oracle.Bfv(n, t, oracle.generate_moduli([62, 62, 62], n), sk_seed=8)
    L = 3
    psi = fh.Context([t], n).psi()[0]
    psi3 = pow(psi, 3, t)   # an odd power of a primitive 2N-th root is one
    sk = bfv.secret()
    plain = fh.Context(bfv.moduli[:2], n)
    ctx = fh.Context(bfv.moduli, n)
    base, other = fh.Codec(ctx, plain, t, sk=sk), fh.Codec(ctx, plain, t, sk=sk, psi_t=psi3)
    same = fh.Codec(ctx, plain, t, sk=sk, psi_t=psi)
    rng = np.random.default_rng(1)
    v = rng.integers(0, t, n).astype(np.uint64)
    dv = fh.to_device(v)
    pts = [fh.empty(L * n) for _ in range(3)]
    for c, out in zip((base, other, same), pts):
        c.encode(dv, 1, out)
    _sync()
    assert np.array_equal(fh.to_host(pts[0]), fh.to_host(pts[2]))
    assert not np.array_equal(fh.to_host(pts[0]), fh.to_host(pts[1]))
    a = fh.to_device(oracle.random_polys(bfv.moduli, n, 1, seed=3))
    e = fh.to_device(np.zeros(n, np.int64).view(np.uint64))
    ct = fh.empty(2 * L * n)
    other.encrypt(pts[1], a, e, ct, 1)
    dec = fh.empty(n)
    other.decrypt(ct, 2, 1, dec)
    _sync()
    assert np.array_equal(fh.to_host(dec), v)
