"""GPU parity at degrees above 32768 (rq::Context accepts any power of two
from 8, math/src/rq/context.rs:42-56): the transforms split rows into halves
until they fit a workgroup (65536 -> 32768 -> 16384), and the BFV operations
run as elementwise passes around them. Every result is compared bit for bit
with the oracle; products and rotations are also decrypted. The plaintext
modulus 786433 = 3 * 2^18 + 1 supports SIMD encoding up to degree 131072."""
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

T = 786433


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


@pytest.mark.parametrize("n,L", [(65536, 1), (65536, 3), (131072, 2)])
def test_wide_ntt_matches_oracle(fh, oracle, n, L):
    print(f"wide ntt n={n} L={L}", flush=True)
    moduli = oracle.generate_moduli([62] * L, n)
    ctx = fh.Context(moduli, n)
    octx = oracle.Context(moduli, n)
    assert ctx.psi() == [int(x) for x in octx.psi()]
    count = 3
    a = oracle.random_polys(moduli, n, count, seed=n + L)
    d = fh.to_device(a)
    ctx.ntt_forward(d, count)
    fwd = fh.to_host(d)
    assert np.array_equal(fwd, octx.forward(a))
    ctx.ntt_backward(d, count)
    assert np.array_equal(fh.to_host(d), a)
    # lazy variant: outputs < 4p and congruent to the reduced transform
    d2 = fh.to_device(a)
    ctx.ntt_forward(d2, count, lazy=True)
    lz = fh.to_host(d2).reshape(count, L, n)
    ref = fwd.reshape(count, L, n)
    for l, p in enumerate(moduli):
        if 4 * p < 2**64:
            assert (lz[:, l] < np.uint64(4 * p)).all()
        assert np.array_equal(lz[:, l] % np.uint64(p), ref[:, l])


# (degree, moduli sizes, ciphertext level, key level): a special modulus, no
# special modulus, and a key two levels above the ciphertext.
MUL = [
    (65536, [60] * 4, 1, 0),
    (65536, [60] * 3, 0, 0),
    (65536, [60] * 5, 2, 0),
    (131072, [60] * 3, 1, 0),
]


@pytest.mark.parametrize("n,sizes,ct_level,key_level", MUL)
def test_wide_multiply_matches_oracle(fh, oracle, n, sizes, ct_level, key_level):
    print(f"wide multiply n={n} sizes={sizes} levels={ct_level},{key_level}", flush=True)
    bfv = oracle.Bfv(n, T, oracle.generate_moduli(sizes, n), sk_seed=n % 97 + ct_level)
    rk = bfv.relin_key(ct_level=ct_level, key_level=key_level, seed=3)
    ksk, cctx = _gpu_key(fh, bfv, rk, ct_level, key_level)
    gm = fh.Multiplicator(rk=ksk, plaintext=T)
    om = oracle.Multiplicator(bfv, rk, level=ct_level)
    rng = random.Random(n + ct_level)
    batch = 2
    vs = [[rng.randrange(T) for _ in range(n)] for _ in range(2 * batch)]
    lhs = np.concatenate([bfv.encrypt(vs[2 * i], level=ct_level, seed=10 + i) for i in range(batch)])
    rhs = np.concatenate([bfv.encrypt(vs[2 * i + 1], level=ct_level, seed=20 + i) for i in range(batch)])
    out = fh.empty(lhs.size)
    gm.multiply(fh.to_device(lhs), fh.to_device(rhs), out, batch)
    _sync()
    got = fh.to_host(out)
    w = lhs.size // batch
    for i in range(batch):
        g = got[i * w:(i + 1) * w]
        assert np.array_equal(g, om.multiply(lhs[i * w:(i + 1) * w], rhs[i * w:(i + 1) * w])), i
        assert [int(x) for x in bfv.decrypt(g, level=ct_level)] == \
            [(a * b) % T for a, b in zip(vs[2 * i], vs[2 * i + 1])]


def test_wide_multiply_without_relinearization(fh, oracle):
    """&ct * &ct (fhe/src/bfv/ops/mod.rs:259-342) at degree 65536."""
    n = 65536
    print(f"wide multiply without relinearization n={n}", flush=True)
    bfv = oracle.Bfv(n, T, oracle.generate_moduli([60] * 2, n), sk_seed=5)
    rng = random.Random(7)
    v1 = [rng.randrange(T) for _ in range(n)]
    v2 = [rng.randrange(T) for _ in range(n)]
    c1, c2 = bfv.encrypt(v1, seed=8), bfv.encrypt(v2, seed=9)
    ext = oracle.Multiplicator(bfv, None).ext_moduli()
    Q = 1
    for x in bfv.level_moduli(0):
        Q *= x
    gm = fh.Multiplicator(base_ctx=fh.Context(bfv.level_moduli(0), n), extended_basis=ext, post=(T, Q))
    out = fh.empty(3 * c1.size // 2)
    gm.multiply(fh.to_device(c1), fh.to_device(c2), out, 1)
    _sync()
    got = fh.to_host(out)
    assert np.array_equal(got, bfv.ct_mul(c1, c2))
    assert [int(x) for x in bfv.decrypt(got)] == [(a * b) % T for a, b in zip(v1, v2)]


@pytest.mark.parametrize("n,sizes,ct_level,key_level", [(65536, [60] * 3, 1, 0), (65536, [60] * 4, 2, 0),
                                                        (131072, [60] * 2, 0, 0)])
def test_wide_galois_matches_oracle(fh, oracle, n, sizes, ct_level, key_level):
    print(f"wide galois n={n} sizes={sizes} levels={ct_level},{key_level}", flush=True)
    bfv = oracle.Bfv(n, T, oracle.generate_moduli(sizes, n), sk_seed=len(sizes))
    rng = random.Random(n)
    v = [rng.randrange(T) for _ in range(n)]
    ct = bfv.encrypt(v, level=ct_level, seed=4)
    for k in [1, 7]:
        e = pow(3, k, 2 * n)
        gk = bfv.galois_key(e, ct_level=ct_level, key_level=key_level, seed=k)
        ksk, cctx = _gpu_key(fh, bfv, gk, ct_level, key_level)
        g = fh.GaloisKey(ksk, e)
        out = fh.empty(ct.size)
        g.relinearize(fh.to_device(ct), out, 1)
        _sync()
        got = fh.to_host(out)
        assert np.array_equal(got, bfv.galois(gk, ct))
        assert [int(x) for x in bfv.decrypt(got, level=ct_level)] == _rotate_columns(v, k)


@pytest.mark.parametrize("n,sizes,level,polys", [(65536, [60] * 3, 0, 2), (131072, [60] * 3, 1, 1)])
def test_wide_switch_down_matches_oracle(fh, oracle, n, sizes, level, polys):
    print(f"wide switch_down n={n} sizes={sizes} level={level}", flush=True)
    moduli = oracle.generate_moduli(sizes, n)
    hi, lo = moduli[:len(moduli) - level], moduli[:len(moduli) - level - 1]
    src = oracle.random_polys(hi, n, polys, seed=n + polys)
    sw = fh.Switcher(fh.Context(hi, n), fh.Context(lo, n))
    out = fh.empty(polys * len(lo) * n)
    sw.switch_down(fh.to_device(src), out, polys)
    _sync()
    assert np.array_equal(fh.to_host(out), oracle.ct_switch_down(oracle.Context(hi, n), oracle.Context(lo, n), src))


def test_wide_relinearize_matches_oracle(fh, oracle):
    n = 65536
    print(f"wide relinearize n={n}", flush=True)
    bfv = oracle.Bfv(n, T, oracle.generate_moduli([60] * 4, n), sk_seed=11)
    rk = bfv.relin_key(ct_level=1, key_level=0, seed=12)
    ksk, cctx = _gpu_key(fh, bfv, rk, 1, 0)
    rng = random.Random(13)
    v1 = [rng.randrange(T) for _ in range(n)]
    v2 = [rng.randrange(T) for _ in range(n)]
    ct3 = bfv.ct_mul(bfv.encrypt(v1, level=1, seed=14), bfv.encrypt(v2, level=1, seed=15), level=1)
    out = fh.empty(2 * ct3.size // 3)
    ksk.relinearize(fh.to_device(ct3), out, 1)
    _sync()
    got = fh.to_host(out)
    assert np.array_equal(got, bfv.relinearize(rk, ct3))
    assert [int(x) for x in bfv.decrypt(got, level=1)] == [(a * b) % T for a, b in zip(v1, v2)]
