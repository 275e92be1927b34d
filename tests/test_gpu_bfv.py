"""GPU parity of the BFV evaluation path against the oracle, bit for bit:
rq::Scaler, key switching, relinearization, Multiplicator::multiply and Galois
rotations, plus decrypt-correctness of the GPU outputs."""
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


# ---- scaler -----------------------------------------------------------------
@pytest.mark.parametrize("n", [16, 4096, 32768])
def test_scaler_extension_and_down(fh, oracle, n):
    q = oracle.generate_moduli([60] * 6, n)
    ext = q + [m for m in oracle.generate_moduli([62] * 8, n) if m not in q][:7]
    cq, cm = fh.Context(q, n), fh.Context(ext, n)
    oq, om = oracle.Context(q, n), oracle.Context(ext, n)
    Q = 1
    for x in q:
        Q *= x
    for ntt in (False, True):
        a = oracle.random_polys(q, n, 3, seed=n + ntt)
        up = fh.Scaler(cq, cm)
        d_out = fh.empty(3 * len(ext) * n)
        up.scale(fh.to_device(a), d_out, 3, ntt)
        _sync()
        assert np.array_equal(fh.to_host(d_out), oracle.PolyScaler(oq, om).scale(a, ntt))
        b = oracle.random_polys(ext, n, 3, seed=n + 7 + ntt)
        down = fh.Scaler(cm, cq, 65537, Q)
        d_out2 = fh.empty(3 * len(q) * n)
        down.scale(fh.to_device(b), d_out2, 3, ntt)
        _sync()
        assert np.array_equal(fh.to_host(d_out2), oracle.PolyScaler(om, oq, 65537, Q).scale(b, ntt))


# source moduli counts of the ciphertext part and the extension part: the
# extension scales nq -> nq + ne by one, the down-scaling nq + ne -> nq by
# t / Q; at most 8 source moduli take MFMA steps of 16 k, more take 32
@pytest.mark.parametrize("nq,ne,n", [(1, 2, 4096), (3, 4, 4096), (6, 7, 4096), (8, 8, 4096), (5, 4, 4096),
                                     (2, 7, 4096), (6, 7, 256), (3, 9, 512)])
def test_scaler_matrix_cores(fh, oracle, nq, ne, n):
    """The scaler's matrix-core form (8-bit integer MFMA over base-256
    digits), taken by reduced inputs (the inverse transform's outputs of
    NTT-domain polynomials), against the oracle and against the per-column
    VALU form. Power-basis columns 0, 1 and 2 hold all zeros, all p - 1 and
    alternating zeros and p - 1. N = 256 is the smallest degree of that form."""
    print(f"scaler matrix cores nq={nq} ne={ne} n={n}", flush=True)
    q = oracle.generate_moduli([60] * nq, n)
    ext = [m for m in oracle.generate_moduli([62] * (nq + ne), n) if m not in q][:ne]
    full = q + ext
    oq, of = oracle.Context(q, n), oracle.Context(full, n)
    cq, cf = fh.Context(q, n), fh.Context(full, n)
    Q = 1
    for x in q:
        Q *= x

    def edge_ntt(moduli, octx, seed):
        a = oracle.random_polys(moduli, n, 3, seed=seed).reshape(3, len(moduli), n)
        for i, p in enumerate(moduli):
            a[:, i, 0] = 0
            a[:, i, 1] = p - 1
            a[:, i, 2] = p - 1 if i % 2 else 0
        return octx.forward(a.reshape(-1))

    cases = [(cq, cf, oq, of, None, None, q), (cf, cq, of, oq, 65537, Q, full)]
    prev = fh.tuning_get(n)
    try:
        for fc, tc, fo, to, num, den, src in cases:
            a = edge_ntt(src, fo, seed=nq * 17 + ne + (num or 0))
            want = oracle.PolyScaler(fo, to, num, den).scale(a, True) if num else \
                oracle.PolyScaler(fo, to).scale(a, True)
            sc = fh.Scaler(fc, tc, num, den) if num else fh.Scaler(fc, tc)
            for mfma in (1, 0):
                fh.tuning_set(n, scaler_mfma=mfma)
                out = fh.empty(3 * tc.L * n)
                sc.scale(fh.to_device(a), out, 3, True)
                _sync()
                assert np.array_equal(fh.to_host(out), want), (num, mfma)
    finally:
        fh.tuning_set(n, **prev)


def test_scaler_general_factors(fh, oracle):
    """The reference's own scaler test grid (math/src/rns/scaler.rs:379-473) on
    polynomials, both directions of context size."""
    n = 64
    q = oracle.generate_moduli([62, 62, 40], n)
    r = q + oracle.generate_moduli([61, 61, 61, 50], n)
    cq, cr = fh.Context(q, n), fh.Context(r, n)
    oq, orr = oracle.Context(q, n), oracle.Context(r, n)
    a = oracle.random_polys(q, n, 2, seed=3)
    for num in [1, 2, 3, 100, 1000, 4611686018326724610]:
        for den in [1, 2, 3, 4, 100, 101, 1000, 1001, 4611686018326724610]:
            for (fc, tc, fo, to) in ((cq, cq, oq, oq), (cq, cr, oq, orr)):
                s = fh.Scaler(fc, tc, num, den)
                out = fh.empty(2 * tc.L * n)
                s.scale(fh.to_device(a), out, 2, False)
                _sync()
                exp = oracle.PolyScaler(fo, to, num, den).scale(a, False)
                assert np.array_equal(fh.to_host(out), exp), (num, den, tc.L)


# ---- key switching --------------------------------------------------------------
def _bfv(oracle, n, sizes, t=65537, seed=1):
    return oracle.Bfv(n, t, oracle.generate_moduli(sizes, n), sk_seed=seed)


def _gpu_key(fh, bfv, key, ct_level, key_level):
    n = bfv.degree
    cctx = fh.Context(bfv.level_moduli(ct_level), n)
    kctx = fh.Context(bfv.level_moduli(key_level), n)
    c0, c1 = key.export()
    return fh.KeySwitchingKey(cctx, kctx, c0, c1), cctx, kctx


@pytest.mark.parametrize("n,L,ct_level,key_level", [(16, 2, 0, 0), (64, 3, 1, 0), (1024, 4, 2, 0),
                                                    (8192, 3, 0, 0), (16384, 7, 1, 0), (32768, 4, 1, 0),
                                                    (32768, 3, 0, 0)])
def test_key_switch_matches_oracle(fh, oracle, n, L, ct_level, key_level):
    print(f"key switch n={n} L={L} levels={ct_level},{key_level}", flush=True)
    bfv = _bfv(oracle, n, [60] * L, seed=L)
    rk = bfv.relin_key(ct_level=ct_level, key_level=key_level, seed=9)
    gk, cctx, kctx = _gpu_key(fh, bfv, rk, ct_level, key_level)
    cnt = 2
    p = oracle.random_polys(cctx.moduli, n, cnt, seed=5)
    out = fh.empty(cnt * 2 * kctx.L * n)
    gk.key_switch(fh.to_device(p), out, cnt)
    _sync()
    got = fh.to_host(out).reshape(cnt, 2, kctx.L * n)
    for i in range(cnt):
        e0, e1 = rk.key_switch(p[i * cctx.L * n:(i + 1) * cctx.L * n])
        assert np.array_equal(got[i, 0], e0)
        assert np.array_equal(got[i, 1], e1)


@pytest.mark.parametrize("n,L,ct_level,key_level", [(16, 3, 0, 0), (32, 3, 1, 0), (64, 4, 2, 0),
                                                    (8192, 3, 0, 0), (32768, 3, 1, 0)])
def test_relinearize_matches_oracle(fh, oracle, n, L, ct_level, key_level):
    bfv = _bfv(oracle, n, [62] * L, seed=L + 1)
    rng = random.Random(n)
    v1 = [rng.randrange(65537) for _ in range(n)]
    v2 = [rng.randrange(65537) for _ in range(n)]
    c1 = bfv.encrypt(v1, level=ct_level, seed=3)
    c2 = bfv.encrypt(v2, level=ct_level, seed=4)
    ct3 = bfv.ct_mul(c1, c2, level=ct_level)
    rk = bfv.relin_key(ct_level=ct_level, key_level=key_level, seed=2)
    gk, cctx, kctx = _gpu_key(fh, bfv, rk, ct_level, key_level)
    out = fh.empty(2 * cctx.L * n)
    gk.relinearize(fh.to_device(ct3), out, 1)
    _sync()
    got = fh.to_host(out)
    assert np.array_equal(got, bfv.relinearize(rk, ct3))
    assert [int(x) for x in bfv.decrypt(got, level=ct_level)] == [(a * b) % 65537 for a, b in zip(v1, v2)]


# ---- multiplication ---------------------------------------------------------------
MUL_CASES = [
    # n, moduli sizes, ct_level, key_level, batch
    (16, [62, 62], 0, 0, 3),
    (32, [62, 62, 62], 1, 0, 2),
    (64, [50, 55, 60, 62], 1, 0, 2),
    (1024, [60] * 4, 2, 0, 2),
    (8192, [62] * 3, 0, 0, 2),        # single-instance configuration shape
    (16384, [60] * 7, 1, 0, 2),       # batched configuration shape: 6 x 60 + special prime
    (32768, [60] * 4, 1, 0, 2),       # unfused pipeline, one special prime
    (32768, [55, 60, 62], 0, 0, 1),   # unfused pipeline, no special prime
]


@pytest.mark.parametrize("n,sizes,ct_level,key_level,batch", MUL_CASES)
def test_multiply_matches_oracle(fh, oracle, n, sizes, ct_level, key_level, batch):
    print(f"multiply n={n} sizes={sizes} levels={ct_level},{key_level} batch={batch}", flush=True)
    bfv = _bfv(oracle, n, sizes, seed=n)
    rk = bfv.relin_key(ct_level=ct_level, key_level=key_level, seed=17)
    om = oracle.Multiplicator(bfv, rk, level=ct_level)
    gk, cctx, kctx = _gpu_key(fh, bfv, rk, ct_level, key_level)
    gm = fh.Multiplicator(rk=gk, plaintext=65537)
    assert gm.extended_basis() == om.ext_moduli()
    L = cctx.L
    lhs = oracle.random_polys(cctx.moduli, n, 2 * batch, seed=100 + n)
    rhs = oracle.random_polys(cctx.moduli, n, 2 * batch, seed=200 + n)
    out = fh.empty(batch * 2 * L * n)
    gm.multiply(fh.to_device(lhs), fh.to_device(rhs), out, batch)
    _sync()
    assert np.array_equal(fh.to_host(out), om.multiply(lhs, rhs))


@pytest.mark.parametrize("n,L", [(16, 2), (8192, 3), (32768, 3)])
def test_multiply_decrypts(fh, oracle, n, L):
    t = 65537
    bfv = _bfv(oracle, n, [62] * L, seed=3)
    rng = random.Random(L)
    v1 = [rng.randrange(t) for _ in range(n)]
    v2 = [rng.randrange(t) for _ in range(n)]
    c1, c2 = bfv.encrypt(v1, seed=8), bfv.encrypt(v2, seed=9)
    rk = bfv.relin_key(seed=10)
    gk, cctx, kctx = _gpu_key(fh, bfv, rk, 0, 0)
    gm = fh.Multiplicator(rk=gk, plaintext=t)
    out = fh.empty(2 * L * n)
    gm.multiply(fh.to_device(c1), fh.to_device(c2), out, 1)
    _sync()
    assert [int(x) for x in bfv.decrypt(fh.to_host(out))] == [(a * b) % t for a, b in zip(v1, v2)]


def test_multiply_without_relinearization(fh, oracle):
    """Multiplicator::new with the per-level extended basis and factors (1, 1, t/Q)
    equals &ct * &ct (fhe/src/bfv/ops/mod.rs:259-342)."""
    n, t = 64, 65537
    bfv = _bfv(oracle, n, [62] * 3, seed=5)
    rng = random.Random(1)
    v1 = [rng.randrange(t) for _ in range(n)]
    v2 = [rng.randrange(t) for _ in range(n)]
    c1, c2 = bfv.encrypt(v1, seed=8), bfv.encrypt(v2, seed=9)
    om = oracle.Multiplicator(bfv, None)
    ext = om.ext_moduli()
    cctx = fh.Context(bfv.level_moduli(0), n)
    Q = 1
    for x in bfv.level_moduli(0):
        Q *= x
    gm = fh.Multiplicator(base_ctx=cctx, extended_basis=ext, post=(t, Q))
    out = fh.empty(3 * 3 * n)
    gm.multiply(fh.to_device(c1), fh.to_device(c2), out, 1)
    _sync()
    got = fh.to_host(out)
    assert np.array_equal(got, bfv.ct_mul(c1, c2))
    assert [int(x) for x in bfv.decrypt(got)] == [(a * b) % t for a, b in zip(v1, v2)]


def test_batch_equals_single_calls(fh, oracle):
    """Each instance's result is the same alone and inside a batch."""
    n, batch = 4096, 5
    bfv = _bfv(oracle, n, [60] * 4, seed=2)
    rk = bfv.relin_key(ct_level=1, key_level=0, seed=3)
    gk, cctx, kctx = _gpu_key(fh, bfv, rk, 1, 0)
    gm = fh.Multiplicator(rk=gk, plaintext=65537)
    L = cctx.L
    lhs = oracle.random_polys(cctx.moduli, n, 2 * batch, seed=1)
    rhs = oracle.random_polys(cctx.moduli, n, 2 * batch, seed=2)
    dl, dr = fh.to_device(lhs), fh.to_device(rhs)
    out = fh.empty(batch * 2 * L * n)
    gm.multiply(dl, dr, out, batch)
    _sync()
    allb = fh.to_host(out).reshape(batch, -1)
    sz = 2 * L * n
    for i in range(batch):
        one = fh.empty(sz)
        gm.multiply(dl[i * sz:(i + 1) * sz], dr[i * sz:(i + 1) * sz], one, 1)
        _sync()
        assert np.array_equal(fh.to_host(one), allb[i])


# ---- Galois -------------------------------------------------------------------
def _rotate_columns(v, k):
    h = len(v) // 2
    return v[k:h] + v[:k] + v[h + k:] + v[h:h + k]


@pytest.mark.parametrize("n,L,ct_level,key_level", [(16, 3, 1, 0), (64, 3, 0, 0), (4096, 4, 2, 0),
                                                    (16384, 7, 1, 0), (32768, 4, 1, 0)])
def test_galois_matches_oracle(fh, oracle, n, L, ct_level, key_level):
    print(f"galois n={n} L={L}", flush=True)
    t = 65537
    bfv = _bfv(oracle, n, [60] * L, t=t, seed=L)
    rng = random.Random(n)
    v = [rng.randrange(t) for _ in range(n)]
    ct = bfv.encrypt(v, level=ct_level, seed=4)
    for k in [1, 5]:
        e = pow(3, k, 2 * n)
        gk = bfv.galois_key(e, ct_level=ct_level, key_level=key_level, seed=k)
        ksk, cctx, kctx = _gpu_key(fh, bfv, gk, ct_level, key_level)
        g = fh.GaloisKey(ksk, e)
        out = fh.empty(ct.size)
        g.relinearize(fh.to_device(ct), out, 1)
        _sync()
        got = fh.to_host(out)
        assert np.array_equal(got, bfv.galois(gk, ct))
        assert [int(x) for x in bfv.decrypt(got, level=ct_level)] == _rotate_columns(v, k)


@pytest.mark.parametrize("n,sizes", [(16, [62, 62, 62]), (4096, [50, 55, 55]), (16384, [62] * 6)])
def test_mul_and_relin_2_matches_oracle(fh, oracle, n, sizes):
    """The second multiplication strategy of fhe/benches/bfv.rs:257-286:
    Multiplicator::new(1, P/Q, basis, t/P) with relinearization."""
    from helpers import mul_and_relin_2_basis
    print(f"mul_and_relin_2 n={n} sizes={sizes}", flush=True)
    t = 65537
    bfv = _bfv(oracle, n, sizes, t=t, seed=6)
    ext, P, Q = mul_and_relin_2_basis(oracle, bfv.level_moduli(0), n)
    rk = bfv.relin_key(seed=7)
    gk, cctx, _ = _gpu_key(fh, bfv, rk, 0, 0)
    om = oracle.Multiplicator(bfv, rk, 0, extended_basis=ext, rhs=(P, Q), post=(t, P))
    gm = fh.Multiplicator(rk=gk, base_ctx=cctx, extended_basis=ext, rhs=(P, Q), post=(t, P))
    assert gm.extended_basis() == ext
    rng = random.Random(n)
    batch = 2
    vs = [[rng.randrange(t) for _ in range(n)] for _ in range(2 * batch)]
    lhs = np.concatenate([bfv.encrypt(vs[2 * i], seed=10 + i) for i in range(batch)])
    rhs = np.concatenate([bfv.encrypt(vs[2 * i + 1], seed=20 + i) for i in range(batch)])
    out = fh.empty(lhs.size)
    gm.multiply(fh.to_device(lhs), fh.to_device(rhs), out, batch)
    _sync()
    got = fh.to_host(out)
    assert np.array_equal(got, om.multiply(lhs, rhs))
    for i in range(batch):
        g = got[i * 2 * len(sizes) * n:(i + 1) * 2 * len(sizes) * n]
        assert [int(x) for x in bfv.decrypt(g)] == [(a * b) % t for a, b in zip(vs[2 * i], vs[2 * i + 1])]


@pytest.mark.parametrize("n,L,level,batch", [(16, 2, 0, 3), (4096, 3, 1, 2), (16384, 6, 0, 2), (32768, 3, 0, 1)])
def test_rgsw_external_product_matches_oracle(fh, oracle, n, L, level, batch):
    """Ciphertext x RGSWCiphertext (fhe/src/bfv/rgsw_ciphertext.rs:123-160)."""
    print(f"rgsw n={n} L={L} level={level} batch={batch}", flush=True)
    t = 65537 if n > 16 else 1153
    bfv = _bfv(oracle, n, [60] * L if n > 16 else [62] * L, t=t, seed=9)
    rng = random.Random(n)
    v2 = [rng.randrange(t) for _ in range(n)]
    rgsw = bfv.rgsw_encrypt(v2, level=level)
    keys = [_gpu_key(fh, bfv, k, level, level)[0] for k in rgsw]
    vs = [[rng.randrange(t) for _ in range(n)] for _ in range(batch)]
    cts = np.concatenate([bfv.encrypt(v, level=level, seed=40 + i) for i, v in enumerate(vs)])
    out = fh.empty(cts.size)
    fh.rgsw_external_product(keys[0], keys[1], fh.to_device(cts), out, batch)
    _sync()
    got = fh.to_host(out)
    assert np.array_equal(got, bfv.rgsw_product(rgsw, cts))
    step = cts.size // batch
    for i, v in enumerate(vs):
        dec = [int(x) for x in bfv.decrypt(got[i * step:(i + 1) * step], level=level)]
        assert dec == [(a * b) % t for a, b in zip(v, v2)]
    with pytest.raises(fh.FheHipError):
        fh.rgsw_external_product(keys[0], keys[1], out, out, batch)


@pytest.mark.parametrize("n,L", [(16, 2), (4096, 3), (32768, 4)])
def test_ciphertext_sub_neg_mul_plain(fh, oracle, n, L):
    """Sub, Neg and Mul<&Plaintext> for &Ciphertext (fhe/src/bfv/ops/mod.rs:
    110-258) against the oracle's modular operations, limb by limb, and
    ct x poly_ntt(v2) decrypting to the slot products."""
    print(f"ciphertext sub/neg/mul_plain n={n} L={L}", flush=True)
    t = 65537 if n > 16 else 1153
    bfv = _bfv(oracle, n, [60] * L if n > 16 else [62] * L, t=t, seed=11)
    moduli = bfv.level_moduli(0)
    ctx = fh.Context(moduli, n)
    num, parts = 3, 2
    a = oracle.random_polys(moduli, n, num * parts, seed=1)
    b = oracle.random_polys(moduli, n, num * parts, seed=2)
    a[:n] = 0   # neg of zero stays zero
    pts = oracle.random_polys(moduli, n, num, seed=3)

    def limbwise(op, x, y=None, rows=num * parts * L):
        out = np.empty_like(x)
        for r in range(rows):
            q = moduli[r % L]
            sl = slice(r * n, (r + 1) * n)
            out[sl] = oracle.modulus_op(q, op, x[sl], None if y is None else y[sl])
        return out

    out = fh.empty(a.size)
    fh.ciphertext_sub(ctx, fh.to_device(a), fh.to_device(b), out, num * parts)
    _sync()
    assert np.array_equal(fh.to_host(out), limbwise("sub", a, b))
    fh.ciphertext_neg(ctx, fh.to_device(a), out, num * parts)
    _sync()
    assert np.array_equal(fh.to_host(out), limbwise("neg", a))
    for stride in (L * n, 0):
        fh.ciphertext_mul_plain(ctx, fh.to_device(a), parts, num, fh.to_device(pts), stride, out)
        _sync()
        per_ct = [pts[(i if stride else 0) * L * n:((i if stride else 0) + 1) * L * n] for i in range(num)]
        want = limbwise("mul", a, np.concatenate([np.tile(p, parts) for p in per_ct]))
        assert np.array_equal(fh.to_host(out), want), stride
    rng = random.Random(n)
    v1, v2 = [rng.randrange(t) for _ in range(n)], [rng.randrange(t) for _ in range(n)]
    ct = fh.to_device(bfv.encrypt(v1, seed=4))
    fh.ciphertext_mul_plain(ctx, ct, 2, 1, fh.to_device(bfv.encode(v2, scaled=False)), 0, ct)   # in place
    _sync()
    assert [int(x) for x in bfv.decrypt(fh.to_host(ct))] == [(x * y) % t for x, y in zip(v1, v2)]
