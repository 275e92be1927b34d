"""GPU parity of EvaluationKey operations (rotations, inner sum) against the
oracle's Galois key switching composed with modular additions, and by
decryption (fhe/src/bfv/keys/evaluation_key.rs:56-131)."""
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


def _add(a, b, moduli, n):
    a = a.reshape(-1, len(moduli), n)
    b = b.reshape(-1, len(moduli), n)
    q = np.array(moduli, dtype=object).reshape(1, -1, 1)
    return ((a.astype(object) + b.astype(object)) % q).astype(np.uint64).reshape(-1)


def _setup(fh, oracle, n, L, ct_level, key_level, exps, t=65537):
    bfv = oracle.Bfv(n, t, oracle.generate_moduli([60] * L, n), sk_seed=n)
    cctx = fh.Context(bfv.level_moduli(ct_level), n)
    kctx = fh.Context(bfv.level_moduli(key_level), n)
    okeys, gkeys = {}, []
    for i, e in enumerate(exps):
        ok = bfv.galois_key(e, ct_level=ct_level, key_level=key_level, seed=100 + i)
        c0, c1 = ok.export()
        gkeys.append(fh.GaloisKey(fh.KeySwitchingKey(cctx, kctx, c0, c1), e))
        okeys[e] = ok
    return bfv, cctx, okeys, fh.EvaluationKey(gkeys)


@pytest.mark.parametrize("n,L,ct_level,key_level", [(16, 3, 1, 0), (1024, 4, 1, 0), (8192, 3, 1, 0)])
def test_inner_sum_matches_oracle_and_decrypts(fh, oracle, n, L, ct_level, key_level):
    print(f"inner sum n={n}", flush=True)
    t = 65537
    exps = [pow(3, 1 << k, 2 * n) for k in range((n // 2).bit_length() - 1)] + [2 * n - 1]
    bfv, cctx, okeys, ek = _setup(fh, oracle, n, L, ct_level, key_level, exps)
    rng = random.Random(n)
    v = [rng.randrange(t) for _ in range(n)]
    ct = bfv.encrypt(v, level=ct_level, seed=3)
    out = fh.empty(ct.size)
    ek.inner_sum(fh.to_device(ct), out, 1)
    _sync()
    got = fh.to_host(out)
    ref = ct.copy()
    for e in exps:
        ref = _add(ref, bfv.galois(okeys[e], ref), cctx.moduli, n)
    assert np.array_equal(got, ref)
    assert [int(x) for x in bfv.decrypt(got, level=ct_level)] == [sum(v) % t] * n


def test_rotate_columns_rows_and_partial_sum(fh, oracle):
    n, L, t = 64, 3, 65537
    exps = [pow(3, s, 2 * n) for s in (1, 2, 4, 8, 5)] + [2 * n - 1]
    bfv, cctx, okeys, ek = _setup(fh, oracle, n, L, 1, 0, exps)
    rng = random.Random(9)
    v = [rng.randrange(t) for _ in range(n)]
    ct = bfv.encrypt(v, level=1, seed=4)
    d = fh.to_device(ct)
    h = n // 2
    for step in (1, 5):
        out = fh.empty(ct.size)
        ek.rotate_columns(d, out, step, 1)
        _sync()
        got = fh.to_host(out)
        assert np.array_equal(got, bfv.galois(okeys[pow(3, step, 2 * n)], ct))
        assert [int(x) for x in bfv.decrypt(got, level=1)] == v[step:h] + v[:step] + v[h + step:] + v[h:h + step]
    out = fh.empty(ct.size)
    ek.rotate_rows(d, out, 1)
    _sync()
    assert [int(x) for x in bfv.decrypt(fh.to_host(out), level=1)] == v[h:] + v[:h]
    # sum over 8 consecutive slots of each row
    out = fh.empty(ct.size)
    ek.rotate_and_sum(d, out, 3, False, 1)
    _sync()
    dec = [int(x) for x in bfv.decrypt(fh.to_host(out), level=1)]
    for r in range(2):
        row = v[r * h:(r + 1) * h]
        for i in range(h):
            assert dec[r * h + i] == sum(row[(i + m) % h] for m in range(8)) % t
    # unsupported step fails cleanly
    with pytest.raises(fh.FheHipError):
        ek.rotate_columns(d, out, 3, 1)


def test_batched_rotation_equals_single(fh, oracle):
    n, L, batch = 4096, 4, 4
    exps = [pow(3, 1, 2 * n)]
    bfv, cctx, okeys, ek = _setup(fh, oracle, n, L, 1, 0, exps)
    cts = oracle.random_polys(cctx.moduli, n, 2 * batch, seed=2)
    d = fh.to_device(cts)
    out = fh.empty(cts.size)
    ek.rotate_columns(d, out, 1, batch)
    _sync()
    allb = fh.to_host(out)
    sz = 2 * cctx.L * n
    for i in range(batch):
        one = fh.empty(sz)
        ek.rotate_columns(d[i * sz:(i + 1) * sz], one, 1, 1)
        _sync()
        assert np.array_equal(fh.to_host(one), allb[i * sz:(i + 1) * sz])
        assert np.array_equal(allb[i * sz:(i + 1) * sz], bfv.galois(okeys[exps[0]], cts[i * sz:(i + 1) * sz]))


def test_ciphertext_add(fh, oracle):
    n = 1024
    moduli = oracle.generate_moduli([62, 60, 50], n)
    ctx = fh.Context(moduli, n)
    a = oracle.random_polys(moduli, n, 4, seed=1)
    b = oracle.random_polys(moduli, n, 4, seed=2)
    da, db = fh.to_device(a), fh.to_device(b)
    fh.ciphertext_add(ctx, da, db, da, 4)
    _sync()
    assert np.array_equal(fh.to_host(da), _add(a, b, moduli, n))
