"""The PIR server path on the GPU, bit for bit against the oracle:
dot_product_scalar (fhe/src/bfv/ops/dot_product.rs:56-143), the column
products of a database (examples/mulpir.rs:170-179) and
EvaluationKey::expands (fhe/src/bfv/keys/evaluation_key.rs:153-193), plus the
decryption of the expanded ciphertexts."""
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# MulPIR's plaintext modulus and moduli (examples/mulpir.rs:35-37)
T_MULPIR = (1 << 20) + (1 << 19) + (1 << 17) + (1 << 16) + (1 << 14) + 1


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


DOT_CASES = [
    # n, moduli sizes, count, parts, ct padding, pt padding (words)
    (16, [62, 62], 17, 2, 0, 0),          # past the u128 accumulator's 16 terms
    (16, [62, 62], 1, 2, 0, 0),
    (1024, [62, 62, 62], 40, 3, 16, 0),
    (8192, [50, 55], 58, 2, 0, 8192),     # MulPIR at level 1, padded plaintext rows
    (4096, [36, 36, 37], 5, 1, 0, 0),
    (32768, [60] * 4, 3, 4, 0, 0),
    (1024, [62, 62], 9, 5, 0, 0),         # two part groups, the second partial
    (4096, [50, 55], 12, 16, 0, 0),       # eight queries' two parts
]


@pytest.mark.parametrize("n,sizes,count,parts,cpad,ppad", DOT_CASES)
def test_dot_product_scalar_matches_oracle(fh, oracle, n, sizes, count, parts, cpad, ppad):
    print(f"dot_product_scalar n={n} sizes={sizes} count={count} parts={parts}", flush=True)
    moduli = oracle.generate_moduli(sizes, n)
    L = len(moduli)
    cs, ps = parts * L * n + cpad, L * n + ppad
    cts = np.zeros(count * cs, np.uint64)
    pts = np.zeros(count * ps, np.uint64)
    rc = oracle.random_polys(moduli, n, count * parts, seed=count).reshape(count, parts * L * n)
    rp = oracle.random_polys(moduli, n, count, seed=count + 1).reshape(count, L * n)
    for i in range(count):
        cts[i * cs:i * cs + parts * L * n] = rc[i]
        pts[i * ps:i * ps + L * n] = rp[i]
    ctx = fh.Context(moduli, n)
    out = fh.empty(parts * L * n)
    fh.dot_product_scalar(ctx, fh.to_device(cts), cs, fh.to_device(pts), ps, count, parts, out)
    _sync()
    want = oracle.dot_product_scalar(oracle.Context(moduli, n), cts, pts, count, parts, cs, ps)
    assert np.array_equal(fh.to_host(out), want)


def test_dot_product_edge_values(fh, oracle):
    """Operands at q - 1 (the largest products) over many terms: the top word
    of the accumulators is in use."""
    n, count = 1024, 300
    moduli = oracle.generate_moduli([62, 62, 62], n)
    L = len(moduli)
    row = np.concatenate([np.full(n, q - 1, np.uint64) for q in moduli])
    cts = np.tile(row, 2 * count)
    pts = np.tile(row, count)
    ctx = fh.Context(moduli, n)
    out = fh.empty(2 * L * n)
    fh.dot_product_scalar(ctx, fh.to_device(cts), 2 * L * n, fh.to_device(pts), L * n, count, 2, out)
    _sync()
    assert np.array_equal(fh.to_host(out),
                          oracle.dot_product_scalar(oracle.Context(moduli, n), cts, pts, count, 2))
    # (q - 1)^2 = 1 mod q: every coefficient is count mod q
    assert all(int(x) == count for x in fh.to_host(out))


COLUMN_CASES = [
    # n, moduli sizes, rows, columns, parts
    (8192, [50, 55], 9, 7, 2),     # column count not a multiple of the column group
    (1024, [62, 62], 20, 5, 3),
    (4096, [36, 36, 37], 6, 1, 2),
    (8192, [50, 55], 58, 16, 2),   # MulPIR's first dimension
    (16, [62, 62], 3, 9, 2),       # blocks straddle limbs
    (8192, [50, 55], 11, 6, 8),    # four queries
    (1024, [62, 62], 5, 3, 7),     # ragged part groups
]


@pytest.mark.parametrize("n,sizes,rows,columns,parts", COLUMN_CASES)
def test_dot_product_columns_match_oracle(fh, oracle, n, sizes, rows, columns, parts):
    print(f"dot_product_columns n={n} rows={rows} columns={columns} parts={parts}", flush=True)
    moduli = oracle.generate_moduli(sizes, n)
    L = len(moduli)
    cts = oracle.random_polys(moduli, n, rows * parts, seed=rows)
    db = oracle.random_polys(moduli, n, rows * columns, seed=columns + 100)
    ctx = fh.Context(moduli, n)
    out = fh.empty(columns * parts * L * n)
    fh.dot_product_columns(ctx, fh.to_device(cts), rows, parts, fh.to_device(db), columns, out)
    _sync()
    got = fh.to_host(out).reshape(columns, parts * L * n)
    octx = oracle.Context(moduli, n)
    for j in range(columns):
        want = oracle.dot_product_scalar(octx, cts, db[j * L * n:], rows, parts, pt_stride=columns * L * n)
        assert np.array_equal(got[j], want), j


def test_dot_product_errors(fh, oracle):
    n = 1024
    moduli = oracle.generate_moduli([62, 62], n)
    ctx = fh.Context(moduli, n)
    buf = fh.empty(5 * 2 * n)
    with pytest.raises(fh.FheHipError):
        fh.dot_product_scalar(ctx, buf, 4 * n, buf, 2 * n, 0, 2, buf)   # empty
    with pytest.raises(fh.FheHipError):
        fh.dot_product_columns(ctx, buf, 0, 2, buf, 3, buf)


def _expansion_key(fh, bfv, levels, ct_level, key_level):
    n = bfv.degree
    okeys, gkeys = [], []
    cctx = fh.Context(bfv.level_moduli(ct_level), n)
    kctx = fh.Context(bfv.level_moduli(key_level), n)
    for l in range(levels):
        e = (n >> l) + 1
        k = bfv.galois_key(e, ct_level=ct_level, key_level=key_level, seed=80 + l)
        c0, c1 = k.export()
        gkeys.append(fh.GaloisKey(fh.KeySwitchingKey(cctx, kctx, c0, c1), e))
        okeys.append(k)
    return okeys, fh.EvaluationKey(gkeys)


EXPAND_CASES = [
    # n, moduli sizes, t, ct_level, key_level, sizes to expand to
    (16, [62, 62], 1153, 0, 0, [1, 2, 3, 8, 16]),
    (1024, [62, 62, 62], 65537, 1, 0, [5, 64]),
    (8192, [50, 55, 55], T_MULPIR, 1, 0, [115]),     # MulPIR: dim1 + dim2 = 58 + 57
    (16384, [60] * 7, 65537, 1, 0, [6]),
    (32768, [60] * 4, 65537, 0, 0, [3]),
]


@pytest.mark.parametrize("n,sizes,t,ct_level,key_level,targets", EXPAND_CASES)
def test_expand_matches_oracle(fh, oracle, n, sizes, t, ct_level, key_level, targets):
    print(f"expand n={n} sizes={sizes} levels={ct_level},{key_level} targets={targets}", flush=True)
    bfv = oracle.Bfv(n, t, oracle.generate_moduli(sizes, n), sk_seed=n + 3)
    levels = max(s - 1 for s in targets).bit_length()
    okeys, ek = _expansion_key(fh, bfv, levels, ct_level, key_level)
    assert ek.expansion_levels() == levels
    L = len(bfv.level_moduli(ct_level))
    rng = random.Random(n)
    for size in targets:
        v = [rng.randrange(t) for _ in range(size)]
        ct = bfv.encrypt(v, level=ct_level, seed=size, poly=True)
        out = fh.empty(size * 2 * L * n)
        ek.expand(fh.to_device(ct), size, out)
        _sync()
        got = fh.to_host(out)
        want = bfv.expands(okeys, ct, size, level=ct_level)
        assert np.array_equal(got, want), size
        lev = (size - 1).bit_length()
        g = got.reshape(size, 2 * L * n)
        for j in [0, size // 2, size - 1]:
            dec = [int(x) for x in bfv.decrypt(g[j], level=ct_level, poly=True)]
            assert dec == [v[j] * (1 << lev) % t] + [0] * (n - 1), (size, j)


def test_expand_errors(fh, oracle):
    n = 1024
    bfv = oracle.Bfv(n, 65537, oracle.generate_moduli([62, 62], n), sk_seed=9)
    okeys, ek = _expansion_key(fh, bfv, 2, 0, 0)
    assert ek.expansion_levels() == 2
    ct = fh.to_device(bfv.encrypt([1, 2], seed=1, poly=True))
    out = fh.empty(8 * 2 * 2 * n)
    with pytest.raises(fh.FheHipError):
        ek.expand(ct, 5, out)            # needs level 3
    with pytest.raises(fh.FheHipError):
        ek.expand(out, 2, out)           # aliasing
    ek.expand(ct, 0, out)                # size 0 gives the input, as in fhe.rs
    _sync()
    assert np.array_equal(fh.to_host(out)[:2 * 2 * n], fh.to_host(ct))


@pytest.mark.parametrize("n,num,size", [(1024, 3, 5), (8192, 4, 115), (16, 2, 16)])
def test_expand_batch_matches_single_expansions(fh, oracle, n, num, size):
    """fhehip_expand_batch of num queries == fhehip_expand of each (and the
    oracle for the first), output j of query q at [j][q]."""
    print(f"expand_batch n={n} num={num} size={size}", flush=True)
    t = 1153 if n == 16 else 65537
    bfv = oracle.Bfv(n, t, oracle.generate_moduli([50, 55, 55] if n == 8192 else [62, 62, 62], n), sk_seed=4)
    levels = (size - 1).bit_length()
    okeys, ek = _expansion_key(fh, bfv, levels, 1, 0)
    L = len(bfv.level_moduli(1))
    sz = 2 * L * n
    qs = np.concatenate([bfv.encrypt([q + 1, 2], level=1, seed=q, poly=True) for q in range(num)])
    out = fh.empty(size * num * sz)
    ek.expand_batch(fh.to_device(qs), num, size, out)
    _sync()
    got = fh.to_host(out).reshape(size, num, sz)
    for q in range(num):
        one = fh.empty(size * sz)
        ek.expand(fh.to_device(qs[q * sz:(q + 1) * sz]), size, one)
        _sync()
        assert np.array_equal(got[:, q, :].reshape(-1), fh.to_host(one)), q
    assert np.array_equal(got[:, 0, :].reshape(-1), bfv.expands(okeys, qs[:sz], size, level=1))


def _pir_setup(fh, oracle, n, dim1, dim2, t=None, seed=31):
    t = t or T_MULPIR
    bfv = oracle.Bfv(n, t, oracle.generate_moduli([50, 55, 55], n), sk_seed=seed)
    m1, m2 = bfv.level_moduli(1), bfv.level_moduli(2)
    level = (dim1 + dim2 - 1).bit_length()
    okeys, ek = _expansion_key(fh, bfv, level, 1, 0)
    ork = bfv.relin_key(ct_level=1, key_level=1, seed=seed + 1)
    c0, c1 = ork.export()
    ctx1, ctx2 = fh.Context(m1, n), fh.Context(m2, n)
    rk = fh.KeySwitchingKey(ctx1, ctx1, c0, c1)
    om = oracle.Multiplicator(bfv, None, level=1)
    mul = fh.Multiplicator(base_ctx=ctx1, extended_basis=om.ext_moduli(), post=(t, int(m1[0]) * int(m1[1])))
    sw = fh.Switcher(ctx1, ctx2)
    return dict(bfv=bfv, t=t, m1=m1, m2=m2, level=level, okeys=okeys, ek=ek, ork=ork, rk=rk, om=om, mul=mul, sw=sw,
                ctx1=ctx1, ctx2=ctx2)


def _pir_oracle_response(oracle, S, db, query, dim1, dim2):
    """The MulPIR server response (examples/mulpir.rs:151-185), step by step
    in the oracle."""
    bfv, m1, n = S["bfv"], S["m1"], S["bfv"].degree
    L = len(m1)
    o1 = oracle.Context(m1, n)
    ex = bfv.expands(S["okeys"], query, dim1 + dim2, level=1).reshape(dim1 + dim2, 2 * L * n)
    acc = None
    qv = np.array(m1, dtype=np.uint64).repeat(n)
    for c in range(dim2):
        dc = oracle.dot_product_scalar(o1, ex[:dim1].reshape(-1), db[c * L * n:], dim1, 2, pt_stride=dim2 * L * n)
        pc = S["om"].multiply(dc, ex[dim1 + c]).reshape(3, L * n)
        acc = pc if acc is None else np.where(acc + pc >= qv, acc + pc - qv, acc + pc)
    return oracle.ct_switch_down(o1, oracle.Context(S["m2"], n), bfv.relinearize(S["ork"], acc.reshape(-1)))


@pytest.mark.parametrize("n,dim1,dim2,num", [(8192, 4, 3, 2), (8192, 5, 1, 5), (16384, 3, 4, 1), (1024, 7, 9, 16)])
def test_pir_respond_end_to_end(fh, oracle, n, dim1, dim2, num):
    """A MulPIR round trip (examples/mulpir.rs:120-215) through
    fhehip_pir_respond: selection queries encrypted by the oracle, a database
    of poly-encoded plaintexts, and responses that decrypt to the selected
    plaintexts; every query's response equals the oracle's, bit for bit."""
    print(f"pir_respond MulPIR n={n} {dim1}x{dim2} database, {num} queries", flush=True)
    S = _pir_setup(fh, oracle, n, dim1, dim2)
    bfv, t, m1, m2 = S["bfv"], S["t"], S["m1"], S["m2"]
    rng = np.random.default_rng(n + num)
    vals = rng.integers(0, t, (dim1 * dim2, n)).astype(np.uint64)
    db = np.concatenate([bfv.encode(vals[k], level=1, poly=True, scaled=False) for k in range(dim1 * dim2)])
    inv = pow(1 << S["level"], -1, t)
    sel = [(int(rng.integers(dim1)), int(rng.integers(dim2))) for _ in range(num)]
    queries = []
    for q, (i, j) in enumerate(sel):
        pt = [0] * (dim1 + dim2)
        pt[i] = pt[dim1 + j] = inv
        queries.append(bfv.encrypt(pt, level=1, seed=10 + q, poly=True))
    out = fh.empty(num * 2 * len(m2) * n)
    fh.pir_respond(S["ek"], S["mul"], S["rk"], S["sw"], fh.to_device(db), dim1, dim2,
                   fh.to_device(np.concatenate(queries)), num, out)
    _sync()
    got = fh.to_host(out).reshape(num, -1)
    for q, (i, j) in enumerate(sel):
        dec = bfv.decrypt(got[q], level=2, poly=True)
        assert np.array_equal(dec, vals[i * dim2 + j]), q
        assert np.array_equal(got[q], _pir_oracle_response(oracle, S, db, queries[q], dim1, dim2)), q


def test_pir_respond_rejects_bad_handles_and_sizes(fh, oracle):
    """fhehip_pir_respond refuses handles of another context, a multiplicator
    that relinearizes or switches moduli, and batches whose products or
    switch-down exceed one launch, before any work."""
    n, dim1, dim2 = 1024, 2, 2
    S = _pir_setup(fh, oracle, n, dim1, dim2, seed=41)
    L1, L2 = len(S["m1"]), len(S["m2"])
    db = fh.empty(dim1 * dim2 * L1 * n)
    q = fh.empty(2 * 2 * L1 * n)
    out = fh.empty(2 * 2 * L2 * n)

    def respond(ek=S["ek"], mul=S["mul"], rk=S["rk"], sw=S["sw"], d1=dim1, d2=dim2, num=2, queries=q):
        fh.pir_respond(ek, mul, rk, sw, db, d1, d2, queries, num, out)

    # another context: a switcher from the level-2 context
    other_sw = fh.Switcher(S["ctx2"], S["ctx2"])
    with pytest.raises(fh.FheHipError) as e:
        respond(sw=other_sw)
    assert e.value.code == fh.ERR_CTX_MISMATCH
    # a relinearizing multiplicator
    relin_mul = fh.Multiplicator(rk=S["rk"], plaintext=S["t"])
    with pytest.raises(fh.FheHipError) as e:
        respond(mul=relin_mul)
    assert e.value.code == fh.ERR_INVALID_ARG
    # a mod-switching multiplicator
    ms = fh.Multiplicator(base_ctx=S["ctx1"], extended_basis=S["om"].ext_moduli(),
                          post=(S["t"], int(S["m1"][0]) * int(S["m1"][1])))
    ms.enable_mod_switching(S["ctx2"])
    with pytest.raises(fh.FheHipError) as e:
        respond(mul=ms)
    assert e.value.code == fh.ERR_INVALID_ARG
    # dim2 x queries above one launch, and more queries than one switch-down takes
    with pytest.raises(fh.FheHipError) as e:
        respond(d2=40000, num=2)
    assert e.value.code == fh.ERR_INVALID_ARG
    with pytest.raises(fh.FheHipError) as e:
        respond(d2=1, num=40000)
    assert e.value.code == fh.ERR_INVALID_ARG
    with pytest.raises(fh.FheHipError):
        respond(d1=0)
