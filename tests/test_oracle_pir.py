"""Pins of the oracle's PIR server path: dot_product_scalar
(fhe/src/bfv/ops/dot_product.rs:56-143, rq/ops.rs:448-550) against Python
integers, and EvaluationKey::expands (evaluation_key.rs:153-193) against the
reference's own property (evaluation_key.rs:680-725): expansion to 2^i
ciphertexts of a poly-encoded v decrypts, for each j, to v_j 2^i at
coefficient 0 and zero elsewhere."""
import random

import numpy as np
import pytest


def _py_dot(moduli, n, cts, pts, count, parts, ct_stride, pt_stride):
    L = len(moduli)
    out = np.zeros(parts * L * n, np.uint64)
    for p in range(parts):
        for m, q in enumerate(moduli):
            for e in range(n):
                s = 0
                for i in range(count):
                    s += int(cts[i * ct_stride + (p * L + m) * n + e]) * int(pts[i * pt_stride + m * n + e])
                out[(p * L + m) * n + e] = s % q
    return out


@pytest.mark.parametrize("sizes,count,parts", [
    ([62, 62], 1, 2), ([62, 62], 15, 2), ([62, 62], 16, 2), ([62, 62], 17, 2),   # u128 path / rq::dot_product path
    ([62, 62], 40, 3), ([36, 36, 37], 33, 2), ([50, 55, 55], 127, 2), ([62], 64, 2)])
def test_dot_product_scalar_vs_python(oracle, sizes, count, parts):
    n = 16
    moduli = oracle.generate_moduli(sizes, n)
    ctx = oracle.Context(moduli, n)
    L = len(moduli)
    cts = oracle.random_polys(moduli, n, count * parts, seed=count)
    pts = oracle.random_polys(moduli, n, count, seed=count + 1000)
    got = oracle.dot_product_scalar(ctx, cts, pts, count, parts)
    assert np.array_equal(got, _py_dot(moduli, n, cts, pts, count, parts, parts * L * n, L * n))


def test_dot_product_scalar_edges(oracle):
    n = 16
    moduli = oracle.generate_moduli([62, 62], n)
    ctx = oracle.Context(moduli, n)
    L = len(moduli)
    # operands at q - 1: the largest products, past the reduction point
    cts = np.array([(q - 1) for q in moduli for _ in range(n)] * 2 * 20, np.uint64)
    pts = np.array([(q - 1) for q in moduli for _ in range(n)] * 20, np.uint64)
    got = oracle.dot_product_scalar(ctx, cts, pts, 20, 2)
    assert np.array_equal(got, _py_dot(moduli, n, cts, pts, 20, 2, 2 * L * n, L * n))
    with pytest.raises(RuntimeError):
        oracle.dot_product_scalar(ctx, cts, pts, 0, 2)


def test_dot_product_strided_columns(oracle):
    """The PIR column product: database [rows][columns][L][N], column c taken
    with `skip(c).step_by(columns)` (mulpir.rs:170-174)."""
    n, rows, columns = 16, 9, 5
    moduli = oracle.generate_moduli([50, 55, 55], n)
    ctx = oracle.Context(moduli, n)
    L = len(moduli)
    cts = oracle.random_polys(moduli, n, rows * 2, seed=3)
    db = oracle.random_polys(moduli, n, rows * columns, seed=4)
    for c in range(columns):
        got = oracle.dot_product_scalar(ctx, cts, db[c * L * n:], rows, 2, pt_stride=columns * L * n)
        col = np.concatenate([db[(r * columns + c) * L * n:(r * columns + c + 1) * L * n] for r in range(rows)])
        assert np.array_equal(got, oracle.dot_product_scalar(ctx, cts, col, rows, 2))


def _expansion_keys(bfv, levels, ct_level, key_level):
    n = bfv.degree
    return [bfv.galois_key((n >> l) + 1, ct_level=ct_level, key_level=key_level, seed=70 + l) for l in range(levels)]


@pytest.mark.parametrize("n,sizes,ct_level,key_level", [(16, [62, 62], 0, 0), (16, [62, 62, 62], 1, 0),
                                                        (16, [62, 62, 62], 2, 1), (32, [50, 55, 55], 1, 1)])
def test_expands_decrypts(oracle, n, sizes, ct_level, key_level):
    t = 1153
    bfv = oracle.Bfv(n, t, oracle.generate_moduli(sizes, n), sk_seed=31)
    rng = random.Random(n + ct_level)
    logn = n.bit_length() - 1
    gks = _expansion_keys(bfv, logn, ct_level, key_level)
    L = len(bfv.level_moduli(ct_level))
    for i in range(0, logn + 1):
        for size in sorted({1 << i, (1 << i) - 1 if i > 1 else 1 << i}):
            v = [rng.randrange(t) for _ in range(1 << i)]
            ct = bfv.encrypt(v, level=ct_level, seed=40 + i, poly=True)
            out = bfv.expands(gks[:i], ct, size, level=ct_level).reshape(size, 2 * L * n)
            for j in range(size):
                dec = [int(x) for x in bfv.decrypt(out[j], level=ct_level, poly=True)]
                assert dec == [v[j] * (1 << i) % t] + [0] * (n - 1), (i, size, j)


def test_expands_errors_and_size_one(oracle):
    n = 16
    bfv = oracle.Bfv(n, 1153, oracle.generate_moduli([62, 62], n), sk_seed=5)
    gks = _expansion_keys(bfv, 2, 0, 0)
    ct = bfv.encrypt([3, 4, 5], seed=2, poly=True)
    assert np.array_equal(bfv.expands([], ct, 1), ct)
    with pytest.raises(RuntimeError):
        bfv.expands(gks, ct, 8)   # needs the key of level 2
