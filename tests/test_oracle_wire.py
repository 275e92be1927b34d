"""Pins of the oracle's Rq wire format (math/src/rq/convert.rs:15-87 over
fhe-util/src/lib.rs:60-133 transcode): each limb's coefficients packed least
significant bit first with the width of q_i - 1, checked against a Python
big-integer packing (sum of v_k 2^(k nbits), little-endian bytes), round
trips in both representations, and the lengths of
Modulus::serialization_length."""
import numpy as np
import pytest


def _py_pack(values, nbits):
    x = 0
    for k, v in enumerate(values):
        x |= int(v) << (k * nbits)
    return np.frombuffer(x.to_bytes(len(values) * nbits // 8, "little"), np.uint8)


@pytest.mark.parametrize("n,primes,sizes", [(8, [17], None), (16, [12289], None), (16, None, [62, 62, 62]),
                                            (64, None, [36, 36, 37]), (1024, None, [50, 55, 55])])
def test_serialize_matches_python_packing(oracle, n, primes, sizes):
    moduli = primes or [int(q) for q in oracle.generate_moduli(sizes, n)]
    ctx = oracle.Context(moduli, n)
    L = len(moduli)
    polys = oracle.random_polys(moduli, n, 2, seed=n)
    polys[:n] = [q - 1 for q in [moduli[0]] * n]   # the widest value of limb 0
    data = oracle.poly_serialize(ctx, polys, 2, ntt=False)
    assert data.size == 2 * oracle.serialization_length(moduli, n)
    want = np.concatenate([_py_pack(polys[(k * L + i) * n:(k * L + i + 1) * n], (moduli[i] - 1).bit_length())
                           for k in range(2) for i in range(L)])
    assert np.array_equal(data, want)
    assert np.array_equal(oracle.poly_deserialize(ctx, data, 2, ntt=False), polys)


def test_ntt_polynomials_serialize_in_power_basis(oracle):
    n = 64
    moduli = oracle.generate_moduli([62, 62], n)
    ctx = oracle.Context(moduli, n)
    pb = oracle.random_polys(moduli, n, 1, seed=3)
    ntt = ctx.forward(pb)
    data = oracle.poly_serialize(ctx, ntt, 1, ntt=True)
    assert np.array_equal(data, oracle.poly_serialize(ctx, pb, 1, ntt=False))
    assert np.array_equal(oracle.poly_deserialize(ctx, data, 1, ntt=True), ntt)
