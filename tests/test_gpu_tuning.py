"""GPU tests of the multiplication autotuner's launch shapes: every choice
of every knob (XCD-aware workgroup order, instances per key load in the key
inner product, values per thread of the fused tensor, lift and finish
kernels, half-row transforms for launches of any size, the scalers' column
sums on the matrix cores or the VALU) gives results
bit-exact to the oracle, and the autotuner returns a
valid choice with its timings."""
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

T = 65537
KNOBS = [("xcd_order", 0), ("xcd_order", 1), ("ks_mac_instances", 2), ("ks_mac_instances", 4),
         ("ks_mac_instances", 8), ("tensor_log2_values", 5), ("lift_log2_values", 5), ("finish_log2_values", 5),
         ("halves", 0), ("halves", 1), ("halves", 2), ("halves", 4), ("halves", 7), ("scaler_mfma", 0),
         ("scaler_mfma", 1)]
DEFAULT = {"xcd_order": 1, "ks_mac_instances": 4, "tensor_log2_values": 4, "lift_log2_values": 4,
           "finish_log2_values": 4, "halves": 1, "scaler_mfma": 1}


@pytest.fixture(scope="module")
def fh():
    import torch
    import fhehip
    assert torch.cuda.is_available()
    fhehip.lib()
    return fhehip


@pytest.fixture
def restore(fh):
    yield
    for n in (1024, 4096, 16384, 32768):
        fh.tuning_set(n, **DEFAULT)


def _sync():
    import torch
    torch.cuda.synchronize()


def _setup(fh, oracle, n, sizes, seed):
    bfv = oracle.Bfv(n, T, oracle.generate_moduli(sizes, n), sk_seed=seed)
    moduli = bfv.level_moduli(1)
    cctx, kctx = fh.Context(moduli, n), fh.Context(bfv.level_moduli(0), n)
    rk = bfv.relin_key(ct_level=1, key_level=0, seed=seed + 1)
    c0, c1 = rk.export()
    mult = fh.Multiplicator(rk=fh.KeySwitchingKey(cctx, kctx, c0, c1), plaintext=T)
    e = pow(3, 1, 2 * n)
    gk = bfv.galois_key(e, ct_level=1, key_level=0, seed=seed + 2)
    c0, c1 = gk.export()
    g = fh.GaloisKey(fh.KeySwitchingKey(cctx, kctx, c0, c1), e)
    return bfv, moduli, mult, oracle.Multiplicator(bfv, rk, level=1), gk, g


@pytest.mark.parametrize("n,sizes", [(4096, [60] * 4), (16384, [60] * 7), (32768, [60] * 3)])
def test_every_choice_matches_oracle(fh, oracle, restore, n, sizes):
    print(f"tuning choices n={n}", flush=True)
    bfv, moduli, mult, om, gk, g = _setup(fh, oracle, n, sizes, seed=n % 31)
    batch = 5   # not a multiple of 2, 4 or 8: every ks_mac grouping has a partial group
    w = 2 * len(moduli) * n
    lhs = oracle.random_polys(moduli, n, 2 * batch, seed=1)
    rhs = oracle.random_polys(moduli, n, 2 * batch, seed=2)
    want = om.multiply(lhs, rhs)
    want_rot = bfv.galois(gk, want[:w])
    dl, dr = fh.to_device(lhs), fh.to_device(rhs)
    # no small-launch half rows: the halves knob alone decides
    prev = fh.split_row_limit(0)
    try:
        _each_choice(fh, n, mult, g, dl, dr, batch, w, want, want_rot)
    finally:
        fh.split_row_limit(prev)


def _each_choice(fh, n, mult, g, dl, dr, batch, w, want, want_rot):
    for knob, value in KNOBS:
        fh.tuning_set(n, **dict(DEFAULT, **{knob: value}))
        out = fh.empty(batch * w)
        mult.multiply(dl, dr, out, batch)
        rot = fh.empty(w)
        g.relinearize(out[:w], rot, 1)
        _sync()
        assert np.array_equal(fh.to_host(out), want), (knob, value)
        assert np.array_equal(fh.to_host(rot), want_rot), (knob, value)


def test_autotune_returns_a_valid_choice(fh, oracle, restore):
    n = 4096
    bfv, moduli, mult, om, gk, g = _setup(fh, oracle, n, [60] * 4, seed=3)
    chosen, ms = fh.multiply_autotune(mult, 16)
    assert chosen["xcd_order"] in (0, 1) and chosen["ks_mac_instances"] in (2, 4, 8)
    assert all(chosen[k] in (4, 5) for k in ("tensor_log2_values", "lift_log2_values", "finish_log2_values"))
    assert 0 <= chosen["halves"] <= 7 and chosen["scaler_mfma"] in (0, 1)
    assert all(v > 0 for v in ms.values()), ms
    assert fh.tuning_get(n) == chosen
    lhs = oracle.random_polys(moduli, n, 6, seed=5)
    rhs = oracle.random_polys(moduli, n, 6, seed=6)
    out = fh.empty(lhs.size)
    mult.multiply(fh.to_device(lhs), fh.to_device(rhs), out, 3)
    _sync()
    assert np.array_equal(fh.to_host(out), om.multiply(lhs, rhs))
    with pytest.raises(fh.FheHipError):
        fh.tuning_set(n, ks_mac_instances=3)
    with pytest.raises(fh.FheHipError):
        fh.tuning_set(1000, xcd_order=0)
