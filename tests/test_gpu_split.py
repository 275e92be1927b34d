"""GPU parity of both ways a row is transformed at degrees 4096 .. 16384: on
one workgroup, and as two halves on two workgroups (the small-launch path,
fhehip_split_row_limit). Multiplication, a Galois rotation, the forward and
inverse NTT and the scaler are compared with the oracle under each."""
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


@pytest.fixture
def restore_limit(fh):
    prev = fh.split_row_limit(128)
    fh.split_row_limit(prev)
    yield
    fh.split_row_limit(prev)


def _sync():
    import torch
    torch.cuda.synchronize()


@pytest.mark.parametrize("n", [4096, 8192, 16384])
@pytest.mark.parametrize("limit", [0, 1 << 20])
def test_whole_and_half_rows_match_oracle(fh, oracle, restore_limit, n, limit):
    print(f"row split limit {limit} n={n}", flush=True)
    fh.split_row_limit(limit)
    bfv = oracle.Bfv(n, T, oracle.generate_moduli([60] * 4, n), sk_seed=n % 37)
    moduli = bfv.level_moduli(1)
    L = len(moduli)
    cctx, kctx = fh.Context(moduli, n), fh.Context(bfv.level_moduli(0), n)
    rk = bfv.relin_key(ct_level=1, key_level=0, seed=3)
    c0, c1 = rk.export()
    mult = fh.Multiplicator(rk=fh.KeySwitchingKey(cctx, kctx, c0, c1), plaintext=T)
    om = oracle.Multiplicator(bfv, rk, level=1)
    batch = 3
    rng = random.Random(n + limit)
    vs = [[rng.randrange(T) for _ in range(n)] for _ in range(2 * batch)]
    lhs = np.concatenate([bfv.encrypt(vs[2 * i], level=1, seed=10 + i) for i in range(batch)])
    rhs = np.concatenate([bfv.encrypt(vs[2 * i + 1], level=1, seed=20 + i) for i in range(batch)])
    out = fh.empty(lhs.size)
    mult.multiply(fh.to_device(lhs), fh.to_device(rhs), out, batch)
    e = pow(3, 2, 2 * n)
    gk = bfv.galois_key(e, ct_level=1, key_level=0, seed=4)
    g0, g1 = gk.export()
    g = fh.GaloisKey(fh.KeySwitchingKey(cctx, kctx, g0, g1), e)
    w = 2 * L * n
    rot = fh.empty(w)
    g.relinearize(out[:w], rot, 1)
    # NTT round trip and forward against the oracle
    polys = oracle.random_polys(moduli, n, 2, seed=7)
    d = fh.to_device(polys)
    cctx.ntt_forward(d, 2)
    _sync()
    fwd = fh.to_host(d)
    cctx.ntt_backward(d, 2)
    # scaler to the key context (extension) in the NTT representation
    sc = fh.Scaler(cctx, kctx)
    sout = fh.empty(2 * (L + 1) * n)
    sc.scale(fh.to_device(polys), sout, 2, True)
    _sync()
    got = fh.to_host(out)
    assert np.array_equal(got, om.multiply(lhs, rhs))
    for i in range(batch):
        dec = [int(x) for x in bfv.decrypt(got[i * w:(i + 1) * w], level=1)]
        assert dec == [(a * b) % T for a, b in zip(vs[2 * i], vs[2 * i + 1])]
    assert np.array_equal(fh.to_host(rot), bfv.galois(gk, got[:w]))
    oc = oracle.Context(moduli, n)
    assert np.array_equal(fwd, oc.forward(polys))
    assert np.array_equal(fh.to_host(d), polys)
    assert np.array_equal(fh.to_host(sout), oracle.PolyScaler(oc, oracle.Context(bfv.level_moduli(0), n)).scale(polys, True))


@pytest.mark.parametrize("polys", [1, 17])
def test_scaler_launch_forms(fh, oracle, polys):
    """Small launches (at most 1024 workgroups) stage the scaler's constants
    in LDS, larger ones read them through the scalar cache: both equal the
    oracle (one and 17 polynomials of degree 16384, 6 -> 13 moduli and the
    t/Q scaling back)."""
    n = 16384
    print(f"scaler launch forms polys={polys}", flush=True)
    qs = oracle.generate_moduli([60] * 6, n)
    ms = qs + [m for m in oracle.generate_moduli([62] * 8, n) if m not in qs][:7]
    cq, cm = fh.Context(qs, n), fh.Context(ms, n)
    oq, om = oracle.Context(qs, n), oracle.Context(ms, n)
    a = oracle.random_polys(qs, n, polys, seed=polys)
    up = fh.Scaler(cq, cm)
    out = fh.empty(polys * len(ms) * n)
    up.scale(fh.to_device(a), out, polys, False)
    b = oracle.random_polys(ms, n, polys, seed=polys + 1)
    Q = 1
    for q in qs:
        Q *= int(q)
    down = fh.Scaler(cm, cq, T, Q)
    out2 = fh.empty(polys * len(qs) * n)
    down.scale(fh.to_device(b), out2, polys, False)
    _sync()
    assert np.array_equal(fh.to_host(out), oracle.PolyScaler(oq, om).scale(a, False))
    assert np.array_equal(fh.to_host(out2), oracle.PolyScaler(om, oq, T, Q).scale(b, False))
