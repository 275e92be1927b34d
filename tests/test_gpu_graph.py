"""GPU tests of HIP-graph capture (fhehip_multiply_graph, fhehip_multiply_sum_graph):
every replay equals the eager calls bit for bit, also with eager work that grows
other streams' scratch between replays, and the first replay equals the oracle."""
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


def _sync():
    import torch
    torch.cuda.synchronize()


def _setup(fh, oracle, n, sizes, log2_slots, seed):
    bfv = oracle.Bfv(n, T, oracle.generate_moduli(sizes, n), sk_seed=seed)
    moduli = bfv.level_moduli(1)
    cctx, kctx = fh.Context(moduli, n), fh.Context(bfv.level_moduli(0), n)
    rk = bfv.relin_key(ct_level=1, key_level=0, seed=seed + 1)
    c0, c1 = rk.export()
    mult = fh.Multiplicator(rk=fh.KeySwitchingKey(cctx, kctx, c0, c1), plaintext=T)
    gks = []
    for k in range(log2_slots):
        e = pow(3, 1 << k, 2 * n)
        c0, c1 = bfv.galois_key(e, ct_level=1, key_level=0, seed=seed + 10 + k).export()
        gks.append(fh.GaloisKey(fh.KeySwitchingKey(cctx, kctx, c0, c1), e))
    ek = fh.EvaluationKey(gks) if gks else None
    om = oracle.Multiplicator(bfv, rk, level=1)
    return bfv, moduli, mult, ek, om, (cctx, kctx)


@pytest.mark.parametrize("n,sizes,num", [(1024, [60] * 4, 3), (32768, [60] * 3, 1)])
def test_multiply_graph_replays(fh, oracle, n, sizes, num):
    import torch
    print(f"multiply graph n={n}", flush=True)
    bfv, moduli, mult, ek, om, ctxs = _setup(fh, oracle, n, sizes, 0, seed=n)
    L = len(moduli)
    w = 2 * L * n
    a = oracle.random_polys(moduli, n, 2 * num, seed=1)
    b = oracle.random_polys(moduli, n, 2 * num, seed=2)
    da, db = fh.to_device(a), fh.to_device(b)
    eager = fh.empty(num * w)
    mult.multiply(da, db, eager, num)
    _sync()
    ref = fh.to_host(eager)
    assert np.array_equal(ref, om.multiply(a, b))
    out = fh.empty(num * w)
    _sync()
    g = fh.Graph(mult, da, db, out, num)
    # a bigger eager batch on the default stream between replays
    big = 4 * num
    ba = fh.to_device(oracle.random_polys(moduli, n, 2 * big, seed=3))
    bb = fh.to_device(oracle.random_polys(moduli, n, 2 * big, seed=4))
    bo = fh.empty(big * w)
    for rep in range(3):
        out.zero_()
        _sync()
        g.launch()
        _sync()
        assert np.array_equal(fh.to_host(out), ref), rep
        mult.multiply(ba, bb, bo, big)
        _sync()
    g.close()


def test_multiply_sum_graph_matches_eager(fh, oracle):
    n, log2_slots, num = 2048, 3, 2
    bfv, moduli, mult, ek, om, ctxs = _setup(fh, oracle, n, [55, 60, 62], log2_slots, seed=7)
    L = len(moduli)
    w = 2 * L * n
    da = fh.to_device(oracle.random_polys(moduli, n, 2 * num, seed=5))
    db = fh.to_device(oracle.random_polys(moduli, n, 2 * num, seed=6))
    prod, eager = fh.empty(num * w), fh.empty(num * w)
    mult.multiply(da, db, prod, num)
    ek.rotate_and_sum(prod, eager, log2_slots, False, num)
    _sync()
    ref = fh.to_host(eager)
    out = fh.empty(num * w)
    _sync()
    g = fh.Graph(mult, da, db, out, num, ek=ek, log2_slots=log2_slots)
    for rep in range(3):
        out.zero_()
        _sync()
        g.launch()
        _sync()
        assert np.array_equal(fh.to_host(out), ref), rep
    g.close()


def test_graph_rejects_missing_rotation(fh, oracle):
    bfv, moduli, mult, ek, om, ctxs = _setup(fh, oracle, 1024, [60] * 3, 1, seed=9)
    d = fh.empty(2 * len(moduli) * 1024)
    with pytest.raises(fh.FheHipError):
        fh.Graph(mult, d, d, fh.empty(d.numel()), 1, ek=ek, log2_slots=2)


def test_capture_while_server_runs(fh, oracle):
    """A graph captured while the request server's threads allocate scratch
    and synchronize (its first batches) leaves both intact: every served
    request equals its direct calls, and the graph's replays the eager
    product."""
    n, log2_slots = 1024, 2
    bfv, moduli, mult, ek, om, ctxs = _setup(fh, oracle, n, [60] * 4, log2_slots, seed=11)
    L = len(moduli)
    w = 2 * L * n
    count = 48
    a = fh.to_device(oracle.random_polys(moduli, n, 2 * count, seed=21))
    b = fh.to_device(oracle.random_polys(moduli, n, 2 * count, seed=22))
    served = fh.empty(count * w)
    _sync()
    srv = fh.Server(max_batch=4, lanes=3)
    pid = srv.add_params(mult, ek, log2_slots)[0]
    for k in range(count):
        srv.submit(pid, a[k * w:(k + 1) * w], b[k * w:(k + 1) * w], served[k * w:(k + 1) * w])
    gout = fh.empty(2 * w)
    g = fh.Graph(mult, a[:2 * w], b[:2 * w], gout, 2)   # captured while the server works
    srv.drain(120)
    st = srv.stats()
    srv.close()
    assert st["completed"] == count and st["failed"] == 0
    eager = fh.empty(2 * w)
    mult.multiply(a[:2 * w], b[:2 * w], eager, 2)
    prod, ref = fh.empty(count * w), fh.empty(count * w)
    mult.multiply(a, b, prod, count)
    ek.rotate_and_sum(prod, ref, log2_slots, False, count)
    _sync()
    assert np.array_equal(fh.to_host(served), fh.to_host(ref))
    for rep in range(2):
        gout.zero_()
        _sync()
        g.launch()
        _sync()
        assert np.array_equal(fh.to_host(gout), fh.to_host(eager)), rep
    g.close()
