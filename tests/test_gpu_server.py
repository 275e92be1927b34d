"""GPU tests of the request server (fhehip_server_*): every request's result is
the same alone and in a mix of concurrently arriving requests of different
parameter sets, bit for bit; it equals the oracle's multiply + relinearize
followed by the slot sum; admission respects the device-memory budget."""
import time

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


class Kind:
    """One request kind: keys on the GPU and in the oracle."""

    def __init__(self, fh, oracle, n, sizes, log2_slots, seed):
        self.n, self.log2_slots = n, log2_slots
        self.bfv = oracle.Bfv(n, T, oracle.generate_moduli(sizes, n), sk_seed=seed)
        self.moduli = self.bfv.level_moduli(1)
        self.L = len(self.moduli)
        cctx = fh.Context(self.moduli, n)
        kctx = fh.Context(self.bfv.level_moduli(0), n)
        self.rk = self.bfv.relin_key(ct_level=1, key_level=0, seed=seed + 1)
        c0, c1 = self.rk.export()
        self.mult = fh.Multiplicator(rk=fh.KeySwitchingKey(cctx, kctx, c0, c1), plaintext=T)
        self.om = oracle.Multiplicator(self.bfv, self.rk, level=1)
        self.gk, gkeys = {}, []
        for k in range(log2_slots):
            e = pow(3, 1 << k, 2 * n)
            ok = self.bfv.galois_key(e, ct_level=1, key_level=0, seed=seed + 10 + k)
            c0, c1 = ok.export()
            gkeys.append(fh.GaloisKey(fh.KeySwitchingKey(cctx, kctx, c0, c1), e))
            self.gk[e] = ok
        self.ek = fh.EvaluationKey(gkeys)

    @property
    def words(self):
        return 2 * self.L * self.n

    def oracle_result(self, a, b):
        ct = self.om.multiply(a, b)
        q = np.array(self.moduli, dtype=object).reshape(1, -1, 1)
        for k in range(self.log2_slots):
            e = pow(3, 1 << k, 2 * self.n)
            r = self.bfv.galois(self.gk[e], ct, e)
            s = (ct.reshape(2, self.L, self.n).astype(object) + r.reshape(2, self.L, self.n).astype(object)) % q
            ct = s.astype(np.uint64).reshape(-1)
        return ct


def _requests(fh, oracle, kind, count, seed):
    out = []
    for k in range(count):
        a = oracle.random_polys(kind.moduli, kind.n, 2, seed=seed + k)
        b = oracle.random_polys(kind.moduli, kind.n, 2, seed=seed + 1000 + k)
        out.append((a, b, fh.to_device(a), fh.to_device(b)))
    _sync()  # the server's streams are not ordered after torch's
    return out


def test_results_same_alone_and_in_mix(fh, oracle):
    kinds = [Kind(fh, oracle, 1024, [60] * 4, 4, seed=1), Kind(fh, oracle, 2048, [55, 60, 62], 3, seed=2)]
    reqs = {0: _requests(fh, oracle, kinds[0], 12, 10), 1: _requests(fh, oracle, kinds[1], 12, 20)}
    # the mix: interleaved arrivals, some while earlier batches run
    srv = fh.Server(max_batch=4, lanes=3)
    ids = [srv.add_params(k.mult, k.ek, k.log2_slots)[0] for k in kinds]
    outs, tickets = [], []
    for i in range(12):
        for p in (0, 1):
            o = fh.empty(kinds[p].words)
            outs.append((p, i, o))
            tickets.append(srv.submit(ids[p], reqs[p][i][2], reqs[p][i][3], o))
        if i % 3 == 2:
            time.sleep(0.002)
    print("mix submitted", flush=True)
    srv.drain(120)
    st = srv.stats()
    assert st["completed"] == 24 and st["failed"] == 0
    for t in tickets:
        state, lat = srv.query(t)
        assert state == 2 and lat > 0
    srv.close()
    # each request alone, on a fresh server
    for p, i, o in outs:
        alone = fh.Server(max_batch=4, lanes=1)
        pid = alone.add_params(kinds[p].mult, kinds[p].ek, kinds[p].log2_slots)[0]
        o1 = fh.empty(kinds[p].words)
        alone.submit(pid, reqs[p][i][2], reqs[p][i][3], o1)
        alone.drain(60)
        alone.close()
        assert np.array_equal(fh.to_host(o1), fh.to_host(o)), (p, i)
        if i < 2:
            assert np.array_equal(fh.to_host(o), kinds[p].oracle_result(reqs[p][i][0], reqs[p][i][1])), (p, i)


def test_admission_respects_budget(fh, oracle):
    kind = Kind(fh, oracle, 1024, [60] * 3, 2, seed=3)
    probe = fh.Server(max_batch=8, lanes=2)
    _, fp = probe.add_params(kind.mult, kind.ek, kind.log2_slots)
    probe.close()
    srv = fh.Server(hbm_budget_bytes=int(2.5 * fp), max_batch=8, lanes=3)
    pid, _ = srv.add_params(kind.mult, kind.ek, kind.log2_slots)
    rq = _requests(fh, oracle, kind, 10, 30)
    outs = []
    for a, b, da, db in rq:
        o = fh.empty(kind.words)
        outs.append(o)
        srv.submit(pid, da, db, o)
    srv.drain(120)
    st = srv.stats()
    assert st["completed"] == 10
    assert st["peak_bytes"] <= int(2.5 * fp)
    assert st["max_batch_seen"] <= 2
    srv.close()
    assert np.array_equal(fh.to_host(outs[0]), kind.oracle_result(rq[0][0], rq[0][1]))


def test_rejects_oversized_and_unknown_requests(fh, oracle):
    kind = Kind(fh, oracle, 1024, [60] * 3, 1, seed=4)
    srv = fh.Server(hbm_budget_bytes=1024, max_batch=2, lanes=1)
    pid, fp = srv.add_params(kind.mult, kind.ek, kind.log2_slots)
    assert fp > 1024
    a = fh.empty(kind.words)
    with pytest.raises(fh.FheHipError) as e:
        srv.submit(pid, a, a, fh.empty(kind.words))
    assert e.value.code == fh.ERR_OOM
    with pytest.raises(fh.FheHipError):
        srv.submit(pid + 1, a, a, fh.empty(kind.words))
    srv.close()


def test_degree_32768_requests(fh, oracle):
    kind = Kind(fh, oracle, 32768, [60] * 3, 2, seed=5)
    small = Kind(fh, oracle, 1024, [60] * 3, 2, seed=6)
    srv = fh.Server(max_batch=4, lanes=2)
    pb = srv.add_params(kind.mult, kind.ek, kind.log2_slots)[0]
    ps = srv.add_params(small.mult, small.ek, small.log2_slots)[0]
    rb = _requests(fh, oracle, kind, 2, 40)
    rs = _requests(fh, oracle, small, 2, 50)
    ob = [fh.empty(kind.words) for _ in rb]
    os_ = [fh.empty(small.words) for _ in rs]
    for i in range(2):
        srv.submit(pb, rb[i][2], rb[i][3], ob[i])
        srv.submit(ps, rs[i][2], rs[i][3], os_[i])
    srv.drain(120)
    srv.close()
    assert np.array_equal(fh.to_host(ob[0]), kind.oracle_result(rb[0][0], rb[0][1]))
    assert np.array_equal(fh.to_host(os_[1]), small.oracle_result(rs[1][0], rs[1][1]))
