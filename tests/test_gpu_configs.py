"""GPU parity at the shapes of the benchmarked workloads, bit for bit against
the oracle (not against other GPU calls):

- the uneven batch's widest parameter sets (N = 32768 with 9 and 12 moduli
  drawn as the benchmark draws them, and 13 moduli of 62 bits, whose t/Q
  scaling runs the 22- to 26-modulus scaler instances), multiply +
  relinearize + a rotation;
- one fhehip_submit holding jobs at every degree of the uneven batch at both
  ends of its modulus-count range;
- the serving workload's request kinds (N = 16384, 6 x 60-bit + special, and
  N = 32768, 8 x 60-bit + special, each a multiply + relinearize + 12
  rotate-and-add steps) through the request server;
- the batched workload at its full 512 instances, spot-checked at instances
  0-3, 255 and 511 (the key loads of ks_mac are shared by groups of four);
- fhe.rs's own benchmark parameters at N = 16384 (default_parameters_128(20),
  fhe/src/bfv/parameters.rs:236-249, used by fhe/benches/bfv.rs:28).
"""
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

T = 65537
# The uneven batch's parameter rule: a set of `count` moduli at degree n has
# bit sizes drawn by numpy's default_rng(n + count).integers(50, 63, count)
# (generate_moduli, fhe/src/bfv/parameters.rs:383-423); the last modulus is
# the key's special prime.
UNEVEN_EXTREMES = [(4096, 2), (4096, 3), (8192, 3), (8192, 5), (16384, 5), (16384, 8), (32768, 8), (32768, 12)]


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


def uneven_sizes(n, count):
    return [int(x) for x in np.random.default_rng(n + count).integers(50, 63, count)]


class Params:
    """A parameter set with its keys on the GPU and in the oracle: the
    ciphertexts one level below the key (one special prime)."""

    def __init__(self, fh, oracle, n, sizes, steps, seed, t=T, ct_level=1, key_level=0):
        self.n, self.t = n, t
        self.bfv = oracle.Bfv(n, t, oracle.generate_moduli(sizes, n), sk_seed=seed)
        self.moduli = self.bfv.level_moduli(ct_level)
        self.L = len(self.moduli)
        self.level = ct_level
        cctx = fh.Context(self.moduli, n)
        kctx = fh.Context(self.bfv.level_moduli(key_level), n)
        self.rk = self.bfv.relin_key(ct_level=ct_level, key_level=key_level, seed=seed + 1)
        c0, c1 = self.rk.export()
        self.mult = fh.Multiplicator(rk=fh.KeySwitchingKey(cctx, kctx, c0, c1), plaintext=t)
        self.om = oracle.Multiplicator(self.bfv, self.rk, level=ct_level)
        self.gk, gkeys = {}, []
        for i, e in enumerate(sorted({pow(3, st, 2 * n) for st in steps})):
            ok = self.bfv.galois_key(e, ct_level=ct_level, key_level=key_level, seed=seed + 10 + i)
            c0, c1 = ok.export()
            gkeys.append(fh.GaloisKey(fh.KeySwitchingKey(cctx, kctx, c0, c1), e))
            self.gk[e] = ok
        self.ek = fh.EvaluationKey(gkeys) if gkeys else None

    @property
    def words(self):
        return 2 * self.L * self.n

    def rotate(self, ct, step):
        e = pow(3, step, 2 * self.n)
        return self.bfv.galois(self.gk[e], ct, e)

    def add(self, a, b):
        q = np.array(self.moduli, dtype=object).reshape(1, -1, 1)
        s = (a.reshape(2, self.L, self.n).astype(object) + b.reshape(2, self.L, self.n).astype(object)) % q
        return s.astype(np.uint64).reshape(-1)

    def job(self, a, b, steps):
        ct = self.om.multiply(a, b)
        for st in steps:
            ct = self.rotate(ct, st)
        return ct

    def slot_sum(self, a, b, log2_slots):
        ct = self.om.multiply(a, b)
        for k in range(log2_slots):
            ct = self.add(ct, self.rotate(ct, 1 << k))
        return ct


@pytest.mark.parametrize("sizes", [uneven_sizes(32768, 9), uneven_sizes(32768, 12), [62] * 13],
                         ids=["9-moduli", "12-moduli", "13x62"])
def test_wide_degree_32768_mul_relin_rotate(fh, oracle, sizes):
    n = 32768
    print(f"N=32768 moduli sizes {sizes}", flush=True)
    ps = Params(fh, oracle, n, sizes, [1], seed=len(sizes))
    rng = random.Random(len(sizes))
    batch = 2
    vs = [[rng.randrange(T) for _ in range(n)] for _ in range(2 * batch)]
    lhs = np.concatenate([ps.bfv.encrypt(vs[2 * i], level=1, seed=10 + i) for i in range(batch)])
    rhs = np.concatenate([ps.bfv.encrypt(vs[2 * i + 1], level=1, seed=20 + i) for i in range(batch)])
    prod = fh.empty(lhs.size)
    ps.mult.multiply(fh.to_device(lhs), fh.to_device(rhs), prod, batch)
    rot = fh.empty(lhs.size)
    ps.ek.rotate_columns(prod, rot, 1, batch)
    _sync()
    got_p, got_r = fh.to_host(prod), fh.to_host(rot)
    w = ps.words
    for i in range(batch):
        ep = ps.om.multiply(lhs[i * w:(i + 1) * w], rhs[i * w:(i + 1) * w])
        assert np.array_equal(got_p[i * w:(i + 1) * w], ep), i
        assert np.array_equal(got_r[i * w:(i + 1) * w], ps.rotate(ep, 1)), i
    h = n // 2
    want = [(a * b) % T for a, b in zip(vs[0], vs[1])]
    want = want[1:h] + want[:1] + want[h + 1:] + want[h:h + 1]
    assert [int(x) for x in ps.bfv.decrypt(got_r[:w], level=1)] == want


def test_scaler_with_32_source_moduli(fh, oracle):
    """The widest scaler instantiation (32 source moduli) in both directions:
    a t/Q-style down-scaling 32 -> 16 and an extension 16 -> 32."""
    n = 1024
    q = oracle.generate_moduli([62] * 16 + [60] * 16, n)
    lo, hi = q[:16], q
    clo, chi = fh.Context(lo, n), fh.Context(hi, n)
    olo, ohi = oracle.Context(lo, n), oracle.Context(hi, n)
    Q = 1
    for x in lo:
        Q *= x
    for ntt in (False, True):
        a = oracle.random_polys(hi, n, 3, seed=31 + ntt)
        down = fh.Scaler(chi, clo, 65537, Q)
        out = fh.empty(3 * len(lo) * n)
        down.scale(fh.to_device(a), out, 3, ntt)
        _sync()
        assert np.array_equal(fh.to_host(out), oracle.PolyScaler(ohi, olo, 65537, Q).scale(a, ntt))
        b = oracle.random_polys(lo, n, 3, seed=41 + ntt)
        up = fh.Scaler(clo, chi)
        out2 = fh.empty(3 * len(hi) * n)
        up.scale(fh.to_device(b), out2, 3, ntt)
        _sync()
        assert np.array_equal(fh.to_host(out2), oracle.PolyScaler(olo, ohi).scale(b, ntt))


def test_submit_every_degree_at_both_ends(fh, oracle):
    """One fhehip_submit with a job at each (degree, modulus count) end of the
    uneven batch, each with 1-4 rotations by steps from {1, 2, 4}, against the
    oracle."""
    rng = random.Random(11)
    sets = [Params(fh, oracle, n, uneven_sizes(n, c), [1, 2, 4], seed=3 * k + 1)
            for k, (n, c) in enumerate(UNEVEN_EXTREMES)]
    print("parameter sets ready", flush=True)
    jobs, expect = [], []
    for k, ps in enumerate(sets):
        for r in range(2):
            a = oracle.random_polys(ps.moduli, ps.n, 2, seed=100 * k + r)
            b = oracle.random_polys(ps.moduli, ps.n, 2, seed=100 * k + 50 + r)
            steps = [rng.choice([1, 2, 4]) for _ in range(rng.randrange(1, 5))]
            o = fh.empty(ps.words)
            jobs.append((ps.mult, ps.ek, fh.to_device(a), fh.to_device(b), o, steps))
            expect.append((ps, a, b, steps))
    order = list(range(len(jobs)))
    rng.shuffle(order)
    fh.submit([jobs[i] for i in order])
    _sync()
    for j, (ps, a, b, steps) in zip(jobs, expect):
        assert np.array_equal(fh.to_host(j[4]), ps.job(a, b, steps)), (ps.n, ps.L, steps)


def test_server_at_serving_shapes(fh, oracle):
    """Requests of both serving kinds (12 rotate-and-add steps), arriving
    together, against the oracle's multiply + relinearize + slot sum."""
    kinds = [Params(fh, oracle, 16384, [60] * 7, [1 << k for k in range(12)], seed=1),
             Params(fh, oracle, 32768, [60] * 9, [1 << k for k in range(12)], seed=2)]
    print("serving kinds ready", flush=True)
    srv = fh.Server(max_batch=8, lanes=3)
    ids = [srv.add_params(k.mult, k.ek, 12)[0] for k in kinds]
    reqs = []
    for i, kind in enumerate([0, 0, 1, 0, 0, 1, 0]):
        ps = kinds[kind]
        a = oracle.random_polys(ps.moduli, ps.n, 2, seed=10 + i)
        b = oracle.random_polys(ps.moduli, ps.n, 2, seed=40 + i)
        reqs.append((kind, a, b, fh.to_device(a), fh.to_device(b), fh.empty(ps.words)))
    _sync()
    for kind, a, b, da, db, o in reqs:
        srv.submit(ids[kind], da, db, o)
    srv.drain(120)
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    for i, (kind, a, b, da, db, o) in enumerate(reqs):
        if i in (0, 1, 2, 5):
            assert np.array_equal(fh.to_host(o), kinds[kind].slot_sum(a, b, 12)), (i, kind)


def test_full_batch_of_512_spot_checked(fh, oracle):
    import torch
    n, batch = 16384, 512
    bfv = oracle.Bfv(n, T, oracle.generate_moduli([60] * 7, n), sk_seed=5)
    rk = bfv.relin_key(ct_level=1, key_level=0, seed=6)
    moduli = bfv.level_moduli(1)
    L = len(moduli)
    cctx, kctx = fh.Context(moduli, n), fh.Context(bfv.level_moduli(0), n)
    c0, c1 = rk.export()
    gm = fh.Multiplicator(rk=fh.KeySwitchingKey(cctx, kctx, c0, c1), plaintext=T)
    om = oracle.Multiplicator(bfv, rk, level=1)
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(3)

    def uniform():
        x = torch.empty((batch, 2, L, n), dtype=torch.int64, device=dev)
        for i, q in enumerate(moduli):
            x[:, :, i, :] = torch.randint(0, q, (batch, 2, n), dtype=torch.int64, device=dev, generator=gen)
        return x.view(-1)

    lhs, rhs = uniform(), uniform()
    out = torch.empty_like(lhs)
    gm.multiply(lhs, rhs, out, batch)
    _sync()
    w = 2 * L * n
    for i in (0, 1, 2, 3, 255, 511):
        a, b = fh.to_host(lhs[i * w:(i + 1) * w]), fh.to_host(rhs[i * w:(i + 1) * w])
        assert np.array_equal(fh.to_host(out[i * w:(i + 1) * w]), om.multiply(a, b)), i


# default_parameters_128's N = 16384 moduli (fhe/src/bfv/parameters.rs:236-249)
REF_16384_MODULI = [0xfffffffd8001, 0xfffffffa0001, 0xfffffff00001, 0x1fffffff68001, 0x1fffffff50001,
                    0x1ffffffee8001, 0x1ffffffea0001, 0x1ffffffe88001, 0x1ffffffe48001]


def test_reference_default_parameters_16384(fh, oracle):
    """Multiplicator::default with RelinearizationKey::new (key and ciphertext
    at level 0) at fhe.rs's own benchmark parameters: degree 16384, the nine
    48/49-bit moduli, t = generate_prime(20, 2N, 2^20 - 1)."""
    n = 16384
    t = oracle.generate_prime(20, 2 * n, (1 << 20) - 1)
    bfv = oracle.Bfv(n, t, REF_16384_MODULI, sk_seed=21)
    rk = bfv.relin_key(seed=22)
    ctx = fh.Context(REF_16384_MODULI, n)
    c0, c1 = rk.export()
    gm = fh.Multiplicator(rk=fh.KeySwitchingKey(ctx, ctx, c0, c1), plaintext=t)
    om = oracle.Multiplicator(bfv, rk, level=0)
    assert gm.extended_basis() == om.ext_moduli()
    rng = random.Random(t)
    batch = 4
    vs = [[rng.randrange(t) for _ in range(n)] for _ in range(2 * batch)]
    lhs = np.concatenate([bfv.encrypt(vs[2 * i], seed=30 + i) for i in range(batch)])
    rhs = np.concatenate([bfv.encrypt(vs[2 * i + 1], seed=40 + i) for i in range(batch)])
    out = fh.empty(lhs.size)
    gm.multiply(fh.to_device(lhs), fh.to_device(rhs), out, batch)
    _sync()
    got = fh.to_host(out)
    assert np.array_equal(got, om.multiply(lhs, rhs))
    w = lhs.size // batch
    for i in range(batch):
        assert [int(x) for x in bfv.decrypt(got[i * w:(i + 1) * w])] == \
            [(a * b) % t for a, b in zip(vs[2 * i], vs[2 * i + 1])]
