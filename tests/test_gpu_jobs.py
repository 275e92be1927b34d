"""GPU parity of fhehip_submit: a batch of jobs with differing parameter sets
(degree, moduli, keys), each a multiply + relinearize followed by up to four
column rotations, equals the same operations run one job at a time, and the
oracle's Multiplicator and Galois key switching, bit for bit."""
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


def _sync():
    import torch
    torch.cuda.synchronize()


class ParamSet:
    """Keys of one parameter set on the GPU and in the oracle."""

    def __init__(self, fh, oracle, n, sizes, ct_level, key_level, steps, seed):
        self.n = n
        self.bfv = oracle.Bfv(n, T, oracle.generate_moduli(sizes, n), sk_seed=seed)
        self.ct_level = ct_level
        self.moduli = self.bfv.level_moduli(ct_level)
        self.L = len(self.moduli)
        cctx = fh.Context(self.moduli, n)
        kctx = fh.Context(self.bfv.level_moduli(key_level), n)
        self.rk = self.bfv.relin_key(ct_level=ct_level, key_level=key_level, seed=seed + 1)
        c0, c1 = self.rk.export()
        self.mult = fh.Multiplicator(rk=fh.KeySwitchingKey(cctx, kctx, c0, c1), plaintext=T)
        self.om = oracle.Multiplicator(self.bfv, self.rk, level=ct_level)
        self.gk = {}
        gkeys = []
        for i, st in enumerate(steps):
            e = pow(3, st, 2 * n)
            ok = self.bfv.galois_key(e, ct_level=ct_level, key_level=key_level, seed=seed + 10 + i)
            c0, c1 = ok.export()
            gkeys.append(fh.GaloisKey(fh.KeySwitchingKey(cctx, kctx, c0, c1), e))
            self.gk[st] = ok
        self.ek = fh.EvaluationKey(gkeys)

    @property
    def ct_words(self):
        return 2 * self.L * self.n


def _single(fh, ps, a, b, steps):
    """The job run as separate single-instance calls."""
    out = fh.empty(ps.ct_words)
    ps.mult.multiply(a, b, out, 1)
    for st in steps:
        nxt = fh.empty(ps.ct_words)
        ps.ek.rotate_columns(out, nxt, st, 1)
        out = nxt
    _sync()
    return fh.to_host(out)


def _oracle(ps, a, b, steps):
    ct = ps.om.multiply(a, b)
    for st in steps:
        ct = ps.bfv.galois(ps.gk[st], ct, pow(3, st, 2 * ps.n))
    return ct


def test_mixed_parameter_jobs_match_single_calls_and_oracle(fh, oracle):
    sets = [ParamSet(fh, oracle, 1024, [60] * 4, 1, 0, [1, 2, 3], seed=1),
            ParamSet(fh, oracle, 2048, [55, 60, 62], 0, 0, [1, 4], seed=2),
            ParamSet(fh, oracle, 4096, [50, 55, 60], 1, 0, [2, 5], seed=3)]
    rng = random.Random(7)
    specs, expect, meta = [], [], []
    # per set, one contiguous block of inputs plus separately allocated ones
    for si, ps in enumerate(sets):
        blk = 3
        a_all = oracle.random_polys(ps.moduli, ps.n, 2 * blk, seed=10 + si)
        b_all = oracle.random_polys(ps.moduli, ps.n, 2 * blk, seed=20 + si)
        da, db = fh.to_device(a_all), fh.to_device(b_all)
        w = ps.ct_words
        for k in range(blk):
            meta.append((ps, da[k * w:(k + 1) * w], db[k * w:(k + 1) * w],
                         a_all[k * w:(k + 1) * w], b_all[k * w:(k + 1) * w]))
        for k in range(3):
            a = oracle.random_polys(ps.moduli, ps.n, 2, seed=100 * si + k)
            b = oracle.random_polys(ps.moduli, ps.n, 2, seed=100 * si + 50 + k)
            meta.append((ps, fh.to_device(a), fh.to_device(b), a, b))
    rng.shuffle(meta)
    outs = []
    for ps, da, db, a, b in meta:
        steps = [rng.choice(sorted(ps.gk)) for _ in range(rng.randrange(0, 5))]
        o = fh.empty(ps.ct_words)
        outs.append(o)
        specs.append((ps.mult, ps.ek, da, db, o, steps))
        expect.append((ps, da, db, a, b, steps))
    print(f"submit {len(specs)} jobs over {len(sets)} parameter sets", flush=True)
    fh.submit(specs)
    _sync()
    for k, (ps, da, db, a, b, steps) in enumerate(expect):
        got = fh.to_host(outs[k])
        assert np.array_equal(got, _single(fh, ps, da, db, steps)), (k, ps.n, steps)
        if k < 4:
            assert np.array_equal(got, _oracle(ps, a, b, steps)), (k, ps.n, steps)


def test_contiguous_group_and_repeated_submits(fh, oracle):
    """A group whose inputs and outputs are contiguous runs without staging;
    two submits back to back (the second reusing the first's scratch) give
    the same results."""
    ps = ParamSet(fh, oracle, 1024, [60] * 3, 1, 0, [1, 2], seed=5)
    cnt, w = 5, ps.ct_words
    a = oracle.random_polys(ps.moduli, ps.n, 2 * cnt, seed=1)
    b = oracle.random_polys(ps.moduli, ps.n, 2 * cnt, seed=2)
    da, db = fh.to_device(a), fh.to_device(b)
    out1, out2 = fh.empty(cnt * w), fh.empty(cnt * w)
    steps = [[], [1], [1, 2], [2, 2, 1], [1, 1, 1, 1]]
    for out in (out1, out2):
        fh.submit([(ps.mult, ps.ek, da[k * w:(k + 1) * w], db[k * w:(k + 1) * w], out[k * w:(k + 1) * w], steps[k])
                   for k in range(cnt)])
    _sync()
    h1, h2 = fh.to_host(out1), fh.to_host(out2)
    assert np.array_equal(h1, h2)
    for k in range(cnt):
        assert np.array_equal(h1[k * w:(k + 1) * w], _single(fh, ps, da[k * w:(k + 1) * w], db[k * w:(k + 1) * w],
                                                             steps[k]))


def test_degree_32768_job(fh, oracle):
    ps = ParamSet(fh, oracle, 32768, [60] * 3, 1, 0, [3], seed=9)
    small = ParamSet(fh, oracle, 1024, [60] * 3, 1, 0, [3], seed=10)
    jobs, exp = [], []
    for ps_, seed in ((ps, 1), (small, 2), (ps, 3)):
        a = oracle.random_polys(ps_.moduli, ps_.n, 2, seed=seed)
        b = oracle.random_polys(ps_.moduli, ps_.n, 2, seed=seed + 100)
        o = fh.empty(ps_.ct_words)
        jobs.append((ps_.mult, ps_.ek, fh.to_device(a), fh.to_device(b), o, [3]))
        exp.append(_oracle(ps_, a, b, [3]))
    fh.submit(jobs)
    _sync()
    for j, e in zip(jobs, exp):
        assert np.array_equal(fh.to_host(j[4]), e)


def test_degree_65536_job(fh, oracle):
    """A wide-degree job (elementwise passes around split transforms) in one
    submit with fused-pipeline jobs."""
    wide = ParamSet(fh, oracle, 65536, [60] * 3, 1, 0, [2], seed=12)
    small = ParamSet(fh, oracle, 4096, [60] * 3, 1, 0, [2], seed=13)
    jobs, exp = [], []
    for ps_, seed in ((wide, 1), (small, 2)):
        a = oracle.random_polys(ps_.moduli, ps_.n, 2, seed=seed)
        b = oracle.random_polys(ps_.moduli, ps_.n, 2, seed=seed + 100)
        o = fh.empty(ps_.ct_words)
        jobs.append((ps_.mult, ps_.ek, fh.to_device(a), fh.to_device(b), o, [2]))
        exp.append(_oracle(ps_, a, b, [2]))
    fh.submit(jobs)
    _sync()
    for j, e in zip(jobs, exp):
        assert np.array_equal(fh.to_host(j[4]), e)


def test_invalid_job_rejected_before_any_work(fh, oracle):
    ps = ParamSet(fh, oracle, 1024, [60] * 3, 1, 0, [1], seed=11)
    a = fh.to_device(oracle.random_polys(ps.moduli, ps.n, 2, seed=1))
    o = fh.empty(ps.ct_words)
    o.zero_()
    with pytest.raises(fh.FheHipError) as e:
        fh.submit([(ps.mult, ps.ek, a, a, o, []), (ps.mult, ps.ek, a, a, o, [2])])  # no key for step 2
    assert e.value.code == fh.ERR_INVALID_ARG
    _sync()
    assert not fh.to_host(o).any()
