"""Leveled operations on the GPU: Ciphertext::switch_down
(fhe/src/bfv/ciphertext.rs:86-97) and Multiplicator::enable_mod_switching
(fhe/src/bfv/ops/mul.rs:146-155, 221-223), bit for bit against the oracle."""
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

T = 65537


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


SWITCH_CASES = [
    # n, moduli sizes, level, polynomials
    (16, [62, 62, 62], 0, 3),
    (64, [50, 55, 60, 62], 0, 3),    # dropping a 62-bit modulus into 50-bit limbs
    (1024, [60] * 4, 1, 5),
    (8192, [62] * 3, 0, 2),
    (16384, [60] * 7, 1, 2),
    (32768, [60] * 4, 0, 2),         # unfused inverse / forward transforms
]


@pytest.mark.parametrize("n,sizes,level,polys", SWITCH_CASES)
def test_switch_down_matches_oracle(fh, oracle, n, sizes, level, polys):
    print(f"switch_down n={n} sizes={sizes} level={level} polys={polys}", flush=True)
    moduli = oracle.generate_moduli(sizes, n)
    hi, lo = moduli[:len(moduli) - level], moduli[:len(moduli) - level - 1]
    ocf, oct_ = oracle.Context(hi, n), oracle.Context(lo, n)
    src = oracle.random_polys(hi, n, polys, seed=n + polys)
    sw = fh.Switcher(fh.Context(hi, n), fh.Context(lo, n))
    out = fh.empty(polys * len(lo) * n)
    sw.switch_down(fh.to_device(src), out, polys)
    _sync()
    assert np.array_equal(fh.to_host(out), oracle.ct_switch_down(ocf, oct_, src))


@pytest.mark.parametrize("n,L", [(1024, 3), (16384, 3)])
def test_switch_down_decrypts(fh, oracle, n, L):
    bfv = oracle.Bfv(n, T, oracle.generate_moduli([60] * L, n), sk_seed=7)
    rng = random.Random(n)
    v = [rng.randrange(T) for _ in range(n)]
    ct = bfv.encrypt(v, level=0, seed=8)
    for level in range(L - 1):
        sw = fh.Switcher(fh.Context(bfv.level_moduli(level), n), fh.Context(bfv.level_moduli(level + 1), n))
        out = fh.empty(2 * (L - level - 1) * n)
        sw.switch_down(fh.to_device(ct), out, 2)
        _sync()
        ct = fh.to_host(out)
        assert [int(x) for x in bfv.decrypt(ct, level=level + 1)] == v


MODSWITCH_CASES = [
    # n, moduli sizes, ct_level, key_level, relinearize, batch
    (1024, [60] * 5, 1, 0, True, 3),
    (16384, [60] * 7, 1, 0, True, 2),   # the batched configuration's shape
    (8192, [62] * 3, 0, 0, True, 2),
    (1024, [62] * 3, 0, 0, False, 2),
    (32768, [60] * 4, 1, 0, True, 1),
]


@pytest.mark.parametrize("n,sizes,ct_level,key_level,relin,batch", MODSWITCH_CASES)
def test_multiply_with_mod_switching(fh, oracle, n, sizes, ct_level, key_level, relin, batch):
    print(f"mod switching n={n} sizes={sizes} levels={ct_level},{key_level} relin={relin}", flush=True)
    bfv = oracle.Bfv(n, T, oracle.generate_moduli(sizes, n), sk_seed=n + 1)
    cm, nm = bfv.level_moduli(ct_level), bfv.level_moduli(ct_level + 1)
    cctx, nctx = fh.Context(cm, n), fh.Context(nm, n)
    rk = bfv.relin_key(ct_level=ct_level, key_level=key_level, seed=19) if relin else None
    om = oracle.Multiplicator(bfv, rk, level=ct_level)
    if relin:
        c0, c1 = rk.export()
        gk = fh.KeySwitchingKey(cctx, fh.Context(bfv.level_moduli(key_level), n), c0, c1)
        gm = fh.Multiplicator(rk=gk, plaintext=T)
    else:
        gm = fh.Multiplicator(base_ctx=cctx, extended_basis=om.ext_moduli(), post=(T, _q(cm)))
    gm.enable_mod_switching(nctx)
    parts = 2 if relin else 3
    rng = random.Random(n)
    vs = [[rng.randrange(T) for _ in range(n)] for _ in range(2 * batch)]
    lhs = np.concatenate([bfv.encrypt(vs[2 * i], level=ct_level, seed=30 + i) for i in range(batch)])
    rhs = np.concatenate([bfv.encrypt(vs[2 * i + 1], level=ct_level, seed=40 + i) for i in range(batch)])
    out = fh.empty(batch * parts * len(nm) * n)
    gm.multiply(fh.to_device(lhs), fh.to_device(rhs), out, batch)
    _sync()
    got = fh.to_host(out)
    ocf, ocn = oracle.Context(cm, n), oracle.Context(nm, n)
    step = 2 * len(cm) * n
    for i in range(batch):
        prod = om.multiply(lhs[i * step:(i + 1) * step], rhs[i * step:(i + 1) * step])
        want = oracle.ct_switch_down(ocf, ocn, prod)
        g = got[i * parts * len(nm) * n:(i + 1) * parts * len(nm) * n]
        assert np.array_equal(g, want), i
        dec = [int(x) for x in bfv.decrypt(g, level=ct_level + 1)]
        assert dec == [(a * b) % T for a, b in zip(vs[2 * i], vs[2 * i + 1])]


def _q(moduli):
    q = 1
    for m in moduli:
        q *= int(m)
    return q


def test_leveled_errors(fh, oracle):
    n = 1024
    m = oracle.generate_moduli([60] * 4, n)
    a, b = fh.Context(m, n), fh.Context(m[:3], n)
    with pytest.raises(fh.FheHipError):
        fh.Switcher(a, fh.Context(m[1:], n))          # not a prefix
    with pytest.raises(fh.FheHipError):
        fh.Switcher(b, a)                              # upwards
    sw = fh.Switcher(a, b)
    buf = fh.empty(2 * 4 * n)
    with pytest.raises(fh.FheHipError):
        sw.switch_down(buf, buf, 2)                    # aliasing
    bfv = oracle.Bfv(n, T, m, sk_seed=3)
    rk = bfv.relin_key(ct_level=1, key_level=0, seed=4)
    c0, c1 = rk.export()
    gm = fh.Multiplicator(rk=fh.KeySwitchingKey(b, a, c0, c1), plaintext=T)
    with pytest.raises(fh.FheHipError):
        gm.enable_mod_switching(fh.Context(m[1:3], n))  # not the next level
    with pytest.raises(fh.FheHipError):
        gm.enable_mod_switching(fh.Context(m[:1], n))   # two levels down
    gm.enable_mod_switching(fh.Context(m[:2], n))
    with pytest.raises(fh.FheHipError):
        fh.submit([(gm, None, buf, buf, buf, [])])


def test_multiply_graph_with_mod_switching(fh, oracle):
    """fhehip_multiply_graph of a mod-switching Multiplicator: every replay
    equals the eager call, with eager work and synchronizations between the
    replays."""
    n, batch = 4096, 3
    print(f"graph with mod switching n={n} batch={batch}", flush=True)
    bfv = oracle.Bfv(n, T, oracle.generate_moduli([60] * 5, n), sk_seed=21)
    rk = bfv.relin_key(ct_level=0, key_level=0, seed=22)
    c0, c1 = rk.export()
    cctx = fh.Context(bfv.level_moduli(0), n)
    gm = fh.Multiplicator(rk=fh.KeySwitchingKey(cctx, cctx, c0, c1), plaintext=T)
    gm.enable_mod_switching(fh.Context(bfv.level_moduli(1), n))
    lhs = fh.to_device(oracle.random_polys(bfv.level_moduli(0), n, 2 * batch, seed=5))
    rhs = fh.to_device(oracle.random_polys(bfv.level_moduli(0), n, 2 * batch, seed=6))
    size = batch * 2 * 4 * n
    eager, out = fh.empty(size), fh.empty(size)
    gm.multiply(lhs, rhs, eager, batch)
    _sync()
    want = fh.to_host(eager)
    g = fh.Graph(gm, lhs, rhs, out, batch)
    try:
        for r in range(3):
            out.zero_()
            g.launch()
            _sync()
            assert np.array_equal(fh.to_host(out), want), r
            gm.multiply(rhs, lhs, eager, batch)
            _sync()
    finally:
        g.close()


@pytest.mark.parametrize("n,sizes,level,target,cts", [(1024, [60] * 5, 0, 3, 2), (16, [62] * 4, 1, 3, 1),
                                                     (16384, [60] * 7, 0, 6, 1), (32768, [60] * 4, 0, 2, 1),
                                                     (64, [62] * 3, 1, 1, 2)])
def test_switch_to_level_matches_oracle(fh, oracle, n, sizes, level, target, cts):
    """Ciphertext::switch_to_level (ciphertext.rs:99-119): one inverse NTT,
    the switch-down steps and one forward NTT equal repeated switch_down;
    the ciphertexts decrypt at the target level (a target equal to the
    level is a copy)."""
    print(f"switch_to_level n={n} sizes={sizes} {level}->{target}", flush=True)
    t = T if n > 16 else 1153
    bfv = oracle.Bfv(n, t, oracle.generate_moduli(sizes, n), sk_seed=5)
    vals = [[i + 1, 2, 3] for i in range(cts)]
    src = np.concatenate([bfv.encrypt(v, level=level, seed=i) for i, v in enumerate(vals)])
    want = src
    for lv in range(level, target):
        want = oracle.ct_switch_down(oracle.Context(bfv.level_moduli(lv), n),
                                     oracle.Context(bfv.level_moduli(lv + 1), n), want)
    sw = fh.Switcher(fh.Context(bfv.level_moduli(level), n), fh.Context(bfv.level_moduli(target), n))
    step = 2 * len(bfv.level_moduli(target)) * n
    out = fh.empty(cts * step)
    sw.switch_down(fh.to_device(src), out, 2 * cts)
    _sync()
    got = fh.to_host(out)
    assert np.array_equal(got, want)
    for i, v in enumerate(vals):
        assert [int(x) for x in bfv.decrypt(got[i * step:(i + 1) * step], level=target)][:3] == v
