"""GPU test of fhehip_multiply_noalloc (the multiplication with its scratch in
the library's fixed device region): through the C ABI, two calls in a row,
the second with a smaller batch, each equal bit for bit to fhehip_multiply on
the same inputs (and to the oracle for one instance); batches whose scratch
exceeds the region are refused before any work."""
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


def _mult(fh, oracle, n, sizes, ct_level, key_level, seed):
    bfv = oracle.Bfv(n, T, oracle.generate_moduli(sizes, n), sk_seed=seed)
    rk = bfv.relin_key(ct_level=ct_level, key_level=key_level, seed=seed + 1)
    moduli = bfv.level_moduli(ct_level)
    cctx, kctx = fh.Context(moduli, n), fh.Context(bfv.level_moduli(key_level), n)
    c0, c1 = rk.export()
    gm = fh.Multiplicator(rk=fh.KeySwitchingKey(cctx, kctx, c0, c1), plaintext=T)
    return bfv, rk, moduli, gm, (cctx, kctx)


@pytest.mark.parametrize("n,sizes,ct_level,key_level,batches",
                         [(16384, [60] * 7, 1, 0, (12, 3)), (8192, [62] * 3, 0, 0, (40, 1)),
                          (1024, [60] * 5, 2, 0, (9, 4))])
def test_noalloc_matches_multiply(fh, oracle, n, sizes, ct_level, key_level, batches):
    import torch
    print(f"noalloc n={n} sizes={sizes} levels={ct_level},{key_level} batches={batches}", flush=True)
    bfv, rk, moduli, gm, ctxs = _mult(fh, oracle, n, sizes, ct_level, key_level, seed=n % 97)
    L = len(moduli)
    w = 2 * L * n
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(n)

    def uniform(b):
        x = torch.empty((b, 2, L, n), dtype=torch.int64, device=dev)
        for i, q in enumerate(moduli):
            x[:, :, i, :] = torch.randint(0, q, (b, 2, n), dtype=torch.int64, device=dev, generator=gen)
        return x.view(-1)

    for rep, b in enumerate(batches):
        lhs, rhs = uniform(b), uniform(b)
        ref, got = torch.empty_like(lhs), torch.full_like(lhs, -1)
        gm.multiply(lhs, rhs, ref, b)
        gm.multiply_noalloc(lhs, rhs, got, b)
        _sync()
        assert torch.equal(got, ref), (rep, b)
        if rep == 1:
            om = oracle.Multiplicator(bfv, rk, level=ct_level)
            a0, b0 = fh.to_host(lhs[:w]), fh.to_host(rhs[:w])
            assert np.array_equal(fh.to_host(got[:w]), om.multiply(a0, b0))


def test_noalloc_refuses_batches_beyond_the_region(fh, oracle):
    bfv, rk, moduli, gm, ctxs = _mult(fh, oracle, 16384, [60] * 7, 1, 0, seed=3)
    L, n = len(moduli), 16384
    b = 14   # 14 x 19.1 MiB > 256 MiB
    import torch
    lhs = torch.zeros(b * 2 * L * n, dtype=torch.int64, device="cuda")
    out = torch.full_like(lhs, 7)
    with pytest.raises(fh.FheHipError) as e:
        gm.multiply_noalloc(lhs, lhs, out, b)
    assert e.value.code == fh.ERR_OOM
    _sync()
    assert bool((out == 7).all())
    gm.multiply_noalloc(lhs, lhs, out, 0)   # empty batch: nothing to do
