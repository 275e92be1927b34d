"""Edge cases of the entry points around the evaluation path: zero-sized
batches are no-ops that leave the output untouched, and oversized batches
are refused with an error rather than launched."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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


def test_zero_sized_calls_are_noops(fh, oracle):
    import torch
    print("zero-sized calls", flush=True)
    n = 1024
    moduli = [int(q) for q in oracle.generate_moduli([62, 62, 62], n)]
    ctx, lo = fh.Context(moduli, n), fh.Context(moduli[:2], n)
    sentinel = 0x5A5A5A5A5A5A5A5A
    out = torch.full((4 * 3 * n,), sentinel, dtype=torch.int64, device="cuda")
    src = fh.to_device(oracle.random_polys(moduli, n, 4, seed=1))
    fh.Switcher(ctx, lo).switch_down(src, out, 0)
    fh.serialize(ctx, src, 0, out, ntt=True)
    fh.deserialize(ctx, out, 0, out, ntt=True)
    fh.dot_product_columns(ctx, src, 2, 2, src, 0, out)   # no columns
    codec = fh.Codec(ctx, fh.Context(moduli[:2], n), 65537, sk=np.zeros(n, np.int64))
    codec.encode(src, 0, out)
    codec.encrypt(src, src, src, out, 0)
    codec.decrypt(src, 2, 0, out)
    _sync()
    assert bool((out == sentinel).all())


def test_oversized_batches_are_refused(fh, oracle):
    import torch
    n = 16
    moduli = [int(q) for q in oracle.generate_moduli([62, 62], n)]
    ctx = fh.Context(moduli, n)
    buf = torch.empty(8, dtype=torch.int64, device="cuda")
    with pytest.raises(fh.FheHipError):
        fh.serialize(ctx, buf, 70000, buf, ntt=False)
    with pytest.raises(fh.FheHipError):
        fh.deserialize(ctx, buf, 70000, buf, ntt=False)
    codec = fh.Codec(ctx, ctx, 1153, sk=np.zeros(n, np.int64))
    with pytest.raises(fh.FheHipError):
        codec.encode(buf, 70000, buf)
    with pytest.raises(fh.FheHipError):
        fh.Switcher(ctx, fh.Context(moduli[:1], n)).switch_down(buf, buf[4:], 70000)
