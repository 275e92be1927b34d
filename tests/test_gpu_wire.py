"""The Rq wire format on the GPU (math/src/rq/convert.rs:15-87): payloads
byte for byte equal to the oracle's, in both representations, and the
inverse, at N = 8 ... 32768 and coefficient widths 5 ... 62 bits."""
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


CASES = [
    # n, explicit primes or None, sizes, polynomials
    (8, [17], None, 3),
    (16, [12289, 40961], None, 2),
    (16, None, [62, 62, 62], 3),
    (4096, None, [36, 36, 37], 2),
    (8192, None, [50, 55, 55], 4),
    (16384, None, [60] * 7, 2),
    (32768, None, [62] * 3, 2),
    (65536, None, [60] * 2, 2),
]


@pytest.mark.parametrize("n,primes,sizes,num", CASES)
def test_wire_matches_oracle(fh, oracle, n, primes, sizes, num):
    import torch
    print(f"wire n={n} primes={primes} sizes={sizes} num={num}", flush=True)
    moduli = primes or [int(q) for q in oracle.generate_moduli(sizes, n)]
    octx, ctx = oracle.Context(moduli, n), fh.Context(moduli, n)
    per = oracle.serialization_length(moduli, n)
    assert fh.serialization_length(ctx) == per
    polys = oracle.random_polys(moduli, n, num, seed=n)
    polys[:n] = moduli[0] - 1
    for ntt in (False, True):
        data = torch.empty(num * per, dtype=torch.uint8, device="cuda")
        fh.serialize(ctx, fh.to_device(polys), num, data, ntt=ntt)
        _sync()
        want = oracle.poly_serialize(octx, polys, num, ntt)
        assert np.array_equal(data.cpu().numpy(), want), ntt
        back = fh.empty(polys.size)
        fh.deserialize(ctx, data, num, back, ntt=ntt)
        _sync()
        assert np.array_equal(fh.to_host(back), polys), ntt
        assert np.array_equal(fh.to_host(back), oracle.poly_deserialize(octx, want, num, ntt)), ntt


def test_wire_unaligned_buffer(fh, oracle):
    """A payload buffer at an odd address takes the byte-wise packing."""
    import torch
    n = 1024
    moduli = [int(q) for q in oracle.generate_moduli([62, 62], n)]
    ctx = fh.Context(moduli, n)
    per = fh.serialization_length(ctx)
    polys = oracle.random_polys(moduli, n, 2, seed=1)
    buf = torch.empty(2 * per + 8, dtype=torch.uint8, device="cuda")
    fh.serialize(ctx, fh.to_device(polys), 2, buf[3:3 + 2 * per], ntt=False)
    _sync()
    assert np.array_equal(buf[3:3 + 2 * per].cpu().numpy(), oracle.poly_serialize(oracle.Context(moduli, n), polys,
                                                                                   2, False))
