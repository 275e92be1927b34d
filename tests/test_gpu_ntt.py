"""GPU parity of the batched NTT kernels against the oracle (bit-exact)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CFG2_MODULI = [4611686018427322369, 4611686018427289601, 4611686018425815041, 4611686018424733697,
               4611686018423881729, 4611686018423390209, 4611686018423062529, 4611686018422669313]


@pytest.fixture(scope="module")
def fh():
    import torch
    import fhehip
    assert torch.cuda.is_available()
    fhehip.lib()
    return fhehip


@pytest.mark.parametrize("n,L", [(8, 1), (16, 2), (32, 3), (64, 2), (1024, 3), (4096, 2), (8192, 3), (16384, 2),
                                 (32768, 1), (32768, 3)])
def test_ntt_matches_oracle(fh, oracle, n, L):
    moduli = oracle.generate_moduli([62] * L, n)
    print(f"ntt parity n={n} L={L}", flush=True)
    ctx = fh.Context(moduli, n)
    octx = oracle.Context(moduli, n)
    assert ctx.psi() == [int(x) for x in octx.psi()]
    count = 3
    a = oracle.random_polys(moduli, n, count, seed=n + L)
    d = fh.to_device(a)
    ctx.ntt_forward(d, count)
    fwd = fh.to_host(d)
    assert np.array_equal(fwd, octx.forward(a))
    ctx.ntt_backward(d, count)
    assert np.array_equal(fh.to_host(d), a)
    # lazy variant: outputs < 4p and congruent to the reduced transform
    d2 = fh.to_device(a)
    ctx.ntt_forward(d2, count, lazy=True)
    lz = fh.to_host(d2).reshape(count, L, n)
    ref = fwd.reshape(count, L, n)
    for l, p in enumerate(moduli):
        assert (lz[:, l] < np.uint64(4 * p)).all() if 4 * p < 2**64 else True
        assert np.array_equal(lz[:, l] % np.uint64(p), ref[:, l])


def _smallest_prime_above(oracle, lo, n):
    p = lo + (1 - lo) % (2 * n)  # first value = 1 mod 2n from lo on
    if p <= lo:
        p += 2 * n
    for _ in range(100000):
        if oracle.is_prime(p):
            return p
        p += 2 * n
    raise AssertionError("no prime found")


@pytest.mark.parametrize("n", [8, 1024, 16384, 32768])
def test_ntt_moduli_around_2_60(fh, oracle, n):
    """Moduli on both sides of 2^60 (the largest prime below it, whose lazy
    operands reach just under 2^62, and the smallest above it), a 62-bit and
    small moduli, with all-(p-1) and delta rows, in both kernel shapes."""
    moduli = oracle.generate_moduli([60, 62, 20, 40], n) + [_smallest_prime_above(oracle, 1 << 60, n)]
    assert moduli[0] < 1 << 60 <= moduli[4]
    L = len(moduli)
    print(f"ntt moduli around 2^60 n={n}", flush=True)
    ctx = fh.Context(moduli, n)
    octx = oracle.Context(moduli, n)
    a = oracle.random_polys(moduli, n, 3, seed=n + 7).reshape(3, L, n)
    a[1] = np.array(moduli, np.uint64)[:, None] - np.uint64(1)
    a[2] = 0
    a[2, :, 0] = 1
    a = a.reshape(-1)
    ref = octx.forward(a)
    for shape in (4, 5):
        fh.ntt_set_shape(n, shape, shape)
        d = fh.to_device(a)
        ctx.ntt_forward(d, 3)
        assert np.array_equal(fh.to_host(d), ref), shape
        ctx.ntt_backward(d, 3)
        assert np.array_equal(fh.to_host(d), a), shape
        d2 = fh.to_device(a)
        ctx.ntt_forward(d2, 3, lazy=True)
        lz = fh.to_host(d2).reshape(3, L, n)
        for l, p in enumerate(moduli):
            assert (lz[:, l] < np.uint64(4 * p)).all()
            assert np.array_equal(lz[:, l] % np.uint64(p), ref.reshape(3, L, n)[:, l])
    fh.ntt_set_shape(n, 4, 4)


def test_ntt_edge_values(fh, oracle):
    n = 4096
    moduli = [4611686018427322369, 1152921504606748673]
    ctx = fh.Context(moduli, n)
    octx = oracle.Context(moduli, n)
    a = np.zeros((3, 2, n), np.uint64)
    a[1] = np.array(moduli, np.uint64)[:, None] - np.uint64(1)  # all p-1
    a[2, :, 0] = 1  # delta
    d = fh.to_device(a.reshape(-1))
    ctx.ntt_forward(d, 3)
    assert np.array_equal(fh.to_host(d), octx.forward(a.reshape(-1)))
    ctx.ntt_backward(d, 3)
    assert np.array_equal(fh.to_host(d), a.reshape(-1))


def test_ntt_cfg2_shape_roundtrip(fh, oracle):
    """256 polynomials x 8 x 62-bit primes x N=16384: forward vs oracle on every
    row, then the inverse returns the input bit for bit."""
    n, count = 16384, 256
    print("ntt cfg2 shape", flush=True)
    ctx = fh.Context(CFG2_MODULI, n)
    octx = oracle.Context(CFG2_MODULI, n)
    a = oracle.random_polys(CFG2_MODULI, n, count, seed=2024)
    d = fh.to_device(a)
    ctx.ntt_forward(d, count)
    assert np.array_equal(fh.to_host(d), octx.forward(a))
    ctx.ntt_backward(d, count)
    assert np.array_equal(fh.to_host(d), a)


def test_ntt_explicit_psi(fh, oracle):
    n, p = 1024, 4611686018427322369
    ps = oracle.primitive_root(n, p)
    other = pow(ps, 3, p)  # another primitive 2N-th root
    ctx = fh.Context([p], n, psi=[other])
    assert ctx.psi() == [other]
    a = oracle.random_polys([p], n, 1, seed=5)
    d = fh.to_device(a)
    ctx.ntt_forward(d, 1)
    assert np.array_equal(fh.to_host(d), oracle.ntt(p, n, a, "forward", psi=other))


def test_ntt_unsupported_modulus(fh):
    with pytest.raises(fh.FheHipError) as e:
        fh.Context([4611686018427387847], 1024)  # prime but not 1 mod 2048
    assert e.value.code == fh.ERR_UNSUPPORTED


def test_ntt_32768_split_rows(fh, oracle):
    """N=32768 runs its first forward stage (last inverse stage) as a separate
    pass over row halves; mixed 50/62-bit primes and edge rows pin the pairing
    of halves to limbs."""
    n = 32768
    moduli = oracle.generate_moduli([62, 50, 60, 55], n)
    ctx = fh.Context(moduli, n)
    octx = oracle.Context(moduli, n)
    a = oracle.random_polys(moduli, n, 2, seed=99).reshape(2, 4, n)
    a[1, 0] = np.uint64(moduli[0] - 1)
    a[1, 1] = 0
    a[1, 1, n - 1] = 1
    a = a.reshape(-1)
    d = fh.to_device(a)
    ctx.ntt_forward(d, 2)
    assert np.array_equal(fh.to_host(d), octx.forward(a))
    ctx.ntt_backward(d, 2)
    assert np.array_equal(fh.to_host(d), a)


@pytest.mark.parametrize("n", [1024, 16384, 32768])
def test_ntt_shapes_and_autotune(fh, oracle, n):
    """Both kernel shapes (16 and 32 values per thread) are bit-exact; the
    autotuner picks one of them per direction."""
    moduli = oracle.generate_moduli([62, 60], n)
    ctx = fh.Context(moduli, n)
    octx = oracle.Context(moduli, n)
    a = oracle.random_polys(moduli, n, 3, seed=n + 1)
    ref = octx.forward(a)
    for fwd, inv in [(4, 5), (5, 4), (5, 5), (4, 4)]:
        fh.ntt_set_shape(n, fwd, inv)
        d = fh.to_device(a)
        ctx.ntt_forward(d, 3)
        assert np.array_equal(fh.to_host(d), ref), (fwd, inv)
        ctx.ntt_backward(d, 3)
        assert np.array_equal(fh.to_host(d), a), (fwd, inv)
    f, i, ms = ctx.ntt_autotune(64)
    assert f in (4, 5) and i in (4, 5) and all(x >= 0 for x in ms)
    d = fh.to_device(a)
    ctx.ntt_forward(d, 3)
    assert np.array_equal(fh.to_host(d), ref)
    fh.ntt_set_shape(n, 4, 4)
