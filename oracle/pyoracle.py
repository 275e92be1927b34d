"""ORACLE (test infrastructure only): ctypes binding of oracle/lib/liboracle.so.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module. It is the checker, never the thing measured or shipped.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "liboracle.so")

_u64p = C.POINTER(C.c_uint64)
_u32p = C.POINTER(C.c_uint32)
_i64p = C.POINTER(C.c_int64)
_dblp = C.POINTER(C.c_double)
_vp = C.c_void_p

_lib = None


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        sig = {
            "or_last_error": (C.c_char_p, []),
            "or_generate_prime": (C.c_uint64, [C.c_uint32, C.c_uint64, C.c_uint64]),
            "or_supports_opt": (C.c_int, [C.c_uint64]),
            "or_is_prime": (C.c_int, [C.c_uint64]),
            "or_generate_moduli": (C.c_int, [_u32p, C.c_uint32, C.c_uint32, _u64p]),
            "or_primitive_root": (C.c_uint64, [C.c_uint32, C.c_uint64]),
            "or_modulus_op": (C.c_int, [C.c_uint64, C.c_uint32, _u64p, _u64p, C.c_uint32, _u64p]),
            "or_ntt": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, _u64p, C.c_uint32]),
            "or_ntt_tables": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint64, _u64p, _u64p]),
            "or_rns_garner": (C.c_int, [_u64p, C.c_uint32, _u32p, C.c_uint32]),
            "or_rns_lift": (C.c_int, [_u64p, C.c_uint32, _u64p, _u32p, C.c_uint32]),
            "or_rns_scaler_new": (_vp, [_u64p, C.c_uint32, _u64p, C.c_uint32, _u32p, C.c_uint32, _u32p, C.c_uint32]),
            "or_rns_scaler_free": (None, [_vp]),
            "or_rns_scale": (C.c_int, [_vp, _u64p, C.c_uint32, _u64p, C.c_uint32, C.c_uint32]),
            "or_rns_scaler_consts": (C.c_int, [_vp, _u64p, _u64p, _u64p, _u64p, _u64p, _u64p]),
            "or_ctx_new": (_vp, [_u64p, C.c_uint32, C.c_uint32]),
            "or_ctx_free": (None, [_vp]),
            "or_ctx_psi": (C.c_int, [_vp, _u64p]),
            "or_ctx_ntt": (C.c_int, [_vp, C.c_uint32, _u64p, C.c_uint32]),
            "or_switch_down": (C.c_int, [_vp, _u64p, _u64p, C.c_uint32]),
            "or_substitute": (C.c_int, [_vp, C.c_uint32, C.c_uint32, _u64p, _u64p, C.c_uint32]),
            "or_poly_scaler_new": (_vp, [_vp, _vp, _u32p, C.c_uint32, _u32p, C.c_uint32]),
            "or_poly_scaler_free": (None, [_vp]),
            "or_poly_scale": (C.c_int, [_vp, C.c_uint32, _u64p, _u64p, C.c_uint32]),
            "or_bfv_new": (_vp, [C.c_uint32, C.c_uint64, _u64p, C.c_uint32, C.c_uint32, C.c_uint64]),
            "or_bfv_free": (None, [_vp]),
            "or_bfv_level_moduli": (C.c_int, [_vp, C.c_uint32, _u64p, _u32p]),
            "or_bfv_secret": (C.c_int, [_vp, _i64p]),
            "or_bfv_encrypt": (C.c_int, [_vp, C.c_uint32, _u64p, C.c_uint32, C.c_uint64, C.c_uint32, _u64p]),
            "or_bfv_decrypt": (C.c_int, [_vp, C.c_uint32, _u64p, C.c_uint32, C.c_uint32, _u64p]),
            "or_dot_product_scalar": (C.c_int, [_vp, _u64p, C.c_uint64, _u64p, C.c_uint64, C.c_uint32, C.c_uint32,
                                                _u64p]),
            "or_mult_new_general": (_vp, [_vp, _vp, C.c_uint32, _u64p, C.c_uint32, C.POINTER(_u32p), _u32p]),
            "or_bfv_encode": (C.c_int, [_vp, C.c_uint32, _u64p, C.c_uint32, C.c_uint32, C.c_uint32, _u64p]),
            "or_bfv_encrypt_with": (C.c_int, [_vp, C.c_uint32, _u64p, _u64p, _i64p, _u64p]),
            "or_bfv_encrypt_pk_with": (C.c_int, [_vp, C.c_uint32, _u64p, _u64p, _i64p, _i64p, _i64p, _u64p]),
            "or_rgsw_key_new": (_vp, [_vp, C.c_uint32, _u64p, C.c_uint32, C.c_uint64]),
            "or_rgsw_external_product": (C.c_int, [_vp, _vp, _u64p, _u64p, C.c_uint32]),
            "or_poly_serialize": (C.c_int, [_vp, _u64p, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint8)]),
            "or_poly_deserialize": (C.c_int, [_vp, C.POINTER(C.c_uint8), C.c_uint32, C.c_uint32, _u64p]),
            "or_expands": (C.c_int, [_vp, C.POINTER(C.c_void_p), _u32p, C.c_uint32, C.c_uint32, _u64p, C.c_uint32,
                                     _u64p]),
            "or_ksk_new": (_vp, [_vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64]),
            "or_ksk_free": (None, [_vp]),
            "or_ksk_from_rows": (_vp, [_vp, C.c_uint32, C.c_uint32, _u64p, _u64p]),
            "or_ksk_export": (C.c_int, [_vp, _u64p, _u64p, _u32p, _u32p]),
            "or_key_switch": (C.c_int, [_vp, _u64p, _u64p, _u64p, C.c_uint32]),
            "or_mult_new": (_vp, [_vp, _vp, C.c_uint32]),
            "or_mult_free": (None, [_vp]),
            "or_mult_ext_moduli": (C.c_int, [_vp, _u64p, _u32p]),
            "or_mult_multiply": (C.c_int, [_vp, _u64p, _u64p, _u64p, C.c_uint32]),
            "or_ct_mul": (C.c_int, [_vp, C.c_uint32, _u64p, _u64p, _u64p]),
            "or_relinearize": (C.c_int, [_vp, _vp, _u64p, _u64p]),
            "or_ct_mul_parts": (C.c_int, [_vp, C.c_uint32, _u64p, C.c_uint32, _u64p, C.c_uint32, _u64p]),
            "or_bfv_mul_moduli": (C.c_int, [_vp, C.c_uint32, _u64p, _u32p]),
            "or_galois": (C.c_int, [_vp, _vp, C.c_uint32, _u64p, _u64p, C.c_uint32]),
            "or_random_polys": (C.c_int, [_u64p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, _u64p]),
            "or_bench_mul_relin": (C.c_int, [_vp, _u64p, _u64p, C.c_uint32, C.c_uint32, _dblp, _dblp]),
            "or_bench_ntt": (C.c_int, [C.c_uint64, C.c_uint32, _u64p, C.c_uint32, C.c_uint32, _dblp]),
            "or_sha256": (None, [C.c_char_p, C.c_size_t, C.POINTER(C.c_uint8)]),
            "or_chacha_block": (None, [_u32p, C.c_int, _u32p]),
            "or_chacha8_stream": (None, [C.c_char_p, C.c_int, C.c_size_t, C.c_void_p]),
            "or_pcg32_seed": (None, [C.c_uint64, C.POINTER(C.c_uint8)]),
            "or_pcg32_output": (C.c_uint32, [C.c_uint64]),
            "or_uniform": (None, [C.c_char_p, C.c_int, C.c_uint64, C.c_uint64, C.c_size_t, _u64p]),
            "or_random_from_seed": (None, [_u64p, C.c_uint32, C.c_uint32, C.c_char_p, _u64p]),
            "or_generate_c1": (None, [_u64p, C.c_uint32, C.c_uint32, C.c_char_p, C.c_uint32, _u64p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p64(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u64p)


def _pi64(a):
    assert a.dtype == np.int64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_i64p)


def _p32(a):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u32p)


def _chk(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().or_last_error().decode())


def _chkp(p):
    if not p:
        raise RuntimeError("oracle: " + lib().or_last_error().decode())
    return p


def u64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


def big_words(x: int):
    """Python int -> little-endian uint32 words."""
    out = []
    while x:
        out.append(x & 0xFFFFFFFF)
        x >>= 32
    return np.array(out or [0], dtype=np.uint32)


def words_big(w) -> int:
    r = 0
    for i, v in enumerate(np.asarray(w, dtype=np.uint64).tolist()):
        r |= int(v) << (32 * i)
    return r


# ---- Zq --------------------------------------------------------------------
def generate_prime(bits, modulo, upper):
    r = lib().or_generate_prime(bits, modulo, upper)
    return None if r == 0 else int(r)


def supports_opt(p):
    return bool(lib().or_supports_opt(p))


def is_prime(p):
    return bool(lib().or_is_prime(p))


def generate_moduli(sizes, degree):
    s = np.array(sizes, dtype=np.uint32)
    out = np.zeros(len(sizes), dtype=np.uint64)
    _chk(lib().or_generate_moduli(_p32(s), len(sizes), degree, _p64(out)))
    return [int(x) for x in out]


def primitive_root(n, p):
    return int(lib().or_primitive_root(n, p))


# ---- pseudo-random streams (oracle_rng.hpp) ----------------------------------
def sha256(data: bytes) -> bytes:
    out = (C.c_uint8 * 32)()
    lib().or_sha256(data, len(data), out)
    return bytes(out)


def chacha_block(state16, rounds):
    a = np.asarray(state16, dtype=np.uint32)
    out = np.zeros(16, dtype=np.uint32)
    lib().or_chacha_block(_p32(a), rounds, _p32(out))
    return out


def chacha8_stream(seed: bytes, kind: str, n: int):
    """ChaCha8Rng::from_seed(seed): n x next_u32 ("u32"), next_u64 ("u64"),
    n bytes of fill_bytes ("bytes"), or next_u32 / next_u64 alternating ("mixed")."""
    assert len(seed) == 32
    k = {"u32": 0, "u64": 1, "bytes": 2, "mixed": 3}[kind]
    if k == 0:
        out = np.zeros(n, dtype=np.uint32)
    elif k == 2:
        out = np.zeros(n, dtype=np.uint8)
    else:
        out = np.zeros(n, dtype=np.uint64)
    lib().or_chacha8_stream(seed, k, n, out.ctypes.data_as(C.c_void_p))
    return out


def pcg32_seed(state: int) -> bytes:
    out = (C.c_uint8 * 32)()
    lib().or_pcg32_seed(state, out)
    return bytes(out)


def pcg32_output(state: int) -> int:
    return int(lib().or_pcg32_output(state))


def uniform(seed: bytes, n: int, high: int, low: int = 0, single: bool = False):
    out = np.zeros(n, dtype=np.uint64)
    lib().or_uniform(seed, 1 if single else 0, low, high, n, _p64(out))
    return out


def random_from_seed(moduli, n, seed: bytes):
    m = u64(moduli)
    out = np.zeros(len(moduli) * n, dtype=np.uint64)
    lib().or_random_from_seed(_p64(m), len(moduli), n, seed, _p64(out))
    return out


def generate_c1(moduli, n, seed: bytes, size: int):
    m = u64(moduli)
    out = np.zeros(size * len(moduli) * n, dtype=np.uint64)
    lib().or_generate_c1(_p64(m), len(moduli), n, seed, size, _p64(out))
    return out


MOD_OPS = {n: i for i, n in enumerate(
    ["add", "sub", "mul", "mul_opt", "shoup", "mul_shoup", "lazy_mul_shoup", "reduce", "lazy_reduce",
     "reduce_u128", "lazy_reduce_u128", "neg", "lazy_reduce_opt", "reduce_opt_u128", "pow", "inv",
     "center", "reduce_i64"])}


def modulus_op(p, op, a, b=None):
    a = u64(a)
    bb = u64(b) if b is not None else None
    out = np.zeros_like(a)
    _chk(lib().or_modulus_op(p, MOD_OPS[op], _p64(a), _p64(bb) if bb is not None else None, len(a), _p64(out)))
    return out


# ---- NTT -------------------------------------------------------------------
NTT_VARIANTS = {"forward": 0, "forward_vt": 1, "forward_vt_lazy": 2, "backward": 3}


def ntt(p, n, a, variant="forward", psi=0):
    a = u64(a).copy()
    count = a.size // n
    _chk(lib().or_ntt(p, n, psi, NTT_VARIANTS[variant], _p64(a), count))
    return a


def ntt_tables(p, n, psi=0):
    out = np.zeros(4 * n, dtype=np.uint64)
    inv = np.zeros(1, dtype=np.uint64)
    _chk(lib().or_ntt_tables(p, n, psi, _p64(out), _p64(inv)))
    return out.reshape(4, n), int(inv[0])


# ---- RNS -------------------------------------------------------------------
def rns_garner(moduli):
    m = u64(moduli)
    stride = 2 * len(moduli) + 4
    out = np.zeros(len(moduli) * stride, dtype=np.uint32)
    _chk(lib().or_rns_garner(_p64(m), len(m), _p32(out), stride))
    return [words_big(out[i * stride:(i + 1) * stride]) for i in range(len(moduli))]


def rns_lift(moduli, rests):
    m, r = u64(moduli), u64(rests)
    stride = 2 * len(moduli) + 4
    out = np.zeros(stride, dtype=np.uint32)
    _chk(lib().or_rns_lift(_p64(m), len(m), _p64(r), _p32(out), stride))
    return words_big(out)


class RnsScaler:
    def __init__(self, from_moduli, to_moduli, num: int, den: int):
        self.f, self.t = u64(from_moduli), u64(to_moduli)
        n, d = big_words(num), big_words(den)
        self.h = _chkp(lib().or_rns_scaler_new(_p64(self.f), len(self.f), _p64(self.t), len(self.t),
                                               _p32(n), len(n), _p32(d), len(d)))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            try:
                _lib.or_rns_scaler_free(self.h)
            except Exception:  # interpreter teardown
                pass

    def scale(self, rests, start=0, length=None):
        r = u64(rests).reshape(-1, len(self.f))
        length = len(self.t) - start if length is None else length
        out = np.zeros((r.shape[0], length), dtype=np.uint64)
        _chk(lib().or_rns_scale(self.h, _p64(r), r.shape[0], _p64(out), start, length))
        return out

    def consts(self):
        nf, nt = len(self.f), len(self.t)
        gamma = np.zeros(nt, np.uint64)
        tg = np.zeros(3, np.uint64)
        om = np.zeros(nt * nf, np.uint64)
        to = np.zeros(3 * nf, np.uint64)
        tga = np.zeros(2 * nf, np.uint64)
        sh = np.zeros(1, np.uint64)
        _chk(lib().or_rns_scaler_consts(self.h, _p64(gamma), _p64(tg), _p64(om), _p64(to), _p64(tga), _p64(sh)))
        return dict(gamma=gamma, theta_gamma=tg, omega=om.reshape(nt, nf), theta_omega=to.reshape(nf, 3),
                    theta_garner=tga.reshape(nf, 2), shift=int(sh[0]))


# ---- rq --------------------------------------------------------------------
class Context:
    def __init__(self, moduli, n):
        self.moduli = [int(x) for x in moduli]
        self.n = n
        m = u64(moduli)
        self.h = _chkp(lib().or_ctx_new(_p64(m), len(m), n))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            try:
                _lib.or_ctx_free(self.h)
            except Exception:  # interpreter teardown
                pass

    @property
    def L(self):
        return len(self.moduli)

    def psi(self):
        out = np.zeros(self.L, np.uint64)
        _chk(lib().or_ctx_psi(self.h, _p64(out)))
        return out

    def _count(self, a):
        return a.size // (self.L * self.n)

    def forward(self, a):
        a = u64(a).copy()
        _chk(lib().or_ctx_ntt(self.h, 0, _p64(a), self._count(a)))
        return a

    def backward(self, a):
        a = u64(a).copy()
        _chk(lib().or_ctx_ntt(self.h, 1, _p64(a), self._count(a)))
        return a

    def switch_down(self, a):
        a = u64(a)
        c = self._count(a)
        out = np.zeros(c * (self.L - 1) * self.n, np.uint64)
        _chk(lib().or_switch_down(self.h, _p64(a), _p64(out), c))
        return out

    def substitute(self, a, exponent, ntt_repr):
        a = u64(a)
        out = np.zeros_like(a)
        _chk(lib().or_substitute(self.h, exponent, 1 if ntt_repr else 0, _p64(a), _p64(out), self._count(a)))
        return out

    def random(self, count, seed):
        return random_polys(self.moduli, self.n, count, seed)


def ct_switch_down(ctx_from: Context, ctx_to: Context, ct):
    """Ciphertext::switch_down (fhe/src/bfv/ciphertext.rs:86-97): every
    NTT-domain polynomial [polys][L][N] of ctx_from to power basis,
    Poly::switch_down (rq/mod.rs:390-449), and into the NTT domain of ctx_to
    (ctx_from without its last modulus)."""
    assert ctx_to.moduli == ctx_from.moduli[:-1] and ctx_to.n == ctx_from.n
    return ctx_to.forward(ctx_from.switch_down(ctx_from.backward(ct)))


class PolyScaler:
    def __init__(self, from_ctx: Context, to_ctx: Context, num=None, den=None):
        self.f, self.t = from_ctx, to_ctx
        if num is None:
            n = np.zeros(1, np.uint32)
            d = np.zeros(1, np.uint32)
            self.h = _chkp(lib().or_poly_scaler_new(from_ctx.h, to_ctx.h, _p32(n), 0, _p32(d), 0))
        else:
            n, d = big_words(num), big_words(den)
            self.h = _chkp(lib().or_poly_scaler_new(from_ctx.h, to_ctx.h, _p32(n), len(n), _p32(d), len(d)))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            try:
                _lib.or_poly_scaler_free(self.h)
            except Exception:  # interpreter teardown
                pass

    def scale(self, a, ntt_repr):
        a = u64(a)
        c = a.size // (self.f.L * self.f.n)
        out = np.zeros(c * self.t.L * self.t.n, np.uint64)
        _chk(lib().or_poly_scale(self.h, 1 if ntt_repr else 0, _p64(a), _p64(out), c))
        return out


def dot_product_scalar(ctx: Context, cts, pts, count, parts, ct_stride=None, pt_stride=None):
    """dot_product_scalar (fhe/src/bfv/ops/dot_product.rs:56-143) of `count`
    ciphertexts [parts][L][N] with NTT-domain plaintext polynomials [L][N],
    each `ct_stride` / `pt_stride` words apart; returns [parts][L][N]."""
    L, n = len(ctx.moduli), ctx.n
    ct_stride = parts * L * n if ct_stride is None else ct_stride
    pt_stride = L * n if pt_stride is None else pt_stride
    cts, pts = u64(cts), u64(pts)
    out = np.zeros(parts * L * n, np.uint64)
    _chk(lib().or_dot_product_scalar(ctx.h, _p64(cts), ct_stride, _p64(pts), pt_stride, count, parts, _p64(out)))
    return out


def serialization_length(moduli, n):
    """Bytes of one polynomial's Rq coefficients (Modulus::serialization_length per limb)."""
    return sum((int(q) - 1).bit_length() * n // 8 for q in moduli)


def poly_serialize(ctx: Context, polys, count, ntt):
    """Rq coefficient payloads (rq/convert.rs:15-39) of `count` polynomials, back to back."""
    out = np.zeros(count * serialization_length(ctx.moduli, ctx.n), np.uint8)
    _chk(lib().or_poly_serialize(ctx.h, _p64(u64(polys)), count, 1 if ntt else 0,
                                 out.ctypes.data_as(C.POINTER(C.c_uint8))))
    return out


def poly_deserialize(ctx: Context, data, count, ntt):
    data = np.ascontiguousarray(data, dtype=np.uint8)
    out = np.zeros(count * len(ctx.moduli) * ctx.n, np.uint64)
    _chk(lib().or_poly_deserialize(ctx.h, data.ctypes.data_as(C.POINTER(C.c_uint8)), count, 1 if ntt else 0,
                                   _p64(out)))
    return out


def random_polys(moduli, n, count, seed):
    m = u64(moduli)
    out = np.zeros(count * len(moduli) * n, np.uint64)
    _chk(lib().or_random_polys(_p64(m), len(m), n, count, seed, _p64(out)))
    return out


# ---- BFV -------------------------------------------------------------------
class Bfv:
    """BfvParameters + one seeded SecretKey."""

    def __init__(self, degree, t, moduli, variance=10, sk_seed=1):
        self.degree, self.t, self.moduli = degree, t, [int(x) for x in moduli]
        m = u64(moduli)
        self.h = _chkp(lib().or_bfv_new(degree, t, _p64(m), len(m), variance, sk_seed))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            try:
                _lib.or_bfv_free(self.h)
            except Exception:  # interpreter teardown
                pass

    def level_moduli(self, level):
        out = np.zeros(len(self.moduli), np.uint64)
        L = np.zeros(1, np.uint32)
        _chk(lib().or_bfv_level_moduli(self.h, level, _p64(out), _p32(L)))
        return [int(x) for x in out[:int(L[0])]]

    def secret(self):
        out = np.zeros(self.degree, np.int64)
        _chk(lib().or_bfv_secret(self.h, _pi64(out)))
        return out

    def encrypt(self, slots, level=0, seed=7, poly=False):
        """SIMD-encoded slots, or with `poly` the plaintext's coefficients
        (Encoding::poly_at_level)."""
        s = u64(slots)
        L = len(self.level_moduli(level))
        out = np.zeros(2 * L * self.degree, np.uint64)
        _chk(lib().or_bfv_encrypt(self.h, level, _p64(s), len(s), seed, 1 if poly else 0, _p64(out)))
        return out

    def decrypt(self, ct, level=0, poly=False):
        ct = u64(ct)
        L = len(self.level_moduli(level))
        parts = ct.size // (L * self.degree)
        out = np.zeros(self.degree, np.uint64)
        _chk(lib().or_bfv_decrypt(self.h, level, _p64(ct), parts, 1 if poly else 0, _p64(out)))
        return out

    def encode(self, values, level=0, poly=False, scaled=True):
        """Plaintext::try_encode: the NTT-domain polynomial [L][N] that
        encryption adds (to_poly, scaled) or Plaintext::poly_ntt."""
        v = u64(values)
        L = len(self.level_moduli(level))
        out = np.zeros(L * self.degree, np.uint64)
        _chk(lib().or_bfv_encode(self.h, level, _p64(v), len(v), 1 if poly else 0, 1 if scaled else 0, _p64(out)))
        return out

    def encrypt_with(self, pt, a, e, level=0):
        """SecretKey::encrypt_poly of pt with the given a (NTT) and e (signed)."""
        L = len(self.level_moduli(level))
        out = np.zeros(2 * L * self.degree, np.uint64)
        e = np.ascontiguousarray(e, dtype=np.int64)
        _chk(lib().or_bfv_encrypt_with(self.h, level, _p64(u64(pt)), _p64(u64(a)), _pi64(e), _p64(out)))
        return out

    def rgsw_encrypt(self, values, level=0, seed=31):
        """SecretKey::try_encrypt into an RGSWCiphertext (rgsw_ciphertext.rs:97-117)
        of SIMD-encoded values: its two KeySwitchingKeys."""
        pt = self.encode(values, level=level, scaled=False)
        ks = []
        for which in (0, 1):
            h = lib().or_rgsw_key_new(self.h, level, _p64(pt), which, seed + which)
            ks.append(KeySwitchingKey(self, h, level, level))
        return ks

    def rgsw_product(self, rgsw, ct):
        """Ciphertext x RGSWCiphertext (rgsw_ciphertext.rs:123-160)."""
        ct = u64(ct)
        out = np.zeros_like(ct)
        L = len(self.level_moduli(rgsw[0].ct_level))
        _chk(lib().or_rgsw_external_product(rgsw[0].h, rgsw[1].h, _p64(ct), _p64(out), ct.size // (2 * L * self.degree)))
        return out

    def encrypt_pk_with(self, pk, pt, u, e1, e2, level=0):
        """PublicKey::try_encrypt of pt with the public key pk [2][L][N] and the given u, e1, e2."""
        L = len(self.level_moduli(level))
        out = np.zeros(2 * L * self.degree, np.uint64)
        u, e1, e2 = [np.ascontiguousarray(x, dtype=np.int64) for x in (u, e1, e2)]
        _chk(lib().or_bfv_encrypt_pk_with(self.h, level, _p64(u64(pk)), _p64(u64(pt)), _pi64(u), _pi64(e1),
                                          _pi64(e2), _p64(out)))
        return out

    def expands(self, gks, ct, size, level=0):
        """EvaluationKey::expands (fhe/src/bfv/keys/evaluation_key.rs:153-193)
        with Galois keys `gks` (from galois_key); returns [size][2][L][N]."""
        ct = u64(ct)
        L = len(self.level_moduli(level))
        arr = (C.c_void_p * max(len(gks), 1))(*[k.h for k in gks])
        exps = np.array([k.exponent for k in gks] or [0], np.uint32)
        out = np.zeros(max(size, 1) * 2 * L * self.degree, np.uint64)
        _chk(lib().or_expands(self.h, arr, _p32(exps), len(gks), level, _p64(ct), size, _p64(out)))
        return out

    def relin_key(self, ct_level=0, key_level=0, seed=11):
        return KeySwitchingKey(self, lib().or_ksk_new(self.h, 0, 0, ct_level, key_level, seed), ct_level, key_level)

    def galois_key(self, exponent, ct_level=0, key_level=0, seed=13):
        k = KeySwitchingKey(self, lib().or_ksk_new(self.h, 1, exponent, ct_level, key_level, seed), ct_level,
                            key_level)
        k.exponent = exponent
        return k

    def key_from_rows(self, c0, c1, ct_level=0, key_level=0):
        """KeySwitchingKey from NTT-domain rows [digits][L_key][N]."""
        c0, c1 = u64(c0).reshape(-1), u64(c1).reshape(-1)
        k = KeySwitchingKey(self, lib().or_ksk_from_rows(self.h, ct_level, key_level, _p64(c0), _p64(c1)),
                            ct_level, key_level)
        return k

    def ct_mul(self, lhs, rhs, level=0):
        L = len(self.level_moduli(level))
        out = np.zeros(3 * L * self.degree, np.uint64)
        _chk(lib().or_ct_mul(self.h, level, _p64(u64(lhs)), _p64(u64(rhs)), _p64(out)))
        return out

    def ct_mul_parts(self, lhs, rhs, level=0):
        """&ct * &ct (ops/mod.rs:259-342) for ciphertexts of any part counts."""
        lhs, rhs = u64(lhs), u64(rhs)
        w = len(self.level_moduli(level)) * self.degree
        lp, rp = lhs.size // w, rhs.size // w
        out = np.zeros((lp + rp - 1) * w, np.uint64)
        _chk(lib().or_ct_mul_parts(self.h, level, _p64(lhs), lp, _p64(rhs), rp, _p64(out)))
        return out

    def mul_moduli(self, level=0):
        """Moduli of the level's multiplication context (parameters.rs:585-612)."""
        out = np.zeros(128, np.uint64)
        n = np.array([128], np.uint32)
        _chk(lib().or_bfv_mul_moduli(self.h, level, _p64(out), _p32(n)))
        return [int(x) for x in out[:int(n[0])]]

    def relinearize(self, rk, ct3):
        L = len(self.level_moduli(rk.ct_level))
        out = np.zeros(2 * L * self.degree, np.uint64)
        _chk(lib().or_relinearize(self.h, rk.h, _p64(u64(ct3)), _p64(out)))
        return out

    def galois(self, gk, ct, exponent=None):
        ct = u64(ct)
        e = gk.exponent if exponent is None else exponent
        L = len(self.level_moduli(gk.ct_level))
        out = np.zeros_like(ct)
        _chk(lib().or_galois(self.h, gk.h, e, _p64(ct), _p64(out), ct.size // (2 * L * self.degree)))
        return out


class KeySwitchingKey:
    def __init__(self, bfv, h, ct_level, key_level):
        self.bfv, self.h = bfv, _chkp(h)
        self.ct_level, self.key_level = ct_level, key_level

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            try:
                _lib.or_ksk_free(self.h)
            except Exception:  # interpreter teardown
                pass

    def export(self):
        d = np.zeros(1, np.uint32)
        kl = np.zeros(1, np.uint32)
        _chk(lib().or_ksk_export(self.h, None, None, _p32(d), _p32(kl)))
        digits, key_L = int(d[0]), int(kl[0])
        n = self.bfv.degree
        c0 = np.zeros(digits * key_L * n, np.uint64)
        c1 = np.zeros_like(c0)
        _chk(lib().or_ksk_export(self.h, _p64(c0), _p64(c1), _p32(d), _p32(kl)))
        return c0.reshape(digits, key_L, n), c1.reshape(digits, key_L, n)

    def key_switch(self, p_pb):
        p = u64(p_pb)
        L = len(self.bfv.level_moduli(self.ct_level))
        KL = len(self.bfv.level_moduli(self.key_level))
        n = self.bfv.degree
        c = p.size // (L * n)
        c0 = np.zeros(c * KL * n, np.uint64)
        c1 = np.zeros_like(c0)
        _chk(lib().or_key_switch(self.h, _p64(p), _p64(c0), _p64(c1), c))
        return c0, c1


class Multiplicator:
    """Multiplicator::default(rk) (or without relinearization when rk is None)."""

    def __init__(self, bfv, rk=None, level=0, extended_basis=None, lhs=None, rhs=None, post=None):
        """With extended_basis: Multiplicator::new_leveled (mul.rs:37-98) with
        the (num, den) factors lhs, rhs, post (None: one)."""
        self.bfv, self.rk, self.level = bfv, rk, level
        rkh = rk.h if rk is not None else None
        if extended_basis is None:
            self.h = _chkp(lib().or_mult_new(bfv.h, rkh, level))
            return
        ext = u64(extended_basis)
        arrs, counts = [], []
        for f in (lhs, rhs, post):
            for x in (f if f is not None else (None, None)):
                w = big_words(x) if x is not None else np.zeros(1, np.uint32)
                arrs.append(w)
                counts.append(len(w) if x is not None else 0)
        ptrs = (_u32p * 6)(*[_p32(a) for a in arrs])
        cnt = np.array(counts, np.uint32)
        self._keep = arrs
        self.h = _chkp(lib().or_mult_new_general(bfv.h, rkh, level, _p64(ext), len(ext), ptrs, _p32(cnt)))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            try:
                _lib.or_mult_free(self.h)
            except Exception:  # interpreter teardown
                pass

    def ext_moduli(self):
        out = np.zeros(64, np.uint64)
        n = np.zeros(1, np.uint32)
        _chk(lib().or_mult_ext_moduli(self.h, _p64(out), _p32(n)))
        return [int(x) for x in out[:int(n[0])]]

    def multiply(self, lhs, rhs):
        lhs, rhs = u64(lhs), u64(rhs)
        L = len(self.bfv.level_moduli(self.level))
        n = self.bfv.degree
        count = lhs.size // (2 * L * n)
        parts = 2 if self.rk is not None else 3
        out = np.zeros(count * parts * L * n, np.uint64)
        _chk(lib().or_mult_multiply(self.h, _p64(lhs), _p64(rhs), _p64(out), count))
        return out

    def bench(self, lhs, rhs, threads, reps):
        wall = np.zeros(1, np.float64)
        lat = np.zeros(reps, np.float64)
        _chk(lib().or_bench_mul_relin(self.h, _p64(u64(lhs)), _p64(u64(rhs)), threads, reps,
                                      wall.ctypes.data_as(_dblp), lat.ctypes.data_as(_dblp)))
        return float(wall[0]), lat


def bench_ntt(p, n, a, threads):
    a = u64(a).copy()
    wall = np.zeros(1, np.float64)
    _chk(lib().or_bench_ntt(p, n, _p64(a), a.size // n, threads, wall.ctypes.data_as(_dblp)))
    return float(wall[0])
