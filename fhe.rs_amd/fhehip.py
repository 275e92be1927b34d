"""ctypes binding of lib/libfhehip.so (the C ABI in include/fhehip.h).

Device buffers are passed as raw device pointers; the helpers accept torch
tensors (int64 storage holding the u64 residues bit for bit) for convenience.
There is no CPU fallback: every call goes to the HIP library, and loading fails
loudly when the library is missing.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfhehip.so")

OK = 0
ERR_INVALID_ARG = -1
ERR_UNSUPPORTED = -2
ERR_BAD_LEVEL = -3
ERR_CTX_MISMATCH = -4
ERR_OOM = -5
ERR_DEVICE = -6

_u64p = C.POINTER(C.c_uint64)
_u32p = C.POINTER(C.c_uint32)
_szp = C.POINTER(C.c_size_t)
_vp = C.c_void_p
_vpp = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); the exported surface of include/fhehip.h
SIGNATURES = {
    "fhehip_last_error": (C.c_char_p, []),
    "fhehip_version": (C.c_char_p, []),
    "fhehip_context_create": (C.c_int, [_u64p, C.c_uint32, C.c_uint32, _u64p, _vpp]),
    "fhehip_context_destroy": (None, [_vp]),
    "fhehip_context_info": (C.c_int, [_vp, _u32p, _u32p]),
    "fhehip_context_psi": (C.c_int, [_vp, _u64p]),
    "fhehip_ntt_forward": (C.c_int, [_vp, _vp, C.c_uint32, C.c_int, _vp]),
    "fhehip_ntt_backward": (C.c_int, [_vp, _vp, C.c_uint32, _vp]),
    "fhehip_ntt_autotune": (C.c_int, [_vp, C.c_uint32, _u32p, _u32p, C.POINTER(C.c_float)]),
    "fhehip_ntt_set_shape": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32]),
    "fhehip_split_row_limit": (C.c_int, [C.c_uint32, _u32p]),
    "fhehip_scaler_create": (C.c_int, [_vp, _vp, _vp, _vpp]),
    "fhehip_scaler_destroy": (None, [_vp]),
    "fhehip_scaler_scale": (C.c_int, [_vp, _vp, _vp, C.c_uint32, C.c_int, _vp]),
    "fhehip_ksk_create": (C.c_int, [_vp, _vp, _u64p, _u64p, C.c_uint32, _vpp]),
    "fhehip_ksk_destroy": (None, [_vp]),
    "fhehip_key_switch": (C.c_int, [_vp, _vp, _vp, C.c_uint32, _vp]),
    "fhehip_relinearize": (C.c_int, [_vp, _vp, _vp, C.c_uint32, _vp]),
    "fhehip_multiplicator_default": (C.c_int, [_vp, C.c_uint64, _vpp]),
    "fhehip_multiplicator_create": (C.c_int, [_vp, _u64p, C.c_uint32, _vp, _vp, _vp, _vpp]),
    "fhehip_multiplicator_enable_relinearization": (C.c_int, [_vp, _vp]),
    "fhehip_multiplicator_enable_mod_switching": (C.c_int, [_vp, _vp]),
    "fhehip_switcher_create": (C.c_int, [_vp, _vp, _vpp]),
    "fhehip_switcher_destroy": (None, [_vp]),
    "fhehip_switch_down": (C.c_int, [_vp, _vp, _vp, C.c_uint32, _vp]),
    "fhehip_multiplicator_destroy": (None, [_vp]),
    "fhehip_multiplicator_extended_basis": (C.c_int, [_vp, _u64p, _u32p]),
    "fhehip_multiply": (C.c_int, [_vp, _vp, _vp, _vp, C.c_uint32, _vp]),
    "fhehip_multiply_noalloc": (C.c_int, [_vp, _vp, _vp, _vp, C.c_uint32, _vp]),
    "fhehip_multiply_autotune": (C.c_int, [_vp, C.c_uint32, _vp, C.POINTER(C.c_float)]),
    "fhehip_tuning_set": (C.c_int, [C.c_uint32, _vp]),
    "fhehip_tuning_get": (C.c_int, [C.c_uint32, _vp]),
    "fhehip_parameters_create": (C.c_int, [C.c_uint32, C.c_uint64, _u64p, C.c_uint32, _vpp]),
    "fhehip_parameters_destroy": (None, [_vp]),
    "fhehip_parameters_info": (C.c_int, [_vp, _u32p, _u32p, _u64p]),
    "fhehip_parameters_context": (C.c_int, [_vp, C.c_uint32, _vpp]),
    "fhehip_parameters_plaintext_context": (C.c_int, [_vp, _vpp]),
    "fhehip_parameters_mul_basis": (C.c_int, [_vp, C.c_uint32, _u64p, _u32p]),
    "fhehip_ciphertext_mul": (C.c_int, [_vp, C.c_uint32, _vp, C.c_uint32, _vp, C.c_uint32, _vp, C.c_uint32, _vp]),
    "fhehip_galois_create": (C.c_int, [_vp, C.c_uint32, _vpp]),
    "fhehip_galois_destroy": (None, [_vp]),
    "fhehip_galois_relinearize": (C.c_int, [_vp, _vp, _vp, C.c_uint32, _vp]),
    "fhehip_evaluation_key_create": (C.c_int, [_vpp, C.c_uint32, _vpp]),
    "fhehip_evaluation_key_destroy": (None, [_vp]),
    "fhehip_rotate_columns": (C.c_int, [_vp, _vp, _vp, C.c_uint32, C.c_uint32, _vp]),
    "fhehip_rotate_rows": (C.c_int, [_vp, _vp, _vp, C.c_uint32, _vp]),
    "fhehip_inner_sum": (C.c_int, [_vp, _vp, _vp, C.c_uint32, _vp]),
    "fhehip_rotate_and_sum": (C.c_int, [_vp, _vp, _vp, C.c_uint32, C.c_int, C.c_uint32, _vp]),
    "fhehip_ciphertext_add": (C.c_int, [_vp, _vp, _vp, _vp, C.c_uint32, _vp]),
    "fhehip_ciphertext_sub": (C.c_int, [_vp, _vp, _vp, _vp, C.c_uint32, _vp]),
    "fhehip_ciphertext_neg": (C.c_int, [_vp, _vp, _vp, C.c_uint32, _vp]),
    "fhehip_ciphertext_mul_plain": (C.c_int, [_vp, _vp, C.c_uint32, C.c_uint32, _vp, C.c_uint64, _vp, _vp]),
    "fhehip_serialization_length": (C.c_int, [_vp, _u64p]),
    "fhehip_serialize": (C.c_int, [_vp, _vp, C.c_uint32, C.c_int, _vp, _vp]),
    "fhehip_deserialize": (C.c_int, [_vp, _vp, C.c_uint32, C.c_int, _vp, _vp]),
    "fhehip_rgsw_external_product": (C.c_int, [_vp, _vp, _vp, _vp, C.c_uint32, _vp]),
    "fhehip_codec_create": (C.c_int, [_vp, _vp, C.c_uint64, _u64p, C.POINTER(C.c_int64), _vpp]),
    "fhehip_codec_destroy": (None, [_vp]),
    "fhehip_encode": (C.c_int, [_vp, _vp, C.c_uint32, C.c_int, C.c_int, _vp, _vp]),
    "fhehip_encrypt_sk": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_uint32, _vp]),
    "fhehip_encrypt_pk": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_uint32, _vp]),
    "fhehip_decrypt": (C.c_int, [_vp, _vp, C.c_uint32, C.c_uint32, C.c_int, _vp, _vp]),
    "fhehip_expand": (C.c_int, [_vp, _vp, C.c_uint32, _vp, _vp]),
    "fhehip_expansion_levels": (C.c_int, [_vp, _u32p]),
    "fhehip_expand_batch": (C.c_int, [_vp, _vp, C.c_uint32, C.c_uint32, _vp, _vp]),
    "fhehip_pir_respond": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_uint32, C.c_uint32, _vp, C.c_uint32, _vp, _vp]),
    "fhehip_dot_product_scalar": (C.c_int, [_vp, _vp, C.c_uint64, _vp, C.c_uint64, C.c_uint32, C.c_uint32, _vp,
                                            _vp]),
    "fhehip_dot_product_columns": (C.c_int, [_vp, _vp, C.c_uint32, C.c_uint32, _vp, C.c_uint32, _vp, _vp]),
    "fhehip_multiply_graph": (C.c_int, [_vp, _vp, _vp, _vp, C.c_uint32, _vpp]),
    "fhehip_multiply_sum_graph": (C.c_int, [_vp, _vp, C.c_uint32, C.c_int, _vp, _vp, _vp, C.c_uint32, _vpp]),
    "fhehip_graph_launch": (C.c_int, [_vp, _vp]),
    "fhehip_graph_destroy": (None, [_vp]),
    "fhehip_submit": (C.c_int, [_vp, C.c_uint32, _vp]),
    "fhehip_generate_moduli": (C.c_int, [_u32p, C.c_uint32, C.c_uint32, _u64p]),
    "fhehip_default_psi": (C.c_int, [C.c_uint32, C.c_uint64, _u64p]),
    "fhehip_ciphertext_to_bytes": (C.c_int, [_vp, C.c_uint32, _vp, C.c_uint32, C.c_char_p, C.c_uint32, C.c_void_p,
                                             C.c_size_t, _szp, _vp]),
    "fhehip_ciphertext_from_bytes": (C.c_int, [_vp, C.c_char_p, C.c_size_t, _vp, C.c_uint32, _u32p, _u32p,
                                               C.c_void_p, C.POINTER(C.c_int), _vp]),
    "fhehip_ksk_from_bytes": (C.c_int, [_vp, C.c_char_p, C.c_size_t, _vpp]),
    "fhehip_ksk_to_bytes": (C.c_int, [_vp, _vp, C.c_int, C.c_void_p, C.c_size_t, _szp]),
    "fhehip_relinearization_key_from_bytes": (C.c_int, [_vp, C.c_char_p, C.c_size_t, _vpp]),
    "fhehip_relinearization_key_to_bytes": (C.c_int, [_vp, _vp, C.c_int, C.c_void_p, C.c_size_t, _szp]),
    "fhehip_galois_key_from_bytes": (C.c_int, [_vp, C.c_char_p, C.c_size_t, _vpp]),
    "fhehip_galois_key_to_bytes": (C.c_int, [_vp, _vp, C.c_int, C.c_void_p, C.c_size_t, _szp]),
    "fhehip_evaluation_key_from_bytes": (C.c_int, [_vp, C.c_char_p, C.c_size_t, _vpp, _u32p]),
    "fhehip_evaluation_key_to_bytes": (C.c_int, [_vp, _vp, C.c_int, C.c_void_p, C.c_size_t, _szp]),
    "fhehip_parameters_from_bytes": (C.c_int, [C.c_char_p, C.c_size_t, _vpp]),
    "fhehip_parameters_to_bytes": (C.c_int, [_vp, C.c_void_p, C.c_size_t, _szp]),
    "fhehip_random_from_seed": (C.c_int, [_u64p, C.c_uint32, C.c_uint32, C.c_char_p, _u64p]),
    "fhehip_generate_c1": (C.c_int, [_u64p, C.c_uint32, C.c_uint32, C.c_char_p, C.c_uint32, _u64p]),
    "fhehip_server_create": (C.c_int, [_vp, _vpp]),
    "fhehip_server_destroy": (None, [_vp]),
    "fhehip_server_add_params": (C.c_int, [_vp, _vp, _vp, C.c_uint32, _u32p, _u64p]),
    "fhehip_server_submit": (C.c_int, [_vp, C.c_uint32, _vp, _vp, _vp, _u64p]),
    "fhehip_server_query": (C.c_int, [_vp, C.c_uint64, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "fhehip_server_drain": (C.c_int, [_vp, C.c_double]),
    "fhehip_server_stats": (C.c_int, [_vp, _vp]),
    "fhehip_profiler_enable": (C.c_int, [C.c_int]),
    "fhehip_profiler_report": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
}


class ScalingFactorC(C.Structure):
    _fields_ = [("num", _u32p), ("num_len", C.c_uint32), ("den", _u32p), ("den_len", C.c_uint32)]


def _words32(x: int):
    w = []
    while x:
        w.append(x & 0xFFFFFFFF)
        x >>= 32
    w = w or [0]
    return (C.c_uint32 * len(w))(*w), len(w)


def scaling_factor(num: int, den: int):
    """Build a fhehip_scaling_factor; keeps the word arrays alive on the struct."""
    n, nl = _words32(num)
    d, dl = _words32(den)
    f = ScalingFactorC(C.cast(n, _u32p), nl, C.cast(d, _u32p), dl)
    f._keep = (n, d)
    return f

_lib = None


class FheHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fhehip error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != OK:
        raise FheHipError(rc, lib().fhehip_last_error().decode())


def _arr64(vals):
    a = (C.c_uint64 * len(vals))(*[int(v) for v in vals])
    return a


def dptr(t):
    """Device pointer of a torch tensor (or an int)."""
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def stream_ptr(stream=None):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return stream if isinstance(stream, int) else stream.cuda_stream


class Context:
    """rq::Context: moduli + degree with one NttOperator per modulus, resident."""

    def __init__(self, moduli, degree, psi=None):
        self.moduli = [int(m) for m in moduli]
        self.degree = int(degree)
        h = C.c_void_p()
        ps = _arr64(psi) if psi is not None else None
        check(lib().fhehip_context_create(_arr64(self.moduli), len(self.moduli), self.degree, ps, C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            _destroy("fhehip_context_destroy", getattr(self, "h", None))
        except Exception:  # module globals already cleared at interpreter exit
            pass

    @property
    def L(self):
        return len(self.moduli)

    def psi(self):
        out = (C.c_uint64 * self.L)()
        check(lib().fhehip_context_psi(self.h, out))
        return [int(x) for x in out]

    def ntt_forward(self, polys, num_polys, lazy=False, stream=None):
        check(lib().fhehip_ntt_forward(self.h, dptr(polys), num_polys, 1 if lazy else 0, stream_ptr(stream)))

    def ntt_autotune(self, rows):
        """Times both NTT shapes for this degree and keeps the faster per
        direction; returns (fwd_log2_values, inv_log2_values, [ms fwd16, fwd32, inv16, inv32])."""
        f, i = C.c_uint32(), C.c_uint32()
        ms = (C.c_float * 4)()
        check(lib().fhehip_ntt_autotune(self.h, rows, C.byref(f), C.byref(i), ms))
        return f.value, i.value, [float(x) for x in ms]

    def ntt_backward(self, polys, num_polys, stream=None):
        check(lib().fhehip_ntt_backward(self.h, dptr(polys), num_polys, stream_ptr(stream)))


# ---- torch helpers ---------------------------------------------------------
def to_device(np_u64, device="cuda"):
    import numpy as np
    import torch
    a = np.ascontiguousarray(np_u64, dtype=np.uint64)
    return torch.from_numpy(a.view(np.int64)).to(device)


def to_host(t):
    import numpy as np
    return t.detach().cpu().numpy().view(np.uint64)


def _message(call):
    """Runs call(out, cap, len_ptr) once for the size, then into a buffer of that size."""
    n = C.c_size_t(0)
    check(call(None, 0, C.byref(n)))
    buf = (C.c_uint8 * max(n.value, 1))()
    check(call(buf, n.value, C.byref(n)))
    return bytes(buf[:n.value])


def empty(n, device="cuda"):
    import torch
    return torch.empty(int(n), dtype=torch.int64, device=device)


def _destroy(name, h):
    if h and _lib is not None:
        try:
            getattr(_lib, name)(h)
        except Exception:  # interpreter teardown: ctypes may already be gone
            pass


class Scaler:
    """rq::scaler::Scaler between two contexts by num/den (None = one)."""

    def __init__(self, from_ctx, to_ctx, num=None, den=None):
        self.from_ctx, self.to_ctx = from_ctx, to_ctx
        h = C.c_void_p()
        f = scaling_factor(num, den) if num is not None else None
        check(lib().fhehip_scaler_create(from_ctx.h, to_ctx.h, C.byref(f) if f is not None else None, C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            _destroy("fhehip_scaler_destroy", getattr(self, "h", None))
        except Exception:  # module globals already cleared at interpreter exit
            pass

    def scale(self, d_in, d_out, num_polys, ntt_repr, stream=None):
        check(lib().fhehip_scaler_scale(self.h, dptr(d_in), dptr(d_out), num_polys, 1 if ntt_repr else 0,
                                        stream_ptr(stream)))


class KeySwitchingKey:
    """KeySwitchingKey from host NTT-domain rows c0, c1 [digits][L_key][N]."""

    def __init__(self, ct_ctx, key_ctx, c0, c1):
        import numpy as np
        c0 = np.ascontiguousarray(c0, dtype=np.uint64)
        c1 = np.ascontiguousarray(c1, dtype=np.uint64)
        digits = c0.shape[0]
        h = C.c_void_p()
        check(lib().fhehip_ksk_create(ct_ctx.h, key_ctx.h, c0.ctypes.data_as(_u64p), c1.ctypes.data_as(_u64p),
                                      digits, C.byref(h)))
        self.h = h
        self.ct_ctx, self.key_ctx = ct_ctx, key_ctx

    def __del__(self):
        try:
            _destroy("fhehip_ksk_destroy", getattr(self, "h", None))
        except Exception:  # module globals already cleared at interpreter exit
            pass

    def key_switch(self, d_p, d_out, num, stream=None):
        check(lib().fhehip_key_switch(self.h, dptr(d_p), dptr(d_out), num, stream_ptr(stream)))

    def relinearize(self, d_ct3, d_out, num, stream=None):
        check(lib().fhehip_relinearize(self.h, dptr(d_ct3), dptr(d_out), num, stream_ptr(stream)))


class Multiplicator:
    """Multiplicator::default(rk) when rk is given, else Multiplicator::new."""

    def __init__(self, rk=None, plaintext=None, base_ctx=None, extended_basis=None, lhs=None, rhs=None,
                 post=None):
        h = C.c_void_p()
        self.rk = rk
        if base_ctx is None:
            check(lib().fhehip_multiplicator_default(rk.h, plaintext, C.byref(h)))
        else:
            self._f = [scaling_factor(*f) if f is not None else None for f in (lhs, rhs, post)]
            refs = [C.byref(f) if f is not None else None for f in self._f]
            check(lib().fhehip_multiplicator_create(base_ctx.h, _arr64(extended_basis), len(extended_basis),
                                                    refs[0], refs[1], refs[2], C.byref(h)))
            if rk is not None:
                check(lib().fhehip_multiplicator_enable_relinearization(h, rk.h))
        self.h = h

    def __del__(self):
        try:
            _destroy("fhehip_multiplicator_destroy", getattr(self, "h", None))
        except Exception:  # module globals already cleared at interpreter exit
            pass

    def extended_basis(self):
        n = C.c_uint32(64)
        out = (C.c_uint64 * 64)()
        check(lib().fhehip_multiplicator_extended_basis(self.h, out, C.byref(n)))
        return [int(out[i]) for i in range(n.value)]

    def multiply(self, d_lhs, d_rhs, d_out, num, stream=None):
        check(lib().fhehip_multiply(self.h, dptr(d_lhs), dptr(d_rhs), dptr(d_out), num, stream_ptr(stream)))

    def multiply_noalloc(self, d_lhs, d_rhs, d_out, num, stream=None):
        """fhehip_multiply_noalloc: scratch in the library's fixed device region."""
        check(lib().fhehip_multiply_noalloc(self.h, dptr(d_lhs), dptr(d_rhs), dptr(d_out), num, stream_ptr(stream)))

    def enable_mod_switching(self, next_ctx):
        """Multiplicator::enable_mod_switching: products land in next_ctx."""
        self.next_ctx = next_ctx
        check(lib().fhehip_multiplicator_enable_mod_switching(self.h, next_ctx.h))


class TuningC(C.Structure):
    """fhehip_tuning of include/fhehip.h."""
    _fields_ = [("xcd_order", C.c_uint32), ("ks_mac_instances", C.c_uint32), ("tensor_log2_values", C.c_uint32),
                ("lift_log2_values", C.c_uint32), ("finish_log2_values", C.c_uint32), ("halves", C.c_uint32),
                ("scaler_mfma", C.c_uint32)]


TUNING_TIMES = ["xcd0", "xcd1", "ks_mac2", "ks_mac4", "ks_mac8", "tensor4", "tensor5", "lift4", "lift5", "finish4",
                "finish5", "scaler_valu", "scaler_mfma", "finish_rows", "finish_halves", "input_rows", "input_halves",
                "tensor_rows", "tensor_halves"]


def tuning_get(degree):
    t = TuningC()
    check(lib().fhehip_tuning_get(degree, C.byref(t)))
    return {k: getattr(t, k) for k, _ in TuningC._fields_}


def tuning_set(degree, **kw):
    """Sets the named knobs (the others keep their current value)."""
    cur = tuning_get(degree)
    cur.update(kw)
    t = TuningC(**cur)
    check(lib().fhehip_tuning_set(degree, C.byref(t)))


def multiply_autotune(mult, num):
    """fhehip_multiply_autotune: (chosen knobs, {candidate: ms per multiply})."""
    t = TuningC()
    ms = (C.c_float * len(TUNING_TIMES))()
    check(lib().fhehip_multiply_autotune(mult.h, num, C.byref(t), ms))
    return {k: getattr(t, k) for k, _ in TuningC._fields_}, dict(zip(TUNING_TIMES, [float(x) for x in ms]))


class Parameters:
    """BfvParameters: level chain and per-level multiplication parameters."""

    def __init__(self, degree, plaintext, moduli):
        self.moduli = [int(q) for q in moduli]
        h = C.c_void_p()
        check(lib().fhehip_parameters_create(degree, plaintext, _arr64(self.moduli), len(self.moduli), C.byref(h)))
        self.h = h
        mx, n, t = C.c_uint32(), C.c_uint32(), C.c_uint64()
        check(lib().fhehip_parameters_info(self.h, C.byref(mx), C.byref(n), C.byref(t)))
        self.max_level, self.degree, self.plaintext = mx.value, n.value, t.value

    def __del__(self):
        _destroy("fhehip_parameters_destroy", getattr(self, "h", None))

    def _ctx(self, h):
        """A Context object over a handle the library returned (every context
        of the parameters holds a prefix of their moduli)."""
        c = Context.__new__(Context)
        c.h = h
        L, n = C.c_uint32(), C.c_uint32()
        check(lib().fhehip_context_info(h, C.byref(L), C.byref(n)))
        c.moduli = self.moduli[:L.value]
        c.degree = n.value
        return c

    def context(self, level):
        h = C.c_void_p()
        check(lib().fhehip_parameters_context(self.h, level, C.byref(h)))
        return self._ctx(h)

    def plaintext_context(self):
        h = C.c_void_p()
        check(lib().fhehip_parameters_plaintext_context(self.h, C.byref(h)))
        return self._ctx(h)

    @classmethod
    def from_bytes(cls, data: bytes):
        """BfvParameters::try_deserialize (fhehip_parameters_from_bytes)."""
        h = C.c_void_p()
        check(lib().fhehip_parameters_from_bytes(data, len(data), C.byref(h)))
        par = cls.__new__(cls)
        par.h = h
        mx, n, t = C.c_uint32(), C.c_uint32(), C.c_uint64()
        check(lib().fhehip_parameters_info(h, C.byref(mx), C.byref(n), C.byref(t)))
        par.max_level, par.degree, par.plaintext = mx.value, n.value, t.value
        par.moduli = par.mul_basis(0)[:par.max_level + 1]  # the level-0 moduli lead the basis
        return par

    def to_bytes(self):
        return _message(lambda out, cap, ln: lib().fhehip_parameters_to_bytes(self.h, out, cap, ln))

    def ciphertext_to_bytes(self, level, d_ct, parts, seed=None, variable_time_mask=0, stream=None):
        """Ciphertext::to_bytes of d_ct [parts][L][N] at `level`."""
        if seed is not None and len(seed) != 32:
            raise ValueError("a seed has 32 bytes")
        return _message(lambda out, cap, ln: lib().fhehip_ciphertext_to_bytes(
            self.h, level, dptr(d_ct), parts, seed, variable_time_mask, out, cap, ln, stream_ptr(stream)))

    def ciphertext_from_bytes(self, data: bytes, stream=None):
        """Ciphertext::from_bytes -> (device [parts][L][N], parts, level, seed or None)."""
        parts, level, seeded = C.c_uint32(), C.c_uint32(), C.c_int()
        seed = (C.c_uint8 * 32)()
        check(lib().fhehip_ciphertext_from_bytes(self.h, data, len(data), None, 0, C.byref(parts), C.byref(level),
                                                 seed, C.byref(seeded), None))
        L = len(self.moduli) - level.value
        out = empty(parts.value * L * self.degree)
        check(lib().fhehip_ciphertext_from_bytes(self.h, data, len(data), dptr(out), parts.value, C.byref(parts),
                                                 C.byref(level), None, None, stream_ptr(stream)))
        return out, parts.value, level.value, (bytes(seed) if seeded.value else None)

    def _ksk(self, fn, data):
        h = C.c_void_p()
        check(getattr(lib(), fn)(self.h, data, len(data), C.byref(h)))
        k = KeySwitchingKey.__new__(KeySwitchingKey)
        k.h = h
        k.ct_ctx = k.key_ctx = None
        return k

    def ksk_from_bytes(self, data: bytes):
        return self._ksk("fhehip_ksk_from_bytes", data)

    def relinearization_key_from_bytes(self, data: bytes):
        return self._ksk("fhehip_relinearization_key_from_bytes", data)

    def ksk_to_bytes(self, ksk, variable_time=True):
        return _message(lambda out, cap, ln: lib().fhehip_ksk_to_bytes(self.h, ksk.h, int(variable_time), out, cap, ln))

    def relinearization_key_to_bytes(self, ksk, variable_time=True):
        return _message(lambda out, cap, ln: lib().fhehip_relinearization_key_to_bytes(
            self.h, ksk.h, int(variable_time), out, cap, ln))

    def galois_key_from_bytes(self, data: bytes):
        h = C.c_void_p()
        check(lib().fhehip_galois_key_from_bytes(self.h, data, len(data), C.byref(h)))
        g = GaloisKey.__new__(GaloisKey)
        g.h, g.ksk, g.exponent = h, None, None
        return g

    def galois_key_to_bytes(self, g, variable_time=True):
        return _message(lambda out, cap, ln: lib().fhehip_galois_key_to_bytes(
            self.h, g.h, int(variable_time), out, cap, ln))

    def evaluation_key_from_bytes(self, data: bytes):
        h = C.c_void_p()
        lvl = C.c_uint32()
        check(lib().fhehip_evaluation_key_from_bytes(self.h, data, len(data), C.byref(h), C.byref(lvl)))
        ek = EvaluationKey.__new__(EvaluationKey)
        ek.h, ek.keys = h, []
        return ek, lvl.value

    def evaluation_key_to_bytes(self, ek, variable_time=True):
        return _message(lambda out, cap, ln: lib().fhehip_evaluation_key_to_bytes(
            self.h, ek.h, int(variable_time), out, cap, ln))

    def mul_basis(self, level):
        n = C.c_uint32(128)
        out = (C.c_uint64 * 128)()
        check(lib().fhehip_parameters_mul_basis(self.h, level, out, C.byref(n)))
        return [int(out[i]) for i in range(n.value)]

    def ct_mul(self, level, d_lhs, lhs_parts, d_rhs, rhs_parts, d_out, num, stream=None):
        check(lib().fhehip_ciphertext_mul(self.h, level, dptr(d_lhs), lhs_parts, dptr(d_rhs), rhs_parts, dptr(d_out),
                                          num, stream_ptr(stream)))


class Switcher:
    """Ciphertext::switch_down from `from_ctx` to `to_ctx` (its first L - 1 moduli)."""

    def __init__(self, from_ctx, to_ctx):
        self.from_ctx, self.to_ctx = from_ctx, to_ctx
        h = C.c_void_p()
        check(lib().fhehip_switcher_create(from_ctx.h, to_ctx.h, C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            _destroy("fhehip_switcher_destroy", getattr(self, "h", None))
        except Exception:  # module globals already cleared at interpreter exit
            pass

    def switch_down(self, d_in, d_out, num_polys, stream=None):
        check(lib().fhehip_switch_down(self.h, dptr(d_in), dptr(d_out), num_polys, stream_ptr(stream)))


def serialization_length(ctx):
    v = C.c_uint64()
    check(lib().fhehip_serialization_length(ctx.h, C.byref(v)))
    return v.value


def serialize(ctx, d_polys, num, d_bytes, ntt=True, stream=None):
    """Rq coefficient payloads of num polynomials (d_bytes: a uint8 device buffer)."""
    check(lib().fhehip_serialize(ctx.h, dptr(d_polys), num, 1 if ntt else 0, dptr(d_bytes), stream_ptr(stream)))


def deserialize(ctx, d_bytes, num, d_polys, ntt=True, stream=None):
    check(lib().fhehip_deserialize(ctx.h, dptr(d_bytes), num, 1 if ntt else 0, dptr(d_polys), stream_ptr(stream)))


def pir_respond(ek, mult, rk, switcher, d_db, dim1, dim2, d_queries, num, d_responses, stream=None):
    """fhehip_pir_respond: the MulPIR server response for num queries."""
    check(lib().fhehip_pir_respond(ek.h, mult.h, rk.h, switcher.h, dptr(d_db), dim1, dim2, dptr(d_queries), num,
                                   dptr(d_responses), stream_ptr(stream)))


def rgsw_external_product(ksk0, ksk1, d_ct, d_out, num, stream=None):
    """Ciphertext x RGSWCiphertext with the RGSW ciphertext's two KeySwitchingKeys."""
    check(lib().fhehip_rgsw_external_product(ksk0.h, ksk1.h, dptr(d_ct), dptr(d_out), num, stream_ptr(stream)))


class Codec:
    """Encoding, secret-key encryption and decryption at the level of `ctx`
    (fhehip_codec_*); sk: the SecretKey's signed coefficients or None."""

    def __init__(self, ctx, plain_ctx, t, sk=None, psi_t=None):
        import numpy as np
        h = C.c_void_p()
        self._sk = None if sk is None else np.ascontiguousarray(sk, dtype=np.int64)
        skp = None if self._sk is None else self._sk.ctypes.data_as(C.POINTER(C.c_int64))
        psi = None if psi_t is None else _arr64([psi_t])
        check(lib().fhehip_codec_create(ctx.h, plain_ctx.h, t, psi, skp, C.byref(h)))
        self.h = h
        self.ctx, self.plain_ctx = ctx, plain_ctx

    def __del__(self):
        try:
            _destroy("fhehip_codec_destroy", getattr(self, "h", None))
        except Exception:  # module globals already cleared at interpreter exit
            pass

    def encode(self, d_values, num, d_out, simd=True, scaled=True, stream=None):
        check(lib().fhehip_encode(self.h, dptr(d_values), num, 1 if simd else 0, 1 if scaled else 0, dptr(d_out),
                                  stream_ptr(stream)))

    def encrypt(self, d_pt, d_a, d_e, d_ct, num, stream=None):
        check(lib().fhehip_encrypt_sk(self.h, dptr(d_pt), dptr(d_a), dptr(d_e), dptr(d_ct), num,
                                      stream_ptr(stream)))

    def encrypt_pk(self, d_pk, d_pt, d_u, d_e1, d_e2, d_ct, num, stream=None):
        check(lib().fhehip_encrypt_pk(self.h, dptr(d_pk), dptr(d_pt), dptr(d_u), dptr(d_e1), dptr(d_e2), dptr(d_ct),
                                      num, stream_ptr(stream)))

    def decrypt(self, d_ct, parts, num, d_values, simd=True, stream=None):
        check(lib().fhehip_decrypt(self.h, dptr(d_ct), parts, num, 1 if simd else 0, dptr(d_values),
                                   stream_ptr(stream)))


class GaloisKey:
    def __init__(self, ksk, exponent):
        h = C.c_void_p()
        check(lib().fhehip_galois_create(ksk.h, exponent, C.byref(h)))
        self.h = h
        self.ksk = ksk
        self.exponent = exponent

    def __del__(self):
        try:
            _destroy("fhehip_galois_destroy", getattr(self, "h", None))
        except Exception:  # module globals already cleared at interpreter exit
            pass

    def relinearize(self, d_ct, d_out, num, stream=None):
        check(lib().fhehip_galois_relinearize(self.h, dptr(d_ct), dptr(d_out), num, stream_ptr(stream)))


class EvaluationKey:
    """EvaluationKey over GaloisKey objects (shares ownership of them)."""

    def __init__(self, galois_keys):
        arr = (C.c_void_p * len(galois_keys))(*[g.h.value for g in galois_keys])
        h = C.c_void_p()
        check(lib().fhehip_evaluation_key_create(arr, len(galois_keys), C.byref(h)))
        self.h = h
        self.keys = list(galois_keys)

    def __del__(self):
        try:
            _destroy("fhehip_evaluation_key_destroy", getattr(self, "h", None))
        except Exception:  # module globals already cleared at interpreter exit
            pass

    def rotate_columns(self, d_ct, d_out, step, num, stream=None):
        check(lib().fhehip_rotate_columns(self.h, dptr(d_ct), dptr(d_out), step, num, stream_ptr(stream)))

    def rotate_rows(self, d_ct, d_out, num, stream=None):
        check(lib().fhehip_rotate_rows(self.h, dptr(d_ct), dptr(d_out), num, stream_ptr(stream)))

    def inner_sum(self, d_ct, d_out, num, stream=None):
        check(lib().fhehip_inner_sum(self.h, dptr(d_ct), dptr(d_out), num, stream_ptr(stream)))

    def rotate_and_sum(self, d_ct, d_out, log2_slots, rows, num, stream=None):
        check(lib().fhehip_rotate_and_sum(self.h, dptr(d_ct), dptr(d_out), log2_slots, 1 if rows else 0, num,
                                          stream_ptr(stream)))

    def expand(self, d_ct, size, d_out, stream=None):
        """EvaluationKey::expands: d_out [max(size, 1)][2][L][N]."""
        check(lib().fhehip_expand(self.h, dptr(d_ct), size, dptr(d_out), stream_ptr(stream)))

    def expand_batch(self, d_cts, num, size, d_out, stream=None):
        """fhehip_expand_batch: d_out [max(size, 1)][num][2][L][N]."""
        check(lib().fhehip_expand_batch(self.h, dptr(d_cts), num, size, dptr(d_out), stream_ptr(stream)))

    def expansion_levels(self):
        v = C.c_uint32()
        check(lib().fhehip_expansion_levels(self.h, C.byref(v)))
        return v.value


def split_row_limit(max_rows):
    """fhehip_split_row_limit: sets the half-row launch limit, returns the old one."""
    prev = C.c_uint32()
    check(lib().fhehip_split_row_limit(max_rows, C.byref(prev)))
    return prev.value


def ntt_set_shape(degree, fwd, inv):
    """log2 values per thread (4 or 5) of the plain NTT kernels for `degree`."""
    check(lib().fhehip_ntt_set_shape(degree, fwd, inv))


def ciphertext_add(ctx, d_a, d_b, d_out, num_polys, stream=None):
    check(lib().fhehip_ciphertext_add(ctx.h, dptr(d_a), dptr(d_b), dptr(d_out), num_polys, stream_ptr(stream)))


def ciphertext_sub(ctx, d_a, d_b, d_out, num_polys, stream=None):
    check(lib().fhehip_ciphertext_sub(ctx.h, dptr(d_a), dptr(d_b), dptr(d_out), num_polys, stream_ptr(stream)))


def ciphertext_neg(ctx, d_a, d_out, num_polys, stream=None):
    check(lib().fhehip_ciphertext_neg(ctx.h, dptr(d_a), dptr(d_out), num_polys, stream_ptr(stream)))


def ciphertext_mul_plain(ctx, d_ct, parts, num, d_pt, pt_stride, d_out, stream=None):
    check(lib().fhehip_ciphertext_mul_plain(ctx.h, dptr(d_ct), parts, num, dptr(d_pt), pt_stride, dptr(d_out),
                                            stream_ptr(stream)))


def dot_product_scalar(ctx, d_ct, ct_stride, d_pt, pt_stride, count, parts, d_out, stream=None):
    """dot_product_scalar over `count` (ciphertext, plaintext) pairs; strides in words."""
    check(lib().fhehip_dot_product_scalar(ctx.h, dptr(d_ct), ct_stride, dptr(d_pt), pt_stride, count, parts,
                                          dptr(d_out), stream_ptr(stream)))


def dot_product_columns(ctx, d_ct, rows, parts, d_db, columns, d_out, stream=None):
    """Every column of a [rows][columns][L][N] database against `rows` ciphertexts."""
    check(lib().fhehip_dot_product_columns(ctx.h, dptr(d_ct), rows, parts, dptr(d_db), columns, dptr(d_out),
                                           stream_ptr(stream)))


class Graph:
    """A captured HIP graph: multiply (and optionally the slot sum) on fixed buffers."""

    def __init__(self, mult, d_lhs, d_rhs, d_out, num, ek=None, log2_slots=None, rows=False):
        h = C.c_void_p()
        if ek is None:
            check(lib().fhehip_multiply_graph(mult.h, dptr(d_lhs), dptr(d_rhs), dptr(d_out), num, C.byref(h)))
        else:
            check(lib().fhehip_multiply_sum_graph(mult.h, ek.h, log2_slots, 1 if rows else 0, dptr(d_lhs),
                                                  dptr(d_rhs), dptr(d_out), num, C.byref(h)))
        self.h = h
        self._keep = (mult, ek, d_lhs, d_rhs, d_out)

    def launch(self, stream=None):
        check(lib().fhehip_graph_launch(self.h, stream_ptr(stream)))

    def close(self):
        if getattr(self, "h", None):
            lib().fhehip_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


JOB_MAX_ROTATIONS = 4


class JobC(C.Structure):
    """fhehip_job of include/fhehip.h."""
    _fields_ = [("mult", _vp), ("ek", _vp), ("d_lhs", _vp), ("d_rhs", _vp), ("d_out", _vp),
                ("num_rotations", C.c_uint32), ("steps", C.c_uint32 * JOB_MAX_ROTATIONS)]


def make_jobs(specs):
    """specs: iterable of (Multiplicator, EvaluationKey or None, lhs, rhs, out, steps)
    with device tensors or pointers; returns a ctypes array of fhehip_job."""
    specs = list(specs)
    arr = (JobC * max(len(specs), 1))()
    for k, (m, ek, a, b, o, steps) in enumerate(specs):
        steps = list(steps or [])
        if len(steps) > JOB_MAX_ROTATIONS:
            raise ValueError("too many rotations in a job")
        arr[k].mult = m.h.value
        arr[k].ek = ek.h.value if ek is not None else None
        arr[k].d_lhs = dptr(a)
        arr[k].d_rhs = dptr(b)
        arr[k].d_out = dptr(o)
        arr[k].num_rotations = len(steps)
        for r, st in enumerate(steps):
            arr[k].steps[r] = st
    return arr, len(specs)


def submit(specs, stream=None):
    """fhehip_submit over (Multiplicator, EvaluationKey, lhs, rhs, out, steps) jobs."""
    arr, n = specs if isinstance(specs, tuple) else make_jobs(specs)
    check(lib().fhehip_submit(arr, n, stream_ptr(stream)))


class ServerConfigC(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_batch", C.c_uint32), ("lanes", C.c_uint32)]


class ServerStatsC(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("batches", C.c_uint64), ("peak_bytes", C.c_uint64), ("max_batch_seen", C.c_uint32)]


class Server:
    """Continuous-batching request server (fhehip_server_*)."""

    def __init__(self, hbm_budget_bytes=0, max_batch=64, lanes=4):
        cfg = ServerConfigC(hbm_budget_bytes, max_batch, lanes)
        h = C.c_void_p()
        check(lib().fhehip_server_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self._keep = []

    def close(self):
        if getattr(self, "h", None):
            lib().fhehip_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_params(self, mult, ek, log2_slots):
        """Returns (id, footprint_bytes)."""
        i = C.c_uint32()
        fp = C.c_uint64()
        check(lib().fhehip_server_add_params(self.h, mult.h, ek.h, log2_slots, C.byref(i), C.byref(fp)))
        self._keep.append((mult, ek))
        return i.value, fp.value

    def submit(self, params, d_lhs, d_rhs, d_out):
        t = C.c_uint64()
        check(lib().fhehip_server_submit(self.h, params, dptr(d_lhs), dptr(d_rhs), dptr(d_out), C.byref(t)))
        return t.value

    def query(self, ticket):
        """(state, latency_us): state 0 queued, 1 running, 2 done, -1 failed."""
        st = C.c_int()
        lat = C.c_double()
        check(lib().fhehip_server_query(self.h, ticket, C.byref(st), C.byref(lat)))
        return st.value, lat.value

    def drain(self, timeout_s=60.0):
        check(lib().fhehip_server_drain(self.h, timeout_s))

    def stats(self):
        s = ServerStatsC()
        check(lib().fhehip_server_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStatsC._fields_}


def generate_moduli(sizes, degree):
    """BfvParametersBuilder::generate_moduli on the library side."""
    sz = (C.c_uint32 * max(len(sizes), 1))(*sizes)
    out = (C.c_uint64 * max(len(sizes), 1))()
    check(lib().fhehip_generate_moduli(sz, len(sizes), degree, out))
    return [int(out[i]) for i in range(len(sizes))]


def default_psi(degree, modulus):
    """fhehip_default_psi: NttOperator::new's root of `modulus` for `degree`."""
    out = C.c_uint64(0)
    check(lib().fhehip_default_psi(degree, modulus, C.byref(out)))
    return int(out.value)


def _moduli_arg(moduli):
    m = (C.c_uint64 * len(moduli))(*[int(q) for q in moduli])
    return C.cast(m, _u64p), m


def random_from_seed(moduli, degree, seed: bytes):
    """fhehip_random_from_seed -> numpy [L * degree] (host)."""
    import numpy as np
    if len(seed) != 32:
        raise ValueError("a seed has 32 bytes")
    out = np.zeros(len(moduli) * degree, dtype=np.uint64)
    mp, _keep = _moduli_arg(moduli)
    check(lib().fhehip_random_from_seed(mp, len(moduli), degree, seed, out.ctypes.data_as(_u64p)))
    return out


def generate_c1(moduli, degree, seed: bytes, count):
    """fhehip_generate_c1 -> numpy [count * L * degree] (host)."""
    import numpy as np
    if len(seed) != 32:
        raise ValueError("a seed has 32 bytes")
    out = np.zeros(count * len(moduli) * degree, dtype=np.uint64)
    mp, _keep = _moduli_arg(moduli)
    check(lib().fhehip_generate_c1(mp, len(moduli), degree, seed, count, out.ctypes.data_as(_u64p)))
    return out


def profiler_enable(on=True):
    check(lib().fhehip_profiler_enable(1 if on else 0))


def profiler_report():
    """{kernel: (launches, total_ms)} accumulated since profiler_enable()."""
    need = C.c_size_t(0)
    check(lib().fhehip_profiler_report(None, 0, C.byref(need)))
    buf = C.create_string_buffer(need.value + 1)
    check(lib().fhehip_profiler_report(buf, len(buf), C.byref(need)))
    out = {}
    for line in buf.value.decode().splitlines():
        name, calls, ms = line.split()
        out[name] = (int(calls), float(ms))
    return out
