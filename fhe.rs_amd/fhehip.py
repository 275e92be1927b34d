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
}

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
        if getattr(self, "h", None) and _lib is not None:
            _lib.fhehip_context_destroy(self.h)
            self.h = None

    @property
    def L(self):
        return len(self.moduli)

    def psi(self):
        out = (C.c_uint64 * self.L)()
        check(lib().fhehip_context_psi(self.h, out))
        return [int(x) for x in out]

    def ntt_forward(self, polys, num_polys, lazy=False, stream=None):
        check(lib().fhehip_ntt_forward(self.h, dptr(polys), num_polys, 1 if lazy else 0, stream_ptr(stream)))

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


def empty(n, device="cuda"):
    import torch
    return torch.empty(int(n), dtype=torch.int64, device=device)
