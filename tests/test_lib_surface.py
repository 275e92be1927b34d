"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
public header declares, and reports errors without a GPU call."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    src = open(os.path.join(ROOT, "include", "fhehip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fhehip_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_header_symbols():
    import ctypes
    import fhehip
    L = ctypes.CDLL(fhehip.LIB_PATH)
    syms = _header_symbols()
    assert len(syms) >= 7
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # the Python binding declares exactly the header surface
    assert sorted(fhehip.SIGNATURES) == syms


def test_version_and_null_checks():
    import ctypes
    import fhehip
    lib = fhehip.lib()
    assert b"gfx950" in lib.fhehip_version()
    h = ctypes.c_void_p()
    rc = lib.fhehip_context_create(None, 0, 16, None, ctypes.byref(h))
    assert rc == fhehip.ERR_INVALID_ARG
    assert lib.fhehip_last_error()


def test_generate_moduli_matches_oracle_and_config_lists(oracle):
    """fhehip_generate_moduli is host-only; it reproduces the oracle's
    generate_moduli (pinned by the reference's KATs) and the configuration
    moduli lists."""
    import fhehip
    cfg3 = [1152921504606748673, 1152921504606683137, 1152921504606584833, 1152921504605962241,
            1152921504604979201, 1152921504600260609, 1152921504599080961]
    assert fhehip.generate_moduli([60] * 7, 16384) == cfg3
    cfg1 = [4611686018427322369, 4611686018427289601, 4611686018426454017]
    assert fhehip.generate_moduli([62] * 3, 8192) == cfg1
    for sizes, n in [([50, 55, 60, 62], 4096), ([62, 50, 62, 50, 40], 32768), ([20, 20], 8)]:
        assert fhehip.generate_moduli(sizes, n) == [int(x) for x in oracle.generate_moduli(sizes, n)]
    with pytest.raises(fhehip.FheHipError):
        fhehip.generate_moduli([63], 1024)
    with pytest.raises(fhehip.FheHipError):
        fhehip.generate_moduli([60], 1000)


def test_counter_summaries_name_their_sources():
    """bench.py reports HBM traffic and VALU figures only from counter
    summaries measured on the current kernel sources: each summary's
    `Kernel sources:` line must parse to a digest of the form bench.py
    computes, and the traffic summary must hold the dominant kernel's row."""
    import sys
    sys.path.insert(0, ROOT)
    import bench
    here = bench.kernel_sources_digest()
    assert re.fullmatch(r"[0-9a-f]{16}", here)
    for name in (bench.TRAFFIC_PROFILE, bench.VALU_PROFILE, "kernels_mul_relin_n16384_b512.md"):
        rows, digest = bench.load_profile(name)
        assert digest is not None and re.fullmatch(r"[0-9a-f]{16}", digest), (name, digest)
    rows, digest = bench.load_profile(bench.TRAFFIC_PROFILE)
    assert bench.load_traffic("tensor_intt", digest) > 0
