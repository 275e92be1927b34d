"""GPU parity of the protobuf messages (fhe/src/proto/bfv.proto,
math/src/proto/rq.proto) written and read by the library.

The independent encoder is the protobuf runtime of this image (google.protobuf)
over message descriptors built here with the schemas' field numbers and
types; its bytes are compared with the library's, and messages it writes
(fields in any order, unknown fields) are read by the library. Coefficient
payloads are the oracle's (Poly -> Rq, convert.rs:17-42), seeded polynomials
the oracle's random_from_seed / generate_c1. The objects read are then used
(multiply, relinearize, rotate) and compared with the oracle bit for bit.
"""
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


@pytest.fixture(scope="module")
def pbm():
    """Message classes for the schemas (field numbers and types of bfv.proto / rq.proto)."""
    from google.protobuf import descriptor_pb2, descriptor_pool, message_factory
    F = descriptor_pb2.FieldDescriptorProto
    fd = descriptor_pb2.FileDescriptorProto(name="fhers_wire_check.proto", package="wirecheck", syntax="proto3")
    en = fd.enum_type.add(name="Representation")
    for name, v in (("UNKNOWN", 0), ("POWERBASIS", 1), ("NTT", 2), ("NTTSHOUP", 3)):
        en.value.add(name=name, number=v)

    def msg(name, fields, oneofs=()):
        m = fd.message_type.add(name=name)
        for o in oneofs:
            m.oneof_decl.add(name=o)
        for fname, num, ftype, rep, tname, oneof in fields:
            f = m.field.add(name=fname, number=num, type=ftype,
                            label=F.LABEL_REPEATED if rep else F.LABEL_OPTIONAL)
            if tname:
                f.type_name = ".wirecheck." + tname
            if oneof is not None:
                f.oneof_index = oneof

    msg("Rq", [("representation", 1, F.TYPE_ENUM, False, "Representation", None),
               ("degree", 2, F.TYPE_UINT32, False, None, None),
               ("coefficients", 3, F.TYPE_BYTES, False, None, None),
               ("allow_variable_time", 4, F.TYPE_BOOL, False, None, None)])
    msg("Ciphertext", [("c", 1, F.TYPE_BYTES, True, None, None), ("seed", 2, F.TYPE_BYTES, False, None, None),
                       ("level", 3, F.TYPE_UINT32, False, None, None)])
    msg("KeySwitchingKey", [("c0", 1, F.TYPE_BYTES, True, None, None), ("c1", 2, F.TYPE_BYTES, True, None, None),
                            ("seed", 3, F.TYPE_BYTES, False, None, None),
                            ("ciphertext_level", 4, F.TYPE_UINT32, False, None, None),
                            ("ksk_level", 5, F.TYPE_UINT32, False, None, None),
                            ("log_base", 6, F.TYPE_UINT32, False, None, None)])
    msg("RelinearizationKey", [("ksk", 1, F.TYPE_MESSAGE, False, "KeySwitchingKey", None)])
    msg("GaloisKey", [("ksk", 1, F.TYPE_MESSAGE, False, "KeySwitchingKey", None),
                      ("exponent", 2, F.TYPE_UINT32, False, None, None)])
    msg("EvaluationKey", [("gk", 2, F.TYPE_MESSAGE, True, "GaloisKey", None),
                          ("ciphertext_level", 3, F.TYPE_UINT32, False, None, None),
                          ("evaluation_key_level", 4, F.TYPE_UINT32, False, None, None)])
    msg("Parameters", [("degree", 1, F.TYPE_UINT32, False, None, None),
                       ("moduli", 2, F.TYPE_UINT64, True, None, None),
                       ("plaintext", 3, F.TYPE_UINT64, False, None, 0),
                       ("variance", 4, F.TYPE_UINT32, False, None, None),
                       ("plaintext_big", 5, F.TYPE_BYTES, False, None, 0)], oneofs=("plaintext_modulus",))
    pool = descriptor_pool.DescriptorPool()
    pool.Add(fd)

    class M:
        pass
    for name in ("Rq", "Ciphertext", "KeySwitchingKey", "RelinearizationKey", "GaloisKey", "EvaluationKey",
                 "Parameters"):
        setattr(M, name, message_factory.GetMessageClass(pool.FindMessageTypeByName("wirecheck." + name)))
    return M


def _sync():
    import torch
    torch.cuda.synchronize()


def _rq(pbm, oracle, moduli, n, ntt_poly, repr_, vt):
    """The Rq message of an NTT-domain polynomial (Poly -> Rq, convert.rs:17-42)."""
    payload = oracle.poly_serialize(oracle.Context(moduli, n), ntt_poly, 1, True).tobytes()
    return pbm.Rq(representation=repr_, degree=n, coefficients=payload, allow_variable_time=vt).SerializeToString()


N, SIZES = 4096, [55, 60, 62, 50]


@pytest.fixture(scope="module")
def setup(fh, oracle):
    moduli = oracle.generate_moduli(SIZES, N)
    bfv = oracle.Bfv(N, T, moduli, sk_seed=5)
    par = fh.Parameters(N, T, moduli)
    return moduli, bfv, par


def test_ciphertext_messages(fh, oracle, pbm, setup):
    moduli, bfv, par = setup
    rng = random.Random(3)
    for level in (0, 1, 3):
        lm = bfv.level_moduli(level)
        w = len(lm) * N
        v = [rng.randrange(T) for _ in range(N)]
        ct = bfv.encrypt(v, level=level, seed=level + 1)
        ct3 = bfv.ct_mul_parts(ct, ct, level)
        for c, parts, mask in ((ct, 2, 0b10), (ct3, 3, 0b101)):
            got = par.ciphertext_to_bytes(level, fh.to_device(c), parts, variable_time_mask=mask)
            exp = pbm.Ciphertext(c=[_rq(pbm, oracle, lm, N, c[i * w:(i + 1) * w], 2, bool((mask >> i) & 1))
                                    for i in range(parts)], level=level).SerializeToString()
            assert got == exp, (level, parts)
            d, p, lv, seed = par.ciphertext_from_bytes(exp)
            _sync()
            assert (p, lv, seed) == (parts, level, None)
            assert np.array_equal(fh.to_host(d), c)
    # the same ciphertext with its fields in another order and an unknown field
    lm = bfv.level_moduli(1)
    w = len(lm) * N
    ct = bfv.encrypt([1] * N, level=1, seed=9)
    m = pbm.Ciphertext(c=[_rq(pbm, oracle, lm, N, ct[i * w:(i + 1) * w], 2, False) for i in range(2)], level=1)
    raw = pbm.Ciphertext(level=1).SerializeToString() + b"\xa8\x06\x05" + pbm.Ciphertext(c=m.c).SerializeToString()
    d, p, lv, _ = par.ciphertext_from_bytes(raw)
    _sync()
    assert (p, lv) == (2, 1) and np.array_equal(fh.to_host(d), ct)


def test_seeded_ciphertext(fh, oracle, pbm, setup):
    moduli, bfv, par = setup
    level = 1
    lm = bfv.level_moduli(level)
    w = len(lm) * N
    seed = bytes(range(40, 72))
    c1 = oracle.random_from_seed(lm, N, seed)
    c0 = oracle.random_polys(lm, N, 1, seed=4)
    ct = np.concatenate([c0, c1])
    got = par.ciphertext_to_bytes(level, fh.to_device(ct), 2, seed=seed)
    exp = pbm.Ciphertext(c=[_rq(pbm, oracle, lm, N, c0, 2, False)], seed=seed, level=level).SerializeToString()
    assert got == exp
    d, p, lv, s = par.ciphertext_from_bytes(got)
    _sync()
    assert (p, lv, s) == (2, level, seed)
    assert np.array_equal(fh.to_host(d), ct)


def test_ciphertext_message_errors(fh, oracle, pbm, setup):
    moduli, bfv, par = setup
    lm = bfv.level_moduli(0)
    w = len(lm) * N
    ct = bfv.encrypt([0] * N, level=0, seed=2)
    good = [_rq(pbm, oracle, lm, N, ct[i * w:(i + 1) * w], 2, False) for i in range(2)]
    pb_repr = _rq(pbm, oracle, lm, N, ct[:w], 1, False)
    cases = {
        "one polynomial": pbm.Ciphertext(c=good[:1]).SerializeToString(),
        "no polynomial": pbm.Ciphertext(seed=bytes(32)).SerializeToString(),
        "power basis": pbm.Ciphertext(c=[pb_repr, good[1]]).SerializeToString(),
        "seed of 31 bytes": pbm.Ciphertext(c=good[:1], seed=bytes(31)).SerializeToString(),
        "level above max": pbm.Ciphertext(c=good, level=4).SerializeToString(),
        "level mismatch": pbm.Ciphertext(c=good, level=1).SerializeToString(),
        "truncated": pbm.Ciphertext(c=good).SerializeToString()[:-7],
        "degree": pbm.Ciphertext(c=[pbm.Rq(representation=2, degree=N // 2,
                                           coefficients=pbm.Rq.FromString(good[0]).coefficients).SerializeToString(),
                                    good[1]]).SerializeToString(),
    }
    for name, raw in cases.items():
        with pytest.raises(fh.FheHipError):
            par.ciphertext_from_bytes(raw)
            pytest.fail(name)


def _ksk_msg(pbm, oracle, key_moduli, rows0, rows1, ct_level, key_level, seed=None):
    c0 = [_rq(pbm, oracle, key_moduli, N, r.reshape(-1), 3, True) for r in rows0]
    c1 = [] if seed else [_rq(pbm, oracle, key_moduli, N, r.reshape(-1), 3, True) for r in rows1]
    return pbm.KeySwitchingKey(c0=c0, c1=c1, seed=seed or b"", ciphertext_level=ct_level, ksk_level=key_level)


def test_relinearization_key_message(fh, oracle, pbm, setup):
    moduli, bfv, par = setup
    ct_level, key_level = 1, 0
    rk = bfv.relin_key(ct_level=ct_level, key_level=key_level, seed=8)
    r0, r1 = rk.export()
    km = _ksk_msg(pbm, oracle, bfv.level_moduli(key_level), r0, r1, ct_level, key_level)
    raw = pbm.RelinearizationKey(ksk=km).SerializeToString()
    k = par.relinearization_key_from_bytes(raw)
    assert par.relinearization_key_to_bytes(k) == raw
    assert par.ksk_to_bytes(k) == km.SerializeToString()
    mult = fh.Multiplicator(rk=k, plaintext=T)
    om = oracle.Multiplicator(bfv, rk, level=ct_level)
    lhs = np.concatenate([bfv.encrypt([3] * N, level=ct_level, seed=s) for s in (1, 2)])
    rhs = np.concatenate([bfv.encrypt([5] * N, level=ct_level, seed=s) for s in (3, 4)])
    out = fh.empty(lhs.size)
    mult.multiply(fh.to_device(lhs), fh.to_device(rhs), out, 2)
    _sync()
    got = fh.to_host(out)
    assert np.array_equal(got, om.multiply(lhs, rhs))
    assert [int(x) for x in bfv.decrypt(got[:got.size // 2], level=ct_level)] == [15] * N
    with pytest.raises(fh.FheHipError):
        par.relinearization_key_from_bytes(pbm.RelinearizationKey().SerializeToString())
    with pytest.raises(fh.FheHipError):  # one c0 row short
        par.ksk_from_bytes(_ksk_msg(pbm, oracle, bfv.level_moduli(key_level), r0[:-1], r1[:-1], ct_level,
                                    key_level).SerializeToString())


def test_seeded_key_switching_key(fh, oracle, pbm, setup):
    moduli, bfv, par = setup
    ct_level, key_level = 2, 1
    km_moduli = bfv.level_moduli(key_level)
    L = len(bfv.level_moduli(ct_level))
    seed = bytes(range(7, 39))
    c1 = oracle.generate_c1(km_moduli, N, seed, L).reshape(L, len(km_moduli), N)
    c0 = oracle.random_polys(km_moduli, N, L, seed=21).reshape(L, len(km_moduli), N)
    km = _ksk_msg(pbm, oracle, km_moduli, c0, None, ct_level, key_level, seed=seed)
    raw = km.SerializeToString()
    k = par.ksk_from_bytes(raw)
    assert par.ksk_to_bytes(k) == raw  # the seed stands for c1 again
    ok = bfv.key_from_rows(c0, c1, ct_level=ct_level, key_level=key_level)
    p = oracle.random_polys(bfv.level_moduli(ct_level), N, 2, seed=5)
    out = fh.empty(2 * 2 * len(km_moduli) * N)
    k.key_switch(fh.to_device(p), out, 2)
    _sync()
    got = fh.to_host(out).reshape(2, 2, len(km_moduli) * N)
    for i in range(2):
        e0, e1 = ok.key_switch(p[i * L * N:(i + 1) * L * N])
        assert np.array_equal(got[i, 0], e0) and np.array_equal(got[i, 1], e1)


def test_galois_and_evaluation_key_messages(fh, oracle, pbm, setup):
    moduli, bfv, par = setup
    level = 1
    lm = bfv.level_moduli(level)
    keys, msgs = {}, []
    for e in (3, 2 * N - 1):
        gk = bfv.galois_key(e, ct_level=level, key_level=level, seed=e % 97)
        r0, r1 = gk.export()
        m = pbm.GaloisKey(ksk=_ksk_msg(pbm, oracle, lm, r0, r1, level, level), exponent=e)
        keys[e] = gk
        msgs.append(m)
        g = par.galois_key_from_bytes(m.SerializeToString())
        assert par.galois_key_to_bytes(g) == m.SerializeToString()
        ct = bfv.encrypt(list(range(N)), level=level, seed=e % 5)
        out = fh.empty(ct.size)
        g.relinearize(fh.to_device(ct), out, 1)
        _sync()
        assert np.array_equal(fh.to_host(out), bfv.galois(gk, ct))
    em = pbm.EvaluationKey(gk=msgs, ciphertext_level=level, evaluation_key_level=level)
    ek, lvl = par.evaluation_key_from_bytes(em.SerializeToString())
    assert lvl == level
    assert par.evaluation_key_to_bytes(ek) == em.SerializeToString()
    ct = bfv.encrypt(list(range(N)), level=level, seed=8)
    out = fh.empty(ct.size)
    ek.rotate_rows(fh.to_device(ct), out, 1)
    _sync()
    assert np.array_equal(fh.to_host(out), bfv.galois(keys[2 * N - 1], ct))
    bad = pbm.EvaluationKey(gk=msgs, ciphertext_level=level, evaluation_key_level=0)
    with pytest.raises(fh.FheHipError):
        par.evaluation_key_from_bytes(bad.SerializeToString())


def test_parameters_messages(fh, oracle, pbm, setup):
    moduli, bfv, par = setup
    # prost writes the oneof after the declared fields: two concatenated
    # messages give that order
    exp = (pbm.Parameters(degree=N, moduli=moduli, variance=10).SerializeToString() +
           pbm.Parameters(plaintext=T).SerializeToString())
    assert par.to_bytes() == exp
    for raw in (exp, pbm.Parameters(degree=N, moduli=moduli, variance=4, plaintext=T).SerializeToString(),
                pbm.Parameters(degree=N, moduli=moduli, variance=4,
                               plaintext_big=T.to_bytes(3, "little")).SerializeToString()):
        p2 = fh.Parameters.from_bytes(raw)
        assert (p2.degree, p2.plaintext, p2.max_level, p2.moduli) == (N, T, len(moduli) - 1, list(moduli))
    assert fh.Parameters.from_bytes(
        pbm.Parameters(degree=N, moduli=moduli, variance=4, plaintext=T).SerializeToString()).to_bytes() == (
        pbm.Parameters(degree=N, moduli=moduli, variance=4).SerializeToString() +
        pbm.Parameters(plaintext=T).SerializeToString())
    big = (1 << 127) - 1
    for raw in (pbm.Parameters(degree=N, moduli=moduli, variance=4).SerializeToString(),
                pbm.Parameters(degree=N, moduli=moduli, variance=0, plaintext=T).SerializeToString(),
                pbm.Parameters(degree=N, moduli=moduli, variance=4,
                               plaintext_big=big.to_bytes(16, "little")).SerializeToString(),
                pbm.Parameters(degree=1000, moduli=moduli, variance=4, plaintext=T).SerializeToString()):
        with pytest.raises(fh.FheHipError):
            fh.Parameters.from_bytes(raw)
