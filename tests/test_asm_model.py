"""CPU model of the gfx950 assembly blocks of the device arithmetic.

The Shoup product, the forward and inverse NTT butterflies
(math/src/ntt/native.rs:256-300, math/src/zq/mod.rs:224-234) and the scaler's
column accumulators are inline assembly in fhe.rs_amd/csrc/device_math.h and
rns_device.h. This test reads those blocks from the headers, executes them
with a small per-lane model of the instructions they use, and checks

- the results against Python integer arithmetic, on edge values (operands at
  2^64 - 1, 4p - 1, 2p - 1, moduli next to 2^62, 2^60 and small ones) and on
  random values;
- that every VALU read of vcc or of a carry SGPR comes at least two wait
  states after the VALU instruction that wrote it (the gfx950 requirement the
  blocks are written to);
- that 64-bit shift amounts stay within v_lshl_add_u64's 3-bit field.

No GPU is needed; the GPU tests check the same code bit for bit against the
oracle on whole transforms.
"""
import os
import random
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "fhe.rs_amd", "csrc")
M32 = (1 << 32) - 1
M64 = (1 << 64) - 1


def asm_block(header, name):
    """The instruction list of macro `name` (its string literals, split on \\n\\t)."""
    text = open(os.path.join(CSRC, header)).read()
    start = text.index("#define " + name)
    lines = []
    for line in text[start:].splitlines()[1:]:
        lines.append(line)
        if not line.rstrip().endswith("\\"):
            break
    body = "".join(re.findall(r'"((?:[^"\\]|\\.)*)"', "\n".join(lines)))
    return [ins.strip() for ins in body.split("\\n\\t") if ins.strip()]


class Lane:
    """One lane: VGPRs v0..v15, named operands, carry SGPRs (one bit per lane)."""

    def __init__(self, operands):
        self.v = {}
        self.ops = dict(operands)   # name -> value (32- or 64-bit)
        self.sgpr = {}              # carry SGPR names ("vcc", "%[k]", ...) -> bit
        self.clock = 0
        self.written = {}           # SGPR name -> wait-state clock of its last VALU write
        self.min_gap = None

    # -- operand access ------------------------------------------------------
    def read(self, tok, width):
        tok = tok.strip()
        m = re.fullmatch(r"v\[(\d+):(\d+)\]", tok)
        if m:
            lo, hi = int(m.group(1)), int(m.group(2))
            assert hi == lo + 1 and lo % 2 == 0, tok
            return self.v[lo] | (self.v[hi] << 32)
        m = re.fullmatch(r"v(\d+)", tok)
        if m:
            return self.v[int(m.group(1))]
        m = re.fullmatch(r"%\[(\w+)\]", tok)
        if m:
            val = self.ops[m.group(1)]
            assert val <= (M64 if width == 64 else M32), (tok, width)
            return val
        if re.fullmatch(r"-?\d+", tok):
            return int(tok) & (M64 if width == 64 else M32)
        raise AssertionError("unknown operand " + tok)

    def write(self, tok, val, width):
        tok = tok.strip()
        m = re.fullmatch(r"v\[(\d+):(\d+)\]", tok)
        if m:
            lo = int(m.group(1))
            self.v[lo], self.v[lo + 1] = val & M32, (val >> 32) & M32
            return
        m = re.fullmatch(r"v(\d+)", tok)
        if m:
            self.v[int(m.group(1))] = val & M32
            return
        m = re.fullmatch(r"%\[(\w+)\]", tok)
        if m:
            self.ops[m.group(1)] = val & (M64 if width == 64 else M32)
            return
        raise AssertionError("unknown destination " + tok)

    def set_carry(self, name, bit):
        self.sgpr[name] = bit
        self.written[name] = self.clock

    def get_carry(self, name):
        gap = self.clock - self.written[name] - 1
        self.min_gap = gap if self.min_gap is None else min(self.min_gap, gap)
        return self.sgpr[name]

    # -- instructions ----------------------------------------------------------
    def run(self, block):
        for ins in block:
            op, _, rest = ins.partition(" ")
            a = [x.strip() for x in rest.split(",")] if rest else []
            if op == "s_nop":
                self.clock += int(a[0]) + 1
                continue
            self.step(op, a)
            self.clock += 1

    def step(self, op, a):
        if op == "v_mov_b32":
            self.write(a[0], self.read(a[1], 32), 32)
        elif op == "v_mov_b64":
            self.write(a[0], self.read(a[1], 64), 64)
        elif op == "v_mul_hi_u32":
            self.write(a[0], (self.read(a[1], 32) * self.read(a[2], 32)) >> 32, 32)
        elif op == "v_mul_lo_u32":
            self.write(a[0], self.read(a[1], 32) * self.read(a[2], 32), 32)
        elif op == "v_mad_u64_u32":
            full = self.read(a[2], 32) * self.read(a[3], 32) + self.read(a[4], 64)
            self.write(a[0], full, 64)
            self.set_carry(a[1], full >> 64)
        elif op == "v_add3_u32":
            self.write(a[0], self.read(a[1], 32) + self.read(a[2], 32) + self.read(a[3], 32), 32)
        elif op == "v_add_u32":
            self.write(a[0], self.read(a[1], 32) + self.read(a[2], 32), 32)
        elif op == "v_and_b32":
            self.write(a[0], self.read(a[1], 32) & self.read(a[2], 32), 32)
        elif op == "v_addc_co_u32":
            r = self.read(a[2], 32) + self.read(a[3], 32) + self.get_carry(a[4])
            self.write(a[0], r, 32)
            self.set_carry(a[1], r >> 32)
        elif op in ("v_sub_co_u32", "v_subrev_co_u32"):
            x, y = self.read(a[2], 32), self.read(a[3], 32)
            if op == "v_subrev_co_u32":
                x, y = y, x
            self.write(a[0], x - y, 32)
            self.set_carry(a[1], 1 if x < y else 0)
        elif op == "v_subb_co_u32":
            r = self.read(a[2], 32) - self.read(a[3], 32) - self.get_carry(a[4])
            self.write(a[0], r, 32)
            self.set_carry(a[1], 1 if r < 0 else 0)
        elif op == "v_cndmask_b32":
            c = self.get_carry(a[3])
            self.write(a[0], self.read(a[2], 32) if c else self.read(a[1], 32), 32)
        elif op == "v_cmp_gt_u64":
            self.set_carry(a[0], 1 if self.read(a[1], 64) > self.read(a[2], 64) else 0)
        elif op == "v_lshl_add_u64":
            sh = int(a[2])
            assert 0 <= sh <= 7, "v_lshl_add_u64 takes a 3-bit shift"
            self.write(a[0], (self.read(a[1], 64) << sh) + self.read(a[3], 64), 64)
        else:
            raise AssertionError("instruction not modelled: " + op)


def halves(x):
    return x & M32, x >> 32


def shoup_ref(a, w, p):
    ws = (w << 64) // p
    q = (a * ws) >> 64
    return (a * w - q * p) & M64


MODULI = [
    4611686018427322369,   # 62-bit NTT prime (2^62 - 2^15 * k + 1)
    4611686018326724609,   # primes.rs KAT
    1152921504606748673,   # 60-bit
    1152921504599080961,   # 60-bit special prime
    1125899906826241,      # 50-bit
    65537,
    12289,
]


def edge_values(bound):
    return sorted({0, 1, 2, bound - 1, bound - 2, bound // 2, (bound // 2) - 1, bound // 3})


def shoup_cases(p):
    rng = random.Random(p)
    ws = [0, 1, p - 1, p - 2, p // 2] + [rng.randrange(p) for _ in range(6)]
    aa = [0, 1, p - 1, 2 * p - 1, 4 * p - 1, M64, M64 - 1, 1 << 63, (1 << 32) - 1] + [rng.randrange(1 << 64)
                                                                                       for _ in range(12)]
    return [(a & M64, w) for a in aa for w in ws]


def run_shoup(block, a, w, p):
    ws = (w << 64) // p
    n = (1 << 64) - p
    a0, a1 = halves(a)
    s0, s1 = halves(ws)
    w0, w1 = halves(w)
    n0, n1 = halves(n)
    lane = Lane(dict(a0=a0, a1=a1, s0=s0, s1=s1, w0=w0, w1=w1, n0=n0, n1=n1))
    lane.run(block)
    return lane.read("v[2:3]", 64), lane


@pytest.mark.parametrize("p", MODULI)
def test_shoup_product(p):
    block = asm_block("device_math.h", "FHEHIP_SHOUP_ASM")
    for a, w in shoup_cases(p):
        r, lane = run_shoup(block, a, w, p)
        assert r == shoup_ref(a, w, p), (p, a, w)
        assert r < 2 * p and r % p == (a * w) % p
        assert lane.min_gap >= 2


def run_fwd(block, x, y, w, p):
    ws = (w << 64) // p
    n, p2 = (1 << 64) - p, 2 * p
    ops = {}
    ops["x0"], ops["x1"] = halves(x)
    ops["a0"], ops["a1"] = halves(y)
    ops["s0"], ops["s1"] = halves(ws)
    ops["w0"], ops["w1"] = halves(w)
    ops["n0"], ops["n1"] = halves(n)
    ops["P0"], ops["P1"] = halves(p2)
    ops["xo"] = ops["yo"] = 0
    lane = Lane(ops)
    lane.run(block)
    return lane.ops["xo"], lane.ops["yo"], lane


@pytest.mark.parametrize("p", MODULI)
def test_forward_butterfly(p):
    block = asm_block("device_math.h", "FHEHIP_FWD_BFLY_ASM")
    rng = random.Random(p + 1)
    xs = edge_values(4 * p) + [rng.randrange(4 * p) for _ in range(8)]
    ys = edge_values(4 * p) + [rng.randrange(4 * p) for _ in range(8)]
    ws = [0, 1, p - 1] + [rng.randrange(p) for _ in range(3)]
    for x in xs:
        for y in ys:
            for w in ws:
                xo, yo, lane = run_fwd(block, x, y, w, p)
                xr = x - 2 * p if x >= 2 * p else x
                t = shoup_ref(y, w, p)
                assert (xo, yo) == (xr + t, xr + 2 * p - t), (p, x, y, w)
                assert xo < 4 * p and yo < 4 * p
                assert (xo - x - y * w) % p == 0 and (yo - x + y * w) % p == 0
                assert lane.min_gap >= 2


def run_inv(block, x, y, z, p):
    zs = (z << 64) // p
    n, p2 = (1 << 64) - p, 2 * p
    m2 = (1 << 64) - p2
    ops = {"x": x, "y": y, "P": p2}
    ops["x0"], ops["x1"] = halves(x)
    ops["y0"], ops["y1"] = halves(y)
    ops["s0"], ops["s1"] = halves(zs)
    ops["w0"], ops["w1"] = halves(z)
    ops["n0"], ops["n1"] = halves(n)
    ops["m0"], ops["m1"] = halves(m2)
    ops["xo"] = ops["yo"] = 0
    lane = Lane(ops)
    lane.run(block)
    return lane.ops["xo"], lane.ops["yo"], lane


@pytest.mark.parametrize("p", MODULI)
def test_inverse_butterfly(p):
    block = asm_block("device_math.h", "FHEHIP_INV_BFLY_ASM")
    rng = random.Random(p + 2)
    xs = edge_values(2 * p) + [rng.randrange(2 * p) for _ in range(8)]
    ys = edge_values(2 * p) + [rng.randrange(2 * p) for _ in range(8)]
    zs = [0, 1, p - 1] + [rng.randrange(p) for _ in range(3)]
    for x in xs:
        for y in ys:
            for z in zs:
                xo, yo, lane = run_inv(block, x, y, z, p)
                u = x + y
                assert xo == (u - 2 * p if u >= 2 * p else u), (p, x, y, z)
                assert yo == shoup_ref(x + 2 * p - y, z, p), (p, x, y, z)
                assert xo < 2 * p and yo < 2 * p
                assert (yo - (x - y) * z) % p == 0
                assert lane.min_gap >= 2


def acc3_value(ops):
    return (ops["a"] + (ops["b"] << 32) + (ops["c"] << 64) + (ops["ka"] << 64) + (ops["kb"] << 96) +
            (ops["kc"] << 128))


def test_column_accumulator():
    first = asm_block("rns_device.h", "FHEHIP_ACC3_FIRST_ASM")
    mac = asm_block("rns_device.h", "FHEHIP_ACC3_ASM")
    rng = random.Random(7)
    cases = [[M64] * 40, [M64, 1, M64, 0] * 5, [(1 << 62) - 1] * 26]
    for _ in range(20):
        cases.append([rng.randrange(1 << 64) for _ in range(rng.randrange(1, 30))])
    for vals in cases:
        ws = [M64 if i % 3 == 0 else rng.randrange(1 << 64) for i in range(len(vals))]
        ops = {k: 0 for k in ("a", "b", "c", "ka", "kb", "kc")}
        total = 0
        for i, (r, w) in enumerate(zip(vals, ws)):
            ops["r0"], ops["r1"] = halves(r)
            ops["w0"], ops["w1"] = halves(w)
            lane = Lane(ops)
            lane.run(first if i == 0 else mac)
            assert lane.min_gap >= 2
            ops = lane.ops
            total += r * w
            assert acc3_value(ops) == total
            assert max(ops["ka"], ops["kb"], ops["kc"]) <= M32
