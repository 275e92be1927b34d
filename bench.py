#!/usr/bin/env python3
"""Headline benchmark: batched ct x ct multiply + relinearize at N = 2^14.

Workload: 512 independent BFV ciphertext pairs at N=16384 with 6 x 60-bit
ciphertext primes and one 60-bit special prime (key at the level above the
ciphertexts), t = 65537, Multiplicator::default with a relinearization key.
One step = one replay of the HIP graph of one fhehip_multiply over the whole
batch (extension, tensor, t/Q down-scaling, key switch, switch-down) at fixed
launch shapes of the fused kernels (HEADLINE_TUNING, the shapes the committed
counter summaries were measured at), inputs already resident in HBM; the
same step as an eager call, and at the autotuner's choice, is reported
beside it.

Also measured: the per-call GPU drop-in (the same entry point called once per
operation), per-kernel durations with HIP events on the launch stream for the
roofline figure, and the CPU baseline: the oracle (a C++ restatement of the
fhe.rs CPU path) timed on the host cores on a bounded sample, which also checks
one GPU output bit for bit.

Prints one JSON line.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fhe.rs_amd"))

METRIC = "ct×ct mul+relin/sec at N=2^14 (batched vs per-call GPU), bit-exact; p99 ms"
UNIT = "mul+relin/s"
N = 16384
# generate_moduli([60] * 7, 16384) of fhe/src/bfv/parameters.rs:383-423 (pinned in
# tests/test_oracle_zq.py::test_config_moduli): 6 ciphertext primes + 1 special.
MODULI = [1152921504606748673, 1152921504606683137, 1152921504606584833, 1152921504605962241,
          1152921504604979201, 1152921504600260609, 1152921504599080961]
T_PLAIN = 65537
HBM_PEAK_GBS = 8000.0
# Launch shapes of the fused kernels for the timed steps (fhehip_tuning_set at
# this degree): 2 instances per key load in the key inner product, 16 values
# per thread, the key-switch finish and the tensor product + t/Q scaling as
# half rows (the latter with the paired scaler on the matrix cores). The
# committed counter summaries are measured at these shapes, so that their
# kernels are the ones timed here; the autotuner's own choice is timed
# beside it.
HEADLINE_TUNING = {"xcd_order": 1, "ks_mac_instances": 2, "tensor_log2_values": 4, "lift_log2_values": 4,
                   "finish_log2_values": 4, "halves": 5, "scaler_mfma": 1}


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q / 100.0
    f = int(k)
    c = min(f + 1, len(xs) - 1)
    return xs[f] + (xs[c] - xs[f]) * (k - f)


def kernel_bytes(name, B, L, Lk, M):
    """Algorithmic HBM bytes of one launch of a kernel of the mul+relin
    pipeline for a batch of B instances (each u64 row = N * 8 bytes read or
    written once)."""
    row = N * 8
    kind = name.split("/")[0]
    if kind == "intt_rows":      # 4 input polys x L rows, read + write
        return B * 4 * L * row * 2
    if kind == "ntt_forward":    # the new extension rows, in place
        return B * 4 * (M - L) * row * 2
    if kind == "tensor_intt":    # read a0,a1,b0,b1 over M limbs, write 3 x M rows
        return B * (4 + 3) * M * row
    if kind == "rns_extend":     # one launch for both operands, 4 polys: read L rows, write M - L rows
        return B * 4 * M * row
    if kind == "rns_down":       # 3 polys: read M rows, write L rows
        return B * 3 * (M + L) * row
    if kind == "lift_ntt":       # read c2 (L rows), write the lifted, transformed digits (L x Lk rows)
        return B * (L + L * Lk) * row
    if kind == "ks_mac":         # read the L x Lk transformed digits, write 2 (L + 1) rows; key rows once
        return B * (L * Lk + 2 * Lk) * row + 2 * L * Lk * row
    if kind == "ntt_backward":   # the dropped limb's two rows, in place
        return B * 2 * row * 2
    if kind == "ks_finish":      # read 2 dropped-limb rows, 2L c0/c1 rows and 2L outputs; write 2L rows
        return B * (2 + 6 * L) * row
    if kind == "ks_finish_paired":  # after ks_z_pair: read 2L z rows and 2L outputs; write 2L rows
        return B * 6 * L * row
    if kind == "ks_z_pair":      # read 2 dropped-limb rows and 2L c0/c1 rows; write 2L z rows
        return B * (2 + 4 * L) * row
    return None


def kernel_sources_digest():
    """SHA-256 (first 16 hex digits) of the library's kernel and host sources
    (fhe.rs_amd/csrc, sorted by name): the counter summaries under profiles/
    record the digest of the sources they were measured on."""
    import hashlib
    h = hashlib.sha256()
    d = os.path.join(ROOT, "fhe.rs_amd", "csrc")
    for name in sorted(os.listdir(d)):
        h.update(name.encode())
        with open(os.path.join(d, name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def load_profile(name):
    """(rows, digest) of a committed rocprofv3 counter summary: its table
    rows split on whitespace, and the digest on its `Kernel sources:` line."""
    rows, digest = [], None
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            for line in f:
                if line.startswith("Kernel sources:"):
                    # the first `…` token (the line may go on to say what it is)
                    m = re.search(r"`([0-9a-f]+)`", line)
                    digest = m.group(1) if m else line.split(":", 1)[1].strip().strip("`")
                else:
                    rows.append(line.split())
    except OSError:
        pass
    return rows, digest


TRAFFIC_PROFILE = "pmc_traffic_mul_relin_n16384_b512.md"
VALU_PROFILE = "pmc_valu_mul_relin_n16384_b512.md"


def load_traffic(kernel, digest):
    """HBM bytes per launch of `kernel` (its name without the degree suffix)
    from the committed counter summary (rows `<kernel> <fetch KiB> <write KiB>
    <bytes>`), only when that summary was measured on the current sources;
    otherwise None (a stale figure is never reported as this run's)."""
    rows, d = load_profile(TRAFFIC_PROFILE)
    if d != digest:
        return None
    for parts in rows:
        if len(parts) == 4 and parts[0] == kernel and parts[3].isdigit():
            return int(parts[3])
    return None


def load_valu(kernel, digest):
    """(VALUBusy %, wave-instructions per CU-cycle) of `kernel` from the
    committed VALU counter summary, when measured on the current sources."""
    rows, d = load_profile(VALU_PROFILE)
    if d != digest:
        return None, None
    for parts in rows:
        if len(parts) >= 6 and parts[0] == kernel:
            try:
                return float(parts[1]), float(parts[5])
            except ValueError:
                return None, None
    return None, None


def profile_provenance(digest):
    """Which counter summaries the roofline's traffic and VALU figures come
    from, and whether they were measured on the sources of this run."""
    out = {"kernel_sources": digest}
    for key, name in (("traffic", TRAFFIC_PROFILE), ("valu", VALU_PROFILE)):
        _, d = load_profile(name)
        out[key] = {"file": "profiles/" + name, "measured_on": d, "current": d == digest}
        if d != digest:
            print(f"bench.py: profiles/{name} was measured on kernel sources {d}, not {digest}; "
                  f"its figures are left out", file=sys.stderr)
    return out


def host_cpu():
    """The host's CPU as the CPU baseline states it: usable cores, nproc and
    lscpu's model fields."""
    info = {"usable_cores": len(os.sched_getaffinity(0)), "nproc": os.cpu_count()}
    try:
        import subprocess
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("Model name", "CPU family", "Model", "Vendor ID", "Thread(s) per core", "Core(s) per socket",
                     "Socket(s)"):
                info[k] = v.strip()
    except (OSError, ValueError):
        pass
    return info


_ORACLE_BUILD = None
ORACLE_GENERIC = "g++ -O3 -march=x86-64-v3 (committed build)"
ORACLE_NATIVE = "g++ -O3 -march=native (built on this host)"


def oracle_native_path():
    """The oracle compiled on this host with -O3 -march=native into a
    temporary directory, or None if the compiler is unavailable."""
    import subprocess
    import tempfile
    out = os.path.join(tempfile.mkdtemp(prefix="fhehip_oracle_"), "liboracle.so")
    cmd = ["g++", "-O3", "-march=native", "-std=c++17", "-fPIC", "-shared", "-o", out,
           os.path.join(ROOT, "oracle", "oracle_capi.cpp"), "-lpthread"]
    try:
        subprocess.run(cmd, check=True, timeout=600, capture_output=True)
        return out
    except (OSError, subprocess.SubprocessError):
        return None


def time_oracle_builds(budget_s=4.0):
    """Single-thread time of one headline-shape Multiplicator::multiply with
    each build of the oracle, each in its own process (a library loads once
    per process): {build: ms}."""
    import subprocess
    script = (
        "import sys, time, numpy as np; sys.path.insert(0, sys.argv[2]); import pyoracle as o; "
        "o.LIB_PATH = sys.argv[1]; o.lib(); "
        f"N = {N}; M = {MODULI!r}; b = o.Bfv(N, {T_PLAIN}, M, sk_seed=1); rk = b.relin_key(1, 0, seed=2); "
        "m = o.Multiplicator(b, rk, level=1); x = o.random_polys(M[:-1], N, 2, seed=3); m.multiply(x, x); "
        "ts = []\n"
        "for _ in range(5):\n t0 = time.perf_counter(); m.multiply(x, x); ts.append(time.perf_counter() - t0)\n"
        "print(sorted(ts)[2] * 1e3)")
    out = {}
    paths = {ORACLE_GENERIC: os.path.join(ROOT, "oracle", "lib", "liboracle.so")}
    native = oracle_native_path()
    if native:
        paths[ORACLE_NATIVE] = native
    for build, path in paths.items():
        try:
            r = subprocess.run([sys.executable, "-c", script, path, os.path.join(ROOT, "oracle")], capture_output=True,
                               text=True, timeout=300, check=True)
            out[build] = (float(r.stdout.strip().splitlines()[-1]), path)
        except (OSError, subprocess.SubprocessError, ValueError, IndexError):
            pass
    return out


def oracle_module():
    """The oracle (C++ restatement of the fhe.rs CPU path) for the CPU
    baseline: of the committed x86-64-v3 build and a -march=native build made
    on this host, the one faster on this host's cores. Returns (module, build,
    {build: single-op ms})."""
    global _ORACLE_BUILD
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    if _ORACLE_BUILD is None:
        timed = time_oracle_builds()
        if timed:
            build = min(timed, key=lambda k: timed[k][0])
            pyoracle.LIB_PATH = timed[build][1]
            _ORACLE_BUILD = (build, {k: v[0] for k, v in timed.items()})
        else:
            _ORACLE_BUILD = (ORACLE_GENERIC, {})
    pyoracle.lib()
    return pyoracle, _ORACLE_BUILD[0], _ORACLE_BUILD[1]


# Single-instance configuration: N = 8192, generate_moduli([62] * 3, 8192), t = 65537,
# relinearization key at the ciphertext level.
CFG1_N = 8192
CFG1_MODULI = [4611686018427322369, 4611686018427289601, 4611686018426454017]
# NTT microbenchmark: 256 polynomials x 8 primes x N = 16384.
NTT_MODULI = [4611686018427322369, 4611686018427289601, 4611686018425815041, 4611686018424733697,
              4611686018423881729, 4611686018423390209, 4611686018423062529, 4611686018422669313]


def single_instance_latency(fhehip, torch, reps):
    """One ct x ct multiply + relinearize per call at N = 8192, 3 x 62-bit."""
    import numpy as np
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(11)
    ctx = fhehip.Context(CFG1_MODULI, CFG1_N)
    L = len(CFG1_MODULI)
    c0 = np.stack([np.stack([rng.integers(0, q, CFG1_N, dtype=np.uint64) for q in CFG1_MODULI]) for _ in range(L)])
    c1 = np.stack([np.stack([rng.integers(0, q, CFG1_N, dtype=np.uint64) for q in CFG1_MODULI]) for _ in range(L)])
    mult = fhehip.Multiplicator(rk=fhehip.KeySwitchingKey(ctx, ctx, c0, c1), plaintext=T_PLAIN)
    per = 2 * L * CFG1_N
    cts = np.concatenate([np.stack([rng.integers(0, q, (2, CFG1_N), dtype=np.uint64) for q in CFG1_MODULI], 1)
                          .reshape(-1) for _ in range(2)])
    d = fhehip.to_device(cts)
    out = torch.empty(per, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()
    for _ in range(5):
        mult.multiply(d[:per], d[per:], out, 1, stream)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    ev[0].record(stream)
    for k in range(reps):
        mult.multiply(d[:per], d[per:], out, 1, stream)
        ev[k + 1].record(stream)
    torch.cuda.synchronize()
    lat = [ev[k].elapsed_time(ev[k + 1]) for k in range(reps)]
    return {"workload": "single instance: N=8192, 3x62-bit, t=65537, 1 ct*ct mul+relin per call",
            "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
            "value": reps / (ev[0].elapsed_time(ev[-1]) / 1e3), "unit": UNIT, "reps": reps}


def ntt_microbench(fhehip, torch, reps):
    """Forward then inverse NTT of 256 x 8 x 16384 residues (bit-exact round trip)."""
    dev = torch.device("cuda:0")
    n, count, L = 16384, 256, len(NTT_MODULI)
    ctx = fhehip.Context(NTT_MODULI, n)
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    d = torch.empty(count * L * n, dtype=torch.int64, device=dev)
    v = d.view(count, L, n)
    for i, q in enumerate(NTT_MODULI):
        v[:, i, :] = torch.randint(0, q, (count, n), dtype=torch.int64, device=dev, generator=g)
    ref = d.clone()
    stream = torch.cuda.current_stream()
    ctx.ntt_forward(d, count, stream=stream)
    ctx.ntt_backward(d, count, stream=stream)
    torch.cuda.synchronize()
    exact = bool(torch.equal(d, ref))
    # steady state: warm up for about as long as the timed region
    for _ in range(reps):
        ctx.ntt_forward(d, count, stream=stream)
        ctx.ntt_backward(d, count, stream=stream)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ev[0].record(stream)
    for _ in range(reps):
        ctx.ntt_forward(d, count, stream=stream)
    ev[1].record(stream)
    for _ in range(reps):
        ctx.ntt_backward(d, count, stream=stream)
    ev[2].record(stream)
    torch.cuda.synchronize()
    f_ms = ev[0].elapsed_time(ev[1]) / reps
    b_ms = ev[1].elapsed_time(ev[2]) / reps
    bytes_one = count * L * n * 16  # read + write once per transform
    rt_ms = f_ms + b_ms
    return {"workload": "NTT 256 polys x 8 x 62-bit primes x N=16384, forward then inverse",
            "forward_us": f_ms * 1e3, "inverse_us": b_ms * 1e3, "roundtrip_us": rt_ms * 1e3,
            "roundtrip_GBps": 2 * bytes_one / (rt_ms / 1e3) / 1e9,
            "hbm_frac": 2 * bytes_one / (rt_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
            "roundtrip_bit_exact": exact}


# Uneven batch: N drawn from these degrees, modulus count (including the one
# special prime of the key) from the range of its degree, 50-62-bit primes,
# 1-4 column rotations per instance with steps drawn from UNEVEN_STEPS.
UNEVEN_DEGREES = [4096, 8192, 16384, 32768]
UNEVEN_MODULI_COUNT = {4096: (2, 3), 8192: (3, 5), 16384: (5, 8), 32768: (8, 12)}
UNEVEN_STEPS = [1, 2, 4]


def _uniform_rows(torch, dev, moduli, shape, gen):
    """Device int64 tensor [..., L, N] with limb i uniform in [0, q_i)."""
    t = torch.empty(shape, dtype=torch.int64, device=dev)
    for i, q in enumerate(moduli):
        t[..., i, :] = torch.randint(0, q, t[..., i, :].shape, dtype=torch.int64, device=dev, generator=gen)
    return t


def _random_ksk(fhehip, np, rng, cctx, kctx, ct_moduli, key_moduli, n):
    c = []
    for _ in range(2):
        rows = np.empty((len(ct_moduli), len(key_moduli), n), dtype=np.uint64)
        for j, q in enumerate(key_moduli):
            rows[:, j, :] = rng.integers(0, q, (len(ct_moduli), n), dtype=np.uint64)
        c.append(rows)
    return fhehip.KeySwitchingKey(cctx, kctx, c[0], c[1])


def uneven_batch(fhehip, torch, instances=1024, reps=3, per_call_jobs=48):
    """1024 independent evaluations of mixed degree and RNS width, each a
    multiply + relinearize followed by 1-4 column rotations, run by one
    fhehip_submit per step (jobs grouped by parameter set, groups on
    concurrent streams); compared with the same jobs issued one call at a time."""
    import numpy as np
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(4)
    inst = []
    for _ in range(instances):
        n = UNEVEN_DEGREES[int(rng.integers(len(UNEVEN_DEGREES)))]
        lo, hi = UNEVEN_MODULI_COUNT[n]
        count = int(rng.integers(lo, hi + 1))
        steps = [int(x) for x in rng.choice(UNEVEN_STEPS, int(rng.integers(1, 5)))]
        inst.append((n, count, steps))
    gen = torch.Generator(device=dev)
    gen.manual_seed(5)
    t0 = time.perf_counter()
    sets = {}
    for key in sorted({(n, c) for n, c, _ in inst}):
        n, count = key
        srng = np.random.default_rng(n + count)
        moduli = fhehip.generate_moduli([int(x) for x in srng.integers(50, 63, count)], n)
        ct_moduli = moduli[:-1]
        cctx, kctx = fhehip.Context(ct_moduli, n), fhehip.Context(moduli, n)
        rk = _random_ksk(fhehip, np, srng, cctx, kctx, ct_moduli, moduli, n)
        mult = fhehip.Multiplicator(rk=rk, plaintext=T_PLAIN)
        gks = [fhehip.GaloisKey(_random_ksk(fhehip, np, srng, cctx, kctx, ct_moduli, moduli, n), pow(3, st, 2 * n))
               for st in UNEVEN_STEPS]
        members = [k for k, (n2, c2, _) in enumerate(inst) if (n2, c2) == key]
        m = len(members)
        L = len(ct_moduli)
        sets[key] = {"mult": mult, "ek": fhehip.EvaluationKey(gks), "ctx": (cctx, kctx), "L": L, "n": n,
                     "lhs": _uniform_rows(torch, dev, ct_moduli, (m, 2, L, n), gen).view(-1),
                     "rhs": _uniform_rows(torch, dev, ct_moduli, (m, 2, L, n), gen).view(-1),
                     "out": torch.empty(m * 2 * L * n, dtype=torch.int64, device=dev), "next": 0}
    setup_s = time.perf_counter() - t0
    specs = []
    for n, count, steps in inst:
        S = sets[(n, count)]
        w = 2 * S["L"] * n
        j = S["next"]
        S["next"] += 1
        sl = slice(j * w, (j + 1) * w)
        specs.append((S["mult"], S["ek"], S["lhs"][sl], S["rhs"][sl], S["out"][sl], steps))
    jobs = fhehip.make_jobs(specs)
    stream = torch.cuda.current_stream()
    fhehip.submit(jobs, stream)
    torch.cuda.synchronize()
    # A half-precision fill on each side of the timed submits marks them in a
    # kernel trace (it runs outside the timed region).
    marker = torch.empty(4, dtype=torch.float16, device=dev)
    marker.fill_(1.5)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    ev[0].record(stream)
    for k in range(reps):
        fhehip.submit(jobs, stream)
        ev[k + 1].record(stream)
    marker.fill_(2.5)
    torch.cuda.synchronize()
    sub_ms = [ev[k].elapsed_time(ev[k + 1]) for k in range(reps)]
    total_ms = ev[0].elapsed_time(ev[-1])
    # the same jobs, one call per operation
    exact = True
    outs = []
    pe = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    pe[0].record(stream)
    for mult, ek, a, b, o, steps in specs[:per_call_jobs]:
        cur = torch.empty_like(o)
        mult.multiply(a, b, cur, 1, stream)
        for st in steps:
            nxt = torch.empty_like(o)
            ek.rotate_columns(cur, nxt, st, 1, stream)
            cur = nxt
        outs.append(cur)
    pe[1].record(stream)
    torch.cuda.synchronize()
    per_ms = pe[0].elapsed_time(pe[1]) / max(1, min(per_call_jobs, len(specs)))
    for (mult, ek, a, b, o, steps), cur in zip(specs, outs):
        exact = exact and bool(torch.equal(cur, o))
    degrees = {str(n): sum(1 for x in inst if x[0] == n) for n in UNEVEN_DEGREES}
    value = instances * reps / (total_ms / 1e3)
    return {"workload": f"{instances} instances, N from {UNEVEN_DEGREES}, 2-12 moduli of 50-62 bits scaled to N, "
                        f"mul+relin then 1-4 column rotations (steps {UNEVEN_STEPS}), one fhehip_submit per step",
            "value": value, "unit": "instances/s", "ms_per_submit": total_ms / reps,
            "p50_submit_ms": percentile(sub_ms, 50), "parameter_sets": len(sets),
            "instances_per_degree": degrees, "rotations": sum(len(x[2]) for x in inst),
            "per_call_ms_per_instance": per_ms, "batched_over_per_call": value * per_ms / 1e3,
            "bit_exact_vs_single_calls": exact, "single_calls_checked": len(outs), "setup_s": setup_s}


# Serving: Poisson arrivals of encrypted 4096-slot sums (ct x ct multiply +
# relinearize, then 12 rotate-and-add steps); 70% at N=16384 with the batched
# configuration's moduli, 30% at N=32768 with 8 x 60-bit primes + 1 special.
SERVE_RATE = 2000.0
SERVE_SECONDS = 3.0
SERVE_SHARE_16K = 0.7
SERVE_LOG2_SLOTS = 12


def serving(fhehip, torch, rate=SERVE_RATE, seconds=SERVE_SECONDS, seed=6, lanes=4, max_batch=64,
            budget_gib=64):
    import numpy as np
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    kinds = []
    t0 = time.perf_counter()
    for n, moduli in ((16384, MODULI), (32768, fhehip.generate_moduli([60] * 9, 32768))):
        ct_moduli = moduli[:-1]
        cctx, kctx = fhehip.Context(ct_moduli, n), fhehip.Context(moduli, n)
        krng = np.random.default_rng(n)
        mult = fhehip.Multiplicator(rk=_random_ksk(fhehip, np, krng, cctx, kctx, ct_moduli, moduli, n),
                                    plaintext=T_PLAIN)
        gks = [fhehip.GaloisKey(_random_ksk(fhehip, np, krng, cctx, kctx, ct_moduli, moduli, n),
                                pow(3, 1 << k, 2 * n)) for k in range(SERVE_LOG2_SLOTS)]
        L = len(ct_moduli)
        pool = 16
        kinds.append({"n": n, "L": L, "w": 2 * L * n, "mult": mult, "ek": fhehip.EvaluationKey(gks),
                      "ctx": (cctx, kctx), "pool": pool,
                      "lhs": _uniform_rows(torch, dev, ct_moduli, (pool, 2, L, n), gen).view(-1),
                      "rhs": _uniform_rows(torch, dev, ct_moduli, (pool, 2, L, n), gen).view(-1)})
    setup_s = time.perf_counter() - t0
    # arrival schedule
    count = int(rate * seconds)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    which = (rng.random(count) >= SERVE_SHARE_16K).astype(int)
    for k, K in enumerate(kinds):
        K["out"] = torch.empty((int((which == k).sum()) + max_batch) * K["w"], dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    srv = fhehip.Server(hbm_budget_bytes=budget_gib << 30, max_batch=max_batch, lanes=lanes)
    pids = []
    for K in kinds:
        pid, fp = srv.add_params(K["mult"], K["ek"], SERVE_LOG2_SLOTS)
        pids.append(pid)
        K["footprint"] = fp

    def inp(K, j):
        i = j % K["pool"]
        w = K["w"]
        return K["lhs"][i * w:(i + 1) * w], K["rhs"][i * w:(i + 1) * w]

    # warm-up: one full batch of each kind per lane sizes every scratch buffer
    for _ in range(lanes):
        for k, K in enumerate(kinds):
            for j in range(max_batch):
                a, b = inp(K, j)
                srv.submit(pids[k], a, b, K["out"][:K["w"]])
    srv.drain(300)
    base = srv.stats()
    # timed: open-loop Poisson arrivals
    slot = [max_batch, max_batch]
    tickets = []
    start = time.perf_counter()
    for j in range(count):
        target = start + arrivals[j]
        while True:
            d = target - time.perf_counter()
            if d <= 0:
                break
            if d > 2e-4:
                time.sleep(d - 1.5e-4)
        k = int(which[j])
        K = kinds[k]
        a, b = inp(K, j)
        o = K["out"][slot[k] * K["w"]:(slot[k] + 1) * K["w"]]
        slot[k] += 1
        tickets.append((k, srv.submit(pids[k], a, b, o), j, o))
    submit_end = time.perf_counter()
    srv.drain(600)
    wall = time.perf_counter() - start
    st = srv.stats()
    lat = {0: [], 1: []}
    for k, t, j, o in tickets:
        state, us = srv.query(t)
        if state == 2:
            lat[k].append(us / 1e3)
    all_lat = lat[0] + lat[1]
    # a few served results recomputed alone with direct calls
    exact = True
    stream = torch.cuda.current_stream()
    for k, t, j, o in tickets[:: max(1, len(tickets) // 6)]:
        K = kinds[k]
        a, b = inp(K, j)
        cur = torch.empty_like(o)
        res = torch.empty_like(o)
        K["mult"].multiply(a, b, cur, 1, stream)
        K["ek"].rotate_and_sum(cur, res, SERVE_LOG2_SLOTS, False, 1, stream)
        torch.cuda.synchronize()
        exact = exact and bool(torch.equal(res, o))
    srv.close()
    done = st["completed"] - base["completed"]
    return {"workload": f"Poisson arrivals, mean {rate:.0f} req/s for {seconds:.0f} s: ct*ct mul+relin + "
                        f"{SERVE_LOG2_SLOTS} rotate-and-add steps (4096-slot sum); "
                        f"{int(SERVE_SHARE_16K * 100)}% N=16384 (6x60-bit + special), "
                        f"{100 - int(SERVE_SHARE_16K * 100)}% N=32768 (8x60-bit + special)",
            "offered_rps": count / arrivals[-1], "value": done / wall, "unit": "req/s",
            "p50_ms": percentile(all_lat, 50), "p99_ms": percentile(all_lat, 99),
            "p50_ms_n16384": percentile(lat[0], 50), "p99_ms_n16384": percentile(lat[0], 99),
            "p50_ms_n32768": percentile(lat[1], 50), "p99_ms_n32768": percentile(lat[1], 99),
            "requests": count, "completed": done, "failed": st["failed"],
            "batches": st["batches"] - base["batches"], "max_batch_seen": st["max_batch_seen"],
            "peak_admitted_gib": st["peak_bytes"] / 2 ** 30, "budget_gib": budget_gib,
            "footprint_mib": [K["footprint"] / 2 ** 20 for K in kinds], "lanes": lanes, "max_batch": max_batch,
            "submit_lag_ms": (submit_end - start - arrivals[-1]) * 1e3,
            "bit_exact_vs_direct_calls": exact, "setup_s": setup_s}


# MulPIR (examples/mulpir.rs:35-37) over the PIR examples' default database
# (examples/pir.rs: 65536 elements of 1024 bytes).
PIR_N = 8192
PIR_T = (1 << 20) + (1 << 19) + (1 << 17) + (1 << 16) + (1 << 14) + 1
PIR_MODULI_SIZES = [50, 55, 55]
PIR_ELEMENTS, PIR_ELEMENT_BYTES = 65536, 1024


def pir_dims():
    """encode_database (examples/util.rs:95-123): plaintexts of
    floor(log2 t) N / (8 element_bytes) elements, as a dim1 x dim2 matrix."""
    per_pt = (PIR_T.bit_length() - 1) * PIR_N // (PIR_ELEMENT_BYTES * 8)
    rows = -(-PIR_ELEMENTS // per_pt)
    d1 = 1
    while d1 * d1 < rows:
        d1 += 1
    return d1, -(-rows // d1)


def _random_rows(np, rng, ct_moduli, key_moduli, n):
    """Key-switching key rows c0, c1 [digits = len(ct_moduli)][len(key_moduli)][N]."""
    c = []
    for _ in range(2):
        rows = np.empty((len(ct_moduli), len(key_moduli), n), dtype=np.uint64)
        for j, q in enumerate(key_moduli):
            rows[:, j, :] = rng.integers(0, q, (len(ct_moduli), n), dtype=np.uint64)
        c.append(rows)
    return c


def pir_server(fhehip, torch, reps=20, check_cpu=True, batches=(16, 64)):
    """MulPIR's server response (examples/mulpir.rs:151-185) on the examples'
    default database, a 58 x 57 matrix of N = 8192 plaintexts at level 1
    (433 MiB in HBM): expansion of the query into dim1 + dim2 ciphertexts,
    the dim2 column dot products over the database in one pass, the dim2
    ct x ct products (no relinearization) summed, one relinearization and
    the switch to the last level, one fhehip_pir_respond call. Inputs and
    keys are seeded uniform rows.
    Then `batches` queries at once (expansions query-minor, so every stage is
    one batch and one database pass serves all of them); query 0's batched
    response is compared with the single one. Returns the result and, with
    check_cpu, the host copies pir_cpu_baseline recomputes the response from."""
    import numpy as np
    dev = torch.device("cuda:0")
    n, t = PIR_N, PIR_T
    d1, d2 = pir_dims()
    moduli = fhehip.generate_moduli(PIR_MODULI_SIZES, n)
    m1, m2 = moduli[:2], moduli[:1]
    L = len(m1)
    per = 2 * L * n
    ctx0, ctx1, ctx2 = fhehip.Context(moduli, n), fhehip.Context(m1, n), fhehip.Context(m2, n)
    rng = np.random.default_rng(11)
    levels = (d1 + d2 - 1).bit_length()
    gk_rows = [((n >> l) + 1, *_random_rows(np, rng, m1, moduli, n)) for l in range(levels)]
    gks = [fhehip.GaloisKey(fhehip.KeySwitchingKey(ctx1, ctx0, c0, c1), e) for e, c0, c1 in gk_rows]
    ek = fhehip.EvaluationKey(gks)
    rk_rows = _random_rows(np, rng, m1, m1, n)
    rk = fhehip.KeySwitchingKey(ctx1, ctx1, *rk_rows)
    ext = fhehip.Multiplicator(rk=rk, plaintext=t).extended_basis()
    mul = fhehip.Multiplicator(base_ctx=ctx1, extended_basis=ext, post=(t, m1[0] * m1[1]))
    sw = fhehip.Switcher(ctx1, ctx2)
    gen = torch.Generator(device=dev)
    gen.manual_seed(12)
    db = _uniform_rows(torch, dev, m1, (d1 * d2, L, n), gen).reshape(-1)
    query = _uniform_rows(torch, dev, m1, (2, L, n), gen).reshape(-1)
    resp = torch.empty(2 * n, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()

    def respond():
        fhehip.pir_respond(ek, mul, rk, sw, db, d1, d2, query, 1, resp, stream)

    for _ in range(3):
        respond()
    torch.cuda.synchronize()
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    evs[0].record(stream)
    for k in range(reps):
        respond()
        evs[k + 1].record(stream)
    torch.cuda.synchronize()
    ms = [evs[k].elapsed_time(evs[k + 1]) for k in range(reps)]
    fhehip.profiler_enable(True)
    respond()
    torch.cuda.synchronize()
    kern = fhehip.profiler_report()
    fhehip.profiler_enable(False)
    dot_name = f"dot_product/N={n}"
    dot_ms = kern[dot_name][1] / kern[dot_name][0] if dot_name in kern else None
    single = fhehip.to_host(resp).copy()

    def batched(nq):
        """The same response for nq queries (query 0 = the single query)."""
        qs = torch.cat([query, _uniform_rows(torch, dev, m1, (nq - 1, 2, L, n), gen).reshape(-1)])
        rs = torch.empty(nq * 2 * n, dtype=torch.int64, device=dev)

        def run():
            fhehip.pir_respond(ek, mul, rk, sw, db, d1, d2, qs, nq, rs, stream)

        run()
        torch.cuda.synchronize()
        b0, b1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        k = max(2, reps // 4)
        b0.record(stream)
        for _ in range(k):
            run()
        b1.record(stream)
        torch.cuda.synchronize()
        bms = b0.elapsed_time(b1) / k
        fhehip.profiler_enable(True)
        run()
        torch.cuda.synchronize()
        kb = fhehip.profiler_report()
        fhehip.profiler_enable(False)
        d_ms = kb[dot_name][1] / kb[dot_name][0] if dot_name in kb else None
        d_bytes = (d1 * d2 * L * n + (d1 + d2) * nq * per) * 8
        return {"queries": nq, "ms_per_batch": bms, "value": nq * 1e3 / bms, "unit": "responses/s",
                "dot_product_ms": d_ms, "dot_product_GBps": d_bytes / d_ms / 1e6 if d_ms else None,
                "query0_equals_single": bool(np.array_equal(fhehip.to_host(rs)[:2 * n], single))}
    dot_bytes = (d1 * d2 * L * n + d1 * per + d2 * per) * 8
    res = {"workload": f"MulPIR server response (examples/mulpir.rs:151-185): database of {PIR_ELEMENTS} x "
                       f"{PIR_ELEMENT_BYTES}-byte elements as {d1} x {d2} plaintexts, N={n}, moduli "
                       f"{PIR_MODULI_SIZES} at level 1, t={t}; expand to {d1 + d2}, {d2} column dot products, "
                       f"{d2} ct x ct products summed, relinearize, switch to level 2",
           "value": 1e3 / percentile(ms, 50), "unit": "responses/s", "p50_ms": percentile(ms, 50),
           "p99_ms": percentile(ms, 99), "reps": reps, "database_bytes": d1 * d2 * L * n * 8,
           "kernels_ms": {k: {"launches": c, "avg_ms": v / c} for k, (c, v) in sorted(kern.items())},
           "dot_product_roofline": {"bound": "hbm", "algorithmic_bytes": dot_bytes, "avg_launch_ms": dot_ms,
                                    "achieved_GBps": dot_bytes / dot_ms / 1e6 if dot_ms else None,
                                    "frac": dot_bytes / dot_ms / 1e6 / HBM_PEAK_GBS if dot_ms else None},
           "batched": [batched(nq) for nq in batches]}
    # host copies for the CPU baseline leg (pir_cpu_baseline)
    handover = {"moduli": moduli, "n": n, "t": t, "d1": d1, "d2": d2, "gk_rows": gk_rows, "rk_rows": rk_rows,
                "db": fhehip.to_host(db), "query": fhehip.to_host(query), "resp": fhehip.to_host(resp),
                "gpu_ms": percentile(ms, 50)} if check_cpu else None
    return res, handover


def codec_throughput(fhehip, torch, batch=512, reps=10):
    """The steps either side of the evaluation path at the headline shape
    (N=16384, the six 60-bit ciphertext moduli, t=65537), batched: SIMD
    encoding with Plaintext::to_poly plus SecretKey::encrypt_poly (samples
    a, e generated on the device), and try_decrypt with SIMD decoding."""
    import numpy as np
    dev = torch.device("cuda:0")
    L = len(MODULI) - 1
    ctx = fhehip.Context(MODULI[:L], N)
    plain = fhehip.Context(MODULI[:2], N)   # 120 bits >= log2(t) + 60 (parameters.rs:479-494)
    rng = np.random.default_rng(3)
    sk = rng.integers(-1, 2, N).astype(np.int64)
    codec = fhehip.Codec(ctx, plain, T_PLAIN, sk=sk)
    gen = torch.Generator(device=dev)
    gen.manual_seed(9)
    vals = torch.randint(0, T_PLAIN, (batch * N,), dtype=torch.int64, device=dev, generator=gen)
    a = _uniform_rows(torch, dev, MODULI[:L], (batch, L, N), gen).reshape(-1)
    e = torch.randint(-16, 17, (batch * N,), dtype=torch.int64, device=dev, generator=gen)
    pt = torch.empty(batch * L * N, dtype=torch.int64, device=dev)
    ct = torch.empty(batch * 2 * L * N, dtype=torch.int64, device=dev)
    dec = torch.empty(batch * N, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()

    def enc():
        codec.encode(vals, batch, pt, simd=True, scaled=True, stream=stream)
        codec.encrypt(pt, a, e, ct, batch, stream)

    def dcr():
        codec.decrypt(ct, 2, batch, dec, simd=True, stream=stream)

    out = {"workload": f"{batch} x (SIMD encode + secret-key encryption) and {batch} x (decryption + SIMD decode), "
                       f"N={N}, {L} x 60-bit moduli, t={T_PLAIN}", "batch": batch}
    for name, fn in (("encode_encrypt", enc), ("decrypt_decode", dcr)):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = {"ms_per_batch": ms, "value": batch / (ms / 1e3), "unit": "ciphertexts/s"}
    out["round_trip_exact"] = bool(torch.equal(dec, vals))
    # Rq wire format of the ciphertexts (both polynomials, NTT domain -> bytes and back)
    per = fhehip.serialization_length(ctx)
    wire = torch.empty(batch * 2 * per, dtype=torch.uint8, device=dev)
    back = torch.empty_like(ct)
    for name, fn in (("serialize", lambda: fhehip.serialize(ctx, ct, 2 * batch, wire, True, stream)),
                     ("deserialize", lambda: fhehip.deserialize(ctx, wire, 2 * batch, back, True, stream))):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = {"ms_per_batch": ms, "value": batch / (ms / 1e3), "unit": "ciphertexts/s",
                     "wire_bytes_per_ciphertext": 2 * per}
    out["wire_round_trip_exact"] = bool(torch.equal(back, ct))
    return out


def pir_cpu_baseline(h):
    """CPU baseline of the PIR leg: the oracle (C++ restatement of the fhe.rs
    CPU path) recomputes the single-query response on one core, for a bit
    comparison with the GPU's and its time."""
    import numpy as np
    oracle, build, _ = oracle_module()
    moduli, n, t, d1, d2 = h["moduli"], h["n"], h["t"], h["d1"], h["d2"]
    m1, m2 = moduli[:2], moduli[:1]
    L = len(m1)
    per = 2 * L * n
    bfv = oracle.Bfv(n, t, moduli, sk_seed=1)
    okeys = []
    for e, c0, c1 in h["gk_rows"]:
        k = bfv.key_from_rows(c0, c1, ct_level=1, key_level=0)
        k.exponent = e
        okeys.append(k)
    ork = bfv.key_from_rows(*h["rk_rows"], ct_level=1, key_level=1)
    o1, o2 = oracle.Context(m1, n), oracle.Context(m2, n)
    om = oracle.Multiplicator(bfv, None, level=1)
    qv = np.array(m1, dtype=np.uint64).repeat(n)
    db_h = h["db"]
    t0 = time.perf_counter()
    ex = bfv.expands(okeys, h["query"], d1 + d2, level=1).reshape(d1 + d2, per)
    acc = None
    for j in range(d2):
        dj = oracle.dot_product_scalar(o1, ex[:d1].reshape(-1), db_h[j * L * n:], d1, 2, pt_stride=d2 * L * n)
        pj = om.multiply(dj, ex[d1 + j]).reshape(3, L * n)
        if acc is None:
            acc = pj
        else:
            acc = acc + pj
            acc = np.where(acc >= qv, acc - qv, acc)
    want = oracle.ct_switch_down(o1, o2, bfv.relinearize(ork, acc.reshape(-1)))
    cpu_s = time.perf_counter() - t0
    return {"bit_exact_vs_oracle": bool(np.array_equal(h["resp"], want)),
            "cpu_reference": {"seconds_per_response": cpu_s, "cores": 1, "kind": "port",
                              "sample": f"one response by the oracle (C++ restatement, {build}, one thread)"},
            "gpu_over_cpu": cpu_s * 1e3 / h["gpu_ms"]}


def cpu_baseline(lhs_h, rhs_h, c0, c1, gpu_out0, budget_s):
    """Oracle (C++ restatement of the fhe.rs CPU path) on every host core this
    process may use, one independent Multiplicator::multiply per thread at a
    time (fhe.rs itself is single-threaded per operation)."""
    import numpy as np
    oracle, build, builds_ms = oracle_module()
    bfv = oracle.Bfv(N, T_PLAIN, MODULI, sk_seed=1)
    rk = bfv.key_from_rows(c0, c1, ct_level=1, key_level=0)
    om = oracle.Multiplicator(bfv, rk, level=1)
    ref = om.multiply(lhs_h, rhs_h)
    bit_exact = bool(np.array_equal(ref, gpu_out0))
    # one op to size the sample
    t0 = time.perf_counter()
    om.multiply(lhs_h, rhs_h)
    one = time.perf_counter() - t0
    host = host_cpu()
    threads = max(1, host["usable_cores"])
    reps = max(1, int(budget_s / max(one, 1e-3) / 1.5))
    wall, lat0 = om.bench(lhs_h, rhs_h, threads, reps)
    return {
        "value": threads * reps / wall,
        "unit": UNIT,
        "cores": threads,
        "kind": "port",
        "sample": f"{threads} threads (every core this process may use) x {reps} Multiplicator::multiply each "
                  f"at the same shape (oracle C++ restatement, {build}); single-op p50 "
                  f"{percentile(list(lat0 * 1e3), 50):.1f} ms",
        "p50_ms_1core": percentile(list(lat0 * 1e3), 50),
        "build": build,
        "single_op_ms_by_build": builds_ms,
        "host": host,
    }, bit_exact


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--per-call-ops", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the NTT autotuning and the secondary workloads (single instance, NTT, uneven, serving)")
    ap.add_argument("--only", choices=["uneven", "serve", "pir", "codec", "ntt"], default=None,
                    help="run only this secondary workload and print its JSON object")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="run one device copy of this many MiB (reference for rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import fhehip

    def copy_calibration():
        """One plain device copy of --copy-calibration MiB, the reference
        dispatch of the rocprofv3 byte counters (it runs last)."""
        if not args.copy_calibration:
            return None
        src = torch.empty(args.copy_calibration << 17, dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        return args.copy_calibration << 20

    if args.only == "ntt":
        res = ntt_microbench(fhehip, torch, args.steps)
        res["copy_calibration_bytes"] = copy_calibration()
        print(json.dumps(res))
        return
    if args.only == "uneven":
        print(json.dumps(uneven_batch(fhehip, torch, reps=args.steps)))
        return
    if args.only == "serve":
        print(json.dumps(serving(fhehip, torch)))
        return
    if args.only == "codec":
        print(json.dumps(codec_throughput(fhehip, torch)))
        return
    if args.only == "pir":
        res, handover = pir_server(fhehip, torch, check_cpu=not args.no_cpu)
        if handover is not None:
            res.update(pir_cpu_baseline(handover))
        print(json.dumps(res))
        return
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev = torch.device("cuda:0")
    B, L, Lk = args.batch, len(MODULI) - 1, len(MODULI)

    # Resident contexts and a relinearization key with seeded uniform rows
    # (the key values do not change the work; they are checked by the oracle).
    rng = np.random.default_rng(2024)
    cctx = fhehip.Context(MODULI[:L], N)
    kctx = fhehip.Context(MODULI, N)
    c0 = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in MODULI]) for _ in range(L)])
    c1 = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in MODULI]) for _ in range(L)])
    ksk = fhehip.KeySwitchingKey(cctx, kctx, c0, c1)
    mult = fhehip.Multiplicator(rk=ksk, plaintext=T_PLAIN)
    M = len(mult.extended_basis())
    # Autotuned shape of the plain NTT kernels at this degree (16 or 32 values
    # per thread, per direction), timed on 2048 rows of the multiplication context.
    # (Skipped with --no-extras, so that profiled runs hold only the pipeline's launches.)
    ntt_tuning = None
    if not args.no_extras:
        tune_f, tune_i, tune_ms = fhehip.Context(mult.extended_basis(), N).ntt_autotune(2048)
        ntt_tuning = {"forward_log2_values_per_thread": tune_f, "inverse_log2_values_per_thread": tune_i,
                      "ms_per_launch": {"fwd16": tune_ms[0], "fwd32": tune_ms[1], "inv16": tune_ms[2],
                                        "inv32": tune_ms[3]}, "rows": 2048}

    # Seeded uniform ciphertext residues, generated on the device.
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    per = 2 * L * N

    def ct_batch(count):
        t = torch.empty(count * per, dtype=torch.int64, device=dev)
        v = t.view(count, 2, L, N)
        for i, q in enumerate(MODULI[:L]):
            v[:, :, i, :] = torch.randint(0, q, (count, 2, N), dtype=torch.int64, device=dev, generator=g)
        return t

    lhs, rhs = ct_batch(B), ct_batch(B)
    out = torch.empty(B * per, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()

    # Autotuned launch shapes of the fused kernels at this batch (skipped with
    # --no-extras, so that profiled runs hold only the pipeline's launches);
    # the timed steps then run at the headline shapes.
    fused_tuning = None
    if not args.no_extras:
        chosen, tms = fhehip.multiply_autotune(mult, B)
        fused_tuning = {"chosen": chosen, "ms_per_multiply": tms, "instances": B}
    fhehip.tuning_set(N, **HEADLINE_TUNING)

    for _ in range(args.warmup):
        mult.multiply(lhs, rhs, out, B, stream)
    torch.cuda.synchronize()

    # Eager steps: K batched fhehip_multiply calls with per-step events and
    # per-kernel events (the roofline figures).
    fhehip.profiler_enable(True)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record(stream)
    for k in range(args.steps):
        mult.multiply(lhs, rhs, out, B, stream)
        evs[k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernels = fhehip.profiler_report()
    fhehip.profiler_enable(False)
    eager_step_ms = [evs[k].elapsed_time(evs[k + 1]) for k in range(args.steps)]
    eager_total_ms = evs[0].elapsed_time(evs[-1])
    eager_value = B * args.steps / (eager_total_ms / 1e3)

    # Timed steps: the same batched step captured once as a HIP graph (its
    # ten launches) and replayed K times; each replay is checked against the
    # eager output.
    gout = torch.empty_like(out)
    torch.cuda.synchronize()
    graph = fhehip.Graph(mult, lhs, rhs, gout, B)
    for _ in range(args.warmup):
        graph.launch(stream)
    torch.cuda.synchronize()
    gev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    gev[0].record(stream)
    for k in range(args.steps):
        graph.launch(stream)
        gev[k + 1].record(stream)
    torch.cuda.synchronize()
    step_ms = [gev[k].elapsed_time(gev[k + 1]) for k in range(args.steps)]
    total_ms = gev[0].elapsed_time(gev[-1])
    graph_equals_eager = bool(torch.equal(gout, out))
    graph.close()
    value = B * args.steps / (total_ms / 1e3)
    # The autotuner's choice, when it differs, timed the same way.
    if fused_tuning is not None:
        fused_tuning["headline_shapes"] = dict(HEADLINE_TUNING)
        if fused_tuning["chosen"] != HEADLINE_TUNING:
            fhehip.tuning_set(N, **fused_tuning["chosen"])
            graph = fhehip.Graph(mult, lhs, rhs, gout, B)
            for _ in range(args.warmup):
                graph.launch(stream)
            torch.cuda.synchronize()
            aev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            aev[0].record(stream)
            for k in range(args.steps):
                graph.launch(stream)
            aev[1].record(stream)
            torch.cuda.synchronize()
            a_ms = aev[0].elapsed_time(aev[1])
            fused_tuning["chosen_shapes"] = {"value": B * args.steps / (a_ms / 1e3), "unit": UNIT,
                                             "ms_per_step": a_ms / args.steps,
                                             "equals_headline_output": bool(torch.equal(gout, out))}
            graph.close()
            fhehip.tuning_set(N, **HEADLINE_TUNING)
    del gout

    # Per-call GPU drop-in: one operation per call, into its own output buffer.
    call_ms, per_call_value = [], None
    per_call_graph = None
    per_call_host = None
    if args.per_call_ops > 0:
        one_o = torch.empty(per, dtype=torch.int64, device=dev)
        for _ in range(3):
            mult.multiply(lhs[:per], rhs[:per], one_o, 1, stream)
        torch.cuda.synchronize()
        pe = [torch.cuda.Event(enable_timing=True) for _ in range(args.per_call_ops + 1)]
        pe[0].record(stream)
        for k in range(args.per_call_ops):
            i = k % B
            mult.multiply(lhs[i * per:(i + 1) * per], rhs[i * per:(i + 1) * per], one_o, 1, stream)
            pe[k + 1].record(stream)
        torch.cuda.synchronize()
        call_ms = [pe[k].elapsed_time(pe[k + 1]) for k in range(args.per_call_ops)]
        per_call_value = args.per_call_ops / (pe[0].elapsed_time(pe[-1]) / 1e3)
        # the same single-instance call captured once as a HIP graph and replayed
        graph = fhehip.Graph(mult, lhs[:per], rhs[:per], one_o, 1)
        for _ in range(3):
            graph.launch(stream)
        torch.cuda.synchronize()
        ge = [torch.cuda.Event(enable_timing=True) for _ in range(args.per_call_ops + 1)]
        ge[0].record(stream)
        for k in range(args.per_call_ops):
            graph.launch(stream)
            ge[k + 1].record(stream)
        torch.cuda.synchronize()
        graph_ms = [ge[k].elapsed_time(ge[k + 1]) for k in range(args.per_call_ops)]
        per_call_graph = {"value": args.per_call_ops / (ge[0].elapsed_time(ge[-1]) / 1e3), "unit": UNIT,
                          "p50_ms": percentile(graph_ms, 50), "p99_ms": percentile(graph_ms, 99)}
        graph.close()
        # The same call as a drop-in behind an API that holds ciphertexts in
        # host memory (as fhe.rs does): pinned host inputs copied in, the
        # product copied back, and the call waited on. Reported beside the
        # device-resident figures; `value` and batched_over_per_call use those.
        h_l = torch.empty(per, dtype=torch.int64, pin_memory=True)
        h_r = torch.empty(per, dtype=torch.int64, pin_memory=True)
        h_o = torch.empty(per, dtype=torch.int64, pin_memory=True)
        d_l = torch.empty(per, dtype=torch.int64, device=dev)
        d_r = torch.empty(per, dtype=torch.int64, device=dev)
        h_l.copy_(lhs[:per])
        h_r.copy_(rhs[:per])
        host_ms = []
        for k in range(args.per_call_ops + 3):
            t1 = time.perf_counter()
            d_l.copy_(h_l, non_blocking=True)
            d_r.copy_(h_r, non_blocking=True)
            mult.multiply(d_l, d_r, one_o, 1, stream)
            h_o.copy_(one_o, non_blocking=True)
            torch.cuda.synchronize()
            if k >= 3:
                host_ms.append((time.perf_counter() - t1) * 1e3)
        per_call_host = {"value": len(host_ms) / (sum(host_ms) / 1e3), "unit": UNIT,
                         "p50_ms": percentile(host_ms, 50), "p99_ms": percentile(host_ms, 99),
                         "bytes_per_call_pcie": (3 * per) * 8,
                         "bit_exact_vs_device_resident": bool(torch.equal(h_o, one_o.cpu()))}

    # Roofline of the dominant kernel. The key-switch finish reads less when
    # ks_z_pair formed its first forward stage (the paired, half-row finish).
    paired = any(k.startswith("ks_z_pair/") for k in kernels)

    def step_bytes(kname):
        if paired and kname.startswith("ks_finish/"):
            kname = "ks_finish_paired"
        return kernel_bytes(kname, B, L, Lk, M)

    dom = max(kernels.items(), key=lambda kv: kv[1][1])
    dname, (dcalls, dms) = dom
    avg_s = dms / dcalls / 1e3
    abytes = step_bytes(dname)
    achieved = abytes / avg_s / 1e9 if abytes else None
    digest = kernel_sources_digest()
    provenance = profile_provenance(digest)
    traffic = load_traffic(dname.split("/")[0], digest)
    valu_busy, valu_rate = load_valu(dname.split("/")[0], digest)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                "kernel": dname, "avg_launch_ms": avg_s * 1e3, "algorithmic_bytes": abytes,
                "valu_busy_pct": valu_busy, "valu_wave_instructions_per_cu_cycle": valu_rate,
                "counter_profiles": provenance}

    # The same figures for every kernel of the step (HBM-bound accounting; the
    # VALU columns come from the committed counter summary).
    kernel_roofline = {}
    for kname, (kc, kms) in sorted(kernels.items()):
        kb = step_bytes(kname)
        ks = kms / kc / 1e3
        vb, vr = load_valu(kname.split("/")[0], digest)
        kernel_roofline[kname] = {"avg_launch_ms": ks * 1e3, "algorithmic_bytes": kb,
                                  "achieved_GBps": kb / ks / 1e9 if kb else None,
                                  "hbm_frac": kb / ks / 1e9 / HBM_PEAK_GBS if kb else None,
                                  "traffic": load_traffic(kname.split("/")[0], digest), "valu_busy_pct": vb}

    extras = {}
    pir_handover = None
    if not args.no_extras:
        extras["single_instance"] = single_instance_latency(fhehip, torch, 200)
        extras["ntt_microbench"] = ntt_microbench(fhehip, torch, 50)
        extras["uneven_batch"] = uneven_batch(fhehip, torch)
        extras["serving"] = serving(fhehip, torch)
        extras["pir_server"], pir_handover = pir_server(fhehip, torch, check_cpu=not args.no_cpu)
        extras["codec"] = codec_throughput(fhehip, torch)
    if args.copy_calibration:
        extras["copy_calibration_bytes"] = copy_calibration()

    cpu = None
    bit_exact = None
    if not args.no_cpu and not args.no_extras and pir_handover is not None:
        extras["pir_server"].update(pir_cpu_baseline(pir_handover))
    if not args.no_cpu:
        out0 = fhehip.to_host(out[:per]).copy()
        cpu, bit_exact = cpu_baseline(fhehip.to_host(lhs[:per]).copy(), fhehip.to_host(rhs[:per]).copy(),
                                      c0, c1, out0, args.cpu_seconds)

    res = {
        "metric": METRIC,
        "value": value,
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: seeded uniform ciphertext residues and relinearization-key rows",
        "config": {"workload": "batched uniform BFV: 512 x ct*ct multiply+relinearize, N=16384, "
                               "6x60-bit ciphertext primes + 1 special prime, t=65537",
                   "batch": B, "degree": N, "ct_moduli": L, "key_moduli": Lk, "mul_moduli": M,
                   "parallelism": "single GPU"},
        "p50_step_ms": percentile(step_ms, 50),
        "p99_step_ms": percentile(step_ms, 99),
        "step": "one HIP-graph replay of the batched fhehip_multiply (captured once before the timed region)",
        "launch_shapes": HEADLINE_TUNING,
        "graph_equals_eager": graph_equals_eager,
        "eager": {"value": eager_value, "unit": UNIT, "ms_per_step": eager_total_ms / args.steps,
                  "p50_step_ms": percentile(eager_step_ms, 50), "p99_step_ms": percentile(eager_step_ms, 99),
                  "note": "fhehip_multiply called per step, with per-kernel timing events"},
        "fused_autotune": fused_tuning,
        "per_call": {"value": per_call_value, "unit": UNIT, "p50_ms": percentile(call_ms, 50),
                     "p99_ms": percentile(call_ms, 99), "ops": args.per_call_ops},
        "per_call_graph": per_call_graph,
        "per_call_host_buffers": per_call_host,
        "batched_over_per_call_host_buffers": (value / per_call_host["value"]) if per_call_host else None,
        "batched_over_per_call": (value / per_call_value) if per_call_value else None,
        "bit_exact_vs_oracle": bit_exact,
        "kernels_ms": {k: {"launches": c, "avg_ms": ms / c} for k, (c, ms) in sorted(kernels.items())},
        "ntt_autotune": ntt_tuning,
        "roofline": roofline,
        "kernel_roofline": kernel_roofline,
        "cpu_baseline": cpu,
        **extras,
        "gpu_over_cpu": (value / cpu["value"]) if cpu else None,
        "host_wall_s": wall,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
