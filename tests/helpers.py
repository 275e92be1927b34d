"""Shared setup of the tests (no test functions)."""


def mul_and_relin_2_basis(oracle, moduli, n):
    """The second multiplication strategy of fhe/benches/bfv.rs:257-276: the
    moduli plus ceil(log q / 62) primes of 62 bits below 2^62 - 1; returns
    (extended basis, P, Q)."""
    logq = sum(int(q).bit_length() for q in moduli)
    ext = [int(q) for q in moduli]
    ub = (1 << 62) - 1
    while len(ext) != -(-logq // 62) + len(moduli):
        ub = oracle.generate_prime(62, 2 * n, ub)
        if ub not in ext:
            ext.append(ub)
    Q = P = 1
    for q in ext[:len(moduli)]:
        Q *= q
    for q in ext[len(moduli):]:
        P *= q
    return ext, P, Q


def plaintext_moduli(moduli, t):
    """The plaintext context of BfvParameters (parameters.rs:479-494): the
    first moduli whose sizes sum to at least log2(t) + 60 (at least one)."""
    acc, count = 0, 0
    for q in moduli:
        acc += int(q).bit_length()
        count += 1
        if acc >= int(t).bit_length() + 60:
            break
    return list(moduli[:max(count, 1)])
