// CPU check of the matrix-core scaler's digit tables (scale_mfma.h) against
// the oracle's RnsScaler::scale (math/src/rns/scaler.rs:249-352): for each
// scaler shape and MFMA depth, the tables the library builds are multiplied
// with the balanced base-256 digits of random reduced columns exactly as the
// kernel does (pass 1 over the residue digits, S -> v, W -> signed w, the
// output tiles' v and w step, the offset M and one reduction), and every
// output is compared with the oracle's. Columns 0..3 hold all zeros, all
// p - 1, alternating zeros and p - 1, and p / 2.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "bfv.h"
#include "scale_mfma.h"

using namespace fhehip;

extern "C" {
void* or_rns_scaler_new(const uint64_t*, uint32_t, const uint64_t*, uint32_t, const uint32_t*, uint32_t,
                        const uint32_t*, uint32_t);
void or_rns_scaler_free(void*);
int or_rns_scale(void*, const uint64_t*, uint32_t, uint64_t*, uint32_t, uint32_t);
}

namespace {

constexpr u64 kBias = 0x8080808080808080ull;

// Byte k of row `row` of the table block at `base` (MFMA steps of kw k).
int8_t table_byte(const std::vector<u64>& t, size_t base, int kw, int row, int k) {
    const int kappa = k / kw, h = (k % kw) / (kw / 2), e = k % (kw / 2);
    const size_t word = base + ((size_t)kappa * 64 + 32 * h + row) * (kw / 16) + e / 8;
    return (int8_t)(uint8_t)(t[word] >> (8 * (e % 8)));
}

void digits(u128 x, int8_t* d, int nd) {  // x + bias below 2^(8 nd)
    u128 bias = 0;
    for (int e = 0; e < nd; e++) bias |= (u128)0x80 << (8 * e);
    const u128 y = x + bias;
    for (int e = 0; e < nd; e++) d[e] = (int8_t)((uint8_t)(y >> (8 * e)) ^ 0x80);
}

struct W256 {
    u64 w[4];
};

W256 combine(const long long* p) {  // sum_b p_b 256^b over 32 rows, mod 2^256
    u32 d[8];
    long long carry = 0;
    for (int u = 0; u < 8; u++) {
        long long acc = carry;
        for (int v = 0; v < 4; v++) acc += p[4 * u + v] * (1ll << (8 * v));
        d[u] = (u32)acc;
        carry = acc >> 32;
    }
    W256 r;
    for (int k = 0; k < 4; k++) r.w[k] = d[2 * k] | ((u64)d[2 * k + 1] << 32);
    return r;
}

u128 shr_lo128(const W256& a, unsigned s) {
    const unsigned ws = s >> 6, bs = s & 63;
    auto word = [&](unsigned k) -> u64 { return k < 4 ? a.w[k] : 0; };
    const u64 x0 = word(ws), x1 = word(ws + 1), x2 = word(ws + 2);
    const u64 r0 = bs ? ((x0 >> bs) | (x1 << (64 - bs))) : x0;
    const u64 r1 = bs ? ((x1 >> bs) | (x2 << (64 - bs))) : x1;
    return ((u128)r1 << 64) | r0;
}

int check(const std::vector<u64>& from, const std::vector<u64>& to, u32 start, u64 num, bool t_over_q, int kw,
          int cols, u32 seed) {
    RnsBasis fb, tb;
    fb.init(from);
    tb.init(to);
    BigUint bn(num), bd(1);
    if (t_over_q)
        for (u64 q : to) bd = bd * BigUint(q);
    ScalerConsts c;
    c.build(fb, tb, bn, bd);
    std::vector<ModInfo> fm(from.size()), tm(to.size());
    for (size_t i = 0; i < from.size(); i++) {
        std::memset(&fm[i], 0, sizeof(ModInfo));
        fm[i].p = from[i];
        barrett128(from[i], fm[i].barrett_lo, fm[i].barrett_hi);
    }
    for (size_t i = 0; i < to.size(); i++) {
        std::memset(&tm[i], 0, sizeof(ModInfo));
        tm[i].p = to[i];
        barrett128(to[i], tm[i].barrett_lo, tm[i].barrett_hi);
    }
    const std::vector<u64> sc = scaler_blob(c, fm, tm, start);
    u64 pmax = 0;
    for (u64 p : from) pmax = std::max(pmax, p);
    ScaleMfmaLayout lay;
    const std::vector<u64> t = scaler_mfma_blob(sc, pmax, lay, kw);
    if (t.empty() || lay.kw != kw) {
        printf("no table for nf %zu kw %d\n", from.size(), kw);
        return 1;
    }
    const int nf = lay.nf, count = lay.count;
    std::vector<uint32_t> nw = {(uint32_t)num, (uint32_t)(num >> 32)}, dw;
    for (BigUint x = bd; !x.is_zero(); x = x >> 32) dw.push_back((uint32_t)x.limb(0));
    void* h = or_rns_scaler_new(from.data(), (uint32_t)from.size(), to.data(), (uint32_t)to.size(), nw.data(), 2,
                                dw.data(), (uint32_t)dw.size());
    if (!h) return 1;
    std::mt19937_64 rng(seed);
    std::vector<u64> r(nf), want(count);
    int bad = 0;
    for (int col = 0; col < cols; col++) {
        for (int j = 0; j < nf; j++) {
            const u64 p = from[j];
            r[j] = col == 0 ? 0 : col == 1 ? p - 1 : col == 2 ? (j & 1 ? p - 1 : 0) : col == 3 ? p / 2 : rng() % p;
        }
        or_rns_scale(h, r.data(), 1, want.data(), start, count);
        std::vector<int8_t> b(kw * lay.ks1, 0);
        for (int j = 0; j < nf; j++) digits(r[j], &b[8 * j], 8);
        long long P[32];
        auto pass1 = [&](int tile) {
            const size_t base = (size_t)tile * lay.ks1 * 64 * lay.wl();
            for (int row = 0; row < 32; row++) {
                long long s = 0;
                for (int k = 0; k < kw * lay.ks1; k++) s += table_byte(t, base, kw, row, k) * (long long)b[k];
                P[row] = s;
            }
        };
        auto step = [&](size_t base, const int8_t* vw, int kmax) {
            for (int row = 0; row < 32; row++)
                for (int k = 0; k < kmax; k++) P[row] += table_byte(t, base, kw, row, k) * (long long)vw[k];
        };
        pass1(0);
        const u128 vv = shr_lo128(combine(P), (unsigned)sc[SC_SHIFT] - 1);
        const u128 v = (vv >> 1) + (vv & 1);
        int8_t vw[32] = {0};
        digits(v, vw, 16);
        if (lay.has_w) {
            pass1(1);
            step(lay.a2(), vw, 16);
            const W256 so = combine(P);
            const bool neg = (so.w[3] >> 63) != 0;
            u128 w;
            if (neg) {
                W256 x;
                for (int i = 0; i < 4; i++) x.w[i] = ~so.w[i];
                w = (shr_lo128(x, 126) + 1) >> 1;
            } else {
                const u128 ww = shr_lo128(so, 126);
                w = (ww >> 1) + (ww & 1);
            }
            digits(neg ? (u128)0 - w : w, vw + 16, 16);
        }
        for (int o = 0; o < lay.nto; o++) {
            pass1(1 + (lay.has_w ? 1 : 0) + o);
            step(lay.a3() + (size_t)o * lay.ks3() * 64 * lay.wl(), vw, lay.has_w ? 32 : 16);
            for (int s = 0; s < 4 && 4 * o + s < count; s++) {
                const int i = 4 * o + s;
                long long acc = 0;
                for (int q = 0; q < 4; q++) acc += P[8 * s + q] * (1ll << (8 * q));
                const u32 lo32 = (u32)acc;
                acc >>= 32;
                for (int q = 0; q < 4; q++) acc += P[8 * s + 4 + q] * (1ll << (8 * q));
                u128 x = ((u128)(u64)(acc >> 32) << 64) | ((u64)lo32 | ((u64)(u32)acc << 32));
                x += (u128)t[lay.m() + 2 * i] | ((u128)t[lay.m() + 2 * i + 1] << 64);
                if ((x >> 80) != 0) return 1;
                const u64 y = (u64)(x % to[start + i]);
                if (y != want[i] && bad++ < 5)
                    printf("nf %d kw %d column %d output %d: %llu, oracle %llu\n", nf, kw, col, i,
                           (unsigned long long)y, (unsigned long long)want[i]);
            }
        }
    }
    or_rns_scaler_free(h);
    printf("nf %d count %d kw %d has_w %d: %d columns, %d mismatches\n", nf, count, kw, (int)lay.has_w, cols, bad);
    return bad != 0;
}

}  // namespace

int main() {
    const u64 n = 16384;
    std::vector<u64> q60, p62;
    for (u64 u = (u64)1 << 60; q60.size() < 16;) q60.push_back(u = generate_prime64(60, 2 * n, u));
    for (u64 u = (u64)1 << 62; p62.size() < 8;) p62.push_back(u = generate_prime64(62, 2 * n, u));
    int rc = 0;
    // (ciphertext moduli, extension moduli): extension by one, then t/Q back
    const int shapes[][2] = {{6, 7}, {1, 2}, {3, 4}, {8, 8}, {5, 4}};
    for (const auto& sh : shapes) {
        std::vector<u64> q(q60.begin(), q60.begin() + sh[0]), full = q;
        full.insert(full.end(), p62.begin(), p62.begin() + sh[1]);
        for (int kw : {16, 32}) {
            rc |= check(q, full, (u32)q.size(), 1, false, kw, 3000, 1 + sh[0]);
            rc |= check(full, q, 0, 65537, true, kw, 3000, 2 + sh[1]);
        }
    }
    // sixteen 60-bit source moduli down to five
    std::vector<u64> five(q60.begin(), q60.begin() + 5);
    for (int kw : {16, 32}) rc |= check(q60, five, 0, 65537, true, kw, 2000, 9);
    printf(rc ? "FAIL\n" : "OK\n");
    return rc;
}
