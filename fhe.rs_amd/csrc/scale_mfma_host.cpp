// Digit tables of the matrix-core scaler (layout in scale_mfma.h).
#include "scale_mfma.h"

#include <cstdint>
#include <stdexcept>

#include "rns_device.h"

namespace fhehip {

namespace {

// Balanced base-256 digits of sign * mag, each in [-128, 127]: positive
// values take digits in [-128, 127] directly; negative ones are the negated
// digits (in [-127, 128]) of the magnitude.
void balanced_digits(u128 mag, bool neg, int nd, int8_t* out) {
    for (int e = 0; e < nd; e++) {
        int d = (int)(mag & 0xff);
        mag >>= 8;
        if (neg ? d > 128 : d >= 128) {
            d -= 256;
            mag += 1;
        }
        out[e] = (int8_t)(neg ? -d : d);
    }
    if (mag != 0) throw std::logic_error("scaler constant wider than its digits");
}

u64 mulmod(u64 a, u64 b, u64 q) { return (u64)(((u128)a * b) % q); }

// Byte table of one tile over one range of k: rows[32][kw].
struct Tile {
    int kw;
    std::vector<int8_t> a;
    explicit Tile(int k) : kw(k), a((size_t)32 * k, 0) {}
    int8_t& at(int row, int k) { return a[(size_t)row * kw + k]; }
    // The words of K step kappa of an MFMA of kstep k values (16 or 32) in
    // lane order, kstep / 16 per lane.
    void emit(int kappa, int kstep, std::vector<u64>& out) const {
        for (int l = 0; l < 64; l++)
            for (int part = 0; part < kstep / 16; part++) {
                u64 w = 0;
                for (int e = 0; e < 8; e++) {
                    const int k = kstep * kappa + (kstep / 2) * (l / 32) + 8 * part + e;
                    w |= (u64)(uint8_t)a[(size_t)(l % 32) * kw + k] << (8 * e);
                }
                out.push_back(w);
            }
    }
};

// rows b of the digit convolution of one residue with a 17-digit constant:
// A[b][k0 + a] = c[b - a].
void conv_rows(Tile& t, int k0, const int8_t* c, int nc, int na) {
    for (int b = 0; b < 32; b++)
        for (int a = 0; a < na; a++)
            if (b - a >= 0 && b - a < nc) t.at(b, k0 + a) = c[b - a];
}

}  // namespace

std::vector<u64> scaler_mfma_blob(const std::vector<u64>& sc, u64 pmax, ScaleMfmaLayout& lay, int kw) {
    lay = ScaleMfmaLayout{};
    const int nf = (int)sc[SC_NF], count = (int)sc[SC_COUNT];
    if (nf < 1 || nf > 16 || count < 1 || count > 8 || pmax >= ((u64)1 << 62) || sc[SC_SHIFT] < 65 ||
        sc[SC_SHIFT] > 128)
        return {};
    lay.nf = nf;
    lay.count = count;
    if (kw != 16 && kw != 32) kw = nf > 8 ? 32 : 16;
    lay.kw = kw;
    lay.ks1 = (8 * nf + kw - 1) / kw;
    lay.nto = (count + 3) / 4;
    lay.has_w = sc[SC_IS_ONE] == 0;
    const int k1 = kw * lay.ks1;
    std::vector<Tile> t1;
    // S: sum_j r_j theta_garner_j
    {
        Tile t(k1);
        for (int j = 0; j < nf; j++) {
            const u128 tg = (u128)sc[SC_HEADER + 2 * j] | ((u128)sc[SC_HEADER + 2 * j + 1] << 64);
            int8_t c[17];
            balanced_digits(tg, false, 17, c);
            conv_rows(t, 8 * j, c, 17, 8);
        }
        t1.push_back(t);
    }
    Tile w2(32);
    if (lay.has_w) {
        // W: sum_j r_j (+-theta_omega_j), plus v (+-theta_gamma) in pass 2
        Tile t(k1);
        const u64* tom = sc.data() + SC_HEADER + 2 * nf;
        for (int j = 0; j < nf; j++) {
            const u128 th = (u128)tom[3 * j] | ((u128)tom[3 * j + 1] << 64);
            int8_t c[17];
            balanced_digits(th, tom[3 * j + 2] != 0, 17, c);
            conv_rows(t, 8 * j, c, 17, 8);
        }
        t1.push_back(t);
        const u128 tgm = (u128)sc[SC_TG_LO] | ((u128)sc[SC_TG_HI] << 64);
        int8_t c[17];
        balanced_digits(tgm, sc[SC_TG_SIGN] == 0, 17, c);  // so -= v theta_gamma unless the sign is set
        conv_rows(w2, 0, c, 17, 16);
    }
    // outputs: row 8 s + t of tile o is digit t of output 4 o + s
    std::vector<Tile> t3;
    const u64* outc = sc.data() + sc_out_base(nf);
    for (int o = 0; o < lay.nto; o++) {
        Tile t(k1), tv(32);
        for (int s = 0; s < 4 && 4 * o + s < count; s++) {
            const u64* oc = outc + (size_t)(4 * o + s) * sc_out_stride(nf);
            const u64 q = oc[0];
            int8_t c[8];
            for (int j = 0; j < nf; j++) {
                u64 x = oc[9 + j] % q;  // omega_ij
                for (int a = 0; a < 8; a++) {
                    balanced_digits(x, false, 8, c);
                    for (int e = 0; e < 8; e++) t.at(8 * s + e, 8 * j + a) = c[e];
                    x = mulmod(x, 256, q);
                }
            }
            u64 g = oc[3] % q, one = 1 % q;  // q - gamma, 1
            for (int a = 0; a < 16; a++) {
                balanced_digits(g, false, 8, c);
                for (int e = 0; e < 8; e++) tv.at(8 * s + e, a) = c[e];
                if (lay.has_w) {
                    balanced_digits(one, false, 8, c);
                    for (int e = 0; e < 8; e++) tv.at(8 * s + e, 16 + a) = c[e];
                }
                g = mulmod(g, 256, q);
                one = mulmod(one, 256, q);
            }
        }
        t1.push_back(t);
        t3.push_back(tv);
    }
    std::vector<u64> out;
    out.reserve(lay.words());
    for (const Tile& t : t1)
        for (int k = 0; k < lay.ks1; k++) t.emit(k, kw, out);
    if (lay.has_w) w2.emit(0, kw, out);
    for (const Tile& t : t3)
        for (int k = 0; k < lay.ks3(); k++) t.emit(k, kw, out);
    for (int i = 0; i < count; i++) {
        const u64 q = outc[(size_t)i * sc_out_stride(nf)];
        const u128 m = (((u128)1 << 78) / q + 1) * q;
        out.push_back((u64)m);
        out.push_back((u64)(m >> 64));
    }
    if (out.size() != lay.words()) throw std::logic_error("scaler digit table size");
    return out;
}

}  // namespace fhehip
