// RNS scaling on the matrix cores (scale_mfma.hip): the exact integer sums of
// RnsScaler::scale (math/src/rns/scaler.rs:249-352) formed by 8-bit integer
// MFMA over balanced base-256 digits.
//
// Every 64-bit input residue r is written as r = sum_a d_a 256^a with eight
// digits d_a in [-128, 127] (the bytes of (r + B) xor B, B = 0x80..80). Each
// sum the scaler needs is then a product of a constant digit matrix with
// the column's digit vector:
// - S = sum_j r_j theta_garner_j (exact, 256 bits) and W = sum_j +-r_j
//   theta_omega_j +- v theta_gamma (exact, signed): rows b of the matrix give
//   the coefficient P_b of 256^b (a digit convolution), S = sum_b P_b 256^b;
// - output i: rows t < 8 hold the digits t of [256^a omega_ij]_q (and of
//   [256^a (q - gamma)]_q for v's digits, [256^a]_q for the signed w's), so
//   that X = sum_t P_t 256^t is congruent to the output modulo q and below
//   2^78 in magnitude; X + M (M a multiple of q above 2^78) is reduced once.
// Each |P| stays below 160 * 128 * 128 < 2^22: the 32-bit accumulators are
// exact.
//
// The products run as v_mfma_i32_32x32x16_i8 (kw = 16: K steps of 16 k,
// lane l supplying the 8 signed bytes k = 16 kappa + 8 (l / 32) + e of row
// or column l % 32) or v_mfma_i32_32x32x32_i8 (kw = 32: K steps of 32 k, 16
// bytes per lane, k = 32 kappa + 16 (l / 32) + e). Table layout in u64
// words, wl = kw / 16 words per lane and K step, lane-major:
//   A1 [tiles1][ks1][64][wl]  pass 1 over the residue digits (k = 8 j + a);
//                             tiles: S, W (when has_w), output tiles
//   A2 [64][wl]               (has_w) W's v step (k < 16: the digits of v)
//   A3 [nto][ks3][64][wl]     output tiles' v and w digits: kw = 32 one
//                             step (k < 16 v, 16 <= k < 32 the signed w);
//                             kw = 16 a v step, then (has_w) a w step
//   M  [count][2]             the offsets M_i (lo, hi)
#pragma once

#include <vector>

#include "host_math.h"

namespace fhehip {

struct ScaleMfmaLayout {
    int nf = 0, count = 0;
    int kw = 32;   // k per MFMA (16 or 32)
    int ks1 = 0;   // K steps of pass 1 (kw / 8 residues each)
    int nto = 0;   // output tiles (four outputs each)
    bool has_w = false;
    int wl() const { return kw / 16; }
    int ks3() const { return kw == 32 ? 1 : (has_w ? 2 : 1); }
    int tiles1() const { return 1 + (has_w ? 1 : 0) + nto; }
    size_t a2() const { return (size_t)tiles1() * ks1 * 64 * wl(); }
    size_t a3() const { return a2() + (has_w ? 64 * wl() : 0); }
    size_t m() const { return a3() + (size_t)nto * ks3() * 64 * wl(); }
    size_t words() const { return m() + 2 * (size_t)count; }
};

// The digit tables of a scaler constant block (scaler_blob, rns_device.h)
// whose inputs are reduced residues of source moduli up to pmax, for MFMA
// steps of kw k values (16 or 32; any other value: 32 above 8 source
// moduli, else 16); empty when the shape has no matrix-core form (more
// than 16 source moduli, more than 8 outputs, pmax of 2^62 or more, or a
// theta_garner shift outside 65 .. 128).
std::vector<u64> scaler_mfma_blob(const std::vector<u64>& sc, u64 pmax, ScaleMfmaLayout& lay, int kw = 0);

}  // namespace fhehip
