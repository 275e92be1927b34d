// Resident context tables (rq::Context + NttOperator per modulus).
#include <cstring>

#include "fhehip_internal.h"
#include "prng.h"

namespace fhehip {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        int code = (e == hipErrorOutOfMemory) ? -5 : -6;
        throw Error(code, std::string(what) + ": " + hipGetErrorString(e));
    }
}

DevBuf::DevBuf(size_t n) { reset(n); }

DevBuf::~DevBuf() {
    if (ptr) (void)hipFree(ptr);
}

void DevBuf::reset(size_t n) {
    if (ptr) {
        (void)hipFree(ptr);
        ptr = nullptr;
    }
    bytes = n;
    if (n) HIPCHK(hipMalloc(&ptr, n));
}

// NttOperator::new tables (math/src/ntt/native.rs:35-73) for every modulus,
// uploaded as interleaved (value, shoup) pairs.
// Constants of the inverse transform's last stage with the final scaling by
// N^-1 (or N^-1 2^64) folded in; z is that stage's twiddle.
static void set_last_stage(ModInfo& mi, u64 z) {
    const u64 p = mi.p;
    mi.last_zn = mulmod64(z, mi.n_inv, p);
    mi.last_zn_shoup = shoup64(mi.last_zn, p);
    mi.last_zr = mulmod64(z, mi.n_inv_r, p);
    mi.last_zr_shoup = shoup64(mi.last_zr, p);
}

// Half tables (see DeviceContext::half) of context c from its host tables.
// Forward: local index k = 2^s + g of an N/2-point transform at stage s maps
// to the global index 2^(s+1) + h 2^s + g. Inverse: local (N' - (N' >> v)) + g
// maps to (N - (N >> v)) + h 2^(logn - 2 - v) + g. A half of 32768 or more
// coefficients gets halves of its own, down to rows that fit a workgroup.
static void attach_half(DeviceContext& c, const std::vector<ulonglong2>& fwd, const std::vector<ulonglong2>& inv) {
    const u32 n = c.N, nh = n / 2;
    const u32 lh = (u32)c.logn - 1;
    auto hc = std::make_shared<DeviceContext>();
    hc->L = 2 * c.L;
    hc->N = nh;
    hc->logn = (int)lh;
    std::vector<ulonglong2> hf((size_t)hc->L * nh), hi((size_t)hc->L * nh);
    std::vector<ulonglong2> sp;
    for (u32 l = 0; l < c.L; l++) {
        const u64 p = c.moduli[l];
        for (u32 h = 0; h < 2; h++) {
            ulonglong2* F = hf.data() + ((size_t)2 * l + h) * nh;
            ulonglong2* I = hi.data() + ((size_t)2 * l + h) * nh;
            F[0] = fwd[(size_t)l * n + 1];
            for (u32 k = 1; k < nh; k++) {
                const u32 sft = 31 - __builtin_clz(k);
                F[k] = fwd[(size_t)l * n + k + ((1u + h) << sft)];
            }
            for (u32 v = 0; v < lh; v++) {
                const u32 lo = nh - (nh >> v), go = n - (n >> v);
                for (u32 g = 0; g < (1u << (lh - 1 - v)); g++)
                    I[lo + g] = inv[(size_t)l * n + go + (h << (lh - 1 - v)) + g];
            }
            I[nh - 1] = inv[(size_t)l * n + n - 1];
            ModInfo mi = c.h_mod[l];
            mi.n_inv = 1;
            mi.n_inv_shoup = shoup64(1, p);
            mi.n_inv_r = (u64)(((u128)1 << 64) % p);
            mi.n_inv_r_shoup = shoup64(mi.n_inv_r, p);
            set_last_stage(mi, I[nh - 2].x);
            hc->h_mod.push_back(mi);
            hc->moduli.push_back(p);
        }
        const u64 ninv = c.h_mod[l].n_inv;
        const u64 zl = mulmod64(inv[(size_t)l * n + n - 2].x, ninv, p);
        sp.push_back(fwd[(size_t)l * n + 1]);
        sp.push_back(make_ulonglong2(zl, shoup64(zl, p)));
        sp.push_back(make_ulonglong2(ninv, shoup64(ninv, p)));
    }
    if (hc->logn >= 15) attach_half(*hc, hf, hi);
    hc->fwd_tw.reset(hf.size() * sizeof(ulonglong2));
    hc->inv_tw.reset(hi.size() * sizeof(ulonglong2));
    hc->mod.reset(hc->h_mod.size() * sizeof(ModInfo));
    HIPCHK(hipMemcpy(hc->fwd_tw.ptr, hf.data(), hc->fwd_tw.bytes, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(hc->inv_tw.ptr, hi.data(), hc->inv_tw.bytes, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(hc->mod.ptr, hc->h_mod.data(), hc->mod.bytes, hipMemcpyHostToDevice));
    c.split.reset(sp.size() * sizeof(ulonglong2));
    HIPCHK(hipMemcpy(c.split.ptr, sp.data(), c.split.bytes, hipMemcpyHostToDevice));
    c.half = hc;
}

std::shared_ptr<DeviceContext> make_context(const std::vector<u64>& moduli, u32 n, const u64* psi) {
    if (moduli.empty()) throw Error(-1, "empty moduli list");
    if (n < 8 || (n & (n - 1))) throw Error(-1, "The degree is not a power of two larger or equal to 8");
    if (n > kMaxDegree) throw Error(-2, "degree above 131072 is not supported");
    auto c = std::make_shared<DeviceContext>();
    c->moduli = moduli;
    c->L = (u32)moduli.size();
    c->N = n;
    c->logn = ilog2_u64(n);
    for (u64 p : moduli) {
        if ((p >> 62) != 0 || p < 2) throw Error(-1, "InvalidModulus");
        if (!supports_ntt64(p, n)) throw Error(-2, "Impossible to construct a Ntt operator");
    }
    try {
        c->rns.init(moduli);
    } catch (const std::invalid_argument& e) {
        throw Error(-1, e.what());
    }
    std::vector<ulonglong2> fwd((size_t)c->L * n), inv((size_t)c->L * n);
    std::vector<u64> pw(n), pwi(n);
    for (u32 l = 0; l < c->L; l++) {
        const u64 p = moduli[l];
        // NttOperator::new's root (native.rs:320-336); the scan of 2, 3, ... only
        // if its 100 draws find none (fhe.rs fails there)
        u64 r = psi ? psi[l] : fhers_primitive_root(n, p);
        if (!r) r = primitive_root64(n, p);
        if (psi && (powmod64(r, n, p) != p - 1)) throw Error(-1, "psi is not a primitive 2N-th root of unity");
        c->psi.push_back(r);
        const u64 rinv = invmod64(r, p);
        pw[0] = 1;
        for (u32 i = 1; i < n; i++) pw[i] = mulmod64(pw[i - 1], r, p);
        pwi[0] = rinv;
        for (u32 i = 1; i < n; i++) pwi[i] = mulmod64(pwi[i - 1], rinv, p);
        for (u32 i = 0; i < n; i++) {
            u32 j = bit_reverse(i, c->logn);
            fwd[(size_t)l * n + i] = make_ulonglong2(pw[j], shoup64(pw[j], p));
            inv[(size_t)l * n + i] = make_ulonglong2(pwi[j], shoup64(pwi[j], p));
        }
        ModInfo mi;
        std::memset(&mi, 0, sizeof mi);
        mi.p = p;
        barrett128(p, mi.barrett_lo, mi.barrett_hi);
        mi.n_inv = invmod64(n % p, p);
        mi.n_inv_shoup = shoup64(mi.n_inv, p);
        u64 pinv = 1;  // p^-1 mod 2^64 by Newton iteration (p odd)
        for (int it = 0; it < 6; it++) pinv *= 2 - p * pinv;
        mi.pinv_neg = (u64)0 - pinv;
        mi.n_inv_r = (u64)(((u128)mi.n_inv << 64) % p);
        mi.n_inv_r_shoup = shoup64(mi.n_inv_r, p);
        mi.r64 = (u64)(((u128)1 << 64) % p);
        mi.r64_shoup = shoup64(mi.r64, p);
        mi.w1s = shoup64(1, p);
        set_last_stage(mi, inv[(size_t)l * n + n - 2].x);
        c->h_mod.push_back(mi);
    }
    if (n >= 4096) attach_half(*c, fwd, inv);
    c->fwd_tw.reset(fwd.size() * sizeof(ulonglong2));
    c->inv_tw.reset(inv.size() * sizeof(ulonglong2));
    c->mod.reset(c->h_mod.size() * sizeof(ModInfo));
    HIPCHK(hipMemcpy(c->fwd_tw.ptr, fwd.data(), c->fwd_tw.bytes, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->inv_tw.ptr, inv.data(), c->inv_tw.bytes, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->mod.ptr, c->h_mod.data(), c->mod.bytes, hipMemcpyHostToDevice));
    return c;
}

}  // namespace fhehip
