// extern "C" boundary of the library (declared in include/fhehip.h).
#include "../../include/fhehip.h"

#include <cstring>

#include "capi_internal.h"

using namespace fhehip;

static thread_local std::string g_last_error;

namespace fhehip {
int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
}  // namespace fhehip

extern "C" {

const char* fhehip_last_error(void) { return g_last_error.c_str(); }

const char* fhehip_version(void) { return "fhehip 0.1.0 (gfx950)"; }

int fhehip_context_create(const uint64_t* moduli, uint32_t num_moduli, uint32_t degree, const uint64_t* psi,
                          fhehip_context** out) {
    FHEHIP_API_BEGIN
    if (!moduli || !out || num_moduli == 0) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    auto c = make_context(std::vector<u64>(moduli, moduli + num_moduli), degree, psi);
    *out = new fhehip_context{c};
    FHEHIP_API_END
}

void fhehip_context_destroy(fhehip_context* ctx) { delete ctx; }

int fhehip_context_info(const fhehip_context* ctx, uint32_t* num_moduli, uint32_t* degree) {
    FHEHIP_API_BEGIN
    if (!ctx) return fail(FHEHIP_ERR_INVALID_ARG, "null context");
    if (num_moduli) *num_moduli = ctx->c->L;
    if (degree) *degree = ctx->c->N;
    FHEHIP_API_END
}

int fhehip_context_psi(const fhehip_context* ctx, uint64_t* psi_out) {
    FHEHIP_API_BEGIN
    if (!ctx || !psi_out) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    std::memcpy(psi_out, ctx->c->psi.data(), ctx->c->psi.size() * sizeof(uint64_t));
    FHEHIP_API_END
}

int fhehip_ntt_forward(const fhehip_context* ctx, uint64_t* d_polys, uint32_t num_polys, int lazy, void* stream) {
    FHEHIP_API_BEGIN
    if (!ctx || (!d_polys && num_polys)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    const DeviceContext& c = *ctx->c;
    RowMap m{d_polys, c.L, c.L, 0, 0};
    launch_ntt_forward(c, m, num_polys * c.L, lazy != 0, (hipStream_t)stream);
    FHEHIP_API_END
}

int fhehip_ntt_backward(const fhehip_context* ctx, uint64_t* d_polys, uint32_t num_polys, void* stream) {
    FHEHIP_API_BEGIN
    if (!ctx || (!d_polys && num_polys)) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    const DeviceContext& c = *ctx->c;
    RowMap m{d_polys, c.L, c.L, 0, 0};
    launch_ntt_backward(c, m, num_polys * c.L, (hipStream_t)stream);
    FHEHIP_API_END
}

int fhehip_ntt_set_shape(uint32_t degree, uint32_t fwd, uint32_t inv) {
    FHEHIP_API_BEGIN
    if (degree < 8 || (degree & (degree - 1))) return fail(FHEHIP_ERR_INVALID_ARG, "degree is not a power of two >= 8");
    set_ntt_shape(31 - __builtin_clz(degree), (int)fwd, (int)inv);
    FHEHIP_API_END
}

int fhehip_split_row_limit(uint32_t max_rows, uint32_t* previous) {
    FHEHIP_API_BEGIN
    if (previous) *previous = split_row_limit();
    set_split_row_limit(max_rows);
    FHEHIP_API_END
}

int fhehip_ntt_autotune(const fhehip_context* ctx, uint32_t rows, uint32_t* fwd, uint32_t* inv, float* ms_out) {
    FHEHIP_API_BEGIN
    if (!ctx || !fwd || !inv || !rows) return fail(FHEHIP_ERR_INVALID_ARG, "null argument");
    int rf = 4, ri = 4;
    float ms[4];
    autotune_ntt(*ctx->c, rows, rf, ri, ms);
    *fwd = (uint32_t)rf;
    *inv = (uint32_t)ri;
    if (ms_out)
        for (int k = 0; k < 4; k++) ms_out[k] = ms[k];
    FHEHIP_API_END
}

}  // extern "C"
