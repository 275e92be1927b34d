// Protobuf wire format of fhe.rs's messages (see proto.h).
#include "proto.h"

#include "fhehip_internal.h"

namespace fhehip {
namespace pb {

void Writer::varint(uint64_t v) {
    while (v >= 0x80) {
        out.push_back((uint8_t)(v | 0x80));
        v >>= 7;
    }
    out.push_back((uint8_t)v);
}

void Writer::bytes_field(uint32_t field, const uint8_t* p, size_t n) {
    key(field, 2);
    varint(n);
    out.insert(out.end(), p, p + n);
}

void Writer::packed_u64(uint32_t field, const std::vector<uint64_t>& v) {
    if (v.empty()) return;
    Writer body;
    for (uint64_t x : v) body.varint(x);
    bytes_field(field, body.out.data(), body.out.size());
}

static void malformed(const char* what) { throw Error(-1, std::string("malformed protobuf: ") + what); }

uint64_t Reader::varint() {
    uint64_t v = 0;
    for (int shift = 0; shift < 70; shift += 7) {
        if (p_ == end_) malformed("truncated varint");
        const uint8_t b = *p_++;
        if (shift == 63 && b > 1) malformed("varint above 64 bits");
        v |= (uint64_t)(b & 0x7f) << shift;
        if (!(b & 0x80)) return v;
    }
    malformed("varint longer than 10 bytes");
    return 0;
}

bool Reader::next(Field& f) {
    if (p_ == end_) return false;
    const uint64_t k = varint();
    f.number = (uint32_t)(k >> 3);
    f.wire = (uint32_t)(k & 7);
    if (f.number == 0 || (k >> 3) > 0x1fffffff) malformed("field number");
    f.value = 0;
    f.bytes = Span{};
    switch (f.wire) {
        case 0: f.value = varint(); break;
        case 1:
        case 5: {
            const size_t w = f.wire == 1 ? 8 : 4;
            if ((size_t)(end_ - p_) < w) malformed("truncated fixed-width field");
            for (size_t i = 0; i < w; i++) f.value |= (uint64_t)p_[i] << (8 * i);
            p_ += w;
            break;
        }
        case 2: {
            const uint64_t n = varint();
            if (n > (uint64_t)(end_ - p_)) malformed("length past the end");
            f.bytes = Span{p_, (size_t)n};
            p_ += n;
            break;
        }
        default: malformed("unsupported wire type");
    }
    return true;
}

// A field of the expected wire type, or malformed.
static void expect(const Field& f, uint32_t wire) {
    if (f.wire != wire) malformed("field with an unexpected wire type");
}

Rq parse_rq(Span s) {
    Rq r;
    Reader rd(s);
    Field f;
    while (rd.next(f)) {
        switch (f.number) {
            case 1: expect(f, 0), r.representation = (uint32_t)f.value; break;
            case 2: expect(f, 0), r.degree = (uint32_t)f.value; break;
            case 3: expect(f, 2), r.coefficients = f.bytes; break;
            case 4: expect(f, 0), r.allow_variable_time = f.value != 0; break;
            default: break;
        }
    }
    return r;
}

std::vector<uint8_t> write_rq(uint32_t representation, uint32_t degree, const uint8_t* coeffs, size_t len,
                              bool allow_variable_time) {
    Writer w;
    w.uint_field(1, representation);
    w.uint_field(2, degree);
    w.bytes_if(3, coeffs, len);
    w.uint_field(4, allow_variable_time ? 1 : 0);
    return std::move(w.out);
}

Ciphertext parse_ciphertext(Span s) {
    Ciphertext c;
    Reader rd(s);
    Field f;
    while (rd.next(f)) {
        switch (f.number) {
            case 1: expect(f, 2), c.c.push_back(f.bytes); break;
            case 2: expect(f, 2), c.seed = f.bytes; break;
            case 3: expect(f, 0), c.level = (uint32_t)f.value; break;
            default: break;
        }
    }
    return c;
}

KeySwitchingKey parse_ksk(Span s) {
    KeySwitchingKey k;
    Reader rd(s);
    Field f;
    while (rd.next(f)) {
        switch (f.number) {
            case 1: expect(f, 2), k.c0.push_back(f.bytes); break;
            case 2: expect(f, 2), k.c1.push_back(f.bytes); break;
            case 3: expect(f, 2), k.seed = f.bytes; break;
            case 4: expect(f, 0), k.ciphertext_level = (uint32_t)f.value; break;
            case 5: expect(f, 0), k.ksk_level = (uint32_t)f.value; break;
            case 6: expect(f, 0), k.log_base = (uint32_t)f.value; break;
            default: break;
        }
    }
    return k;
}

std::vector<uint8_t> write_ksk(const std::vector<std::vector<uint8_t>>& c0, const std::vector<std::vector<uint8_t>>& c1,
                               const uint8_t* seed, uint32_t ciphertext_level, uint32_t ksk_level, uint32_t log_base) {
    Writer w;
    for (const auto& r : c0) w.bytes_field(1, r.data(), r.size());
    for (const auto& r : c1) w.bytes_field(2, r.data(), r.size());
    if (seed) w.bytes_field(3, seed, 32);
    w.uint_field(4, ciphertext_level);
    w.uint_field(5, ksk_level);
    w.uint_field(6, log_base);
    return std::move(w.out);
}

WrappedKsk parse_wrapped_ksk(Span s) {
    WrappedKsk k;
    Reader rd(s);
    Field f;
    while (rd.next(f)) {
        if (f.number == 1) {
            expect(f, 2);
            k.has_ksk = true;
            k.ksk = f.bytes;
        } else if (f.number == 2) {
            expect(f, 0);
            k.exponent = (uint32_t)f.value;
        }
    }
    return k;
}

EvaluationKey parse_evaluation_key(Span s) {
    EvaluationKey e;
    Reader rd(s);
    Field f;
    while (rd.next(f)) {
        switch (f.number) {
            case 2: expect(f, 2), e.gk.push_back(f.bytes); break;
            case 3: expect(f, 0), e.ciphertext_level = (uint32_t)f.value; break;
            case 4: expect(f, 0), e.evaluation_key_level = (uint32_t)f.value; break;
            default: break;
        }
    }
    return e;
}

Parameters parse_parameters(Span s) {
    Parameters p;
    Reader rd(s);
    Field f;
    while (rd.next(f)) {
        switch (f.number) {
            case 1: expect(f, 0), p.degree = (uint32_t)f.value; break;
            case 2:
                if (f.wire == 0) {
                    p.moduli.push_back(f.value);
                } else {
                    expect(f, 2);
                    // a packed run is a sequence of varints
                    const uint8_t* q = f.bytes.data;
                    const uint8_t* e = q + f.bytes.len;
                    while (q != e) {
                        uint64_t v = 0;
                        int shift = 0;
                        for (;;) {
                            if (q == e || shift > 63) malformed("packed varint");
                            const uint8_t b = *q++;
                            v |= (uint64_t)(b & 0x7f) << shift;
                            shift += 7;
                            if (!(b & 0x80)) break;
                        }
                        p.moduli.push_back(v);
                    }
                }
                break;
            case 3: expect(f, 0), p.plaintext_kind = 1, p.plaintext = f.value; break;
            case 4: expect(f, 0), p.variance = (uint32_t)f.value; break;
            case 5: expect(f, 2), p.plaintext_kind = 2, p.plaintext_big = f.bytes; break;
            default: break;
        }
    }
    return p;
}

std::vector<uint8_t> write_parameters(uint32_t degree, const std::vector<uint64_t>& moduli, uint64_t plaintext,
                                      uint32_t variance) {
    // prost writes the oneof (plaintext = 3) after the declared fields
    Writer w;
    w.uint_field(1, degree);
    w.packed_u64(2, moduli);
    w.uint_field(4, variance);
    w.key(3, 0);
    w.varint(plaintext);
    return std::move(w.out);
}

}  // namespace pb
}  // namespace fhehip
