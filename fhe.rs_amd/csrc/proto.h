// Protobuf wire format (proto3) of fhe.rs's messages, on the host: the
// Rq message (math/src/proto/rq.proto) and the bfv messages (fhe/src/proto/
// bfv.proto). Writers emit fields the way prost encodes these structs:
// declared fields in order, a oneof after them, proto3 defaults (0, empty,
// false) left out, repeated bytes always written. Readers accept any field
// order, skip unknown fields and keep the last value of a singular field.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace fhehip {
namespace pb {

struct Span {
    const uint8_t* data = nullptr;
    size_t len = 0;
};

class Writer {
  public:
    std::vector<uint8_t> out;
    void varint(uint64_t v);
    void key(uint32_t field, uint32_t wire) { varint(((uint64_t)field << 3) | wire); }
    void uint_field(uint32_t field, uint64_t v) {  // uint32 / uint64 / enum / bool
        if (v) key(field, 0), varint(v);
    }
    void bytes_field(uint32_t field, const uint8_t* p, size_t n);            // repeated: always
    void bytes_if(uint32_t field, const uint8_t* p, size_t n) {             // singular: non-empty
        if (n) bytes_field(field, p, n);
    }
    void packed_u64(uint32_t field, const std::vector<uint64_t>& v);
};

struct Field {
    uint32_t number = 0, wire = 0;
    uint64_t value = 0;  // wire 0 (and 1, 5)
    Span bytes;          // wire 2
};

// Iterates the fields of one message; throws Error(FHEHIP_ERR_INVALID_ARG)
// on a truncated or malformed encoding.
class Reader {
  public:
    Reader(const uint8_t* p, size_t n) : p_(p), end_(p + n) {}
    explicit Reader(Span s) : Reader(s.data, s.len) {}
    bool next(Field& f);

  private:
    uint64_t varint();
    const uint8_t* p_;
    const uint8_t* end_;
};

// rq.proto: Representation enum values.
enum Representation : uint32_t { REPR_UNKNOWN = 0, REPR_POWERBASIS = 1, REPR_NTT = 2, REPR_NTTSHOUP = 3 };

struct Rq {
    uint32_t representation = 0, degree = 0;
    Span coefficients;
    bool allow_variable_time = false;
};
Rq parse_rq(Span s);
std::vector<uint8_t> write_rq(uint32_t representation, uint32_t degree, const uint8_t* coeffs, size_t len,
                              bool allow_variable_time);

struct Ciphertext {
    std::vector<Span> c;
    Span seed;
    uint32_t level = 0;
};
Ciphertext parse_ciphertext(Span s);

struct KeySwitchingKey {
    std::vector<Span> c0, c1;
    Span seed;
    uint32_t ciphertext_level = 0, ksk_level = 0, log_base = 0;
};
KeySwitchingKey parse_ksk(Span s);
std::vector<uint8_t> write_ksk(const std::vector<std::vector<uint8_t>>& c0, const std::vector<std::vector<uint8_t>>& c1,
                               const uint8_t* seed, uint32_t ciphertext_level, uint32_t ksk_level, uint32_t log_base);

// RelinearizationKey { ksk = 1 }; GaloisKey { ksk = 1, exponent = 2 }.
struct WrappedKsk {
    bool has_ksk = false;
    Span ksk;
    uint32_t exponent = 0;
};
WrappedKsk parse_wrapped_ksk(Span s);

struct EvaluationKey {
    std::vector<Span> gk;
    uint32_t ciphertext_level = 0, evaluation_key_level = 0;
};
EvaluationKey parse_evaluation_key(Span s);

struct Parameters {
    uint32_t degree = 0, variance = 0;
    std::vector<uint64_t> moduli;
    int plaintext_kind = 0;  // 0 absent, 1 plaintext (u64), 2 plaintext_big (little-endian bytes)
    uint64_t plaintext = 0;
    Span plaintext_big;
};
Parameters parse_parameters(Span s);
std::vector<uint8_t> write_parameters(uint32_t degree, const std::vector<uint64_t>& moduli, uint64_t plaintext,
                                      uint32_t variance);

}  // namespace pb
}  // namespace fhehip
