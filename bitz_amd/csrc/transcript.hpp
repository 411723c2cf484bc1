// Host transcript object behind bitz_transcript (see transcript.cpp).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "gf128.hpp"
#include "keccak.hpp"

namespace bitz {

void sha256(const uint8_t* msg, size_t len, uint8_t out[32]);

// Streaming SHA-256 (FIPS 180-4).
struct Sha256 {
    uint32_t h[8];
    uint8_t buf[64];
    size_t fill = 0;
    uint64_t total = 0;
    Sha256();
    void update(const void* data, size_t len);
    void update_u64(uint64_t v);  // 8 little-endian bytes
    void finish(uint8_t out[32]);
};
uint64_t splitmix_next(uint64_t& state);

// BLAKE3 of at most 64 bytes (false beyond); the reference's proof-of-work
// check and search (crates/transcript/src/pow.rs:21-52); grinding_bits
// (crates/common/src/security.rs:25-32).
bool blake3_small(const uint8_t* in, size_t len, uint8_t out[32]);
bool pow_valid(const uint8_t seed[16], uint64_t nonce, uint32_t bits);
uint64_t pow_find(const uint8_t seed[16], uint32_t bits);
uint32_t grinding_bits(uint32_t security_bits, uint64_t coefficient);
constexpr uint32_t kMaxGrindingBits = 32;

struct Transcript {
    enum class Type { kCallbacks, kSeeded, kDuplex, kKeccak } type;
    void* cb_ctx = nullptr;
    void (*cb_absorb)(void*, int32_t, const uint8_t*, size_t) = nullptr;
    void (*cb_squeeze)(void*, uint8_t*) = nullptr;
    // ProverState::grind handed over whole (the caller's own transcript), or
    // null: grind() frames it over absorb and squeeze.
    void (*cb_grind)(void*, const uint8_t*, size_t, uint32_t) = nullptr;
    uint64_t seed = 0;
    uint8_t state[32] = {0};
    KeccakState keccak{};  // kKeccak: the duplex state (keccak.hpp)
    std::vector<uint8_t> narg;
    // The hint stream (ProverState::hint, crates/transcript/src/prover.rs:64-68):
    // prover data beside the narg string that the sponge never sees.
    std::vector<uint8_t> hints;

    void absorb(int kind, const uint8_t* bytes, size_t len);
    void absorb_elems(int kind, const f128* e, size_t n) {
        absorb(kind, reinterpret_cast<const uint8_t*>(e), n * sizeof(f128));
    }
    f128 squeeze();
    // ProverState::grind (crates/transcript/src/prover.rs:29-41): nothing at
    // zero bits; else the pow.rs:11-16 header as public messages (the tag,
    // le64(label length), the label, le32(bits)), a squeezed seed, and the
    // first valid nonce as an 8-byte prover message.  bits <= 32.
    void grind(const uint8_t* label, size_t len, uint32_t bits);
};

}  // namespace bitz

// The C ABI's opaque transcript handle.
struct bitz_transcript {
    bitz::Transcript t;
};
