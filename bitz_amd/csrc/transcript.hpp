// Host transcript object behind bitz_transcript (see transcript.cpp).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "gf128.hpp"

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

struct Transcript {
    enum class Type { kCallbacks, kSeeded, kDuplex } type;
    void* cb_ctx = nullptr;
    void (*cb_absorb)(void*, int32_t, const uint8_t*, size_t) = nullptr;
    void (*cb_squeeze)(void*, uint8_t*) = nullptr;
    uint64_t seed = 0;
    uint8_t state[32] = {0};
    std::vector<uint8_t> narg;

    void absorb(int kind, const uint8_t* bytes, size_t len);
    void absorb_elems(int kind, const f128* e, size_t n) {
        absorb(kind, reinterpret_cast<const uint8_t*>(e), n * sizeof(f128));
    }
    f128 squeeze();
};

}  // namespace bitz
