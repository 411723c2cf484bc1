// Proof bytes: the transcript's Proof { narg_string, hints }
// (crates/transcript/src/proof.rs:13-16) and the host container that carries
// it (crates/host/src/wire_proof.rs): a 32-byte header, then the narg stream,
// then the hint stream.
//
//   offset size field
//   0      8    magic "bitzPCS\0"
//   8      2    wire_version = 1      (little-endian)
//   10     2    header_len = 32
//   12     4    flags = 0
//   16     8    narg_byte_len
//   24     8    hint_byte_len
//   32     ..   narg bytes, then hint bytes
//
// Decoding checks the fields in that order, rejects anything but the exact
// encoding, and never sizes an allocation from the input.
#include <string.h>

#include "../../include/bitz_hip.h"
#include "transcript.hpp"

namespace bitz {
int32_t report_error(int32_t code, const std::string& msg);
}


namespace {
const uint8_t kMagic[8] = {'b', 'i', 't', 'z', 'P', 'C', 'S', 0};
constexpr uint16_t kWireVersion = 1;
constexpr size_t kHeaderLen = 32;

uint64_t le(const uint8_t* p, int n) {
    uint64_t v = 0;
    for (int i = n - 1; i >= 0; --i) v = v << 8 | p[i];
    return v;
}
void put(uint8_t* p, uint64_t v, int n) {
    for (int i = 0; i < n; ++i) p[i] = (uint8_t)(v >> (8 * i));
}
}  // namespace

extern "C" int32_t bitz_proof_encode(const uint8_t* narg, size_t narg_len, const uint8_t* hints, size_t hints_len,
                                     uint8_t* out, size_t cap, size_t* len) {
    if (!len || (narg_len && !narg) || (hints_len && !hints)) return bitz::report_error(BITZ_ERR_ARG, "null pointer");
    if (narg_len > SIZE_MAX - kHeaderLen || hints_len > SIZE_MAX - kHeaderLen - narg_len)
        return bitz::report_error(BITZ_ERR_PROOF_OVERFLOW, "the lengths do not fit the address space");
    *len = kHeaderLen + narg_len + hints_len;
    if (!out) return BITZ_OK;  // a length query
    if (cap < *len) return bitz::report_error(BITZ_ERR_ARG, "output buffer too small");
    memcpy(out, kMagic, 8);
    put(out + 8, kWireVersion, 2);
    put(out + 10, kHeaderLen, 2);
    put(out + 12, 0, 4);
    put(out + 16, narg_len, 8);
    put(out + 24, hints_len, 8);
    if (narg_len) memcpy(out + kHeaderLen, narg, narg_len);
    if (hints_len) memcpy(out + kHeaderLen + narg_len, hints, hints_len);
    return BITZ_OK;
}

extern "C" int32_t bitz_proof_decode(const uint8_t* bytes, size_t len, size_t* narg_offset, size_t* narg_len,
                                     size_t* hints_offset, size_t* hints_len, uint64_t* detail) {
    if ((len && !bytes) || !narg_offset || !narg_len || !hints_offset || !hints_len || !detail)
        return bitz::report_error(BITZ_ERR_ARG, "null pointer");
    *detail = 0;
    // Field by field, as wire_proof.rs:108-155 reads them: a field that runs
    // past the input is OutOfInput.
    if (len < 8) return bitz::report_error(BITZ_ERR_PROOF_SHORT, "proof is too short");
    if (memcmp(bytes, kMagic, 8) != 0) return bitz::report_error(BITZ_ERR_PROOF_MAGIC, "bad magic");
    if (len < 10) return bitz::report_error(BITZ_ERR_PROOF_SHORT, "proof is too short");
    const uint64_t version = le(bytes + 8, 2);
    if (version != kWireVersion) {
        *detail = version;
        return bitz::report_error(BITZ_ERR_PROOF_VERSION, "unsupported wire version");
    }
    if (len < 12) return bitz::report_error(BITZ_ERR_PROOF_SHORT, "proof is too short");
    const uint64_t header_len = le(bytes + 10, 2);
    if (header_len != kHeaderLen) {
        *detail = header_len;
        return bitz::report_error(BITZ_ERR_PROOF_HEADER_LEN, "the declared header length is not 32");
    }
    if (len < 16) return bitz::report_error(BITZ_ERR_PROOF_SHORT, "proof is too short");
    const uint64_t flags = le(bytes + 12, 4);
    if (flags != 0) {
        *detail = flags;
        return bitz::report_error(BITZ_ERR_PROOF_FLAGS, "a reserved flag bit is set");
    }
    if (len < kHeaderLen) return bitz::report_error(BITZ_ERR_PROOF_SHORT, "proof is too short");
    const uint64_t n = le(bytes + 16, 8), h = le(bytes + 24, 8);
    if (n > SIZE_MAX || h > SIZE_MAX || n > SIZE_MAX - kHeaderLen || h > SIZE_MAX - kHeaderLen - n)
        return bitz::report_error(BITZ_ERR_PROOF_OVERFLOW, "the declared lengths do not fit the address space");
    const size_t declared = kHeaderLen + (size_t)n + (size_t)h;
    if (declared != len) {
        *detail = declared;
        return bitz::report_error(BITZ_ERR_PROOF_LENGTH, "length mismatch");
    }
    *narg_offset = kHeaderLen;
    *narg_len = (size_t)n;
    *hints_offset = kHeaderLen + (size_t)n;
    *hints_len = (size_t)h;
    return BITZ_OK;
}

extern "C" int32_t bitz_transcript_hint(bitz_transcript* t, const uint8_t* bytes, size_t len) {
    if (!t || (len && !bytes)) return bitz::report_error(BITZ_ERR_ARG, "null pointer");
    if (t->t.type == bitz::Transcript::Type::kCallbacks)
        return bitz::report_error(BITZ_ERR_ARG, "a callback transcript keeps its own hint stream");
    t->t.hints.insert(t->t.hints.end(), bytes, bytes + len);
    return BITZ_OK;
}

extern "C" size_t bitz_transcript_hints(const bitz_transcript* t, uint8_t* out, size_t cap) {
    if (!t) return 0;
    const size_t n = t->t.hints.size();
    if (out && n) memcpy(out, t->t.hints.data(), n < cap ? n : cap);
    return n;
}

extern "C" int32_t bitz_transcript_proof(const bitz_transcript* t, uint8_t* out, size_t cap, size_t* len) {
    if (!t) return bitz::report_error(BITZ_ERR_ARG, "null pointer");
    return bitz_proof_encode(t->t.narg.data(), t->t.narg.size(), t->t.hints.data(), t->t.hints.size(), out, cap,
                             len);
}
