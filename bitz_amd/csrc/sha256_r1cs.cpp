// SHA-256 R1CS instances: the matrices M, A, B, C of BitZ's SHA-256 circuits
// and their witnesses (f, h), produced on the host for the GF(2^128) prover.
//
// One gadget description (compress and its sums) runs over two backends:
//  - Symbolic records the constraint system the way ConstraintGenerator does
//    (crates/circuit/src/constraints.rs:578-790): Boolean witnesses are the
//    inputs followed by every hint bit; each bitz call records one integer
//    witness (a row of M over [1 | f]); assert_r1c appends one row to A, B
//    and C with integer coefficients, the constant in column 0 and witness k
//    in column k + 1.
//  - Concrete evaluates the same calls the way Witgen does
//    (crates/circuit/src/witgen.rs:425-500): f is the inputs then the hint
//    bits, h is 1 then every recorded bit, both packed least significant bit
//    first in 64-bit words (PackedWitness, witgen.rs:207-305).
// Running one description through both keeps the columns of A, B, C and the
// bits of h in the same order by construction.
//
// Gadgets follow crates/circuit/src/sha256.rs: words are little-endian bits
// with a lifted integer expression (UInt, :135-180); sums of words are hinted
// as WIDTH bits and constrained by one linear row (sum_32, :220-265); the
// round additions use the doubled Ch / Maj form (sum_doubled_32, :267-315,
// choice_twice :318-327, majority_twice :330-337); compress is :344-470.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <initializer_list>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/bitz_hip.h"

namespace bitz {
int32_t report_error(int32_t code, const std::string& msg);
}

namespace {

// FIPS 180-4 section 4.2.2 and 5.3.3.
constexpr uint32_t kK[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2,
};
constexpr uint32_t kIV[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                             0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};

constexpr int kMaxWidth = 35;  // the widest hinted sum (sum_doubled_32)

struct Failure {
    int32_t code;
    std::string msg;
};

// ------------------------------------------------------------ symbolic --
// An F2 combination of Boolean witnesses: c + w[0] + ... + w[n-1], indices
// ascending.  SHA-256 needs at most three (xor3 of single witnesses).
struct SymBit {
    uint32_t w[3];
    uint8_t n = 0;
    bool c = false;
};

SymBit sym_xor(const SymBit& a, const SymBit& b) {
    SymBit r;
    r.c = a.c != b.c;
    int i = 0, j = 0;
    auto put = [&](uint32_t v) {
        if (r.n == 3) throw Failure{BITZ_ERR_STATE, "F2 combination wider than three witnesses"};
        r.w[r.n++] = v;
    };
    while (i < a.n && j < b.n) {
        if (a.w[i] < b.w[j]) put(a.w[i++]);
        else if (b.w[j] < a.w[i]) put(b.w[j++]);
        else ++i, ++j;  // shared witness cancels
    }
    while (i < a.n) put(a.w[i++]);
    while (j < b.n) put(b.w[j++]);
    return r;
}

struct SymWord {
    int n = 0;
    SymBit b[kMaxWidth];
};

// Integer linear combination over the recorded witnesses: constant c plus
// (index, coefficient) terms, merged when a row is emitted.
struct Lin {
    int64_t c = 0;
    std::vector<std::pair<uint32_t, int64_t>> t;
    Lin() = default;
    explicit Lin(int64_t v) : c(v) {}
    Lin& operator+=(const Lin& o) {
        c += o.c;
        t.insert(t.end(), o.t.begin(), o.t.end());
        return *this;
    }
    Lin& operator-=(const Lin& o) {
        c -= o.c;
        for (const auto& e : o.t) t.emplace_back(e.first, -e.second);
        return *this;
    }
    Lin& operator*=(int64_t k) {
        c *= k;
        for (auto& e : t) e.second *= k;
        return *this;
    }
    friend Lin operator+(Lin a, const Lin& b) { return a += b; }
    friend Lin operator-(Lin a, const Lin& b) { return a -= b; }
    friend Lin operator*(Lin a, int64_t k) { return a *= k; }
};

struct Csr {
    std::vector<uint64_t> starts{0};
    std::vector<uint32_t> cols;
    std::vector<int64_t> coef;
    // LinearCombination::into_sparse_row (constraints.rs:412-431): witness k
    // in column k + 1, ascending, equal columns summed, zero sums dropped,
    // a nonzero constant first in column 0.
    void push(Lin& v) {
        std::sort(v.t.begin(), v.t.end(),
                  [](const std::pair<uint32_t, int64_t>& x, const std::pair<uint32_t, int64_t>& y) {
                      return x.first < y.first;
                  });
        if (v.c != 0) {
            cols.push_back(0);
            coef.push_back(v.c);
        }
        for (size_t i = 0; i < v.t.size();) {
            size_t j = i;
            int64_t s = 0;
            while (j < v.t.size() && v.t[j].first == v.t[i].first) s += v.t[j++].second;
            if (s != 0) {
                cols.push_back(v.t[i].first + 1);
                coef.push_back(s);
            }
            i = j;
        }
        starts.push_back(cols.size());
    }
};

struct Symbolic {
    using Word = SymWord;
    using Z = Lin;
    uint32_t next_witness;      // Boolean witnesses allocated so far
    std::vector<uint64_t> m_starts{0, 1};
    std::vector<uint32_t> m_pos{0};  // M row 0 is the constant one
    Csr a, b, c;

    explicit Symbolic(uint32_t inputs) : next_witness(inputs) {}

    static Word input(uint32_t first, int n) {
        Word w;
        w.n = n;
        for (int i = 0; i < n; ++i) {
            w.b[i].w[0] = first + (uint32_t)i;
            w.b[i].n = 1;
        }
        return w;
    }
    static Word constant(uint64_t v, int n) {
        Word w;
        w.n = n;
        for (int i = 0; i < n; ++i) w.b[i].c = i < 64 && ((v >> i) & 1);
        return w;
    }
    static Word xor_(const Word& x, const Word& y) {
        Word w;
        w.n = x.n;
        for (int i = 0; i < x.n; ++i) w.b[i] = sym_xor(x.b[i], y.b[i]);
        return w;
    }
    static Word rotr(const Word& x, int k) {
        Word w;
        w.n = x.n;
        for (int i = 0; i < x.n; ++i) w.b[i] = x.b[(i + k) % x.n];
        return w;
    }
    static Word shr(const Word& x, int k) {
        Word w;
        w.n = x.n;
        for (int i = 0; i < x.n; ++i)
            if (i + k < x.n) w.b[i] = x.b[i + k];
        return w;
    }
    static Word low(const Word& x, int n) {
        Word w = x;
        w.n = n;
        return w;
    }
    static uint64_t value(const Word&) { return 0; }  // hints are not evaluated here
    static SymBit bit(const Word& x, int i) { return x.b[i]; }
    Word hint(int n, uint64_t) {
        const uint32_t first = next_witness;
        next_witness += (uint32_t)n;
        return input(first, n);
    }
    // bool_sparse_row (constraints.rs:700-715): the constant in column 0,
    // Boolean witness k in column k + 1.
    uint32_t record(const SymBit& v) {
        if (v.c) m_pos.push_back(0);
        for (int i = 0; i < v.n; ++i) m_pos.push_back(v.w[i] + 1);
        m_starts.push_back(m_pos.size());
        return (uint32_t)(m_starts.size() - 3);  // integer witness index (row 0 excluded)
    }
    Z bitz(const SymBit& v) {
        Z z;
        z.t.emplace_back(record(v), 1);
        return z;
    }
    // bitz_unsigned (constraints.rs:748-773): one record per bit, 2^i.
    void bitz_unsigned(const Word& x, Z& full, Z& low32) {
        full = Z();
        full.t.reserve(x.n);
        for (int i = 0; i < x.n; ++i) full.t.emplace_back(record(x.b[i]), (int64_t)1 << i);
        low32 = Z();
        low32.t.assign(full.t.begin(), full.t.begin() + std::min(x.n, 32));
    }
    static Z zconst(int64_t v) { return Z(v); }
    void assert_r1c(Z za, Z zb, Z zc) {
        a.push(za);
        b.push(zb);
        c.push(zc);
    }
};

// ------------------------------------------------------------ concrete --
struct Packed {
    std::vector<uint64_t> words;
    size_t len = 0;
    void append(uint64_t v, int n) {  // PackedWitness::append_word (witgen.rs:268-290)
        if (n <= 0) return;
        if (n < 64) v &= ((uint64_t)1 << n) - 1;
        const int off = (int)(len % 64);
        if (off == 0) words.push_back(v);
        else {
            words.back() |= v << off;
            if (off + n > 64) words.push_back(v >> (64 - off));
        }
        len += (size_t)n;
    }
};

struct ValWord {
    uint64_t v = 0;
    int n = 0;
};

struct Concrete {
    using Word = ValWord;
    using Z = int64_t;
    Packed f, h;
    bool satisfied = true;
    Concrete() { h.append(1, 1); }
    static uint64_t mask(int n) { return n >= 64 ? ~(uint64_t)0 : (((uint64_t)1 << n) - 1); }
    static Word constant(uint64_t v, int n) { return {v & mask(n), n}; }
    static Word xor_(const Word& x, const Word& y) { return {x.v ^ y.v, x.n}; }
    static Word rotr(const Word& x, int k) {
        k %= x.n;
        return {k ? (((x.v >> k) | (x.v << (x.n - k))) & mask(x.n)) : x.v, x.n};
    }
    static Word shr(const Word& x, int k) { return {k >= x.n ? 0 : x.v >> k, x.n}; }
    static Word low(const Word& x, int n) { return {x.v & mask(n), n}; }
    static uint64_t value(const Word& x) { return x.v; }
    static bool bit(const Word& x, int i) { return (x.v >> i) & 1; }
    Word hint(int n, uint64_t v) {
        f.append(v, n);
        return {v & mask(n), n};
    }
    Z bitz(bool v) {
        h.append(v, 1);
        return v;
    }
    void bitz_unsigned(const Word& x, Z& full, Z& low32) {
        h.append(x.v, x.n);
        full = (int64_t)x.v;
        low32 = (int64_t)(x.v & 0xffffffffu);
    }
    static Z zconst(int64_t v) { return v; }
    void assert_r1c(Z za, Z zb, Z zc) {
        if (za * zb != zc) satisfied = false;
    }
};

// ------------------------------------------------------------- gadgets --
template <class CS>
struct UInt {
    typename CS::Word word;
    typename CS::Z full, low;
};

template <class CS>
UInt<CS> uint_from_word(CS& cs, const typename CS::Word& w) {
    UInt<CS> u{w, {}, {}};
    cs.bitz_unsigned(w, u.full, u.low);
    return u;
}

template <class CS>
UInt<CS> uint_constant(uint32_t v) {
    return {CS::constant(v, 32), CS::zconst(v), CS::zconst(v)};
}

template <class CS>
void assert_zero(CS& cs, typename CS::Z v) {
    cs.assert_r1c(CS::zconst(0), CS::zconst(0), std::move(v));
}

// sum_32 (sha256.rs:220-265): hint the sum as `width` bits, record them,
// constrain sum(terms) - wide = 0, keep the low 32 bits.
template <class CS>
UInt<CS> sum_32(CS& cs, int width, std::initializer_list<const UInt<CS>*> terms) {
    uint64_t v = 0;
    for (const auto* t : terms) v += CS::value(t->word);
    const UInt<CS> wide = uint_from_word(cs, cs.hint(width, v));
    typename CS::Z input = CS::zconst(0);
    for (const auto* t : terms) input += t->full;
    assert_zero(cs, input - wide.full);
    return {CS::low(wide.word, 32), wide.low, wide.low};
}

enum class Doubled { kChoice, kMajority };
template <class CS>
struct DoubledTerm {
    typename CS::Z twice;  // the Z expression of 2 Ch or 2 Maj
    Doubled kind;
    const typename CS::Word *x, *y, *z;
    uint64_t value() const {
        const uint32_t a = (uint32_t)CS::value(*x), b = (uint32_t)CS::value(*y), c = (uint32_t)CS::value(*z);
        return kind == Doubled::kChoice ? ((a & b) ^ (~a & c)) : ((a & b) ^ (a & c) ^ (b & c));
    }
};

// sum_doubled_32 (sha256.rs:267-315): hint half = sum(ordinary) + sum(Ch/Maj
// values) as 35 bits; constrain 2 sum(ordinary) + sum(doubled) - 2 half = 0.
template <class CS>
UInt<CS> sum_doubled_32(CS& cs, std::initializer_list<const UInt<CS>*> ordinary,
                        std::initializer_list<const DoubledTerm<CS>*> doubled) {
    uint64_t v = 0;
    for (const auto* t : ordinary) v += CS::value(t->word);
    for (const auto* d : doubled) v += d->value();
    const UInt<CS> half = uint_from_word(cs, cs.hint(kMaxWidth, v));
    typename CS::Z ord = CS::zconst(0), dbl = CS::zconst(0);
    for (const auto* t : ordinary) ord += t->full;
    for (const auto* d : doubled) dbl += d->twice;
    assert_zero(cs, ord * 2 + dbl - half.full * 2);
    return {CS::low(half.word, 32), half.low, half.low};
}

template <class CS>
typename CS::Word xor3(const typename CS::Word& a, const typename CS::Word& b, const typename CS::Word& c) {
    return CS::xor_(CS::xor_(a, b), c);
}

// compress (sha256.rs:344-470).
template <class CS>
void compress(CS& cs, const typename CS::Word block[16], const typename CS::Word state_in[8],
              UInt<CS> out[8]) {
    using W = typename CS::Word;
    std::vector<UInt<CS>> state, sched;
    state.reserve(8);
    sched.reserve(64);
    for (int i = 0; i < 8; ++i) state.push_back(uint_from_word(cs, state_in[i]));
    for (int i = 0; i < 16; ++i) sched.push_back(uint_from_word(cs, block[i]));
    for (int i = 16; i < 64; ++i) {
        const W& w15 = sched[i - 15].word;
        const UInt<CS> s0 = uint_from_word(cs, xor3<CS>(CS::rotr(w15, 7), CS::rotr(w15, 18), CS::shr(w15, 3)));
        const W& w2 = sched[i - 2].word;
        const UInt<CS> s1 = uint_from_word(cs, xor3<CS>(CS::rotr(w2, 17), CS::rotr(w2, 19), CS::shr(w2, 10)));
        sched.push_back(sum_32(cs, 34, {&sched[i - 16], &s0, &sched[i - 7], &s1}));
    }
    UInt<CS> a = state[0], b = state[1], c = state[2], d = state[3];
    UInt<CS> e = state[4], f = state[5], g = state[6], h = state[7];
    for (int i = 0; i < 64; ++i) {
        const UInt<CS> S1 = uint_from_word(cs, xor3<CS>(CS::rotr(e.word, 6), CS::rotr(e.word, 11), CS::rotr(e.word, 25)));
        // choice_twice: y + z - (x ^ y) + (x ^ z)
        const UInt<CS> xy = uint_from_word(cs, CS::xor_(e.word, f.word));
        const UInt<CS> xz = uint_from_word(cs, CS::xor_(e.word, g.word));
        const DoubledTerm<CS> ch{f.full + g.full - xy.full + xz.full, Doubled::kChoice, &e.word, &f.word, &g.word};
        const UInt<CS> S0 = uint_from_word(cs, xor3<CS>(CS::rotr(a.word, 2), CS::rotr(a.word, 13), CS::rotr(a.word, 22)));
        // majority_twice: x + y + z - (x ^ y ^ z)
        const UInt<CS> xyz = uint_from_word(cs, xor3<CS>(a.word, b.word, c.word));
        const DoubledTerm<CS> maj{a.full + b.full + c.full - xyz.full, Doubled::kMajority, &a.word, &b.word, &c.word};
        const UInt<CS> K = uint_constant<CS>(kK[i]);
        UInt<CS> new_e = sum_doubled_32(cs, {&d, &h, &S1, &K, &sched[i]}, {&ch});
        UInt<CS> new_a = sum_doubled_32(cs, {&h, &S1, &K, &sched[i], &S0}, {&ch, &maj});
        h = g;
        g = f;
        f = e;
        e = std::move(new_e);
        d = c;
        c = b;
        b = a;
        a = std::move(new_a);
    }
    const UInt<CS>* working[8] = {&a, &b, &c, &d, &e, &f, &g, &h};
    for (int i = 0; i < 8; ++i) out[i] = sum_32(cs, 33, {&state[i], working[i]});
}

// The public-bit constraint of the CLI's statements (tooling/cli/src/
// circuits.rs:212-217): 1 * bit = expected.
template <class CS, class Bit>
void constrain_bit(CS& cs, const Bit& bit, bool expected) {
    cs.assert_r1c(CS::zconst(1), cs.bitz(bit), CS::zconst(expected ? 1 : 0));
}

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

// Plain SHA-256 compression for the statements' expected digests.
void compress_plain(uint32_t st[8], const uint8_t* block) {
    uint32_t w[64];
    auto rotr = [](uint32_t x, int k) { return (x >> k) | (x << (32 - k)); };
    for (int i = 0; i < 16; ++i) w[i] = be32(block + 4 * i);
    for (int i = 16; i < 64; ++i) {
        const uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
        const uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
        w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint32_t a = st[0], b = st[1], c = st[2], d = st[3], e = st[4], f = st[5], g = st[6], h = st[7];
    for (int i = 0; i < 64; ++i) {
        const uint32_t t1 = h + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) + kK[i] + w[i];
        const uint32_t t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
        h = g;
        g = f;
        f = e;
        e = d + t1;
        d = c;
        c = b;
        b = a;
        a = t1 + t2;
    }
    st[0] += a, st[1] += b, st[2] += c, st[3] += d, st[4] += e, st[5] += f, st[6] += g, st[7] += h;
}

}  // namespace

struct bitz_sha256 {
    uint32_t kind = 0;
    size_t blocks = 0;
    size_t inputs = 0;          // Boolean input witnesses
    size_t witness_bits = 0;    // |f|: inputs + hint bits
    size_t rows = 0;            // R1CS rows
    size_t columns = 0;         // |h| (constant included) = rows of M
    std::vector<uint64_t> m_starts;
    std::vector<uint32_t> m_pos;
    Csr mat[3];
    std::vector<uint8_t> statement;  // BITZ_SHA256_CHAIN: the message the rows are built for
};

namespace {

// The circuit for `kind` over a backend: inputs are the message bits (and the
// chaining value for the compression chain); see the kinds in bitz_hip.h.
template <class CS>
void synthesize(CS& cs, uint32_t kind, size_t blocks, const typename CS::Word* in_words,
                const typename CS::Word iv_words[8], const uint8_t* statement, uint8_t digest_out[32]) {
    using W = typename CS::Word;
    W state[8];
    UInt<CS> out[8];
    if (kind == BITZ_SHA256_CHAIN) {
        // circuits.rs:167-209: constrain every input bit, compress from the
        // FIPS IV, constrain every output bit.
        uint32_t st[8];
        memcpy(st, kIV, sizeof st);
        for (size_t b = 0; b < blocks; ++b) compress_plain(st, statement + 64 * b);
        for (size_t b = 0; b < blocks; ++b)
            for (int w = 0; w < 16; ++w) {
                const uint32_t word = be32(statement + 64 * b + 4 * w);
                for (int j = 0; j < 32; ++j) constrain_bit(cs, CS::bit(in_words[16 * b + w], j), (word >> j) & 1);
            }
        for (int i = 0; i < 8; ++i) state[i] = CS::constant(kIV[i], 32);
        for (size_t b = 0; b < blocks; ++b) {
            compress(cs, in_words + 16 * b, state, out);
            for (int i = 0; i < 8; ++i) state[i] = out[i].word;
        }
        for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 32; ++j) constrain_bit(cs, CS::bit(state[i], j), (st[i] >> j) & 1);
    } else if (kind == BITZ_SHA256_BLOCK_ALIGNED) {
        // sha256.rs:503-550: message blocks from the FIPS IV, then the
        // padding block of constants.
        for (int i = 0; i < 8; ++i) state[i] = CS::constant(kIV[i], 32);
        for (size_t b = 0; b < blocks; ++b) {
            compress(cs, in_words + 16 * b, state, out);
            for (int i = 0; i < 8; ++i) state[i] = out[i].word;
        }
        const uint64_t len = (uint64_t)blocks * 512;
        W pad[16];
        for (int w = 0; w < 16; ++w) pad[w] = CS::constant(0, 32);
        pad[0] = CS::constant(0x80000000u, 32);
        pad[14] = CS::constant((uint32_t)(len >> 32), 32);
        pad[15] = CS::constant((uint32_t)len, 32);
        compress(cs, pad, state, out);
        for (int i = 0; i < 8; ++i) state[i] = out[i].word;
    } else {
        // compression_circuit (sha256.rs:476-490) chained: block b's state
        // is block b-1's output words, block 0's is the input chaining value.
        for (int i = 0; i < 8; ++i) state[i] = iv_words[i];
        for (size_t b = 0; b < blocks; ++b) {
            compress(cs, in_words + 16 * b, state, out);
            for (int i = 0; i < 8; ++i) state[i] = out[i].word;
        }
    }
    if (digest_out)
        for (int i = 0; i < 8; ++i) {
            const uint32_t v = (uint32_t)CS::value(state[i]);
            for (int k = 0; k < 4; ++k) digest_out[4 * i + k] = (uint8_t)(v >> (24 - 8 * k));
        }
}

size_t input_bits(uint32_t kind, size_t blocks) {
    return 512 * blocks + (kind == BITZ_SHA256_COMPRESSION_CHAIN ? 256 : 0);
}

// Block b, word w as input witnesses.  Chain kinds: bit j of the word is the
// word's bit j (word_bits, circuits.rs:219-227).  Block-aligned: the message
// is a bit stream, bytes in order, each most significant bit first, and bit j
// of word w is stream bit 32 w + 31 - j (sha256.rs:520-526).
Symbolic::Word sym_input_word(uint32_t kind, size_t b, int w) {
    Symbolic::Word x;
    x.n = 32;
    for (int j = 0; j < 32; ++j) {
        const size_t idx = kind == BITZ_SHA256_BLOCK_ALIGNED ? 512 * b + 32 * (size_t)w + (31 - j)
                                                             : 512 * b + 32 * (size_t)w + (size_t)j;
        x.b[j].w[0] = (uint32_t)idx;
        x.b[j].n = 1;
    }
    return x;
}

}  // namespace

extern "C" int32_t bitz_sha256_circuit(bitz_sha256** out, uint32_t kind, size_t blocks, const uint8_t* statement) {
    if (!out) return bitz::report_error(BITZ_ERR_ARG, "null pointer");
    if (kind > BITZ_SHA256_BLOCK_ALIGNED) return bitz::report_error(BITZ_ERR_ARG, "unknown SHA-256 circuit kind");
    if (blocks == 0 && kind != BITZ_SHA256_BLOCK_ALIGNED)
        return bitz::report_error(BITZ_ERR_ARG, "the chain needs at least one block");
    if (blocks > 4096) return bitz::report_error(BITZ_ERR_ARG, "at most 4096 blocks");
    if (kind == BITZ_SHA256_CHAIN && !statement)
        return bitz::report_error(BITZ_ERR_ARG, "the sha256-chain statement needs its message");
    auto* c = new (std::nothrow) bitz_sha256;
    if (!c) return bitz::report_error(BITZ_ERR_ALLOC, "host allocation failed");
    try {
        c->kind = kind;
        c->blocks = blocks;
        c->inputs = input_bits(kind, blocks);
        if (kind == BITZ_SHA256_CHAIN) c->statement.assign(statement, statement + 64 * blocks);
        Symbolic cs((uint32_t)c->inputs);
        std::vector<Symbolic::Word> words(16 * blocks + 1);
        for (size_t b = 0; b < blocks; ++b)
            for (int w = 0; w < 16; ++w) words[16 * b + w] = sym_input_word(kind, b, w);
        Symbolic::Word iv[8];
        for (int i = 0; i < 8; ++i) iv[i] = Symbolic::input((uint32_t)(512 * blocks + 32 * i), 32);
        synthesize(cs, kind, blocks, words.data(), iv, statement, nullptr);
        c->witness_bits = cs.next_witness;
        c->columns = cs.m_starts.size() - 1;
        c->rows = cs.a.starts.size() - 1;
        c->m_starts = std::move(cs.m_starts);
        c->m_pos = std::move(cs.m_pos);
        c->mat[0] = std::move(cs.a);
        c->mat[1] = std::move(cs.b);
        c->mat[2] = std::move(cs.c);
    } catch (const Failure& f) {
        delete c;
        return bitz::report_error(f.code, f.msg);
    } catch (const std::bad_alloc&) {
        delete c;
        return bitz::report_error(BITZ_ERR_ALLOC, "host allocation failed");
    }
    *out = c;
    return BITZ_OK;
}

extern "C" int32_t bitz_sha256_shape(const bitz_sha256* c, size_t* rows, size_t* columns, size_t* witness_bits,
                                     size_t nnz[3], size_t* m_positions) {
    if (!c) return bitz::report_error(BITZ_ERR_ARG, "null pointer");
    if (rows) *rows = c->rows;
    if (columns) *columns = c->columns;
    if (witness_bits) *witness_bits = c->witness_bits;
    if (nnz)
        for (int m = 0; m < 3; ++m) nnz[m] = c->mat[m].cols.size();
    if (m_positions) *m_positions = c->m_pos.size();
    return BITZ_OK;
}

extern "C" int32_t bitz_sha256_export(const bitz_sha256* c, uint64_t* const row_starts[3], uint32_t* const cols[3],
                                      int64_t* const coefficients[3], bitz_f128* const parity[3],
                                      uint64_t* m_row_starts, uint32_t* m_positions) {
    if (!c) return bitz::report_error(BITZ_ERR_ARG, "null pointer");
    for (int m = 0; m < 3; ++m) {
        const Csr& x = c->mat[m];
        if (row_starts && row_starts[m]) memcpy(row_starts[m], x.starts.data(), x.starts.size() * 8);
        if (cols && cols[m]) memcpy(cols[m], x.cols.data(), x.cols.size() * 4);
        if (coefficients && coefficients[m]) memcpy(coefficients[m], x.coef.data(), x.coef.size() * 8);
        if (parity && parity[m])
            for (size_t k = 0; k < x.coef.size(); ++k) parity[m][k] = bitz_f128{(uint64_t)(x.coef[k] & 1), 0};
    }
    if (m_row_starts) memcpy(m_row_starts, c->m_starts.data(), c->m_starts.size() * 8);
    if (m_positions) memcpy(m_positions, c->m_pos.data(), c->m_pos.size() * 4);
    return BITZ_OK;
}

extern "C" int32_t bitz_sha256_witness(const bitz_sha256* c, const uint8_t* message, const uint32_t iv[8],
                                       uint64_t* f_words, uint64_t* h_words, uint8_t digest[32],
                                       int32_t* satisfied) {
    if (!c || !message) return bitz::report_error(BITZ_ERR_ARG, "null pointer");
    if (iv && c->kind != BITZ_SHA256_COMPRESSION_CHAIN)
        return bitz::report_error(BITZ_ERR_ARG, "only the compression chain takes a chaining value");
    try {
        Concrete cs;
        const size_t blocks = c->blocks;
        cs.f.words.reserve((c->witness_bits + 63) / 64);
        cs.h.words.reserve((c->columns + 63) / 64);
        std::vector<ValWord> words(16 * blocks + 1);
        for (size_t b = 0; b < blocks; ++b)
            for (int w = 0; w < 16; ++w) words[16 * b + w] = {be32(message + 64 * b + 4 * w), 32};
        ValWord ivw[8];
        for (int i = 0; i < 8; ++i) ivw[i] = {iv ? iv[i] : kIV[i], 32};
        // f starts with the inputs, in the order the symbolic run numbered them.
        for (size_t b = 0; b < blocks; ++b)
            for (int w = 0; w < 16; ++w) {
                const uint32_t v = (uint32_t)words[16 * b + w].v;
                if (c->kind == BITZ_SHA256_BLOCK_ALIGNED)
                    for (int j = 31; j >= 0; --j) cs.f.append((v >> j) & 1, 1);
                else cs.f.append(v, 32);
            }
        if (c->kind == BITZ_SHA256_COMPRESSION_CHAIN)
            for (int i = 0; i < 8; ++i) cs.f.append(ivw[i].v, 32);
        uint8_t dg[32];
        synthesize(cs, c->kind, blocks, words.data(), ivw,
                   c->kind == BITZ_SHA256_CHAIN ? c->statement.data() : message, dg);
        if (cs.f.len != c->witness_bits || cs.h.len != c->columns)
            return bitz::report_error(BITZ_ERR_STATE, "witness length differs from the circuit's");
        if (f_words) memcpy(f_words, cs.f.words.data(), cs.f.words.size() * 8);
        if (h_words) memcpy(h_words, cs.h.words.data(), cs.h.words.size() * 8);
        if (digest) memcpy(digest, dg, 32);
        if (satisfied) *satisfied = cs.satisfied ? 1 : 0;
    } catch (const std::bad_alloc&) {
        return bitz::report_error(BITZ_ERR_ALLOC, "host allocation failed");
    }
    return BITZ_OK;
}

extern "C" int32_t bitz_sha256_prepare(const bitz_sha256* c, bitz_matrices** out, uint8_t digest[32]) {
    if (!c || (!out && !digest)) return bitz::report_error(BITZ_ERR_ARG, "null pointer");
    std::vector<bitz_f128> vals[3];
    const uint64_t* starts[3];
    const uint32_t* cols[3];
    const bitz_f128* values[3];
    try {
        for (int m = 0; m < 3; ++m) {
            const Csr& x = c->mat[m];
            vals[m].resize(std::max<size_t>(x.coef.size(), 1));
            for (size_t k = 0; k < x.coef.size(); ++k) vals[m][k] = bitz_f128{(uint64_t)(x.coef[k] & 1), 0};
            starts[m] = x.starts.data();
            cols[m] = x.cols.data();
            values[m] = vals[m].data();
        }
    } catch (const std::bad_alloc&) {
        return bitz::report_error(BITZ_ERR_ALLOC, "host allocation failed");
    }
    auto vars = [](size_t n) {
        uint32_t v = 0;
        while (((size_t)1 << v) < std::max<size_t>(n, 1)) ++v;
        return v;
    };
    int32_t rc;
    if (digest &&
        (rc = bitz_constraint_matrix_digest(c->columns, c->witness_bits + 1, c->m_starts.data(), c->m_pos.data(),
                                            c->rows, c->columns, starts, cols, values, digest)))
        return rc;
    if (out && (rc = bitz_matrices_prepare(out, c->rows, vars(c->rows), vars(c->columns), starts, cols, values)))
        return rc;
    return BITZ_OK;
}

extern "C" void bitz_sha256_free(bitz_sha256* c) { delete c; }
