"""Independent pure-Python models that pin the oracle (test infrastructure).

- GF(2^128) by definition: carry-less products term by term on Python ints
  and reduction by explicit long division, shaped unlike the oracle's
  word-level pipeline (the reference pins its own `portable` module the same
  way: crates/field/src/gf128/portable.rs:144-214).
- AES-128 (FIPS-197), only to recover GHASH's subkey and tag mask from the
  NIST GCM vectors (crates/field/tests/ghash_vectors.rs:129-161).  AES lives in
  GF(2^8) under another polynomial, so it shares no arithmetic with the field
  under test.
"""
from __future__ import annotations

MASK64 = (1 << 64) - 1
POLY = (1 << 128) | 0x87  # X^128 + X^7 + X^2 + X + 1


def clmul_def(a: int, b: int) -> int:
    """Polynomial product over GF(2): coefficient k is the parity of
    #{(i, j): i + j = k, a_i = b_j = 1}."""
    out = 0
    i = 0
    while a:
        if a & 1:
            out ^= b << i
        a >>= 1
        i += 1
    return out


def reduce_longdiv(x: int) -> int:
    """Long division by X^128 + X^7 + X^2 + X + 1: clear the top set bit."""
    while x.bit_length() > 128:
        x ^= POLY << (x.bit_length() - 129)
    return x


def gf_mul(a: int, b: int) -> int:
    return reduce_longdiv(clmul_def(a, b))


def gf_pow(a: int, e: int) -> int:
    r = 1
    while e:
        if e & 1:
            r = gf_mul(r, a)
        a = gf_mul(a, a)
        e >>= 1
    return r


def gf_inv(a: int) -> int:
    assert a != 0
    return gf_pow(a, (1 << 128) - 2)


def words(x: int) -> tuple[int, int]:
    return x & MASK64, x >> 64


# ---- AES-128 (FIPS-197) ----------------------------------------------------------


def _xtime(b: int) -> int:
    b <<= 1
    return (b ^ 0x11B) & 0xFF if b & 0x100 else b


def _gmul8(a: int, b: int) -> int:
    r = 0
    while b:
        if b & 1:
            r ^= a
        a = _xtime(a)
        b >>= 1
    return r


def _sbox() -> list[int]:
    box = [0] * 256
    for x in range(256):
        inv = 0 if x == 0 else next(y for y in range(1, 256) if _gmul8(x, y) == 1)
        s = inv
        for k in range(1, 5):
            s ^= ((inv << k) | (inv >> (8 - k))) & 0xFF
        box[x] = s ^ 0x63
    return box


SBOX = _sbox()


def _expand(key: bytes) -> list[list[int]]:
    w = [list(key[4 * i: 4 * i + 4]) for i in range(4)]
    rcon = 1
    for i in range(4, 44):
        t = list(w[i - 1])
        if i % 4 == 0:
            t = t[1:] + t[:1]
            t = [SBOX[b] for b in t]
            t[0] ^= rcon
            rcon = _xtime(rcon)
        w.append([w[i - 4][j] ^ t[j] for j in range(4)])
    return w


def aes128_encrypt(key: bytes, block: bytes) -> bytes:
    w = _expand(key)
    s = [block[i] ^ w[i // 4][i % 4] for i in range(16)]
    for rnd in range(1, 11):
        s = [SBOX[b] for b in s]
        # ShiftRows: state is column-major (byte i is row i % 4, column i // 4)
        s = [s[(i + 4 * (i % 4)) % 16] for i in range(16)]
        if rnd != 10:
            t = []
            for c in range(4):
                a = s[4 * c: 4 * c + 4]
                t += [
                    _gmul8(a[0], 2) ^ _gmul8(a[1], 3) ^ a[2] ^ a[3],
                    a[0] ^ _gmul8(a[1], 2) ^ _gmul8(a[2], 3) ^ a[3],
                    a[0] ^ a[1] ^ _gmul8(a[2], 2) ^ _gmul8(a[3], 3),
                    _gmul8(a[0], 3) ^ a[1] ^ a[2] ^ _gmul8(a[3], 2),
                ]
            s = t
        s = [s[i] ^ w[4 * rnd + i // 4][i % 4] for i in range(16)]
    return bytes(s)


def _rev8(b: int) -> int:
    return int(f"{b:08b}"[::-1], 2)


def from_nist(block: bytes) -> int:
    """NIST's bit order (MSB of byte 0 = X^0) to this field's: reverse the
    bits within each byte, read 16 bytes little-endian
    (crates/field/tests/ghash_vectors.rs:23-29)."""
    return int.from_bytes(bytes(_rev8(b) for b in block), "little")


def to_nist(x: int) -> bytes:
    return bytes(_rev8(b) for b in x.to_bytes(16, "little"))


def parse_rsp(text: str) -> list[dict]:
    cases, cur, tag_bytes, in_rec = [], {}, 0, False
    for line in text.splitlines():
        line = line.strip()
        if not line or line.startswith("#"):
            if in_rec:
                cases.append(cur)
                in_rec = False
            continue
        if line.startswith("[") and line.endswith("]"):
            p = line[1:-1]
            if p.startswith("Taglen = "):
                tag_bytes = int(p.split("=")[1]) // 8
            continue
        name, _, value = line.partition("=")
        name, value = name.strip(), value.strip()
        if name == "Count":
            cur = {"tag_bytes": tag_bytes}
            in_rec = True
        elif name in ("Key", "IV", "AAD", "CT", "Tag"):
            cur[name] = bytes.fromhex(value)
    if in_rec:
        cases.append(cur)
    return cases
