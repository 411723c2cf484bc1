"""The Rust binding (crates/hip-backend) against the header and its
documentation, on the host: every C function the crate declares exists in
include/bitz_hip.h with as many parameters, is exported by the library, and
the crate's blocks quoted in INTEGRATION.md are verbatim copies."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CRATE = os.path.join(ROOT, "crates", "hip-backend")


def read(*parts):
    with open(os.path.join(*parts)) as f:
        return f.read()


def params(sig: str) -> int:
    sig = re.sub(r"/\*.*?\*/", "", sig, flags=re.S).strip()
    if sig in ("", "void"):
        return 0
    depth, n = 0, 1
    for ch in sig:
        depth += ch in "([<"
        depth -= ch in ")]>"
        n += ch == "," and depth == 0
    return n - sig.rstrip().endswith(",")


def test_ffi_declarations_match_the_header_and_the_library():
    import bitz_amd
    header = read(ROOT, "include", "bitz_hip.h")
    ffi = read(CRATE, "src", "ffi.rs")
    decls = re.findall(r"pub fn (bitz_\w+)\(([^)]*)\)", ffi, flags=re.S)
    assert len(decls) >= 25
    lib = bitz_amd.lib()
    for name, args in decls:
        m = re.search(r"\b" + name + r"\(([^;]*?)\);", header, flags=re.S)
        assert m, f"{name} is not declared in the header"
        assert params(args) == params(m.group(1)), name
        assert hasattr(lib, name), name


def test_integration_quotes_the_crate_verbatim():
    doc = read(ROOT, "INTEGRATION.md")
    part = doc[doc.index("## Cargo.toml"): doc.index("## Batches")]
    blocks = re.findall(r"```(?:rust|toml)\n(.*?)```", part, flags=re.S)
    assert len(blocks) >= 8
    sources = "\n".join(read(CRATE, *p) for p in (("Cargo.toml",), ("build.rs",), ("src", "ffi.rs"),
                                                   ("src", "lib.rs")))
    for b in blocks:
        assert b in sources, b[:200]
