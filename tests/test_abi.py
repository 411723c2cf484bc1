"""CPU checks of the C ABI library: it loads, exports every function that
include/bitz_hip.h declares, and rejects bad dimensions before touching the
device or the transcript (crates/spartan/src/sumcheck.rs:258-269)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bitz_hip.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(bitz_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import bitz_amd
    lib = bitz_amd.lib()
    names = declared()
    assert len(names) > 30
    for name in names:
        assert hasattr(lib, name), name


def test_dimension_errors_need_no_device():
    import bitz_amd
    L = bitz_amd.lib()
    one = (C.c_uint64 * 2)(1, 0)
    tr = C.c_void_p(0)
    # mismatched product dims -> InvalidProductDimensions
    rc = L.bitz_prove_outer_sumcheck(tr, bitz_amd.F128(0, 0), one, 0, one, 0, one, 0, one, 1, one, 0, 0,
                                     one, one, one, one)
    assert rc in (-1, -7)  # a null transcript may be caught first
    rc = L.bitz_outer_begin(C.byref(C.c_void_p()), one, 0, one, 0, one, 0, one, 1, one, 0, 0)
    assert rc == -1
    rc = L.bitz_outer_begin(C.byref(C.c_void_p()), one, 1, one, 0, one, 0, one, 0, one, 0, 0)
    assert rc == -2
    assert b"eq factor" in L.bitz_last_error()


def test_short_challenge_arrays_are_rejected_before_any_call():
    """The queued provers read exactly num_vars challenges; a shorter array
    raises in the wrapper instead of reading past its end."""
    import numpy as np

    import bitz_amd as G
    n = 3
    t = np.zeros((1 << n, 2), np.uint64)
    lo, hi = np.zeros((2, 2), np.uint64), np.zeros((4, 2), np.uint64)
    for short in (np.zeros((2, 2), np.uint64), []):
        with pytest.raises(G.BitzError) as e:
            G.prove_outer_sumcheck_with_challenges(short, 0, (lo, hi), G.R1csProductMles(t, t, t))
        assert e.value.variant == "InvalidArgument" and "challenges" in str(e.value)
        with pytest.raises(G.BitzError) as e:
            G.prove_inner_sumcheck_with_challenges(short, 0, t, t)
        assert e.value.variant == "InvalidArgument" and "challenges" in str(e.value)
