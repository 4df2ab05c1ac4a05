"""The no-alloc Heff entry point (itg_heff2_apply_noalloc) against the entry
point with a caller work buffer (itg_heff2_apply), both called through the C
ABI with ctypes: two calls in a row, the second at a smaller shape (the
scratch array keeps the first call's contents), in each arithmetic mode.
Both run the same kernels on the same inputs, so the outputs agree bit for
bit; the FP64 output is also checked against the oracle's product."""
import ctypes as C

import numpy as np
import pytest

from oracle import dmrg as odmrg
from oracle import tensor_ops as ops

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from itg import _lib, ops as d
    _lib.ctx()
    return d


def _operands(dev, chi, w, d, seed):
    rng = np.random.default_rng(seed)
    L = rng.standard_normal((chi, w, chi))
    R = rng.standard_normal((chi, w, chi))
    W1 = rng.standard_normal((w, w, d, d))
    W2 = rng.standard_normal((w, w, d, d))
    psi = rng.standard_normal((chi, d, d, chi))
    host = (L, W1, W2, R, psi)
    return host, [dev.to_device(ops.to_flat(a)) for a in host]


@pytest.mark.parametrize("mode", ["native", "fp32", "bf16x3"])
def test_heff2_apply_noalloc_matches_work_buffer_entry(dev, mode):
    from itg import _lib
    lib, c = _lib.lib(), _lib.ctx()
    w, d = 5, 2
    for call, chi in enumerate((120, 48)):
        print("noalloc case: mode=%s call=%d chi=%d" % (mode, call, chi), flush=True)
        host, (L, W1, W2, R, psi) = _operands(dev, chi, w, d, seed=10 + chi)
        n = chi * d * d * chi
        work = dev.heff2_work(chi, w, w, w, d, chi)
        y_ref = dev.empty(n)
        y = dev.empty(n)
        args = (c.ptr, _lib.F64, _lib.MODES[mode], chi, w, w, w, d, chi, dev.ptr(L), dev.ptr(W1),
                dev.ptr(W2), dev.ptr(R), dev.ptr(psi))
        c.check(lib.itg_heff2_apply(*args, dev.ptr(y_ref), dev.ptr(work)))
        c.check(lib.itg_heff2_apply_noalloc(*args, dev.ptr(y)))
        got, want = dev.to_host(y), dev.to_host(y_ref)
        np.testing.assert_array_equal(got, want)
        if mode == "native":
            ref = ops.to_flat(odmrg.heff_apply(*host))
            err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            assert err < 1e-12, err


def test_heff2_apply_noalloc_rejects_shapes_beyond_scratch(dev):
    """A shape whose work buffer exceeds the fixed scratch is refused on the
    host, before any launch, with a message naming the limit."""
    from itg import _lib
    lib, c = _lib.lib(), _lib.ctx()
    cap = lib.itg_heff2_noalloc_max_bytes()
    assert cap == 64 << 20
    chi, w, d = 1024, 5, 2
    assert lib.itg_heff2_work_elems(chi, w, w, w, d, chi) * 4 > cap
    dummy = dev.empty(16)
    p = dev.ptr(dummy)
    rc = lib.itg_heff2_apply_noalloc(c.ptr, _lib.F64, _lib.MODE_NATIVE, chi, w, w, w, d, chi,
                                     p, p, p, p, p, p)
    assert rc == _lib.ITG_EUNSUPPORTED
    msg = lib.itg_last_error(c.ptr).decode()
    assert "no-alloc scratch" in msg and str(cap) in msg
