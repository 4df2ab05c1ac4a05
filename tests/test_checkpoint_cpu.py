"""CPU tier: the checkpoint file format (itg/checkpoint.py) round-trips
arrays and metadata exactly, and refuses damaged or foreign files."""
import os

import numpy as np
import pytest

from itg import checkpoint as ck


def test_roundtrip_exact(tmp_path):
    rng = np.random.default_rng(0)
    arrays = [("site.0", rng.standard_normal(1 * 2 * 3), (1, 2, 3)),
              ("site.1", rng.standard_normal(3 * 2 * 5).astype(np.float32), (3, 2, 5)),
              ("empty", np.zeros(0), (0,))]
    meta = {"sweeps_done": 3, "energy": -1.2345678901234567, "schedule": {"maxdim": [10, 20]}}
    p = str(tmp_path / "a.itg")
    ck.write_arrays(p, arrays, meta)
    got, m = ck.read_arrays(p)
    assert m == meta
    for (n0, a0, d0), (n1, a1, d1) in zip(arrays, got):
        assert n0 == n1 and tuple(d0) == d1 and a0.dtype == a1.dtype
        np.testing.assert_array_equal(a0, a1)
    assert not os.path.exists(p + ".tmp")


def test_rejects_corruption_truncation_and_foreign_files(tmp_path):
    p = str(tmp_path / "b.itg")
    ck.write_arrays(p, [("x", np.arange(100, dtype=np.float64), (10, 10))])
    raw = bytearray(open(p, "rb").read())
    bad = bytearray(raw)
    bad[-3] ^= 0x40                       # flip a bit in the data
    open(p, "wb").write(bad)
    with pytest.raises(ck.CheckpointError, match="CRC"):
        ck.read_arrays(p)
    open(p, "wb").write(raw[:-16])        # cut the last array short
    with pytest.raises(ck.CheckpointError, match="truncated"):
        ck.read_arrays(p)
    open(p, "wb").write(b"PK\x03\x04" + raw[4:])
    with pytest.raises(ck.CheckpointError, match="not a checkpoint"):
        ck.read_arrays(p)


def test_rejects_size_mismatch(tmp_path):
    with pytest.raises(ck.CheckpointError):
        ck.write_arrays(str(tmp_path / "c.itg"), [("x", np.zeros(5), (2, 3))])


def test_sweeps_schedule_roundtrip_and_tail():
    from itg.mps import Sweeps
    sw = Sweeps(5, maxdim=[10, 20, 100], cutoff=[1e-10], noise=[1e-6, 0.0])
    d = sw.as_dict()
    s2 = Sweeps.from_dict(d)
    assert s2.as_dict() == d
    t = sw.tail(2)
    assert t.nsweep == 3 and t.maxdim == [100, 100, 100] and t.noise == [0.0, 0.0, 0.0]
