"""MPS / MPO / sweep-state checkpoints.

The reference declares HDF5 (Project.toml:12) for ITensors' MPS files; this
image has no HDF5, so the file format here is its own and self-describing:

    bytes 0..7     magic b"ITGCKPT1"
    bytes 8..15    little-endian uint64 H: length of the JSON header
    next H bytes   JSON header (UTF-8)
    then           the arrays, each starting at a 64-byte aligned offset
                   from the start of the data section

The header holds {"format": 1, "arrays": [{"name", "dtype", "dims",
"offset", "nbytes", "crc32"}, ...], "meta": {...}}. Arrays are stored flat
in column-major order (first index fastest), the layout of Dense storage, so
a device buffer is written and read back without any reordering. A file is
written to `path + ".tmp"` and renamed over `path`, so a crash during a
write leaves the previous checkpoint intact. Reads check the magic, the
sizes and every array's CRC-32, and raise on any mismatch.
"""
import json
import os
import struct
import zlib

import numpy as np

MAGIC = b"ITGCKPT1"
ALIGN = 64


class CheckpointError(RuntimeError):
    pass


def write_arrays(path, arrays, meta=None):
    """arrays: list of (name, flat numpy array, dims). Atomic replace."""
    entries, off = [], 0
    for name, a, dims in arrays:
        a = np.ascontiguousarray(a)
        if a.dtype not in (np.float64, np.float32):
            raise CheckpointError("unsupported dtype %s for %s" % (a.dtype, name))
        n = 1
        for d in dims:
            n *= int(d)
        if a.size != n:
            raise CheckpointError("%s: %d elements for dims %s" % (name, a.size, dims))
        off = (off + ALIGN - 1) // ALIGN * ALIGN
        entries.append({"name": name, "dtype": str(a.dtype), "dims": [int(d) for d in dims],
                        "offset": off, "nbytes": int(a.nbytes),
                        "crc32": zlib.crc32(a.view(np.uint8)) & 0xFFFFFFFF})
        off += a.nbytes
    header = json.dumps({"format": 1, "arrays": entries, "meta": meta or {}}).encode()
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<Q", len(header)))
        f.write(header)
        base = f.tell()
        for (name, a, dims), e in zip(arrays, entries):
            f.seek(base + e["offset"])
            f.write(np.ascontiguousarray(a).tobytes())
        f.flush()
        os.fsync(f.fileno())
    os.replace(tmp, path)


def read_arrays(path):
    """Returns (list of (name, flat array, dims), meta)."""
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise CheckpointError("%s: not a checkpoint file" % path)
        raw = f.read(8)
        if len(raw) != 8:
            raise CheckpointError("%s: truncated header" % path)
        (hlen,) = struct.unpack("<Q", raw)
        hdr = f.read(hlen)
        if len(hdr) != hlen:
            raise CheckpointError("%s: truncated header" % path)
        h = json.loads(hdr.decode())
        if h.get("format") != 1:
            raise CheckpointError("%s: unknown format %r" % (path, h.get("format")))
        base = f.tell()
        out = []
        for e in h["arrays"]:
            f.seek(base + e["offset"])
            buf = f.read(e["nbytes"])
            if len(buf) != e["nbytes"]:
                raise CheckpointError("%s: array %s truncated" % (path, e["name"]))
            if zlib.crc32(buf) & 0xFFFFFFFF != e["crc32"]:
                raise CheckpointError("%s: array %s fails its CRC-32" % (path, e["name"]))
            out.append((e["name"], np.frombuffer(buf, dtype=np.dtype(e["dtype"])).copy(),
                        tuple(e["dims"])))
    return out, h["meta"]


# ------------------------------------------------------------ device objects
def save_mps(path, psi, meta=None):
    """MPS site tensors (l, s, r) plus its orthogonality limits and `meta`."""
    from . import ops
    arrays = [("site.%d" % j, ops.to_host(t), dm) for j, (t, dm) in
              enumerate(zip(psi.data, psi.dims))]
    m = dict(meta or {})
    m.update({"kind": "mps", "N": len(psi), "llim": psi.llim, "rlim": psi.rlim})
    write_arrays(path, arrays, m)


def load_mps(path, dtype=None):
    from . import ops
    from .mps import MPS
    arrays, meta = read_arrays(path)
    if meta.get("kind") != "mps":
        raise CheckpointError("%s holds %r, not an MPS" % (path, meta.get("kind")))
    sites = sorted((a for a in arrays if a[0].startswith("site.")),
                   key=lambda a: int(a[0].split(".")[1]))
    if len(sites) != meta["N"]:
        raise CheckpointError("%s: %d site tensors for N=%d" % (path, len(sites), meta["N"]))
    data = [ops.to_device(a, dtype or str(a.dtype)) for _, a, _ in sites]
    psi = MPS(data, [dm for _, _, dm in sites], meta["llim"], meta["rlim"])
    return psi, meta


def save_mpo(path, H, meta=None):
    """MPO site tensors (wl, wr, s', s)."""
    from . import ops
    arrays = [("site.%d" % j, ops.to_host(t), dm) for j, (t, dm) in
              enumerate(zip(H.data, H.dims))]
    m = dict(meta or {})
    m.update({"kind": "mpo", "N": len(H)})
    write_arrays(path, arrays, m)


def load_mpo(path, dtype=None):
    from .mps import MPO
    arrays, meta = read_arrays(path)
    if meta.get("kind") != "mpo":
        raise CheckpointError("%s holds %r, not an MPO" % (path, meta.get("kind")))
    sites = sorted(arrays, key=lambda a: int(a[0].split(".")[1]))
    return MPO([np.reshape(a, dm, order="F") for _, a, dm in sites],
               dtype or str(sites[0][1].dtype)), meta
