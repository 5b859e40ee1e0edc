"""SIRT checkpoint and resume.

The state of a SIRT run is (x_k, k): the iterate after k iterations and the
count.  A checkpoint is a header and the volume it names:

  <stem>.json      header: format version, k, shape, dtype, the volume's file
                   name, byte count and CRC-32, and the geometry and run
                   parameters the caller wants to keep (lambda, bounds, seed ...)
  <stem>.k<k>.raw  float32 volume [nz][ns][nt], t fastest (the library's layout)

The header is the commit point.  A save writes and fsyncs the new volume
under its own name, then replaces the header atomically, and only then
removes the volume the old header named.  A crash at any point leaves a
header whose volume is complete: the old state or the new one, never the
new volume under the old count.  Loading checks the named volume's size and
CRC-32 and raises on a mismatch.

Resuming passes x_k as the initial volume (x0) of the next SIRT call and
continues the count from k; SIRT is a fixed-point iteration, so n + m
iterations equal n iterations, a checkpoint, and m more from it.
"""
import json
import os
import zlib

import numpy as np

FORMAT = "ctr-sirt-state/2"


def _fsync_dir(path):
    try:
        fd = os.open(os.path.dirname(os.path.abspath(path)) or ".", os.O_RDONLY)
    except OSError:
        return
    try:
        os.fsync(fd)
    except OSError:
        pass
    finally:
        os.close(fd)


def _write_durable(path, data):
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(data)
        f.flush()
        os.fsync(f.fileno())
    os.replace(tmp, path)
    _fsync_dir(path)


def _volume_name(stem, k):
    return f"{os.path.basename(stem)}.k{int(k)}.raw"


def write_volume(stem, a, k):
    """Write the volume file of state k (the first half of a save); returns
    (file name, byte count, CRC-32)."""
    data = np.ascontiguousarray(a, dtype="<f4").tobytes()
    name = _volume_name(stem, k)
    _write_durable(os.path.join(os.path.dirname(os.path.abspath(stem)), name), data)
    return name, len(data), zlib.crc32(data) & 0xFFFFFFFF


def _read_header(stem):
    with open(stem + ".json") as f:
        header = json.load(f)
    if header.get("format") != FORMAT:
        raise ValueError(f"{stem}.json: not a {FORMAT} header")
    return header


def save_sirt_state(stem, x, k, shape, meta=None):
    """Write x (any array of nz*ns*nt float32 values, host or CUDA tensor) and its header."""
    nz, ns, nt = (int(v) for v in shape)
    if hasattr(x, "detach"):   # a torch tensor, possibly on the GPU
        x = x.detach().cpu().numpy()
    a = np.asarray(x, dtype=np.float32).reshape(-1)
    if a.size != nz * ns * nt:
        raise ValueError(f"volume has {a.size} values, shape {shape} needs {nz * ns * nt}")
    if k < 0:
        raise ValueError("iteration count must be non-negative")
    old = None
    if os.path.exists(stem + ".json"):
        try:
            old = _read_header(stem).get("volume")
        except (ValueError, OSError, json.JSONDecodeError):
            old = None
    name, nbytes, crc = write_volume(stem, a, k)
    header = {"format": FORMAT, "k": int(k), "shape": [nz, ns, nt], "dtype": "float32", "order": "z,s,t (t fastest)",
              "volume": name, "bytes": nbytes, "crc32": crc, "meta": meta or {}}
    _write_durable(stem + ".json", json.dumps(header, indent=1).encode())
    if old and old != name:
        try:
            os.remove(os.path.join(os.path.dirname(os.path.abspath(stem)), old))
        except OSError:
            pass


def load_sirt_state(stem):
    """(x as float32 [nz*ns*nt], k, header) from a checkpoint written by save_sirt_state."""
    header = _read_header(stem)
    nz, ns, nt = header["shape"]
    path = os.path.join(os.path.dirname(os.path.abspath(stem)), header["volume"])
    with open(path, "rb") as f:
        data = f.read()
    if len(data) != header["bytes"] or len(data) != 4 * nz * ns * nt:
        raise ValueError(f"{path}: {len(data)} bytes, header says {header['bytes']} ({nz}x{ns}x{nt} float32)")
    if zlib.crc32(data) & 0xFFFFFFFF != header["crc32"]:
        raise ValueError(f"{path}: CRC-32 does not match the header")
    x = np.frombuffer(data, dtype="<f4").astype(np.float32)
    return x, int(header["k"]), header
