"""SIRT checkpoint and resume.

The state of a SIRT run is (x_k, k): the iterate after k iterations and the
count.  A checkpoint is two files next to each other:

  <stem>.raw   float32 volume [nz][ns][nt], t fastest (the library's layout)
  <stem>.json  header: format version, k, shape, dtype, the geometry and run
               parameters the caller wants to keep (lambda, bounds, seed ...)

Resuming passes x_k as the initial volume (x0) of the next SIRT call and
continues the count from k; SIRT is a fixed-point iteration, so n + m
iterations equal n iterations, a checkpoint, and m more from it.
"""
import json
import os

import numpy as np

FORMAT = "ctr-sirt-state/1"


def save_sirt_state(stem, x, k, shape, meta=None):
    """Write x (any array of nz*ns*nt float32 values, host or CUDA tensor) and its header."""
    nz, ns, nt = (int(v) for v in shape)
    if hasattr(x, "detach"):   # a torch tensor, possibly on the GPU
        x = x.detach().cpu().numpy()
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(-1))
    if a.size != nz * ns * nt:
        raise ValueError(f"volume has {a.size} values, shape {shape} needs {nz * ns * nt}")
    if k < 0:
        raise ValueError("iteration count must be non-negative")
    header = {"format": FORMAT, "k": int(k), "shape": [nz, ns, nt], "dtype": "float32", "order": "z,s,t (t fastest)",
              "meta": meta or {}}
    tmp = stem + ".raw.tmp"
    a.tofile(tmp)
    os.replace(tmp, stem + ".raw")
    with open(stem + ".json.tmp", "w") as f:
        json.dump(header, f, indent=1)
    os.replace(stem + ".json.tmp", stem + ".json")


def load_sirt_state(stem):
    """(x as float32 [nz*ns*nt], k, header) from a checkpoint written by save_sirt_state."""
    with open(stem + ".json") as f:
        header = json.load(f)
    if header.get("format") != FORMAT:
        raise ValueError(f"{stem}.json: not a {FORMAT} header")
    nz, ns, nt = header["shape"]
    x = np.fromfile(stem + ".raw", dtype="<f4")
    if x.size != nz * ns * nt:
        raise ValueError(f"{stem}.raw: {x.size} values, header says {nz}x{ns}x{nt}")
    return x, int(header["k"]), header
