"""Velocity snapshots on disk: the array the reference saves, in raw binary.

The reference writes the final DNS velocity with ``save_object(file, u.data |>
Array)`` (warmup.jl:86-88) and reads it back with ``load_object``
(dns_aid.jl:40-42): an ``Array{Float64,4}`` of size (n, n, n, 3), column-major,
x fastest, the component last.  JLD2 (HDF5) has no reader in this
environment, so a snapshot here is that array's memory as raw little-endian
bytes in ``<path>.bin`` plus a JSON sidecar ``<path>.json`` with the shape,
element type and grid.  In Julia the same data is

    u = Array{Float64}(undef, n, n, n, 3); read!("<path>.bin", u)

Fields move between HBM and the file in slabs of z-planes, so host memory
stays at one slab per direction whatever the grid size.
"""
from __future__ import annotations

import json
import os

import numpy as np

FORMAT = "exactclosure-velocity-v1"
_NP = {"float64": np.float64, "float32": np.float32}


def _paths(path):
    base = path[:-4] if path.endswith(".bin") else path
    return base + ".bin", base + ".json"


def write_sidecar(path, n, dtype="float64", **meta):
    """The JSON description of a (n, n, n, 3) snapshot at ``path``."""
    _, js = _paths(path)
    doc = {"format": FORMAT, "shape": [n, n, n, 3], "order": "F", "dtype": dtype,
           "byteorder": "little", "components": "last"}
    doc.update(meta)
    with open(js, "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
    return doc


def read_sidecar(path):
    _, js = _paths(path)
    with open(js) as f:
        doc = json.load(f)
    if doc.get("format") != FORMAT:
        raise ValueError(f"{js}: not a {FORMAT} snapshot")
    n = doc["shape"][0]
    if doc["shape"] != [n, n, n, 3] or doc.get("order") != "F" or doc.get("dtype") not in _NP:
        raise ValueError(f"{js}: unsupported layout {doc.get('shape')} {doc.get('order')} {doc.get('dtype')}")
    return doc


def save_arrays(path, comps, dtype="float64", **meta):
    """Host arrays (three Julia-ordered [i, j, k] arrays) to a snapshot."""
    bin_, _ = _paths(path)
    n = comps[0].shape[0]
    le = "<" + np.dtype(_NP[dtype]).str[1:]
    with open(bin_, "wb") as f:
        for x in comps:
            assert x.shape == (n, n, n)
            # column-major bytes of the component: x fastest, then y, then z
            np.asfortranarray(x, dtype=_NP[dtype]).ravel(order="F").astype(le, copy=False).tofile(f)
    return write_sidecar(path, n, dtype, **meta)


def load_arrays(path):
    """A snapshot as three Julia-ordered [i, j, k] host arrays (Fortran order)."""
    doc = read_sidecar(path)
    n = doc["shape"][0]
    bin_, _ = _paths(path)
    raw = np.fromfile(bin_, dtype="<" + np.dtype(_NP[doc["dtype"]]).str[1:])
    if raw.size != 3 * n ** 3:
        raise ValueError(f"{bin_}: {raw.size} values, expected {3 * n ** 3}")
    return [raw[a * n ** 3:(a + 1) * n ** 3].reshape((n, n, n), order="F") for a in range(3)], doc


def save_velocity(path, u, dtype="float64", slab=64, **meta):
    """Save a device velocity (three torch tensors [k, j, i] or Fields) to
    ``path``; ``dtype`` is the file's element type (the reference's is
    float64).  ``meta`` goes to the sidecar (e.g. l, t, visc)."""
    bin_, _ = _paths(path)
    data = [getattr(x, "data", x) for x in u]
    n = data[0].shape[0]
    if hasattr(u[0], "grid"):
        meta.setdefault("l", u[0].grid.l)
    with open(bin_, "wb") as f:
        for t in data:
            assert tuple(t.shape) == (n, n, n) and t.is_contiguous()
            for k0 in range(0, n, slab):
                # [k, j, i] slabs in k order are the column-major bytes of the component
                f.write(t[k0:k0 + slab].to("cpu", dtype=_torch_dtype(dtype)).numpy().tobytes())
    return write_sidecar(path, n, dtype, **meta)


def load_velocity(path, dtype=None, slab=64):
    """Load a snapshot into three new device tensors [k, j, i] (``dtype``:
    torch dtype, default the file's); returns (tensors, sidecar)."""
    import torch

    doc = read_sidecar(path)
    n = doc["shape"][0]
    fdt = _NP[doc["dtype"]]
    tdt = dtype or _torch_dtype(doc["dtype"])
    bin_, _ = _paths(path)
    item = np.dtype(fdt).itemsize
    if os.path.getsize(bin_) != 3 * n ** 3 * item:
        raise ValueError(f"{bin_}: size {os.path.getsize(bin_)} != {3 * n ** 3 * item}")
    out = [torch.empty((n, n, n), dtype=tdt, device="cuda") for _ in range(3)]
    with open(bin_, "rb") as f:
        for t in out:
            for k0 in range(0, n, slab):
                kc = min(slab, n - k0)
                buf = np.fromfile(f, dtype="<" + np.dtype(fdt).str[1:], count=kc * n * n)
                t[k0:k0 + kc].copy_(torch.from_numpy(buf.reshape(kc, n, n)).to(tdt))
    return out, doc


def _torch_dtype(name):
    import torch

    return {"float64": torch.float64, "float32": torch.float32}[name]
