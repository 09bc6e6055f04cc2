"""Velocity snapshots: the file bytes are the memory of the reference's saved
(n, n, n, 3) Float64 array (warmup.jl:86-88: save_object(file, u.data |>
Array)), and device fields make the round trip bit for bit."""
import json

import numpy as np
import pytest
import torch


def test_file_is_the_julia_array_memory(ec, tmp_path):
    """The .bin holds Array{Float64,4}(n, n, n, 3) in column-major order: byte
    offset 8 * (i + n (j + n (k + n a))) holds u_a[i, j, k] (0-based)."""
    S = ec.snapshot
    n = 6
    rng = np.random.default_rng(3)
    comps = [np.asfortranarray(rng.standard_normal((n, n, n))) for _ in range(3)]
    doc = S.save_arrays(str(tmp_path / "u"), comps, l=2.0, t=0.5)
    raw = np.fromfile(tmp_path / "u.bin", dtype="<f8")
    julia = np.stack(comps, axis=3)  # (n, n, n, 3) [i, j, k, a]
    assert np.array_equal(raw, julia.ravel(order="F"))
    for (i, j, k, a) in [(0, 0, 0, 0), (1, 0, 0, 0), (0, 1, 0, 1), (2, 3, 4, 2), (5, 5, 5, 2)]:
        assert raw[i + n * (j + n * (k + n * a))] == comps[a][i, j, k]
    side = json.load(open(tmp_path / "u.json"))
    assert side["shape"] == [n, n, n, 3] and side["order"] == "F" and side["dtype"] == "float64"
    assert side["l"] == 2.0 and side["t"] == 0.5 and doc == side
    back, _ = S.load_arrays(str(tmp_path / "u"))
    for x, y in zip(back, comps):
        assert np.array_equal(x, y)


def test_bad_files_raise(ec, tmp_path):
    S = ec.snapshot
    n = 4
    comps = [np.zeros((n, n, n)) for _ in range(3)]
    S.save_arrays(str(tmp_path / "v"), comps)
    with open(tmp_path / "v.bin", "ab") as f:
        f.write(b"\0" * 8)
    with pytest.raises(ValueError):
        S.load_arrays(str(tmp_path / "v"))
    json.dump({"format": "other"}, open(tmp_path / "w.json", "w"))
    with pytest.raises(ValueError):
        S.read_sidecar(str(tmp_path / "w"))


@pytest.mark.gpu
@pytest.mark.parametrize("n,dtype,fdt", [(16, torch.float64, "float64"), (48, torch.float32, "float64"),
                                         (40, torch.float32, "float32")])
def test_device_round_trip(gpu, tmp_path, n, dtype, fdt):
    """Device fields -> file -> device, bit for bit (FP32 fields widen exactly
    to the reference's Float64 and come back); the file is the host-side
    column-major array of the same values; slabs smaller than n cover the
    whole field."""
    S = gpu.snapshot
    NS = gpu.navierstokes
    g = NS.Grid(1.0, n)
    gen = torch.Generator(device="cuda").manual_seed(n)
    u = tuple(NS.Field(g, torch.randn((n, n, n), device="cuda", generator=gen, dtype=dtype)) for _ in range(3))
    S.save_velocity(str(tmp_path / "s"), u, dtype=fdt, slab=7, visc=1e-3)
    host, doc = S.load_arrays(str(tmp_path / "s"))
    assert doc["l"] == 1.0 and doc["visc"] == 1e-3 and doc["dtype"] == fdt
    for x, h in zip(u, host):
        assert np.array_equal(x.numpy(), h.astype(np.float64))
    back, _ = S.load_velocity(str(tmp_path / "s"), dtype=dtype, slab=5)
    for x, y in zip(u, back):
        assert torch.equal(x.data, y)
