"""Persisted Graph format (graphgp_amd.store): a save / load round trip keeps every field of
the reference's Graph (graph.py:20-41) bit for bit, a damaged file is refused, and on the
GPU a loaded graph (with its stored dataflow schedule) generates bit-identical values."""
import os

import numpy as np
import pytest
import torch

from graphgp_amd.graph import Graph
from graphgp_amd.store import load_graph, save_graph


def _cpu_graph(n=5000, d=3, n0=100, k=8, seed=0, indices=True):
    rng = np.random.default_rng(seed)
    pts = torch.from_numpy(rng.standard_normal((n, d)).astype(np.float32))
    nb = torch.from_numpy(rng.integers(0, n0, (n - n0, k)).astype(np.int32))
    idx = torch.from_numpy(rng.permutation(n).astype(np.int32)) if indices else None
    return Graph(pts, nb, (n0, 2000, n), idx)


@pytest.mark.parametrize("indices", [True, False])
def test_round_trip_cpu(tmp_path, indices):
    g = _cpu_graph(indices=indices)
    path = str(tmp_path / "g.ggp")
    save_graph(path, g)
    h = load_graph(path, device="cpu")
    assert h.offsets == g.offsets
    assert torch.equal(h.points, g.points) and h.points.dtype == torch.float32
    assert torch.equal(h.neighbors, g.neighbors) and h.neighbors.dtype == torch.int32
    assert (h.indices is None) == (g.indices is None)
    if indices:
        assert torch.equal(h.indices, g.indices)
    with open(path, "rb") as f:
        assert f.read(8) == b"GGPGRAF1"


def test_round_trip_schedule_and_f64_cpu(tmp_path):
    g = _cpu_graph(d=2, seed=1)
    g = Graph(g.points.double(), g.neighbors, g.offsets, g.indices)
    order = torch.arange(g.neighbors.shape[0], dtype=torch.int32).flip(0)
    g._ggp_schedule = ("key", (order, g.neighbors.flip(0).contiguous(), g.points.flip(0).contiguous()))
    path = str(tmp_path / "g.ggp")
    save_graph(path, g)
    h = load_graph(path, device="cpu")
    assert h.points.dtype == torch.float64 and torch.equal(h.points, g.points)
    o, nbv, ps = h._ggp_schedule[1]
    assert torch.equal(o, order) and torch.equal(nbv, g.neighbors.flip(0)) and torch.equal(ps, g.points.flip(0))
    assert h._ggp_schedule[0][4] == g.offsets
    h2 = load_graph(path, device="cpu", schedule=False)
    assert getattr(h2, "_ggp_schedule", None) is None


def test_damaged_files_are_refused(tmp_path):
    g = _cpu_graph(seed=2)
    path = str(tmp_path / "g.ggp")
    save_graph(path, g)
    raw = bytearray(open(path, "rb").read())
    bad = str(tmp_path / "bad.ggp")
    from graphgp_amd.store import _read_header

    header, base = _read_header(path)
    raw2 = bytearray(raw)
    raw2[base + header["arrays"]["neighbors"]["offset"] + 100] ^= 0x40  # inside the neighbours
    open(bad, "wb").write(raw2)
    with pytest.raises(ValueError, match="CRC"):
        load_graph(bad, device="cpu")
    load_graph(bad, device="cpu", verify=False)  # unchecked load still reads it
    open(bad, "wb").write(raw[: len(raw) // 2])
    with pytest.raises(ValueError, match="truncated"):
        load_graph(bad, device="cpu")
    open(bad, "wb").write(b"NOTAGRAPH" + raw[9:])
    with pytest.raises(ValueError, match="not a graphgp_amd graph"):
        load_graph(bad, device="cpu")


@pytest.mark.gpu
def test_loaded_graph_generates_bit_identical(tmp_path, monkeypatch):
    import sys

    import graphgp_amd as gp
    from helpers import clustered_points

    R = sys.modules["graphgp_amd.graph"]
    monkeypatch.setattr(R, "SCHEDULE_MIN_REFINED", 0)  # a schedule for this small graph too
    p = clustered_points(120_000, d=3, seed=41)
    g = gp.build_graph(p, n0=1000, k=30)
    assert g._ggp_schedule[1] is not None
    cov = gp.extras.matern_kernel(p=1, variance=1.0, cutoff=0.05, r_min=1e-6, r_max=2, n_bins=2001, jitter=1e-5)
    xi = torch.randn(120_000, device="cuda", generator=torch.Generator(device="cuda").manual_seed(9))
    path = str(tmp_path / "g.ggp")
    gp.save_graph(path, g)
    h = gp.load_graph(path)
    assert h.points.is_cuda and h.offsets == g.offsets
    for a, b in ((h.points, g.points), (h.neighbors, g.neighbors), (h.indices, g.indices)):
        assert torch.equal(a, b)
    sched_before = h._ggp_schedule[1]
    for prec in ("fp32_increment", "fp32", "fp64"):
        a = gp.generate(g, cov, xi, precision=prec)
        b = gp.generate(h, cov, xi, precision=prec)
        assert torch.equal(a, b), prec
    assert h._ggp_schedule[1] is sched_before  # the stored schedule was used, not recomputed
    assert torch.equal(gp.generate_inv(h, cov, a, precision="fp64"), gp.generate_inv(g, cov, a, precision="fp64"))
    assert float(gp.generate_logdet(h, cov)) == float(gp.generate_logdet(g, cov))
