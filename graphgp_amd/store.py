"""Persisted ``Graph`` format: a graph (and its dataflow schedule) saved to one file and
loaded back into HBM, so that a serving process or a restart does not rebuild it.

The reference's ``Graph`` (graph.py:20-41) is an in-memory pytree of ``points``,
``neighbors``, the static ``offsets`` tuple and the optional ``indices``; it has no file
format.  This one stores exactly those fields, plus the optional schedule arrays that
``generate`` keeps on large graphs.

Layout (little-endian):

    bytes 0..7     magic  b"GGPGRAF1"
    bytes 8..15    header length H (uint64)
    bytes 16..     H bytes of UTF-8 JSON: format version, n, d, k, n0, the offsets tuple,
                   and per array its dtype, shape, byte offset, byte length and CRC-32
    then           the arrays, each starting on a 4096-byte boundary (raw, C order)

Arrays: ``points`` (n, d) float32/float64, ``neighbors`` (n - n0, k) int32, ``indices``
(n,) int32 if present; with the schedule: ``schedule`` (n - n0,) int32,
``schedule_neighbors`` (n - n0, k) int32 and ``schedule_points`` (n, d).  Arrays move
between the device and the file in chunks through one pinned host buffer, so a graph
larger than host memory can be saved and loaded; ``load_graph`` maps the file and checks
every array's CRC-32 unless ``verify=False``.
"""
from __future__ import annotations

import json
import os
import struct
import zlib
from typing import Optional

import numpy as np
import torch

from .graph import Graph

MAGIC = b"GGPGRAF1"
VERSION = 1
ALIGN = 4096
CHUNK = 64 << 20

_DTYPES = {"float32": (torch.float32, np.float32), "float64": (torch.float64, np.float64),
           "int32": (torch.int32, np.int32)}


def _dtype_name(t: torch.Tensor) -> str:
    for name, (td, _) in _DTYPES.items():
        if t.dtype == td:
            return name
    raise TypeError(f"unsupported dtype {t.dtype} in a graph")


def _staging(nbytes: int, pin: bool) -> torch.Tensor:
    return torch.empty(min(CHUNK, max(nbytes, 1)), dtype=torch.uint8, pin_memory=pin)


def _write_tensor(f, t: torch.Tensor) -> int:
    """Write t's bytes at the file's position in chunks; returns their CRC-32."""
    flat = t.detach().contiguous().view(-1).view(torch.uint8)
    nbytes = flat.numel()
    crc = 0
    if flat.device.type == "cpu":
        buf = flat.numpy()
        for s in range(0, nbytes, CHUNK):
            b = buf[s:s + CHUNK]
            crc = zlib.crc32(b, crc)
            f.write(memoryview(b))
        return crc
    stage = _staging(nbytes, True)
    for s in range(0, nbytes, CHUNK):
        e = min(nbytes, s + CHUNK)
        h = stage[:e - s]
        h.copy_(flat[s:e])  # synchronous device-to-host copy into pinned memory
        b = h.numpy()
        crc = zlib.crc32(b, crc)
        f.write(memoryview(b))
    return crc


def save_graph(path: str, graph: Graph, *, schedule: bool = True) -> None:
    """Save ``graph`` to ``path`` (overwritten).  With ``schedule`` the dataflow schedule
    that ``generate`` keeps on the graph is stored too, when it has one."""
    pts, nb = graph.points, graph.neighbors
    if pts.ndim != 2 or nb.ndim != 2:
        raise ValueError("graph.points must be (N, d) and graph.neighbors (N - n0, k)")
    n, d = pts.shape
    n_ref, k = nb.shape
    offsets = tuple(int(x) for x in graph.offsets)
    arrays = {"points": pts, "neighbors": nb}
    if graph.indices is not None:
        arrays["indices"] = graph.indices
    sched = getattr(graph, "_ggp_schedule", None)
    if schedule and sched is not None and sched[1] is not None:
        order, nbv, pts_s = sched[1]
        arrays.update(schedule=order, schedule_neighbors=nbv, schedule_points=pts_s)
    meta = {}
    pos = 0
    for name, t in arrays.items():
        nbytes = t.numel() * t.element_size()
        meta[name] = {"dtype": _dtype_name(t), "shape": list(t.shape), "offset": pos, "nbytes": nbytes, "crc32": 0}
        pos += (nbytes + ALIGN - 1) // ALIGN * ALIGN
    header = {"format": "graphgp_amd graph", "version": VERSION, "n": n, "d": d, "k": k, "n0": n - n_ref,
              "offsets": list(offsets), "arrays": meta}

    def header_bytes():
        h = json.dumps(header, sort_keys=True).encode()
        return h + b" " * (-(len(MAGIC) + 8 + len(h)) % ALIGN)

    # the CRCs are known only after writing: reserve the header's space (CRCs as 10-digit
    # fields), write the arrays, then rewrite the header in place
    for m in meta.values():
        m["crc32"] = 0xFFFFFFFF
    hb = header_bytes()
    base = len(MAGIC) + 8 + len(hb)
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(MAGIC + struct.pack("<Q", len(hb)) + hb)
        for name, t in arrays.items():
            f.seek(base + meta[name]["offset"])
            meta[name]["crc32"] = _write_tensor(f, t)
        f.truncate(base + pos)
        hb2 = json.dumps(header, sort_keys=True).encode()
        if len(hb2) > len(hb):
            raise RuntimeError("graph header grew while writing")
        f.seek(len(MAGIC) + 8)
        f.write(hb2 + b" " * (len(hb) - len(hb2)))
    os.replace(tmp, path)


def _read_header(path: str):
    with open(path, "rb") as f:
        head = f.read(len(MAGIC) + 8)
        if len(head) < len(MAGIC) + 8 or head[:len(MAGIC)] != MAGIC:
            raise ValueError(f"{path}: not a graphgp_amd graph file")
        (hlen,) = struct.unpack("<Q", head[len(MAGIC):])
        header = json.loads(f.read(hlen).decode())
    if header.get("version") != VERSION:
        raise ValueError(f"{path}: unsupported graph file version {header.get('version')}")
    return header, len(MAGIC) + 8 + hlen


def load_graph(path: str, *, device: Optional[torch.device] = None, verify: bool = True,
               schedule: bool = True) -> Graph:
    """Load a graph saved by ``save_graph`` onto ``device`` (default: the current GPU).
    The stored schedule, if any, is attached so that ``generate`` uses it at once."""
    header, base = _read_header(path)
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device()) if torch.cuda.is_available() else torch.device("cpu")
    device = torch.device(device)
    size = os.path.getsize(path)
    mm = np.memmap(path, dtype=np.uint8, mode="r")
    out = {}
    for name, m in header["arrays"].items():
        if not schedule and name.startswith("schedule"):
            continue
        td, nd = _DTYPES[m["dtype"]]
        start, nbytes = base + m["offset"], m["nbytes"]
        if start + nbytes > size:
            raise ValueError(f"{path}: array {name} is truncated")
        dst = torch.empty(m["shape"], dtype=td, device=device)
        flat = dst.view(-1).view(torch.uint8)
        crc = 0
        stage = _staging(nbytes, device.type == "cuda")
        for s in range(0, nbytes, CHUNK):
            e = min(nbytes, s + CHUNK)
            src = mm[start + s:start + e]
            if verify:
                crc = zlib.crc32(src, crc)
            if device.type == "cpu":
                flat[s:e].numpy()[:] = src
            else:
                h = stage[:e - s]
                h.numpy()[:] = src
                flat[s:e].copy_(h)  # synchronous: the staging buffer is reused next chunk
        if verify and crc != m["crc32"]:
            raise ValueError(f"{path}: array {name} fails its CRC-32 check (corrupted file)")
        out[name] = dst
    del mm
    n, d, k, n0 = header["n"], header["d"], header["k"], header["n0"]
    if tuple(out["points"].shape) != (n, d) or tuple(out["neighbors"].shape) != (n - n0, k):
        raise ValueError(f"{path}: array shapes do not match the header")
    g = Graph(out["points"], out["neighbors"], tuple(int(x) for x in header["offsets"]), out.get("indices"))
    if "schedule" in out:
        pts, nb = g.points, g.neighbors
        key = (pts.data_ptr(), pts._version, nb.data_ptr(), nb._version, g.offsets)
        g._ggp_schedule = (key, (out["schedule"], out["schedule_neighbors"], out["schedule_points"]))
    return g
