"""Host emulation of the dataflow kernel's protocol (apply_kernel in
graphgp_amd/csrc/refine_impl.hpp), run on the CPU under adversarial scheduling.

What is emulated, step for step as the kernel does it:
  * persistent workgroups of 4 wavefronts x 64 / P points (P = 4 lanes per point up to
    k = 16, 8 beyond, as the launch chooses), XCD groups that take blocks of consecutive
    tiles (32 in the kernel; 1, 2 and 32 here) from one counter per group, workgroup g in
    group g % groups, a tile's ticket taken when the workgroup's previous tile is finished;
  * per lane: rows j = P r + lane, rows j >= k masked (never pending, zero weight and
    value), a pending bit per row; each poll round reads every pending row with a plain
    read, and re-reads it coherently when the plain read returns the sentinel;
  * a group (point) becomes ready when the ballot shows all P of its lanes with no pending
    row; it publishes once;
  * memory: a plain read of a published slot may still return the sentinel (a stale line
    in the reader's L2), a coherent read returns the sentinel until the store has become
    visible, which takes a random number of scheduler steps;
  * limited residency: only some workgroups are resident; the others start when a resident
    one exits.  As on the device, workgroups start in index order and at least `groups` of
    them are resident (the launch caps the grid to the resident capacity under a CU budget,
    and one index per group comes first), so every group keeps a running workgroup.
The scheduler picks a random resident wavefront to advance by one step.  Checked: every
run ends within a bounded number of steps (no deadlock), and the published values equal
the sequential evaluation of the same recurrence (v_i = v_0 + w_i0 v_0 + sum_{j>0} w_ij
(u_j - v_0) + sx_i, the kernel's sum in increments of neighbour 0).
"""
import numpy as np
import pytest

SENT = None  # the unpublished sentinel
WAVES = 4  # wavefronts per workgroup (of 64 lanes: 64 / P points each, P lanes per point)


def make_graph(rng, n0, nref, k, batch):
    """Refined point i (value slot n0 + i) depends on k distinct earlier slots, all in
    earlier batches (a batched dependency structure like the graph's offsets)."""
    nbrs = np.empty((nref, k), np.int64)
    for i in range(nref):
        b = i // batch
        hi = n0 + b * batch  # every slot below the start of this batch is earlier
        nbrs[i] = rng.choice(hi, k, replace=False)
    return nbrs


def reference(vals0, nbrs, w, sx, n0):
    v = list(vals0) + [0.0] * len(nbrs)
    for i, row in enumerate(nbrs):
        v0 = v[row[0]]
        v[n0 + i] = v0 + w[i, 0] * v0 + sum(w[i, j] * (v[row[j]] - v0) for j in range(1, len(row))) + sx[i]
    return np.array(v)


class Memory:
    def __init__(self, vals0, nref, rng, delay_max):
        self.true = list(vals0) + [SENT] * nref
        self.visible_at = [0] * len(self.true)
        self.stale = set()
        self.rng = rng
        self.delay_max = delay_max
        self.clock = 0

    def publish(self, slot, v):
        assert self.true[slot] is SENT, "a slot was published twice"
        self.true[slot] = v
        self.visible_at[slot] = self.clock + int(self.rng.integers(0, self.delay_max + 1))
        if self.rng.random() < 0.5:
            self.stale.add(slot)  # some reader's L2 keeps the old line for plain reads

    def plain(self, slot):
        if slot in self.stale or self.clock < self.visible_at[slot]:
            return SENT
        return self.true[slot]

    def coherent(self, slot):
        return SENT if self.clock < self.visible_at[slot] else self.true[slot]


class Wave:
    def __init__(self, wg, wave, tile, m, a):
        self.points = []
        P = a["P"]
        ppw = 64 // P
        for g in range(ppw):
            q = tile * WAVES * ppw + wave * ppw + g
            valid = q < a["nref"]
            qc = q if valid else a["nref"] - 1
            lanes = []
            for lane in range(P):
                rows = []
                for r in range(a["RM"]):
                    j = P * r + lane
                    live = valid and j < a["k"]
                    rows.append({"j": j, "nb": a["nbrs"][qc, j] if live else 0,
                                 "w": a["w"][qc, j] if live else 0.0, "u": 0.0, "pend": live})
                lanes.append(rows)
            self.points.append({"q": q, "valid": valid, "lanes": lanes, "done": not valid})
        self.polls = 0

    def step(self, m, a, out):
        """One pass of the poll loop: reads of the pending rows, readiness by ballot, publish."""
        for pnt in self.points:
            if pnt["done"]:
                continue
            for rows in pnt["lanes"]:
                for row in rows:
                    if row["pend"]:
                        v = m.plain(row["nb"])
                        if v is SENT:
                            v = m.coherent(row["nb"])
                        if v is not SENT:
                            row["u"], row["pend"] = v, False
        for pnt in self.points:
            if pnt["done"]:
                continue
            if all(not row["pend"] for rows in pnt["lanes"] for row in rows):
                u0 = pnt["lanes"][0][0]["u"]  # neighbour 0 is lane 0's first row
                part = 0.0
                for rows in pnt["lanes"]:
                    for row in rows:
                        uj = row["u"] - u0 if row["j"] != 0 else row["u"]
                        part += row["w"] * uj
                v = u0 + part + a["sx"][pnt["q"]]
                m.publish(a["n0"] + pnt["q"], v)
                out[a["n0"] + pnt["q"]] = v
                pnt["done"] = True
        self.polls += 1
        return all(p["done"] for p in self.points)


class Workgroup:
    def __init__(self, b, a, counters):
        self.b = b
        self.group = b % a["groups"]
        self.a = a
        self.counters = counters
        self.tile = None
        self.waves = None

    def take(self):
        t = self.counters[self.group]
        self.counters[self.group] += 1
        return t

    def tile_of(self, t):
        a = self.a
        gt = a["group_tiles"]
        return ((t // gt) * a["groups"] + self.group) * gt + t % gt

    def step(self, m, out, rng):
        """Advance one random wavefront; returns False when the workgroup has exited."""
        a = self.a
        if self.waves is None:
            self.tile = self.tile_of(self.take())  # thread 0's ticket, then the barrier
            if self.tile >= a["ntiles"]:
                return False
            self.waves = [Wave(self, w, self.tile, m, a) for w in range(WAVES)]
            self.finished = [False] * WAVES
            return True
        live = [i for i in range(WAVES) if not self.finished[i]]
        i = live[int(rng.integers(len(live)))]
        self.finished[i] = self.waves[i].step(m, a, out)
        if all(self.finished):  # every wavefront left the poll loop: next ticket
            self.waves = None
        return True


@pytest.mark.parametrize("seed,k,batch,groups,resident,delay,group_tiles", [
    (0, 5, 40, 8, 8, 3, 2), (1, 7, 17, 3, 3, 10, 1), (2, 30, 100, 8, 16, 5, 2), (3, 3, 1, 2, 2, 0, 1),
    (4, 9, 64, 5, 40, 20, 2), (5, 16, 1, 1, 1, 4, 32), (6, 30, 100, 8, 8, 5, 32), (7, 7, 3, 3, 4, 8, 2)])
def test_dataflow_protocol_completes_and_matches_sequential(seed, k, batch, groups, resident, delay, group_tiles):
    rng = np.random.default_rng(seed)
    n0, nref = max(k, 8), 700
    nbrs = make_graph(rng, n0, nref, k, batch)
    w = rng.standard_normal((nref, k)) / k
    sx = rng.standard_normal(nref)
    vals0 = rng.standard_normal(n0)
    P = 8 if k > 16 else 4  # lanes per point, as the launch chooses them
    rm = -(-k // P)  # rows per lane
    gpt = WAVES * 64 // P
    ntiles = (nref + gpt - 1) // gpt
    grid = min(ntiles, 3 * groups + 1)
    a = dict(nbrs=nbrs, w=w, sx=sx, n0=n0, nref=nref, k=k, P=P, RM=rm, ntiles=ntiles, groups=min(groups, grid),
             group_tiles=group_tiles)
    m = Memory(vals0, nref, rng, delay)
    out = {}
    counters = [0] * a["groups"]
    pending = list(range(grid))
    live = []
    steps = 0
    while pending or live:
        while pending and len(live) < resident:
            live.append(Workgroup(pending.pop(0), a, counters))
        i = int(rng.integers(len(live)))
        if not live[i].step(m, out, rng):
            live.pop(i)
        m.clock += 1
        steps += 1
        assert steps < 2_000_000, "the emulated dataflow did not finish"
    ref = reference(vals0, nbrs, w, sx, n0)
    got = np.array([out[n0 + i] for i in range(nref)])
    assert np.allclose(got, ref[n0:], rtol=1e-12, atol=1e-12)
