"""The dependency-driven sweep's work lists (mx_sweep_describe, host only)
and a simulation of its scheduling protocol, on the CPU.

The lists are checked against what the structure fixes, recomputed here
from the tasks' waits and signals rather than read from the builder:
* every counter a task waits on is signalled by exactly one task, and the
  wait's target equals the number of (tile, wave) signals that task makes
  per column tile, so a consumer never reads a block before its last write;
* every producer tile ranks below its consumer, and each tile's per-queue
  "last producer" index is the largest index of its producers in that queue;
* every task of the fused schedule appears once per row tile.

The simulation runs the claim/help/wait/signal protocol of x3_sweep_kernel
(mx_bf16x3.hip) with workgroups interleaved at random, a limited number of
resident workgroups, and two sweeps competing for the same slots; it must
always finish, run every work item once, and never pass a wait early.
"""
import collections

import numpy as np
import pytest

from conftest import make_problem
from matrox_ru_amd import hgemm

NQ, WAVES = 8, 4

CASES = {
    "h2b": dict(n=2048, d=8, leaf=64, rank=48, tol=1e-6, budget=0.2, h=1.0),
    "h2b_far": dict(n=3000, d=4, leaf=48, rank=24, tol=1e-6, budget=0.05, h=1.2, seed=3),
    "hss": dict(n=1500, d=6, leaf=50, rank=30, tol=1e-6, budget=0.0, h=0.9, seed=5),
    # ranks and leaves past 128 rows: several row tiles per task
    "h2b_tall": dict(n=4096, d=6, leaf=256, rank=160, tol=1e-10, budget=0.1, h=1.0, seed=7),
}


def _describe(case, direction):
    X, hm = make_problem(**CASES[case])
    return hm, hgemm.sweep_describe(hm, direction)


def _queues(d):
    return [d["tiles"][d["qinfo"][z, 0]:d["qinfo"][z, 0] + d["qinfo"][z, 1]] for z in range(NQ)]


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("direction", ["forward", "adjoint"])
def test_sweep_lists_consistent(case, direction):
    hm, d = _describe(case, direction)
    tasks, waits = d["tasks"], d["waits"]
    qs = _queues(d)
    # the writer of every signalled block, and its row tiles
    writer = {}
    for t, row in enumerate(tasks):
        if row[5] >= 0:
            assert row[5] not in writer, "two tasks signal one block"
            writer[row[5]] = t
    rowtiles = collections.Counter()
    where = {}
    for z, q in enumerate(qs):
        assert (np.diff(q[:, 2]) > 0).all(), "a queue out of rank order"
        for j, tl in enumerate(q):
            key = (int(tl[0]), int(tl[1]))
            assert key not in where, "a tile in two places"
            where[key] = (z, j, int(tl[2]))
            rowtiles[int(tl[0])] += 1
    for t, n in rowtiles.items():
        assert sorted(r for (tt, r) in where if tt == t) == list(range(n))
    for t, row in enumerate(tasks):
        for wi in range(row[0], row[3]):
            b, target = waits[wi]
            assert b in writer, "a wait on a block nobody writes"
            assert target == WAVES * rowtiles[writer[b]]
        # early waits precede late waits, and a late segment exists iff late waits do
        assert row[0] <= row[1] == row[2] <= row[3]
        assert (row[4] >= 0) == (row[3] > row[2])
    for z, q in enumerate(qs):
        for j, tl in enumerate(q):
            row = tasks[tl[0]]
            expect = [-1] * NQ
            for wi in range(row[0], row[3]):
                p = writer[waits[wi][0]]
                for r in range(rowtiles[p]):
                    pz, pj, pg = where[(p, r)]
                    assert pg < tl[2], "a producer ranked after its consumer"
                    expect[pz] = max(expect[pz], pj)
            assert list(tl[4:12]) == expect
    # every non-empty output of the structure is covered: skeleton blocks of
    # nodes with a rank (twice: w^ and u^ for H2-b; u^ may be absent at the
    # top) and every leaf block of Y
    nleaf = int(hm.scalars["nleaf"])
    y_tasks = [t for t in rowtiles if tasks[t][5] < 0]
    assert len(y_tasks) == sum(1 for l in range(nleaf) if hm.arrays["Ddim"][l] > 0)


class _Sweep:
    """One sweep's shared state: queue heads and counters."""

    def __init__(self, d, ncols):
        self.d, self.ncols = d, ncols
        self.q = _queues(d)
        self.len = [len(q) * ncols for q in self.q]
        self.heads = [0] * NQ
        self.cnt = collections.Counter()
        self.done = set()
        self.passed = []   # (block, ct, target) of every wait passed
        self.total = sum(self.len)

    def item(self, z, pos):
        return self.q[z][pos // self.ncols], pos % self.ncols

    def rank(self, z):
        h = self.heads[z]
        if h >= self.len[z]:
            return None
        tl, ct = self.item(z, h)
        return int(tl[2]) * self.ncols + ct

    def claim(self, home):
        """sweep_claim: fetch-and-add on the home queue's head, or once it
        is exhausted on the head of the queue whose next item ranks lowest."""
        cand = home
        while True:
            got = self.heads[cand]
            self.heads[cand] += 1
            if got < self.len[cand]:
                return cand, got
            ranks = [(self.rank(z), z) for z in range(NQ) if self.rank(z) is not None]
            if not ranks:
                return None
            cand = min(ranks)[1]

    def producers_claimed(self, got):
        tl, ct = self.item(*got)
        return all(m < 0 or self.heads[y] > m * self.ncols + ct for y, m in enumerate(tl[4:12]))

    def claim_lowest(self, limit):
        """sweep_claim_lowest: the untaken item of lowest rank, if it ranks
        below `limit` (the held item's rank)."""
        ranks = [(self.rank(z), z) for z in range(NQ) if self.rank(z) is not None]
        if not ranks or (limit is not None and min(ranks)[0] >= limit):
            return None
        z = min(ranks)[1]
        self.heads[z] += 1
        return z, self.heads[z] - 1

    def ready(self, b, e, ct):
        w = self.d["waits"]
        return all(self.cnt[(int(w[i][0]), ct)] >= w[i][1] for i in range(b, e))

    def note_pass(self, b, e, ct):
        for i in range(b, e):
            self.passed.append((int(self.d["waits"][i][0]), ct, int(self.d["waits"][i][1])))


class _WG:
    """A workgroup: idle -> early wait -> late wait -> signal -> idle."""

    bounded = True   # sweep_claim_lowest's rank bound (a planted mistake drops it)

    def __init__(self, sw, home):
        self.sw, self.home, self.state, self.cur = sw, home, "idle", None
        self.held, self.fresh = None, True

    def start(self, got):
        if got is None:
            self.state = "exit"
            return
        tl, ct = self.sw.item(*got)
        self.cur = (got, tl, ct)
        self.state = "early"

    def step(self):
        """Advance one step; False when blocked.  Claiming, checking the
        held item's producers and claiming a helper are separate steps, as
        in x3_sweep_kernel's wave 0, so other workgroups act in between."""
        sw = self.sw
        if self.state == "idle":
            if self.fresh:
                self.held = sw.claim(self.home)
            self.fresh = True
            if self.held is None:
                self.start(None)
            elif sw.producers_claimed(self.held):
                self.start(self.held)
            else:
                self.state = "help"
            return True
        if self.state == "help":
            tl, ct = sw.item(*self.held)
            got = sw.claim_lowest(int(tl[2]) * sw.ncols + ct if self.bounded else None)
            if got is not None:
                self.fresh = False
                self.start(got)
            else:
                self.start(self.held)
            return True
        got, tl, ct = self.cur
        tk = sw.d["tasks"][tl[0]]
        if self.state == "early":
            if not sw.ready(tk[0], tk[1], ct):
                return False
            sw.note_pass(tk[0], tk[1], ct)
            self.state = "late"
            return True
        if self.state == "late":
            if tk[4] >= 0:
                if not sw.ready(tk[2], tk[3], ct):
                    return False
                sw.note_pass(tk[2], tk[3], ct)
            self.state = "signal"
            return True
        if self.state == "signal":
            if tk[5] >= 0:
                sw.cnt[(int(tk[5]), ct)] += WAVES
            assert got not in sw.done, "an item ran twice"
            sw.done.add(got)
            self.state = "idle"
            return True
        return False


def _simulate(sweeps, slots, rng, wg=None):
    """Run the sweeps' workgroups (64 each, homes round robin) on `slots`
    resident slots; a workgroup holds its slot until it exits."""
    wg = wg or _WG
    pending = [wg(sw, k % NQ) for sw in sweeps for k in range(64)]
    rng.shuffle(pending)
    live = []
    steps = 0
    while True:
        while len(live) < slots and pending:
            live.append(pending.pop())
        moved = False
        for i in rng.permutation(len(live)):
            wg = live[i]
            if wg.step():
                moved = True
                steps += 1
                break
        live = [w for w in live if w.state != "exit"]
        if not live and not pending:
            break
        assert moved, "deadlock: every resident workgroup is waiting"
    for sw in sweeps:
        assert len(sw.done) == sw.total
        for b, ct, target in sw.passed:   # a wait passes only at the final count
            assert sw.cnt[(b, ct)] == target
    return steps


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("direction", ["forward", "adjoint"])
def test_sweep_protocol_finishes(case, direction):
    hm, d = _describe(case, direction)
    rng = np.random.default_rng(11)
    for slots, ncols in ((1, 1), (3, 2), (16, 3), (128, 1), (128, 2), (200, 2)):
        _simulate([_Sweep(d, ncols)], slots, rng)


def test_two_sweeps_share_slots():
    """Two sweeps (say two plans on two streams) competing for few slots."""
    _, d1 = _describe("h2b", "forward")
    _, d2 = _describe("hss", "adjoint")
    rng = np.random.default_rng(12)
    for slots in (1, 2, 5, 40):
        _simulate([_Sweep(d1, 2), _Sweep(d2, 1)], slots, rng)


def test_claim_rule_is_needed():
    """Planted mistake: running a claimed item without first helping its
    unclaimed producers deadlocks with few resident slots."""
    _, d = _describe("h2b", "forward")

    class Greedy(_Sweep):
        def producers_claimed(self, got):
            return True

    rng = np.random.default_rng(13)
    hit = False
    for _ in range(20):
        try:
            _simulate([Greedy(d, 2)], 2, rng)
        except AssertionError as e:
            assert "deadlock" in str(e)
            hit = True
            break
    assert hit


def test_helper_bound_is_needed():
    """Planted mistake: without the rank bound on helpers, a workgroup can
    claim a helper that depends on the item it holds (its held item's
    producer having been claimed by another workgroup between the check and
    the helper's claim), and with few slots that deadlocks."""
    _, d = _describe("h2b_tall", "forward")

    class Unbounded(_WG):
        bounded = False

    rng = np.random.default_rng(14)
    hit = False
    for _ in range(20):
        try:
            _simulate([_Sweep(d, 1)], 128, rng, wg=Unbounded)
        except AssertionError as e:
            assert "deadlock" in str(e)
            hit = True
            break
    assert hit
