"""The multi-block moment exchange of the wide kernels, as a schedule model.

Rows of 512 and 1024 bins split a z range's team over several blocks
(cosenu_pair_wide in stage_pair.hip, cosenu_quad_wide in stage_quad.hip).
Every wavefront publishes its partial moments of a row into a device-memory
ring of kRing rows (slot = row mod kRing, tagged with the row), on every z
step whatever the row.  In each partner block one loader wavefront per ring
(and, in teams of 4 blocks, per partner) reads a row's slots about a z step
before its block needs the row's totals and passes the values to the block
through an LDS staging buffer of two rows.  A
reader waits until the slot carries the row's tag.  Nothing but that data
dependency (and each block's own barriers) orders the blocks, so the design
rests on two claims:

* no slot is overwritten with a later row before every partner wavefront
  that needs the earlier row has read it (a reader would otherwise wait for
  a tag that never comes back);
* the waits never form a cycle (every block finishes).

This model replays each kernel's publish / read / barrier sequence per
wavefront (the kernels' loop bounds, ring depths, barrier spacing and read
conditions), with wavefronts as independent agents that drift between their
block's barriers, under random interleavings, on grids at the edges of the
kernels' range (the shortest z ranges, ragged ranges, the open boundary's
ghost rows), and checks both claims.

The same model checks the block's own LDS moment rings (single-block teams
included) and the staging buffer: a row's values must be written and a
barrier passed before any wavefront reads them (no wait protects an LDS
read), and a slot must not be rewritten before every wavefront has read the
row it holds.  Those are what the barrier spacing (every 4th z step in the
single-block two-stage kernel, 3 of every 5 in the single-block four-stage
one, every z step in the multi-block ones) and the ring depths (16 and 8 rows) guarantee.
"""
import random

import pytest

W = 3  # wavefronts per block in the single-block models (the kernels have 8)
WQ = 4  # in the four-stage model: one staging wavefront per ring


def pair_waves(nb):
    """Wavefronts per block in the multi-block two-stage model: a loader per
    ring and partner, plus one that only consumes."""
    return 2 * (nb - 1) + 1


def quad_program(j0, jend, L, nz, openb, wide=True, wave=0):
    """Wavefront `wave`'s action list in the four-stage kernel (quad_body):
    ('pub', ring, row) / ('read', ring, row) on the device rings (wide only),
    ('lpub', ring, row) / ('lread', ring, row) on the block's LDS rings,
    ('spub', ring, row, k) / ('sread', ring, row, k) on the LDS staging
    buffer of partner k's partials (wide only; wavefront r < 4 stages ring
    r from partner 1),
    ('bar',)."""
    ibeg, iend = j0 - 6, j0 + L + 6

    def need(s, r):  # xneed, and the cells stage s + 1 computes
        m = 6 - 2 * s
        return j0 - m <= r < jend + m and not (openb and (r < 0 or r >= nz))

    def stage(row):  # the staging wavefront: check the partner's words, write the buffer
        if wide and wave < 4 and need(wave, row):
            return [("read", wave, row), ("spub", wave, row, 1)]
        return []

    prog = [("lpub", 0, ibeg), ("lpub", 0, ibeg + 1)]
    if wide:
        prog += [("pub", 0, ibeg), ("pub", 0, ibeg + 1)]
    prog += stage(ibeg - 2 * wave)
    for i in range(ibeg, iend):
        # barriers at iterations 0, 2, 4 of each group of 5; on every
        # iteration with a staging buffer
        if wide or (i - ibeg) % 5 in (0, 2, 4):
            prog.append(("bar",))
        c = [i, i - 2, i - 4, i - 6]
        # the four stages' partials from LDS (and the partner's, staged), at
        # the top of the z step
        prog += [("lread", s, c[s]) for s in range(4) if need(s, c[s])]
        if wide:
            prog += [("sread", s, c[s], 1) for s in range(4) if need(s, c[s])]
        prog.append(("lpub", 0, i + 2))
        if wide:
            prog.append(("pub", 0, i + 2))
        for s in range(3):  # stages 1-3: their outputs' partials
            if need(s, c[s]):
                prog.append(("lpub", s + 1, c[s]))
            if wide:
                prog.append(("pub", s + 1, c[s]))
        prog += stage(i + 1 - 2 * wave)  # after stage 4: the next z step's row
    return prog


def pair_program(j0, jend, L, nz, openb, fd, wide=True, wave=0, nb=2):
    """Wavefront `wave`'s action list in the two-stage kernel (pair_body),
    in a team of nb blocks: wavefront w < 2 (nb - 1) stages ring
    w // (nb - 1) from partner k = w % (nb - 1) + 1 (('read', ring, row, k)
    on partner block role ^ k's device ring, then ('spub', ring, row, k));
    every wavefront reads the staged rows it needs (('sread', ...))."""
    ibeg, iend = j0 - 5, j0 + L + 4
    last = ibeg + 2 * ((iend - ibeg + 1) // 2) if fd else iend  # FD: pairs of iterations
    nx = nb - 1
    lrg, lk = wave // nx, wave % nx + 1
    loader = wide and wave < 2 * nx

    def need_a(r):  # the first stage's cells (and xneed_a)
        return j0 - 4 <= r < jend + 4 and not (openb and (r < 0 or r >= nz))

    def need_b(r):
        return j0 <= r < jend

    prog = [("lpub", 0, ibeg + 1), ("lpub", 0, ibeg + 2), ("lpub", 0, ibeg + 3)]
    if wide:
        prog += [("pub", 0, ibeg + 1), ("pub", 0, ibeg + 2), ("pub", 0, ibeg + 3)]
    for i in range(ibeg, last):
        prog.append(("lpub", 0, i + 4))
        if wide:
            prog.append(("pub", 0, i + 4))
        if wide or ((i - ibeg) & 3) == 0:  # kSync = 4 in single-block teams
            prog.append(("bar",))
        jb = i - 4
        # both stages' partials from LDS, right after the barrier, and the
        # partners' staged partials of the first stage's row
        if need_a(i):
            prog.append(("lread", 0, i))
        if need_b(jb):
            prog.append(("lread", 1, jb))
        if wide and need_a(i):
            prog += [("sread", 0, i, k) for k in range(1, nb)]
        if need_a(i):
            prog.append(("lpub", 1, i))
        if wide:
            prog.append(("pub", 1, i))
        if wide and need_b(jb):
            prog += [("sread", 1, jb, k) for k in range(1, nb)]
        # the loader: check the next iteration's row, stage it
        if loader:
            row = i + 1 if lrg == 0 else i - 3
            if (need_a if lrg == 0 else need_b)(row):
                prog += [("read", lrg, row, lk), ("spub", lrg, row, lk)]
    return prog


def run_team(programs, nb, kring, rng, dev_kring=None):
    """Interleave the agents (block, wavefront) of `programs` at random, with
    LDS rings of kring rows and device rings of dev_kring rows (default:
    kring).  Returns None when every agent finished, else a description of
    the failure."""
    dev_kring = dev_kring or kring
    agents = sorted(programs)
    waves = {b: [w for (bb, w) in agents if bb == b] for b in range(nb)}
    pc = {a: 0 for a in agents}
    bar_count = {a: 0 for a in agents}        # barriers passed
    slot = {}                                 # (block, wave, ring, row % kring) -> row
    lds = {}                                  # (block, wave, ring, row % kring) -> (row, barriers)
    stg = {}                                  # (block, ring, k, row % 2) -> (row, barriers)
    # LDS reads still to come, per (kind, block, ring, row): how many wavefronts will read it
    pending = {}
    for a in agents:
        for op in programs[a]:
            if op[0] in ("lread", "sread"):
                key = (op[0], a[0]) + op[1:]
                pending[key] = pending.get(key, 0) + 1
    while True:
        runnable = []
        for a in agents:
            prog = programs[a]
            if pc[a] >= len(prog):
                continue
            op = prog[pc[a]]
            b, _ = a
            if op[0] == "bar":
                # passes once every wavefront of its block has reached this barrier
                if all(bar_count[(b, v)] > bar_count[a] or
                       (pc[(b, v)] < len(programs[(b, v)]) and
                        programs[(b, v)][pc[(b, v)]][0] == "bar" and
                        bar_count[(b, v)] == bar_count[a]) for v in waves[b]):
                    runnable.append(a)
            elif op[0] == "read":
                ring, row = op[1], op[2]
                ok = True
                # ('read', ring, row, k): partner block b ^ k only
                for pb in (range(nb) if len(op) == 3 else (b ^ op[3],)):
                    if pb == b:
                        continue
                    for v in waves[pb]:
                        have = slot.get((pb, v, ring, row % dev_kring))
                        if have is not None and have > row:
                            return (f"block {pb} wave {v} overwrote ring {ring} row {row} "
                                    f"with row {have} before block {b} read it")
                        if have != row:
                            ok = False
                if ok:
                    runnable.append(a)
            else:
                runnable.append(a)
        if not runnable:
            if all(pc[a] >= len(programs[a]) for a in agents):
                return None
            stuck = [(a, programs[a][pc[a]]) for a in agents if pc[a] < len(programs[a])]
            return f"deadlock: {stuck[:4]}"
        a = rng.choice(runnable)
        op = programs[a][pc[a]]
        b, w = a
        if op[0] == "bar":
            bar_count[a] += 1
        elif op[0] == "pub":
            _, ring, row = op
            slot[(b, w, ring, row % dev_kring)] = row
        elif op[0] == "lpub":
            _, ring, row = op
            old = lds.get((b, w, ring, row % kring))
            if old is not None and pending.get(("lread", b, ring, old[0]), 0) > 0:
                return (f"block {b} wave {w} rewrote LDS ring {ring} row {old[0]} with row {row} "
                        f"before every wavefront read it")
            lds[(b, w, ring, row % kring)] = (row, bar_count[a])
        elif op[0] == "spub":
            _, ring, row, k = op
            old = stg.get((b, ring, k, row % 2))
            if old is not None and pending.get(("sread", b, ring, old[0], k), 0) > 0:
                return (f"block {b} wave {w} rewrote staged ring {ring} row {old[0]} with row "
                        f"{row} before every wavefront read it")
            stg[(b, ring, k, row % 2)] = (row, bar_count[a])
        elif op[0] == "sread":
            _, ring, row, k = op
            have = stg.get((b, ring, k, row % 2))
            if have is None or have[0] != row:
                return f"block {b} wave {w} read staged ring {ring} row {row} holding {have}"
            if have[1] >= bar_count[a]:
                return (f"block {b} wave {w} read staged ring {ring} row {row} with no barrier "
                        f"after it was written")
            pending[("sread", b, ring, row, k)] -= 1
        elif op[0] == "lread":
            _, ring, row = op
            for v in waves[b]:
                have = lds.get((b, v, ring, row % kring))
                if have is None or have[0] != row:
                    return (f"block {b} wave {w} read LDS ring {ring} row {row} holding "
                            f"{have and have[0]} (wave {v})")
                if have[1] >= bar_count[a]:
                    return (f"block {b} wave {w} read LDS ring {ring} row {row} with no barrier "
                            f"after wave {v} wrote it")
            pending[("lread", b, ring, row)] -= 1
        pc[a] += 1


def team_ranges(nz, nteams, sample=False):
    """The teams' z ranges (j0, jend, L); sample: the first three, the last
    three and two in between (the open boundary's ghost rows lie in the
    first and last teams; the others differ only by a shift)."""
    ts = range(nteams)
    if sample and nteams > 8:
        ts = sorted({0, 1, 2, nteams // 3, 2 * nteams // 3, nteams - 3, nteams - 2, nteams - 1})
    for t in ts:
        yield t * nz // nteams, (t + 1) * nz // nteams, -(-nz // nteams)


GRIDS = [(16, 8), (16, 3), (100, 50), (100, 7), (64, 32), (200, 9)]  # (nz, teams)


@pytest.mark.parametrize("openb", [False, True])
@pytest.mark.parametrize("nz,nteams", GRIDS)
def test_quad_exchange_never_overwrites_or_deadlocks(nz, nteams, openb):
    rng = random.Random(nz * 131 + nteams * 7 + openb)
    for j0, jend, L in team_ranges(nz, nteams):
        programs = {(b, w): quad_program(j0, jend, L, nz, openb, wave=w)
                    for b in range(2) for w in range(WQ)}
        for _ in range(4):
            assert run_team(programs, 2, 8, rng) is None, (nz, nteams, j0, jend)


@pytest.mark.parametrize("fd", [False, True])
@pytest.mark.parametrize("openb", [False, True])
@pytest.mark.parametrize("nz,nteams", GRIDS)
def test_pair_exchange_never_overwrites_or_deadlocks(nz, nteams, openb, fd):
    rng = random.Random(nz * 131 + nteams * 7 + openb * 3 + fd)
    for nb in (2, 4):
        for j0, jend, L in team_ranges(nz, nteams, sample=nb == 4):
            programs = {(b, w): pair_program(j0, jend, L, nz, openb, fd, wave=w, nb=nb)
                        for b in range(nb) for w in range(pair_waves(nb))}
            for _ in range(2):
                assert run_team(programs, nb, 16, rng) is None, (nb, nz, nteams, j0, jend)


@pytest.mark.parametrize("openb", [False, True])
@pytest.mark.parametrize("nz,nteams", [(16, 8), (100, 7), (64, 32), (300, 2)])
def test_single_block_lds_rings_are_race_free(nz, nteams, openb):
    """Single-block teams (rows of at most 256 bins): the LDS rings alone."""
    rng = random.Random(nz * 17 + nteams + openb)
    for j0, jend, L in team_ranges(nz, nteams):
        for prog, kring in ((quad_program(j0, jend, L, nz, openb, wide=False), 8),
                            (pair_program(j0, jend, L, nz, openb, True, wide=False), 16),
                            (pair_program(j0, jend, L, nz, openb, False, wide=False), 16)):
            programs = {(0, w): prog for w in range(W)}
            for _ in range(3):
                assert run_team(programs, 1, kring, rng) is None, (nz, nteams, j0, jend)


def test_model_detects_a_missing_barrier():
    """The LDS checks are not vacuous: without the barrier at iteration 4 of
    each group of 5, the four-stage schedule lets a wavefront read a row's
    partials before a barrier separates them from their writes."""
    rng = random.Random(7)
    prog = quad_program(0, 64, 64, 64, False, wide=False)
    # drop every third barrier (those at iteration 4 of a group)
    thinned, n = [], 0
    for op in prog:
        if op == ("bar",):
            n += 1
            if n % 3 == 0:
                continue
        thinned.append(op)
    programs = {(0, w): thinned for w in range(W)}
    failures = [run_team(programs, 1, 8, rng) for _ in range(10)]
    assert any(f is not None and "no barrier" in f for f in failures)


@pytest.mark.parametrize("kind", ["quad", "pair"])
def test_model_detects_a_too_shallow_device_ring(kind):
    """The model is not vacuous: with 2-row device rings (the LDS rings as
    built) a block overwrites a slot its partner still needs."""
    rng = random.Random(5)
    if kind == "quad":
        programs = {(b, w): quad_program(0, 64, 64, 64, False, wave=w) for b in range(2)
                    for w in range(WQ)}
        kring = 8
    else:
        programs = {(b, w): pair_program(0, 64, 64, 64, False, False, wave=w)
                    for b in range(2) for w in range(pair_waves(2))}
        kring = 16
    failures = [run_team(programs, 2, kring, rng, dev_kring=2) for _ in range(20)]
    assert any(f is not None and "overwrote" in f for f in failures)


def test_model_detects_a_wait_that_never_ends():
    """...and a read of a row no partner ever publishes (beyond the last row
    its ring receives) is reported as a deadlock."""
    rng = random.Random(6)
    programs = {}
    for b in range(2):
        for w in range(WQ):
            prog = quad_program(0, 64, 64, 64, False, wave=w)
            programs[(b, w)] = [("read", op[1], 200) if op[0] == "read" and op[2] == 5 else op
                                for op in prog]
    f = run_team(programs, 2, 8, rng)
    assert f is not None and f.startswith("deadlock")
