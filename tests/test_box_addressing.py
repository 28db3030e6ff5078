"""Host replay of the box DILU sweep's FULL-group addressing (csrc/dilu.hip,
boxSweep): in a group of 16 steps where every in-tile row is valid and every
prefetch stays inside the sweep, each step addresses its streams through one
scalar offset per array plus per-lane constants.  For every such step of many
tile grids (every tile width, both sweep directions) the replay checks that

- the FULL condition makes every in-tile row valid;
- the records / input of step i + 4, the inbox row of step i + 3 and the
  output row of step t are the rows the general loop body addresses
  (its clamped step indices), at non-negative offsets inside the sweep;
- every value a tile delivers lands in its consumer's inbox, inside the
  consumer's steps, at the consumer's step that holds the same cell row.

The tile grid follows layout.cpp boxWaves (steps = nA + actB - 1 + actC - 1).
"""
import itertools
import random

import pytest

kBoxAhead, kBoxInAhead, kGroup = 4, 3, 16
kBoxAheadLone = 8   # lone sweeps (no inbox) prefetch records further ahead


def tangential(a):
    return {0: (1, 2), 1: (0, 2), 2: (0, 1)}[a]


def tile_grid(n, TB):
    a = 0
    for q in (1, 2):
        if n[q] > n[a]:
            a = q
    b, c = tangential(a)
    TC = 64 // TB
    nTB, nTC = -(-n[b] // TB), -(-n[c] // TC)
    tiles = {}
    for tc in range(nTC):
        for tb in range(nTB):
            actB, actC = min(TB, n[b] - tb * TB), min(TC, n[c] - tc * TC)
            tiles[(tb, tc)] = dict(nA=n[a], actB=actB, actC=actC, steps=n[a] + actB - 1 + actC - 1)
    return tiles, TC


def replay(n, TB, fwd):
    tiles, TC = tile_grid(n, TB)
    full_groups = 0
    for (tb, tc), m in tiles.items():
        nA, actB, actC, nSteps = m["nA"], m["actB"], m["actC"], m["steps"]
        toB = tiles.get((tb + 1, tc) if fwd else (tb - 1, tc))
        toC = tiles.get((tb, tc + 1) if fwd else (tb, tc - 1))
        dB, dC = (1 - TB, 1 - TC) if fwd else (TB - 1, TC - 1)
        dA, dI = (kBoxAhead, kBoxInAhead) if fwd else (-kBoxAhead, -kBoxInAhead)
        dBe, dCe = (dB if toB else 0), (dC if toC else 0)
        mOff = min(min(dI, 0), min(dBe, dCe))
        tLo = actB - 1 + actC - 1
        for i0 in range(0, nSteps, kGroup):
            tMin = i0 if fwd else nSteps - kGroup - i0
            tMax = i0 + kGroup - 1 if fwd else nSteps - 1 - i0
            if not (tMin >= tLo and tMax <= nA - 1 and i0 + kGroup - 1 + kBoxAhead < nSteps):
                continue
            full_groups += 1
            for i in range(i0, i0 + kGroup):
                t = i if fwd else nSteps - 1 - i
                step_of = lambda k: (min(k, nSteps - 1) if fwd else nSteps - 1 - min(k, nSteps - 1))
                # records and input: scalar offset of step tm, lane constant dA (fwd)
                tm = t if fwd else t + dA
                assert tm >= 0 and tm + (dA if fwd else 0) == step_of(i + kBoxAhead)
                # output of step t: the same scalar offset, lane constant -dA (bwd)
                assert tm + (0 if fwd else -dA) == t and 0 <= t < nSteps
                # inbox row of step i + 3: scalar offset t + mOff, lane constant dI - mOff
                assert t + mOff >= 0 and dI - mOff >= 0
                assert (t + mOff) + (dI - mOff) == step_of(i + kBoxInAhead)
                for bl, cl in itertools.product(range(TB), range(TC)):
                    if not (bl < actB and cl < actC):
                        continue
                    ia = t - bl - cl
                    assert 0 <= ia < nA
                    send_b = (bl == TB - 1 if fwd else bl == 0) and toB is not None
                    send_c = (cl == TC - 1 if fwd else cl == 0) and toC is not None
                    if send_b:
                        assert dB - mOff >= 0
                        row = (t + mOff) + (dB - mOff)
                        assert 0 <= row < toB["steps"]
                        assert row - (0 if fwd else TB - 1) - cl == ia and cl < toB["actC"]
                    if send_c:
                        assert dC - mOff >= 0
                        row = (t + mOff) + (dC - mOff)
                        assert 0 <= row < toC["steps"]
                        assert row - bl - (0 if fwd else TC - 1) == ia and bl < toC["actB"]
    return full_groups


def segment(nA, s, segs):
    """dilu.hpp boxSegment: rows [lo, hi) of segment s of segs."""
    return s * nA // segs, (s + 1) * nA // segs


def replay_segments(n, TB, fwd, segs):
    """The lone (tile-local) sweep of each segment of each tile: its steps
    tFirst..tLast, the FULL condition and the addresses of a FULL step (lone
    sweeps prefetch kBoxAheadLone steps ahead)."""
    tiles, TC = tile_grid(n, TB)
    full_groups = 0
    for m in tiles.values():
        nA, actB, actC, nSteps = m["nA"], m["actB"], m["actC"], m["steps"]
        tLo = actB - 1 + actC - 1
        dA = kBoxAheadLone if fwd else -kBoxAheadLone
        covered = set()
        for sg in range(segs):
            lo, hi = segment(nA, sg, segs)
            tFirst, tLast = lo, hi - 1 + tLo
            nLocal = tLast - tFirst + 1
            # every step of the segment is a step of the tile
            assert nLocal <= 0 or (0 <= tFirst and tLast < nSteps)
            # each row of the tile falls in exactly one segment's steps
            for ia in range(lo, hi):
                assert ia not in covered
                covered.add(ia)
            for i0 in range(0, max(nLocal, 0), kGroup):
                tMin = tFirst + i0 if fwd else tLast - (i0 + kGroup - 1)
                tMax = tFirst + i0 + kGroup - 1 if fwd else tLast - i0
                if not (tMin >= lo + tLo and tMax <= hi - 1 and i0 + kGroup - 1 + kBoxAheadLone < nLocal):
                    continue
                full_groups += 1
                for i in range(i0, i0 + kGroup):
                    t = tFirst + i if fwd else tLast - i
                    # records / input of the step kBoxAheadLone later, inside the segment's steps
                    j = t + dA
                    assert tFirst <= j <= tLast and (t if fwd else j) >= 0
                    for bl, cl in itertools.product(range(actB), range(actC)):
                        assert lo <= t - bl - cl < hi
        assert covered == set(range(nA))
    return full_groups


def shapes():
    rng = random.Random(7)
    out = [(160, 80, 80), (200, 100, 19), (37, 23, 29), (41, 27, 7), (12, 14, 40), (5, 5, 5), (64, 1, 1),
           (3, 7, 200), (100, 2, 3)]
    out += [tuple(rng.randint(1, 60) for _ in range(3)) for _ in range(8)]
    return out


@pytest.mark.parametrize("TB", [1, 2, 4, 8, 16, 32, 64])
def test_full_group_addressing(TB):
    groups = 0
    for n in shapes():
        for fwd in (True, False):
            groups += replay(n, TB, fwd)
    assert groups > 0


def test_duct_fluid_tiles_run_mostly_full_groups():
    tiles, _ = tile_grid((160, 80, 80), 8)
    per_tile = replay((160, 80, 80), 8, True) / len(tiles)
    assert per_tile == 9   # of 11 groups (174 steps)


@pytest.mark.parametrize("segs", [2, 3])
def test_segmented_lone_sweep_addressing(segs):
    groups = 0
    for n in shapes():
        for TB in (1, 4, 8, 16, 64):
            for fwd in (True, False):
                groups += replay_segments(n, TB, fwd, segs)
    assert groups > 0
