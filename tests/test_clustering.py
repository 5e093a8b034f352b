"""Clustering of flagged cells into patches (Mesh::splitRectangle and its
helpers, Mesh.cpp:386-792), pinned by layouts worked out by hand.

Each case is a small flag set chosen to drive one branch of the reference's
recursion; the expected boxes were derived on paper by following
Mesh.cpp:644-792 step by step (the derivation is written beside each case).
Both the oracle (oracle/oracle_amr.cpp) and the host library that builds the
GPU hierarchies (veritas_amd/csrc/host/amr.cpp, an independent rewrite) must
give exactly these boxes, in this order.  The cases also run with x and p
exchanged, which drives the other axis's branch (for the inflection tie-break
that is the p-branch with its 32767 start distance, Mesh.cpp:552), except the
square bisection, whose mirror image is derived as a case of its own (equal
sides always bisect in p)."""
import ctypes

import numpy as np
import pytest

import veritas_amd as va
from _oracle import lib


def _call(f, box, flags, eff):
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                       ctypes.c_int]
    a = np.ascontiguousarray(np.array(flags, dtype=np.int32).reshape(-1, 2))
    n = f(*box, a.ctypes.data, len(a), eff, None, 0)
    out = np.zeros(max(n, 1), dtype=np.int32)
    assert f(*box, a.ctypes.data, len(a), eff, out.ctypes.data, n) == n
    return [tuple(int(v) for v in out[k:k + 4]) for k in range(0, n, 4)]


def oracle_split(box, flags, eff):
    return _call(lib().vxo_split, box, flags, eff)


def host_split(box, flags, eff):
    return _call(va.host_lib().vxh_split_box, box, flags, eff)


def cells(x0, x1, p0, p1):
    return [(x, p) for x in range(x0, x1 + 1) for p in range(p0, p1 + 1)]


CASES = {
    # 14 of 16 cells flagged: 0.875 > 0.6, the box is accepted whole (Mesh.cpp:650-660)
    "efficiency": ((0, 0, 3, 3), [c for c in cells(0, 3, 0, 3) if c not in ((0, 0), (3, 3))], 0.6,
                   [(0, 0, 3, 3)]),
    # 10/16 < 0.7; sigX = [2,2,0,0,0,2,2,2]: first zero at 2 -> start 1, first
    # non-zero after it at 5 -> finish 5; no p hole; cut in x (Mesh.cpp:699-704):
    # (0,0)-(1,1) and (5,0)-(7,1), each full
    "x_hole": ((0, 0, 7, 1), cells(0, 1, 0, 1) + cells(5, 7, 0, 1), 0.7,
               [(0, 0, 1, 1), (5, 0, 7, 1)]),
    # holes in both: sigX = [2,2,0,0,0,0,0,2,2,2] -> (1, 7), length 6;
    # sigP = [2,2,0,0,3,3] -> (1, 4), length 3; 6 > 3 cuts x (Mesh.cpp:681-697).
    # (0,0)-(1,5): 4/12; sigP = [2,2,0,0,0,0] -> start 1, the zeros run to the
    # end -> finish 5: (0,0)-(1,1) kept, (0,5)-(1,5) has no flags.
    # (7,0)-(9,5): 6/18; sigP = [0,0,0,0,3,3], the hole starts the signature ->
    # start 0, finish 4: (7,0)-(9,0) has no flags, (7,4)-(9,5) kept
    "both_holes": ((0, 0, 9, 5), cells(0, 1, 0, 1) + cells(7, 9, 4, 5), 0.7,
                   [(0, 0, 1, 1), (7, 4, 9, 5)]),
    # no hole, 12/16 < 0.8.  sigX = [2,2,2,2,1,1,1,1], delSigX = [0,0,-1,1,0,0]:
    # one sign change at index 2 (size 2); delSigP is empty (one p edge):
    # cut x at 0 + 2 + 1 = 3 (Mesh.cpp:582-590, 715-720).  (0,0)-(3,1) full;
    # (4,0)-(7,1): 4/8, sigP = [4,0] -> p hole (0, 1): (4,0)-(7,0) kept
    "inflection": ((0, 0, 7, 1), cells(0, 3, 0, 1) + cells(4, 7, 0, 0), 0.8,
                   [(0, 0, 3, 1), (4, 0, 7, 0)]),
    # 16/20 < 0.85.  sigX = [2,2,2,2,1,1,1,1,2,2], delSigX =
    # [0,0,-1,1,0,0,1,-1]: sign changes at 2 and 6, both of size 2; the centre
    # is 8/2 = 4, both are 2 away and the strict comparison keeps the first
    # (Mesh.cpp:480-490): cut x at 3.  (0,0)-(3,1) full.  (4,0)-(9,1): 8/12,
    # sigX = [1,1,1,1,2,2], delSigX = [0,0,1,-1]: one change at 2 -> cut at 7;
    # (4,0)-(7,1): 4/8, sigP = [4,0] -> (4,0)-(7,0); (8,0)-(9,1) full
    "inflection_tie": ((0, 0, 9, 1), cells(0, 3, 0, 1) + cells(4, 7, 0, 0) + cells(8, 9, 0, 1), 0.85,
                       [(0, 0, 3, 1), (4, 0, 7, 0), (8, 0, 9, 1)]),
    # 16/24 < 0.7, columns of height [3,3,1,1,3,3,1,1], sigP = [8,4,4] (no
    # hole).  delSigX = [-2,2,2,-2,-2,2]: sign changes at 0, 2 and 4, all of
    # size 4; the centre is 6/2 = 3, distances 3, 1, 1: index 2 is the first
    # at distance 1 and a later equal distance does not replace it -> cut x at
    # 3.  (0,0)-(3,2) and (4,0)-(7,2): 8/12 each, delSigX of length 2 and
    # delSigP of length 1 -> too small to cut, kept.  (Taking index 4 instead
    # would keep (0,0)-(5,2), 14/18, and (6,0)-(7,0).)
    "inflection_tie_three": ((0, 0, 7, 2), [(x, p) for x, h in enumerate([3, 3, 1, 1, 3, 3, 1, 1])
                                             for p in range(h)], 0.7, [(0, 0, 3, 2), (4, 0, 7, 2)]),
    # the triangle x >= p in 6 x 5: 20/30 < 0.7, no hole.  sigX = [1,2,3,4,5,5]
    # gives delSigX = [0,0,0,-1] (no pair of non-zero entries), sigP =
    # [6,5,4,3,2] gives delSigP = [0,0,0]: both longer than 2 without an
    # inflection -> bisect the long side (Mesh.cpp:722-741): lengthx 5 > 4, cut
    # at floor(5/2) = 2.  (0,0)-(2,4): 6/15, sigP = [3,2,1,0,0] -> p hole
    # (2, 4): (0,0)-(2,2), 6/9 < 0.7, signatures of length 3 give one-entry
    # Laplacians -> too small to cut, kept (Mesh.cpp:591-594, 743-748).
    # (3,0)-(5,4): 14/15 kept
    "bisect": ((0, 0, 5, 4), [(x, p) for x in range(6) for p in range(5) if x >= p], 0.7,
               [(0, 0, 2, 2), (3, 0, 5, 4)]),
    # two diagonals (i, i) and (i, i+2 mod 5) in 5 x 5: every row and column
    # holds 2, 10/25 < 0.6; flat signatures, bisect; lengths equal -> cut p at 2.
    # (0,0)-(4,2): 6 flags, sigX = [2,1,1,1,1] -> delSigX = [1,0,0] has no
    # inflection (bisect flag) but delSigP has one entry -> no inflection in
    # either and not both bisecting -> kept (Mesh.cpp:591-594).
    # (0,3)-(4,4): 4/10, sigX = [0,1,1,1,1] -> x hole (0, 1): (1,3)-(4,4),
    # 4/8, signatures too short -> kept
    "bisect_square": ((0, 0, 4, 4), [(i, i) for i in range(5)] + [(i, (i + 2) % 5) for i in range(5)], 0.6,
                      [(0, 0, 4, 2), (1, 3, 4, 4)]),
}

# Equal side lengths bisect in p whichever axis the pattern leans on, so the
# mirror image of "bisect_square" is its own case, derived the same way:
# bisect p at 2.  (0,0)-(4,2): 6/15, sigX = [1,1,2,1,1], delSigX = [1,-2,1]:
# two changes of size 3 at 0 and 1, centre 3/2 = 1 -> index 1, cut x at 2;
# (0,0)-(2,2): 4/9, one-entry Laplacians -> kept; (3,0)-(4,2): sigP = [0,1,1]
# -> p hole (0, 1): (3,1)-(4,2), 2/4, kept.  (0,3)-(4,4): sigX = [1,1,0,1,1]
# -> x hole (1, 3): (0,3)-(1,4) and (3,3)-(4,4), 2/4 each, kept.
UNMIRRORED = {
    "bisect_square_mirror": ((0, 0, 4, 4), [(i, i) for i in range(5)] + [((i + 2) % 5, i) for i in range(5)],
                             0.6, [(0, 0, 2, 2), (3, 1, 4, 2), (0, 3, 1, 4), (3, 3, 4, 4)]),
}


def transpose(case):
    box, flags, eff, want = case
    return ((box[1], box[0], box[3], box[2]), [(p, x) for x, p in flags], eff,
            [(b[1], b[0], b[3], b[2]) for b in want])


def all_cases():
    for name, case in CASES.items():
        yield pytest.param(case, id=name)
        if name != "bisect_square":
            yield pytest.param(transpose(case), id=name + "_transposed")
    for name, case in UNMIRRORED.items():
        yield pytest.param(case, id=name)


def check_oracle_cases():
    """Every hand-derived case against the oracle (the pin that
    tests/test_oracle_plants.py runs on planted builds)."""
    for p in all_cases():
        box, flags, eff, want = p.values[0]
        assert oracle_split(box, flags, eff) == want, p.id


@pytest.mark.parametrize("case", list(all_cases()))
def test_oracle_clustering_gives_the_hand_derived_boxes(case):
    box, flags, eff, want = case
    assert oracle_split(box, flags, eff) == want


@pytest.mark.parametrize("case", list(all_cases()))
def test_host_clustering_gives_the_hand_derived_boxes(case):
    box, flags, eff, want = case
    assert host_split(box, flags, eff) == want


def test_host_and_oracle_clustering_agree_on_random_flags():
    """Beyond the hand-derived cases: random flag sets on random boxes."""
    rng = np.random.default_rng(3)
    for _ in range(400):
        w, h = int(rng.integers(1, 40)), int(rng.integers(1, 40))
        x0, p0 = int(rng.integers(0, 50)), int(rng.integers(0, 50))
        n = int(rng.integers(1, w * h + 1))
        idx = rng.choice(w * h, size=n, replace=False)
        flags = [(x0 + int(k) // h, p0 + int(k) % h) for k in idx]
        box = (min(f[0] for f in flags), min(f[1] for f in flags), max(f[0] for f in flags),
               max(f[1] for f in flags))
        eff = float(rng.uniform(0.3, 0.95))
        assert host_split(box, flags, eff) == oracle_split(box, flags, eff)
