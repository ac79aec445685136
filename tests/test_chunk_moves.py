"""The chunk step's move lists (the cycles of a chunk-matrix transpose).

The reference walks the cycles of its permutation on the host
(src/permute.cu:18-49); the library builds the lists on the device
(ipt_block_cycle.hip, chunk_moves_device) and keeps the host walk behind
ipt_debug_chunk_moves as a cross-check.  Both builders' lists are checked
against what the chunk kernels rely on, not against each other's layout:

- items partition the moves, each 1..ell moves long;
- a saved head appears only as the first move of an item;
- every move sends position x to P(x) = (x mod cc) * rc + x / cc, and the
  destinations are every position P moves, once;
- an item reads only positions it also writes (so items need no ordering
  once the heads are saved);
- applying the lists to labelled chunks gives the transposed chunk matrix.

Both builders must also agree on the number of moves and saved heads (one
head per segment of a cycle longer than ell).  Plain loops over small cases
would retype the builder; these checks only use P.
"""
import numpy as np
import pytest

FLAG = np.uint32(0x80000000)


def check_lists(rc, cc, ell, items, msrc, mdst, hpos):
    n = rc * cc
    q = np.arange(n, dtype=np.int64)
    perm = (q % cc) * rc + q // cc
    moving = np.nonzero(perm != q)[0]
    nm = len(msrc)
    assert nm == len(mdst) == len(moving)
    if nm == 0:
        assert len(items) == 0 and len(hpos) == 0
        return
    starts, lens = items[:, 0].astype(np.int64), items[:, 1].astype(np.int64)
    order = np.argsort(starts)
    starts, lens = starts[order], lens[order]
    assert starts[0] == 0 and np.all(starts[1:] == starts[:-1] + lens[:-1])
    assert starts[-1] + lens[-1] == nm
    assert lens.min() >= 1 and lens.max() <= ell
    item_of = np.repeat(np.arange(len(starts)), lens)  # item of each move

    flagged = (msrc & FLAG) != 0
    first = np.zeros(nm, dtype=bool)
    first[starts] = True
    assert not np.any(flagged & ~first), "a saved head is not an item's first move"
    slot = (msrc[flagged] & ~FLAG).astype(np.int64)
    assert np.array_equal(np.sort(slot), np.arange(len(hpos))), "each head used once"
    src = msrc.astype(np.int64)
    src[flagged] = hpos[slot]
    assert np.array_equal(mdst.astype(np.int64), perm[src]), "a move does not follow P"
    assert np.array_equal(np.sort(mdst.astype(np.int64)), moving)

    writer = np.full(n, -1, dtype=np.int64)
    writer[mdst] = item_of
    reads = ~flagged
    assert np.array_equal(writer[src[reads]], item_of[reads]), "an item reads outside its writes"

    data = q * 7 + 3
    out = data.copy()
    out[mdst] = data[src]
    exp = np.empty(n, dtype=np.int64)
    exp[perm] = data
    assert np.array_equal(out, exp)


def cases():
    fixed = [(2, 2, 1), (2, 3, 1), (3, 2, 2), (2, 5, 3), (7, 11, 128), (13, 17, 4), (4, 4, 8),
             (64, 64, 2), (5, 7, 1), (2, 1000, 128), (1000, 2, 5), (101, 103, 7), (3, 3, 3),
             (32, 4096, 16), (7, 65536, 32), (1024, 1024, 128), (255, 257, 9)]
    rng = np.random.default_rng(11)
    rand = [(int(rng.integers(2, 400)), int(rng.integers(2, 400)),
             int(rng.choice([1, 2, 3, 4, 7, 16, 33, 64, 128]))) for _ in range(40)]
    return fixed + rand


@pytest.mark.gpu
def test_device_and_host_move_lists_are_valid(ipt):
    for rc, cc, ell in cases():
        dev = ipt.chunk_moves(rc, cc, ell, device=True)
        host = ipt.chunk_moves(rc, cc, ell, device=False)
        check_lists(rc, cc, ell, *dev)
        check_lists(rc, cc, ell, *host)
        assert len(dev[1]) == len(host[1]), (rc, cc, ell)
        assert len(dev[3]) == len(host[3]), (rc, cc, ell)


@pytest.mark.gpu
def test_device_move_lists_at_the_chunk_limit(ipt):
    """Chunk matrices of just under 2^23 positions (the plan limit): one with
    short cycles of up to 23 (32 x 262143), one whose longest cycle is a few
    thousand (7 x 1198371), and a square one (all cycles of two)."""
    for rc, cc, ell in [(32, 262143, 16), (7, 1198371, 64), (2896, 2896, 128)]:
        check_lists(rc, cc, ell, *ipt.chunk_moves(rc, cc, ell, device=True))


def walk_lists(rc, cc, ell):
    """Lists of the chunk transpose by a plain cycle walk (each cycle from its
    least position; cycles longer than ell cut into head-saved segments,
    shorter ones one item each): only to exercise check_lists on the CPU."""
    n = rc * cc
    perm = lambda x: (x % cc) * rc + x // cc
    seen = np.zeros(n, dtype=bool)
    items, msrc, mdst, hpos = [], [], [], []
    for q0 in range(n):
        if seen[q0]:
            continue
        cyc, x = [], q0
        while not seen[x]:
            seen[x] = True
            cyc.append(x)
            x = perm(x)
        if len(cyc) == 1:
            continue
        L = len(cyc)
        if L <= ell:
            items.append((len(msrc), L))
            msrc += cyc
            mdst += [cyc[(j + 1) % L] for j in range(L)]
            continue
        for a in range(0, L, ell):
            seg = min(ell, L - a)
            items.append((len(msrc), seg))
            msrc += [int(FLAG) | len(hpos)] + cyc[a + 1:a + seg]
            mdst += [cyc[(a + j + 1) % L] for j in range(seg)]
            hpos.append(cyc[a])
    u = lambda v: np.array(v, dtype=np.uint32)
    return np.array(items, dtype=np.uint32).reshape(-1, 2), u(msrc), u(mdst), u(hpos)


def test_check_lists_accepts_a_walk_and_rejects_planted_mistakes():
    for rc, cc, ell in [(3, 5, 2), (7, 11, 4), (16, 9, 128), (2, 9, 1)]:
        check_lists(rc, cc, ell, *walk_lists(rc, cc, ell))
    items, msrc, mdst, hpos = walk_lists(7, 11, 4)
    bad = [
        (items, msrc, np.roll(mdst, 1), hpos),                       # moves rotated
        (items, msrc[::-1].copy(), mdst, hpos),                      # sources permuted
        (np.vstack([items[1:], items[:1]]) * np.array([1, 2], dtype=np.uint32), msrc, mdst, hpos),
    ]
    flag_bad = msrc.copy()
    j = int(np.nonzero((flag_bad & FLAG) == 0)[0][1])
    flag_bad[j] |= FLAG                                              # head flag mid-item
    bad.append((items, flag_bad, mdst, hpos))
    merged = items.copy()                                            # an item longer than ell
    merged = np.vstack([[merged[0][0], merged[0][1] + merged[1][1]], merged[2:]]).astype(np.uint32)
    bad.append((merged, msrc, mdst, hpos))
    for k, lists in enumerate(bad):
        with pytest.raises(AssertionError):
            check_lists(7, 11, 4, *lists)
            pytest.fail(f"planted mistake {k} was accepted")
