"""CPU model of the persistent batch kernel's work shares and stage counting
(convfpga_amd/csrc/kernels/batch_kernel.h, batch_kernel()): work-group w of G
runs items w, w + G, ... of each stage; on leaving stages it counts itself
(and its tiles) out of each on its XCD's 64-bit counter, and the last
work-group of an XCD out of a stage publishes the XCD's tiles.  Checked on
random stage sizes, grids and XCD assignments, with the work-groups
interleaved at random: no deadlock, every item run exactly once, every stage
counter reaching its total, one publication per XCD per stage."""
import random

MAX_STAGES = 5  # kBatchMaxStages


def _run(stage_items, tiles, G, xcc_of, rng):
    nst = len(stage_items)
    first = [0]
    for c in stage_items:
        first.append(first[-1] + c)
    first += [first[-1]] * (MAX_STAGES + 1 - len(first))
    nitems = first[-1]
    need = [sum(tiles[s]) if s < nst else 0 for s in range(MAX_STAGES)]
    stage_of = [s for s in range(nst) for _ in range(stage_items[s])]
    tile_of = [t for s in range(nst) for t in tiles[s]]
    wgs = {}
    for b in range(G):
        wgs[xcc_of[b]] = wgs.get(xcc_of[b], 0) + 1
    seqs = [[it for s in range(MAX_STAGES) for it in range(first[s] + b, first[s + 1], G)] for b in range(G)]
    xs, stage_ctr, published, runs = {}, [0] * MAX_STAGES, [0] * MAX_STAGES, [0] * nitems
    pos, cur, done, left, fin = [0] * G, [-1] * G, [0] * G, [0] * G, [False] * G
    steps = 0
    while not all(fin):
        steps += 1
        assert steps < 10 ** 6, "deadlock"
        b = rng.randrange(G)
        if fin[b]:
            continue
        item = seqs[b][pos[b]] if pos[b] < len(seqs[b]) else nitems
        stage = stage_of[item] if item < nitems else MAX_STAGES
        if stage != cur[b]:
            x = xcc_of[b]
            for s in range(left[b], min(stage, MAX_STAGES)):
                if need[s] == 0:
                    continue
                mine = done[b] if s == cur[b] else 0
                old = xs.get((x, s), 0)
                xs[(x, s)] = old + (1 << 32) + mine
                if (old >> 32) + 1 == wgs[x]:
                    stage_ctr[s] += (old & 0xFFFFFFFF) + mine
                    published[s] += 1
            left[b] = max(left[b], stage)
            if item < nitems and stage > 0 and stage_ctr[stage - 1] < need[stage - 1]:
                continue  # waiting for the previous stage
            cur[b], done[b] = stage, 0
        if item >= nitems:
            fin[b] = True
            continue
        runs[item] += 1
        done[b] += tile_of[item]
        pos[b] += 1
    assert runs == [1] * nitems
    for s in range(nst):
        assert stage_ctr[s] == need[s]
        assert published[s] == len(wgs)


def test_stage_counting_model():
    rng = random.Random(7)
    for _ in range(400):
        nst = rng.randint(1, MAX_STAGES)
        stage_items = [rng.randint(1, 200) for _ in range(nst)]
        tiles = [[rng.randint(1, 4) for _ in range(c)] for c in stage_items]
        G = rng.randint(1, min(sum(stage_items), 150))
        nx = rng.choice([1, 2, 8])
        xcc_of = [rng.randrange(nx) if rng.random() < 0.3 else b % nx for b in range(G)]
        _run(stage_items, tiles, G, xcc_of, rng)


def _place(j, S):
    """Tile group of the segment's item j with sibling groups of S items
    (batch_kernel.h, the item's place in its segment)."""
    if S <= 1:
        return j
    x, r = j % 8, j // 8
    return ((r // S) * 8 + x) * S + r % S


def test_sibling_placement():
    """Items of a segment with sibling groups of S: every tile group is run
    exactly once, and the S items of one sibling group land on work-groups of
    one XCD (w % 8, with G a multiple of 8) in the same round of the static
    shares -- whatever the segment's offset in its stage."""
    rng = random.Random(11)
    for _ in range(300):
        S = rng.choice([2, 4])
        items = 8 * S * rng.randint(1, 40)
        offset = rng.randint(0, 500)  # first item of the segment within its stage
        G = 8 * rng.randint(1, 100)
        placed = [_place(j, S) for j in range(items)]
        assert sorted(placed) == list(range(items))
        where = {}
        for j, t in enumerate(placed):
            r = offset + j
            where[t] = ((r % G) % 8, r // G)  # (XCD of the work-group, round)
        for grp in range(items // S):
            xcds = {where[grp * S + k][0] for k in range(S)}
            assert len(xcds) == 1, (S, items, offset, G, grp)
            rounds = {where[grp * S + k][1] for k in range(S)}
            assert max(rounds) - min(rounds) <= 1 + (S * 8) // G


def _list_share(base, K, first, b, G, nitems):
    """Positions base .. base+K-1 of work-group b's share, as the kernel lists
    them in LDS (batch_kernel.h, list_share): per stage, ceil((end - start)/G)
    items from start = first[s] + b, step G; past the end, nitems."""
    out = []
    for t in range(K):
        p, item = base + t, nitems
        for s in range(MAX_STAGES):
            f, e = first[s] + b, first[s + 1]
            cnt = (e - f + G - 1) // G if f < e else 0
            if p < cnt:
                item = f + p * G
                break
            p -= cnt
        out.append(item)
    return out


def test_listed_shares_in_chunks():
    """The work-group's items listed K positions at a time, chunk after chunk,
    are exactly its static share in stage order, then nitems."""
    rng = random.Random(13)
    for _ in range(300):
        nst = rng.randint(1, MAX_STAGES)
        stage_items = [rng.randint(0, 3000) for _ in range(nst)]
        first = [0]
        for c in stage_items:
            first.append(first[-1] + c)
        first += [first[-1]] * (MAX_STAGES + 1 - len(first))
        nitems = first[-1]
        if nitems == 0:
            continue
        G = rng.randint(1, min(nitems, 600))
        K = rng.choice([8, 64, 512])
        b = rng.randrange(G)
        share = [it for s in range(MAX_STAGES) for it in range(first[s] + b, first[s + 1], G)]
        listed = []
        base = 0
        while True:
            chunk = _list_share(base, K, first, b, G, nitems)
            listed += chunk
            if chunk[-1] == nitems:
                break
            base += K
        n = len(share)
        assert listed[:n] == share
        assert all(x == nitems for x in listed[n:])
