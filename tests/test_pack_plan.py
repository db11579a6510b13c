"""The forward path's work plan on the CPU: the planner of
acctricnt_amd/csrc/fwd_plan.h (the header the HIP library compiles), built
here with g++, and an emulation of the wave-pack kernel's index arithmetic
(fwd_pack_unit in acctricnt_amd/csrc/k_forward.hip) over its output, and of
the windowed probe searches of acctricnt_amd/csrc/flat_window.h that the
forward and per-edge kernels share (lds_search over a 512-entry prefix in
LDS, lane_window over the lanes of a wave) and of its per-probe hit adds
(add_row_hits), and the merge path's padded window search
(PaddedWin in acctricnt_amd/csrc/k_paths.hip).

For every pack the emulation recomputes what each lane of the kernel
computes -- the sources' upper-slot, lower-entry and hash-table prefixes, each
probe's upper slot, list start and length, each lower entry's source and
table -- and checks that every index stays inside its array, that every
probe that reads elements has a hash table to look them up in (a source with
upper slots and an empty N<(b) has none: the kernel gives its probes length 0,
or the lookup would never meet an empty entry), that the packs respect the
limits the kernel's LDS layout assumes (64 sources, 64 upper slots, 256
lower entries, 512 table entries), and that the hits found through this
arithmetic equal the triangles counted from the definition.  Workgroup units
must tile each remaining source's upper slots.  No GPU is used."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import bruteforce as BF
from acctricnt_amd import graph as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "acctricnt_amd", "csrc")

WAVE_CAP, WAVE_TABLE, WAVE_SLOTS = 256, 512, 64  # accfwd::kWaveCap, 1 << kWaveTableBits, kWaveSlots
LDS_CAP, BITMAP_IDS, BITMAP_MAX_LOW = 2048, 1 << 17, 65535


class Unit(C.Structure):
    _fields_ = [("b", C.c_int32), ("j0", C.c_uint32), ("j1", C.c_uint32), ("nlow", C.c_uint32)]


@pytest.fixture(scope="module")
def planner(tmp_path_factory):
    out = tmp_path_factory.mktemp("plan") / "libplan.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-pthread", "-I" + CSRC,
                    os.path.join(ROOT, "tests", "plan", "plan_driver.cpp"), "-o", str(out)],
                   check=True, capture_output=True, text=True)
    lib = C.CDLL(str(out))
    u32p, i32p, up = C.POINTER(C.c_uint32), C.POINTER(C.c_int32), C.POINTER(Unit)
    lib.plan_forward.argtypes = [C.c_uint32, u32p, u32p, i32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                 C.c_uint32, up, u32p, up, u32p, up, u32p]
    lib.plan_forward.restype = C.c_int
    return lib


def vertex_meta(g):
    """hs[v]: first slot of N(v) above v; lo[v] = min(N(v) and v) (k_vertex_meta)."""
    off = g.node_off.astype(np.int64)
    src = g.sources()
    low = np.bincount(src[g.edge_dst < src], minlength=g.n)
    hs = (off[:-1] + low).astype(np.uint32)
    lo = np.arange(g.n, dtype=np.int32)
    has = np.diff(off) > 0
    lo[has] = np.minimum(g.edge_dst[off[:-1][has]], lo[has])
    return hs, lo


def plan(lib, g, wave_slots=64, wave_low=256, chunk=512, device_waves=8192):
    hs, lo = vertex_meta(g)
    cap = g.m + g.n + 1
    bufs = [(Unit * cap)() for _ in range(3)]
    counts = [C.c_uint32() for _ in range(3)]
    off = np.ascontiguousarray(g.node_off, np.uint32)
    rc = lib.plan_forward(g.n, off.ctypes.data_as(C.POINTER(C.c_uint32)), hs.ctypes.data_as(C.POINTER(C.c_uint32)),
                          lo.ctypes.data_as(C.POINTER(C.c_int32)), wave_slots, wave_low, chunk, device_waves, cap,
                          bufs[0], C.byref(counts[0]), bufs[1], C.byref(counts[1]), bufs[2], C.byref(counts[2]))
    assert rc == 0
    return [[(u.b, u.j0, u.j1, u.nlow) for u in bufs[i][:counts[i].value]] for i in range(3)], hs, lo


def lane_search(X, e):
    """Largest k in [0, 64) with X[k] <= e (k_forward.hip lane_search)."""
    k = 0
    for st in (32, 16, 8, 4, 2, 1):
        if X[k + st] <= e:
            k += st
    return k


def excl_scan(v):
    return np.concatenate([[0], np.cumsum(v)[:-1]]).astype(np.int64)


def pack_bits(nl):
    bits = 4
    while (1 << bits) < 2 * nl:
        bits += 1
    return bits


def emulate_pack(g, hs, rev, pack, guard=True):
    """One wave pack as fwd_pack_unit computes it; returns the triangles found
    (hits) and raises AssertionError on any index outside its range."""
    off, adj = g.node_off.astype(np.int64), g.edge_dst
    b0, b1, nup_tot, nl_tot = pack
    nv = b1 - b0
    assert 1 <= nv <= 64 and 1 <= nup_tot <= WAVE_SLOTS and nl_tot <= WAVE_CAP
    o0, h, nupv, nlv, tsz, bits = (np.zeros(64, np.int64) for _ in range(6))
    for lane in range(nv):
        v = b0 + lane
        o0[lane], h[lane] = off[v], hs[v]
        nupv[lane] = off[v + 1] - hs[v]
        if nupv[lane]:
            nlv[lane] = hs[v] - off[v]
            if nlv[lane]:
                bits[lane] = pack_bits(nlv[lane])
                tsz[lane] = 1 << bits[lane]
    Q, NL, TB = excl_scan(nupv), excl_scan(nlv), excl_scan(tsz)
    tbl_tot = TB[63] + tsz[63]
    assert nupv.sum() == nup_tot and nlv.sum() == nl_tot, "pack totals differ from its sources"
    assert tbl_tot <= WAVE_TABLE, f"hash tables of {tbl_tot} entries exceed the wave's {WAVE_TABLE}"
    # lower entries: four per lane, each into its source's table
    tables = {}
    for r in range(4):
        for lane in range(64):
            e = lane + 64 * r
            k = lane_search(NL, e)
            if e >= nl_tot:
                continue
            lpos = e - NL[k]
            assert 0 <= lpos < nlv[k] and bits[k] > 0
            a = adj[o0[k] + lpos]
            assert a < b0 + k, "a lower entry that is not below its source"
            assert TB[k] + tsz[k] <= tbl_tot
            tables.setdefault(k, set()).add(int(a))
    for k, t in tables.items():
        assert len(t) == nlv[k] and 2 * len(t) <= tsz[k], "a table without a free entry never ends a lookup"
    # probes: one per lane, laid end to end
    hits = 0
    for lane in range(nup_tot):
        pk = lane_search(Q, lane)
        j = h[pk] + (lane - Q[pk])
        v = b0 + pk
        assert hs[v] <= j < off[v + 1], "a probe slot outside its source's upper slots"
        c = adj[j]
        p = rev[j]
        cb = off[c]
        if guard:
            ln = p - cb if (p != 0xFFFFFFFF and bits[pk] > 0) else 0
        else:
            ln = p - cb if p != 0xFFFFFFFF else 0
        if ln == 0:
            continue
        assert bits[pk] > 0, "a probe with elements into a source without a table (the lookup never ends)"
        assert off[c] <= cb + ln - 1 < off[c + 1] and adj[cb + ln - 1] < v, "a probe element at or above b"
        hits += sum(1 for a in adj[cb:cb + ln] if int(a) in tables.get(pk, ()))
    return hits


def middle_triangles(g, v):
    """Triangles whose middle vertex is v, from the definition."""
    nb = g.edge_dst[g.node_off[v]:g.node_off[v + 1]]
    lower = set(int(x) for x in nb if x < v)
    t = 0
    for c in nb[nb > v]:
        nc = g.edge_dst[g.node_off[c]:g.node_off[c + 1]]
        t += sum(1 for a in nc if a < v and int(a) in lower)
    return t


def empty_low_source_graph():
    """Source 1 has no lower neighbor but an upper neighbor 5 whose list holds
    0 < 1: its probe of N(5) is non-empty while N<(1) is empty."""
    s = [0, 1, 0, 2, 3, 2, 4, 6]
    d = [5, 5, 2, 5, 5, 3, 6, 7]
    return G.from_edges(8, s, d, name="empty-low")


GRAPHS = {
    "rmat11": lambda: G.gen_rmat(11, 16, seed=3),
    "rmat11_raw": lambda: G.gen_rmat(11, 8, seed=4, relabel=False),
    "er3000": lambda: G.gen_er(3000, 12, seed=5),
    "chunglu4000": lambda: G.gen_chung_lu(4000, 2.1, 10, 600, seed=6),
    "empty_low": empty_low_source_graph,
    "wheel9": lambda: BF.wheel(9),
    "K300": lambda: BF.complete(300),
}


@pytest.mark.parametrize("params", [dict(), dict(wave_slots=16, wave_low=64, chunk=128, device_waves=64),
                                    dict(wave_slots=0), dict(wave_low=0, chunk=64),
                                    dict(device_waves=10 ** 7)], ids=["default", "small_bins", "no_packs",
                                                                      "empty_low_only", "one_source_packs"])
@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_plan_bounds_and_pack_arithmetic(planner, name, params):
    g = GRAPHS[name]()
    (waves, blocks, bigs), hs, lo = plan(planner, g, **params)
    rev = g.reverse_slots()
    off = g.node_off.astype(np.int64)
    covered = np.zeros(g.n, np.int64)  # upper slots planned per source
    for pack in waves:
        b0, b1, nup_tot, nl_tot = pack
        for v in range(b0, b1):
            nup = off[v + 1] - hs[v]
            nl = hs[v] - off[v]
            if nup:
                assert nup <= params.get("wave_slots", 64) and nl <= params.get("wave_low", 256)
            covered[v] += nup
        hits = emulate_pack(g, hs, rev, pack)
        assert hits == sum(middle_triangles(g, v) for v in range(b0, b1)), pack
    for unit, lds in [(u, True) for u in blocks] + [(u, False) for u in bigs]:
        b, j0, j1, nlow = unit
        assert hs[b] <= j0 < j1 <= off[b + 1] and j1 - j0 <= params.get("chunk", 512)
        assert nlow == hs[b] - off[b]
        fits = nlow <= LDS_CAP or (b - lo[b] <= BITMAP_IDS and nlow <= BITMAP_MAX_LOW)
        assert fits == lds
        covered[b] += j1 - j0
    assert np.array_equal(covered, off[1:] - hs.astype(np.int64)), "upper slots not planned exactly once"


def test_empty_low_source_needs_the_guard(planner):
    """Without the kernel's rule (no table: probe length 0), source 1's probe
    of N(5) would look up 0 in a table that does not exist; the emulation
    reports it.  The kernel source still carries the rule."""
    g = empty_low_source_graph()
    (waves, _, _), hs, _ = plan(planner, g)
    rev = g.reverse_slots()
    pack = next(p for p in waves if p[0] <= 1 < p[1])
    emulate_pack(g, hs, rev, pack)
    with pytest.raises(AssertionError, match="without a table"):
        emulate_pack(g, hs, rev, pack, guard=False)
    src = open(os.path.join(CSRC, "k_forward.hip")).read()
    assert "len = p != ~0u && p_bits > 0 ? p - cb : 0;" in src


def test_pack_limits_are_checked(planner):
    """Packs beyond the kernel's LDS layout are reported: 257 lower entries,
    or tables of more than 512 entries."""
    g = BF.complete(300)
    hs, _ = vertex_meta(g)
    rev = g.reverse_slots()
    with pytest.raises(AssertionError):
        emulate_pack(g, hs, rev, (257, 258, 42, 257))
    # sources 2..9 of K300: 2..9 lower entries each, 16-entry tables; claim fewer than they hold
    with pytest.raises(AssertionError, match="totals"):
        emulate_pack(g, hs, rev, (2, 10, 64, 10))


def window_search(P, e0, U, T, first_step=lambda w, hp: w + 1 - hp):
    """k(e) of every element of the round at e0 as lds_search (flat_window.h)
    computes it (a two-level ballot search for the round's first and last probes,
    then a binary search of that window per element); P is s_P (512 entries,
    T beyond the chunk's probes).  Returns (k per element row, klo, khi) and
    raises AssertionError on a read outside [0, 512) or outside the window."""
    assert len(P) == 512 and e0 < T
    e1 = min(e0 + 64 * U - 1, T - 1)
    coarse = P[0:512:8]
    glo = 8 * (int(np.count_nonzero(coarse <= e0)) - 1)
    ghi = 8 * (int(np.count_nonzero(coarse <= e1)) - 1)
    assert 0 <= glo <= ghi <= 504
    klo = glo + sum(1 for j in range(1, 8) if P[glo + j] <= e0)
    khi = ghi + sum(1 for j in range(1, 8) if P[ghi + j] <= e1)
    w = khi - klo
    ks = np.full((U, 64), klo, np.int64)
    for q in range(U):
        for lane in range(64):
            e = e0 + 64 * q + lane
            k = klo
            if w:
                hp = 1 << (w.bit_length() - 1)
                d = klo + first_step(w, hp)
                assert klo <= d <= khi, "a first step outside the window"
                if P[d] <= e:
                    k = d
                st = hp >> 1
                while st:
                    assert klo <= k + st <= khi, "a read outside the window"
                    if P[k + st] <= e:
                        k += st
                    st >>= 1
            ks[q, lane] = k
    return ks, klo, khi


def chunk_prefix(lengths):
    """s_P of a chunk: exclusive prefix of the probe lengths, T beyond them."""
    P = np.zeros(512, np.int64)
    ex = np.concatenate([[0], np.cumsum(lengths)])
    P[:len(lengths)] = ex[:-1]
    P[len(lengths):] = ex[-1]
    return P, int(ex[-1])


def chunk_cases():
    rng = np.random.default_rng(11)
    yield "one_long", [5000]
    yield "all_ones", [1] * 512
    yield "zeros_between", [0, 3, 0, 0, 70, 0, 1, 0, 0, 0, 300, 0]
    yield "leading_zeros", [0] * 40 + [9] * 30 + [0] * 20
    yield "full_chunk_skew", list((rng.pareto(1.2, 512) * 8).astype(int))
    yield "sparse", list(rng.integers(0, 3, 512))
    yield "short", [2]
    yield "mid_sizes", list(rng.integers(0, 700, 37))


@pytest.mark.parametrize("U", [1, 2, 4])
@pytest.mark.parametrize("case", [c[0] for c in chunk_cases()])
def test_block_window_search(case, U):
    """Every element's probe equals the largest k with s_P[k] <= e, and lanes
    past the chunk's last element get the window's last probe (so a round's
    probe indices never decrease across lanes)."""
    lengths = dict(chunk_cases())[case]
    P, T = chunk_prefix(lengths)
    for e0 in range(0, T, 64 * U):
        ks, klo, khi = window_search(P, e0, U, T)
        for q in range(U):
            e = e0 + 64 * q + np.arange(64)
            want = np.searchsorted(P, e, side="right") - 1
            valid = e < T
            assert np.array_equal(ks[q][valid], want[valid]), (case, e0, q)
            assert np.all(ks[q][~valid] == khi)
            assert np.all(np.diff(ks[q]) >= 0)
            k = ks[q][valid]
            assert np.all(P[k] <= e[valid]) and np.all(P[np.minimum(k + 1, 511)][k < 511] > e[valid][k < 511])


def test_block_window_search_catches_a_short_first_step():
    """A first step one entry short leaves the window's last probe out of
    reach: the emulation reports wrong probes."""
    P, T = chunk_prefix([0, 3, 0, 0, 70, 0, 1, 0, 0, 0, 300, 0])
    bad = 0
    for e0 in range(0, T, 64):
        ks, _, _ = window_search(P, e0, 1, T, first_step=lambda w, hp: w - hp if w > hp else 1)
        e = e0 + np.arange(64)
        valid = e < T
        bad += int(np.count_nonzero(ks[0][valid] != (np.searchsorted(P, e, side="right") - 1)[valid]))
    assert bad > 0
    src = open(os.path.join(CSRC, "flat_window.h")).read()
    assert "const uint32_t d = klo + w + 1 - hp;" in src


def lane_window(X, e0, U, T):
    """lane_window (flat_window.h) for the round at e0: X holds the 64
    lanes' exclusive prefix of the pack's probe lengths (T beyond its last
    probe).  Returns k per element row; AssertionError on a lane read outside
    [klo, khi]."""
    assert len(X) == 64 and e0 < T
    e1 = min(e0 + 64 * U - 1, T - 1)
    klo = int(np.count_nonzero(X <= e0)) - 1
    khi = int(np.count_nonzero(X <= e1)) - 1
    w = khi - klo
    ks = np.full((U, 64), klo, np.int64)
    for q in range(U):
        for lane in range(64):
            e = e0 + 64 * q + lane
            k = klo
            if w:
                hp = 1 << (w.bit_length() - 1)
                d = klo + w + 1 - hp
                if X[d] <= e:
                    k = d
                st = hp >> 1
                while st:
                    assert klo <= k + st <= khi, "a lane read outside the window"
                    if X[k + st] <= e:
                        k += st
                    st >>= 1
            ks[q, lane] = k
    return ks, khi


@pytest.mark.parametrize("U", [1, 2, 4])
@pytest.mark.parametrize("seed", range(6))
def test_lane_window_search(seed, U):
    """The wave pack's per-element probe equals lane_search's (the largest
    lane k with X_k <= e) for every element, with zero-length probes and lanes
    past the pack's probes (length 0, X = T) included."""
    rng = np.random.default_rng(100 + seed)
    nup = int(rng.integers(1, 65))
    lengths = np.zeros(64, np.int64)
    lengths[:nup] = rng.integers(0, [3, 40, 300, 5, 2000, 64][seed], nup)
    if lengths.sum() == 0:
        lengths[nup - 1] = 1
    X = np.concatenate([[0], np.cumsum(lengths)[:-1]])
    T = int(lengths.sum())
    for e0 in range(0, T, 64 * U):
        ks, khi = lane_window(X, e0, U, T)
        for q in range(U):
            for lane in range(64):
                e = e0 + 64 * q + lane
                want = lane_search(X, e) if e < T else khi
                assert ks[q, lane] == want, (seed, e0, q, lane)


def add_row_hits(k, hit):
    """add_row_hits (flat_window.h) for one row of 64 lanes: the LDS adds it
    issues, as {probe: amount}, and the number of adds."""
    k = [int(x) for x in k]
    hits = sum(1 << l for l in range(64) if hit[l])
    if not hits:
        return {}, 0
    if k[0] == k[63]:
        return {k[0]: bin(hits).count("1")}, 1
    heads = sum(1 << l for l in range(64) if l == 0 or k[l] != k[l - 1])
    adds, n = {}, 0
    M = (1 << 64) - 1
    for l in range(64):
        if not (heads >> l) & 1:
            continue
        above = heads & ~(((2 << l) & M) - 1) & M
        seg = ((above & -above) - 1 if above else M) & ~((1 << l) - 1) & M
        c = bin(hits & seg).count("1")
        if c:
            assert k[l] not in adds, "two adds on one probe"
            adds[k[l]] = c
            n += 1
    return adds, n


@pytest.mark.parametrize("seed", range(20))
def test_row_hits_by_segment(seed):
    """The per-segment popcounts add up to the per-hit counts of every probe,
    with at most one add per probe of the row."""
    rng = np.random.default_rng(200 + seed)
    for _ in range(50):
        k = np.sort(rng.integers(0, [1, 2, 5, 64][seed % 4] + 1, 64)) + int(rng.integers(0, 400))
        hit = rng.random(64) < [0.05, 0.3, 0.9][seed % 3]
        adds, n = add_row_hits(k, hit)
        want = {}
        for l in range(64):
            if hit[l]:
                want[int(k[l])] = want.get(int(k[l]), 0) + 1
        assert adds == want
        assert n == len(want)


INT32_MAX = 2**31 - 1


def padded_win_has(window, log, a, first_read=lambda pos, step: pos + step - 1):
    """PaddedWin<log>::has_n for one element a: the window (sorted, len <=
    2^log) padded with INT32_MAX, log branch-free steps, then one compare."""
    s = list(window) + [INT32_MAX] * ((1 << log) - len(window))
    pos, step = 0, 1 << (log - 1)
    while step:
        i = first_read(pos, step)
        assert 0 <= i < (1 << log), "a read outside the padded window"
        if s[i] < a:
            pos += step
        step >>= 1
    assert 0 <= pos < (1 << log)
    return a >= 0 and s[pos] == a


@pytest.mark.parametrize("log", [3, 8, 11])
def test_padded_window_search(log):
    """Membership equals the set's for full, partial and empty windows, for
    ids below, inside, between and above the window's (the last window of a
    list is partial and every window's slice may reach past its last id)."""
    rng = np.random.default_rng(300 + log)
    for fill in (0, 1, 3, (1 << log) - 1, 1 << log):
        for _ in range(20):
            w = np.sort(rng.choice(5 * (1 << log) + 10, size=fill, replace=False)).astype(np.int64)
            ws = set(int(x) for x in w)
            for a in list(range(-1, 5 * (1 << log) + 12, max(1, (1 << log) // 64))) + [int(x) for x in w]:
                assert padded_win_has(w, log, a) == (a >= 0 and a in ws), (log, fill, a)


def test_padded_window_search_catches_an_off_by_one():
    w = list(range(0, 40, 2))
    # reading one entry too far at every step (kept inside the window) misses ids
    bad = sum(padded_win_has(w, 5, a, first_read=lambda pos, step: min(pos + step, 31)) != (a in w)
              for a in range(40))
    assert bad > 0
    src = open(os.path.join(CSRC, "k_paths.hip")).read()
    assert "v[q] = s[pos[q] + step - 1];" in src and "INT32_MAX" in src
