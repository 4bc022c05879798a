"""LDS bank conflicts of the tile product's staging (fw_cuda_amd/csrc/fwx_tile.hpp),
checked on the CPU against the gfx950 bank model measured with
fw_cuda_amd/microbench/lds_banks.hip (rocprofv3 SQ_LDS_BANK_CONFLICT per
pattern; profiles/lds_banks_mi355x.md):

* 32 banks of 4 bytes;
* 4-byte accesses serve 32 lanes per cycle (stride 8 bytes: 2-way conflict,
  stride 128 bytes: 32-way);
* 8-byte reads serve 32 lanes per cycle over 256 bytes (a lane stride of 16
  bytes conflicts 2-way; lanes reading the same address are free);
* 16-byte reads serve 16 lanes per cycle over 256 bytes;
* 8-byte writes serve 16 lanes and 16-byte writes 8 lanes per cycle, over
  128 bytes;
* a paired 4-byte write (ds_write2_b32) is two 4-byte writes: on 8-byte
  slots it conflicts 2-way (4 extra cycles per instruction, measured).

A pass of lanes is conflict-free when its distinct addresses fall in
distinct bank groups modulo the pass width.  The test walks every LDS
access of the staging -- A stash (16-byte writes), B stash (8-byte writes
of pair-interleaved slots), A reads (16 bytes), B reads (8 bytes) -- for the
512-thread 128-wide product and the 256-thread 32/64/128-wide ones, with
the same index formulas as the kernel, and asserts that no pass conflicts.
Planted mistakes are flagged: the B stash as two 4-byte writes (what the
compiler emitted before the staging arrays were 16-byte aligned), 16-deep A
rows padded like 32-deep ones, and 32-wide B rows without the odd-row
rotation.
"""
import pytest

# (bytes per access, is_write) -> (lanes per pass, bytes per pass)
PASS = {(4, False): (32, 128), (4, True): (32, 128), (8, False): (32, 256), (16, False): (16, 256),
        (8, True): (16, 128), (16, True): (8, 128)}


def conflicts(addrs, width, write):
    """Extra cycles of one wave's access: addrs[l] = byte address of lane l
    (None = lane inactive)."""
    lanes, span = PASS[(width, write)]
    extra = 0
    for p0 in range(0, 64, lanes):
        distinct = {a for a in addrs[p0:p0 + lanes] if a is not None}
        groups = {}
        for a in distinct:
            g = (a % span) // width
            groups[g] = groups.get(g, 0) + 1
        ways = max(groups.values()) if groups else 1
        extra += ways - 1
    return extra


class Shape:
    """Index math of TileProduct<BM, BN, KC, VEC, NTX> (fwx_tile.hpp)."""

    def __init__(self, BM, BN, KC, NTX):
        self.BM, self.BN, self.KC, self.NTX = BM, BN, KC, NTX
        self.NT = NTX * 16
        self.TM, self.TN = BM // 16, BN // NTX
        self.LDA = 16 if KC == 16 else KC + 4  # tile_lda
        self.SKEW = 1 if BN // 4 < 16 else 0
        self.G = min(4, self.TN)
        self.AG = KC // 4
        self.AROWS = self.NT // self.AG
        self.APASS = (BM + self.AROWS - 1) // self.AROWS
        self.BU = (KC // 2) * (BN // 4)
        self.BPASS = (self.BU + self.NT - 1) // self.NT
        self.a_bytes = 2 * BM * self.LDA * 4  # the b arrays follow the a arrays

    def colmap(self, tx, c):
        G = self.G
        return G * (tx + self.NTX * (c // G)) + c % G

    def bpos(self, j):
        G = self.G
        return (j % G) * (self.BN // G) + j // G

    def a_stash(self, tid, p):
        r = p * self.AROWS + tid // self.AG
        k = (tid % self.AG) * 4
        return (r * self.LDA + k) * 4 if r < self.BM else None

    def b_stash(self, tid, u, e):
        idx = u * self.NT + tid
        if idx >= self.BU:
            return None
        pr, j = idx // (self.BN // 4), (idx % (self.BN // 4)) * 4
        return self.a_bytes + self.bslot(pr, j + e) * 8

    def bslot(self, pr, j):
        if self.SKEW:
            return pr * self.BN + ((self.bpos(j) + (pr & 1)) & (self.BN - 1))
        return pr * self.BN + self.bpos(j)

    def a_read(self, tid, q, r):
        ty = tid // self.NTX
        return ((ty + 16 * r) * self.LDA + 4 * q) * 4

    def b_read(self, tid, q, c, second):
        tx = tid % self.NTX
        return self.a_bytes + self.bslot(2 * q + second, self.colmap(tx, c)) * 8


SHAPES = {"128 wide, 512 threads (8x4 cells)": (128, 128, 32, 32),
          "128 wide, 256 threads (8x8 cells)": (128, 128, 32, 16),
          "128 wide, 256 threads, 16-deep chunks": (128, 128, 16, 16),
          "64 wide, 256 threads (4x4 cells)": (64, 64, 32, 16),
          "64 wide, 256 threads, 16-deep chunks": (64, 64, 16, 16),
          "32 wide, 256 threads (2x2 cells)": (32, 32, 32, 16)}


def waves(sh):
    return [list(range(w * 64, w * 64 + 64)) for w in range(sh.NT // 64)]


@pytest.mark.parametrize("name", list(SHAPES))
def test_staging_and_operand_reads_are_conflict_free(name):
    sh = Shape(*SHAPES[name])
    total = 0
    for lanes in waves(sh):
        for p in range(sh.APASS):
            total += conflicts([sh.a_stash(t, p) for t in lanes], 16, True)
        for u in range(sh.BPASS):
            for e in range(4):
                total += conflicts([sh.b_stash(t, u, e) for t in lanes], 8, True)
        for q in range(sh.KC // 4):
            for r in range(sh.TM):
                total += conflicts([sh.a_read(t, q, r) for t in lanes], 16, False)
            for c in range(sh.TN):
                for second in (0, 1):
                    total += conflicts([sh.b_read(t, q, c, second) for t in lanes], 8, False)
    assert total == 0, f"{name}: {total} conflict cycles per chunk"


def test_model_reproduces_the_measured_patterns():
    """The model gives the probe's measured extra cycles per instruction."""
    lane = list(range(64))
    assert conflicts([4 * l for l in lane], 4, False) == 0
    assert conflicts([8 * l for l in lane], 4, False) == 2        # measured 2.0
    assert conflicts([128 * l for l in lane], 4, False) == 62     # measured 62.0
    assert conflicts([8 * l for l in lane], 8, False) == 0
    assert conflicts([8 * (l % 32) for l in lane], 8, False) == 0
    assert conflicts([16 * l for l in lane], 8, False) == 2       # measured 2.0
    assert conflicts([144 * ((l >> 5) & 1) for l in lane], 16, False) == 0
    assert conflicts([16 * l for l in lane], 16, False) == 0
    slots = [(l & 31) + ((l >> 5) & 1) * 128 for l in lane]
    assert conflicts([8 * s for s in slots], 8, True) == 0
    w2 = conflicts([8 * s for s in slots], 4, True) + conflicts([8 * s + 4 for s in slots], 4, True)
    assert w2 == 4                                                 # measured 4.0
    assert conflicts([(l >> 3) * 144 + (l & 7) * 16 for l in lane], 16, True) == 0


def test_b_slots_cover_each_pair_row_once():
    """Every (pair row, column) gets its own slot inside its pair row, for
    every shape (the rotation of odd rows in 32-wide blocks included)."""
    for BM, BN, KC, NTX in SHAPES.values():
        sh = Shape(BM, BN, KC, NTX)
        for pr in range(KC // 2):
            slots = {sh.bslot(pr, j) for j in range(BN)}
            assert slots == set(range(pr * BN, (pr + 1) * BN))


def test_planted_mistakes_are_flagged():
    """A 16-deep chunk with the 32-deep padding (row stride 20 dwords), and a
    32-wide block without the odd-row rotation, conflict."""
    sh = Shape(128, 128, 16, 16)
    sh.LDA = 20
    assert sum(conflicts([sh.a_stash(t, p) for t in range(64)], 16, True) for p in range(sh.APASS)) > 0
    sh = Shape(32, 32, 32, 16)
    sh.SKEW = 0
    assert sum(conflicts([sh.b_stash(t, 0, e) for t in range(64)], 8, True) for e in range(4)) > 0


def test_planted_dword_pair_stash_is_flagged():
    """The B stash written as two 4-byte stores per slot (what the compiler
    emitted while the staging arrays were only 4-byte aligned) conflicts."""
    sh = Shape(*SHAPES["128 wide, 512 threads (8x4 cells)"])
    lanes = waves(sh)[0]
    extra = 0
    for e in range(4):
        lo = [sh.b_stash(t, 0, e) for t in lanes]
        extra += conflicts(lo, 4, True) + conflicts([a + 4 for a in lo], 4, True)
    assert extra > 0
