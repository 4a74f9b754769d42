"""Host-side bounds check of the batched launches of the skipped polled-batch
case (the five scenes of test_batch_matches_single_scenes_bitwise): every
block of a batched launch is mapped to its scene and scene-local block as
the device does (batch_scene / batch_block, rheon_amd/csrc/common.hpp:380-393),
every lane's four-cell chunk to its row, plane and cell (FOR_QUADS with the
multiply-high division, rheon_amd/csrc/mg_quad.hpp:16-28, common.hpp:44-62),
and every address the four-cell outer residual and finest-level sweeps read
or write (load_quad, load4/store4, the mask word, mg_quad.hpp:30-76;
b_mg_outer_residual_q, mg_quad_kernels.hip:26-76) and every partial slot
they publish (block_publish, common.hpp:238-249) is checked against the
buffer it lands in.  The block counts are the host's (quad_blocks,
rheon_amd/csrc/runtime.hip:41-45; Batch::build's qblk,
batch.hip:563-566), and the polled schedule's slices of the program
(Batch::run_polled, batch.hip:1016-1052) are checked to cover it in order
for every cycle at which the host may stop enqueuing."""
import numpy as np
import pytest

K_BLOCK = 256          # kLanesX * kRowsPerBlock
K_GROUPS = 2           # kGroupsPerBlock
K_MAX_BLOCKS = 16384   # kMaxBlocks: the partial arrays hold this many slots per quantity
SHAPES = [(32, 32, 32), (40, 24, 16), (64, 32, 48), (48, 48, 48), (32, 64, 40)]


def clamp_blocks(blocks):
    return int(min(max(blocks, 1), K_MAX_BLOCKS))


def quad_blocks(quads_per_row, rows):
    return clamp_blocks(-(-(quads_per_row * rows) // (K_GROUPS * K_BLOCK)))


def make_fast_div(d):
    l = 0
    while (1 << l) < d:
        l += 1
    return ((1 << 32) * ((1 << l) - d)) // d + 1, l


def fast_div(n, f):
    m, l = f
    n = n.astype(np.uint64)
    return ((((np.uint64(m) * n) >> np.uint64(32)) + n) >> np.uint64(l)).astype(np.int64)


def batch_scene(prefix, n, b):
    lo, hi = 0, n
    while hi - lo > 1:
        mid = (lo + hi) >> 1
        if prefix[mid] <= b:
            lo = mid
        else:
            hi = mid
    return lo


def block_span(units, b, nb):
    return b * units // nb, (b + 1) * units // nb


def check_quad_launch(shapes, block_div=1):
    """One batched four-cell launch over `shapes` (the outer residual, or a
    finest-level sweep: both iterate FOR_QUADS over the grid's cells)."""
    n = len(shapes)
    blocks = [max(1, quad_blocks(nx // 4, ny * nz) // block_div) for nx, ny, nz in shapes]
    prefix = np.concatenate([[0], np.cumsum(blocks)]).astype(np.int64)
    total = int(prefix[-1])
    covered = [np.zeros(nx // 4 * ny * nz, np.int32) for nx, ny, nz in shapes]
    published = [set() for _ in shapes]
    tid = np.arange(K_BLOCK)
    for b in range(total):
        s = batch_scene(prefix, n, b)
        assert prefix[s] <= b < prefix[s + 1], (b, s)
        vb, nb = b - int(prefix[s]), int(prefix[s + 1] - prefix[s])
        nx, ny, nz = shapes[s]
        cells, cpr = nx * ny * nz, nx // 4
        begin, end = block_span(cpr * ny * nz, vb, nb)
        # every lane's chunks: t = begin + tid, + kBlock, ... < end
        t = (begin + tid[None, :] + K_BLOCK * np.arange(max(1, -(-(end - begin) // K_BLOCK)))[:, None]).ravel()
        t = t[t < end]
        row = fast_div(t, make_fast_div(cpr))
        assert np.array_equal(row, t // cpr)
        q = t - row * cpr
        k = fast_div(row, make_fast_div(ny))
        assert np.array_equal(k, row // ny)
        j = row - k * ny
        i0 = 4 * q
        c = i0 + nx * row
        np.add.at(covered[s], t, 1)
        # four-cell vectors at c (p, b, b0, x0; the mask word) and the neighbours
        addr = [c, np.where(j > 0, c - nx, c), np.where(j + 1 < ny, c + nx, c),
                np.where(k > 0, c - nx * ny, c), np.where(k + 1 < nz, c + nx * ny, c)]
        for a in addr:
            assert np.all(a >= 0) and np.all(a + 3 < cells), (shapes[s], vb)
            assert np.all(a % 4 == 0)  # 16-byte float4 / 32-byte double pair alignment
        left = np.where(i0 > 0, c - 1, c)
        right = np.where(i0 + 4 < nx, c + 4, c + 3)
        assert np.all(left >= 0) and np.all(right < cells)
        assert vb < K_MAX_BLOCKS
        published[s].add(vb)
    for s, (nx, ny, nz) in enumerate(shapes):
        assert np.all(covered[s] == 1), shapes[s]  # every chunk exactly once: no write overlaps another block's
        assert published[s] == set(range(blocks[s]))  # the finisher reads exactly these partials
    return blocks


@pytest.mark.parametrize("block_div", [1, 2, 3, 8])
def test_batched_four_cell_launches_stay_in_bounds(block_div):
    blocks = check_quad_launch(SHAPES, block_div)
    assert blocks[0] == max(1, 16 // block_div)  # 32^3: 8192 chunks, 512 per block


def test_batched_launches_with_idle_scenes():
    """Scenes that take no part in a stage have no blocks (prefix[s] ==
    prefix[s + 1]); the search still lands every block in a scene that has it."""
    prefix = np.array([0, 0, 5, 5, 9, 9], np.int64)
    for b in range(9):
        s = batch_scene(prefix, 5, b)
        assert prefix[s] <= b < prefix[s + 1]


def test_fast_division_covers_every_index_range():
    """The multiply-high division equals floor division for every divisor a
    level can use up to 4096 and numerators up to the 32-bit index bound."""
    rng = np.random.default_rng(0)
    n = np.concatenate([np.arange(0, 1 << 16), rng.integers(0, (1 << 31) - 1, 1 << 16),
                        np.array([(1 << 31) - 1, (1 << 31) - 2])]).astype(np.int64)
    for d in range(1, 4097):
        assert np.array_equal(fast_div(n, make_fast_div(d)), n // d), d


def polled_slices(program, cycle_at, post_at, max_cycles, stop_after, poll=2):
    """Batch::run_polled's enqueued ranges when the host stops after the
    batch `stop_after` (None: never stops early)."""
    out = [(0, cycle_at[1] if max_cycles >= 1 else post_at)]
    it, pending, b = 1, -1, 0
    while it <= max_cycles:
        last = min(max_cycles, it + poll - 1)
        out.append((cycle_at[it], cycle_at[last + 1] if last < max_cycles else post_at))
        it = last + 1
        cur = b & 1
        if pending >= 0 and stop_after is not None and b - 1 >= stop_after:
            break
        pending = cur
        b += 1
    out.append((post_at, program))
    return out


@pytest.mark.parametrize("max_cycles", [1, 2, 3, 20, 30])
def test_polled_schedule_covers_the_program_in_order(max_cycles):
    prologue, cycle_len, outer_len, tail = 9, 37, 3, 14
    cycle_at = [0] * (max_cycles + 1)
    pos = prologue + outer_len
    for it in range(1, max_cycles + 1):
        cycle_at[it] = pos
        pos += cycle_len + outer_len
    post_at, program = pos, pos + tail
    for stop in [None] + list(range(max_cycles)):
        sl = polled_slices(program, cycle_at, post_at, max_cycles, stop)
        assert sl[0][0] == 0 and sl[-1] == (post_at, program)
        for (a0, a1), (b0, b1) in zip(sl[:-2], sl[1:-1]):
            assert a1 == b0  # contiguous up to the stop
        assert all(0 <= a <= b <= program for a, b in sl)
        stopped_at = sl[-2][1]
        assert stopped_at == post_at or stopped_at in cycle_at[1:]  # stops only on a cycle boundary
        if stop is None:
            assert stopped_at == post_at
