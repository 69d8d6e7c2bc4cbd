"""Batched placement on the GPU: blocks run together (grouped launches, one graph per run of passes) end exactly
where their own device loops end (bit for bit), whatever tile layouts the autotuner picks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def states():
    from autoplace_amd import batch
    blocks = batch.make_blocks(6, 2000, 24000, seed=3)
    grids = sorted({batch.block_grid(b.num_mov) for b in blocks})
    print("blocks: cells %s, grids %s" % ([b.num_mov for b in blocks], grids))
    assert len(grids) >= 2
    return [batch.start_state(b) for b in blocks]


def _same(a, b):
    """Every block's state and placement bit for bit -- but for `accept` / `is_best`, the flags of the last pass,
    which every pass after the stop resets (so they depend on how many passes ran after it)."""
    from autoplace_amd.gpu_loop import FIELDS
    for k, (ra, rb, pa, pb) in enumerate(zip(a["blocks"], b["blocks"], a["positions"], b["positions"])):
        diff = [(FIELDS[i], x, y) for i, (x, y) in enumerate(zip(ra["state"], rb["state"]))
                if x != y and FIELDS[i] not in ("accept", "is_best")]
        assert not diff, (k, diff)
        assert np.array_equal(pa, pb), (k, float(np.abs(pa - pb).max()))


def test_batch_equals_sequential_runs(states):
    from autoplace_amd import batch
    seq = batch.run_sequential(states)
    bat = batch.run_batched(states)
    print("sequential: %.3f s, batched: %.3f s; iterations %s; stops %s; tile layouts %s"
          % (seq["seconds"], bat["seconds"], [r["iters"] for r in seq["blocks"]],
             [r["stop"] for r in seq["blocks"]], bat["tile_layout"]))
    assert all(r["stop"] == "converged" for r in seq["blocks"])
    _same(seq, bat)


def test_autotuned_batch_gives_the_same_results(states):
    from autoplace_amd import batch
    plain = batch.run_batched(states)
    tuned = batch.run_batched(states, autotune=True, reps=2)
    print("tile layouts: default %s, autotuned %s" % (plain["tile_layout"], tuned["tile_layout"]))
    _same(plain, tuned)
