"""Launch orders of a tone batch (host logic, no GPU): ToneBatch.longest_first_order
and ToneBatch.tone_order on the block table of a 200 ms tone batch.  The
kernels give identical results in any order (tests/test_gpu_warmstart.py,
test_lockstep_grouping_does_not_change_results); these checks pin what the
orders are: permutations of the table, longest block first, and in tone order
the higher-frequency tone first within a length."""
import math
from types import SimpleNamespace

import numpy as np

from cochlearmodel_amd import batch, solver

FS = 44100.0


def _stub(n_tones, seed):
    """The attributes of a ToneBatch the two orders read, built as ToneBatch
    builds them (without device buffers)."""
    rng = np.random.default_rng(seed)
    interval = int(round(FS * 0.2))
    block = int(math.ceil(np.float32(0.02) * np.float32(FS)))
    DF = interval // block
    TB = n_tones * DF
    ov = 220
    bufP1 = ov + block * (TB + 1)
    return SimpleNamespace(items=solver.block_table(TB, bufP1, ov, DF), TB=TB, DF=DF,
                           freqs=np.exp(rng.uniform(np.log(250.0), np.log(8000.0), n_tones)))


def test_longest_first_order():
    tb = _stub(37, 1)
    o = batch.ToneBatch.longest_first_order(tb)
    lens = tb.items["mesi"].astype(np.int64) - tb.items["start"]
    assert sorted(o.tolist()) == list(range(tb.TB))
    assert np.all(np.diff(lens[o]) <= 0)
    # every tone's last block ends at the tone's boundary without the overlap
    # run-out, so it is the short one (the table's final block keeps it)
    last = np.arange(tb.DF - 1, tb.TB - 1, tb.DF)
    assert np.all(lens[last] < lens.max())
    assert np.all(np.delete(lens, last) == lens.max())
    # ties in table order
    for L in np.unique(lens):
        sel = o[lens[o] == L]
        assert np.all(np.diff(sel) > 0)


def test_tone_order():
    tb = _stub(37, 2)
    o = batch.ToneBatch.tone_order(tb)
    lens = tb.items["mesi"].astype(np.int64) - tb.items["start"]
    tone = np.arange(tb.TB) // tb.DF
    assert o.dtype == np.int32
    assert sorted(o.tolist()) == list(range(tb.TB))
    assert np.all(np.diff(lens[o]) <= 0)
    for L in np.unique(lens):
        sel = o[lens[o] == L]
        f = tb.freqs[tone[sel]]
        assert np.all(np.diff(f) <= 0)            # higher frequency first
        same = np.diff(f) == 0                     # blocks of one tone: table order
        assert np.all(np.diff(sel)[same] > 0)
    # a tone's long blocks stay together, so a group of 16 spans at most 3 tones
    long_part = o[lens[o] == lens.max()]
    for g in range(0, len(long_part), 16):
        assert len(np.unique(tone[long_part[g:g + 16]])) <= 3


def test_tone_order_low_first():
    tb = _stub(29, 3)
    o = batch.ToneBatch.tone_order(tb, high_first=False)
    hi = batch.ToneBatch.tone_order(tb)
    lens = tb.items["mesi"].astype(np.int64) - tb.items["start"]
    tone = np.arange(tb.TB) // tb.DF
    assert sorted(o.tolist()) == list(range(tb.TB))
    assert np.all(np.diff(lens[o]) <= 0)
    for L in np.unique(lens):
        sel = o[lens[o] == L]
        f = tb.freqs[tone[sel]]
        assert np.all(np.diff(f) >= 0)            # lower frequency first
        # the same tones as high-first, in reverse, each tone's blocks in table order
        rev = hi[lens[hi] == L]
        assert np.array_equal(np.unique(tone[sel]), np.unique(tone[rev]))
        assert np.all(np.diff(sel)[np.diff(f) == 0] > 0)
