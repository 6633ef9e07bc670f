"""Batched runs (many grids of uneven sizes in one call): every grid's result
must equal the single-grid path and the one-thread oracle, for every launch
shape the autotuner may pick, across arena-chunk splits, and on every replay
of a captured hipGraph."""
import numpy as np
import pytest

import oracle
import skewengine_amd as ske
from skewengine_amd import _lib, synth

pytestmark = pytest.mark.gpu

SIDES = [64, 96, 128, 64, 72]  # 72: a partial last block of 16 deskew rows


@pytest.fixture(scope="module")
def dems():
    return [synth.dem(n, 0.7 + 0.05 * i, 100 + i, 3000.0) for i, n in enumerate(SIDES)]


@pytest.fixture(scope="module")
def refs(dems):
    return [oracle.execute(oracle.SDEM, g, nthreads=1)[0] for g in dems]


def test_batch_viewshed_matches_oracle(dems, refs):
    with ske.SkewBatch(dems) as b:
        b.run(ske.KERNEL_SDEM)
        outs = b.outputs()
    for o, r in zip(outs, refs):
        assert np.array_equal(o, r)


@pytest.mark.parametrize("waves,pers,occ", [(2, 0, 0), (4, 0, 0), (8, 0, 0), (4, 4, 0), (2, 16, 8), (8, 8, 8),
                                            (4, 8, 8)])
def test_batch_launch_shapes_bit_exact(dems, refs, waves, pers, occ):
    with ske.SkewBatch(dems) as b:
        b.set_option(_lib.OPT_SDEM_WAVES, waves)
        b.set_option(_lib.OPT_SDEM_PERSISTENT, pers)
        b.set_option(_lib.OPT_SDEM_OCCUPANCY, occ)
        b.run(ske.KERNEL_SDEM)
        outs = b.outputs()
    for o, r in zip(outs, refs):
        assert np.array_equal(o, r)


def test_batch_chunk_splits(dems, refs):
    with ske.SkewBatch(dems) as b:
        b.set_option(_lib.OPT_ARENA_BYTES, 1 << 20)  # grids split across many chunks
        b.set_option(_lib.OPT_SDEM_PERSISTENT, 4)
        b.run(ske.KERNEL_SDEM)
        outs = b.outputs()
    for o, r in zip(outs, refs):
        assert np.array_equal(o, r)


def test_batch_graph_replays_match_eager(dems, refs):
    with ske.SkewBatch(dems) as b:
        b.set_option(_lib.OPT_SDEM_PERSISTENT, 8)
        b.run(ske.KERNEL_SDEM)
        eager = b.outputs()
        b.set_option(_lib.OPT_GRAPH, 1)
        for rep in range(4):
            b.run(ske.KERNEL_SDEM)
            outs = b.outputs()
            for o, e in zip(outs, eager):
                assert np.array_equal(o, e), rep
            # eager work between replays: a different operator on the same batch
            b.set_option(_lib.OPT_GRAPH, 0)
            b.run(ske.KERNEL_IDENTITYDEM)
            b.synchronize()
            b.set_option(_lib.OPT_GRAPH, 1)
    for o, r in zip(eager, refs):
        assert np.array_equal(o, r)


def test_batch_other_operators_match_single(dems):
    imgs = [synth.blurred_texture(n=n, seed=7 + n)[0] for n in (64, 96)]
    for kernel, grids in ((ske.KERNEL_IDENTITYDEM, dems), (ske.KERNEL_RADON, imgs), (ske.KERNEL_CEPSTRUM, imgs)):
        with ske.SkewBatch(grids) as b:
            b.run(kernel)
            outs = b.outputs()
        for g, o in zip(grids, outs):
            with ske.SkewEngine(g) as e:
                single = e.run(kernel)
            assert np.array_equal(o, single), kernel


def test_autotune_keeps_results(dems, refs):
    with ske.SkewBatch(dems) as b:
        waves, pers, occ, ms = b.autotune()
        assert waves in (2, 4, 8) and 0 <= pers <= 64 and occ in (0, 8) and ms > 0
        b.run(ske.KERNEL_SDEM)
        outs = b.outputs()
    for o, r in zip(outs, refs):
        assert np.array_equal(o, r)
