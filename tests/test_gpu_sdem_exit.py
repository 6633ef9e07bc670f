"""GPU: the viewshed scans' early exit (a wave stops once every lane is past
its row end or has an upper bound of every later elevation angle at or below
its running maximum) gives the same maps as scanning every step.

References: the one-thread oracle on small terrains built to trigger the exit
at many places (cones, walls, pits, spikes), and the per-step kernel (option
OPT_SDEM_VARIANT = 2: quotient by the two-term reciprocal, no early exit), itself
bit-exact against the oracle, on large fBm DEMs up to the 4096-cell row limit
of the bit-block kernels.
"""
import numpy as np
import pytest

import oracle
import skewengine_amd as ske
from skewengine_amd import _lib, synth

pytestmark = pytest.mark.gpu


def _terrains(n):
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    c = (n - 1) / 2.0
    r = np.hypot(x - c, y - c)
    rng = np.random.default_rng(12)
    cone = np.maximum(0.0, 60.0 - 0.8 * r)
    walls = np.where((np.abs(x - 0.3 * n) < 2) | (np.abs(y - 0.7 * n) < 3), 80.0, 0.0) + 0.01 * x
    pit = 40.0 * (r / r.max()) ** 2
    spikes = np.where(rng.random((n, n)) < 0.002, rng.uniform(20, 120, (n, n)), 0.0)
    steps = 5.0 * np.floor(x / 17.0) - 3.0 * np.floor(y / 29.0)
    fbm = synth.dem(n, 0.7, 5, 3000.0).astype(np.float64)
    out = {"cone": cone, "walls": walls, "pit": pit, "spikes": spikes, "steps": steps, "fbm_plus_cone": fbm + cone}
    return {k: np.ascontiguousarray(v, dtype=np.float32) for k, v in out.items()}


@pytest.mark.parametrize("group", [1, 4])
@pytest.mark.parametrize("name", ["cone", "walls", "pit", "spikes", "steps", "fbm_plus_cone"])
def test_exit_terrains_bit_exact_vs_oracle(name, group):
    grid = _terrains(160)[name]
    print("terrain", name, "group", group, flush=True)
    with ske.SkewEngine(grid) as e:
        e.set_option(_lib.OPT_SDEM_GROUP, group)
        out = e.run(ske.KERNEL_SDEM, povh=ske.POVH_DEFAULT)
    ref, _ = oracle.execute(oracle.SDEM, grid, nthreads=1, povh=ske.POVH_DEFAULT)
    assert np.array_equal(out, ref), name


@pytest.mark.parametrize("n,hurst", [(2048, 0.8), (4000, 0.7)])
def test_exit_large_dem_matches_per_step_kernel(n, hurst):
    grid = synth.dem(n, hurst, 9, 3000.0)
    print("dem", n, hurst, flush=True)
    ranges = [(0, 2), (44, 47), (89, 91), (134, 136), (178, 180)]
    with ske.SkewEngine(grid) as e:
        e.set_option(_lib.OPT_SDEM_GROUP, 1)  # one chunk per wave, 128-cell guard bands
        fast = [e.run(ske.KERNEL_SDEM, s0, s1) for s0, s1 in ranges]
        e.set_option(_lib.OPT_SDEM_GROUP, 4)  # lane compaction (the default)
        packed = [e.run(ske.KERNEL_SDEM, s0, s1) for s0, s1 in ranges]
        e.set_option(_lib.OPT_SDEM_VARIANT, 2)
        full = [e.run(ske.KERNEL_SDEM, s0, s1) for s0, s1 in ranges]
    for (s0, s1), a, b, c in zip(ranges, fast, packed, full):
        assert np.array_equal(a, c), (n, s0, s1, "group 1")
        assert np.array_equal(b, c), (n, s0, s1, "group 4")


@pytest.mark.parametrize("waves,pers,occ,group", [(4, 0, 0, 1), (4, 8, 8, 1), (2, 16, 8, 1), (8, 4, 0, 1),
                                                  (4, 0, 0, 4), (4, 8, 8, 4), (2, 16, 8, 4), (8, 4, 0, 4),
                                                  (4, 16, 8, 4)])
def test_exit_batch_launch_shapes_match_per_step_kernel(waves, pers, occ, group):
    dems = [synth.dem(n, 0.75, 30 + i, 3000.0) for i, n in enumerate((1000, 1500, 700))]
    print("batch shape", waves, pers, occ, group, flush=True)
    with ske.SkewBatch(dems) as b:
        b.set_option(_lib.OPT_SDEM_WAVES, waves)
        b.set_option(_lib.OPT_SDEM_PERSISTENT, pers)
        b.set_option(_lib.OPT_SDEM_OCCUPANCY, occ)
        b.set_option(_lib.OPT_SDEM_GROUP, group)
        b.run(ske.KERNEL_SDEM)
        outs = b.outputs()
    for g, o in zip(dems, outs):
        with ske.SkewEngine(g) as e:
            e.set_option(_lib.OPT_SDEM_VARIANT, 2)
            ref = e.run(ske.KERNEL_SDEM)
        assert np.array_equal(o, ref)


@pytest.mark.parametrize("group", [1, 4])
def test_step_counter_counts_without_changing_results(group):
    grid = synth.dem(500, 0.8, 1, 1500.0)
    _, steps = ske.work(500, 500)
    with ske.SkewEngine(grid) as e:
        e.set_option(_lib.OPT_SDEM_GROUP, group)
        plain = e.run(ske.KERNEL_SDEM)
        assert e.counter("op_sdem_lane_steps") == 0.0  # not counting
        e.set_option(_lib.OPT_COUNT_STEPS, 1)
        counted = e.run(ske.KERNEL_SDEM)
        n1 = e.counter("op_sdem_lane_steps")
        e.run(ske.KERNEL_SDEM)
        n2 = e.counter("op_sdem_lane_steps")
        e.profile_reset()
        n0 = e.counter("op_sdem_lane_steps")
    assert np.array_equal(plain, counted)
    assert 0 < n1 < 1.5 * steps and n1 % 2048 == 0  # whole 32-step blocks of 64 lanes
    assert n2 == 2 * n1 and n0 == 0.0


def test_group_option_round_trip_and_errors():
    grid = synth.dem(64, 0.8, 2, 1500.0)
    with ske.SkewEngine(grid) as e:
        assert e.get_option(_lib.OPT_SDEM_GROUP) == 4  # default: lane compaction
        e.set_option(_lib.OPT_SDEM_GROUP, 1)
        assert e.get_option(_lib.OPT_SDEM_GROUP) == 1
        with pytest.raises(_lib.SkeError):
            e.set_option(_lib.OPT_SDEM_GROUP, 3)
        assert e.get_option(_lib.OPT_SDEM_GROUP) == 1
        out = e.run(ske.KERNEL_SDEM)
    ref, _ = oracle.execute(oracle.SDEM, grid, nthreads=1)
    assert np.array_equal(out, ref)
