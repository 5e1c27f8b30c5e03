"""FrameRenderer's cut of a frame into lanes (render.lane_plan): whole pixels, equal pixel counts up
to one, covering the frame; the automatic lane count; the paths per launch sequence."""
import numpy as np

from gonzales_amd.render import lane_plan


def test_automatic_lanes_and_sequence_size():
    # the Cornell frame (512^2 x 16 spp): two lanes of 2,097,152 paths
    assert lane_plan(512 * 512 * 16, 16) == ([0, 2097152, 4194304], 2097152)
    # 256^2 x 16 spp, below 4M paths: one lane, one sequence of the whole frame
    assert lane_plan(256 * 256 * 16, 16) == ([0, 1048576], 1048576)
    # the 5M-triangle frame (1920 x 1080 x 64 spp) with 2^28 paths of device state: two halves,
    # each one launch sequence
    assert lane_plan(1920 * 1080 * 64, 64, 0, 1 << 28) == ([0, 66355200, 132710400], 66355200)
    # the default 2^25 paths of device state: 2^24 per lane
    assert lane_plan(1920 * 1080 * 64, 64)[1] == 1 << 24


def test_explicit_lanes_small_frames_and_sequences():
    assert lane_plan(10 * 4, 4, 3) == ([0, 12, 24, 40], 16)
    assert lane_plan(2 * 4, 4, 5) == ([0, 4, 8], 4)  # never more lanes than pixels
    assert lane_plan(0, 4) == ([0, 0], 4)
    # 192 paths of device state over two lanes of 16-spp pixels: 96 paths, 6 whole pixels, per sequence
    assert lane_plan(40 * 40 * 16, 16, 2, 192) == ([0, 12800, 25600], 96)
    assert lane_plan(40 * 40 * 16, 16, 2, 200) == ([0, 12800, 25600], 96)  # 100 paths: still 6 pixels
    # a sequence is at least one pixel
    assert lane_plan(40 * 40 * 64, 64, 2, 8)[1] == 64


def test_bounds_properties_on_random_frames():
    rng = np.random.default_rng(11)
    for _ in range(500):
        spp = int(rng.choice([1, 2, 4, 16, 64]))
        pixels = int(rng.integers(1, 5000))
        lanes = int(rng.integers(0, 9))
        batch = int(rng.integers(1, 1 << 20))
        bounds, per = lane_plan(pixels * spp, spp, lanes, batch)
        assert bounds[0] == 0 and bounds[-1] == pixels * spp
        assert all(b % spp == 0 for b in bounds)
        sizes = np.diff(bounds) // spp
        assert sizes.min() >= 1 and sizes.max() - sizes.min() <= 1
        assert len(sizes) == (min(lanes, pixels) if lanes > 0 else 1)  # these frames stay below 4M paths
        assert per % spp == 0 and spp <= per <= max(spp, sizes.max() * spp)
