"""Pins the CPU oracle's Z-Sobol sampler (restating ZSobolSampler.swift and SobolGenerator).

Checks: the reference's known answers (SobolSamplerTests.swift:9-140), the generator matrices
against SciPy's independent Joe-Kuo implementation, the Owen-scrambling structure (output prefix
depends only on the input prefix and permutes it), per-pixel stratification of Z-Sobol samples,
bijectivity of getSampleIndex, the dimension-wrap rule, and the committed golden vectors.
"""
import ctypes
from pathlib import Path

import numpy as np
import pytest

import oracle

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def table():
    return oracle.sobol_matrices()


def test_reverse_bits_known_answers():
    L = oracle.lib()
    assert L.orc_reverse_bits32(0) == 0
    assert L.orc_reverse_bits32(0xFFFFFFFF) == 0xFFFFFFFF
    assert L.orc_reverse_bits32(0x80000000) == 1
    for v in (0, 1, 42, 0xDEADBEEF, 0x12345678, 0xFFFFFFFF):
        assert L.orc_reverse_bits32(L.orc_reverse_bits32(v)) == v
        assert L.orc_reverse_bits32(v) == int(f"{v:032b}"[::-1], 2)


def test_morton_known_answers():
    L = oracle.lib()
    assert [L.orc_encode_morton2(x, y) for x, y in ((0, 0), (1, 0), (0, 1), (1, 1))] == [0, 1, 2, 3]
    for x, y in ((5, 9), (1919, 1079), (0xFFFF, 0), (12345, 54321)):
        m = L.orc_encode_morton2(x, y)
        bits = sum(((x >> i) & 1) << (2 * i) | ((y >> i) & 1) << (2 * i + 1) for i in range(32))
        assert m == bits


def test_next_power_of_two_known_answers():
    L = oracle.lib()
    assert [L.orc_next_pow2(v) for v in (16, 1, 1024, 3, 5, 17, 100)] == [16, 1, 1024, 4, 8, 32, 128]


def test_mix_bits_and_owen_determinism():
    L = oracle.lib()
    assert L.orc_mix_bits(12345) == L.orc_mix_bits(12345)
    assert len({L.orc_mix_bits(0), L.orc_mix_bits(1), L.orc_mix_bits(99999)}) == 3
    assert L.orc_owen(12345, 42) == L.orc_owen(12345, 42)
    assert L.orc_owen(12345, 1) != L.orc_owen(12345, 2)
    # mixBits folds the high word into the low one before hashing (a 32-bit result).
    assert L.orc_mix_bits((7 << 32) | 5) == L.orc_mix_bits(7 ^ 5)


def test_table_matches_scipy_joe_kuo(table):
    import scipy.stats._sobol as sob
    from scipy.stats import qmc

    qmc.Sobol(d=2, scramble=False)  # loads SciPy's direction numbers
    v = np.zeros((1024, 64), dtype=np.uint64)
    sob._initialize_v(v, 1024, 64)
    ref = (v[:, :52] >> np.uint64(32)).astype(np.uint32)
    assert np.array_equal(table.reshape(1024, 52), ref)


def test_direction_number_file_equals_the_reference_table():
    # The committed data file (written from SciPy's copy of Joe & Kuo's table) holds exactly the
    # first 1024 lines of the reference's Resources/new-joe-kuo-6.21201.
    ref = Path("/root/reference/Sources/libgonzales/Resources/new-joe-kuo-6.21201")
    if not ref.exists():
        pytest.skip("reference tree not mounted")
    ours = oracle.JOE_KUO.read_text().split("\n")
    theirs = ref.read_text().split("\n")[:1024]
    assert [l.split() for l in ours[:1024]] == [l.split() for l in theirs]


def test_dimension_zero_is_van_der_corput(table):
    t = table.reshape(1024, 52)
    assert all(int(t[0, i]) == 1 << (31 - i) for i in range(32)) and not t[0, 32:].any()


def test_owen_scrambler_is_a_nested_permutation():
    # Output bits above k depend only on input bits above k and permute them (Owen scrambling).
    L = oracle.lib()
    g = np.random.default_rng(0)
    for seed in (0, 1, 0xDEADBEEF, int(g.integers(0, 2**32))):
        for k in (1, 2, 5, 8):
            prefixes = {}
            for p in range(1 << k):
                outs = {L.orc_owen((p << (32 - k)) | int(low), seed) >> (32 - k)
                        for low in g.integers(0, 1 << (32 - k), 4)}
                assert len(outs) == 1
                prefixes[p] = outs.pop()
            assert sorted(prefixes.values()) == list(range(1 << k))


def test_sampler_values_in_range(table):
    # SobolSamplerTests.swift:96-140 (spp 4, 64x64, seed 0).
    for pat, px in (([1] * 20, 10), ([2] * 10, 5), ([3] * 10, 5)):
        v = oracle.zsobol_batch(4, 64, 64, 0, [px], [px], [0], pat, table=table)
        assert (v >= 0).all() and (v < 1).all()


@pytest.mark.parametrize("spp,seed", [(16, 0), (16, 1234), (64, 7), (4, 3)])
def test_pixel_samples_are_stratified(table, spp, seed):
    m = int(np.log2(spp))
    for px, py in ((0, 0), (5, 5), (10, 10), (1919, 1079)):
        v = oracle.zsobol_batch(spp, 1920, 1080, seed, [px] * spp, [py] * spp, np.arange(spp), [2] + [1] * 6,
                                table=table)
        # every 1-D projection: one sample per interval of width 1/spp
        for d in range(v.shape[1]):
            assert sorted(np.floor(v[:, d] * spp).astype(int).tolist()) == list(range(spp))
        # dims 0, 1 (van der Corput + Pascal, a (0,2)-sequence): a (0, m, 2)-net
        for a in range(m + 1):
            cells = np.floor(v[:, 0] * (1 << a)).astype(int) * (1 << (m - a)) + np.floor(
                v[:, 1] * (1 << (m - a))).astype(int)
            assert len(set(cells.tolist())) == spp


@pytest.mark.parametrize("spp", [1, 2, 4, 8, 16])
def test_sample_index_is_a_bijection(spp):
    L = oracle.lib()
    st = oracle.ZSobolState()
    L.orc_zsobol_init(ctypes.byref(st), spp, 8, 8, 0)
    for dim in (0, 3, 17):
        seen = set()
        for py in range(8):
            for px in range(8):
                for i in range(spp):
                    L.orc_zsobol_start(ctypes.byref(st), px, py, i, dim)
                    seen.add(L.orc_zsobol_sample_index(ctypes.byref(st)))
        assert seen == set(range(64 * spp))


def test_dimension_wraps_to_two(table):
    # get2D at dim 1023: index and hash come from dim 1023, values from dims 2 and 3
    # (ZSobolSampler.swift:221-244).
    L = oracle.lib()
    st = oracle.ZSobolState()
    L.orc_zsobol_init(ctypes.byref(st), 16, 64, 64, 5)
    L.orc_zsobol_start(ctypes.byref(st), 3, 4, 7, 1023)
    idx = L.orc_zsobol_sample_index(ctypes.byref(st))
    out = np.zeros(2, np.float32)
    L.orc_zsobol_get2d(ctypes.byref(st), table.ctypes.data, out.ctypes.data)
    assert st.dim == 4
    h = L.orc_mix_bits(1023 ^ 5)
    assert out[0] == np.float32(L.orc_sobol_sample(idx, 2, h, table.ctypes.data))
    assert out[1] == np.float32(L.orc_sobol_sample(idx, 3, 0, table.ctypes.data))


def test_golden_vectors(table):
    data = np.load(GOLDEN / "zsobol_vectors.npz")
    for name in sorted({k.rsplit("_", 1)[0] for k in data.files}):
        args = data[name + "_args"]
        spp, rx, ry, seed, start = (int(x) for x in args[:5])
        pat = data[name + "_pattern"]
        px, py, si = data[name + "_px"], data[name + "_py"], data[name + "_si"]
        got = oracle.zsobol_batch(spp, rx, ry, seed, px, py, si, pat, start_dim=start, table=table)
        assert np.array_equal(got.view(np.uint32), data[name + "_out"].view(np.uint32)), name
