"""CPU checks that pin the genotype oracle (decode, pack, stats, geometry, generator)."""
import numpy as np
import pytest

from oracle import geno, synth


def test_bed_lut_matches_plink_spec_every_code_every_position():
    # PLINK BED: 00 hom A1 (2 copies), 01 missing, 10 het (1), 11 hom A2 (0).
    expected = {0b00: 2, 0b01: 3, 0b10: 1, 0b11: 0}
    for pos in range(4):
        for code, value in expected.items():
            n = 4
            byte = np.uint8(code << (2 * pos))
            raw = np.array([0x6C, 0x1B, 0x01, byte], dtype=np.uint8)
            got = geno.bed_codes(raw, n_full=n, m=1)[:, 0]
            want = np.full(n, 2, np.uint8)  # other positions hold 00 -> 2
            want[pos] = value
            np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 13])
def test_bed_round_trip_and_padding(n):
    rng = np.random.default_rng(n)
    codes = rng.integers(0, 4, size=(n, 9)).astype(np.uint8)
    raw = geno.codes_to_bed(codes)
    assert raw.size == 3 + 9 * ((n + 3) // 4)
    np.testing.assert_array_equal(geno.bed_codes(raw, n, 9), codes)
    keep = np.flatnonzero(rng.random(n) < 0.6)
    np.testing.assert_array_equal(geno.bed_codes(raw, n, 9, keep), codes[keep])


def test_bed_magic_and_size_are_checked():
    with pytest.raises(ValueError):
        geno.bed_codes(np.array([0x6C, 0x1B, 0x00, 0], np.uint8), 4, 1)
    with pytest.raises(ValueError):
        geno.bed_codes(np.array([0x6C, 0x1B, 0x01], np.uint8), 4, 1)


@pytest.mark.parametrize("w", [1, 2, 3, 4, 5, 7, 8, 9])
def test_pack_every_code_every_slot(w):
    # sample i holds code (i + t) % 4 at SNP t: every code visits every bit position
    n = 8
    codes = (np.arange(n)[:, None] + np.arange(w)[None, :]) % 4
    packed = geno.pack_codes(codes)
    assert packed.shape == (n, (w + 3) // 4)
    for i in range(n):
        for t in range(w):
            assert (packed[i, t // 4] >> (2 * (t % 4))) & 3 == codes[i, t]
    # padding bits are zero (geno_stream.py:1713)
    if w % 4:
        assert np.all(packed[:, -1] >> (2 * (w % 4)) == 0)
    np.testing.assert_array_equal(geno.unpack_codes(packed, w), codes)


def test_int8_lut_and_raw_stats_follow_reference_rules():
    # missing sentinel -9; values outside {0,1,2,miss} pack to 0 but count in stats
    block = np.array([[0, 1, 2, -9, 5]], dtype=np.int8)
    lut = geno.int8_pack_lut(-9)
    np.testing.assert_array_equal(lut[block.view(np.uint8)[0]], [0, 1, 2, 3, 0])
    cnt, s1, s2 = geno.int8_stats_varmaj(block, -9)
    assert (cnt[0], s1[0], s2[0]) == (4, 8, 30)


def test_f32_stats_closed_form_and_edge_cases():
    cnt = np.array([4, 0, 5, 10], np.int64)
    s1 = np.array([4, 0, 10, 3], np.int64)      # means 1, 0, 2, 0.3
    s2 = np.array([6, 0, 20, 3], np.int64)      # var 0.5, 0, 0, 0.21
    mean, inv = geno.f32_stats(cnt, s1, s2)
    assert mean.dtype == np.float32 and inv.dtype == np.float32
    np.testing.assert_array_equal(mean[:3], np.float32([1.0, 0.0, 2.0]))
    assert inv[1] == 0 and inv[2] == 0          # cnt == 0 / var == 0 -> invalid
    np.testing.assert_allclose(inv[0], 1 / np.sqrt(0.5), rtol=1e-7)
    np.testing.assert_allclose(inv[3], 1 / np.sqrt(0.3 - 0.09), rtol=1e-5)


def _reference_fixture():
    # tests/test_sparse_prediction.py:146-174 (reference genotype fixture, data)
    train = np.array([
        [0, 0, 1, 2, 0, 1], [1, 0, 2, 1, 1, 0], [2, 1, 0, 0, 2, 1], [0, 2, 1, 1, 0, 2],
        [1, 1, 2, 0, 1, 2], [2, 2, 0, 2, 2, 0], [0, 1, 1, 0, 2, 2], [1, 2, 0, 1, 0, 1],
        [2, 0, 2, 2, 1, 0], [0, 2, 0, 1, 2, 1], [1, 0, 1, 2, 0, 2], [2, 1, 2, 0, 1, 1],
    ], dtype=np.int8)
    return train


def test_reference_fixture_standardization_matches_fp64_formula():
    # tests/test_sparse_prediction.py:130-143, 225-243: (x - mean) * inv_sd,
    # population variance, 1e-6 floor, rtol = atol = 2e-5.
    x = _reference_fixture()
    panel = geno.OraclePanel.from_codes(x.astype(np.uint8))
    mean = x.astype(np.float64).mean(axis=0)
    var = x.astype(np.float64).var(axis=0)
    inv = np.where(var > 0, 1 / np.sqrt(np.maximum(var, 1e-6)), 0)
    np.testing.assert_allclose(panel.standardized(), (x - mean) * inv, rtol=2e-5, atol=2e-5)


def test_component_partition_order_matches_reference_test():
    # tests/test_sparse_prediction.py:180-200: groups [4,0,2],[5,3,1] -> cache [0,2,4,1,3,5]
    sizes, order = geno.normalize_component_variant_indices(6, [[4, 0, 2], [5, 3, 1]])
    assert sizes == (3, 3)
    np.testing.assert_array_equal(order, [0, 2, 4, 1, 3, 5])
    with pytest.raises(ValueError):
        geno.normalize_component_variant_indices(6, [[0, 1], [1, 2]])
    with pytest.raises(IndexError):
        geno.normalize_component_variant_indices(6, [[0, 9]])


def test_call_geometry_never_straddles_components():
    starts, widths, comps, offs = geno.call_geometry(10, 3, [4, 6])
    np.testing.assert_array_equal(starts, [0, 3, 4, 7])
    np.testing.assert_array_equal(widths, [3, 1, 3, 3])
    np.testing.assert_array_equal(comps, [0, 0, 1, 1])
    np.testing.assert_array_equal(offs, [0, 2, 4])


def test_philox_known_answer_vectors():
    # Random123 philox4x32_10 known-answer tests (kat_vectors)
    def run(c, k):
        out = synth.philox4x32(*[np.array([x], np.uint32) for x in c], *k)
        return [int(o[0]) for o in out]

    assert run((0, 0, 0, 0), (0, 0)) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert run((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF)) == [0x408F276D, 0x41C83B0E,
                                                                0xA20BC7C6, 0x6D5451FD]
    assert run((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0)) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_generator_hwe_frequencies_and_missing_rate():
    p = np.array([0.05, 0.2, 0.5])
    t0, t1, tm = synth.thresholds(p, 0.01)
    codes = synth.genotype_codes(7, 40000, np.arange(3), t0, t1, tm)
    obs = codes != 3
    assert abs(1 - obs.mean() - 0.01) < 0.002
    for j, pj in enumerate(p):
        g = codes[obs[:, j], j]
        for val, prob in ((0, (1 - pj) ** 2), (1, 2 * pj * (1 - pj)), (2, pj * pj)):
            assert abs((g == val).mean() - prob) < 0.01
