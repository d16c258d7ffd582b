"""GPU: resident panel ingest paths are bit-exact against the oracle."""
import numpy as np
import pytest

from oracle import geno, synth
from tests.helpers import cache_codes, int8_block, random_codes

pytestmark = pytest.mark.gpu


def _streamer(codes, **kw):
    from coherit_amd.geno_stream import GenoBlockStreamer
    from coherit_amd.sources import ArraySource

    return GenoBlockStreamer(ArraySource(int8_block(codes), missing_val=-9), **kw)


def _check_packed(st, cache):
    for c in range(st._n_calls):
        s0 = int(st._call_snp_starts[c])
        w = int(st._call_true_widths[c])
        got = st._pop_cached(c).cpu().numpy()
        np.testing.assert_array_equal(got, geno.pack_codes(cache[:, s0 : s0 + w]),
                                      err_msg=f"call {c}")


def _check_stats(st, cache):
    o = geno.OraclePanel.from_codes(cache, st._component_block_sizes)
    mean, inv, cnt, s1, s2 = st.panel.stats()
    cnt_o, s1_o, s2_o = geno.code_stats(cache.T)
    np.testing.assert_array_equal(cnt, cnt_o)
    np.testing.assert_array_equal(s1, s1_o)
    np.testing.assert_array_equal(s2, s2_o)
    # bit-exact float32 epilogue (geno_stream.py:1780-1792)
    np.testing.assert_array_equal(mean.view(np.uint32), o.mean.view(np.uint32))
    np.testing.assert_array_equal(inv.view(np.uint32), o.inv_sd.view(np.uint32))
    assert st._eff_m_host == o.atoms.eff
    assert st._mean_diag_atom_host == pytest.approx(o.atoms.mean_diag, rel=1e-12)
    np.testing.assert_array_equal(st._component_eff_m_host, o.atoms.component_eff)
    np.testing.assert_allclose(st._component_mean_diag_atoms_host, o.atoms.component_mean_diag,
                               rtol=1e-6)


@pytest.mark.parametrize("n,m,cw", [(1003, 777, 100), (4, 9, 3), (257, 1030, 1024), (130, 64, 131072)])
def test_int8_ingest_packs_and_stats_bit_exact(n, m, cw):
    print(f"case n={n} m={m} call_width={cw}", flush=True)
    codes = random_codes(n, m, seed=n + m)
    st = _streamer(codes, call_width=cw)
    _check_packed(st, codes)
    _check_stats(st, codes)
    st.close()


def test_partitioned_block_sizes_and_variant_indices():
    codes = random_codes(301, 500, seed=5)
    st = _streamer(codes, call_width=64, component_block_sizes=[123, 200, 177])
    _check_packed(st, codes)
    _check_stats(st, codes)
    st.close()
    rng = np.random.default_rng(3)
    perm = rng.permutation(500)
    groups = [perm[:171], perm[171:]]
    st = _streamer(codes, call_width=50, component_variant_indices=groups)
    cache = cache_codes(codes, groups)
    _check_packed(st, cache)
    _check_stats(st, cache)
    st.close()


def test_degenerate_snps_constant_and_all_missing():
    codes = random_codes(64, 12, seed=9)
    codes[:, 0] = 2          # constant -> var 0 -> inv 0
    codes[:, 1] = 3          # all missing -> cnt 0
    codes[:, 2] = 0
    codes[5, 2] = 1          # rare variant
    st = _streamer(codes, call_width=8)
    _check_stats(st, codes)
    assert st._inv_sds_host[0] == 0 and st._inv_sds_host[1] == 0 and st._inv_sds_host[2] > 0
    st.close()


def test_bed_ingest_with_sample_mask(tmp_path):
    from coherit_amd.geno_stream import GenoBlockStreamer
    from coherit_amd.sources import write_bed

    codes = random_codes(203, 150, seed=11)
    prefix = str(tmp_path / "panel")
    write_bed(prefix, codes)
    raw = np.fromfile(prefix + ".bed", dtype=np.uint8)
    np.testing.assert_array_equal(geno.bed_codes(raw, 203, 150), codes)
    mask = np.random.default_rng(1).random(203) < 0.7
    st = GenoBlockStreamer(bed_prefix=prefix, call_width=40, sample_mask=mask)
    kept = codes[mask]
    assert st.n == int(mask.sum())
    _check_packed(st, kept)
    _check_stats(st, kept)
    st.close()


def test_packed_call_upload_round_trip():
    from coherit_amd.panel import ResidentPanel
    import torch

    codes = random_codes(99, 70, seed=2)
    starts, widths, comps, _ = geno.call_geometry(70, 30, None)
    p = ResidentPanel(99, starts, widths, comps, 1)
    for c in range(len(starts)):
        blk = geno.pack_codes(codes[:, starts[c] : starts[c] + widths[c]])
        p.ingest_packed_call(c, torch.as_tensor(blk, device="cuda"))
    p.finalize()
    mean, inv, cnt, s1, s2 = p.stats()
    o = geno.OraclePanel.from_codes(codes)
    np.testing.assert_array_equal(inv.view(np.uint32), o.inv_sd.view(np.uint32))
    for c in range(len(starts)):
        np.testing.assert_array_equal(p.read_packed_call(c).cpu().numpy(),
                                      geno.pack_codes(codes[:, starts[c] : starts[c] + widths[c]]))
    p.close()


def test_device_generator_matches_cpu_generator():
    from coherit_amd.geno_stream import GenoBlockStreamer
    from coherit_amd.synth import SyntheticSource

    n, m, seed = 517, 203, 20261203
    p = np.random.default_rng(0).uniform(0.05, 0.5, m)
    src = SyntheticSource(n, p, seed)
    groups = [np.flatnonzero(p < 0.2), np.flatnonzero(p >= 0.2)]
    st = GenoBlockStreamer(src, call_width=64, component_variant_indices=groups)
    t0, t1, tm = synth.thresholds(p, 0.01)
    codes = synth.genotype_codes(seed, n, np.arange(m), t0, t1, tm)
    cache = cache_codes(codes, groups)
    _check_packed(st, cache)
    _check_stats(st, cache)
    st.close()


def test_standardization_override_and_extract_columns():
    codes = random_codes(77, 40, seed=4)
    o = geno.OraclePanel.from_codes(codes)
    mo = (o.mean + 0.125).astype(np.float32)
    io = (o.inv_sd * 0.5).astype(np.float32)
    st = _streamer(codes, call_width=16, standardization_override=(mo, io))
    np.testing.assert_array_equal(st._means_host, mo)
    np.testing.assert_array_equal(st._inv_sds_host, io)
    ov = geno.OraclePanel.from_codes(codes, override=(mo, io))
    assert st._mean_diag_atom_host == pytest.approx(ov.atoms.mean_diag, rel=1e-12)
    cols = np.array([3, 0, 39, 17, 3])
    got = st.extract_standardized_columns(cols)
    np.testing.assert_allclose(got, ov.standardized()[:, cols], rtol=2e-5, atol=2e-5)
    st.close()
