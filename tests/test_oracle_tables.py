"""The oracle's closed-form tables against the reference's constants
(tests/golden/reference_tables.json, extracted from src/TileProcessor.cuh and
src/dtt8x8.cuh) and the 8-point transforms against scipy."""
import numpy as np
import scipy.fft as sf


def test_dtt_matrices_match_scipy(orc):
    M = orc.table(0)
    E = np.eye(8)
    # columns = transform of unit vectors
    dct4 = sf.dct(E, type=4, norm="ortho", axis=0)
    dst4 = sf.dst(E, type=4, norm="ortho", axis=0)
    np.testing.assert_allclose(M[0], dct4, atol=1e-14)
    np.testing.assert_allclose(M[1], dst4, atol=1e-14)
    # DCT-II / DST-II: orthonormal with the first (DCT) / last (DST) row
    # multiplied by sqrt(2) ("nonortho", dtt8x8.cuh:382)
    dct2 = sf.dct(E, type=2, norm="ortho", axis=0)
    dct2[0] *= np.sqrt(2)
    dst2 = sf.dst(E, type=2, norm="ortho", axis=0)
    dst2[7] *= np.sqrt(2)
    np.testing.assert_allclose(M[2], dct2, atol=1e-14)
    np.testing.assert_allclose(M[3], dst2, atol=1e-14)


def test_dtt_constants_match_reference(ref_tables):
    """The reference's butterfly constants are the closed forms at 6 dp."""
    r = ref_tables
    np.testing.assert_allclose(r["COSPI_1_8_SQRT2"], np.cos(np.pi / 8) * np.sqrt(2), atol=6e-7)
    np.testing.assert_allclose(r["COSPI_3_8_SQRT2"], np.cos(3 * np.pi / 8) * np.sqrt(2), atol=6e-7)
    np.testing.assert_allclose(r["SQRT1_8"], np.sqrt(1 / 8), atol=6e-7)
    k = np.arange(4)
    np.testing.assert_allclose(r["COSN2"], np.cos((2 * k + 1) * np.pi / 32), atol=6e-7)
    np.testing.assert_allclose(r["SINN2"], np.sin((2 * k + 1) * np.pi / 32), atol=6e-7)
    k = np.arange(2)
    np.testing.assert_allclose(r["COSN1"], np.cos((2 * k + 1) * np.pi / 16), atol=6e-7)
    np.testing.assert_allclose(r["SINN1"], np.sin((2 * k + 1) * np.pi / 16), atol=6e-7)


def test_windows_match_reference(orc, ref_tables):
    w = orc.table(2)
    for i, name in enumerate(["HWINDOW", "HWINDOW2", "HWINDOW_SQ", "HWINDOW_SQi"]):
        # printed with %f: 6 decimals
        np.testing.assert_allclose(w[i], ref_tables[name], atol=5.01e-7, rtol=0, err_msg=name)


def test_fold_map_reproduces_reference_fold_tables(orc, ref_tables):
    """fold_indx2 (start offsets of Bayer rows 0/1 in the 8x8 block) and
    fold_inc (packed 4-bit row increments going down 2 rows) follow from
    the oracle's fold map rind."""
    rind = orc.table(1)[0].astype(int)
    fi2 = np.array([[rind[c] + 8 * rind[r] for c in range(16)] for r in (0, 1)])
    np.testing.assert_array_equal(fi2.ravel(), ref_tables["fold_indx2"])
    for row0 in (0, 1):
        inc = [rind[2 * i + 2 + row0] - rind[2 * i + row0] for i in range(7)]
        packed = 0
        for d in reversed(inc):
            packed = (packed << 4) | (d & 0xF)
        assert packed == ref_tables["fold_inc"][row0]
    # walking the reference's packed increments from fold_indx2 visits
    # rind[row] for rows row0, row0 + 2, ...
    for row0 in (0, 1):
        off, inc = int(ref_tables["fold_indx2"][row0 * 16]), int(ref_tables["fold_inc"][row0])
        for i in range(8):
            assert off >> 3 == rind[row0 + 2 * i]
            off = (off + ((inc & 0xF) << 3)) & 0x3F
            inc >>= 4


def test_fold_signs_are_the_lapped_transform_signs(orc):
    """The fold map turns the 16-sample lapped kernel cos/sin(pi/8 (n+4.5)(k+1/2))
    into +-DCT-IV/DST-IV kernels of the folded index."""
    rind, rcs, rss = orc.table(1).astype(int)
    n = np.arange(16)[:, None]
    k = np.arange(8)[None, :]
    lc = np.cos(np.pi / 8 * (n + 4.5) * (k + 0.5))
    ls = np.sin(np.pi / 8 * (n + 4.5) * (k + 0.5))
    fc = np.cos(np.pi / 8 * (rind[:, None] + 0.5) * (k + 0.5))
    fs = np.sin(np.pi / 8 * (rind[:, None] + 0.5) * (k + 0.5))
    np.testing.assert_allclose(lc, rcs[:, None] * fc, atol=1e-12)
    np.testing.assert_allclose(ls, rss[:, None] * fs, atol=1e-12)


def test_imclt_index_and_signs_match_reference(orc, ref_tables):
    rind, rcs, rss = orc.table(1).astype(int)
    # imclt_indx9[c]: folded (row 4, column rind[c]) with 9-float rows
    assert list(ref_tables["imclt_indx9"]) == [4 * 9 + rind[c] for c in range(16)]
    # idct_signs[q][row block][column block]: row sign from the vertical
    # branch (sine for q & 2), column sign from the horizontal (sine for q & 1)
    sg = ref_tables["idct_signs"].reshape(4, 4, 4)
    for q in range(4):
        sv = rss if q & 2 else rcs
        sh = rss if q & 1 else rcs
        exp = np.array([[sv[4 * rb] * sh[4 * cb] for cb in range(4)] for rb in range(4)])
        np.testing.assert_array_equal(sg[q], exp)


def test_lpf_tables_match_reference(orc, ref_tables):
    lp = orc.table(3)
    ref = ref_tables["lpf_data"].reshape(4, 64)
    for i in range(4):
        np.testing.assert_allclose(lp[i], ref[i], atol=5.01e-9, rtol=0)
    np.testing.assert_allclose(lp[4], ref_tables["lpf_rb_corr"], atol=5.01e-9, rtol=0)
    np.testing.assert_allclose(lp[5], ref_tables["lpf_corr"], atol=5.01e-9, rtol=0)


def test_alpha_fade_matches_reference(orc, ref_tables):
    a = orc.table(4)
    np.testing.assert_allclose(a, ref_tables["alphaFade"].reshape(16, 256), atol=5.01e-7, rtol=0)


def test_pairs_match_reference(ref_tables):
    from tile_processor_gpu_amd.geometry import PAIRS
    np.testing.assert_array_equal(np.array(PAIRS).ravel(), ref_tables["pairs"])
