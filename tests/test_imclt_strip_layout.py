"""The LDS and output indexing of k_imclt_strip (csrc/k_imclt.hpp), emulated
on the CPU with the kernel's constants read from its source: each plane
pixel is written exactly once, as the sum of exactly the tiles of the grid
that cover it (rows 8 ty .. 8 ty + 15, columns 8 tx .. 8 tx + 15), no LDS
index leaves the 5 half-slots, and a unit's transposition staging stays in
its own wavefront's part of the half-slot that its unfold overwrites.

This checks the work split and the half-slot rotation; the arithmetic is
checked on the GPU against the FP64 oracle (tests/test_gpu_parity.py)."""
import os
import re

import pytest

SRC = os.path.join(os.path.dirname(__file__), "..", "tile_processor_gpu_amd", "csrc", "k_imclt.hpp")


def constants():
    text = open(SRC).read()
    tiles = int(re.search(r"IMS_TILES = (\d+)", text).group(1))
    band = int(re.search(r"IMS_BAND = (\d+)", text).group(1))
    return tiles, band


def emulate(TX, TY, band, tiles):
    own, row = tiles - 1, 16 * tiles
    hs_size = 8 * row
    PW, PH = 8 * TX + 8, 8 * TY + 8
    nstrip = (TX + 1 + own - 1) // own
    nband = (TY + 1 + band - 1) // band
    written = {}
    for strip in range(nstrip):
        for bnd in range(nband):
            G0, J0 = strip * own, bnd * band
            hs = [None] * (5 * hs_size)
            carry = 0
            for s in range((band + 1) // 2):
                Ta = J0 - 1 + 2 * s
                for wave in range(4):
                    for unit in range(8):
                        lt = (wave & 1) * 8 + unit
                        tx, ty = G0 - 1 + lt, Ta + (wave >> 1)
                        top = (carry + 1 + 2 * (wave >> 1)) % 5
                        bot = (carry + 2 + 2 * (wave >> 1)) % 5
                        ok = 0 <= ty < TY and 0 <= tx < TX
                        col = (wave & 1) * (row // 2) + 8 * unit  # staging column within row `unit`
                        assert (wave & 1) * (row // 2) <= col and col + 64 <= ((wave & 1) + 1) * (row // 2)
                        for r in range(16):
                            for c in range(16):
                                i = (top if r < 8 else bot) * hs_size + 16 * lt + (r & 7) * row + c
                                assert 0 <= i < 5 * hs_size
                                hs[i] = (ty, tx, r, c) if ok else None
                up0, lo0 = carry * hs_size, ((carry + 1) % 5) * hs_size
                up1, lo1 = ((carry + 2) % 5) * hs_size, ((carry + 3) % 5) * hs_size
                for t in range(256):
                    xl, hrow = t & 127, t >> 7
                    cg, xc = min(xl >> 3, own - 1), xl & 7
                    X = 8 * G0 + xl
                    xok = xl < 8 * own and X < PW
                    gl, gr = 16 * cg + xc + 8, 16 * (cg + 1) + xc
                    for gsel in range(2):
                        grp = Ta + gsel
                        if grp < J0:
                            continue
                        up, lo = (up1, lo1) if gsel else (up0, lo0)
                        for k in range(4):
                            yr = hrow + 2 * k
                            idx = [b + yr * row + g for b in (up, lo) for g in (gl, gr)]
                            assert all(0 <= i < 5 * hs_size for i in idx)
                            Y = 8 * grp + yr
                            if xok and grp < J0 + band and Y < PH:
                                assert (Y, X) not in written
                                written[(Y, X)] = {hs[i] for i in idx if hs[i] is not None}
                carry = (carry + 4) % 5
    assert len(written) == PW * PH
    for (Y, X), got in written.items():
        exp = {(ty, tx, Y - 8 * ty, X - 8 * tx)
               for ty in (Y // 8 - 1, Y // 8) for tx in (X // 8 - 1, X // 8)
               if 0 <= ty < TY and 0 <= tx < TX}
        assert got == exp, (Y, X)


@pytest.mark.parametrize("TX,TY", [(40, 30), (31, 17), (15, 9), (14, 16), (3, 2)])
def test_strip_layout_covers_every_pixel_once(TX, TY):
    tiles, band = constants()
    assert band % 2 == 1 and tiles == 16
    emulate(TX, TY, band, tiles)


@pytest.mark.parametrize("band", [3, 9])
def test_strip_layout_other_bands(band):
    emulate(33, 21, band, constants()[0])
