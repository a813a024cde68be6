"""Extract the reference's constant tables as numbers into
reference_tables.json (golden vectors for the closed-form table generators).

Reads, as text, the __constant__ initialisers of
  src/TileProcessor.cuh (HWINDOW*, fold_indx2, fold_inc, imclt_indx9,
  idct_signs, lpf_data, lpf_rb_corr, lpf_corr, pairs, alphaFade) and
  src/dtt8x8.cuh (COSPI_*_SQRT2, SQRT*, COSN*, SINN*)
of mfkiwl/tile_processor_gpu.  Only numbers are kept.

Usage: python tests/golden/make_reference_tables.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_tables.json")

NUM = r"[-+]?(?:0x[0-9a-fA-F]+|\d+\.?\d*(?:[eE][-+]?\d+)?)f?"


def numbers(text):
    vals = []
    for tok in re.findall(NUM, text):
        if tok.lower().lstrip("+-").startswith("0x"):
            vals.append(int(tok, 16))
        else:
            vals.append(float(tok.rstrip("f")))
    return vals


def initializer(src, name):
    """Numbers of the first `__constant__ <type> name...= {...};` (comments stripped)."""
    m = re.search(r"__constant__\s+\w+\s+" + re.escape(name) + r"\s*(\[[^=]*\])?\s*=", src)
    if not m:
        raise KeyError(name)
    body = src[m.end():src.index(";", m.end())]
    body = re.sub(r"//[^\n]*", "", body)
    return numbers(body)


def main():
    tp = open(os.path.join(REF, "src", "TileProcessor.cuh")).read()
    dt = open(os.path.join(REF, "src", "dtt8x8.cuh")).read()
    tp = re.sub(r"//[^\n]*", "", re.sub(r"/\*.*?\*/", "", tp, flags=re.S))
    dt = re.sub(r"//[^\n]*", "", re.sub(r"/\*.*?\*/", "", dt, flags=re.S))
    out = {}
    for name in ["HWINDOW", "HWINDOW2", "HWINDOW_SQ", "HWINDOW_SQi", "fold_indx2", "fold_inc",
                 "imclt_indx9", "idct_signs", "lpf_data", "lpf_rb_corr", "lpf_corr", "pairs",
                 "alphaFade"]:
        out[name] = initializer(tp, name)
    for name in ["COSPI_1_8_SQRT2", "COSPI_3_8_SQRT2", "SQRT_2", "SQRT1_2", "SQRT1_8",
                 "COSN1", "COSN2", "SINN1", "SINN2"]:
        out[name] = initializer(dt, name)
    sizes = dict(HWINDOW=8, HWINDOW2=8, HWINDOW_SQ=8, HWINDOW_SQi=8, fold_indx2=32, fold_inc=2,
                 imclt_indx9=16, idct_signs=64, lpf_data=256, lpf_rb_corr=64, lpf_corr=64,
                 pairs=12, alphaFade=4096)
    for k, n in sizes.items():
        assert len(out[k]) == n, (k, len(out[k]))
    with open(OUT, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", OUT, {k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
