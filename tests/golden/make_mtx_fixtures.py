"""Copies the reference's benchmark inputs (benchmarks/dump/Gx_<case>.mtx,
the power-flow Jacobian, and Gu_<case>.mtx, its right-hand sides) into
tests/golden/mtx/ as gzip files, byte for byte inside the gzip stream.  Only
the cases whose Jacobian the reference ships are taken (the three largest
Gx files are missing from it).  Run in the container that holds the
reference; the GPU box reads only the copies."""
import gzip
import os
import sys

SRC = "/root/reference/benchmarks/dump"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mtx")
CASES = ["case9", "case57", "case118", "case300", "case1354pegase", "case_ACTIVSg2000",
         "case2869pegase", "case9241pegase", "case_ACTIVSg10k"]


def main():
    os.makedirs(DST, exist_ok=True)
    for case in CASES:
        for kind in ("Gx", "Gu"):
            src = os.path.join(SRC, f"{kind}_{case}.mtx")
            with open(src, "rb") as f:
                data = f.read()
            dst = os.path.join(DST, f"{kind}_{case}.mtx.gz")
            with open(dst, "wb") as raw, gzip.GzipFile(filename="", mode="wb", fileobj=raw, mtime=0,
                                                       compresslevel=9) as g:
                g.write(data)
            print(f"{dst}: {len(data)} bytes")


if __name__ == "__main__":
    sys.exit(main())
