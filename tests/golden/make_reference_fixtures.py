"""Copies data fixtures that the reference ships with its GH code into
tests/golden/reference/, gzip-compressed.  They are data (input rings and the
reference's result files), read by the tests as golden vectors:

  readPolygon/s.txt, readPolygon/c.txt   the large pair (count line, then one
                                         "x,y" vertex per line)
  results/Fig-large-R.poly               its result: 11 components
                                         (GH-CUDA/ReadMe.txt:59)
  Fig18-P/Q.poly, results/Fig18-R.poly   multi-component inputs and result
  Fig19-P/Q.poly, results/Fig19-R.poly

Run once in a tree where the reference is present:
    python tests/golden/make_reference_fixtures.py [/root/reference]
"""
import gzip
import os
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference")
FILES = ["GH-CUDA/data/readPolygon/s.txt", "GH-CUDA/data/readPolygon/c.txt",
         "GH-CUDA/results/Fig-large-R.poly",
         "GH-CUDA/data/Fig18-P.poly", "GH-CUDA/data/Fig18-Q.poly", "GH-CUDA/results/Fig18-R.poly",
         "GH-CUDA/data/Fig19-P.poly", "GH-CUDA/data/Fig19-Q.poly", "GH-CUDA/results/Fig19-R.poly"]

os.makedirs(DST, exist_ok=True)
for rel in FILES:
    with open(os.path.join(SRC, rel), "rb") as f:
        data = f.read()
    out = os.path.join(DST, os.path.basename(rel) + ".gz")
    with gzip.GzipFile(out, "wb", mtime=0) as g:
        g.write(data)
    print("%s -> %s (%d -> %d bytes)" % (rel, out, len(data), os.path.getsize(out)))
