"""Regenerates the data fixtures under tests/golden/data from the reference tree.

Run in the container that has the reference mounted:

    python tests/golden/make_fixtures.py [/root/reference]

What it writes (all data, no code):
  * data/*.fas, data/*.fasta, data/*.txt  -- the reference's test inputs
    (tests/testdata/), byte-for-byte.
  * data/matrices/<name>.txt -- the eight built-in score tables of
    src/matrices.c:38-290 rendered in the same text format the reference parses
    (header line of column symbols, one row per symbol).
  * data/benchmark_queries/<id>.fasta -- the reference's benchmark queries
    (benchmark/data/*).
"""
import os
import re
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(HERE, "data")


def main(ref):
    os.makedirs(os.path.join(DATA, "matrices"), exist_ok=True)
    os.makedirs(os.path.join(DATA, "benchmark_queries"), exist_ok=True)
    td = os.path.join(ref, "tests", "testdata")
    for name in sorted(os.listdir(td)):
        if name.startswith("."):
            continue
        shutil.copyfile(os.path.join(td, name), os.path.join(DATA, name))
    bd = os.path.join(ref, "benchmark", "data")
    for name in sorted(os.listdir(bd)):
        shutil.copyfile(os.path.join(bd, name), os.path.join(DATA, "benchmark_queries", name + ".fasta"))

    src = open(os.path.join(ref, "src", "matrices.c")).read()
    for m in re.finditer(r"const char mat_(\w+)\[\] =\s*\n(.*?);\n", src, re.S):
        name, body = m.group(1), m.group(2)
        text = "".join(re.findall(r'"((?:[^"\\]|\\.)*)"', body, re.S))
        text = text.replace("\\\n", "").replace("\\n", "\n").replace('\\"', '"')
        with open(os.path.join(DATA, "matrices", name + ".txt"), "w") as f:
            f.write(text)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
