#!/usr/bin/env python3
"""Writes the golden fixtures in tests/golden/ from the compiled reference.

The reference's CPU path (cpu.cxx, arraysizes.cxx, adler32.c) is compiled in
place by ``make -f oracle/ref.mk`` into ``oracle/_ref/ref_harness``; this script
runs that binary and stores its outputs as JSON data (inputs are described by
their generator and seed, outputs are checksums and sampled J bytes). The
fixtures are committed; the reference itself never is.

Fixtures:
  codec.json        decode of all 256 bytes, encode of all (re, im) in [-8, 7]^2
                    through the reference ucomplex4 (icomplex4.hxx:22-40)
  adler32.json      reference adler32.c over seeded byte buffers
  ref_pattern.json  reference form_beams on setup() inputs at the reference's
                    compiled shape: adler32 of J plus sampled J bytes
  ref_seeded.json   reference form_beams on seeded-uniform E, A (oracle
                    generator) and setup() gains at the same shape

Usage: python tests/golden/make_golden.py   (takes about a minute on 8 cores)
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def run(*args):
    out = subprocess.run([HARNESS, *map(str, args)], check=True, capture_output=True, text=True).stdout
    # the reference's setup() prints progress lines on stdout; the JSON is last
    return json.loads([ln for ln in out.splitlines() if ln.startswith("{")][-1])


def dump(name, obj):
    path = os.path.join(HERE, name)
    with open(path, "w") as fh:
        json.dump(obj, fh, separators=(",", ":"))
        fh.write("\n")
    print("wrote", path)


def main():
    subprocess.run(["make", "-s", "-f", os.path.join(ROOT, "oracle", "ref.mk")], check=True)
    if not os.path.exists(HARNESS):
        sys.exit("reference harness not built (reference sources absent?)")
    dump("codec.json", run("codec"))
    dump("adler32.json", {"cases": [run("adler32", s, n) for s, n in
                                    [(1, 0), (1, 1), (2, 100), (3, 5551), (4, 5552), (5, 65536), (6, 1 << 20),
                                     (7, 3 * (1 << 20) + 17)]]})
    dump("ref_pattern.json", run("pattern", 2048))
    seeded = [run("seeded", seed, 2048) for seed in (1, 20240611)]
    dump("ref_seeded.json", {"runs": seeded})


if __name__ == "__main__":
    main()
