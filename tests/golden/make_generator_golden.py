"""Write the fixture of the .boid -> .grid generator test from the reference's own generator.

The reference generator (generator/*.cpp, GeneratorForMph <case>: reads <case>.boid, writes
<case>.grid, generator/generator.cpp:37-47) is built under oracle/_ref by oracle/Makefile.
This script runs it on the two sample .boid files the reference ships
(results/Bridge_l00100_deg000_re35_mu01_g98/box.boid, results/SquareDropSkip_l00400_rv35_rp17/
box.boid) in a temporary directory and stores, as data, each input .boid file and the .grid
file the generator wrote, byte for byte, in generator_samples.npz.

Run from the repository root in the container that has the reference:
    make -C oracle && python tests/golden/make_generator_golden.py
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
GEN = os.path.join(ROOT, "oracle", "_ref", "GeneratorForMph")
SAMPLES = {
    "bridge": "/root/reference/results/Bridge_l00100_deg000_re35_mu01_g98/box.boid",
    "square_drop": "/root/reference/results/SquareDropSkip_l00400_rv35_rp17/box.boid",
}


def main():
    out = {}
    for name, boid in SAMPLES.items():
        with tempfile.TemporaryDirectory() as tmp:
            shutil.copy(boid, os.path.join(tmp, "box.boid"))
            # the generator reports progress on stderr; a bounded run (it loops on a truncated file)
            subprocess.run([GEN, "box"], cwd=tmp, check=True, timeout=60, stderr=subprocess.DEVNULL)
            with open(os.path.join(tmp, "box.boid"), "rb") as f:
                out[name + "_boid"] = np.frombuffer(f.read(), np.uint8)
            with open(os.path.join(tmp, "box.grid"), "rb") as f:
                out[name + "_grid"] = np.frombuffer(f.read(), np.uint8)
        print(name, len(out[name + "_boid"]), "bytes of .boid,", len(out[name + "_grid"]), "bytes of .grid")
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "generator_samples.npz"), **out)


if __name__ == "__main__":
    sys.exit(main())
