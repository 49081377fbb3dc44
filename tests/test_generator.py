"""The .boid -> .grid generator (mphexplicit_amd.decks.parse_boid / grid_text_from_boid /
boid_to_grid) against the reference's own generator (generator/generator.cpp, built under
oracle/_ref by oracle/Makefile).

* the fixture tests/golden/generator_samples.npz holds the reference's two sample .boid
  files and the .grid files its generator wrote for them (make_generator_golden.py): ours
  must be byte-identical;
* the sample decks (decks.sample_bridge, decks.sample_square_drop) are built from cuboid
  lists that must equal what the parser reads from those files;
* where the reference generator is built (this container), a .boid file laid out unusually
  (tokens of one cuboid across lines, a comment, text after EndCuboid, a line the generator
  ignores) must also give its output byte for byte.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from mphexplicit_amd import decks

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "generator_samples.npz")
GEN = os.path.join(ROOT, "oracle", "_ref", "GeneratorForMph")


@pytest.mark.parametrize("name", ["bridge", "square_drop"])
def test_grid_bytes_match_reference_generator(name):
    g = np.load(GOLDEN, allow_pickle=False)
    boid = decks.parse_boid(bytes(g[name + "_boid"]).decode())
    ours = decks.grid_text_from_boid(boid).encode()
    assert ours == bytes(g[name + "_grid"])


def test_sample_decks_use_the_sample_cuboids():
    g = np.load(GOLDEN, allow_pickle=False)
    for name, cubs in (("bridge", decks.SAMPLE_BRIDGE_CUBOIDS), ("square_drop", decks.SAMPLE_SQUARE_DROP_CUBOIDS)):
        assert decks.parse_boid(bytes(g[name + "_boid"]).decode()).cuboids == cubs, name
    d = decks.sample_bridge()
    assert d.n == 3200 and set(d.type.tolist()) == {1, 2}


ODD = """# unusual layout
ParticleDistance 0.001
LowerDomain -0.01 -0.01 -0.001
UpperDomain  0.01  0.01  0.001
this line names no keyword and is ignored
StartCuboid   ignored text after the keyword
  Spacing 0.001 Type 0
  Lower -0.003
  -0.002 -0.0005 Upper 0.003 0.0025
  0.0005
  Velocity 0.1 -0.2 0.0
EndCuboid
StartCuboid
Spacing 0.0013 Type 3 Lower 0.004 0.0 -0.0005 Upper 0.0091 0.0033 0.0005 Velocity 0 0 0 EndCuboid
# the end
"""


def test_unusual_layout_matches_reference_generator(tmp_path):
    if not os.path.exists(GEN):
        pytest.skip("reference generator not built here")
    case = str(tmp_path / "odd")
    with open(case + ".boid", "w") as f:
        f.write(ODD)
    subprocess.run([GEN, case], check=True, timeout=60, stderr=subprocess.DEVNULL)
    with open(case + ".grid", "rb") as f:
        ref = f.read()
    assert decks.grid_text_from_boid(decks.parse_boid(ODD)).encode() == ref
    # and the command-line form writes the same file
    os.remove(case + ".grid")
    subprocess.run([sys.executable, "-m", "mphexplicit_amd.decks", case], check=True, cwd=ROOT, timeout=60)
    with open(case + ".grid", "rb") as f:
        assert f.read() == ref


def test_parse_errors():
    with pytest.raises(ValueError):
        decks.parse_boid("StartCuboid\nSpacing 0.1 Type 1\nEndCuboid\n")          # no Lower/Upper/Velocity
    with pytest.raises(ValueError):
        decks.parse_boid("StartCuboid\nSpacing 0.1 Colour 3\nEndCuboid\n")        # unknown token
