"""The CPU oracle and the library's .data/.grid parsers under AddressSanitizer and
UndefinedBehaviorSanitizer (host code only): a harness (tests/sanitize/oracle_harness.cpp)
built with g++ -fsanitize=address,undefined reads decks with the parsers and steps the
oracle through rebuilds, the virial and the neighbour export.  Any sanitizer report fails."""
import os
import shutil
import subprocess

import pytest

from mphexplicit_amd import decks

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not shutil.which("g++") or not shutil.which("gcc"):
        pytest.skip("no host compiler")
    out = tmp_path_factory.mktemp("asan")
    flags = ["-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
    obj = str(out / "mph_oracle.o")
    subprocess.run(["gcc", "-c", "-std=gnu11", "-ffp-contract=off", *flags, "-fopenmp",
                    os.path.join(ROOT, "oracle", "mph_oracle.c"), "-o", obj], check=True)
    exe = str(out / "oracle_harness")
    subprocess.run(["g++", "-std=c++17", "-ffp-contract=off", *flags, "-fopenmp", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "sanitize", "oracle_harness.cpp"),
                    os.path.join(ROOT, "mphexplicit_amd", "csrc", "mph_io.cpp"), obj, "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("factory,steps", [(decks.small_droplet_2d, 30), (decks.small_moving_wall_2d, 30),
                                           (decks.sample_bridge, 10), (decks.small_two_fluid_3d, 4)])
def test_oracle_and_parsers_clean_under_sanitizers(harness, factory, steps, tmp_path):
    d = factory()
    dpath, gpath = d.write(str(tmp_path), "deck")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="1")
    r = subprocess.run([harness, dpath, gpath, str(d.dim), str(steps)], capture_output=True, text=True, env=env,
                       timeout=300)
    print(r.stdout.strip())
    assert r.returncode == 0, r.stderr[-3000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-3000:]
    assert r.stdout.startswith("n %d " % d.n)
