"""CPU: every planted mistake in the oracle must be caught by some pin.

The oracle is rebuilt with -DORACLE_MUTATION=k (oracle/ugu_oracle.cc) into a
temporary directory; the pins of oracle_pins.py run against that build and at
least one must fail. Mistakes that only move results by float rounding (the
Eigen reduction order, divide-then-multiply in Project) cannot be caught this
way; for those the oracle's parity is unpinned.
"""
import os
import subprocess

import pytest

import oracle as O
import oracle_pins as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "ugu_oracle.cc")

MUTATIONS = {
    1: "edge function operands transposed (w0 = E(v2, v1, p))",
    2: "back-face test sign flipped",
    3: "depth test <= instead of < (ties go to the highest id)",
    4: "affine instead of perspective-correct interpolation",
    5: "face 0 shaded (fid >= 0)",
    6: "u and v swapped in the colourizer",
    7: "nearest-neighbour texture uses floor instead of round",
    8: "bilinear weights lu/lv swapped",
    9: "Lambertian coefficient sign flipped",
    10: "projection adds cx to y",
    12: "normal output rotated by R^T",
    13: "back-face culling ignored at resolve",
    14: "G-buffer barycentrics u and v swapped",
    15: "G-buffer camera position rotated by R^T",
    16: "raytracer barycentric conversion not applied (u, v swapped)",
    17: "raytracer keeps the farthest hit",
    18: "raytracer culling sign flipped",
    19: "raytracer depth = ray parameter t instead of camera z",
    20: "raytracer ray origin ignored",
    21: "visibility: occlusion check skipped",
    22: "visibility: nearest pixel by floor instead of round",
    23: "visibility: viewing angle from +n . ray",
    24: "visibility: bilinear weights swapped",
    25: "texture mapping: the first overlapping face wins",
    26: "texture mapping: padding reads the texture being padded",
    27: "texture mapping: bilinear weights swapped",
    28: "diff model: affine instead of perspective-correct weights (1/Z dropped)",
    29: "diff model: u and v swapped",
    30: "diff model: texel row not flipped (v instead of 1 - v)",
    31: "diff model: bilinear weights lu/lv swapped",
    32: "diff backward: the weights' 1/Z dependence dropped from the position gradient",
    33: "diff backward: texture gradient ignores the colour factor",
}


@pytest.fixture(scope="module")
def builds(tmp_path_factory):
    d = tmp_path_factory.mktemp("mut")
    out = {}
    procs = []
    for k in MUTATIONS:
        so = str(d / f"liboracle_m{k}.so")
        cmd = ["g++", "-O1", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared",
               f"-DORACLE_MUTATION={k}", "-o", so, SRC]
        procs.append((k, so, subprocess.Popen(cmd)))
    for k, so, p in procs:
        assert p.wait() == 0, f"mutant {k} failed to build"
        out[k] = so
    return out


@pytest.mark.parametrize("k", sorted(MUTATIONS), ids=lambda k: f"m{k}")
def test_mutation_is_caught(builds, k):
    O.use_library(builds[k])
    try:
        failures = []
        for pin in P.ALL_PINS:
            try:
                ok, msg = pin()
            except Exception as e:  # a pin that breaks on a mutant also catches it
                ok, msg = False, str(e)
            if not ok:
                failures.append(f"{pin.__name__}: {msg}")
        assert failures, f"planted mistake not caught by any pin: {MUTATIONS[k]}"
        print(f"m{k} ({MUTATIONS[k]}) caught by {failures[0]}")
    finally:
        O.use_library(None)
