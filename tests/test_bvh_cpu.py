"""CPU: the raytracer's hierarchy prunes nothing the float64 test would hit.
tests/bvh_check.cc (compiled here with g++ and the kernels' own header
ugu_amd/csrc/rt_math.h) replays the kernels' traversal on seeded soups and
adversarial rays and compares every closest t with the all-faces answer."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bvh_traversal_equals_brute_force(tmp_path):
    exe = str(tmp_path / "bvh_check")
    src = [os.path.join(ROOT, "tests", "bvh_check.cc"),
           os.path.join(ROOT, "ugu_amd", "csrc", "bvh_build.cpp")]
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-o", exe,
                    *src], check=True)
    for seed in (1, 2, 3):
        r = subprocess.run([exe, str(seed), "12"], capture_output=True, text=True, timeout=600)
        print(r.stdout)
        assert r.returncode == 0, r.stdout
        assert " mismatches 0 " in r.stdout
