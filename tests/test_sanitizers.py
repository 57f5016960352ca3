"""CPU: the host C++ under AddressSanitizer + UndefinedBehaviorSanitizer.

tests/sanitize/Makefile builds, with g++ -fsanitize=address,undefined and
-fno-sanitize-recover=all (the first report aborts):
  * the I/O and statistics drivers (ugu_amd/cpp: mesh_io, image_io, texturing,
    ugu_core), run through the unchanged tests/test_io.py and
    tests/test_texturing_host.py (UGU_IO_TOOL / UGU_STATS_TOOL select them);
  * tests/bvh_check.cc over ugu_amd/csrc/bvh_build.cpp;
  * the CPU oracle, loaded by tests/test_oracle_pins.py, test_host_math.py and
    test_visibility_host.py in a Python with the ASan runtime preloaded
    (UGU_ORACLE_LIB selects it);
  * tests/sanitize/fuzz_io.cc, a seeded fuzz of the PNG decoder and the
    OBJ/MTL loader that also checks their contracts (decoded size, index
    ranges of a loaded mesh).
A sanitizer report makes the child exit non-zero and is printed here.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
OUT = os.path.join(SAN, "_build")
DATA = os.path.join(ROOT, "tests", "golden", "data")

SAN_ENV = {
    "ASAN_OPTIONS": "abort_on_error=0:halt_on_error=1:detect_leaks=1:exitcode=23",
    "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
}


@pytest.fixture(scope="module")
def san_build():
    r = subprocess.run(["make", "-j8", "-C", SAN], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return OUT


def _env(**extra):
    env = dict(os.environ)
    env.update(SAN_ENV)
    env.update(extra)
    return env


def _pytest(files, env):
    cmd = [sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
           *[os.path.join(ROOT, "tests", f) for f in files]]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=900, env=env, cwd=ROOT)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "passed" in r.stdout and "failed" not in r.stdout
    return r.stdout


def test_io_tool_under_sanitizers(san_build):
    print(_pytest(["test_io.py"], _env(UGU_IO_TOOL=os.path.join(san_build, "ugu_io_tool"))))


def test_stats_tool_under_sanitizers(san_build):
    print(_pytest(["test_texturing_host.py"],
                  _env(UGU_STATS_TOOL=os.path.join(san_build, "ugu_stats_tool"))))


def test_bvh_build_under_sanitizers(san_build):
    for seed in (1, 2, 3):
        r = subprocess.run([os.path.join(san_build, "bvh_check"), str(seed), "12"],
                           capture_output=True, text=True, timeout=900, env=_env())
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
        assert " mismatches 0 " in r.stdout


def _preload_free():
    return not os.environ.get("LD_PRELOAD") and not os.path.exists("/etc/ld.so.preload")


@pytest.mark.skipif(not _preload_free(),
                    reason="another library is preloaded here; the ASan runtime must come first")
def test_oracle_under_sanitizers(san_build):
    libasan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True,
                             text=True, check=True).stdout.strip()
    assert os.path.isabs(libasan), libasan
    # the interpreter itself is not instrumented: its own allocations are
    # not leaks of the oracle, so leak checking is off for this child
    env = _env(LD_PRELOAD=libasan, UGU_ORACLE_LIB=os.path.join(san_build, "liboracle_san.so"),
               ASAN_OPTIONS="halt_on_error=1:detect_leaks=0:exitcode=23")
    print(_pytest(["test_oracle_pins.py", "test_host_math.py", "test_visibility_host.py"], env))


@pytest.mark.parametrize("seed", [11, 12])
def test_fuzz_png_decoder(san_build, tmp_path, seed):
    r = subprocess.run([os.path.join(san_build, "fuzz_io"), "png", str(seed), "2000",
                        str(tmp_path)], capture_output=True, text=True, timeout=900, env=_env())
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    ok = int(r.stdout.split(" ok ")[1].split()[0])
    assert ok > 20  # some inputs decode, so the fuzz reaches past the header


@pytest.mark.parametrize("seed", [21, 22])
def test_fuzz_obj_loader(san_build, tmp_path, seed):
    objs = [os.path.join(DATA, f) for f in ("icosphere3_smart_uv.obj", "spot_quadrangulated.obj")]
    r = subprocess.run([os.path.join(san_build, "fuzz_io"), "obj", str(seed), "1500",
                        str(tmp_path), *objs], capture_output=True, text=True, timeout=900,
                       env=_env())
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    ok = int(r.stdout.split(" ok ")[1].split()[0])
    assert ok > 100
