"""The CPU oracle and the host-side input generation under AddressSanitizer and
UndefinedBehaviorSanitizer (gcc; UB aborts, -fno-sanitize-recover=undefined).
The `sanitize` targets of oracle/Makefile and eulerian-liquid_amd/Makefile
build them; each check runs in a child Python process with libasan preloaded.

- The oracle's CPU suites (pins, golden vectors, extrapolation, advection,
  reinitialisation) pass against the sanitized oracle (ELQ_ORACLE_LIB).
- The scene generators (drop, spheres, clamp and serving-request scenes, all
  through elq_scene.cpp) give the same bytes from the sanitized build of
  elq_scene.cpp as from the shipped library.
"""
import hashlib
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "eulerian-liquid_amd")

SCENES = """
import ctypes as C, hashlib, sys
sys.path[:0] = [{root!r}, {pkg!r}]
import elq
L = elq.load()
if {sanitized!r}:
    san = C.CDLL({lib!r})

    class Scenes:
        def __getattr__(self, k):
            if not k.startswith("elq_scene_"):
                return getattr(L, k)
            f, g = getattr(san, k), getattr(L, k)
            f.argtypes, f.restype = g.argtypes, g.restype
            return f

    proxy = Scenes()
    elq.load = lambda: proxy
h = hashlib.sha256()
for s in (elq.scene_drop(32), elq.scene_spheres(40, seed=2, n_spheres=5), elq.scene_clamp(),
          elq.scene_request(7), elq.scene_request(11)):
    for a in (s.phi, s.u, s.v, s.w, s.uw, s.vw, s.ww):
        h.update(a.tobytes())
print(h.hexdigest())
"""


def _asan():
    p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(p) or not os.path.exists(p):
        pytest.skip("gcc's libasan is not installed")
    return p


def _env(asan, **extra):
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", **extra)
    return env


@pytest.fixture(scope="module")
def asan():
    lib = _asan()
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    subprocess.run(["make", "-s", "-C", PKG, "sanitize"], check=True)
    return lib


def test_oracle_suites_under_sanitizers(asan):
    lib = os.path.join(ROOT, "oracle", "_asan", "liboracle_elq.so")
    files = ["test_oracle_pins.py", "test_oracle_advection.py", "test_oracle_extrapolation.py",
             "test_oracle_reinit.py", "test_golden.py"]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu"] +
                       [os.path.join(ROOT, "tests", f) for f in files],
                       capture_output=True, text=True, timeout=600, cwd=ROOT, env=_env(asan, ELQ_ORACLE_LIB=lib))
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "passed" in r.stdout and "ERROR: AddressSanitizer" not in r.stderr


def test_scene_generation_under_sanitizers(asan):
    lib = os.path.join(PKG, "_asan", "libelq_scene.so")
    plain = subprocess.run([sys.executable, "-c", SCENES.format(root=ROOT, pkg=PKG, sanitized=False, lib=lib)],
                           capture_output=True, text=True, timeout=300, check=True).stdout.split()[-1]
    r = subprocess.run([sys.executable, "-c", SCENES.format(root=ROOT, pkg=PKG, sanitized=True, lib=lib)],
                       capture_output=True, text=True, timeout=600, env=_env(asan))
    assert r.returncode == 0, r.stderr[-3000:]
    assert r.stdout.split()[-1] == plain
    assert len(bytes.fromhex(plain)) == hashlib.sha256().digest_size
