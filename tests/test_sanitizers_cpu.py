"""The library's host C (plan.c, tables.c, pynufht.c) built with
AddressSanitizer and UndefinedBehaviorSanitizer against a host-only stand-in
of the device engine (tests/sanitize/stub_engine.c), then driven through
the C ABI (tests/sanitize/driver.c: planning, partition, tables, scratch,
argument errors) and through the _pynufht Python entry points (argument
checks, array conversions, plan and execution errors)."""
import os
import shutil
import subprocess
import sys
import sysconfig

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "nufht_amd", "csrc")
HERE = os.path.join(ROOT, "tests", "sanitize")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
       "-fno-omit-frame-pointer", "-g", "-O1"]
INCS = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
SRCS = [os.path.join(CSRC, "plan.c"), os.path.join(CSRC, "tables.c"),
        os.path.join(HERE, "stub_engine.c")]


def _gcc():
    gcc = shutil.which("gcc")
    if not gcc:
        pytest.skip("gcc not available")
    return gcc


def _asan_env():
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    env.pop("NUFHT_TABLE_CACHE", None)
    return env


def test_host_c_under_sanitizers(tmp_path):
    exe = str(tmp_path / "driver")
    subprocess.run([_gcc(), *SAN, "-std=gnu11", *INCS, *SRCS, os.path.join(HERE, "driver.c"),
                    "-o", exe, "-lm", "-lpthread"], check=True, capture_output=True, text=True)
    r = subprocess.run([exe, str(tmp_path / "t.cache")], env=_asan_env(), capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all cases ran clean" in r.stdout
    assert os.path.getsize(tmp_path / "t.cache") == 2_669_080


def test_pynufht_under_sanitizers(tmp_path):
    import numpy
    gcc = _gcc()
    libasan = subprocess.run([gcc, "-print-file-name=libasan.so"], capture_output=True,
                             text=True).stdout.strip()
    if not os.path.isabs(libasan):
        pytest.skip("libasan not found")
    ext = str(tmp_path / ("_pynufht" + (sysconfig.get_config_var("EXT_SUFFIX") or ".so")))
    subprocess.run([gcc, *SAN, "-std=gnu11", "-shared", "-fPIC", *INCS,
                    "-I", sysconfig.get_paths()["include"], "-I", numpy.get_include(),
                    os.path.join(CSRC, "pynufht.c"), *SRCS, "-o", ext, "-lm", "-lpthread"],
                   check=True, capture_output=True, text=True)
    code = r'''
import importlib.util, sys
import numpy as np
spec = importlib.util.spec_from_file_location("_pynufht", sys.argv[1])
pn = importlib.util.module_from_spec(spec); spec.loader.exec_module(pn)
r = np.linspace(0.1, 50, 300); k = np.linspace(0.1, 40, 200)
cs = np.random.default_rng(1).standard_normal((300, 3))
def raises(exc, f, *a, **kw):
    try:
        f(*a, **kw)
    except exc:
        return
    raise AssertionError(f"{f.__name__} did not raise {exc.__name__}")
raises(ValueError, pn.nufht_batch, 0.0, r, cs[:, 0], k)           # cs not 2-D
raises(ValueError, pn.nufht_batch, 0.0, r, cs[:200], k)           # row count
raises(RuntimeError, pn.nufht_batch, 0.3, r, cs, k)               # order refused
raises(RuntimeError, pn.nufht_batch, 0.0, r, np.asfortranarray(cs), k, tol=1e-8)   # no device
raises(RuntimeError, pn.nufht_batch, 1.0, r, cs.astype(np.float32), k)            # converted, no device
# the per-column entry points go through the legacy nufht(), which (as the
# reference's) reports an execution failure on stderr and returns zeros
g = pn.nufht_batch_loop(2.0, r[::-1].copy(), cs, k)
assert g.shape == (200, 3) and not g.any()
g1 = pn.nufht(0.0, r, cs[:, 0].copy(), k)
assert g1.shape == (200,)
raises(TypeError, pn.nufht_batch, 0.0, r)
print("pynufht paths ran clean")
'''
    env = _asan_env()
    env["LD_PRELOAD"] = libasan
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"   # the interpreter's own allocations
    r = subprocess.run([sys.executable, "-c", code, ext], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "pynufht paths ran clean" in r.stdout
