"""The table cache writer: the file has the reference's layout and size,
its entries equal the reference's own cache slice (tests/golden) and the
oracle's restated generator, and a plan built with NUFHT_TABLE_CACHE naming
the written file gets the same parameters as one built without it."""
import os
import subprocess
import sys

import numpy as np

import oracle
from nufht_amd import tables

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_written_cache_matches_reference_slice_and_oracle(tmp_path):
    path = str(tmp_path / "nufht_tables.cache")
    tables.write(path)
    assert os.path.getsize(path) == 2_669_080
    t = tables.load(path)
    d = np.load(os.path.join(GOLDEN, "ref_tables_slice.npz"))
    assert list(t["header"]) == list(d["header"])
    assert np.array_equal(t["asy_coef"][:81], d["asy_coef"])
    np.testing.assert_array_equal(t["wimp_k"][:41, :12, :10], d["wimp_k"][:, :, :10])
    rz = d["asy_z"][:, :, :10]
    assert np.max(np.abs(t["asy_z"][:41, :12, :10] - rz) / rz) < 5e-11
    # slot K = 10 and the unused rows / tolerance columns stay empty, as written by the reference
    assert np.all(t["asy_z"][:, :, 10:] == 0) and np.all(t["wimp_k"][:, :, 10:] == 0)
    assert np.all(t["asy_z"][201:] == 0) and np.all(t["asy_z"][:, 12:] == 0)
    a, z, k = oracle.tables()
    assert np.array_equal(t["asy_coef"], a)
    np.testing.assert_array_equal(t["wimp_k"][:, :, :10], k[:, :, :10])
    nz = z[:201, :12, :10]
    assert np.max(np.abs(t["asy_z"][:201, :12, :10] - nz) / nz) < 1e-12


def test_plan_parameters_with_written_cache(tmp_path):
    path = str(tmp_path / "nufht_tables.cache")
    tables.write(path)
    code = ("import numpy as np, nufht_amd as na\n"
            "for nu, tol in ((0, 1e-8), (5, 1e-8), (1, 1e-6), (10, 1e-10), (35, 1e-12)):\n"
            "    p = na.Plan(float(nu), np.linspace(0.1, 50, 300), np.linspace(0.1, 50, 200), tol=tol)\n"
            "    print(nu, tol, p.K_asy, repr(p.z_split), p.K_loc)\n"
            "    p.close()\n")
    env = dict(os.environ, PYTHONPATH=ROOT)
    env.pop("NUFHT_TABLE_CACHE", None)
    plain = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                           check=True).stdout
    env["NUFHT_TABLE_CACHE"] = path
    cached = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                            check=True).stdout
    assert plain.strip() and plain == cached


def test_cache_found_where_the_reference_looks(tmp_path):
    """Without NUFHT_TABLE_CACHE the cache is read from data/ under the
    working directory, else under its parent (src/fast_hankel_nufht.c:95-160):
    a cache whose z entry for (nu = 5, tol 1e-8, K_asy = 4) is altered shows
    up in the plan built there, and nowhere else."""
    path = str(tmp_path / "ref.cache")
    tables.write(path)
    raw = bytearray(open(path, "rb").read())
    # asy_z[5][j = 4 (tol 1e-8)][slot 4] <- 12.5
    off = 24 + 401 * 64 * 8 + ((5 * 32 + 4) * 16 + 4) * 8
    raw[off:off + 8] = np.float64(12.5).tobytes()
    code = ("import numpy as np, nufht_amd as na\n"
            "p = na.Plan(5.0, np.linspace(0.1, 50, 300), np.linspace(0.1, 50, 200), tol=1e-8)\n"
            "print(repr(p.z_split))\n")
    env = dict(os.environ, PYTHONPATH=ROOT)
    env.pop("NUFHT_TABLE_CACHE", None)
    outs = {}
    for where in ("cwd", "parent", "none"):
        base = tmp_path / where
        work = base / "work"
        work.mkdir(parents=True)
        if where != "none":
            d = (work if where == "cwd" else base) / "data"
            d.mkdir()
            open(d / "nufht_tables.cache", "wb").write(raw)
        outs[where] = float(subprocess.run([sys.executable, "-c", code], env=env, cwd=str(work),
                                           capture_output=True, text=True, check=True).stdout)
    assert outs["cwd"] == 12.5 and outs["parent"] == 12.5
    assert abs(outs["none"] - 11.666739153830589) < 1e-9
