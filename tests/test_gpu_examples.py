"""The application examples end to end on the GPU: the Fresnel lensing
integral (100 transforms of different nodes in one multi-transform batch, and
the notebook's per-w loop through _pynufht) against the oracle's FP64 direct
sums, and the correlation function (order 1/2) against the closed form
J_(1/2)(x) = sqrt(2 / (pi x)) sin x summed directly."""
import numpy as np
import pytest

import oracle
from nufht_amd import examples
from tests.conftest import rel_l2, say

pytestmark = pytest.mark.gpu


def test_fresnel_batch_and_loop_against_direct_sums(device):
    say("Fresnel F(w, y): 100 w x 150 y, one multi-transform batch")
    out = examples.fresnel()
    say(f"  batch {out['seconds'] * 1e3:.2f} ms")
    y = out["y"]
    for q in (0, 33, 66, 99):
        ref = oracle.direct(0.0, out["nodes"][q], out["columns"][q], y)
        got = np.column_stack([out["G"][q].real, out["G"][q].imag])
        err = rel_l2(got, ref)
        say(f"  w={out['w'][q]:.2f}: rel l2 {err:.2e}")
        assert err < 2e-8
    # F(w, 0) = 1 + (1 / (i w)) sum c_j (J_0(0) = 1)
    s0 = np.array([c[:, 0].sum() + 1j * c[:, 1].sum() for c in out["columns"]])
    assert np.max(np.abs(out["F"][:, 0] - (1.0 + s0 / (1j * out["w"])))) < 1e-8 * np.max(np.abs(s0))
    say("Fresnel F(w, y): per-w _pynufht loop")
    loop = examples.fresnel(loop=True)
    assert rel_l2(loop["G"], out["G"]) < 1e-12


def test_correlation_function_against_closed_form(device):
    say("xi(r): 10000-point ln-k rule, 200 r, 5 Omega_m, order 1/2")
    out = examples.correlation()
    k, cs, r = out["k"], out["cs"], out["r"]
    ref = (np.sqrt(2.0 / (np.pi * np.outer(r, k))) * np.sin(np.outer(r, k))) @ cs
    for i in range(cs.shape[1]):
        err = rel_l2(out["G"][:, i], ref[:, i])
        say(f"  Omega_m={out['omega_m'][i]:.2f}: rel l2 {err:.2e}")
        assert err < 2e-8
    assert np.all(out["xi"][:40] > 0)     # positive on small scales
