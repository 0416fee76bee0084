"""The blend computes 1/den as v_rcp_f32 plus one FMA Newton step (rcp_rn in
csrc/fhtgs_device.h) and relies on it being the IEEE quotient, and scales its
exponential with v_ldexp_f32 (fh_expf_m64_gfx) where the oracle multiplies
by a power of two; rcp_check compares rcp_rn with IEEE 1/x over all 2^32
float inputs, and the two exponentials on every float of [-64, 0] and NaN,
on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "foveatedhtgscudabackend_amd", "_build", "rcp_check")


def test_rcp_rn_is_ieee_reciprocal_for_all_floats():
    assert os.path.exists(EXE), "rcp_check is not built (make -C foveatedhtgscudabackend_amd/csrc)"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
