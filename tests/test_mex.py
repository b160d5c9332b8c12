"""The MEX gateways (cugridsearch_amd/mex/) through the MEX API stand-in of
tests/mex/, on the CPU: their argument errors (the reference gateway's
MATLAB:cudaAdd:* identifiers, src/gpu/cudaGridSearch_mex.cu:52-74), which
leave before any GPU call.  The GPU half is tests/test_gpu_mex.py."""
import os
import subprocess

import pytest

from cugridsearch_amd.build import MEX_CHECKS


@pytest.mark.parametrize("gateway", sorted(MEX_CHECKS))
def test_mex_gateway_argument_errors(gateway):
    exe = MEX_CHECKS[gateway]
    if not os.path.exists(exe):
        pytest.fail(f"{exe} missing: run __graft_entry__.build()")
    r = subprocess.run([exe, "--errors-only"], capture_output=True, text=True, timeout=60,
                       env=dict(os.environ, HIP_VISIBLE_DEVICES=""))
    assert r.returncode == 0, r.stdout + r.stderr
    assert " ok" in r.stdout
    if gateway == "cgs_mex.c":
        for id_ in ("MATLAB:cudaAdd:inputmismatch", "MATLAB:cudaAdd:outputmismatch", "MATLAB:cudaAdd:sizemismatch"):
            assert id_ in r.stderr
