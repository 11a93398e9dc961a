"""The C-ABI smoke program (C) and the C++ API test program, built by
cybermeshgenerator_amd.build, run against libcmg_hip on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native")


@pytest.mark.parametrize("prog", ["smoke_c_abi", "test_cpp_api"])
def test_native_program(gpu, prog):
    exe = os.path.join(HERE, prog)
    assert os.path.exists(exe), f"{exe} not built (run cybermeshgenerator_amd.build)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
