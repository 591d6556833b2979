"""The C ABI from plain C: mdt_fpga-gpu_pipeline_amd/examples/client_dump.c
(built with gcc by the package Makefile) runs the client's set-up and
per-dump calls (MDT_OneStation_client.cu:584-705, 1083-1128, 1170) at the
deployed station shape for 3 dumps and checks them against the fused native
path itself; exit status 0 means they agree to 1e-5."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mdt_fpga-gpu_pipeline_amd",
                   "examples", "client_dump")


def test_c_client_dumps():
    if not os.path.exists(EXE):
        pytest.fail("examples/client_dump not built: run __graft_entry__.build()")
    print("c client: 38 inputs x 1024 channels x 8 taps, 256 spectra, 3 dumps", flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().splitlines()[-1].startswith("ok: 3 dumps")
