"""The C++ Solver drop-in (compat/) driven like AxiSim's GUI drives its
Solver: worker thread, residual callbacks, solutions map, on the GPU."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

DEMO = Path(__file__).resolve().parent.parent / "axisim-software_amd" / "lib" / "compat_demo"


def test_compat_solver_converges_pipe(require_gpu):
    assert DEMO.exists(), "compat_demo not built"
    res = subprocess.run([str(DEMO), "20000"], capture_output=True, text=True, timeout=300)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "iterations=" in res.stdout
