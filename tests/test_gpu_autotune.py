"""Autotuner (build extension): workgroup sizes chosen by timing change how
fast the kernels run, never what they compute."""
import numpy as np
import pytest

import axs

pytestmark = pytest.mark.gpu


def test_autotuned_run_matches_default_run(require_gpu):
    case = axs.nozzle_case(mesh=axs.nozzle_mesh(axs.nozzle_spacing_for(20000)), max_iter=8)
    s = axs.make_solver(case)
    s.run()
    d, t = s.autotune(5)
    print("autotune: default", d, "ms, tuned", t, "ms")
    assert 0.0 < t <= d * 1.05
    case.config.max_iter = 8  # a continued run adds max_iter iterations
    s.configure(case.config)
    s.run(continue_run=True)
    ref = axs.make_solver(axs.nozzle_case(mesh=case.mesh, max_iter=16))
    ref.run()
    assert s.stats().iterations == ref.stats().iterations == 16
    for name in ("u", "v", "p", "pp", "mdot"):
        np.testing.assert_array_equal(s.field(name), ref.field(name), err_msg=name)
