"""Learned initial guess on the GPU (build extension): the device network's
forward pass and gradient against the CPU restatement (oracle/net_reference.py),
training on solver-generated pairs, and what the trained network gains."""
import sys
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import axs
import learned_guess as lg
from oracle_binding import Oracle

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import net_reference as nr  # noqa: E402

pytestmark = pytest.mark.gpu


def small_system(iters=3):
    case = axs.pipe_case(nz=32, nr=8)
    case.config.max_iter = iters
    o = Oracle(case.mesh, case.bcs, case.config)
    o.run()
    fs, fn = o.csr()
    AC, AF, b = o.field("pp_ac") * 1.02, o.field("pp_af"), o.field("pp_b")
    x = spla.spsolve(lg.csr_matrix(fs, fn, AC, AF).tocsc(), b)
    return fs, fn, AC, AF, b, x


def test_forward_matches_reference(require_gpu):
    fs, fn, AC, AF, b, _ = small_system()
    net = axs.Net(layers=4, seed=0)
    p = nr.init_params(4, seed=11).astype(np.float32)
    net.set_params(p)
    got = net.predict(fs, fn, AC, AF, b)
    g = nr.Graph(fs, fn, AC, AF)
    ref = nr.guess(p, g, b, 4, exact=False)
    exact = nr.guess(p, g, b, 4, exact=True)
    e_ref = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    e_exact = np.linalg.norm(got - exact) / np.linalg.norm(exact)
    print("forward rel err vs device-precision reference", e_ref, "vs exact", e_exact)
    assert e_ref < 1e-2 and e_exact < 3e-2


def test_gradient_matches_reference(require_gpu):
    fs, fn, AC, AF, b, x = small_system()
    net = axs.Net(layers=3, seed=0)
    p = nr.init_params(3, seed=4).astype(np.float32)
    net.set_params(p)
    i = net.add_sample(fs, fn, AC, AF, b, x)
    loss, grad = net.loss_grad(i)
    rl, rg = nr.loss_grad(p, nr.Graph(fs, fn, AC, AF), b, x, 3, exact=False)
    print("loss", loss, "reference", rl)
    assert abs(loss - rl) <= 1e-3 * abs(rl) + 1e-6
    sizes = [int(np.prod(s)) for s in nr.param_shapes(3)]
    o = 0
    for s in sizes:
        a, r = grad[o:o + s], rg[o:o + s]
        err = np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-30)
        print("block", o, s, "rel err", err)
        assert err < 3e-2
        o += s


def _pairs(seeds, cells=6000, its=(2, 5, 10, 20)):
    out = []
    for s in seeds:
        r = np.random.default_rng(s)
        mesh = axs.nozzle_mesh(axs.nozzle_spacing_for(cells), seed_points=s, seed_jitter=s + 1,
                               r_throat=float(r.uniform(0.0017, 0.0023)), cone_deg=float(r.uniform(16, 24)))
        case = axs.nozzle_case(mesh=mesh, re_throat=float(r.uniform(300, 700)))
        fs, fn = mesh.arrays["cell_face_start"], lg.slot_neighbours(mesh)
        mg = axs.Multigrid(fs, fn)
        hier, k = lg.Hierarchy(mg), lg.guess_level(mg, cells=2048)
        for AC, AF, b in lg.solver_systems(case, list(its)):
            out.append(lg.coarse_pair(hier, fs, fn, AC, AF, b, k))
    return out


def _captured(net, pairs):
    """Mean share of the coarse solution's energy the scaled guess captures."""
    v = []
    for fsk, fnk, ACk, AFk, bk, xk in pairs:
        xh = net.predict(fsk, fnk, ACk, AFk, bk)
        A = lg.csr_matrix(fsk, fnk, ACk, AFk)
        e = xk - xh
        v.append(1.0 - (e @ (A @ e)) / (xk @ bk))
    return float(np.mean(v)), float(np.min(v))


def test_training_gain_on_held_out_systems(require_gpu):
    """Trained on pairs from seeded nozzles, evaluated on other seeds: the share
    of the solution's energy the guess captures goes up with training, and the
    guess never starts a solve further from the solution than zero does."""
    train, test = _pairs(range(5000, 5006)), _pairs(range(5100, 5102))
    net = axs.Net(layers=4, seed=2)
    for q in train:
        net.add_sample(*q)
    before, _ = _captured(net, test)
    losses = net.train(150, batch=8, lr=1e-3, seed=0)
    after, worst = _captured(net, test)
    print(f"train loss {losses[0]:.5f} -> {losses[-1]:.5f}; held-out captured energy {before:.5f} -> {after:.5f}")
    assert losses[-1] < losses[0]
    assert worst >= -1e-9  # alpha-scaling: never worse than the zero guess
    assert after > before


def test_trained_guess_changes_iterations_not_the_answer(require_gpu):
    """A network trained on seeded nozzles' p' systems, then used as the initial
    guess of every outer iteration's first p' solve on a nozzle it has not seen:
    both runs converge to 1e-6 and end on the same fields (to that tolerance);
    the iteration counts are reported."""
    net = axs.Net(layers=4, seed=3)
    for q in _pairs(range(5000, 5006)):
        net.add_sample(*q)
    losses = net.train(150, batch=8, lr=1e-3, seed=0)
    assert losses[-1] < losses[0]
    mesh = axs.nozzle_mesh(axs.nozzle_spacing_for(6000), seed_points=5301, seed_jitter=5302)
    s0 = axs.make_solver(axs.nozzle_case(mesh=mesh))
    s0.run()
    s1 = axs.make_solver(axs.nozzle_case(mesh=mesh))
    s1.set_learned_guess(net, -1)
    s1.run()
    st0, st1 = s0.stats(), s1.stats()
    print(f"train loss {losses[0]:.4f} -> {losses[-1]:.4f}; outer iterations to 1e-6: plain {st0.iterations}, "
          f"with the trained guess {st1.iterations}")
    assert st0.converged and st1.converged
    uscale = np.linalg.norm(s0.field("u"))
    for name, scale in (("u", uscale), ("v", uscale), ("p", np.linalg.norm(s0.field("p")))):
        assert np.linalg.norm(s1.field(name) - s0.field(name)) <= 1e-5 * scale, name
