"""The learned SGS initial guess on the GPU: forward and gradient against the
numpy checker (oracle/mlp.py), training reduces the loss, the plain method
judges and bounds the guess (fallback), and what the trained network gains:
symmetric sweeps needed to reach the same linear tolerance with and without
the guess, on Newton steps it was not trained on."""
import numpy as np
import pytest

from flowpsi_amd import workloads
from flowpsi_amd.capi import Solver
from oracle import mlp

pytestmark = pytest.mark.gpu


def _solver(n=8):
    g, q, prm = workloads.box(n)
    s = Solver(g, prm)
    s.set_state(q)
    return g, s


def test_forward_and_gradient_match_numpy():
    g, s = _solver(6)
    s.mlp_init(seed=3)
    s.assemble()
    slot = s.mlp_harvest(20)
    out, grad, loss = s.mlp_eval(slot)
    x, y, off, idx = s.mlp_sample(slot)
    p = s.mlp_params()
    l_ref, g_ref, o_ref = mlp.loss_and_grad(p, x, y, off, idx, round_bf16=True)
    e_out = np.abs(out[:, :5] - o_ref[:, :5]).max() / (np.abs(o_ref[:, :5]).max() + 1e-30)
    e_grad = np.abs(grad - g_ref).max() / (np.abs(g_ref).max() + 1e-30)
    print(f"\nmlp out err {e_out:.2e} loss {loss:.6e} vs {l_ref:.6e} grad err {e_grad:.2e}")
    assert e_out < 2e-2
    assert abs(loss - l_ref) <= 2e-2 * l_ref
    assert e_grad < 3e-2


def test_training_reduces_loss_and_guess_saves_sweeps():
    """Trained on (system, 30-sweep solution) pairs from eight Newton steps
    of the 16^3 Euler box, the guess saves sweeps on the four steps after
    them: strictly fewer symmetric sweeps to the same linear tolerance, at
    least two fewer over the four steps (measured: 24 plain, 20 learned)."""
    g, s = _solver(16)
    s.mlp_init(seed=7)
    s.newton(5)  # past the start transient
    for _ in range(8):
        s.assemble()
        s.mlp_harvest(30)
        s.update()
    l0 = s.mlp_train(1, lr=0.0)
    l1 = s.mlp_train(300, lr=2e-3)
    print(f"\nloss before {l0:.4e} after {l1:.4e}")
    assert l1 < 0.5 * l0
    # held-out steps: same linear tolerance with and without the guess
    plain, learned = [], []
    q = s.get_state()
    for use in (False, True):
        s.set_state(q)
        _, lin = s.newton_adaptive(4, 1e-3, 60, use_guess=use)
        (learned if use else plain).append(lin)
    sw_plain = sum(l["sweeps"] for l in plain[0])
    sw_learn = sum(l["sweeps"] for l in learned[0])
    print(f"sweeps to 1e-3: plain {sw_plain}, learned {sw_learn}; "
          f"guess kept {[l['used_guess'] for l in learned[0]]}, r_guess/r0 "
          f"{[round(l['rguess'] / l['r0'], 3) for l in learned[0]]}, alpha "
          f"{[round(l['alpha'], 3) for l in learned[0]]}")
    for l in learned[0]:
        assert l["rguess"] <= l["r0"] * (1 + 1e-12)
        assert l["used_guess"] == 1
    for l in plain[0] + learned[0]:
        assert l["rfinal"] <= 1e-3 * l["r0"] or l["sweeps"] == 60
    assert sw_learn <= sw_plain - 2


def test_guess_under_line_sgs():
    """The same network under line SGS (the boundary-layer SST workload,
    dataflow sweeps): harvest, training and the guess run; the guess is never
    worse than the line-Jacobi start and the solve meets its tolerance."""
    g, q, prm = workloads.c5_bl(nx=10, ny=24, nz=8, sweep="persistent", precision="fp32")
    s = Solver(g, prm)
    s.set_state(q)
    s.mlp_init(seed=5)
    for _ in range(4):
        s.assemble()
        s.mlp_harvest(20)
        s.update()
    l0 = s.mlp_train(1, lr=0.0)
    l1 = s.mlp_train(100, lr=2e-3)
    assert l1 < l0
    _, lin = s.newton_adaptive(3, 0.3, 40, use_guess=True)
    for l in lin:
        assert l["rguess"] <= l["r0"] * (1 + 1e-12)
        assert l["rfinal"] <= 0.3 * l["r0"] or l["sweeps"] == 40
    print(f"\nline SGS guess: r_guess/r0 {[round(l['rguess'] / l['r0'], 3) for l in lin]}, "
          f"sweeps {[l['sweeps'] for l in lin]}")


def test_bad_guess_is_never_worse_than_plain_start():
    """A guess far from the solution: the step along it is scaled to the
    residual-minimising alpha, so the start is never worse than the plain
    one; when it is not better the plain start is kept bit for bit."""
    g, s = _solver(6)
    s.mlp_init(seed=1)
    s.assemble()
    s.mlp_harvest(5)
    p = s.mlp_params()
    o, n = mlp.LAYOUT["b3"]
    p[o:o + 5] = 1e3  # a guess far from the solution
    s.mlp_set_params(p)
    s.assemble()
    st = s.linear_solve_tol(1e-3, 40, use_guess=True)
    s.assemble()
    st0 = s.linear_solve_tol(1e-3, 40, use_guess=False)
    print(f"\nbad guess: {st}; plain: {st0}")
    assert st["rguess"] <= st["r0"] * (1 + 1e-12)
    assert abs(st["alpha"]) < 1e-2
    assert st["rfinal"] <= 1e-3 * st["r0"]
    if not st["used_guess"]:
        assert st["sweeps"] == st0["sweeps"] and st["rfinal"] == st0["rfinal"]


def test_guess_gain_on_held_out_steps_of_the_hybrid_grid():
    """converge.guess_gain on a small prism-layer + hex grid (SA, block SGS,
    dataflow sweep with the scalar riding in the fluid's sweeps): both paths
    reach the tolerance on every held-out step, the kept guesses lower the
    start's residual, and the report adds up."""
    from flowpsi_amd import converge
    g, q, prm = workloads.c4_hybrid(nx=10, nz=10, n_prism=8, n_hex=8)
    r = converge.guess_gain(g, q, dict(prm, precision="fp32", sweep="persistent"), transient=3, harvest_steps=2,
                            epochs=60, eval_steps=3, rtol=1e-3, max_sweeps=60)
    print(r)
    for path in ("learned", "plain"):
        assert all(r[path]["reached"])
        assert r[path]["total_sweeps"] == sum(r[path]["sweeps"])
    for used, ratio in zip(r["learned"]["used_guess"], r["learned"]["rguess_over_r0"]):
        assert (ratio < 1.0) == bool(used)
    assert r["sweeps_saved"] == r["plain"]["total_sweeps"] - r["learned"]["total_sweeps"]
    assert r["training"]["loss_after"] < r["training"]["loss_before"]
