"""Learned preconditioner on the GPU: the MLP kernels against the numpy oracle
(bfloat16 weights and activations, FP32 accumulation), training, and the
guarded solve against the analytic preconditioner."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _case(tag):
    print(f"\n[gpu-case] {tag}", flush=True)


def _theta(seed=0, scale=0.25):
    import mlp as OM
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(OM.n_params()) * scale).astype(np.float32)


def test_mlp_forward_matches_oracle():
    _case("mlp forward vs oracle n=1000")
    import mlp as OM
    from hpmp.learned import MLP
    th = _theta(1)
    X = np.random.default_rng(2).standard_normal((1000, 11)).astype(np.float32)
    net = MLP(0)
    net.set_weights(th)
    Y = net.forward(X)
    Yo, _, _ = OM.forward(th, X)
    # identical roundings; only the FP32 summation order and tanh differ
    assert np.abs(Y - Yo).max() <= 2e-3 * np.abs(Yo).max()


def test_mfma_bf16_lane_layout():
    _case("v_mfma_f32_16x16x16_bf16 lane layout")
    from hpmp import _lib
    L = _lib.load()
    rng = np.random.default_rng(9)
    # small integers: exact in bfloat16, and every sum of 16 products is exact in FP32
    A = rng.integers(-8, 9, (16, 16)).astype(np.float32)
    B = rng.integers(-8, 9, (16, 16)).astype(np.float32)
    out = np.zeros(256, dtype=np.float32)
    assert L.hpg_selftest_mfma_bf16(_lib.ptr(A), _lib.ptr(B), _lib.ptr(out)) == 0
    D = A.astype(np.float64) @ B.astype(np.float64)
    # register i of lane l holds row 4*(l//16) + i, column l % 16
    exp = np.array([[D[4 * (l // 16) + i, l % 16] for i in range(4)] for l in range(64)]).ravel()
    assert np.array_equal(out, exp.astype(np.float32))


@pytest.mark.parametrize("n", [1, 15, 16, 17, 100003])
def test_mlp_forward_ragged_sizes(n):
    _case(f"mlp forward (MFMA tiles of 16) n={n}")
    import mlp as OM
    from hpmp.learned import MLP
    th = _theta(7)
    X = np.random.default_rng(n).standard_normal((n, 11)).astype(np.float32)
    net = MLP(0)
    net.set_weights(th)
    Y = net.forward(X)
    Yo, _, _ = OM.forward(th, X)
    assert np.all(np.isfinite(Y))
    # Both round every activation to bfloat16 the same way, but an FP32 sum that lies
    # within summation-order noise of a bfloat16 rounding boundary rounds to the
    # neighbouring value (one bfloat16 ulp, 2^-8 relative): over 100 003 x 192 hidden
    # activations a few hundred samples carry such a flip.  Every other sample agrees to
    # FP32 summation order; a flipped one stays within a few bfloat16 ulps.
    err = np.abs(Y - Yo).max(axis=1)
    scale = np.abs(Yo).max()
    assert np.mean(err > 2e-3 * scale) <= 1e-2
    assert err.max() <= 1e-2 * scale


def test_mlp_gradient_matches_oracle():
    _case("mlp loss/gradient vs oracle n=777")
    import mlp as OM
    from hpmp.learned import MLP
    th = _theta(3)
    rng = np.random.default_rng(4)
    X = rng.standard_normal((777, 11)).astype(np.float32)
    T = rng.standard_normal((777, 9)).astype(np.float32)
    net = MLP(0)
    net.set_weights(th)
    loss, g = net.loss_grad(X, T)
    lo, go = OM.loss_and_grad(th, X, T)
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    print("grad rel err", rel_err(g, go), flush=True)
    assert rel_err(g, go) < 2e-3


def test_mlp_training_reduces_loss_and_is_reproducible():
    _case("mlp training on a smooth target n=8192")
    from hpmp.learned import MLP
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, (8192, 11)).astype(np.float32)
    T = np.stack([np.sin(2 * X[:, 0]) * X[:, 1], X[:, 2] ** 2 - 0.3, np.tanh(X[:, 3] + X[:, 4]),
                  X[:, 5] * X[:, 6], np.cos(X[:, 7]) - 0.7, 0.5 * X[:, 8], X[:, 9] ** 3, X[:, 10] - X[:, 0],
                  np.sin(X[:, 1] + X[:, 2])], 1).astype(np.float32)
    losses = []
    ws = []
    for _ in range(2):
        net = MLP(7)
        l0, _ = net.loss_grad(X, T)
        last = net.train(X, T, epochs=30, batch=1024, lr=3e-3)
        losses.append((l0, last))
        ws.append(net.weights())
    print("loss before/after", losses[0], flush=True)
    assert losses[0][1] < 0.1 * losses[0][0]
    assert np.array_equal(ws[0], ws[1])


def _problem(seed, n=6, log2p=1, mu=5e-2):
    from hpmp import mesh as M, workloads as W
    m = M.unit_square(n, 0.2, 15)
    return m, W.noisy_freestream(m, 1 << log2p, seed=seed, scale=0.3)


def test_learned_preconditioner_converges_to_the_same_solution_faster():
    _case("learned preconditioner vs analytic n=6 log2p=1 mu=5e-2")
    from hpmp.learned import MLP, GuardedSolver, calibrate, training_data
    from hpmp.tri_hp import TriHpIns
    m, st_test = _problem(16)
    G = TriHpIns(m, log2p=1, mu=5e-2)
    train_states = [_problem(s)[1] for s in (101, 102, 103, 104)]
    X, T, off = training_data(G, train_states)
    net = MLP(11)
    net.train(X, T, epochs=200, batch=256, lr=2e-3)
    c, rates = calibrate(G, net, train_states[0])
    print("calibrated scale", c, "rates", rates, flush=True)

    G.set_state(*st_test)
    it_a, hist_a = GuardedSolver(G, None).solve(tol_rel=1e-10, max_cycles=6000)
    sol_a = G.get_state()
    G.set_state(*st_test)
    gs = GuardedSolver(G, net)
    it_l, hist_l = gs.solve(tol_rel=1e-10, max_cycles=6000)
    print("cycles analytic", it_a, "learned", it_l, "fell back at", gs.fell_back_at, flush=True)
    assert hist_a[-1][1] <= 1e-10 * hist_a[0][1] and hist_l[-1][1] <= 1e-10 * hist_l[0][1]
    from hpmp.tri_hp import rel_diff_mod_pressure
    err = rel_diff_mod_pressure(G.get_state(), sol_a)
    print("solution rel diff (mod pressure constant)", err, flush=True)
    assert err < 1e-8
    assert gs.fell_back_at is None and it_l < it_a


def test_guard_falls_back_from_an_unstable_network():
    _case("learned preconditioner fallback n=6 log2p=1")
    from hpmp.learned import MLP, GuardedSolver, shift_output
    from hpmp.tri_hp import TriHpIns
    m, st = _problem(17)
    G = TriHpIns(m, log2p=1, mu=5e-2)
    net = MLP(3)
    w = net.weights()
    w[:] = 0
    net.set_weights(w)
    shift_output(net, [-2.0] * 9)     # pseudo-time steps 7.4x the analytic ones
    G.set_state(*st)
    gs = GuardedSolver(G, net, check_every=10)
    it, hist = gs.solve(tol_rel=1e-6, max_cycles=6000)
    print("fell back at", gs.fell_back_at, "cycles", it, flush=True)
    assert gs.fell_back_at is not None and hist[-1][1] <= 1e-6 * hist[0][1]


def test_learned_on_every_level_cuts_w_cycles():
    """The learned form as the time-to-tolerance bench runs it, at a size a test affords:
    64 x 64 squares, p = 4, mu = 8e-3 (the cell Reynolds number of 512 x 512 at 1e-3),
    W(2,2,2) cycle over p 4 -> 2 -> 1 and the meshes to 16 x 16.  train_learned (200
    seeded states on 32 x 32, device training, calibration on this mesh size) must give
    a network with which the guarded solve reaches 1e-10 in fewer cycles than the
    analytic preconditioner, without falling back, at the same solution."""
    _case("learned on every level vs analytic, 64x64 p=4 mu=8e-3 W(2,2,2)")
    from hpmp import mesh as M, multigrid as MG, workloads as W
    from hpmp.learned import GuardedSolver, train_learned
    from hpmp.tri_hp import TriHpIns, rel_diff_mod_pressure
    n, mu = 64, 8e-3
    net, info = train_learned(mu, n, ntrain=200, calib_n=64, log=lambda m: print(m, flush=True))
    m = M.unit_square(n, 0.2, 1)
    st = W.fourier_state(m, 4, seed=2)
    meshes, links = MG.square_hierarchy(m, n, 3)
    g = TriHpIns(m, log2p=2, mu=mu)
    grids = MG.build_stack(TriHpIns, g, meshes, links, mu=mu)
    res = {}
    for tag, mlp in (("analytic", None), ("learned", net)):
        g.set_state(*st)
        gs = GuardedSolver(g, mlp, grids=grids, cyc=(2, 2, 2), check_every=5)
        it, hist = gs.solve(tol_rel=1e-10, max_cycles=600)
        res[tag] = (it, hist[-1][1] / hist[0][1], gs.fell_back_at, g.get_state())
        print(tag, "cycles", it, "final", res[tag][1], "fell back at", gs.fell_back_at, flush=True)
    assert res["analytic"][1] <= 1e-10 and res["learned"][1] <= 1e-10
    assert res["learned"][2] is None
    assert res["learned"][0] < res["analytic"][0]
    err = rel_diff_mod_pressure(res["learned"][3], res["analytic"][3])
    print("solution rel diff (mod pressure constant)", err, "type factors", info["type_factors"], flush=True)
    assert err < 1e-8
    for x in grids[:0:-1]:
        x.close()
    g.close()
    net.close()
