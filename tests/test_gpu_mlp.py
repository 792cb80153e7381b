"""Learned selector on the GPU: BF16-MFMA forward/backward against the CPU
oracle (oracle/mlp.py), training, the selection step against an oracle
restatement, and the gain over plain heat-bath selection."""
import numpy as np
import pytest

import pyci_amd as pyci
from oracle import mlp as M
from pyci_amd.selection import hartreefock_genci, run_selection
from pyci_amd.synthetic import seeded_integrals, spin_orbital

pytestmark = pytest.mark.gpu


def genci_problem(n, seed=1003):
    """The selected-CI integral recipe (h_ii = -3 + 0.2 i, 64 vectors decaying
    as exp(-|i-j|/6)) at n spatial orbitals, as spin-orbital integrals."""
    _, h, g = seeded_integrals(n, 64, seed, hdiag0=-3.0, hstep=0.2, decay=6.0)
    return spin_orbital(h, g)


def hf_sd_wfn(n, m):
    wfn = pyci.genci_wfn(2 * n, 2 * m, 0)
    hf = hartreefock_genci(n, m, m)
    wfn.add_det(hf)
    pyci.add_excitations(wfn, 1, 2, ref=hf)
    return wfn, hf


def random_genci_wfn(nb, ne, count, seed):
    rng = np.random.default_rng(seed)
    wfn = pyci.genci_wfn(nb, ne, 0)
    while len(wfn) < count:
        occ = rng.choice(nb, ne, replace=False)
        det = np.array([sum(1 << int(i) for i in occ)], dtype=np.uint64)
        wfn.add_det(det)
    return wfn


def perturbed_params(sel, seed):
    """Network parameters with nonzero biases, so that every term of the
    backward is exercised."""
    rng = np.random.default_rng(seed)
    p = np.array(sel.params, dtype=np.float32)
    p[M.OB1:M.OW2] = rng.normal(0, 0.2, M.HID)
    p[M.OB2:M.OW3] = rng.normal(0, 0.2, M.HID)
    p[M.OB3] = -3.0
    return p


def test_params_roundtrip_and_init():
    sel = pyci.mlp_selector(20, seed=5)
    p = np.array(sel.params)
    assert p.shape == (pyci.MLP_NPARAM,) and pyci.MLP_NPARAM == M.NPARAM
    W1 = p[M.OW1:M.OB1].reshape(M.FPAD, M.HID)
    assert np.all(W1[20:] == 0) and np.all(np.abs(W1[:20]) <= np.sqrt(6 / 20))
    assert np.all(p[M.OB1:M.OW2] == 0)
    q = perturbed_params(sel, 1)
    sel.params = q
    assert np.array_equal(np.array(sel.params), q)
    # deterministic initialisation
    assert np.array_equal(np.array(pyci.mlp_selector(20, seed=5).params), p)
    with pytest.raises(ValueError):
        pyci.mlp_selector(65)


@pytest.mark.parametrize("count", [1, 200, 5000])
def test_forward_matches_oracle(count):
    wfn = random_genci_wfn(24, 8, count, seed=count)
    sel = pyci.mlp_selector(24, seed=11)
    sel.params = perturbed_params(sel, 2)
    y = np.array(sel.predict(wfn))
    X = M.features("genci", 24, wfn.to_det_array())
    ref, _ = M.forward(np.array(sel.params), X)
    err = np.abs(y - ref)
    # FP32 accumulation order differs from the float64 oracle; a bf16 rounding of
    # an activation can then land one bf16 ulp (2^-8 relative) apart
    assert np.max(err) < 2e-2 * (1 + np.max(np.abs(ref)))
    assert np.median(err) < 2e-4 * (1 + np.max(np.abs(ref)))


def test_forward_fullci_features():
    wfn = pyci.fullci_wfn(8, 3, 2)
    wfn.add_all_dets()
    sel = pyci.mlp_selector(16, seed=3)
    sel.params = perturbed_params(sel, 3)
    y = np.array(sel.predict(wfn))
    ref, _ = M.forward(np.array(sel.params), M.features("fullci", 8, wfn.to_det_array()))
    assert np.max(np.abs(y - ref)) < 2e-2 * (1 + np.max(np.abs(ref)))
    with pytest.raises(ValueError):
        pyci.mlp_selector(8).predict(wfn)


@pytest.mark.parametrize("count", [300, 3000])
def test_loss_grad_matches_oracle(count):
    wfn = random_genci_wfn(24, 8, count, seed=100 + count)
    sel = pyci.mlp_selector(24, seed=12)
    sel.params = perturbed_params(sel, 4)
    rng = np.random.default_rng(count)
    t = rng.normal(-4.0, 1.0, count).astype(np.float32)
    loss, g = sel.loss_grad(wfn, t)
    g = np.array(g, dtype=np.float64)
    X = M.features("genci", 24, wfn.to_det_array())
    lref, gref = M.loss_grad(np.array(sel.params), X, t)
    assert abs(loss - lref) <= 1e-3 * lref
    # per parameter block, relative to the block's norm (same tolerance argument
    # as the forward: FP32 accumulation order and bf16 rounding boundaries)
    for a, b in [(M.OW1, M.OB1), (M.OB1, M.OW2), (M.OW2, M.OB2), (M.OB2, M.OW3), (M.OW3, M.OB3), (M.OB3, M.NPARAM)]:
        ref = gref[a:b]
        assert np.linalg.norm(g[a:b] - ref) <= 2e-2 * np.linalg.norm(ref) + 1e-7, (a, b)
    # the padded input rows get no gradient
    assert np.all(g[M.OW1 + 24 * M.HID:M.OB1] == 0)


def test_training_fits_synthetic_amplitudes():
    wfn = random_genci_wfn(24, 8, 20000, seed=9)
    X = M.features("genci", 24, wfn.to_det_array())
    # a smooth function of the occupations: log10|c| falls with the excitation energy
    w = -0.2 * np.arange(24) / 8.0
    t = X[:, :24] @ w - 1.0
    c = 10.0 ** t
    sel = pyci.mlp_selector(24, seed=1)
    y0 = np.array(sel.predict(wfn))
    loss = sel.train(wfn, c, steps=400, batch=4096, lr=2e-3, seed=3)
    y1 = np.array(sel.predict(wfn))
    before = np.mean((y0 - t) ** 2)
    after = np.mean((y1 - t) ** 2)
    print(f"mse before {before:.4f} after {after:.5f} last minibatch {loss:.5f}")
    assert after < 0.02 * before and after < 5e-3
    # training is deterministic
    sel2 = pyci.mlp_selector(24, seed=1)
    sel2.train(wfn, c, steps=400, batch=4096, lr=2e-3, seed=3)
    assert np.array_equal(np.array(sel2.params), np.array(sel.params))


def _heat_bath_scores(oracle, nb, ne, dets, cand, one, two, c):
    allv = np.concatenate([dets, cand])
    R = oracle.csr_build("genci", nb, ne, 0, allv, one, two, symmetric=False, row0=dets.size, nrow=allv.size,
                         ncol=dets.size).scipy()
    Rc = abs(R.multiply(c[None, :])).tocsr()
    return np.asarray(Rc.max(axis=1).todense()).ravel()


@pytest.mark.parametrize("score_mode", [0, 1])
def test_selection_step_matches_oracle(oracle, score_mode):
    n, m = 7, 2
    one, two = genci_problem(n)
    ham = pyci.hamiltonian(0.0, one, two)
    wfn, _ = hf_sd_wfn(n, m)
    rng = np.random.default_rng(21)
    # amplitudes spread over orders of magnitude: the pool (eps/4) is about twice
    # the plain selection (eps)
    c = rng.normal(size=len(wfn)) * np.exp(-rng.random(len(wfn)) * 5) * 1e-2
    sel = pyci.mlp_selector(2 * n, seed=4)
    sel.params = perturbed_params(sel, 5)
    dets = wfn.to_det_array().reshape(-1).copy()
    eps, pool = 5e-4, 4.0
    plain = oracle.add_hci("genci", 2 * n, 2 * m, 0, dets, one, two, c, eps)
    cand = oracle.add_hci("genci", 2 * n, 2 * m, 0, dets, one, two, c, eps / pool)
    assert plain.size > 0 and cand.size > plain.size
    # the network's scores of the pool (each row is evaluated independently)
    tmp = pyci.genci_wfn(2 * n, 2 * m, 0)
    for d in cand:
        tmp.add_det(np.array([d], dtype=np.uint64))
    y = np.array(sel.predict(tmp), dtype=np.float32)
    if score_mode == 1:
        hb = _heat_bath_scores(oracle, 2 * n, 2 * m, dets, cand, one, two, c)
        y = y + np.log10(np.maximum(hb, 1e-300)).astype(np.float32)
    expect = cand[M.topk_select(y, plain.size)]
    added = pyci.add_hci(ham, wfn, c, eps=eps, selector=sel, pool=pool, score_mode=score_mode)
    assert added == plain.size
    got = wfn.to_det_array().reshape(-1)[dets.size:]
    if score_mode == 0:
        np.testing.assert_array_equal(got, expect)
    else:
        # log10 of the heat-bath score is rounded on two different libraries:
        # at most a boundary swap
        assert np.intersect1d(got, expect).size >= expect.size - 2
    # appended determinants are indexed
    for k in range(0, added, max(1, added // 5)):
        assert wfn.index_det(wfn[dets.size + k]) == dets.size + k
    # nadd caps the count
    wfn2, _ = hf_sd_wfn(n, m)
    assert pyci.add_hci(ham, wfn2, c, eps=eps, selector=sel, pool=pool, nadd=5) == 5


def test_selector_gains_over_heat_bath(oracle):
    """Equal-count comparison: each run appends per iteration as many
    determinants as plain add_hci(eps) would; the learned ranking reaches the
    FCI energy with fewer determinants."""
    n, m = 8, 3
    one, two = genci_problem(n)
    ham = pyci.hamiltonian(0.0, one, two)
    # FCI energy of the Ms = 0 sector = FullCI with the spatial integrals
    _, h, g = seeded_integrals(n, 64, 1003, hdiag0=-3.0, hstep=0.2, decay=6.0)
    fdets = oracle.all_dets("fullci", n, m, m)
    efci = float(oracle.lowest_eigs(oracle.csr_build("fullci", n, m, m, fdets, h, g), 1)[0][0])
    schedule = [1e-2, 3e-3, 1e-3, 3e-4]
    w0, _ = hf_sd_wfn(n, m)
    hb = run_selection(ham, w0, schedule)
    w1, _ = hf_sd_wfn(n, m)
    sel = pyci.mlp_selector(2 * n, seed=0)
    ml = run_selection(ham, w1, schedule, selector=sel, pool=5.0, train_steps=300, batch=2048)
    for name, tr in (("heat-bath", hb), ("learned", ml)):
        print(name, [(t["ndet"], round(t["energy"] - efci, 8)) for t in tr])

    def dets_to_reach(tr, delta):
        for t in tr:
            if t["energy"] - efci < delta:
                return t["ndet"]
        return np.inf

    for t in hb + ml:
        assert t["energy"] >= efci - 1e-9  # variational
    assert dets_to_reach(ml, 1e-4) < dets_to_reach(hb, 1e-4)
