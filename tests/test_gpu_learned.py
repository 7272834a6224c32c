"""The learned preconditioner's GPU kernels through the C-ABI: the network's
loss and gradient against the CPU reference (tests/mlp_ref.py, whose
gradient is pinned by finite differences), inference, training, the
features and the training statistics against the same quantities formed on
the host from the oracle's SD step, and the learned factors applied as an
explicit preconditioner."""
import sys

import numpy as np
import pytest

import cpsd_amd
import mlp_ref as M
import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sample_set(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, size=(n, M.NF)).astype(np.float32)
    A = rng.uniform(0.1, 1.0, size=n)
    B = A * np.exp(rng.uniform(-1, 2, size=n))
    return X, A, B


def test_mlp_loss_grad_vs_cpu(gpu):
    say("mlp loss/grad")
    X, A, B = sample_set(5000, 1)
    p = M.init_params(3)
    p = (p + np.random.default_rng(4).normal(0, 0.1, size=p.size)).astype(np.float32)
    L, g = cpsd_amd.mlp_loss_grad(X, A, B, p)
    Lr, gr = M.loss_grad(X.astype(np.float64), A.astype(np.float32).astype(np.float64),
                         B.astype(np.float32).astype(np.float64), p.astype(np.float64))
    # FP32 sums over 5000 samples against float64
    assert abs(L - Lr) < 1e-4 * abs(Lr)
    assert np.max(np.abs(g - gr)) < 1e-4 * np.max(np.abs(gr))


def test_mlp_predict_vs_cpu(gpu):
    say("mlp predict")
    X, _, _ = sample_set(3000, 2)
    p = (M.init_params(5) + np.random.default_rng(6).normal(0, 0.2, size=M.NP)).astype(np.float32)
    y = cpsd_amd.mlp_predict(X, p)
    yr, _, _ = M.forward(X.astype(np.float64), p.astype(np.float64))
    assert np.max(np.abs(y - yr)) < 1e-5 * max(1.0, np.max(np.abs(yr)))


def test_mlp_training_fits_and_is_deterministic(gpu):
    """samples whose best factor K* = B/A is a smooth function of x: Adam
    on the GPU brings the loss near its minimum -sum B^2/A, and two runs
    give the same parameters bit for bit"""
    say("mlp training")
    rng = np.random.default_rng(8)
    X = rng.uniform(-1, 1, size=(20000, M.NF)).astype(np.float32)
    A = rng.uniform(0.5, 1.5, size=20000)
    kstar = np.exp(1.5 * np.tanh(2 * X[:, 0].astype(np.float64)) + 0.5 * X[:, 1])
    B = A * kstar
    Lmin = -np.sum(B * B / A)
    p0 = M.init_params(9)
    L0 = M.loss(X, A, B, p0)
    p1, hist = cpsd_amd.mlp_train(X, A, B, p0, 800, 3e-3)
    p2, _ = cpsd_amd.mlp_train(X, A, B, p0, 800, 3e-3)
    assert np.array_equal(p1, p2)
    L1 = M.loss(X, A, B, p1)
    assert abs(hist[0] - L0) < 1e-3 * abs(L0)
    assert (L1 - Lmin) < 0.02 * (L0 - Lmin)


def packed_index(cell):
    """grid index of each packed coefficient: the masked G in the
    reference's storage order (every row's masked kx form one interval)"""
    m = cell.masker.reshape(-1, 2)[:, 0]
    return np.nonzero(m)[0]


def test_features_vs_host(gpu):
    say("precond features")
    cell = cpsd_amd.water_cell(2, 12.0, 24, seed=7)
    dev = cpsd_amd.DeviceSD(cell, 8)
    dev.set_orbitals(cell.random_orbitals(8, seed=3, decay=0.05))
    X = dev.precond_features()
    gi = packed_index(cell)
    sz = dev.sizes()
    ng, ngp, nproj = sz["ng"], sz["ngp"], sz["nproj"]
    assert X.shape == (ng, 4) and gi.size == ng
    T = cell.tg[gi]
    Bp = dev.fetch(15)
    assert Bp.size == nproj * ngp * 2
    Bp = Bp.reshape(nproj, ngp, 2)[:, :ng]
    pn = np.concatenate([cell.vnlnrm[k * 9: k * 9 + cell.lmmax[k]] for k in cell.katm])
    D = -2.0 / cell.omega * np.sum((Bp[:, :, 0] ** 2 + Bp[:, :, 1] ** 2) / pn[:, None], axis=0)
    va = dev.fetch(13).reshape(ngp, 2)[:ng]
    V = np.sqrt(va[:, 0] ** 2 + va[:, 1] ** 2) / cell.omega
    ref = np.stack([np.log1p(T) / 5, T / (1 + T), np.arcsinh(D), np.arcsinh(10 * V)], axis=1)
    assert np.max(np.abs(X - ref)) < 2e-6 * max(1.0, np.max(np.abs(ref)))
    dev.close()


def test_stats_vs_oracle(gpu):
    """A, B of one call's last state against the same sums formed from the
    oracle's step: R = cpsi + c1 hml with cpsi = (c2 - c1)/dte - c1 lmd from
    what the reference leaves after a call"""
    say("precond stats")
    nfft, L, ne = 24, 12.0, 8
    cell = cpsd_amd.water_cell(2, L, nfft, seed=7)
    psi = cell.random_orbitals(ne, seed=3, decay=0.05)
    cinf = cell.random_orbitals(ne, seed=4, decay=0.05)
    st = O.State(O.Setup(1, L, nfft, cell.katm, cell.zv), ne, cell.lmmax, cell.vl, cell.vnl,
                 cell.vnlnrm, cell.r, psi, cell.amass)
    dev = cpsd_amd.DeviceSD(cell, ne)
    dev.set_orbitals(psi)
    dev.stats_begin(cinf)
    dev.run(3)
    dev.stats_end()
    A, B = dev.stats_read()
    dev.run(2)                       # after _end, nothing more is added
    A2, B2 = dev.stats_read()
    assert np.array_equal(A, A2) and np.array_equal(B, B2)
    dev.close()
    st.step(3)
    dte = 5.8 / np.sqrt(400000.0)
    n2 = cell.n2ft3d
    c1 = st.c1.reshape(ne, n2)
    lmd = st.lmd.reshape(ne, ne).T       # column-major
    hml = st.hml.reshape(ne, ne).T
    cpsi = st.cpsi.reshape(ne, n2) / dte - (c1.T @ lmd).T
    R = dte * (cpsi + (c1.T @ hml).T)
    Dl = cinf.reshape(ne, n2) - c1
    gi = packed_index(cell)
    Rc = R.reshape(ne, -1, 2)[:, gi]
    Dc = Dl.reshape(ne, -1, 2)[:, gi]
    Aref = np.sum(Rc ** 2, axis=(0, 2))
    Bref = np.sum(Rc * Dc, axis=(0, 2))
    assert np.max(np.abs(A - Aref)) < 1e-9 * np.max(Aref)
    assert np.max(np.abs(B - Bref)) < 1e-9 * np.max(np.abs(Bref))


def test_learned_factors_are_the_explicit_preconditioner(gpu):
    """set_learned_preconditioner = set_preconditioner with the clamped
    factors alpha exp(y) of the network's output on the cell's features"""
    say("learned = explicit")
    cell = cpsd_amd.water_cell(2, 12.0, 24, seed=7)
    psi = cell.random_orbitals(8, seed=3, decay=0.05)
    p = (M.init_params(11) + np.random.default_rng(12).normal(0, 0.3, size=M.NP)).astype(np.float32)
    alpha, kmin, kmax = 3.0, 0.5, 8.0
    out = []
    for mode in ("learned", "explicit"):
        dev = cpsd_amd.DeviceSD(cell, 8)
        dev.set_orbitals(psi)
        if mode == "learned":
            dev.set_learned_preconditioner(p, alpha, kmin, kmax)
        else:
            y = cpsd_amd.mlp_predict(dev.precond_features(), p).astype(np.float64)
            K = np.zeros(cell.nfft3d)
            K[packed_index(cell)] = np.clip(alpha * np.exp(np.minimum(y, M.YMAX)), kmin, kmax)
            dev.set_preconditioner(K)
        Es = [dev.run(5)[0] for _ in range(3)]
        out.append((np.array(Es), dev.get(c2=True)["c2"]))
        dev.close()
    assert np.max(np.abs(out[0][0] - out[1][0])) < 1e-11
    assert np.max(np.abs(out[0][1] - out[1][1])) < 1e-11


def test_learned_reaches_plain_sd_minimum_on_held_out_cell(gpu):
    """the committed model on a held-out cell (16 waters, 80^3, seed 202: a
    step-scale validation cell that no training step saw), mode 1: plain FP64
    SD and the learned run each to |dE| < 1e-7 Ha between evaluations from
    the same Gaussian start; the learned run takes under a quarter of the
    steps and its energy is within 1e-6 Ha per atom of plain SD's"""
    import learned
    say("learned on a held-out cell")
    model = learned.load_model()
    nmol, L, nfft, seed = learned.VALID_CELLS[1]
    assert (nmol, nfft, seed) == (16, 80, 202)
    cell = cpsd_amd.water_cell(nmol, L, nfft, seed=seed)
    psi = cell.random_orbitals(4 * nmol, seed=seed + 1000, decay=0.0)
    dev = cpsd_amd.DeviceSD(cell, 4 * nmol)
    ok0, n0, E0, _ = learned.converge(dev, psi)
    dev.set_learned_preconditioner(model["params"], model["alpha"], model["kmin"], model["kmax"])
    dev.set_precision(1)
    ok1, n1, E1, _ = learned.converge(dev, psi)
    fb = dev.fallback_info()["fallbacks"]
    dev.close()
    say("plain %d steps E %.9f, learned %d steps E %.9f" % (n0, E0, n1, E1))
    assert ok0 and ok1 and fb == 0
    assert n1 < n0 / 4
    assert abs(E1 - E0) / (3 * nmol) < 1e-6
