"""GPU checks of the proposal network's hand-written BF16 MFMA kernels
against the CPU oracle (oracle/mlp_oracle.py, FP64 with the kernels' bf16
rounding points).  Tolerances: GEMM within FP32 accumulation error of the
exact product of the bf16 operands; forward outputs within 1e-4 of the
output scale but for rare rows; gradients within 5e-3
relative in norm (FP32 accumulation order, and rare bf16 rounding ties
decided the other way); Adam within FP32 rounding."""
import numpy as np
import pytest
import torch

import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def _bf16_dev(x):
    return torch.from_numpy(mo.bf16_bits(x).view(np.int16)).cuda().view(torch.bfloat16)


@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (70, 45, 50), (130, 256, 96), (5, 3, 7),
                                   (256, 64, 4100)])
@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_gemm_lane_layout(engine, mode, M, N, K):
    """C = A B for A stored [M][K] or transposed [K][M] and B stored [K][N] or
    transposed [N][K], ragged edges included: pins the MFMA operand and
    result lane layout."""
    print(f"gemm mode={mode} M={M} N={N} K={K}", flush=True)
    rng = np.random.default_rng(mode * 1000 + M + N + K)
    A = mo.bf16_round(rng.standard_normal((M, K)))
    B = mo.bf16_round(rng.standard_normal((K, N)))
    As = A.T.copy() if mode & 2 else A
    Bs = B.T.copy() if mode & 1 else B
    lda, ldb = As.shape[1], Bs.shape[1]
    Cg = engine.mlp_gemm(mode, _bf16_dev(As), _bf16_dev(Bs), M, N, K, lda, ldb).cpu().numpy()
    ref = A @ B
    bound = 1e-5 * (np.abs(A) @ np.abs(B)) + 1e-30
    assert np.all(np.abs(Cg - ref) <= bound)


def test_gemm_unaligned_operand(engine):
    """An operand whose base is not 16-byte aligned takes the scalar loads."""
    rng = np.random.default_rng(9)
    M, N, K = 40, 33, 64
    A = mo.bf16_round(rng.standard_normal((M, K)))
    B = mo.bf16_round(rng.standard_normal((K, N)))
    flat = _bf16_dev(np.concatenate([[0.0], A.ravel()]))
    Cg = engine.mlp_gemm(0, flat[1:], _bf16_dev(B), M, N, K, K, N).cpu().numpy()
    assert np.all(np.abs(Cg - A @ B) <= 1e-5 * (np.abs(A) @ np.abs(B)) + 1e-30)


def _params(seed):
    rng = np.random.default_rng(seed)
    P = mo.init_params(seed).astype(np.float64)
    o = mo.OFFSETS["b1"][0]
    P[o:] += rng.standard_normal(mo.N_PARAMS - o) * 0.05
    return P.astype(np.float32).astype(np.float64)


def _dev_params(engine, P):
    pt = torch.from_numpy(P.astype(np.float32)).cuda()
    pb = torch.empty(engine.mlp_param_bf16_count(), dtype=torch.bfloat16, device="cuda")
    engine.mlp_sync_bf16(pt, pb)
    return pt, pb


def test_param_layout(engine):
    assert engine.mlp_param_count() == mo.N_PARAMS
    assert engine.mlp_param_bf16_count() == mo.N_PARAMS_BF16
    P = _params(1)
    pt, pb = _dev_params(engine, P)
    bits = pb.view(torch.int16).cpu().numpy().view(np.uint16)
    assert np.array_equal(bits, mo.bf16_bits(P[:mo.N_PARAMS_BF16]))


@pytest.mark.parametrize("B", [1, 300, 4096])
def test_forward_matches_oracle(engine, B):
    print(f"mlp forward B={B}", flush=True)
    rng = np.random.default_rng(B)
    P = _params(2)
    X = mo.bf16_round(rng.standard_normal((B, mo.IN)))
    pt, pb = _dev_params(engine, P)
    y = engine.mlp_forward(pt, pb, _bf16_dev(X)).cpu().numpy()
    yo, _ = mo.forward(P, X, emulate=True)
    scale = max(1.0, float(np.sqrt(np.mean(yo ** 2))))
    err = np.abs(y - yo) / scale
    # FP32 accumulation order: ~1e-7.  A hidden activation whose FP32 value
    # sits at a bf16 rounding tie may round the other way than in FP64 and
    # move its row's output by up to ~1e-2 of the output scale; such rows
    # are rare.
    assert np.median(err) <= 1e-5
    assert np.mean(err <= 1e-4) >= 0.995
    assert err.max() <= 1e-2


@pytest.mark.parametrize("B", [777, 9000])
def test_gradient_matches_oracle(engine, B):
    """9000 rows span 36 split-K reduction chunks of 256 rows."""
    print(f"mlp grad B={B}", flush=True)
    rng = np.random.default_rng(B)
    P = _params(3)
    X = mo.bf16_round(rng.standard_normal((B, mo.IN)))
    t = rng.standard_normal(B)
    pt, pb = _dev_params(engine, P)
    g, sq = engine.mlp_grad(pt, pb, _bf16_dev(X), torch.from_numpy(t.astype(np.float32)).cuda())
    g = g.cpu().numpy().astype(np.float64)
    loss_o, go, sqo = mo.loss_and_grad(P, X, t.astype(np.float32), emulate=True)
    assert abs(sq.cpu().numpy().mean() - loss_o) <= 1e-4 * loss_o
    _, greal, _ = mo.loss_and_grad(P, X, t.astype(np.float32), emulate=False)
    for name, (o, shape) in mo.OFFSETS.items():
        n = int(np.prod(shape))
        a, b = g[o:o + n], go[o:o + n]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(f"  {name}: |g - oracle| / |oracle| = {rel:.2e}", flush=True)
        assert rel <= 5e-3, name
        # and within bf16 precision of the real-valued gradient
        assert np.linalg.norm(a - greal[o:o + n]) <= 0.1 * np.linalg.norm(greal[o:o + n]), name


def test_gradient_deterministic(engine):
    rng = np.random.default_rng(11)
    B = 10000
    P = _params(4)
    X = _bf16_dev(rng.standard_normal((B, mo.IN)))
    t = torch.from_numpy(rng.standard_normal(B).astype(np.float32)).cuda()
    pt, pb = _dev_params(engine, P)
    g1, _ = engine.mlp_grad(pt, pb, X, t)
    g2, _ = engine.mlp_grad(pt, pb, X, t)
    assert torch.equal(g1, g2)


def test_adam_matches_oracle(engine):
    rng = np.random.default_rng(12)
    n = mo.N_PARAMS
    P = rng.standard_normal(n).astype(np.float32)
    m = (rng.standard_normal(n) * 0.01).astype(np.float32)
    v = (rng.random(n) * 1e-4).astype(np.float32)
    g = rng.standard_normal(n).astype(np.float32)
    pt, mt, vt = (torch.from_numpy(a.copy()).cuda() for a in (P, m, v))
    pb = torch.empty(mo.N_PARAMS_BF16, dtype=torch.bfloat16, device="cuda")
    engine.mlp_adam(pt, pb, torch.from_numpy(g).cuda(), mt, vt, 7, 1e-3)
    Po, mo_, vo = mo.adam(P.astype(np.float64), g.astype(np.float64), m.astype(np.float64),
                          v.astype(np.float64), 7, 1e-3)
    assert np.max(np.abs(pt.cpu().numpy() - Po)) <= 1e-6
    assert np.max(np.abs(mt.cpu().numpy() - mo_)) <= 1e-6
    # beta2 = 0.999 rounded to FP32 is off by 1.3e-8, i.e. 1.3e-5 relative in
    # (1 - beta2), which scales every g^2 term of v
    assert np.max(np.abs(vt.cpu().numpy() - vo) / np.maximum(vo, 1e-12)) <= 5e-5
    bits = pb.view(torch.int16).cpu().numpy().view(np.uint16)
    assert np.array_equal(bits, mo.bf16_bits(pt.cpu().numpy()[:mo.N_PARAMS_BF16]))


def test_features_kernel(engine):
    """The feature rows restate the documented layout from the prefix walk's
    results."""
    p, q, z = 8, 7, 1.0
    n, K = 500, 100
    rng = np.random.default_rng(13)
    D = np.zeros((n, p + q), np.int32)
    D[:, :p] = rng.integers(0, 5, (n, p))
    D[:, p:] = rng.integers(0, 3, (n, q))
    D[D.sum(1) == 0, 0] = 1
    S = np.tile(np.array([1.0] * p + [2.0] * q), (n, 1))
    pre = engine.walk_f32x2(p, q, z, D, S, K, K // 4)
    X = engine.mlp_features(p, q, z, D, S, pre, K)
    got = X.float().cpu().numpy().astype(np.float64)
    r = engine.results_to_numpy(pre)
    f = np.zeros((n, 64))
    for lo, cnt, slot in ((0, p, 0), (p, q, 8)):
        for i in range(n):
            pairs = sorted(zip(D[i, lo:lo + cnt], S[i, lo:lo + cnt]), reverse=True)
            f[i, slot:slot + cnt] = [s_ / 16.0 for _, s_ in pairs]
            f[i, 16 + slot:16 + slot + cnt] = [d_ / 4.0 for d_, _ in pairs]
    f[:, 32] = z
    v = r["valid"] == 1
    sdir = np.abs(D).sum(1)
    f[:, 33] = v
    f[:, 34] = (r["n_fwd"] + r["n_inv"]) / (K * sdir)
    ls, dl, pr, lm = r["log_scale"], r["delta"], r["precision"], r["limit"]
    with np.errstate(invalid="ignore"):
        f[:, 35] = np.where(v & np.isfinite(ls), ls / K, 0)
        f[:, 36] = np.where(v & np.isfinite(dl), np.clip(dl, -4, 4) / 4, 0)
        f[:, 37] = v & np.isfinite(dl)
        f[:, 38] = np.where(v, np.where(np.isfinite(pr), np.minimum(pr, 32) / 16, 2), 0)
        f[:, 39] = np.where(v & np.isfinite(lm), np.tanh(lm), 0)
    f[:, 40] = np.where(v, r["n_limits"] / 256, 0)
    f[:, 41] = sdir / 64
    with np.errstate(divide="ignore", invalid="ignore"):
        lq = np.log(np.abs(r["q"])) + ls
        lp = np.log(np.abs(r["p"])) + ls
    f[:, 42] = np.where(v & np.isfinite(lq), np.clip(lq / 16, -4, 4), 0)
    f[:, 43] = v & np.isfinite(lq) & (lq > 0)
    f[:, 44] = np.where(v & np.isfinite(lp), np.clip(lp / 16, -4, 4), 0)
    assert v.sum() > 0
    ref = mo.bf16_round(f)
    # one bf16 ulp of slack for the float transcendental (tanh) and divisions
    assert np.all(np.abs(got - ref) <= 2.0 ** -7 * np.abs(ref) + 1e-30)


def test_features_permutation_invariant(engine):
    """Candidates that differ by a permutation of the x axes (and of the y
    axes) get identical feature rows: the pFq walk is symmetric under such
    permutations of its parameters, and the zeta start is too."""
    p, q, z = 6, 5, 1.0
    rng = np.random.default_rng(21)
    n = 64
    D = np.zeros((2 * n, p + q), np.int32)
    D[:n, :p] = rng.integers(0, 4, (n, p))
    D[:n, p:] = rng.integers(0, 3, (n, q))
    D[:n][D[:n].sum(1) == 0, 0] = 1
    for i in range(n):
        D[n + i, :p] = rng.permutation(D[i, :p])
        D[n + i, p:] = rng.permutation(D[i, p:])
    S = np.tile(np.array([1.0] * p + [2.0] * q), (2 * n, 1))
    pre = engine.walk_f64_fast(p, q, z, D, S, 100, 25)
    X = engine.mlp_features(p, q, z, D, S, pre, 100).float().cpu().numpy()
    # the structural part (slots 0..33, 41) is identical; the walk outcome
    # agrees to rounding (the axis order inside a step differs)
    assert np.array_equal(X[:n, :34], X[n:, :34]) and np.array_equal(X[:n, 41], X[n:, 41])


def test_training_reduces_loss_deterministically(engine):
    """A few hundred Adam steps on a smooth synthetic target cut the loss by
    more than 10x, and two runs from the same seed end bit-identical."""
    rng = np.random.default_rng(14)
    B = 8192
    Xn = rng.standard_normal((B, mo.IN)) * 0.5
    t = np.tanh(Xn[:, :8].sum(1)) + 0.5 * Xn[:, 8] * Xn[:, 9]
    X = _bf16_dev(Xn)
    tt = torch.from_numpy(t.astype(np.float32)).cuda()
    finals = []
    for run in range(2):
        pt, pb = _dev_params(engine, mo.init_params(0).astype(np.float64))
        m = torch.zeros_like(pt)
        v = torch.zeros_like(pt)
        ws = engine.mlp_workspace(B)
        losses = []
        for step in range(1, 301):
            g, sq = engine.mlp_grad(pt, pb, X, tt, workspace=ws)
            if step == 1 or step == 300:
                losses.append(sq.mean().item())
            engine.mlp_adam(pt, pb, g, m, v, step, 2e-3)
        print(f"  run {run}: loss {losses[0]:.4f} -> {losses[1]:.5f}", flush=True)
        assert losses[1] < 0.1 * losses[0]
        finals.append(pt.clone())
    assert torch.equal(finals[0], finals[1])
