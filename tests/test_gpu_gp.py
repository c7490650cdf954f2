"""GPU tests of the GP loss and its hand-written backward (libmpchol through
the C ABI) against the FP64 oracle (oracle.gp_loss_grad_f64, itself pinned by
central finite differences in tests/test_oracle.py)."""
import numpy as np
import pytest
import torch

from cholesky_gpu_amd import _lib as L
from cholesky_gpu_amd import solver as S

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def gp_case(n, d, seed=5):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, d)).astype(np.float32)
    ls = rng.uniform(0.5, 2.0, d)
    b = rng.standard_normal(n)
    return X, ls, b


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_probes_bit_exact(handle, oracle):
    say("probes")
    Zg = S.probes(handle, 1000, 5, 1234).cpu().numpy()
    assert np.array_equal(Zg, oracle.probes(1000, 5, 1234))
    E = S.probes(handle, 50, 50, 0, unit=True).cpu().numpy()
    assert np.array_equal(E, np.eye(50))


def test_rbf_assemble64_matches_fp64(handle, oracle):
    say("rbf64")
    X, ls, _ = gp_case(700, 8)
    Xd = torch.from_numpy(X).cuda()
    K, K32 = S.rbf_assemble64(handle, Xd, ls, 1.3, 1e-2, want32=True)
    A, _, _ = oracle.gp_matrix_f64(X, ls, 1.3, 1e-2)
    assert rel(K.cpu().numpy(), A) < 1e-14
    assert np.array_equal(K32.cpu().numpy(), K.cpu().numpy().astype(np.float32))
    assert torch.equal(K, K.T)


def test_pcg64_matches_fp64_solve(handle, oracle):
    """FP64 operator, mixed-precision FP32 factor as preconditioner: the
    solution matches the FP64 solve of the FP64 matrix, not of its FP32 copy."""
    say("pcg64")
    X, ls, b = gp_case(900, 8)
    Xd = torch.from_numpy(X).cuda()
    K64, K32 = S.rbf_assemble64(handle, Xd, ls, 1.0, 1e-2, want32=True)
    Lf = K32.clone()
    rc, fi, _, _ = S.factor(handle, Lf, S.factor_opts(nb=128))
    assert rc == 0
    B = torch.from_numpy(np.stack([b, np.cos(np.arange(900.0))])).cuda()
    Xs, pi, rc = S.pcg64(handle, K64, Lf, B, nb=128, opts=S.pcg_opts(tol=1e-12, maxit=300))
    assert rc == 0, S.info_dict(pi)
    A = K64.cpu().numpy()
    for j in range(2):
        ref = np.linalg.solve(A, B[j].cpu().numpy())
        assert rel(Xs[j].cpu().numpy(), ref) < 1e-9


@pytest.mark.parametrize("n,d,ard,P,nb", [(512, 8, True, 8, 128), (1000, 4, False, 16, 256),
                                         (777, 8, True, 32, 64)])
def test_gp_gradient_matches_fp64_same_probes(handle, oracle, n, d, ard, P, nb):
    """Same Rademacher probes on both sides: the GPU gradient of every
    parameter is within 1e-6 relative of the FP64 evaluation."""
    say("gp grad", n, d, ard, P, nb)
    X, ls, b = gp_case(n, d, seed=n)
    if not ard:
        ls = ls[:1]
    var, noise, seed = 1.2, 1e-2, 77
    # residual floor of an FP64-operator solve is about u64 * cond(A) (<= 4e4 here)
    o = S.gp_opts(n_probes=P, probe_seed=seed, factor=S.factor_opts(nb=nb),
                  pcg=S.pcg_opts(tol=1e-11, maxit=400))
    loss, gb, gls, gvar, gno, info = S.gp_loss_grad(handle, torch.from_numpy(X).cuda(), ls, var,
                                                    noise, torch.from_numpy(b).cuda(), o)
    Z = oracle.probes(n, P, seed)
    rl, rb, rls, rvar, rno = oracle.gp_loss_grad_f64(X, ls, var, noise, b, Z=Z)
    assert rel(gb.cpu().numpy(), rb) < 1e-6
    assert rel(gls, rls) < 1e-6, (gls, rls)
    assert abs(gvar - rvar) <= 1e-6 * abs(rvar)
    assert abs(gno - rno) <= 1e-6 * abs(rno)
    # data-fit term from the FP64-operator solve; the logdet comes from the
    # mixed-precision preconditioner factor, so it only approximates logdet(A)
    _, ldet_exact = np.linalg.slogdet(oracle.gp_matrix_f64(X, ls, var, noise)[0])
    assert abs(info.datafit - (rl - 0.5 * ldet_exact)) <= 1e-9 * abs(info.datafit)
    assert abs(info.logdet - ldet_exact) <= 1e-2 * abs(ldet_exact)


def test_gp_exact_trace_gradient(handle, oracle):
    """Unit probes (exact trace): the GPU gradient equals the exact FP64
    gradient (which the finite-difference test pins) to 1e-6."""
    say("gp exact trace")
    n, d = 300, 8
    X, ls, b = gp_case(n, d, seed=3)
    o = S.gp_opts(probe_kind=1, factor=S.factor_opts(nb=64), pcg=S.pcg_opts(tol=1e-12, maxit=400))
    loss, gb, gls, gvar, gno, info = S.gp_loss_grad(handle, torch.from_numpy(X).cuda(), ls, 0.9,
                                                    5e-2, torch.from_numpy(b).cuda(), o)
    rl, rb, rls, rvar, rno = oracle.gp_loss_grad_f64(X, ls, 0.9, 5e-2, b)
    assert rel(gls, rls) < 1e-6
    assert abs(gvar - rvar) <= 1e-6 * abs(rvar)
    assert abs(gno - rno) <= 1e-6 * abs(rno)


def test_gp_autograd_function(handle, oracle):
    say("gp autograd")
    n, d = 400, 8
    X, ls, b = gp_case(n, d, seed=9)
    lst = torch.tensor(ls, dtype=torch.float64, requires_grad=True)
    var = torch.tensor(1.1, dtype=torch.float64, requires_grad=True)
    noise = torch.tensor(2e-2, dtype=torch.float64, requires_grad=True)
    bt = torch.tensor(b, dtype=torch.float64, device="cuda", requires_grad=True)
    o = S.gp_opts(n_probes=8, probe_seed=5, pcg=S.pcg_opts(tol=1e-12, maxit=400),
                  factor=S.factor_opts(nb=128))
    out = S.gp_loss(lst, var, noise, bt, torch.from_numpy(X).cuda(), handle, o)
    (3.0 * out).backward()
    Z = oracle.probes(n, 8, 5)
    _, rb, rls, rvar, rno = oracle.gp_loss_grad_f64(X, ls, 1.1, 2e-2, b, Z=Z)
    assert rel(lst.grad.numpy(), 3.0 * rls) < 1e-6
    assert abs(var.grad.item() - 3.0 * rvar) <= 1e-6 * abs(3.0 * rvar)
    assert abs(noise.grad.item() - 3.0 * rno) <= 1e-6 * abs(3.0 * rno)
    assert rel(bt.grad.cpu().numpy(), 3.0 * rb) < 1e-9
