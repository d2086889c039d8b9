"""torch.autograd wrapper (icm_amd.torch_ops) over the C-ABI forward and adjoint."""
import math

import numpy as np
import pytest

import icm_amd
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_autograd_matches_c_abi():
    import torch
    from icm_amd.torch_ops import icm_equity
    rng = np.random.default_rng(21)
    n, k = 3000, 450
    S_np = np.exp(rng.uniform(0.0, math.log(1e5), n))
    p_np = 0.9 ** np.arange(k)
    p_np /= p_np.sum()
    v_np = rng.normal(size=n)
    S = torch.tensor(S_np, dtype=torch.float64, device="cuda", requires_grad=True)
    p = torch.tensor(p_np, dtype=torch.float64, device="cuda")
    E = icm_equity(S, p, 256)
    (g,) = torch.autograd.grad((E * torch.tensor(v_np, device="cuda")).sum(), S)
    eq_ref, ds_ref = icm_amd.equity_vjp(S_np, p_np, v_np, Q=256)
    cpu = O.equity(S_np, p_np, Q=256, fft_min=64)
    assert np.max(np.abs(E.detach().cpu().numpy() - cpu) / np.abs(cpu)) < 1e-8
    np.testing.assert_allclose(g.cpu().numpy(), ds_ref, rtol=1e-12, atol=1e-20)


def test_full_jacobian_n64_v1():
    """Full 64 x 64 Jacobian by rows (64 adjoint passes) vs the V1 closed form."""
    import torch
    from icm_amd.torch_ops import icm_equity
    rng = np.random.default_rng(64)
    n = 64
    S_np = rng.uniform(1.0, 100.0, n)
    p_np = np.arange(n, 0, -1, dtype=float)
    S = torch.tensor(S_np, dtype=torch.float64, device="cuda", requires_grad=True)
    p = torch.tensor(p_np, dtype=torch.float64, device="cuda")
    E = icm_equity(S, p, 256)
    rows = []
    for i in range(n):
        e_i = torch.zeros(n, dtype=torch.float64, device="cuda")
        e_i[i] = 1.0
        (r,) = torch.autograd.grad(E, S, grad_outputs=e_i, retain_graph=True)
        rows.append(r.cpu().numpy())
    J = np.stack(rows)                       # J[i, j] = dE_i / dS_j
    Si, Sj = S_np[:, None], S_np[None, :]
    den = (Si + Sj) ** 2
    exact = -Si / den
    np.fill_diagonal(exact, 0.0)
    np.fill_diagonal(exact, (np.where(~np.eye(n, dtype=bool), Sj / den, 0.0)).sum(axis=1))
    err = np.max(np.abs(J - exact)) / np.max(np.abs(exact))
    print(f"jacobian 64x64 max rel err {err:.2e}")
    assert err < 1e-8
