"""Backward kernel: the link gradient of S = Re <chi, D psi> and the psi
gradient, through torch.autograd.  Checked (FP64) against
  * the exact directional derivative built from the oracle Dslash: S is
    real-linear in U, so along dU = X U the derivative is Re <chi, D(X U) psi>
    with the non-unitary "gauge" X U (the reference never assumes unitarity);
  * central finite differences with step 1e-6 along random su(3) directions,
    U -> exp(eps X) U;
  * brute-force finite differences of single packed entries on a tiny lattice;
  * psi gradient = D^dagger chi from the oracle.
"""
import numpy as np
import pytest
import torch

from sse_wilson_dslash_amd.autograd import su3_force, wilson_dslash
from sse_wilson_dslash_amd.device import gen_gauge_host, gen_spinor_host
from sse_wilson_dslash_amd.layout import volume
from tests import refimpl as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")


def random_su3_algebra(n, rng):
    a = rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3))
    x = 0.5 * (a - a.conj().transpose(0, 2, 1))
    tr = np.trace(x, axis1=1, axis2=2) / 3.0
    return x - tr[:, None, None] * np.eye(3)


def expm_ah(x):
    """exp of anti-Hermitian matrices via the Hermitian eigendecomposition of -i x."""
    w, v = np.linalg.eigh(-1j * x)
    return (v * np.exp(1j * w)[:, None, :]) @ v.conj().transpose(0, 2, 1)


def action(oracle, u, psi, chi, latt, isign, cb):
    r = oracle.dslash(u, psi, np.zeros_like(psi), latt, isign, cb)
    return float(np.sum(chi * r))  # Re <chi, D psi> as a real dot product


def setup(latt, seed):
    u = gen_gauge_host(latt, 4, seed=seed)
    psi = gen_spinor_host(latt, 4, seed=seed, field=1)
    chi = gen_spinor_host(latt, 4, seed=seed, field=2)
    return u, psi, chi


@pytest.mark.parametrize("latt", [(4, 4, 4, 4), (16, 16, 16, 32)])
@pytest.mark.parametrize("isign,cb", [(1, 0), (-1, 1), (1, 1)])
def test_force_directional_derivative(oracle, latt, isign, cb):
    """Config: 16^3 x 32, FP64."""
    print(f"[force] latt={latt} isign={isign} cb={cb}", flush=True)
    u, psi, chi = setup(latt, 3)
    V = volume(latt)
    rng = np.random.default_rng(17)
    X = random_su3_algebra(V * 4, rng).reshape(V, 4, 3, 3)
    U = R.links(u)
    grad, force = su3_force(torch.from_numpy(u).cuda(), torch.from_numpy(chi).cuda(),
                            torch.from_numpy(psi).cuda(), latt, isign, cb)
    grad, force = grad.cpu().numpy(), force.cpu().numpy()
    F = force[..., 0] + 1j * force[..., 1]
    # F is traceless anti-Hermitian
    assert np.max(np.abs(F + F.conj().swapaxes(-1, -2))) < 1e-12
    assert np.max(np.abs(np.trace(F, axis1=-2, axis2=-1))) < 1e-12
    # exact derivative along dU = X U from the oracle with the non-unitary field X U
    dU = R.pack_links(X @ U)
    exact = action(oracle, np.ascontiguousarray(dU), psi, chi, latt, isign, cb)
    via_force = float(np.sum((X.conj() * F).real))          # sum Re tr(X^dag F)
    via_grad = float(np.sum(grad * dU))                      # linearity in the packed reals
    assert abs(via_force - exact) <= 1e-11 * abs(exact)
    assert abs(via_grad - exact) <= 1e-11 * abs(exact)
    # central finite differences, U -> exp(eps X) U, eps = 1e-6
    eps = 1e-6
    E = expm_ah(eps * X.reshape(-1, 3, 3)).reshape(V, 4, 3, 3)
    Em = expm_ah(-eps * X.reshape(-1, 3, 3)).reshape(V, 4, 3, 3)
    sp = action(oracle, np.ascontiguousarray(R.pack_links(E @ U)), psi, chi, latt, isign, cb)
    sm = action(oracle, np.ascontiguousarray(R.pack_links(Em @ U)), psi, chi, latt, isign, cb)
    fd = (sp - sm) / (2 * eps)
    print(f"  exact={exact:.12e} force={via_force:.12e} fd={fd:.12e}", flush=True)
    assert abs(fd - via_force) <= 1e-6 * abs(via_force)


def test_autograd_gradients(oracle):
    """torch.autograd through the Dslash: dL/dpsi = D^dagger chi, dL/du from the
    force kernel, checked against brute-force finite differences of packed
    entries on a tiny lattice."""
    latt = (2, 2, 2, 4)
    u, psi, chi = setup(latt, 5)
    isign, cb = -1, 0
    tu = torch.tensor(u, device="cuda", requires_grad=True)
    tp = torch.tensor(psi, device="cuda", requires_grad=True)
    tc = torch.tensor(chi, device="cuda")
    out = wilson_dslash(tu, tp, latt, isign, cb)
    L = (out * tc).sum()
    L.backward()
    assert abs(L.item() - action(oracle, u, psi, chi, latt, isign, cb)) < 1e-12 * abs(L.item())
    # psi gradient: D^dagger chi restricted to what feeds the target cb
    vh = volume(latt) // 2
    chi_t = chi.copy()
    chi_t[cb * vh:(cb + 1) * vh] = 0.0
    want = oracle.dslash(u, chi_t, np.zeros_like(chi), latt, -isign, 1 - cb)
    assert R.rel_err(tp.grad.cpu().numpy(), want) < 1e-13
    # link gradient: finite differences of individual packed reals
    g = tu.grad.cpu().numpy().ravel()
    rng = np.random.default_rng(2)
    h = 1e-6
    for k in rng.choice(u.size, size=40, replace=False):
        up, um = u.copy().ravel(), u.copy().ravel()
        up[k] += h
        um[k] -= h
        fd = (action(oracle, up.reshape(u.shape), psi, chi, latt, isign, cb)
              - action(oracle, um.reshape(u.shape), psi, chi, latt, isign, cb)) / (2 * h)
        assert abs(fd - g[k]) < 1e-7 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_reference_built_gradient(oracle):
    """The full packed gradient equals the one assembled entry by entry from the
    oracle Dslash (S is linear in each packed real: dS/du_k = S(e_k))."""
    latt = (2, 2, 2, 2)
    u, psi, chi = setup(latt, 9)
    isign, cb = 1, 1
    grad, _ = su3_force(torch.from_numpy(u).cuda(), torch.from_numpy(chi).cuda(),
                        torch.from_numpy(psi).cuda(), latt, isign, cb)
    grad = grad.cpu().numpy().ravel()
    built = np.empty_like(grad)
    for k in range(u.size):
        e = np.zeros(u.size)
        e[k] = 1.0
        built[k] = action(oracle, e.reshape(u.shape), psi, chi, latt, isign, cb)
    assert np.max(np.abs(built - grad)) < 1e-13 * np.max(np.abs(grad))
