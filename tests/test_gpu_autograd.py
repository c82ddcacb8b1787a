"""E_RHF(R) through the libtorch autograd operator (torch.ops.m2.rhf_energy)
and the RHF gradient on the 12-atom random C/H/N/O chain: backward against
the device gradient, the oracle's compute_forces restatement and central
finite differences (h = 1e-4 bohr) of the operator's own energy."""
import numpy as np
import pytest
import torch

from molecular2_amd import molecules as M

pytestmark = pytest.mark.gpu

TIGHT = dict(enconverge=1e-12, densconverge=1e-10, maxiter=120, precision=1e-14)
# the chain's SCF has a noise floor near |dD| ~ 1e-10 (FP64 atomics in the
# J/K accumulation), so it converges to one decade above it
CHAIN = dict(enconverge=1e-11, densconverge=1e-9, maxiter=120, precision=1e-14)


def _distorted_water():
    return M.Molecule([8, 1, 1], np.array([[0.0, 0.05, -0.1], [1.45, 1.05, 0.2], [-1.5, 1.2, -0.1]]))


def test_autograd_energy_and_backward_water(engine):
    from molecular2_amd.autograd import rhf_energy
    mol = _distorted_water()
    print("case: autograd water forward/backward", flush=True)
    R = torch.tensor(mol.xyz, dtype=torch.float64, requires_grad=True)
    E = rhf_energy(R, mol, **TIGHT)
    assert E.dim() == 0 and E.dtype == torch.float64
    r = engine.rhf(mol, gradient=True, **TIGHT, guess="soad")   # the operator's default guess
    assert abs(E.item() - r["energy"]) < 1e-10
    (2.0 * E).backward()
    g = R.grad.numpy() / 2.0
    assert np.max(np.abs(g - r["grad"])) < 1e-9, np.max(np.abs(g - r["grad"]))
    # central differences of the operator's energy
    h = 1e-4
    for atom, axis in ((0, 1), (1, 0), (2, 2)):
        Rp = torch.tensor(mol.xyz, dtype=torch.float64)
        Rm = Rp.clone()
        Rp[atom, axis] += h
        Rm[atom, axis] -= h
        fd = (rhf_energy(Rp, mol, **TIGHT).item() - rhf_energy(Rm, mol, **TIGHT).item()) / (2 * h)
        assert abs(g[atom, axis] - fd) < 2e-7, (atom, axis, g[atom, axis], fd)


def test_autograd_on_device_tensor_and_no_grad(engine):
    from molecular2_amd.autograd import rhf_energy
    mol = _distorted_water()
    print("case: autograd water, R on cuda:0", flush=True)
    R = torch.tensor(mol.xyz, dtype=torch.float64, device="cuda:0", requires_grad=True)
    E = rhf_energy(R, mol, guess="hcore", **TIGHT)
    assert E.device == R.device
    E.backward()
    assert R.grad.device == R.device and R.grad.shape == (3, 3)
    assert abs(R.grad.sum(dim=0)).max().item() < 1e-8          # translational invariance
    with torch.no_grad():
        E2 = rhf_energy(R, mol, guess="hcore", **TIGHT)
        E3 = rhf_energy(R, mol, guess="soad", **TIGHT)
    assert abs(E3.item() - E.item()) < 1e-10                   # converged energy does not depend on the guess
    assert abs(E2.item() - E.item()) < 1e-11


def test_random_chain_gradient(oracle, engine):
    """12-atom bonded C/H/N/O chain (seed 12): device gradient vs the oracle's
    compute_forces at the device density, and vs central differences."""
    mol = M.random_chain(12, seed=12)
    print("case: random chain gradient natom=12", flush=True)
    r = engine.rhf(mol, pure=True, gradient=True, **CHAIN, guess="hcore")
    assert r["converged"]
    b = oracle.Basis(mol.shells("cc-pvdz"), pure=True)
    nocc = mol.nel // 2
    Co = r["C"][:, :nocc]
    W = Co @ np.diag(r["eps"][:nocc]) @ Co.T
    go = b.grad(mol, r["D"], W, precision=1e-14)
    g = r["grad"]
    assert np.max(np.abs(g - go)) < 1e-8, np.max(np.abs(g - go))
    assert np.max(np.abs(g.sum(axis=0))) < 1e-7
    h = 1e-4
    for atom, axis in ((0, 0), (5, 1), (11, 2)):
        ep = engine.rhf(mol.displaced(atom, axis, h), **CHAIN, guess="hcore")["energy"]
        em = engine.rhf(mol.displaced(atom, axis, -h), **CHAIN, guess="hcore")["energy"]
        assert abs(g[atom, axis] - (ep - em) / (2 * h)) < 5e-7, (atom, axis)


def test_random_chain_mixed_precision_error(engine):
    """Mixed-precision SCF + gradient against FP64 on the chain: |dE| < 1 uEh,
    max |dg| < 1e-6 Eh/bohr."""
    mol = M.random_chain(12, seed=12)
    print("case: random chain mixed vs fp64", flush=True)
    kw = dict(enconverge=1e-10, densconverge=1e-8, maxiter=120)
    r64 = engine.rhf(mol, gradient=True, **kw, guess="hcore")
    rmx = engine.rhf(mol, gradient=True, mixed=1, **kw, guess="hcore")
    assert r64["converged"] and rmx["converged"]
    assert abs(rmx["energy"] - r64["energy"]) < 1e-6
    assert np.max(np.abs(rmx["grad"] - r64["grad"])) < 1e-6
