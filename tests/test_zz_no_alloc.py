"""The allocation-free direct J/K entry point m2_jk_direct_device_noalloc
against m2_jk_direct_device, through the C ABI, on device buffers: two calls
in a row, the second on a smaller basis (so the static scratch array keeps
the first call's larger contents), FP64 and mixed precision; a basis too
large for the scratch array is refused with an error."""
import numpy as np
import pytest
import torch

from molecular2_amd import molecules as M

pytestmark = pytest.mark.gpu


def _sym(n, seed):
    rng = np.random.default_rng(seed)
    C = rng.normal(size=(n, max(2, n // 4)))
    return 2.0 * C @ C.T / n


def _both(engine, D, mixed):
    n = D.shape[0]
    dD = torch.tensor(D, dtype=torch.float64, device="cuda:0")
    out = [torch.full((n, n), float("nan"), dtype=torch.float64, device="cuda:0") for _ in range(4)]
    torch.cuda.synchronize()
    engine.jk_device(dD.data_ptr(), out[0].data_ptr(), out[1].data_ptr(), 1.0, 1e-12, mixed)
    engine.jk_device_noalloc(dD.data_ptr(), out[2].data_ptr(), out[3].data_ptr(), 1.0, 1e-12, mixed)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out]


@pytest.mark.parametrize("mixed", [False, True])
def test_noalloc_matches_allocating_entry_point(engine, mixed):
    for mol, seed in ((M.water_cluster(4, seed=4), 1), (M.water(), 2)):
        print("case: noalloc J/K natom=%d mixed=%s" % (mol.natom, mixed), flush=True)
        n = engine.set_basis(mol.shells("cc-pvdz"), pure=True)
        J, K, Jn, Kn = _both(engine, _sym(n, seed), mixed)
        assert np.isfinite(Jn).all() and np.isfinite(Kn).all()
        # the same kernels on the same quartets; only the order of the FP64
        # atomic additions differs between the two calls
        assert np.max(np.abs(Jn - J)) < 1e-12, np.max(np.abs(Jn - J))
        assert np.max(np.abs(Kn - K)) < 1e-12, np.max(np.abs(Kn - K))


def test_noalloc_refuses_basis_beyond_scratch(engine):
    """2100 s functions need 2 * 2100^2 doubles of accumulators, more than the
    64 MiB static array: the call fails cleanly before any kernel runs."""
    from molecular2_amd.engine import M2Error
    print("case: noalloc size limit", flush=True)
    sh = [M.Shell(0, [1.0], [1.0], (60.0 * i, 0.0, 0.0), i) for i in range(2100)]
    n = engine.set_basis(sh, pure=True)
    assert n == 2100
    bufs = [torch.zeros((n, n), dtype=torch.float64, device="cuda:0") for _ in range(3)]
    torch.cuda.synchronize()
    with pytest.raises(M2Error, match="bytes of scratch"):
        engine.jk_device_noalloc(*(b.data_ptr() for b in bufs))
    engine.set_basis(M.water().shells("cc-pvdz"), pure=True)
