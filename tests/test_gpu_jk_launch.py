"""Launch shapes of the direct J/K build against each other and the oracle.

  * grids bounded from the basis (the default: nothing is read back before
    the end of a build) against grids from host-read work-item totals, FP64
    and mixed, on water (against the oracle's formJKdirect restatement,
    src/fockbuild.cpp:64-164) and on (H2O)4;
  * the grid bound is exact for an unscreened build: at precision 0 every
    bra pairs with every ket it may pair with, so the device-written item
    totals must add up to the bounds the host computed from the basis;
  * other chunks-per-item choices (register and table-driven classes) give
    the same J/K;
  * graph replay: one capture per key, replays interleaved with eager
    builds and other stream work, every replay equal to the eager build;
    a new key (precision, mixed) captures again; a new basis drops them;
  * the autotuner keeps candidate values and leaves J/K unchanged.
"""
import numpy as np
import pytest
import torch

from molecular2_amd import molecules as M

pytestmark = pytest.mark.gpu


def _density(n, seed):
    rng = np.random.default_rng(seed)
    C = rng.normal(size=(n, max(2, n // 4)))
    return 2.0 * C @ C.T / n


def _dev(A):
    return torch.tensor(A, dtype=torch.float64, device="cuda:0")


def _jk_dev(engine, dD, precision=1e-12, mixed=False):
    n = dD.shape[0]
    J = torch.full((n, n), float("nan"), dtype=torch.float64, device="cuda:0")
    K = torch.full((n, n), float("nan"), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    engine.jk_device(dD.data_ptr(), J.data_ptr(), K.data_ptr(), 1.0, precision, mixed)
    return J.cpu().numpy(), K.cpu().numpy()


@pytest.fixture
def restore_launch(engine):
    yield
    engine.set_jk_graphs(False)
    engine.set_jk_launch(host_totals=False)
    for X in range(6):
        for Y in range(X + 1):
            engine.set_jk_block_tasks(X, Y, 0)


def test_bounded_grids_match_host_totals_and_oracle(oracle, engine, restore_launch):
    mol = M.water()
    sh = mol.shells("cc-pvdz")
    ob = oracle.Basis(sh, pure=True)
    n = engine.set_basis(sh, pure=True)
    D = _density(n, 3)
    Jo, Ko = ob.jk_direct(D, ob.schwarz(), 1e-12, 1.0)
    dD = _dev(D)
    for mixed in (False, True):
        print("case: water bounded vs host totals, mixed=%s" % mixed, flush=True)
        engine.set_jk_launch(host_totals=True)
        Jh, Kh = _jk_dev(engine, dD, mixed=mixed)
        items_h = engine.last_stats()[2]
        engine.set_jk_launch(host_totals=False)
        Jb, Kb = _jk_dev(engine, dD, mixed=mixed)
        items_b = engine.last_stats()[2]
        assert items_h == items_b > 0
        assert np.max(np.abs(Jb - Jh)) < 1e-12 and np.max(np.abs(Kb - Kh)) < 1e-12
        tol = 1e-7 if mixed else 1e-10
        assert np.max(np.abs(Jb - Jo)) < tol and np.max(np.abs(Kb - Ko)) < tol
    mol = M.water_cluster(4, seed=4)
    print("case: (H2O)4 bounded vs host totals", flush=True)
    n = engine.set_basis(mol.shells("cc-pvdz"), pure=True)
    dD = _dev(_density(n, 4))
    engine.set_jk_launch(host_totals=True)
    Jh, Kh = _jk_dev(engine, dD)
    engine.set_jk_launch(host_totals=False)
    Jb, Kb = _jk_dev(engine, dD)
    assert np.max(np.abs(Jb - Jh)) < 1e-12 and np.max(np.abs(Kb - Kh)) < 1e-12


def test_grid_bound_is_exact_unscreened(engine, restore_launch):
    for mol in (M.water(), M.water_cluster(3, seed=3)):
        print("case: grid bound at precision 0, natom=%d" % mol.natom, flush=True)
        n = engine.set_basis(mol.shells("cc-pvdz"), pure=True)
        dD = _dev(_density(n, 5))
        _jk_dev(engine, dD, precision=0.0)
        items = engine.last_stats()[2]
        host_totals, shape = engine.jk_launch()
        assert not host_totals
        bound = sum(gb for _, gb in shape.values())
        assert items == bound, (items, bound)
        # screened builds stay within the bound
        _jk_dev(engine, dD, precision=1e-10)
        assert 0 < engine.last_stats()[2] <= bound


def test_chunks_per_item_do_not_change_jk(engine, restore_launch):
    mol = M.water_cluster(4, seed=4)
    n = engine.set_basis(mol.shells("cc-pvdz"), pure=True)
    dD = _dev(_density(n, 6))
    J0, K0 = _jk_dev(engine, dD)
    _, shape0 = engine.jk_launch()
    heavy = [(5, 5)]   # table-driven; (dp|dp), (dd|pp), (dd|ds), (dd|dp) run split register code
    for reg_bt, heavy_bt in ((16, 3), (128, 1), (1000, 16)):
        print("case: chunks per item %d (heavy %d)" % (reg_bt, heavy_bt), flush=True)
        for X, Y in shape0:
            engine.set_jk_block_tasks(X, Y, heavy_bt if (X, Y) in heavy else reg_bt)
        _, shape = engine.jk_launch()
        for (X, Y), (bt, gb) in shape.items():
            want = heavy_bt if (X, Y) in heavy else max(reg_bt, 16)
            assert bt == want, ((X, Y), bt, want)
        J, K = _jk_dev(engine, dD)
        assert np.max(np.abs(J - J0)) < 1e-12 and np.max(np.abs(K - K0)) < 1e-12


def test_graph_replay_matches_eager(engine, restore_launch):
    mol = M.water_cluster(3, seed=3)
    n = engine.set_basis(mol.shells("cc-pvdz"), pure=True)
    dD = _dev(_density(n, 7))
    J = torch.empty((n, n), dtype=torch.float64, device="cuda:0")
    K = torch.empty((n, n), dtype=torch.float64, device="cuda:0")
    c0, r0, _ = engine.jk_graph_stats()
    for step in range(5):
        # new density contents in the same buffer, eager reference first
        dD.copy_(_dev(_density(n, 100 + step)))
        engine.set_jk_graphs(False)
        Je, Ke = _jk_dev(engine, dD)
        engine.set_jk_graphs(True)
        J.fill_(float("nan"))
        K.fill_(float("nan"))
        torch.cuda.synchronize()
        engine.jk_device(dD.data_ptr(), J.data_ptr(), K.data_ptr(), 1.0, 1e-12, False)
        Jg, Kg = J.cpu().numpy(), K.cpu().numpy()
        print("case: graph replay %d max|dJ| %.1e max|dK| %.1e" % (step, np.max(np.abs(Jg - Je)),
                                                                  np.max(np.abs(Kg - Ke))), flush=True)
        assert np.max(np.abs(Jg - Je)) < 1e-12 and np.max(np.abs(Kg - Ke)) < 1e-12
        # other work on the device between replays
        torch.matmul(dD, dD).sum().item()
    # set_jk_graphs(False) dropped the graphs, so every step captured; now
    # the same key replays
    engine.set_jk_graphs(True)
    for step in range(4):
        engine.jk_device(dD.data_ptr(), J.data_ptr(), K.data_ptr(), 1.0, 1e-12, False)
        assert np.max(np.abs(J.cpu().numpy() - Je)) < 1e-12
    c1, r1, held = engine.jk_graph_stats()
    assert c1 - c0 == 5 and r1 - r0 == 4 and held == 1, (c1 - c0, r1 - r0, held)
    # three keys (precision, mixed) held side by side, each replay equal to
    # its eager build
    engine.set_jk_graphs(False)
    ref = {(1e-12, False): _jk_dev(engine, dD, 1e-12, False), (1e-10, False): _jk_dev(engine, dD, 1e-10, False),
           (1e-12, True): _jk_dev(engine, dD, 1e-12, True)}
    engine.set_jk_graphs(True)
    c1, r1, _ = engine.jk_graph_stats()
    for rep in range(2):
        for (prec, mixed), (Jr, Kr) in ref.items():
            engine.jk_device(dD.data_ptr(), J.data_ptr(), K.data_ptr(), 1.0, prec, mixed)
            torch.cuda.synchronize()
            assert np.max(np.abs(J.cpu().numpy() - Jr)) < 1e-12 and np.max(np.abs(K.cpu().numpy() - Kr)) < 1e-12
    c2, r2, held = engine.jk_graph_stats()
    assert (c2 - c1, r2 - r1, held) == (3, 3, 3)
    engine.set_basis(mol.shells("cc-pvdz"), pure=True)
    assert engine.jk_graph_stats()[2] == 0


def test_autotune_keeps_results(engine, restore_launch):
    mol = M.water_cluster(4, seed=4)
    n = engine.set_basis(mol.shells("cc-pvdz"), pure=True)
    dD = _dev(_density(n, 8))
    J0, K0 = _jk_dev(engine, dD)
    before, after = engine.jk_autotune(dD.data_ptr(), 1e-12, False, repeats=2)
    print("case: autotune (H2O)4: %.3f ms -> %.3f ms" % (before, after), flush=True)
    assert before > 0 and after > 0
    _, shape = engine.jk_launch()
    heavy = {(5, 5)}
    for (X, Y), (bt, _) in shape.items():
        assert bt in ((1, 2, 4, 8, 16) if (X, Y) in heavy else (16, 32, 64, 128, 256)), ((X, Y), bt)
    J, K = _jk_dev(engine, dD)
    assert np.max(np.abs(J - J0)) < 1e-12 and np.max(np.abs(K - K0)) < 1e-12
