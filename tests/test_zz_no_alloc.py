"""gh_cluster_forces_static (forces of one free cluster without allocation)
against gh_compute on a context without a box: two calls in a row, the second
with a smaller cluster, each bit-identical to the context's forces, per-atom
energies and total (FP64 and mixed: the same kernels and arithmetic)."""
import ctypes as C

import numpy as np
import pytest

from gems_amd import clusters as CL

pytestmark = pytest.mark.gpu
hip = pytest.importorskip("gems_amd.hip")
torch = pytest.importorskip("torch")


def _agcu(n, seed):
    rng = np.random.default_rng(seed)
    a = CL.AGCU["r0"] * np.sqrt(2.0)
    x = CL.fcc_sphere(n, a, rng.uniform(-0.5, 0.5, 3) * a) + rng.normal(0.0, 0.05, (n, 3))
    lab = np.zeros(n, np.int32)
    lab[rng.permutation(n)[: n // 2]] = 1
    return x, lab


def _context(prec, x, t):
    with hip.Context(0, prec) as c:
        c.set_box(None)
        c.set_atoms(x, t)
        c.set_masses([CL.MASS_AG, CL.MASS_CU])
        c.set_smatb(2, *CL.agcu_params())
        c.set_skin(1.0)
        return c.compute()


def _static(L, prec, x, t):
    n = len(x)
    dev = torch.device("cuda:0")
    xd = torch.from_numpy(np.ascontiguousarray(x, np.float64).reshape(-1)).to(dev)
    td = torch.from_numpy(np.ascontiguousarray(t, np.int32)).to(dev)
    F = torch.full((3 * n,), np.nan, dtype=torch.float64, device=dev)
    e = torch.full((n,), np.nan, dtype=torch.float64, device=dev)
    ep = torch.full((1,), np.nan, dtype=torch.float64, device=dev)
    st = torch.full((1,), -99, dtype=torch.int32, device=dev)
    mats = [np.ascontiguousarray(m, np.float64) for m in CL.agcu_params()]
    dp = C.POINTER(C.c_double)
    stream = torch.cuda.current_stream(dev).cuda_stream
    rc = L.gh_cluster_forces_static(
        0, prec, n, 2, C.c_void_p(xd.data_ptr()), C.c_void_p(td.data_ptr()),
        *[m.ctypes.data_as(dp) for m in mats], 1.0, C.c_void_p(F.data_ptr()),
        C.c_void_p(e.data_ptr()), C.c_void_p(ep.data_ptr()), C.c_void_p(st.data_ptr()),
        C.c_void_p(stream))
    assert rc == 0
    torch.cuda.synchronize(dev)
    assert int(st.item()) == 0
    return F.cpu().numpy().reshape(n, 3), e.cpu().numpy(), float(ep.item())


@pytest.mark.parametrize("prec", [hip.GH_FP64, hip.GH_MIXED], ids=["fp64", "mixed"])
def test_static_forces_equal_context_forces_twice(prec):
    L = hip.load()
    for n, seed in ((1500, 3), (201, 4)):  # the second call smaller than the first
        x, t = _agcu(n, seed)
        Fc, ec, etc = _context(prec, x, t)
        Fs, es, ets = _static(L, prec, x, t)
        print("n=%d max|dF| %.3e" % (n, np.max(np.abs(Fs - Fc))))
        assert np.array_equal(Fs, Fc)
        assert np.array_equal(es, ec)
        assert ets == pytest.approx(etc, rel=1e-14)


def test_static_rejects_sizes_beyond_its_maximum():
    L = hip.load()
    dp = C.POINTER(C.c_double)
    one = np.ones(4)
    args = [one.ctypes.data_as(dp)] * 7
    for n, nt in ((2049, 1), (10, 3), (0, 1)):
        rc = L.gh_cluster_forces_static(0, 0, n, nt, C.c_void_p(1), C.c_void_p(1), *args, 1.0,
                                        C.c_void_p(1), C.c_void_p(1), C.c_void_p(1),
                                        C.c_void_p(1), None)
        assert rc == -1
