"""Ragged Helmholtz batch (persistent kernel) vs the oracle."""

import numpy as np
import pytest

from oracle import spectral as O

pytestmark = pytest.mark.gpu


def _case(Ns, seed=11):
    from jaxfun_amd import workloads as W

    rng = np.random.default_rng(seed)
    alphas = 10.0 ** rng.uniform(-2, 4, size=len(Ns))
    loads = [W.dirichlet_load(W.forcing_coefficients(rng, int(N)), W.legendre_norms(int(N))) for N in Ns]
    return np.asarray(Ns), alphas, loads


def _ref(N, alpha, F):
    if N <= 40:
        sp1 = O.Space1D("legendre", N, True)
        M, K = sp1.mass(), O.legendre_dirichlet_stiffness(N)
        return O.kron_solve([(alpha, [M, M]), (-1.0, [K, M]), (-1.0, [M, K])], F)
    return O.helmholtz_partial_fd(N, alpha, F)


EDGE_NS = [4, 5, 6, 7, 33, 34, 64, 97, 130, 200, 255, 256, 257, 258]


def _residual(N, alpha, U, F):
    sp1 = O.Space1D("legendre", N, True)
    M, K = sp1.mass(), O.legendre_dirichlet_stiffness(N)
    R = alpha * M @ U @ M - K @ U @ M - M @ U @ K - F
    return np.linalg.norm(R) / np.linalg.norm(F)


# two FP64 implementations of the same solve differ by ~cond * eps (cond ~ N^4 / alpha);
# the FP64 residual is bounded by 10x the oracle's own residual
# nwaves selects the kernel: fp64 4/8 waves; mixed 4 = compact (two workgroups per CU, one column parity at
# a time), 8 = pair (one workgroup per CU, both column parities at once, whole-row I/O); mixed16 = the
# compact kernel with the 2-term fp16 split
@pytest.mark.parametrize("mode,nwaves,tol,rtol", [("fp64", 4, 1e-8, None), ("fp64", 8, 1e-8, None),
                                                  ("mixed", 4, 1e-6, 1e-6), ("mixed", 8, 1e-6, 1e-6),
                                                  ("mixed16", 4, 1e-6, 1e-6)])
def test_ragged_vs_oracle(dev, mode, nwaves, tol, rtol):
    import torch

    from jaxfun_amd.helmholtz import HelmholtzBatch

    Ns, alphas, loads = _case(EDGE_NS)
    hb = HelmholtzBatch(Ns, alphas, mode=mode, nwaves=nwaves, blocks_per_cu=1)
    U = hb.unpack(hb.solve(hb.pack(loads)))
    torch.cuda.synchronize()
    for N, a, F, Ub in zip(Ns, alphas, loads, U):
        Ur = _ref(int(N), a, F)
        err = O.rel_l2(Ub, Ur)
        assert err < tol, (mode, N, a, err)
        res = _residual(int(N), a, Ub, F)
        bound = rtol if rtol is not None else 10 * max(_residual(int(N), a, Ur, F), 1e-15)
        assert res < bound, (mode, N, a, res, bound)


@pytest.mark.parametrize("nwaves,bpc", [(4, 1), (4, 3), (8, 1), (8, 2)])
def test_launch_shapes_agree(dev, nwaves, bpc):
    from jaxfun_amd.helmholtz import HelmholtzBatch

    Ns, alphas, loads = _case([32, 77, 128, 256, 190], seed=12)
    base = HelmholtzBatch(Ns, alphas, mode="mixed", nwaves=nwaves, blocks_per_cu=2)
    hb = HelmholtzBatch(Ns, alphas, mode="mixed")
    hb.set_launch(nwaves, bpc)
    F = base.pack(loads)
    a = base.solve(F).cpu().numpy()
    b = hb.solve(F).cpu().numpy()
    # the same kernel computes each work item identically whatever the grid size
    assert np.array_equal(a, b)
    # the two mixed kernels differ in rounding only (each is checked against the oracle)
    other = HelmholtzBatch(Ns, alphas, mode="mixed", nwaves=12 - nwaves, blocks_per_cu=1)
    c = other.solve(F).cpu().numpy()
    assert np.linalg.norm(a - c) <= 2e-6 * np.linalg.norm(a)


def test_graph_replay(dev):
    import torch

    from jaxfun_amd.helmholtz import HelmholtzBatch

    Ns, alphas, loads = _case([40, 100, 256, 63], seed=13)
    hb = HelmholtzBatch(Ns, alphas, mode="mixed")
    F = hb.pack(loads)
    U = torch.empty_like(F)
    eager = hb.solve(F).cpu().numpy()
    hb.capture(F, U)
    for _ in range(3):
        U.fill_(0)
        hb.replay()
        torch.cuda.synchronize()
        assert np.array_equal(U.cpu().numpy(), eager)
        hb.solve(F)  # eager work between replays


def test_rejects_bad_sizes(dev):
    from jaxfun_amd.helmholtz import HelmholtzBatch

    with pytest.raises(ValueError):
        HelmholtzBatch([300], [1.0])
    with pytest.raises(ValueError):
        HelmholtzBatch([3], [1.0])
    with pytest.raises(ValueError):
        HelmholtzBatch([40], [-1.0])


@pytest.mark.parametrize("mode,nwaves", [("mixed", 4), ("mixed", 8), ("mixed16", 4)])
def test_mixed_fp64_io_vs_oracle(dev, mode, nwaves):
    import torch

    from jaxfun_amd.helmholtz import HelmholtzBatch

    Ns, alphas, loads = _case([4, 5, 34, 97, 129, 256, 258], seed=14)
    hb = HelmholtzBatch(Ns, alphas, mode=mode, io_dtype=torch.float64, nwaves=nwaves, blocks_per_cu=1)
    U = hb.unpack(hb.solve(hb.pack(loads)))
    for N, a, F, Ub in zip(Ns, alphas, loads, U):
        err = O.rel_l2(Ub, _ref(int(N), a, F))
        assert err < 1e-6, (mode, nwaves, N, a, err)


def test_io_checks(dev):
    import torch

    from jaxfun_amd.helmholtz import HelmholtzBatch

    Ns, alphas, loads = _case([34, 40], seed=15)
    hb = HelmholtzBatch(Ns, alphas, mode="mixed")
    F = hb.pack(loads)
    with pytest.raises(ValueError):
        hb.solve(F, out=torch.empty(hb.total - 1, dtype=torch.float32, device=F.device))
    with pytest.raises(TypeError):
        hb.solve(F, out=torch.empty(hb.total, dtype=torch.float64, device=F.device))
    with pytest.raises(ValueError):
        hb.solve(F[:-1].contiguous())
    with pytest.raises(ValueError):
        hb.capture(F, torch.empty(hb.total + 1, dtype=torch.float32, device=F.device))
    with pytest.raises(ValueError):
        hb.solve(F, out=F)


@pytest.mark.parametrize("mode,tol", [("fp64", 1e-10), ("mixed", 1e-6)])
def test_solve_ragged_from_inner_systems(dev, mode, tol):
    # systems assembled through the reference's operator API, solved in one
    # ragged launch, against each system's own TPMatrices.solve
    from jaxfun_amd import Div, FunctionSpace, Grad, Legendre, TensorProduct, TestFunction, TrialFunction, inner
    from jaxfun_amd.helmholtz import solve_ragged

    rng = np.random.default_rng(16)
    systems, rhs, refs = [], [], []
    for N, alpha in [(34, 0.5), (63, 1e4), (100, 0.02), (129, 7.0), (200, 300.0), (256, 0.0)]:
        D = FunctionSpace(N, Legendre, {"left": {"D": 0}, "right": {"D": 0}})
        T = TensorProduct(D, D)
        v, u = TestFunction(T), TrialFunction(T)
        A = inner(v * (alpha * u - Div(Grad(u))), sparse=True)
        b = rng.standard_normal((N - 2, N - 2))
        systems.append(A)
        rhs.append(b)
        refs.append(A.solve(b).cpu().numpy())
    U = solve_ragged(systems, rhs, mode=mode)
    for Ub, Ur in zip(U, refs):
        assert Ub.dtype.is_floating_point and Ub.shape == Ur.shape
        assert O.rel_l2(Ub.cpu().numpy(), Ur) < tol


@pytest.mark.parametrize("io", ["f32", "f64"])
def test_mixed16_power_of_two_scaling(dev, io):
    # the fp16 split scales every operand block by a power of two from its
    # largest entry, so scaling the loads by 2^k scales the solution by 2^k
    # bit for bit (FP64 recurrences and FP32 storage commute with it too);
    # a zero load gives exactly zero, and loads far from unit size keep the
    # 1e-6 accuracy
    import torch

    from jaxfun_amd.helmholtz import HelmholtzBatch

    Ns, alphas, loads = _case([6, 40, 129, 256], seed=17)
    loads[1] = np.zeros_like(loads[1])
    dt = torch.float32 if io == "f32" else torch.float64
    hb = HelmholtzBatch(Ns, alphas, mode="mixed16", io_dtype=dt)
    base = hb.solve(hb.pack(loads)).cpu().numpy()
    assert not np.any(base[hb.offsets[1]:hb.offsets[2]])
    for k in (-60, -20, 40):  # FP32 I/O stays in its normal range
        scaled = hb.solve(hb.pack([np.ldexp(F, k) for F in loads])).cpu().numpy()
        assert np.array_equal(scaled, np.ldexp(base, k)), k
    U = hb.unpack(hb.solve(hb.pack([np.ldexp(F, -60) for F in loads])))
    for i, (N, a, F) in enumerate(zip(Ns, alphas, loads)):
        if i != 1:
            assert O.rel_l2(np.ldexp(U[i], 60), _ref(int(N), a, F)) < 1e-6


def test_mixed16_launch_shapes_and_graph(dev):
    import torch

    from jaxfun_amd.helmholtz import HelmholtzBatch

    Ns, alphas, loads = _case([32, 77, 128, 256, 190], seed=18)
    hb = HelmholtzBatch(Ns, alphas, mode="mixed16", blocks_per_cu=1)
    F = hb.pack(loads)
    a = hb.solve(F).cpu().numpy()
    hb.set_launch(4, 3)
    assert np.array_equal(hb.solve(F).cpu().numpy(), a)
    with pytest.raises(ValueError):
        hb.set_launch(8, 1)
    # the two splits agree to their accuracy
    ref = HelmholtzBatch(Ns, alphas, mode="mixed").solve(F).cpu().numpy()
    assert np.linalg.norm(a - ref) <= 2e-6 * np.linalg.norm(ref)
    U = torch.empty_like(F)
    hb.capture(F, U)
    for _ in range(3):
        U.fill_(0)
        hb.replay()
        torch.cuda.synchronize()
        assert np.array_equal(U.cpu().numpy(), a)
        hb.solve(F)
