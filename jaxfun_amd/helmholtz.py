"""Ragged batches of 2D Legendre-Dirichlet Helmholtz solves on one GPU.

``HelmholtzBatch(Ns, alphas)`` solves, for every b,

    (alpha_b M (x) M - K (x) M - M (x) K) U_b = F_b,     U_b, F_b of shape (N_b-2, N_b-2)

which is ``inner(v*(alpha*u - Div(Grad(u))))`` on ``TensorProduct(D, D)`` with
``D = FunctionSpace(N, Legendre, bcs=Dirichlet 0)``, solved by
``TPMatrices.solve`` in the reference (la/tpmatrix.py:429-587).  All solves run
in one persistent kernel launch (``jfx_helm_ragged``) fed by a device work
counter; the launch can be captured in a HIP graph and replayed.

Set-up (once per distinct N, as the reference's ``lu_factor`` caches): the
parity blocks of (K, M) are diagonalised on the host in FP64 with
``scipy.linalg.eigh`` (la/tpmatrix.py:1173), uploaded, and split into bf16
slices on the device.  Work items (solve, row parity) are grouped into one
queue per XCD by N, so items sharing tables run on the same L2, and ordered
largest-first within each queue; each item processes both column parities in
turn.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as sla
import torch

from . import _lib
from . import kernels as K

ITEM_DTYPE = np.dtype([("foff", "<i8"), ("voff", "<i8"), ("alpha", "<f8"), ("n", "<i4"), ("p", "<i4"),
                       ("q", "<i4"), ("a", "<i4"), ("b", "<i4"), ("AP", "<i4"), ("loff", "<i4"),
                       ("texp", "<i4")])

MODES = {"fp64": 0, "mixed": 1, "mixed16": 2}
NQUEUES = 8  # work queues: one per XCD of an MI355X (workgroups are dispatched round-robin over them)


def _parity_blocks(N: int):
    """Host FP64 eigenpairs of the parity blocks of the Legendre-Dirichlet (K, M)."""
    n = N - 2
    k = np.arange(N)
    h = 2.0 / (2 * k + 1)
    j = np.arange(n)
    # S diag(h) S^T with phi_j = P_j - P_{j+2}: bands 0 and +-2
    M = np.diag(h[:n] + h[2:n + 2]) - np.diag(h[2:n], 2) - np.diag(h[2:n], -2)
    Kd = -(4.0 * j + 6.0)
    out = []
    for p in (0, 1):
        Mp = M[p::2, p::2]
        Kp = np.diag(Kd[p::2])
        if Mp.shape[0] == 0:
            out.append((np.zeros(0), np.zeros((0, 0))))
            continue
        lam, V = sla.eigh(Kp, Mp)
        out.append((lam, V))
    return out


class HelmholtzBatch:
    """Ragged batch of Legendre-Dirichlet Helmholtz solves (see module doc).

    Args:
        Ns: per-solve number of Legendre modes N_b (unknowns (N_b-2)^2), 4 <= N_b <= 258.
        alphas: per-solve Helmholtz shifts alpha_b >= 0.
        mode: ``"fp64"`` (FP64 MFMA), ``"mixed"`` (3-term bf16 split MFMA,
            FP32 accumulation, FP64 tridiagonal recurrences with FP32-stored
            sweep values and multipliers) or ``"mixed16"`` (as ``mixed`` with
            a 2-term fp16 split of operands scaled by powers of two: three
            MFMA products instead of six; 4-wave launches only).
        io_dtype: dtype of F/U (``mixed`` defaults to float32, ``fp64`` to float64).
        nwaves, blocks_per_cu: launch shape (see :func:`autotune`); in mixed mode
            nwaves 4 runs one column parity at a time (two workgroups fit a CU),
            nwaves 8 both at once with whole-row reads of F and writes of U
            (one workgroup per CU).
        queues: number of work queues (1..16); items sharing tables share a queue.
    """

    def __init__(self, Ns, alphas, mode: str = "mixed", io_dtype=None, nwaves: int = 4,
                 blocks_per_cu: int = 2, queues: int = NQUEUES):
        self.dev = K.device()
        self.Ns = np.asarray(Ns, dtype=np.int64)
        self.alphas = np.asarray(alphas, dtype=np.float64)
        if self.Ns.shape != self.alphas.shape or self.Ns.ndim != 1:
            raise ValueError("Ns and alphas must be 1D arrays of equal length")
        if np.any(self.Ns < 4) or np.any(self.Ns > 258):
            raise ValueError("every N must be in [4, 258] (parity blocks of at most 128 unknowns)")
        if np.any(self.alphas < 0) or not np.all(np.isfinite(self.alphas)):
            raise ValueError("alphas must be finite and >= 0")
        if mode not in MODES:
            raise ValueError(f"mode must be one of {list(MODES)}")
        self.mode = mode
        self.io_dtype = io_dtype or (torch.float64 if mode == "fp64" else torch.float32)
        if mode == "fp64" and self.io_dtype != torch.float64:
            raise ValueError("fp64 mode needs float64 I/O")
        self.nwaves = nwaves
        self.blocks_per_cu = blocks_per_cu
        ns = self.Ns - 2
        self.sizes = ns * ns
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.total = int(self.offsets[-1])
        if not 1 <= queues <= 16:
            raise ValueError("queues must be in [1, 16]")
        self.queues = queues
        self._build_tables()
        self._build_items()
        self.set_launch(nwaves, blocks_per_cu)

    def set_launch(self, nwaves: int, blocks_per_cu: int) -> None:
        """Choose the launch shape (waves per workgroup, workgroups per CU)."""
        if nwaves not in (4, 8) or blocks_per_cu < 1:
            raise ValueError("nwaves must be 4 or 8 and blocks_per_cu >= 1")
        if self.mode == "mixed16" and nwaves != 4:
            raise ValueError("mixed16 runs with nwaves 4 only")
        self.nwaves, self.blocks_per_cu = nwaves, blocks_per_cu
        cus = torch.cuda.get_device_properties(self.dev).multi_processor_count
        self.nblocks = int(cus * blocks_per_cu)
        per = _lib.load().jfx_helm_scratch_doubles_per_block()
        self.scratch = torch.empty(self.nblocks * per, dtype=torch.float64, device=self.dev)
        self._graph = None

    # ------------------------------------------------------------------ set-up
    def _build_tables(self):
        uniq = sorted(set(int(N) for N in self.Ns))
        voffs, loffs = {}, {}
        Vs, VTs, lams = [], [], []
        vcur = lcur = 0
        for N in uniq:
            for p, (lam, V) in enumerate(_parity_blocks(N)):
                a = lam.shape[0]
                AP = max(32, ((a + 31) // 32) * 32)
                Vp = np.zeros((AP, AP))
                Vp[:a, :a] = V
                voffs[(N, p)] = (vcur, AP)
                loffs[(N, p)] = lcur
                Vs.append(Vp.ravel())
                VTs.append(Vp.T.copy().ravel())
                lams.append(lam)
                vcur += AP * AP
                lcur += a
        self._voffs, self._loffs = voffs, loffs
        self._texp = {}
        self.Vd = torch.tensor(np.concatenate(Vs), device=self.dev)
        self.VTd = torch.tensor(np.concatenate(VTs), device=self.dev)
        self.lam = torch.tensor(np.concatenate(lams), device=self.dev)
        self.table_bytes = (self.Vd.numel() + self.VTd.numel()) * 8
        if self.mode == "mixed":
            self.Vs = torch.empty(3 * vcur, dtype=torch.int16, device=self.dev)
            self.VTs = torch.empty(3 * vcur, dtype=torch.int16, device=self.dev)
            for (N, p), (off, AP) in voffs.items():
                for src, dst in ((self.Vd, self.Vs), (self.VTd, self.VTs)):
                    _lib.call("jfx_split_matrix", K._p(src[off:off + AP * AP]), AP, AP, 0, 3,
                              K._p(dst[3 * off:3 * off + 3 * AP * AP]), K._p(None), K._s())
        elif self.mode == "mixed16":
            # V * 2^texp (max below 2^14) as fp16 high and low terms, per table
            Vh = np.empty((2 * vcur,), dtype=np.float16)
            VTh = np.empty((2 * vcur,), dtype=np.float16)
            Vall = np.concatenate(Vs)
            for (N, p), (off, AP) in voffs.items():
                V = Vall[off:off + AP * AP]
                m = float(np.max(np.abs(V)))
                texp = 14 - int(np.frexp(m)[1]) if m > 0 else 0
                self._texp[(N, p)] = texp
                for src, dst in ((V.reshape(AP, AP), Vh), (V.reshape(AP, AP).T, VTh)):
                    x = np.ldexp(np.ascontiguousarray(src).ravel(), texp)
                    hi = x.astype(np.float16)
                    lo = (x - hi.astype(np.float64)).astype(np.float16)
                    dst[2 * off:2 * off + AP * AP] = hi
                    dst[2 * off + AP * AP:2 * off + 2 * AP * AP] = lo
            self.Vs = torch.tensor(Vh.view(np.int16), device=self.dev)
            self.VTs = torch.tensor(VTh.view(np.int16), device=self.dev)
        else:
            self.Vs = self.VTs = None

    def _build_items(self):
        items = []
        for bi, (N, alpha) in enumerate(zip(self.Ns, self.alphas)):
            n = int(N) - 2
            for p in (0, 1):
                a = (n - p + 1) // 2
                if a == 0:
                    continue
                off, AP = self._voffs[(int(N), p)]
                # one item per (solve, row parity); the kernel runs both column parities
                items.append((int(self.offsets[bi]), off, float(alpha), n, p, 0, a, (n + 1) // 2, AP,
                              self._loffs[(int(N), p)], self._texp.get((int(N), p), 0)))
        arr = np.array(items, dtype=ITEM_DTYPE)
        cost = arr["a"].astype(np.float64) ** 2 * arr["n"] + 40.0 * arr["a"] * arr["n"]
        # One queue per XCD: the items of one N (which share eigenvector tables)
        # go to one queue, split into pieces only when they alone outweigh a
        # queue's share; pieces are dealt largest first onto the lightest
        # queue.  Within a queue, largest items first.
        nq = min(self.queues, len(arr))
        load = np.zeros(nq)
        qof = np.empty(len(arr), dtype=np.int64)
        Ns = arr["n"]
        share = cost.sum() / max(nq, 1)
        pieces = []
        for v in np.unique(Ns):
            idx = np.flatnonzero(Ns == v)
            k = int(min(len(idx), max(1, np.ceil(cost[idx].sum() / share - 1e-9))))
            pieces += [(float(cost[c].sum()), c) for c in np.array_split(idx, k)]
        for c, idx in sorted(pieces, key=lambda t: -t[0]):
            g = int(np.argmin(load))
            load[g] += c
            qof[idx] = g
        order = np.lexsort((-cost, qof))
        arr = arr[order]
        bounds = np.searchsorted(qof[order], np.arange(nq + 1)).astype(np.int32)
        if ITEM_DTYPE.itemsize != _lib.load().jfx_helm_item_size():
            raise RuntimeError("HelmItem layout mismatch between Python and libjfx")
        self.items_host = arr
        self.items = torch.tensor(arr.view(np.uint8), device=self.dev)
        self.nitems = len(arr)
        self.nqueues = max(nq, 1)
        self.qbounds = torch.tensor(bounds if nq else np.zeros(2, np.int32), device=self.dev)
        self.counter = torch.zeros(self.nqueues, dtype=torch.int32, device=self.dev)

    # ------------------------------------------------------------------ packing
    def pack(self, arrays) -> torch.Tensor:
        """Concatenate per-solve (n_b, n_b) arrays into the packed device layout."""
        flat = np.concatenate([np.asarray(a, dtype=np.float64).ravel() for a in arrays])
        if flat.size != self.total:
            raise ValueError("packed size mismatch")
        t = torch.tensor(flat, device=self.dev)
        return K.cast(t, self.io_dtype)

    def unpack(self, packed: torch.Tensor) -> list[np.ndarray]:
        h = packed.detach().to(torch.float64).cpu().numpy()
        return [h[self.offsets[i]:self.offsets[i + 1]].reshape(int(self.Ns[i]) - 2, int(self.Ns[i]) - 2)
                for i in range(len(self.Ns))]

    # ------------------------------------------------------------------ solve
    def _launch(self, F: torch.Tensor, U: torch.Tensor) -> None:
        _lib.call("jfx_helm_ragged", MODES[self.mode], 1 if self.io_dtype == torch.float32 else 0,
                  self.nwaves, K._p(self.items), self.nitems, K._p(self.counter), K._p(self.qbounds),
                  self.nqueues, K._p(self.Vd),
                  K._p(self.VTd), K._p(self.Vs), K._p(self.VTs), K._p(self.lam), K._p(F), K._p(U),
                  K._p(self.scratch), self.nblocks, K._s())

    def _check_io(self, F: torch.Tensor, U: torch.Tensor | None) -> None:
        """F (and U) must be contiguous device tensors of the I/O dtype holding
        exactly the packed batch: the kernel takes no sizes and trusts them."""
        for name, t in (("F", F), ("U", U)):
            if t is None:
                continue
            K._req(t, self.io_dtype, name)
            if t.numel() != self.total:
                raise ValueError(f"{name} has {t.numel()} entries, the batch needs {self.total}")
        if U is not None and U.data_ptr() == F.data_ptr():
            raise ValueError("U must not alias F (the kernel reads F while it writes U)")

    def solve(self, F: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        self._check_io(F, out)
        U = torch.empty_like(F) if out is None else out
        self._launch(F, U)
        return U

    def capture(self, F: torch.Tensor, U: torch.Tensor) -> None:
        """Capture the whole batch (counter reset + persistent kernel) in a HIP graph."""
        self._check_io(F, U)
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self._launch(F, U)  # warm-up outside capture
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._launch(F, U)
        self._graph = g

    def replay(self) -> None:
        if self._graph is None:
            raise RuntimeError("call capture() first")
        self._graph.replay()

    # ------------------------------------------------------------------ accounting
    def algorithmic_flops(self) -> float:
        """Sum of 4 n^3 + 14 n^2 per solve (partial diagonalisation count)."""
        n = (self.Ns - 2).astype(np.float64)
        return float(np.sum(4 * n**3 + 14 * n**2))

    def mfma_flops(self) -> float:
        """Flops of the MFMA instructions one launch issues: per (item, column
        parity) two GEMMs of ceil(a/16) x ceil(b/16) output tiles over AP/32
        k-slabs (bf16 16x16x32, six products each, in mixed mode; fp16
        16x16x32, three products each, in mixed16 mode; FP64 16x16x4 over AP/4
        k-steps in fp64 mode), padding included."""
        it = self.items_host
        tot = 0.0
        for q in (0, 1):
            b = (it["n"] - q + 1) // 2
            tiles = ((it["a"] + 15) // 16) * ((b + 15) // 16)
            if self.mode != "fp64":
                nprod = 6 if self.mode == "mixed" else 3
                tot += float(np.sum(2 * (it["AP"] // 32) * tiles * nprod)) * 16 * 16 * 32 * 2
            else:
                tot += float(np.sum(2 * (it["AP"] // 4) * tiles)) * 16 * 16 * 4 * 2
        return tot

    def algorithmic_bytes(self) -> float:
        """F read + U written once each."""
        return float(2 * self.total * torch.tensor([], dtype=self.io_dtype).element_size())


# ---------------------------------------------------------------------- operator route
def legendre_dirichlet_pair(n: int) -> tuple[np.ndarray, np.ndarray]:
    """Closed-form mass M = S diag(h) S^T and stiffness K = (phi_k, phi_l'') =
    -diag(4k + 6) of phi_k = P_k - P_{k+2}, k < n (galerkin/Legendre.py:439-443,
    galerkin/composite.py:330-337)."""
    N = n + 2
    h = 2.0 / (2 * np.arange(N) + 1)
    M = np.diag(h[:n] + h[2:n + 2]) - np.diag(h[2:n], 2) - np.diag(h[2:n], -2)
    Kd = np.diag(-(4.0 * np.arange(n) + 6.0))
    return M, Kd


def helmholtz_form(A, rtol: float = 1e-11) -> tuple[float, float, int]:
    """Recognise a 2D ``TPMatrices`` / ``TPMatrix`` as
    ``kappa * (alpha M (x) M - K (x) M - M (x) K)`` in the Legendre-Dirichlet
    pair (M, K) of :func:`legendre_dirichlet_pair`: every 1D factor must be a
    combination ``a M + b K`` (checked to ``rtol``), no ``K (x) K`` term may
    remain, the two mixed terms must be equal and ``alpha >= 0``.  Returns
    ``(alpha, kappa, n)``; raises ``SolverNotApplicable`` otherwise (the
    operator of ``inner(v*(alpha*u - Div(Grad(u))), sparse=True)`` on a
    Legendre-Dirichlet tensor space, la/tpmatrix.py:386-509)."""
    from .la.matrix import SolverNotApplicable

    tps = list(getattr(A, "tpmats", [A]))
    if not tps or any(len(tp.mats) != 2 for tp in tps):
        raise SolverNotApplicable("the ragged Helmholtz batch takes 2D tensor-product operators")
    n = tps[0].mats[0].shape[0]
    if any(m.shape != (n, n) for tp in tps for m in tp.mats):
        raise SolverNotApplicable("every factor must be square and of one size")
    M, Kd = legendre_dirichlet_pair(n)
    G = np.stack([M.ravel(), Kd.ravel()], axis=1)
    coef = np.zeros((2, 2))
    for tp in tps:
        ab = []
        for m in tp.mats:
            D = np.asarray(m.todense_host())
            if np.iscomplexobj(D):
                raise SolverNotApplicable("complex factor")
            c, *_ = np.linalg.lstsq(G, D.ravel(), rcond=None)
            if np.linalg.norm(G @ c - D.ravel()) > rtol * max(np.linalg.norm(D), 1e-300):
                raise SolverNotApplicable("a factor is not a combination of the Legendre-Dirichlet mass and stiffness")
            ab.append(c)
        coef += complex(tp.coefficient).real * np.outer(ab[0], ab[1])
    mu, k01, k10, kk = coef[0, 0], coef[0, 1], coef[1, 0], coef[1, 1]
    big = np.abs(coef).max()
    if big == 0 or abs(kk) > rtol * big or abs(k01 - k10) > rtol * big or k01 == 0:
        raise SolverNotApplicable("not of the form alpha M(x)M - K(x)M - M(x)K")
    kappa = -k01
    alpha = 0.0 if abs(mu) <= rtol * big else mu / kappa  # a zero shift up to rounding is zero
    if alpha < 0 or not np.isfinite(alpha):
        raise SolverNotApplicable(f"shift alpha = {alpha} is negative")
    return float(alpha), float(kappa), int(n)


def solve_ragged(systems, rhs, mode: str = "mixed", **kw) -> list[torch.Tensor]:
    """Solve ``A_b U_b = rhs_b`` for a list of 2D Legendre-Dirichlet Helmholtz
    operators (e.g. built by ``inner(v*(alpha*u - Div(Grad(u))), sparse=True)``,
    any sizes and shifts) in one ragged launch; returns FP64 (n_b, n_b) device
    arrays, as ``TPMatrices.solve`` would one at a time.  Keyword arguments go
    to :class:`HelmholtzBatch`."""
    if len(systems) != len(rhs):
        raise ValueError("one right-hand side per system")
    forms = [helmholtz_form(A) for A in systems]
    alphas = [f[0] for f in forms]
    Ns = [f[2] + 2 for f in forms]
    hb = HelmholtzBatch(Ns, alphas, mode=mode, **kw)
    F64 = torch.empty(hb.total, dtype=torch.float64, device=hb.dev)
    for i, ((_, kappa, n), r) in enumerate(zip(forms, rhs)):
        r = K.cast(torch.as_tensor(r, device=hb.dev).to(torch.float64).contiguous(), torch.float64)
        if tuple(r.shape) != (n, n):
            raise ValueError(f"right-hand side {i} has shape {tuple(r.shape)}, the system needs {(n, n)}")
        seg = F64[int(hb.offsets[i]):int(hb.offsets[i + 1])].view(n, n)
        K.lincomb([(1.0 / kappa, r)], out=seg)  # A = kappa (...): the batch solves (...) U = rhs / kappa
    U = hb.solve(K.cast(F64, hb.io_dtype))
    U = K.cast(U, torch.float64)
    return [U[int(hb.offsets[i]):int(hb.offsets[i + 1])].view(f[2], f[2]) for i, f in enumerate(forms)]


def autotune(batch: HelmholtzBatch, F: torch.Tensor, candidates=None, reps: int = 3) -> tuple[int, int]:
    """Time each (waves per workgroup, workgroups per CU) launch shape on the
    batch itself, keep the fastest on ``batch`` and return it.  In mixed mode
    the shapes are the two kernels at their occupancy: (4, 2) two 4-wave
    workgroups per CU (one column parity at a time) and (8, 1) one 8-wave
    workgroup per CU (both column parities at once), plus (4, 1); mixed16
    has the 4-wave kernel only."""
    if candidates is None:
        candidates = {"mixed": ((4, 2), (8, 1), (4, 1)), "mixed16": ((4, 2), (4, 1))}.get(
            batch.mode, ((4, 1), (8, 1), (4, 2), (8, 2)))
    U = torch.empty_like(F)
    best, best_t = None, float("inf")
    for nw, bpc in candidates:
        batch.set_launch(nw, bpc)
        batch.solve(F, U)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            batch.solve(F, U)
        e1.record()
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) / reps
        if t < best_t:
            best, best_t = (nw, bpc), t
    batch.set_launch(*best)
    return best
