"""MI355X-native H2-matrix kernel action with FMCA's interface.

Python face of the C ABI (``include/fmca_amd.h``), mirroring FMCA's pybind11
module (py/FMCA.cpp:448-468):

    K = fmca_amd.CovarianceKernel("MATERN32", 0.5)
    H = fmca_amd.H2Matrix(K, Pr, Pc, 3, 0.8)      # Pr: (d, m), Pc: (d, n) float64
    y = H.action(x)                              # x: (n, nrhs) float64 -> (m, nrhs)

plus the C++-side name ``H2kernelActor(kernel, Peval, P, deg, eta).action(rhs)``
(FMCA/src/H2Matrix/H2kernelActor.h:21-38) and device-resident forms that take
torch tensors.  Interpolation is tensor-product Chebyshev
(FMCA/src/Interpolator/TensorProductInterpolator.h).

The compiled extension is required: importing this package without it raises
ImportError (there is no CPU fallback in the product path).
"""
from __future__ import annotations

import numpy as np

# The HIP runtime of this process must be the one torch brings when torch is
# installed: importing torch first makes the dynamic loader hand the same
# libamdhip64 to the extension (a second runtime would hide the GPU from torch).
try:  # pragma: no cover - depends on the installation
    import torch as _torch  # noqa: F401
except ImportError:
    _torch = None

try:
    from . import _fmca_amd_py as _C
except ImportError as e:  # pragma: no cover - exercised only on a broken install
    raise ImportError(
        "fmca_amd: the compiled extension is missing; run `python fmca_amd/build.py`"
    ) from e

FP64 = _C.FP64
FP32 = _C.FP32
FP32_BF16 = _C.FP32_BF16
ROWS = _C.ROWS
COLS = _C.COLS


def device_memory(reset_peak: bool = False):
    """(live, peak) device bytes held by the library's plans and workspaces."""
    return _C.device_memory(reset_peak)


def set_device_memory_limit(limit_bytes: int):
    """Cap the library's device memory (0: no cap).  Allocations beyond it
    fail with an out-of-memory error, as a full HBM would."""
    _C.set_device_memory_limit(int(limit_bytes))


KERNEL_TYPES = (
    "EXPONENTIAL", "MATERN32", "MATERN52", "MATERN72", "MATERN92", "GAUSSIAN",
    "INVMULTIQUADRIC", "MULTIQUADRIC", "BIHARMONIC2D", "BIHARMONIC3D", "TRIHARMONIC3D",
    "TPS2D",
)


class CovarianceKernel:
    """FMCA::CovarianceKernel(ktype, l=1, c=1, nu=1) (CovarianceKernel.h:50-121).

    Parameters are held by value (FMCA's copy keeps reading the original
    object's lengthscale through captured ``this``, CovarianceKernel.h:57-116).
    """

    def __init__(self, ktype: str = "EXPONENTIAL", l: float = 1.0, c: float = 1.0,
                 nu: float = 1.0):
        kt = str(ktype).upper()
        if kt not in KERNEL_TYPES:
            raise ValueError(f"desired kernel not implemented: {ktype!r}")
        self._ktype = kt
        self.l = float(l)
        self.c = float(c)
        self.nu = float(nu)
        self._dist = "EUCLIDEAN"

    def kernelType(self) -> str:
        return self._ktype

    def setDistanceType(self, dist_type: str):
        """CovarianceKernel::setDistanceType (CovarianceKernel.h:123-137):
        "EUCLIDEAN" (the default) or "GEODESIC", acos(clamp(x . y, -1, 1)) as in
        SphereClusterTree::geodesicDistance (SphereClusterTree.h:65-69).  Actors
        built from this kernel (or given it by set_kernel) take it over; their
        geodesic entries run in FP64."""
        d = str(dist_type).upper()
        if d not in ("EUCLIDEAN", "GEODESIC"):
            raise ValueError(f"desired distance not implemented: {dist_type!r}")
        self._dist = d

    def distanceType(self) -> str:
        return self._dist

    def __repr__(self):
        geo = ", GEODESIC" if self._dist == "GEODESIC" else ""
        return f"CovarianceKernel({self._ktype!r}, l={self.l}, c={self.c}, nu={self.nu}{geo})"


def _points(P, name):
    P = np.asarray(P)
    if P.dtype != np.float64:
        raise TypeError(f"{name} must be a float64 array")
    if P.ndim != 2:
        raise ValueError(f"{name} must be a (dim, npts) array")
    return np.asfortranarray(P)


INTERPOLATORS = {"tensor_product": _C.INTERP_TENSOR_PRODUCT,
                 "total_degree": _C.INTERP_TOTAL_DEGREE}
PLANNERS = {"host": _C.PLAN_HOST, "device": _C.PLAN_DEVICE}
_AUTO_PLANNER = None


def _planner(name: str) -> int:
    """"auto" plans on the GPU when one is visible, else on the host."""
    global _AUTO_PLANNER
    if name == "auto":
        if _AUTO_PLANNER is None:
            try:
                import torch
                _AUTO_PLANNER = "device" if torch.cuda.is_available() else "host"
            except Exception:
                _AUTO_PLANNER = "host"
        name = _AUTO_PLANNER
    if name not in PLANNERS:
        raise ValueError(f"planner must be 'auto' or one of {sorted(PLANNERS)}")
    return PLANNERS[name]


class H2kernelActor:
    """H2kernelActor(kernel, Peval, P, mpole_deg, eta) -> .action(rhs).

    Peval: (dim, m) targets, P: (dim, n) sources.  ``action`` takes an
    (n, nrhs) float64 array and returns the (m, nrhs) product in natural order.
    ``interpolator`` is the Interpolator of NystromMoments<Interpolator>
    (H2kernelActor.h:17-36): "tensor_product" (TensorProductInterpolator) or
    "total_degree" (TotalDegreeInterpolator, as in FMCA's own actor test,
    tests/test_H2kernelActor.cpp:20-24).  ``planner`` "device" builds the
    cluster trees and expands the block tree on the current GPU (the same plan
    as "host", which restates FMCA's ClusterTreeInitializer and computePattern
    on the CPU); "auto" (the default) takes the GPU when one is visible.
    ``plan_report()`` says where each part ran and how long planning took.
    """

    def __init__(self, kernel: CovarianceKernel, Peval, P, mpole_deg: int = 3,
                 eta: float = 0.8, interpolator: str = "tensor_product",
                 planner: str = "auto"):
        Peval = _points(Peval, "Peval")
        P = _points(P, "P")
        if interpolator not in INTERPOLATORS:
            raise ValueError(f"interpolator must be one of {sorted(INTERPOLATORS)}")
        planner_id = _planner(planner)
        self.kernel = kernel
        self.deg = int(mpole_deg)
        self.eta = float(eta)
        self.interpolator = interpolator
        self._a = _C.Actor(kernel.kernelType(), kernel.l, kernel.c, kernel.nu, Peval, P,
                           self.deg, self.eta, INTERPOLATORS[interpolator],
                           planner_id)
        self._a.set_distance_type(kernel.distanceType())
        self.m = self._a.m
        self.n = self._a.n
        self.dim = self._a.dim

    def project(self, M, precision: int = FP64):
        """H2MatrixProjector(h2mat).project(M) (H2MatrixProjector.h:21-75):
        the dense m x n matrix M (natural order) projected onto this plan's H2
        format -- M's near blocks, and per far block the Frobenius-orthogonal
        projection onto the cluster bases.  Later forward actions of that
        precision apply the projected matrix (release_assembled() returns to
        the kernel).  Returns FMCA's dual-basis check on the row and column
        trees (FMCA asserts each < 5e-10)."""
        return self._a.project(np.asfortranarray(M, dtype=np.float64), int(precision))

    def assemble(self, precision: int = FP64) -> int:
        """H2Matrix::computeH2Matrix: store every coupling matrix, dense block
        and leaf basis in device memory (returns the bytes held); the plain
        actions of that precision then stream them (operator*)."""
        return self._a.assemble(precision)

    def release_assembled(self) -> None:
        self._a.release_assembled()

    def plan_report(self) -> dict:
        """Planning seconds (tree_s, pattern_s) and where the row tree, the
        column tree and the block expansion ran (rows_tree, cols_tree, blocks:
        "host", "device", or "host-fallback" where the device was asked for
        but equal coordinates, or an admissibility test within rounding of
        eta, sent that part to the host)."""
        d = dict(self._a.plan_report())
        names = {0: "host", 1: "device", 2: "host-fallback"}
        for k in ("rows_tree", "cols_tree", "blocks"):
            d[k] = names[d[k]]
        return d

    # -- host-memory forms (FMCA semantics) ---------------------------------
    def action(self, rhs, precision: int = FP64) -> np.ndarray:
        return self._a.action(rhs, precision, False)

    def action_cluster_order(self, rhs, precision: int = FP64) -> np.ndarray:
        """H2MatrixBase::action: rhs in column-tree order, result in row-tree order."""
        return self._a.action(rhs, precision, True)

    # -- device-memory forms (torch tensors on the GPU) ----------------------
    def action_device(self, rhs, out=None, precision: int = FP64, cluster_order: bool = False,
                      stream=None):
        """rhs: float64 CUDA tensor (n,) or (n, nrhs); enqueued on the current stream."""
        import torch
        x = _as_device_matrix(rhs, self.n)
        nrhs = x.shape[1]
        if out is None:
            out = torch.empty((nrhs, self.m), dtype=torch.float64, device=x.device).t()
        _check_out(out, self.m, nrhs)
        s = stream if stream is not None else torch.cuda.current_stream(x.device)
        self._a.action_device(x.data_ptr(), x.stride(1) if nrhs > 1 else self.n, nrhs,
                              out.data_ptr(), out.stride(1) if nrhs > 1 else self.m,
                              precision, cluster_order, s.cuda_stream)
        return out if rhs.dim() == 2 else out[:, 0]

    def action_transpose_device(self, rhs, out=None, precision: int = FP64, stream=None):
        """K^T rhs: rhs (m,) or (m, nrhs) -> (n, nrhs), natural order."""
        import torch
        x = _as_device_matrix(rhs, self.m)
        nrhs = x.shape[1]
        if out is None:
            out = torch.empty((nrhs, self.n), dtype=torch.float64, device=x.device).t()
        _check_out(out, self.n, nrhs)
        s = stream if stream is not None else torch.cuda.current_stream(x.device)
        self._a.action_transpose_device(x.data_ptr(), x.stride(1) if nrhs > 1 else self.m, nrhs,
                                        out.data_ptr(), out.stride(1) if nrhs > 1 else self.n,
                                        precision, s.cuda_stream)
        return out if rhs.dim() == 2 else out[:, 0]

    def set_kernel(self, kernel: CovarianceKernel):
        """Swap the kernel (type and parameters); the plan is kept."""
        self._a.set_kernel(kernel.kernelType(), kernel.l, kernel.c, kernel.nu)
        self._a.set_distance_type(kernel.distanceType())
        self.kernel = kernel

    # -- solver ----------------------------------------------------------------
    def cg(self, b, x0=None, ridge: float = 0.0, tol: float = 1e-8, precision: int = FP32,
           refine_factor: float = 0.1, max_iter: int = 100000, max_outer: int = 1000,
           method: str = "cg", max_basis: int = 0, stream=None):
        """Solve (K + ridge I) x = b for every column of b, to
        ||b - (K + ridge I) x|| <= tol ||b|| (FMCA's pCG, PCG.h:15-44, identity
        preconditioner).

        method="cg": FMCA's CG.  precision=FP64: plain CG.  precision=FP32 or
        FP32_BF16: reliable updates -- the lower-precision matvecs drive the
        recurrence, FP64 vectors and inner products, and the iterated residual
        is replaced by the FP64 residual each time it has dropped by
        refine_factor (all columns together; the search direction is
        re-orthogonalised against the new residual).

        method="gcr": GCR over a window of up to max_basis stored pairs
        (p, (K + ridge I) p) (0: as many as fit in half the free device memory,
        at most 256), the residual minimised over the window, with the same
        replacement rule; a replaced residual is projected on the window.  The
        pairs stay consistent under lower-precision matvecs, so FP32 forms
        converge in about as many iterations as FP64.

        b, x0: float64 CUDA tensors (n,) or (n, nrhs); x0 defaults to b as in
        pCG.  Returns (x, info) with info['rel_residual'] measured with the
        FP64 action.
        """
        methods = {"cg": 0, "gcr": 1}
        if method not in methods:
            raise ValueError(f"method must be one of {sorted(methods)}")
        import torch
        bm = _as_device_matrix(b, self.n)
        nrhs = bm.shape[1]
        x = torch.empty((nrhs, self.n), dtype=torch.float64, device=bm.device).t()
        x.copy_(bm if x0 is None else _as_device_matrix(x0, self.n))
        s = stream if stream is not None else torch.cuda.current_stream(bm.device)
        info = self._a.cg_device(bm.data_ptr(), bm.stride(1) if nrhs > 1 else self.n, nrhs,
                                 x.data_ptr(), x.stride(1) if nrhs > 1 else self.n, float(ridge),
                                 float(tol), int(precision), float(refine_factor), int(max_iter),
                                 int(max_outer), method=methods[method],
                                 max_basis=int(max_basis), stream=s.cuda_stream)
        return (x if b.dim() == 2 else x[:, 0]), info

    # -- plan queries ---------------------------------------------------------
    def upload(self):
        self._a.upload()

    def indices(self, which: int) -> np.ndarray:
        return self._a.indices(which)

    def rindices(self) -> np.ndarray:
        return self._a.indices(ROWS)

    def cindices(self) -> np.ndarray:
        return self._a.indices(COLS)

    def statistics(self) -> np.ndarray:
        return self._a.statistics()

    def info(self) -> dict:
        return self._a.info()

    def set_profiling(self, enable: bool = True):
        """Record HIP events around the six phases of every action."""
        self._a.set_profiling(bool(enable))

    def phase_times(self, reset: bool = False, precision: int | None = None) -> dict:
        """{phase: (total ms, number of actions)} of the profiled actions, of
        one precision (FP64 / FP32) or of both (None)."""
        return self._a.phase_times(reset, -1 if precision is None else int(precision))

    def clusters(self, which: int) -> dict:
        d = self._a.clusters(which)
        bb = d.pop("bb")
        d["bbmin"], d["bbmax"], d["bbsize"] = bb[:, 0, :], bb[:, 1, :], bb[:, 2, :]
        return d

    def blocks(self):
        r, c, lr = self._a.blocks()
        return r, c, lr.astype(bool)

    def interpolation(self) -> dict:
        """The interpolator's data: q, multi_indices (q, dim) in FMCA's order,
        Xi (dim, q) nodes on [0,1]^dim and invV (q, q); see
        fmca_h2_interpolation."""
        d = self._a.interpolation()
        d["interpolator"] = {v: k for k, v in INTERPOLATORS.items()}[d["interpolator"]]
        return d

    def launch_order(self, which: str) -> np.ndarray:
        """(workgroups, 4): target (cluster id for "coupling", leaf index for
        "leaf_output"), list begin, list end, flags (1 split, 2 far field) of
        each workgroup; rows of -1 are idle workgroups."""
        return self._a.launch_order({"coupling": 0, "leaf_output": 1}[which])


class H2Matrix(H2kernelActor):
    """FMCA.H2Matrix(CovarianceKernel, Pr, Pc, p=3, eta=0.8) (py/FMCA.cpp:209-250).

    ``action(rhs)`` permutes into the column tree, applies H2MatrixBase::action
    and permutes back, i.e. it returns K(Pr, Pc) @ rhs in natural order.  As in
    FMCA's binding (py/FMCA.cpp:37-38) the moments are total-degree Legendre
    moments (TotalDegreeInterpolator) unless ``interpolator`` says otherwise.
    """

    def __init__(self, cov: CovarianceKernel, Pr, Pc, p: int = 3, eta: float = 0.8,
                 interpolator: str = "total_degree", planner: str = "auto"):
        super().__init__(cov, Pr, Pc, p, eta, interpolator, planner)


class MultipoleFunctionEvaluator:
    """FMCA::MultipoleFunctionEvaluator(kernel, P, Peval, eta=0, mpole_deg=3)
    (FMCA/src/KernelInterpolation/MultipoleFunctionEvaluator.h:17-67):
    out-of-sample evaluation K(Peval, P) C through the H2 action with
    TotalDegreeInterpolator moments.  ``evaluate(P, Peval, C)`` returns the
    (m, ncols) values in natural order; point sets other than the ones of
    ``init`` are planned anew."""

    def __init__(self, kernel: CovarianceKernel = None, P=None, Peval=None, eta: float = 0.0,
                 mpole_deg: int = 3):
        self._h = None
        if kernel is not None:
            self.init(kernel, P, Peval, eta, mpole_deg)

    def init(self, kernel: CovarianceKernel, P, Peval, eta: float = 0.0, mpole_deg: int = 3):
        self.kernel = kernel
        self.eta = float(eta) if eta >= 0 else 0.0  # MultipoleFunctionEvaluator.h:37
        self.mpole_deg = int(mpole_deg)
        self._P = _points(P, "P").copy(order="F")
        self._Peval = _points(Peval, "Peval").copy(order="F")
        self._h = H2kernelActor(kernel, self._Peval, self._P, self.mpole_deg, self.eta,
                                interpolator="total_degree")

    def evaluate(self, P, Peval, C, precision: int = FP64) -> np.ndarray:
        if self._h is None:
            raise RuntimeError("MultipoleFunctionEvaluator: init() first")
        P, Peval = _points(P, "P"), _points(Peval, "Peval")
        if not (np.array_equal(P, self._P) and np.array_equal(Peval, self._Peval)):
            self.init(self.kernel, P, Peval, self.eta, self.mpole_deg)
        return self._h.action(C, precision)


class SampletTree:
    """FMCA's Python SampletTree(P, dtilde) (py/FMCA.cpp:51-118): an
    H2SampletTree on the ClusterTree with TotalDegree moments of degree
    2 (dtilde - 1) and samplets with dtilde vanishing moments.  Built on the
    host; the compression's kernel blocks run on the GPU."""

    def __init__(self, P=None, dtilde: int = 1):
        self._t = None
        if P is not None:
            P = _points(P, "P")
            self._t = _C.SampletTreeHandle(P, int(dtilde))
            self.dtilde = max(int(dtilde), 1)
            self.p = 2 * (self.dtilde - 1)
            self._perm = self._t.indices()
            self._map = None

    @property
    def n(self) -> int:
        return self._t.n

    def indices(self) -> np.ndarray:
        """Cluster position -> point index (pySampletTree::indices)."""
        return self._perm.copy()

    def levels(self) -> np.ndarray:
        """Level of each samplet coefficient (sampletLevelMapper)."""
        return self._t.levels()

    def toClusterOrder(self, mat) -> np.ndarray:
        mat = np.asarray(mat, dtype=np.float64)
        return np.asfortranarray(mat[self._perm])

    def toNaturalOrder(self, mat) -> np.ndarray:
        mat = np.asarray(mat, dtype=np.float64)
        out = np.empty_like(mat)
        out[self._perm] = mat
        return np.asfortranarray(out)

    def clusters(self) -> dict:
        """Per cluster: begin, size, sons, qdim, nscal, start, plus q and mdtilde."""
        return self._t.clusters()

    def basis(self, cluster: int):
        """(Q, V) of one cluster: samplet basis and multiscale cluster basis."""
        return self._t.basis(int(cluster))

    def coeff2indices(self, i: int) -> np.ndarray:
        """Point indices of the cluster that owns coefficient i
        (pySampletTree::coeff2indices)."""
        if self._map is None:
            cl = self.clusters()
            owner = np.empty(self.n, dtype=np.int64)
            for c in range(len(cl["begin"])):
                s0 = int(cl["start"][c])
                cnt = int(cl["qdim"][c] - cl["nscal"][c]) + (int(cl["nscal"][c]) if c == 0 else 0)
                owner[s0:s0 + cnt] = c
            self._map = (owner, cl["begin"], cl["size"])
        owner, begin, size = self._map
        c = owner[int(i)]
        return self._perm[begin[c]:begin[c] + size[c]].astype(np.int64)


def sampletTransform(tree: SampletTree, data) -> np.ndarray:
    """Samplet coefficients of data given in cluster order (n x k)."""
    return tree._t.transform(np.asfortranarray(np.asarray(data, dtype=np.float64)), False)


def inverseSampletTransform(tree: SampletTree, data) -> np.ndarray:
    """Data in cluster order from samplet coefficients (n x k)."""
    return tree._t.transform(np.asfortranarray(np.asarray(data, dtype=np.float64)), True)


class SampletKernelCompressor:
    """FMCA's Python SampletKernelCompressor(hst, ker, P, eta=0.8, thres=0)
    (py/FMCA.cpp:256-330): the kernel matrix in samplet coordinates,
    compressed with CompareCluster at eta; entries |v| <= thres dropped (the
    diagonal kept), upper triangle.  P is accepted for the signature's sake;
    the tree holds the points."""

    def __init__(self, hst: SampletTree = None, ker: CovarianceKernel = None, P=None,
                 eta: float = 0.8, thres: float = 0.0):
        self._m = None
        if hst is not None:
            self.init(hst, ker, P, eta, thres)

    def init(self, hst: SampletTree, ker: CovarianceKernel, P=None, eta: float = 0.8,
             thres: float = 0.0):
        if P is not None and _points(P, "P").shape[1] != hst.n:
            raise ValueError("P does not match the tree")
        self.n = hst.n
        self.eta = float(eta)
        self.thres = float(thres)
        if ker.distanceType() != "EUCLIDEAN":
            raise ValueError("the samplet compressor takes the Euclidean distance")
        self._m = hst._t.compress(ker.kernelType(), ker.l, ker.c, ker.nu, self.eta, self.thres)
        return self

    def triplets(self):
        """(rows, cols, values) in samplet coordinates, rows <= cols."""
        return self._m.triplets()

    def aposteriori_triplets_fast(self, thres: float):
        """aposteriori_triplets_fast: the triplets kept at relative threshold thres."""
        r, c, v = self._m.triplets()
        keep = self._m.aposteriori(float(thres))
        return r[keep], c[keep], v[keep]

    def statistics(self) -> dict:
        keys = ["pattern_entries", "dense_jobs", "far_jobs", "job_elements", "program_nodes",
                "plan_s", "upload_s", "gpu_s", "download_s", "combine_s", "triplets_s",
                "triplets"]
        return dict(zip(keys, self._m.stats()))

    def matrix(self):
        """The compressed matrix as scipy.sparse.csc_matrix (upper triangle)."""
        import scipy.sparse as sp
        r, c, v = self._m.triplets()
        return sp.csc_matrix((v, (r, c)), shape=(self.n, self.n))


def _as_device_matrix(t, rows):
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float64:
        raise TypeError("expected a float64 CUDA tensor")
    x = t if t.dim() == 2 else t.reshape(-1, 1)
    if x.shape[0] != rows:
        raise ValueError(f"expected {rows} rows, got {x.shape[0]}")
    if x.shape[1] > 1 and (x.stride(0) != 1 or x.stride(1) < rows):
        x = x.t().contiguous().t()  # column-major copy
    elif x.shape[1] == 1 and not x.is_contiguous():
        x = x.contiguous()
    return x


def _check_out(out, rows, nrhs):
    import torch
    if out.dtype != torch.float64 or not out.is_cuda or tuple(out.shape) != (rows, nrhs):
        raise ValueError("out must be a float64 CUDA tensor of shape (rows, nrhs)")
    if nrhs > 1 and (out.stride(0) != 1 or out.stride(1) < rows):
        raise ValueError("out must be column-major")


__all__ = ["CovarianceKernel", "H2Matrix", "H2kernelActor", "FP64", "FP32", "FP32_BF16", "ROWS", "COLS",
           "KERNEL_TYPES", "INTERPOLATORS", "MultipoleFunctionEvaluator", "SampletTree",
           "SampletKernelCompressor", "sampletTransform", "inverseSampletTransform"]
