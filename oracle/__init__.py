"""FP64 CPU oracle for the H2-matrix kernel action (TEST INFRASTRUCTURE).

A restatement of FMCA's matrix-free H2 action in C++ (``h2_oracle.cpp``),
loaded here through ctypes.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` may import this package, and only as the
checker: the product path (``fmca_amd``) never calls it.

Reference anchors (FMCA repository paths):
  H2kernelActor.action        FMCA/src/H2Matrix/H2kernelActor.h:38-80
  H2MatrixBase.action         FMCA/src/H2Matrix/H2MatrixBase.h:113-152
  ClusterTree                 FMCA/src/Clustering/ClusterTreeInitializer_ClusterTree.h:24-132
  computePattern              FMCA/src/H2Matrix/H2MatrixBase.h:160-191
  CovarianceKernel            FMCA/src/CovarianceKernel/CovarianceKernel.h:50-151
  pCG                         FMCA/src/util/PCG.h:15-44

Pinning status of each oracle function (what checks it in tests/test_oracle.py):
  kernels          closed forms, scipy's Matern-nu formula, k(0) = 1
  geodesic         the great-circle distance 2 asin(|x - y| / 2) on the unit
                   sphere, orthogonal / antipodal / coincident points, the
                   dense matrix against numpy's arccos(clip(P^T P)), and the H2
                   action against it, small and falling in degree
  interpolation    closed-form nodes/weights, Lagrange delta property,
                   partition of unity, polynomial reproduction
  tree             bijective permutation, leaf sizes in [q, 2q-1], children
                   partition the parent range, split-axis order, tight boxes
                   with the flat-box rule, level-order ids, hand-checked tiny tree
  transfers        FMCA's nestedness check < 1e-13 (the reference's own KAT)
  pattern          every (row point, column point) pair covered exactly once;
                   admissible blocks satisfy the eta condition
  action           dense K*x error small and decreasing in degree and eta;
                   cluster-order and natural-order forms agree; per-leaf
                   restriction equals the full action
  pCG              residual of the returned solution against the action
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

KERNELS = [
    "EXPONENTIAL", "MATERN32", "MATERN52", "MATERN72", "MATERN92", "GAUSSIAN",
    "INVMULTIQUADRIC", "MULTIQUADRIC", "BIHARMONIC2D", "BIHARMONIC3D",
    "TRIHARMONIC3D", "TPS2D",
]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ -fopenmp)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []),
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, dp, ip, lp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), \
            ctypes.POINTER(ctypes.c_int), ctypes.c_long
        L.orc_parse_kernel.argtypes = [ctypes.c_char_p]
        L.orc_parse_kernel.restype = ctypes.c_int
        L.orc_geodesic_flag.argtypes = []
        L.orc_geodesic_flag.restype = ctypes.c_int
        L.orc_kernel_eval.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, vp, lp, vp]
        L.orc_interp.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp,
                                 vp, vp]
        L.orc_interp.restype = ctypes.c_int
        L.orc_eval_polynomials.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, lp,
                                           vp]
        L.orc_create.argtypes = [ctypes.c_int, lp, vp, lp, vp, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                 ctypes.c_double, ctypes.c_int]
        L.orc_create.restype = vp
        L.orc_destroy.argtypes = [vp]
        L.orc_setup_seconds.argtypes = [vp]
        L.orc_setup_seconds.restype = ctypes.c_double
        L.orc_q.argtypes = [vp]
        L.orc_q.restype = ctypes.c_int
        L.orc_num_clusters.argtypes = [vp, ctypes.c_int]
        L.orc_num_clusters.restype = ctypes.c_int
        L.orc_indices.argtypes = [vp, ctypes.c_int, vp]
        L.orc_clusters.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp, vp]
        L.orc_num_blocks.argtypes = [vp]
        L.orc_num_blocks.restype = lp
        L.orc_blocks.argtypes = [vp, vp, vp, vp]
        L.orc_leaf_basis.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
        L.orc_leaf_basis.restype = ctypes.c_int
        L.orc_transfer.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
        L.orc_transfer.restype = ctypes.c_int
        L.orc_check_transfer.argtypes = [vp, ctypes.c_int]
        L.orc_check_transfer.restype = ctypes.c_double
        L.orc_action.argtypes = [vp, vp, lp, vp, ctypes.c_int]
        L.orc_action.restype = ctypes.c_int
        L.orc_action_cluster_order.argtypes = [vp, vp, lp, vp, ctypes.c_int]
        L.orc_action_cluster_order.restype = ctypes.c_int
        L.orc_action_leaves.argtypes = [vp, vp, lp, vp, ctypes.c_int, vp, ctypes.c_int]
        L.orc_action_leaves.restype = ctypes.c_int
        L.orc_dense_action.argtypes = [ctypes.c_int, lp, vp, lp, vp, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       vp, lp, vp, ctypes.c_int]
        L.orc_dense_matrix.argtypes = [ctypes.c_int, lp, vp, lp, vp, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double, vp]
        L.orc_pcg.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                              vp, ctypes.c_int]
        L.orc_pcg.restype = ctypes.c_int
        L.orc_entry_counts.argtypes = [vp, vp]
        L.orc_action_sample.argtypes = [vp, vp, lp, ctypes.c_double, ctypes.c_int, vp]
        L.orc_action_sample.restype = ctypes.c_double
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def kernel_id(name: str, distance: str = "EUCLIDEAN") -> int:
    """Kernel type id; distance "GEODESIC" sets the oracle's geodesic bit
    (CovarianceKernel::setDistanceType, CovarianceKernel.h:123-137)."""
    k = lib().orc_parse_kernel(name.encode())
    if k < 0:
        raise ValueError(f"unknown kernel {name!r}")
    d = distance.upper()
    if d == "GEODESIC":
        return k | lib().orc_geodesic_flag()
    if d != "EUCLIDEAN":
        raise ValueError(f"unknown distance {distance!r}")
    return k


def kernel_eval(name: str, r, l=1.0, c=1.0, nu=1.0) -> np.ndarray:
    r = np.ascontiguousarray(r, dtype=np.float64).ravel()
    out = np.empty_like(r)
    lib().orc_kernel_eval(kernel_id(name), l, c, nu, _ptr(r), r.size, _ptr(out))
    return out


SCHEMES = {"tensor": 0, "total_degree": 1}


def num_basis(dim: int, deg: int, scheme: str = "tensor") -> int:
    """q: (deg+1)^dim (tensor) or binomial(deg+dim, dim) (total degree)."""
    if scheme == "tensor":
        return (deg + 1) ** dim
    return math.comb(deg + dim, dim)


def interp(dim: int, deg: int):
    """Chebyshev nodes, weights, multi-indices (q x dim) and Xi (dim x q)."""
    q = (deg + 1) ** dim
    xi = np.empty(deg + 1)
    w = np.empty(deg + 1)
    mi = np.empty((q, dim), dtype=np.int32)
    Xi = np.empty((q, dim))  # row k = column k of FMCA's Xi
    lib().orc_interp(dim, deg, 0, _ptr(xi), _ptr(w), _ptr(mi), _ptr(Xi), None, None)
    return xi, w, mi, Xi.T.copy()


def interp_total_degree(dim: int, deg: int):
    """Total-degree data: multi-indices (q x dim), DeMarchi points Xi (dim x q),
    V (q x q, V[i, a] = P_a(Xi[:, i])) and invV."""
    q = num_basis(dim, deg, "total_degree")
    mi = np.empty((q, dim), dtype=np.int32)
    Xi = np.empty((q, dim))
    V = np.empty((q, q), order="F")
    invV = np.empty((q, q), order="F")
    lib().orc_interp(dim, deg, 1, None, None, _ptr(mi), _ptr(Xi), _ptr(V), _ptr(invV))
    return mi, Xi.T.copy(), V, invV


def eval_polynomials(dim: int, deg: int, pts: np.ndarray, scheme: str = "tensor") -> np.ndarray:
    """pts: (dim, npts) -> (q, npts): Lagrange products (tensor) or Legendre
    products (total degree)."""
    pts = np.asfortranarray(pts, dtype=np.float64)
    q = num_basis(dim, deg, scheme)
    out = np.empty((q, pts.shape[1]), order="F")
    lib().orc_eval_polynomials(dim, deg, SCHEMES[scheme], _ptr(pts), pts.shape[1], _ptr(out))
    return out


def dense_action(P, Peval, kernel, l, rhs, c=1.0, nu=1.0, nthreads=0, distance="EUCLIDEAN"):
    P = np.asfortranarray(P, dtype=np.float64)
    Peval = np.asfortranarray(Peval, dtype=np.float64)
    rhs = np.asfortranarray(rhs.reshape(P.shape[1], -1), dtype=np.float64)
    out = np.empty((Peval.shape[1], rhs.shape[1]), order="F")
    lib().orc_dense_action(P.shape[0], P.shape[1], _ptr(P), Peval.shape[1], _ptr(Peval),
                           kernel_id(kernel, distance), l, c, nu, _ptr(rhs), rhs.shape[1],
                           _ptr(out), nthreads)
    return out


def dense_matrix(P, Peval, kernel, l, c=1.0, nu=1.0, distance="EUCLIDEAN"):
    P = np.asfortranarray(P, dtype=np.float64)
    Peval = np.asfortranarray(Peval, dtype=np.float64)
    out = np.empty((Peval.shape[1], P.shape[1]), order="F")
    lib().orc_dense_matrix(P.shape[0], P.shape[1], _ptr(P), Peval.shape[1], _ptr(Peval),
                           kernel_id(kernel, distance), l, c, nu, _ptr(out))
    return out


class OracleH2:
    """FMCA's H2kernelActor(kernel, Peval, P, deg, eta) restated on the CPU.
    scheme: "tensor" (NystromMoments<TensorProductInterpolator>) or
    "total_degree" (NystromMoments<TotalDegreeInterpolator>)."""

    def __init__(self, P, Peval, kernel: str, l: float, deg: int, eta: float,
                 c: float = 1.0, nu: float = 1.0, scheme: str = "tensor",
                 distance: str = "EUCLIDEAN"):
        self.P = np.asfortranarray(P, dtype=np.float64)
        self.Peval = np.asfortranarray(Peval, dtype=np.float64)
        self.dim = self.P.shape[0]
        self.n = self.P.shape[1]
        self.m = self.Peval.shape[1]
        L = lib()
        self._h = L.orc_create(self.dim, self.n, _ptr(self.P), self.m, _ptr(self.Peval),
                               kernel_id(kernel, distance), l, c, nu, deg, eta,
                               SCHEMES[scheme])
        if not self._h:
            raise ValueError("oracle: invalid arguments")
        self.q = L.orc_q(self._h)
        self.setup_seconds = L.orc_setup_seconds(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    # which: 0 = rows (Peval), 1 = columns (P)
    def indices(self, which: int) -> np.ndarray:
        n = self.m if which == 0 else self.n
        out = np.empty(n, dtype=np.uint32)
        lib().orc_indices(self._h, which, _ptr(out))
        return out

    def clusters(self, which: int) -> dict:
        nc = lib().orc_num_clusters(self._h, which)
        begin = np.empty(nc, dtype=np.int64)
        size = np.empty(nc, dtype=np.int64)
        level = np.empty(nc, dtype=np.int32)
        sons = np.empty((nc, 2), dtype=np.int32)
        bb = np.empty((nc, 3, self.dim))  # [min, max, size] x dim
        lib().orc_clusters(self._h, which, _ptr(begin), _ptr(size), _ptr(level), _ptr(sons),
                           _ptr(bb))
        return dict(begin=begin, size=size, level=level, sons=sons,
                    bbmin=bb[:, 0, :].copy(), bbmax=bb[:, 1, :].copy(),
                    bbsize=bb[:, 2, :].copy())

    def blocks(self):
        nb = lib().orc_num_blocks(self._h)
        row = np.empty(nb, dtype=np.int32)
        col = np.empty(nb, dtype=np.int32)
        lr = np.empty(nb, dtype=np.int8)
        lib().orc_blocks(self._h, _ptr(row), _ptr(col), _ptr(lr))
        return row, col, lr.astype(bool)

    def leaf_basis(self, which: int, block_id: int, size: int) -> np.ndarray:
        out = np.empty((self.q, size), order="F")
        if lib().orc_leaf_basis(self._h, which, block_id, _ptr(out)) != 0:
            raise ValueError("not a leaf")
        return out

    def transfer(self, which: int, block_id: int, son: int) -> np.ndarray:
        out = np.empty((self.q, self.q), order="F")
        if lib().orc_transfer(self._h, which, block_id, son, _ptr(out)) != 0:
            raise ValueError("not an inner cluster")
        return out

    def check_transfer(self, which: int) -> float:
        return lib().orc_check_transfer(self._h, which)

    def action(self, rhs, nthreads: int = 0) -> np.ndarray:
        rhs = np.asfortranarray(np.asarray(rhs, dtype=np.float64).reshape(self.n, -1))
        out = np.empty((self.m, rhs.shape[1]), order="F")
        lib().orc_action(self._h, _ptr(rhs), rhs.shape[1], _ptr(out), nthreads)
        return out

    def action_cluster_order(self, srhs, nthreads: int = 0) -> np.ndarray:
        srhs = np.asfortranarray(np.asarray(srhs, dtype=np.float64).reshape(self.n, -1))
        out = np.empty((self.m, srhs.shape[1]), order="F")
        lib().orc_action_cluster_order(self._h, _ptr(srhs), srhs.shape[1], _ptr(out), nthreads)
        return out

    def action_leaves(self, rhs, leaves, nthreads: int = 0) -> np.ndarray:
        """Rows of the action for the given target leaves (row block ids), cluster order."""
        rhs = np.asfortranarray(np.asarray(rhs, dtype=np.float64).reshape(self.n, -1))
        leaves = np.ascontiguousarray(leaves, dtype=np.int32)
        cl = self.clusters(0)
        total = int(cl["size"][leaves].sum())
        out = np.empty(total * rhs.shape[1])
        rc = lib().orc_action_leaves(self._h, _ptr(rhs), rhs.shape[1], _ptr(leaves),
                                     leaves.size, _ptr(out), nthreads)
        if rc != 0:
            raise ValueError("a requested cluster is not a leaf")
        # unpack leaf-major blocks into (rows, nrhs)
        res = []
        off = 0
        for t in leaves:
            s = int(cl["size"][t])
            res.append(out[off:off + s * rhs.shape[1]].reshape(rhs.shape[1], s).T)
            off += s * rhs.shape[1]
        return np.concatenate(res, axis=0)

    def pcg(self, b, ridge: float, tol: float = 1e-8, max_iter: int = 1000, nthreads: int = 0):
        b = np.ascontiguousarray(b, dtype=np.float64).ravel()
        x = np.empty_like(b)
        err = np.zeros(1)
        it = lib().orc_pcg(self._h, _ptr(b), _ptr(x), ridge, tol, max_iter, _ptr(err), nthreads)
        return x, it, float(err[0])

    def entry_counts(self) -> dict:
        out = np.empty(4)
        lib().orc_entry_counts(self._h, _ptr(out))
        return dict(near_entries=out[0], far_entries=out[1], far_blocks=int(out[2]),
                    near_blocks=int(out[3]))

    def action_sample(self, rhs, entry_budget: float, nthreads: int = 0):
        rhs = np.asfortranarray(np.asarray(rhs, dtype=np.float64).reshape(self.n, -1))
        times = np.zeros(3)
        done = lib().orc_action_sample(self._h, _ptr(rhs), rhs.shape[1], entry_budget,
                                       nthreads, _ptr(times))
        return done, times
