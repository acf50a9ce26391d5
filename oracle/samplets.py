"""FMCA's samplet tree and samplet matrix compression restated in numpy
(TEST INFRASTRUCTURE: only tests/ may import this module, as the checker).

Follows, on the oracle's own cluster tree and total-degree interpolation
(OracleH2 with scheme "total_degree" at degree 2 (dtilde - 1)):
  samplet moments     MinNystromSampletMoments<MonomialInterpolator>
                      (FMCA/src/Moments/MinNystromSampletMoments.h:23-57):
                      monomials of P - box midpoint, degree < dtilde,
                      MultiIndexSet<TotalDegree> order
  moment shift        SampletHelper::monomialMomentShifter (SampletHelper.h:38-63)
  samplet bases       H2SampletTree::computeSamplets (H2SampletTree.h:101-133):
                      Householder QR of the moments' transpose in Eigen's
                      HouseholderQR convention (makeHouseholder), Q its
                      householderQ(), moments of the scaling functions R^T
  coefficient order   sampletMapper (SampletMapper.h:19-31)
  multiscale bases    H2SampletTreeBase::computeMultiscaleClusterBasis
                      (H2SampletTreeBase.h:68-86)
  transforms          SampletTreeBase::sampletTransform / inverse
                      (SampletTreeBase.h:68-85, 160-208)
  compression         SampletMatrixCompressor (samplet_matrix_compressor.h):
                      init (:37-71) with CompareCluster, compress (:73-186),
                      recursivelyComputeBlock (:330-388), triplets (:222-256),
                      aposteriori_triplets_fast (:258-290), with
                      NystromEvaluator's blocks (NystromEvaluator.h:39-67)

Pins (tests/test_samplets.py): orthogonality of the transform and vanishing
moments (the quantities FMCA's tests/test_SampletBasis.cpp prints); with eta
= 0 (no admissible block) the compressed upper triangle equals T K T^T of
the dense kernel matrix and the dense transform matrix T; with eta > 0 the
compression error against T K T^T falls with dtilde (as
tests/test_H2SampletCompressor.cpp prints it); the Householder QR agrees with
LAPACK's up to the signs of Q's columns and R's rows; the moment shift is
checked against moments computed directly about the new centre.
"""
from __future__ import annotations

import math

import numpy as np

from . import OracleH2, interp_total_degree, kernel_eval

EPS = np.finfo(np.float64).eps


def total_degree_set(dim: int, deg: int) -> list:
    """MultiIndexSet<TotalDegree>: by l1 norm, then lexicographically."""
    out = []

    def rec(prefix):
        if len(prefix) == dim:
            if sum(prefix) <= deg:
                out.append(tuple(prefix))
            return
        for a in range(deg + 1):
            rec(prefix + [a])
    rec([])
    out.sort(key=lambda a: (sum(a), a))
    return out


def monomials(idx, pts):
    """(len(idx), npts): prod_k pts[k]^a[k] (0^0 = 1)."""
    out = np.ones((len(idx), pts.shape[1]))
    for j, a in enumerate(idx):
        for k, e in enumerate(a):
            if e:
                out[j] *= pts[k] ** e
    return out


def multinomial_matrix(idx):
    m = len(idx)
    C = np.zeros((m, m))
    for i, b in enumerate(idx):
        for j, a in enumerate(idx):
            C[j, i] = np.prod([math.comb(a[k], b[k]) if b[k] <= a[k] else 0
                               for k in range(len(a))])
    return C


def shift_matrix(idx, C, shift):
    if np.linalg.norm(shift) < EPS:
        return np.eye(len(idx))
    S = C.copy()
    for i, b in enumerate(idx):
        for j, a in enumerate(idx):
            if S[j, i]:
                for k in range(len(shift)):
                    if a[k] - b[k]:
                        S[j, i] *= shift[k] ** (a[k] - b[k])
    return S


def householder_qr(A):
    """Eigen HouseholderQR on a copy of A (rows > cols): returns (Q, R) with Q
    rows x rows (householderQ) and R cols x cols (upper triangle)."""
    A = np.array(A, dtype=np.float64, copy=True)
    rows, cols = A.shape
    tau = np.zeros(cols)
    for k in range(cols):
        c0 = A[k, k]
        tail = float(np.dot(A[k + 1:, k], A[k + 1:, k]))
        if tail <= np.finfo(np.float64).tiny:
            t, beta = 0.0, c0
            A[k + 1:, k] = 0
        else:
            beta = math.sqrt(c0 * c0 + tail)
            if c0 >= 0:
                beta = -beta
            A[k + 1:, k] = A[k + 1:, k] / (c0 - beta)
            t = (beta - c0) / beta
        A[k, k] = beta
        tau[k] = t
        if t != 0 and k + 1 < cols:
            v = np.concatenate(([1.0], A[k + 1:, k]))
            w = v @ A[k:, k + 1:]
            A[k:, k + 1:] -= t * np.outer(v, w)
    Q = np.eye(rows)
    for k in range(cols - 1, -1, -1):
        if tau[k] == 0:
            continue
        v = np.concatenate(([1.0], A[k + 1:, k]))
        w = v @ Q[k:, k:]
        Q[k:, k:] -= tau[k] * np.outer(v, w)
    return Q, np.triu(A[:cols, :cols])


class OracleSampletTree:
    """H2SampletTree(mom, samp_mom, 0, P) of FMCA's Python SampletTree(P, dtilde)."""

    def __init__(self, P, dtilde: int, kernel: str = "EXPONENTIAL", l: float = 1.0,
                 c: float = 1.0, nu: float = 1.0):
        self.P = np.asfortranarray(P, dtype=np.float64)
        self.dim, self.n = self.P.shape
        self.dtilde = max(int(dtilde), 1)
        self.p = 2 * (self.dtilde - 1)
        self.H = OracleH2(self.P, self.P, kernel, l, self.p, 0.0, c, nu, scheme="total_degree")
        self.perm = self.H.indices(1)
        cl = self.H.clusters(1)
        self.begin, self.size, self.sons = cl["begin"], cl["size"], cl["sons"]
        self.level = cl["level"]
        self.bbmin, self.bbmax, self.bbsize = cl["bbmin"], cl["bbmax"], cl["bbsize"]
        self.ncl = len(self.begin)
        self.mi, self.Xi, _, self.invV = interp_total_degree(self.dim, self.p)
        self.q = self.Xi.shape[1]
        self.idx = total_degree_set(self.dim, self.dtilde - 1)
        self.m = len(self.idx)
        self.C = multinomial_matrix(self.idx)
        self.pts = self.P[:, self.perm]
        self._build()

    def is_leaf(self, c):
        return self.sons[c, 0] < 0

    def _build(self):
        ncl, m = self.ncl, self.m
        self.Q = [None] * ncl
        self.nscal = np.zeros(ncl, dtype=np.int64)
        self.nsamp = np.zeros(ncl, dtype=np.int64)
        mom = [None] * ncl
        for c in range(ncl - 1, -1, -1):  # level order: sons after fathers
            mid = 0.5 * (self.bbmin[c] + self.bbmax[c])
            if self.is_leaf(c):
                b, s = self.begin[c], self.size[c]
                M = monomials(self.idx, self.pts[:, b:b + s] - mid[:, None])
            else:
                parts = []
                for s in self.sons[c]:
                    shift = 0.5 * (self.bbmin[s] - self.bbmin[c] + self.bbmax[s] - self.bbmax[c])
                    parts.append(shift_matrix(self.idx, self.C, shift) @ mom[s])
                M = np.concatenate(parts, axis=1)
            cols = M.shape[1]
            if m < cols:
                Q, R = householder_qr(M.T)
                self.Q[c] = Q
                self.nscal[c] = m
                mom[c] = R.T
            else:
                self.Q[c] = np.eye(cols)
                self.nscal[c] = cols
                mom[c] = M
            self.nsamp[c] = cols - self.nscal[c]
        self.start = np.zeros(ncl, dtype=np.int64)
        s = 0
        for c in range(ncl):
            self.start[c] = s
            s += self.nsamp[c] + (self.nscal[c] if c == 0 else 0)
        assert s == self.n
        # multiscale cluster bases
        self.V = [None] * ncl
        for c in range(ncl - 1, -1, -1):
            if self.is_leaf(c):
                self.V[c] = self.H.leaf_basis(1, c, int(self.size[c])) @ self.Q[c]
            else:
                parts = [self.H.transfer(1, c, k) @ self.V[s][:, :self.nscal[s]]
                         for k, s in enumerate(self.sons[c])]
                self.V[c] = np.concatenate(parts, axis=1) @ self.Q[c]

    # --- transforms (data in cluster order) ---------------------------------
    def transform(self, data):
        data = np.asarray(data, dtype=np.float64).reshape(self.n, -1)
        out = np.zeros_like(data)
        scal = [None] * self.ncl
        for c in range(self.ncl - 1, -1, -1):
            if self.is_leaf(c):
                X = data[self.begin[c]:self.begin[c] + self.size[c]]
            else:
                X = np.concatenate([scal[s] for s in self.sons[c]], axis=0)
            Y = self.Q[c].T @ X
            ns = self.nscal[c]
            sh = ns if c == 0 else 0
            out[self.start[c] + sh:self.start[c] + sh + self.nsamp[c]] = Y[ns:]
            scal[c] = Y[:ns]
            if c == 0:
                out[:ns] = Y[:ns]
        return out

    def inverse_transform(self, coef):
        coef = np.asarray(coef, dtype=np.float64).reshape(self.n, -1)
        out = np.zeros_like(coef)
        scal = [None] * self.ncl
        for c in range(self.ncl):
            ns = self.nscal[c]
            sh = ns if c == 0 else 0
            top = coef[:ns] if c == 0 else scal[c]
            Y = np.concatenate([top, coef[self.start[c] + sh:self.start[c] + sh + self.nsamp[c]]])
            X = self.Q[c] @ Y
            if self.is_leaf(c):
                out[self.begin[c]:self.begin[c] + self.size[c]] = X
            else:
                off = 0
                for s in self.sons[c]:
                    scal[s] = X[off:off + self.nscal[s]]
                    off += self.nscal[s]
        return out

    def transform_matrix(self):
        """T with coefficients = T @ data (cluster order), n x n."""
        return self.transform(np.eye(self.n))


class OracleSampletCompressor:
    """SampletMatrixCompressor<H2SampletTree> (CompareCluster) with
    NystromEvaluator<NystromMoments<TotalDegree>, CovarianceKernel>."""

    def __init__(self, T: OracleSampletTree, kernel: str, l: float, eta: float,
                 threshold: float = 0.0, c: float = 1.0, nu: float = 1.0):
        self.T, self.kernel, self.l, self.c, self.nu = T, kernel, l, c, nu
        self.eta = eta
        self.threshold = threshold
        self.radius = 0.5 * np.linalg.norm(T.bbsize, axis=1)
        self._pattern()
        self._compress()

    def admissible(self, a, b):
        T = self.T
        v = 0.5 * (T.bbmin[a] - T.bbmin[b] + T.bbmax[a] - T.bbmax[b])
        dist = math.sqrt(float(np.dot(v, v))) - self.radius[a] - self.radius[b]
        dist = dist if dist > 0 else 0.0
        rad = max(self.radius[a], self.radius[b])
        return not (rad > self.eta * dist)

    def _pattern(self):
        T = self.T
        self.pattern = {}
        for pc in range(T.ncl):
            stack = [0]
            while stack:
                pr = stack.pop()
                if not T.is_leaf(pr):
                    for s in T.sons[pr]:
                        if not self.admissible(s, pc):
                            stack.append(s)
                if pc >= pr:
                    self.pattern[(pr, pc)] = None

    def _kmat(self, X, Y):
        d = np.sqrt(((X[:, :, None] - Y[:, None, :]) ** 2).sum(0))
        return kernel_eval(self.kernel, d.ravel(), self.l, self.c, self.nu).reshape(d.shape)

    def _nodes(self, c):
        T = self.T
        return T.Xi * T.bbsize[c][:, None] + T.bbmin[c][:, None]

    def _rec(self, R, C):
        T = self.T
        if self.admissible(R, C):
            S = T.invV @ self._kmat(self._nodes(R), self._nodes(C)) @ T.invV.T
            return T.V[R].T @ S @ T.V[C]
        lr, lc = T.is_leaf(R), T.is_leaf(C)
        if lr and lc:
            bR, sR, bC, sC = T.begin[R], T.size[R], T.begin[C], T.size[C]
            K = self._kmat(T.pts[:, bR:bR + sR], T.pts[:, bC:bC + sC])
            return T.Q[R].T @ K @ T.Q[C]
        if lr:
            buf = np.concatenate([self._rec(R, s)[:, :T.nscal[s]] for s in T.sons[C]], axis=1)
            return buf @ T.Q[C]
        if lc:
            buf = np.concatenate([self._rec(s, C)[:T.nscal[s], :] for s in T.sons[R]], axis=0)
            return T.Q[R].T @ buf
        rows = []
        for ri in T.sons[R]:
            r1 = np.concatenate([self._rec(ri, cj)[:, :T.nscal[cj]] for cj in T.sons[C]],
                                axis=1) @ T.Q[C]
            rows.append(r1[:T.nscal[ri], :])
        return T.Q[R].T @ np.concatenate(rows, axis=0)

    def _compress(self):
        T = self.T
        levels = {}
        for (pr, pc) in self.pattern:
            levels.setdefault(int(T.level[pr] + T.level[pc]), []).append((pr, pc))
        for ll in sorted(levels, reverse=True):
            for (pr, pc) in levels[ll]:
                lr, lc = T.is_leaf(pr), T.is_leaf(pc)
                if lr and lc:
                    B = self._rec(pr, pc)
                elif lc:
                    parts = []
                    for s in T.sons[pr]:
                        X = self.pattern.get((s, pc))
                        X = X if X is not None else self._rec(s, pc)
                        parts.append(X[:T.nscal[s], :])
                    B = T.Q[pr].T @ np.concatenate(parts, axis=0)
                else:
                    parts = []
                    for s in T.sons[pc]:
                        X = self.pattern.get((pr, s))
                        X = X if X is not None else self._rec(pr, s)
                        parts.append(X[:, :T.nscal[s]])
                    B = np.concatenate(parts, axis=1) @ T.Q[pc]
                self.pattern[(pr, pc)] = B

    def triplets(self):
        T = self.T
        rows, cols, vals = [], [], []
        for (r, c), B in self.pattern.items():
            if r != 0 and c != 0:
                W = B[-T.nsamp[r]:, -T.nsamp[c]:] if T.nsamp[r] and T.nsamp[c] else B[:0, :0]
            elif c != 0:
                W = B[:, B.shape[1] - T.nsamp[c]:]
            else:
                W = B
            sr, sc = T.start[r], T.start[c]
            for k in range(W.shape[1]):
                for j in range(W.shape[0]):
                    v = W[j, k]
                    if (sr + j <= sc + k and abs(v) > self.threshold) or (sr == sc and j == k):
                        rows.append(sr + j)
                        cols.append(sc + k)
                        vals.append(v)
        return np.array(rows, dtype=np.int64), np.array(cols, dtype=np.int64), np.array(vals)


def aposteriori_fast(vals, rows, cols, thres):
    """Indices kept by aposteriori_triplets_fast (in FMCA's emission order)."""
    entry = np.abs(vals)
    with np.errstate(divide="ignore"):
        val = -np.floor(np.log(entry) / math.log(10.0))
    b = np.where(val < 0, 0, np.where(val >= 16, 16, np.nan_to_num(val, posinf=16))).astype(int)
    norms2 = np.array([np.sum(entry[b == i] ** 2) for i in range(17)])
    fnorm2 = norms2.sum()
    cut, cut_off = 0.0, 17
    for i in range(16, -1, -1):
        cut += norms2[i]
        if math.sqrt(cut / fnorm2) >= thres:
            break
        cut_off -= 1
    keep = []
    for i in range(cut_off):
        keep.extend(np.nonzero(b == i)[0].tolist())
    for i in range(cut_off, 17):
        keep.extend([t for t in np.nonzero(b == i)[0].tolist() if rows[t] == cols[t]])
    return np.array(keep, dtype=np.int64)
