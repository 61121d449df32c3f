"""Block Lanczos for M^{1/2} Z and Brownian displacements over a batch of
right-hand sides: the caller of the batched apply (one apply of s <= 16
vectors per Krylov step, all sharing one P build).  No counterpart in the
reference.

For an SPD operator M (the full RPY mobility: real space, overlap, self and
wave space) and a block Z of s vectors (rows of a [s][n] array):

    Z = Q_1 R_0                                 (Cholesky QR, twice)
    W = M Q_j - Q_{j-1} B_{j-1}^T,  A_j = Q_j^T W,  W -= Q_j A_j
    W -= Q (Q^T W),  Q = [Q_1 .. Q_j]            (full reorthogonalisation)
    W = Q_{j+1} B_j                             (Cholesky QR, twice)
    T_m = block tridiagonal (A_j on the diagonal, B_j below, B_j^T above)
    M^{1/2} Z ~ [Q_1 .. Q_m] (T_m^{1/2})[:, :s] R_0

T_m^{1/2} comes from the eigendecomposition of the small (s m) x (s m)
matrix on the host; the iteration stops when the coefficient block changes
by less than tol (relative Frobenius norm) from one step to the next.  The
blocks Q_i are consecutive rows of one basis array, so the
reorthogonalisation is one Gram product and one update over all earlier
rows (classical Gram-Schmidt as a second pass after the recurrence's own
projections), and the result one update.  The tall-skinny products run as
the library's HIP kernels (spme_block_gram, spme_block_update: up to 4096
basis rows per call), the operator as the engine's batched apply.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from .lib import SPME_FP32, SPME_FULL, SpmeError, check, load


TALL = 4096   # most rows of X in one spme_block_gram / spme_block_update


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


class HipBlockOps:
    """Block Gram products and updates by the library's HIP kernels, on the
    current torch stream.  X, Y: [s][ld] device tensors (rows contiguous)."""

    def __init__(self, precision: int, n: int, device="cuda"):
        import torch
        self.torch = torch
        self.precision = int(precision)
        self.n = int(n)
        self.device = device
        lib = load()
        ws = lib.spme_block_gram_worksize(self.n, 16, 16)
        self._work = torch.empty(int(ws), dtype=torch.float64, device=device)
        self._C = torch.empty(256, dtype=torch.float64, device=device)

    def _stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def _workspace(self, sx: int, sy: int):
        ws = int(load().spme_block_gram_worksize(self.n, sx, sy))
        if ws > self._work.numel():
            self._work = self.torch.empty(ws, dtype=self.torch.float64, device=self.device)
        return self._work

    def gram_dev(self, X, Y, out=None):
        """X Y^T as an FP64 device tensor [sx][sy], without waiting for it
        (sx <= 4096 rows of X, sy <= 16 of Y)."""
        sx, sy = X.shape[0], Y.shape[0]
        if out is None or out.numel() < sx * sy:
            out = self.torch.empty(sx * sy, dtype=self.torch.float64, device=self.device)
        C = out
        self._workspace(sx, sy)
        check(load().spme_block_gram(self.precision, self.n, sx, sy, _ptr(X), X.stride(0),
                                     _ptr(Y), Y.stride(0), _ptr(C), _ptr(self._work),
                                     self._work.numel(), self._stream()), "spme_block_gram")
        return C[:sx * sy].view(sx, sy)

    def gram(self, X, Y) -> np.ndarray:
        return self.gram_dev(X, Y, self._C).cpu().numpy().copy()

    def update(self, X, A, beta: float, Y) -> None:
        """Y = beta Y + A^T X in the rows' sense; A: numpy or FP64 device tensor."""
        sx, sy = X.shape[0], Y.shape[0]
        if self.torch.is_tensor(A):
            dA = A.reshape(sx, sy).to(self.torch.float64).contiguous()
        else:
            A = np.ascontiguousarray(A, dtype=np.float64).reshape(sx, sy)
            dA = self.torch.tensor(A, device=self.device)
        check(load().spme_block_update(self.precision, self.n, sx, sy, _ptr(X), X.stride(0),
                                       _ptr(dA), float(beta), _ptr(Y), Y.stride(0),
                                       self._stream()), "spme_block_update")

    def empty(self, s: int):
        dt = self.torch.float32 if self.precision == SPME_FP32 else self.torch.float64
        return self.torch.empty(s, self.n, dtype=dt, device=self.device)

    @property
    def eps(self) -> float:
        # singular values come from the Gram matrix, which squares them:
        # rank loss shows down to about the square root of the rounding
        return 1e-5 if self.precision == SPME_FP32 else 1e-7


@dataclass
class LanczosInfo:
    iterations: int
    change: float        # last relative change of the coefficient block
    converged: bool
    min_eig: float       # smallest eigenvalue of T_m (a Ritz value of M)
    max_eig: float
    dimension: int       # dimension of the Krylov space built
    invariant: bool      # the space became invariant (the result is exact)


def _eigh(T: np.ndarray):
    """Eigendecomposition of the symmetric T; on the GPU (FP64) once T is
    large enough for the host's O(n^3) to matter."""
    if T.shape[0] >= 512:
        try:
            import torch
            if torch.cuda.is_available():
                w, Q = torch.linalg.eigh(torch.tensor(T, device="cuda"))
                return w.cpu().numpy(), Q.cpu().numpy()
        except ImportError:
            pass
    return np.linalg.eigh(T)


def _sqrt_first_cols(T: np.ndarray, k0: int):
    w, Q = _eigh(0.5 * (T + T.T))
    if w[0] < -1e-8 * max(abs(w[-1]), 1e-300):
        raise SpmeError(f"operator is not positive definite (Ritz value {w[0]:.3e})")
    r = np.sqrt(np.clip(w, 0.0, None))
    return (Q * r) @ Q[:k0, :].T, float(w[0]), float(w[-1])


class BlockLanczos:
    """M^{1/2} Z for blocks Z of s <= 16 vectors of length n.

    apply(X, Y) must set Y = M X for [k][n] blocks, k <= s; ops provides
    gram, update, empty(k) and eps (the relative singular value below which
    a direction of a block counts as lost; the Gram matrix squares singular
    values, so about 1e-7 for FP64 vectors and 1e-5 for FP32).
    Blocks deflate: directions whose singular value falls below eps times
    the operator's scale are dropped, so the block width can shrink, and an
    empty block ends the iteration with the exact result on an invariant
    space."""

    def __init__(self, apply, ops, s: int, tol: float = 1e-6, max_iter: int = 60,
                 reorth: bool = True):
        if not 1 <= s <= 16:
            raise SpmeError("block width must be in [1, 16]")
        self.apply, self.ops, self.s = apply, ops, int(s)
        self.tol, self.max_iter, self.reorth = float(tol), int(max_iter), bool(reorth)

    def _orth(self, W, scale: float, dst=None):
        """V ([k][n], orthonormal rows) and R (k x w) with W = V^T R in the
        rows' sense (w_b = sum_a v_a R[a][b]); k = numerical rank of W.  V is
        written to the first k rows of dst when given."""
        ops = self.ops
        G = ops.gram(W, W)
        sig2, U = np.linalg.eigh(0.5 * (G + G.T))
        sig = np.sqrt(np.clip(sig2, 0.0, None))
        keep = sig > ops.eps * max(scale, sig.max(initial=0.0), 1e-300)
        k = int(keep.sum())
        if k == 0:
            return None, None
        Uk, sk = U[:, keep], sig[keep]
        V1 = ops.empty(k)
        ops.update(W, Uk / sk, 0.0, V1)
        R = sk[:, None] * Uk.T
        # second pass (Cholesky QR) restores orthogonality to rounding
        G2 = ops.gram(V1, V1)
        try:
            R2 = np.linalg.cholesky(0.5 * (G2 + G2.T)).T
        except np.linalg.LinAlgError as exc:
            raise SpmeError("block Lanczos: orthogonalisation failed") from exc
        V = dst[:k] if dst is not None else ops.empty(k)
        ops.update(V1, np.linalg.inv(R2), 0.0, V)
        return V, R2 @ R

    @staticmethod
    def _coefficients(As, Bs, R0):
        """(T^{1/2})[:, :k0] R0 for the block tridiagonal T of As, Bs."""
        sizes = [a.shape[0] for a in As]
        off = np.concatenate([[0], np.cumsum(sizes)])
        T = np.zeros((off[-1], off[-1]))
        for i, a in enumerate(As):
            T[off[i]:off[i + 1], off[i]:off[i + 1]] = a
            if i + 1 < len(As):
                T[off[i + 1]:off[i + 2], off[i]:off[i + 1]] = Bs[i]
                T[off[i]:off[i + 1], off[i + 1]:off[i + 2]] = Bs[i].T
        S, lo, hi = _sqrt_first_cols(T, R0.shape[0])
        return S @ R0, lo, hi, off

    def sqrt(self, Z, out=None):
        """Return (U, info) with U ~ M^{1/2} Z (rows of Z are the vectors)."""
        s, ops = self.s, self.ops
        if Z.shape[0] != s:
            raise SpmeError(f"Z must have {s} rows")
        # Q_1, Q_2, ... as consecutive rows of one array (at most s rows per
        # step and max_iter steps)
        basis = ops.empty(s * self.max_iter)
        Q, R0 = self._orth(Z, 0.0, basis)
        U = out if out is not None else ops.empty(s)
        if Q is None:  # Z = 0
            ops.update(Z, np.zeros((s, s)), 0.0, U)
            return U, LanczosInfo(0, 0.0, True, 0.0, 0.0, 0, True)
        Qs, As, Bs = [Q], [], []
        rows = Q.shape[0]     # basis rows in use
        scale, coef_prev, change = 0.0, None, np.inf
        invariant = False
        next_check = 1
        coef, lo, hi, checked = None, 0.0, 0.0, 0
        for j in range(self.max_iter):
            kj = Qs[j].shape[0]
            W = ops.empty(kj)
            self.apply(Qs[j], W)
            if j > 0:
                ops.update(Qs[j - 1], -Bs[j - 1].T, 1.0, W)
            A = ops.gram(Qs[j], W)
            A = 0.5 * (A + A.T)
            ops.update(Qs[j], -A, 1.0, W)
            As.append(A)
            scale = max(scale, float(np.abs(np.linalg.eigvalsh(A)).max()))
            if self.reorth:
                # W -= Q (Q^T W) over all basis rows, TALL rows per call; with
                # device-side Gram products nothing here waits on the host
                dev = getattr(ops, "gram_dev", None)
                for a in range(0, rows, TALL):
                    Qa = basis[a:min(rows, a + TALL)]
                    ops.update(Qa, -(dev(Qa, W) if dev else ops.gram(Qa, W)), 1.0, W)
            m = j + 1
            # the (s m)^3 eigendecomposition is checked every step while it is
            # cheap, then every m/4 steps (the change is then measured over
            # several steps, a stricter test)
            if m >= next_check or m == self.max_iter:
                coef, lo, hi, off = self._coefficients(As, Bs, R0)
                checked = m
                if coef_prev is not None:
                    pad = np.zeros_like(coef)
                    pad[:coef_prev.shape[0]] = coef_prev
                    change = float(np.linalg.norm(coef - pad) / np.linalg.norm(coef))
                coef_prev = coef
                next_check = m + (1 if m < 8 else max(1, m // 4))
                # a check due within m/8 steps of the cap moves to the cap,
                # which is checked anyway
                if next_check < self.max_iter and self.max_iter - next_check < m // 8:
                    next_check = self.max_iter
                if change < self.tol or m == self.max_iter:
                    break
            Qn, B = self._orth(W, scale, basis[rows:rows + kj])
            if Qn is None:       # invariant space: the result is exact
                invariant = True
                break
            Qs.append(Qn)
            Bs.append(B)
            rows += Qn.shape[0]
        if checked != len(As):
            coef, lo, hi, off = self._coefficients(As, Bs, R0)
        m = len(As)
        info = LanczosInfo(m, change, invariant or change < self.tol, lo, hi, int(off[-1]),
                           invariant)
        used = int(off[-1])   # rows of Q_1 .. Q_m
        for a in range(0, used, TALL):
            b = min(used, a + TALL)
            ops.update(basis[a:b], coef[a:b], 0.0 if a == 0 else 1.0, U)
        return U, info


def engine_lanczos(engine, s: int, tol: float = 1e-6, max_iter: int = 100,
                   parts: int = SPME_FULL) -> BlockLanczos:
    """Block Lanczos over an engine's batched apply (full operator)."""
    if s > engine.max_rhs:
        raise SpmeError(f"block width {s} exceeds the engine's max_rhs {engine.max_rhs}")
    if engine.dtype != engine.out_dtype:
        raise SpmeError("block Lanczos needs an engine whose inputs and outputs share one "
                        "precision (SPME_FP32 or SPME_FP64)")
    n = 3 * engine.info().N
    ops = HipBlockOps(engine.precision, n)

    def apply(X, Y):
        engine.apply(X, Y, parts=parts)

    return BlockLanczos(apply, ops, s, tol=tol, max_iter=max_iter)


def brownian_displacements(engine, F, Z, dt: float, kT: float, tol: float = 1e-4,
                           max_iter: int = 100):
    """dX = dt M F + sqrt(2 kT dt) M^{1/2} Z for s independent realisations
    (rows of F and Z, [s][3N] device tensors of the engine's precision;
    mobility in units of 1/(6 pi eta a)).  Returns (dX, info)."""
    s = F.shape[0]
    bl = engine_lanczos(engine, s, tol=tol, max_iter=max_iter)
    U, info = bl.sqrt(Z)
    dX = bl.ops.empty(s)
    engine.apply(F, dX, alpha=dt, parts=SPME_FULL)
    bl.ops.update(U, np.sqrt(2.0 * kT * dt) * np.eye(s), 1.0, dX)
    return dX, info


__all__ = ["BlockLanczos", "HipBlockOps", "LanczosInfo", "engine_lanczos",
           "brownian_displacements"]
