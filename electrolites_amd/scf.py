"""Device-resident DF-RHF driver on the MI355X DF path.

The SCF loop PySCF/GPU4PySCF runs around ``df_jk.get_jk`` (the reference's
only caller of the hot path, gpu4pyscf ``scf.hf.kernel`` -> ``_DFHF.get_veff``
-> ``DF.get_jk``), restated with every matrix kept in HBM:

* one-electron Hamiltonian ``H = T + V`` and overlap from ``mdf_int1e``;
* the cderi from the GPU build (:mod:`electrolites_amd.fastdfcderi`);
* per cycle: ``J, K`` from :func:`electrolites_amd.fastdf.get_jk` on a
  ``mo_coeff``-tagged density (one-factor exchange), ``F = H + J - K/2``,
  ``E = tr(D (H + F)) / 2 + E_nuc``, DIIS on the orthogonalised ``FDS - SDF``,
  and a FP64 diagonalisation of ``X^T F X`` (``X = S^-1/2``, canonical).

The initial guess is the core Hamiltonian (PySCF's MINAO guess needs a basis
set that is not available offline; the converged energy does not depend on
the guess).

Precision schedule (``k_precision``): 'fp64' runs every cycle in FP64;
'mixed' builds K (step 1, the occupied transform, and step 2) in the
``low_precision`` mode -- 'fp32' (FP32 MFMA) or 'bf16x3' (split-BF16
three-term MFMA), FP32 accumulation either way -- while the DIIS error is
above ``switch_err`` (1e-5), then restarts DIIS and finishes in FP64; 'fp32'
and 'bf16x3' stay in that mode throughout.  J, the metric, the
diagonalisation and the energy sums are FP64 throughout.  The per-cycle
precision is recorded in ``history``.
"""
import math
import threading

import numpy as np
from scipy.linalg import lapack as scipy_lapack
from threadpoolctl import threadpool_limits

from . import _dev
from . import _lib
from . import eigsolve
from . import fastdf
from . import ints
from .df import DF


_HOST_LAPACK = threading.Lock()
_single_blas = 0          # number of open single_thread_blas contexts
_single_lim = None


class single_thread_blas:
    """Context: BLAS/LAPACK at one thread for the whole process (the first
    context to open sets it, the last to close restores it), so that
    :func:`host_eigh` calls with ``concurrent=True`` from many threads run
    side by side while one is open."""

    def __enter__(self):
        global _single_blas, _single_lim
        with _HOST_LAPACK:
            if _single_blas == 0:
                _single_lim = threadpool_limits(1, user_api='blas')
            _single_blas += 1
        return self

    def __exit__(self, *exc):
        global _single_blas, _single_lim
        with _HOST_LAPACK:
            _single_blas -= 1
            if _single_blas == 0:
                _single_lim.restore_original_limits()
                _single_lim = None


def host_eigh(a, concurrent=False):
    """LAPACK ``eigh`` on one host thread.  The thread count of a threaded
    BLAS is process-wide and changes the last bits of the result, so every
    host diagonalisation here is single-threaded: with ``concurrent`` the
    caller guarantees an open :class:`single_thread_blas` context and calls
    run side by side; otherwise they are serialised under one lock.  The same
    matrix gives the same eigenpairs either way."""
    if concurrent:
        if not _single_blas:
            raise RuntimeError('host_eigh(concurrent=True) outside single_thread_blas')
        return np.linalg.eigh(a)
    with _HOST_LAPACK, threadpool_limits(1, user_api='blas'):
        return np.linalg.eigh(a)


def projector_factor(a, rank, concurrent=False):
    """W (n x rank) with a = W W^T for a symmetric projector a of that rank,
    by LAPACK's pivoted Cholesky (``dpstrf``, stopped at pivots below 1e-8,
    on one host thread as :func:`host_eigh`), or None when a's numerical
    rank differs.  For a projector W^T W = I."""
    if concurrent:
        if not _single_blas:
            raise RuntimeError('projector_factor(concurrent=True) outside single_thread_blas')
        c, piv, r, info = scipy_lapack.dpstrf(a, tol=1e-8, lower=1)
    else:
        with _HOST_LAPACK, threadpool_limits(1, user_api='blas'):
            c, piv, r, info = scipy_lapack.dpstrf(a, tol=1e-8, lower=1)
    if info < 0 or r != rank:
        return None
    W = np.zeros((a.shape[0], rank))
    W[piv - 1] = np.tril(c)[:, :rank]
    return W


def _cholesky_inverse(M):
    """``(L^-1, ok)`` for the SPD device matrix M = L L^T (``mdf_potrf`` and
    ``mdf_trtri``, the build's own kernels); ok False for a pivot below
    1e-10 of the largest diagonal (M near singular)."""
    from . import _lib
    t = _dev.torch()
    n = M.shape[0]
    low = M.clone()
    info = t.zeros(1, dtype=t.int32, device=M.device)
    work = t.empty(_lib.query('mdf_potrf_workspace', n), dtype=t.uint8, device=M.device)
    _lib.call('mdf_potrf', _dev.ptr(low), n, n, _dev.ptr(info), _dev.ptr(work), work.numel(),
              _dev.stream_ptr())
    dmin, dmax, bad = t.stack([low.diagonal().min(), low.diagonal().max(), info[0].double()]).tolist()
    if bad != 0 or not dmin > 1e-5 * dmax:
        return None, False
    linv = t.empty((n, n), dtype=t.float64, device=M.device)
    w2 = t.empty(_lib.query('mdf_trtri_workspace', n), dtype=t.uint8, device=M.device)
    _lib.call('mdf_trtri', _dev.ptr(low), n, n, _dev.ptr(work), _dev.ptr(linv), n, _dev.ptr(w2), w2.numel(),
              _dev.stream_ptr())
    return linv, True


class RHF:
    def __init__(self, mol, conv_tol=1e-10, conv_tol_grad=None, max_cycle=100, diis_space=8,
                 df=None, verbose=0, k_precision='fp64', switch_err=1e-5, low_precision='fp32',
                 host_diag_max=0, precision_report=False, purify_min=None, conv_check=False):
        if k_precision not in ('fp64', 'fp32', 'bf16x3', 'mixed'):
            raise ValueError(f'k_precision {k_precision!r}')
        if low_precision not in ('fp32', 'bf16x3'):
            raise ValueError(f'low_precision {low_precision!r}')
        # after convergence on an FP64 J/K, one more FP64 J/K of the new
        # density gives the energy (as PySCF's conv_check re-runs a cycle);
        # off, the converged cycle's FP64 energy is the result.  A run that
        # converged in FP32 or split-BF16 always takes the extra FP64 J/K.
        self.conv_check = conv_check
        self.k_precision = k_precision
        self.low_precision = low_precision
        # diagnostic: every low-precision cycle also builds the FP64 J/K of
        # the same density and records the relative errors in ``history``
        self.precision_report = precision_report
        self.host_diag_max = host_diag_max
        # orthogonal dimensions above this take the occupied space of each
        # cycle's Fock matrix by purification instead of a diagonalisation
        self.purify_min = purify_min
        self.concurrent_diag = False      # set by a caller that holds single_thread_blas
        self.switch_err = switch_err
        self.mol = mol
        self.conv_tol = conv_tol
        self.conv_tol_grad = conv_tol_grad if conv_tol_grad is not None else math.sqrt(conv_tol)
        self.max_cycle = max_cycle
        self.diis_space = diis_space
        self.with_df = df or DF(mol)
        self.verbose = verbose
        self.e_tot = None
        self.converged = False
        self.cycles = 0
        self.history = []

    # ------------------------------------------------------------ setup
    def build(self):
        t = _dev.torch()
        _dev.require_gpu()
        mol = self.mol
        if self.with_df._cderi is None:
            self.with_df.build()
        ao = getattr(mol, '_mdf_dev_ao', None) or ints.DeviceBasis(mol.ao)
        mol._mdf_dev_ao = ao
        bound = getattr(mol, '_mdf_pair_bound', None)
        if bound is None:
            bound = ints.pair_bound(mol.ao)
            mol._mdf_pair_bound = bound
        self.S, T, V = ints.int1e(ao, mol, bound)
        self.H = T + V
        w, v = eigsolve.eigh(self.S)
        keep = w > 1e-9
        self.X = (v[:, keep] / t.sqrt(w[keep])).contiguous()
        self.enuc = mol.energy_nuc()
        self.nocc = mol.nocc
        return self

    def _density(self, C):
        """The closed-shell density of the first ``nocc`` columns of C,
        tagged with them (one-factor exchange)."""
        t = _dev.torch()
        nocc = self.nocc
        co = C[:, :nocc].contiguous()
        dm = 2. * co @ co.T
        occ = t.full((nocc,), 2., dtype=t.float64, device=C.device)
        return fastdf.tag_array(dm, mo_coeff=co, mo_occ=occ)

    def _purify_occ(self, A):
        """An orthonormal basis (n x nocc) of the space of the nocc lowest
        eigenvalues of the symmetric ``A`` without diagonalising it, or None.

        SP2 purification (Niklasson, Phys. Rev. B 66, 155115 (2002)): the
        Gershgorin bounds map A's spectrum into [0, 1] reversed, X0 = (e_max
        - A) / (e_max - e_min); each step takes X^2 or 2X - X^2, whichever
        trace lies closer to nocc, and drives the eigenvalues to 1 (the nocc
        lowest of A) or 0, until X is the spectral projector P (FP64
        products only; tr(X^2) = ||X||_F^2 for the symmetric X, so each step
        is one kernel).  The basis is
        P R (R^T P R)^-1/2 by a Cholesky factor (``mdf_potrf`` /
        ``mdf_trtri``), R the previous cycle's occupied orbitals, or, when P R
        loses rank, from P alone by a pivoted Cholesky (``_projector_basis``):
        no device solver library is called.  None (the caller diagonalises)
        when the steps do not converge (a vanishing gap at nocc) or P does
        not have rank nocc."""
        t = _dev.torch()
        n, nocc = A.shape[0], self.nocc
        R = getattr(self, '_q_occ', None)
        if R is None or R.shape != (n, nocc) or nocc == 0:
            return None
        d = A.diagonal()
        off = A.abs().sum(dim=1) - d.abs()
        emin, emax = t.stack([(d - off).min(), (d + off).max()]).tolist()
        c = self._evecs
        if c is not None:
            # Gershgorin again in the eigenbasis c of the last diagonalised
            # Fock matrix, where A is nearly diagonal once the solve settles:
            # c^T A c has A's eigenvalues (c orthogonal), and its discs are
            # bounds as rigorous as A's own but often ten times narrower
            # (every halving of the width saves about two steps)
            Ac = c.T @ (A @ c)
            dc = Ac.diagonal()
            oc = Ac.abs().sum(dim=1) - dc.abs()
            lo, hi = t.stack([(dc - oc).min(), (dc + oc).max()]).tolist()
            pad = 1e-12 * max(abs(lo), abs(hi))
            emin, emax = max(emin, lo - pad), min(emax, hi + pad)
        if not emax > emin:
            return None
        # two iterates and their per-tile (tr, ||.||_F^2) sums, ping-pong
        Xs = t.empty((2, n, n), dtype=t.float64, device=A.device)
        part = t.empty((2, _lib.query('mdf_sp2_tiles', n), 2), dtype=t.float64, device=A.device)
        t.mul(A, -1. / (emax - emin), out=Xs[0])
        Xs[0].diagonal().add_(emax / (emax - emin))
        st = _dev.stream_ptr()
        _lib.call('mdf_sp2_begin', n, _dev.ptr(Xs[0]), n, _dev.ptr(part[0]), st)
        # one kernel a step (``mdf_sp2_step``: the step choice from the
        # previous sums, the lower tiles of the product, the update and the
        # next sums); convergence is read back once per PURIFY_BLOCK steps,
        # and the steps run until a block ends with tr(X - X^2) below 1e-12
        # (further steps leave a converged projector as it is)
        done, cur = False, 0
        for _ in range(0, self.PURIFY_MAXIT, self.PURIFY_BLOCK):
            for _ in range(self.PURIFY_BLOCK):
                _lib.call('mdf_sp2_step', n, float(nocc), _dev.ptr(Xs[cur]), n, _dev.ptr(part[cur]),
                          _dev.ptr(Xs[1 - cur]), n, _dev.ptr(part[1 - cur]), st)
                cur = 1 - cur
            trX, trX2 = part[cur].sum(dim=0).tolist()
            if abs(trX - trX2) < 1e-12:
                done = True
                break
        if not done:
            return None
        if abs(trX - nocc) > 1e-6:
            return None
        X = Xs[cur]
        PR = X @ R
        M = (R.T @ PR).contiguous()
        L, ok = _cholesky_inverse(M)
        if ok:
            Q = PR @ L.T
            # a nearly singular R^T P R passes the pivot test but leaves Q
            # far from orthonormal: then the basis comes from P alone
            if float((Q.T @ Q - t.eye(nocc, dtype=Q.dtype, device=Q.device)).abs().max()) < 1e-10:
                return Q
        # R misses part of the occupied space (early in a solve it rotates by
        # more than the previous orbitals cover): the basis from P alone
        return self._projector_basis(X)

    def _projector_basis(self, P):
        """An orthonormal basis (n x nocc) of the range of the projector P,
        or None: LAPACK's pivoted Cholesky (``dpstrf``, one host thread)
        stops at rank nocc with P = W W^T, and for a projector W^T W = I;
        one Cholesky-QR pass on the device (``mdf_potrf`` / ``mdf_trtri``)
        removes the rounding left in W^T W."""
        t = _dev.torch()
        nocc = self.nocc
        W = projector_factor(P.cpu().numpy(), nocc, self.concurrent_diag)
        if W is None:
            return None
        W = t.from_numpy(W).to(P.device)
        L, ok = _cholesky_inverse((W.T @ W).contiguous())
        if not ok:
            return None
        return W @ L.T

    #: most SP2 steps before falling back to a diagonalisation
    PURIFY_MAXIT = 96
    #: SP2 steps between reads of the convergence measure
    PURIFY_BLOCK = 4

    # ------------------------------------------------------------- DIIS
    def _diis_reset(self):
        self._diis_n = 0

    def _diis(self, F, err):
        """Pulay DIIS over the last ``diis_space`` (Fock, error) pairs, kept
        in device ring buffers; the error overlaps are updated one row per
        cycle as plain products and sums (a GEMM of shape 8 x nao^2 x 8 or a
        stack of the history would cost milliseconds per cycle)."""
        t = _dev.torch()
        m = self.diis_space
        if getattr(self, '_fh', None) is None or self._fh.shape[1:] != F.shape:
            self._fh = t.empty((m,) + tuple(F.shape), dtype=F.dtype, device=F.device)
            self._eh = t.empty_like(self._fh)
            self._bm = t.zeros((m, m), dtype=F.dtype, device=F.device)
        k = self._diis_n % m
        self._diis_n += 1
        n = min(self._diis_n, m)
        self._fh[k].copy_(F)
        self._eh[k].copy_(err)
        row = (self._eh[:n] * err).sum(dim=(1, 2))
        self._bm[k, :n] = row
        self._bm[:n, k] = row
        if n < 2:
            return F
        # the (n+1) x (n+1) Pulay system on the host (LAPACK; a system this
        # small runs on one thread): no solver-library call from the serving
        # loop's worker threads
        B = -np.ones((n + 1, n + 1))
        B[n, n] = 0.
        B[:n, :n] = self._bm[:n, :n].cpu().numpy()
        rhs = np.zeros(n + 1)
        rhs[n] = -1.
        try:
            coef = np.linalg.solve(B, rhs)[:n]
        except np.linalg.LinAlgError:
            # linearly dependent history (error vectors that repeat near
            # convergence): the least-squares solution
            coef = np.linalg.lstsq(B, rhs, rcond=None)[0][:n]
        coef = t.from_numpy(coef).to(F.device)
        return (coef.view(-1, 1, 1) * self._fh[:n]).sum(dim=0)

    def _occupied(self, A):
        """The next cycle's occupied AO coefficients (nao x nocc) from the
        DIIS-extrapolated Fock matrix in the orthogonal basis: purified
        (``purify_min``) or from a diagonalisation, which also gives the
        next purifications their spectral bounds; the occupied orbitals in
        the orthogonal basis are kept (the next purification and DIIS error
        start from them)."""
        t = _dev.torch()
        purify = self.purify_min is not None and self.X.shape[1] > self.purify_min
        if purify:
            A = .5 * (A + A.T)
            q = self._purify_occ(A)
            if q is not None:
                self._q_occ = q
                self.purified = getattr(self, 'purified', 0) + 1
                self.mo_energy = None
                self.mo_coeff = self.X @ q
                return self.mo_coeff
        e, C, c = self._diag_orth(A)
        self._q_occ = c[:, :self.nocc].contiguous()
        self.mo_energy, self.mo_coeff = e, C
        if purify:
            self._evecs = c
        return C

    def _diag(self, F, orth=False):
        """FP64 eigenpairs of ``X^T F X``: on the device, or with LAPACK on
        the host when the orthogonal dimension is at most ``host_diag_max``
        (small symmetric eigenproblems are latency-bound on the GPU; a serving
        loop runs many of them side by side on host cores).  Either way the
        choice depends on the molecule only."""
        e, C, c = self._diag_orth(self.X.T @ F @ self.X)
        return (e, C, c) if orth else (e, C)

    def _diag_orth(self, A):
        """``(e, X c, c)`` with ``(e, c)`` the eigenpairs of ``A``, a Fock
        matrix in the orthogonal basis."""
        t = _dev.torch()
        if A.shape[0] <= self.host_diag_max:
            e, c = host_eigh(A.cpu().numpy(), self.concurrent_diag)
            e = t.from_numpy(e).to(A.device)
            c = t.from_numpy(c).to(A.device)
        else:
            # through the solver thread (eigsolve): concurrent hipSOLVER syevd
            # calls from a serving loop's worker threads crashed the process;
            # there, the requests diagonalising at once share one batched call
            e, c = eigsolve.eigh(A)
        return e, self.X @ c, c

    # ------------------------------------------------------------- loop
    def kernel(self, dm0=None):
        """Run to convergence with J/K from ``with_df.get_jk``; returns E."""
        self.start(dm0)
        while not self.done:
            self.with_df.k_precision = self.prec
            vj, vk = self.with_df.get_jk(self.dm, hermi=1)
            report = None
            if self.precision_report and self.prec != 'fp64':
                report = self._precision_errors(vj, vk)
            self.absorb(vj, vk)
            if report is not None:
                self.history[-1].update(report)
        self.with_df.k_precision = 'fp64'
        return self.e_tot

    def _precision_errors(self, vj, vk):
        """``max|X - X_fp64| / max|X_fp64|`` for J and K of the current
        density, X from the low-precision call and X_fp64 from an FP64 call
        on the same cderi (J is FP64 arithmetic in every mode, so its error
        is that of the cderi's storage type alone)."""
        t = _dev.torch()
        prec = self.with_df.k_precision
        self.with_df.k_precision = 'fp64'
        try:
            vj64, vk64 = self.with_df.get_jk(self.dm, hermi=1)
        finally:
            self.with_df.k_precision = prec
        err = t.stack([(vj - vj64).abs().max() / vj64.abs().max(),
                       (vk - vk64).abs().max() / vk64.abs().max()]).tolist()
        return {'j_rel_err': err[0], 'k_rel_err': err[1]}

    def start(self, dm0=None):
        """Initial density; afterwards the caller alternates J/K of ``self.dm``
        at precision ``self.prec`` with :meth:`absorb` until ``self.done``.
        This split lets a serving loop compute many solves' J/K in one
        grouped launch while each solve follows exactly this sequence."""
        if not hasattr(self, 'H'):
            self.build()
        # the eigenvectors of the last diagonalised Fock matrix (SP2 bounds)
        self._evecs = None
        if dm0 is None:
            _, C, c = self._diag(self.H, orth=True)
            self._q_occ = c[:, :self.nocc].contiguous()
            dm0 = self._density(C)
        else:
            self._q_occ = None              # dm0 is not 2 X q q^T X^T of a known q
        self.dm = dm0
        self._diis_reset()
        self._e_old = 0.
        self._final = False
        self.converged = False
        self.done = False
        self.cycles = 0
        self.prec = {'fp64': 'fp64', 'fp32': 'fp32', 'bf16x3': 'bf16x3',
                     'mixed': self.low_precision}[self.k_precision]
        return self

    def absorb(self, vj, vk):
        """One SCF cycle from the J/K of ``self.dm``: Fock, energy, DIIS,
        diagonalisation and the next density.  Converged on an FP64 J/K, the
        cycle's energy is the result (``conv_check``: one more FP64 J/K of
        the new density gives it); converged in a narrower precision or at
        ``max_cycle``, one more FP64 J/K gives the energy of the final
        density."""
        t = _dev.torch()
        F = self.H + vj - .5 * vk
        e_el = .5 * t.sum(self.dm * (self.H + F))          # stays on the device
        if self._final:
            self.e_tot = float(e_el) + self.enuc
            self.done = True
            return
        self.cycles = cycle = self.cycles + 1
        prec = self.prec
        # the Fock matrix and the DIIS error in the orthogonal basis (X^T S X
        # = I); with D = 2 X q q^T X^T, X^T (F D S - S D F) X is
        # 2 (Fo q q^T - q q^T Fo): two thin products instead of four square
        Fo = self.X.T @ F @ self.X
        q = self._q_occ
        if q is not None:
            M = (Fo @ q) @ q.T
            err = 2. * (M - M.T)
        else:
            err = F @ self.dm @ self.S
            err = self.X.T @ (err - err.T) @ self.X
        Fo = self._diis(Fo, err)
        # diagonalise first; the scalars come back in one transfer afterwards
        self._occupied(Fo)
        e_el, err_max = t.stack([e_el, err.abs().max()]).tolist()
        e_tot = e_el + self.enuc
        de = e_tot - self._e_old
        self._e_old = e_tot
        self.history.append({'cycle': cycle, 'e_tot': e_tot, 'de': de, 'err': err_max,
                             'k_precision': prec})
        if self.verbose:
            print(f'cycle {cycle}: E = {e_tot:.12f}  dE = {de:.3e}  |err| = {err_max:.3e}')
        self.dm = self._density(self.mo_coeff)
        if prec != 'fp64' and self.k_precision == 'mixed' and err_max < self.switch_err:
            self.prec = 'fp64'              # finishing cycles in FP64, fresh DIIS
            self._diis_reset()
        elif abs(de) < self.conv_tol and err_max < self.conv_tol_grad:
            self.converged = True           # (an all-FP32 run converges in FP32)
            if prec == 'fp64' and not self.conv_check:
                self.e_tot = e_tot
                self.done = True
                return
            self._final = True
        if cycle >= self.max_cycle:
            self._final = True
        if self._final:
            self.prec = 'fp64'
