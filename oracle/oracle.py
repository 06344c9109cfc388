"""CPU FP64 oracle for the density-fitted J/K path.

TEST INFRASTRUCTURE ONLY: this module is the checker that the HIP path is
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU
baseline leg of ``bench.py`` may import it; the shipped package never does.

It restates, in numpy/scipy FP64, the arithmetic of the reference's DF path
(paths relative to the Electrolites checkout):

* ``sparse_dm``      -- electrolites/fastdf.py:674-686 (pair-vector gather and
                        symmetry factors);
* ``get_j``          -- electrolites/fastdf.py:348-386 (two contractions over
                        the packed cderi) and :688-745 (J-only branch);
* ``get_k``          -- electrolites/fastdf.py:974-1005 (unpack, step 1
                        ``Lij,njk->nikL``, step 2 ``nikL,njkL->nij``);
* ``get_jk``         -- electrolites/fastdf.py:805-1026 (shapes, J-only
                        return ``(vj, None)``, ``nocc == 0`` early return);
* ``cholesky_eri``   -- electrolites/fastdfcderi.py:173-215 (metric Cholesky
                        and its eigen-filter fallback) and :352-693
                        (``cderi = L^-1 (P|pq)``, output ``(naux, npairs)``,
                        pair addresses ``i*nao + j``);
* integrals          -- ``oracle/ints.c`` (McMurchie-Davidson), the role of
                        GPU4PySCF's int3c2e/int2c2e evaluators.

The DF-RHF/UHF driver restates PySCF's ``scf.hf`` loop (Fock build, DIIS on
``FDS - SDF``, energy) for the total-energy oracle.
"""
import ctypes
import os

import numpy as np
import scipy.linalg

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, '_build', 'liboracle.so')
_lib = None

LINEAR_DEP_THR = 1e-9   # eigen-filter threshold for a non-PD metric


def build():
    """Compile ``liboracle.so`` (gcc, OpenMP)."""
    import subprocess
    subprocess.check_call(['make', '-s', '-C', _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            build()
        _lib = ctypes.CDLL(_LIBPATH)
    return _lib


def _p(a, ctype=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _shell_args(sh):
    l = np.ascontiguousarray(sh['l'], dtype=np.int32)
    npr = np.ascontiguousarray(sh['nprim'], dtype=np.int32)
    off = np.ascontiguousarray(sh['prim_off'], dtype=np.int32)
    aoo = np.ascontiguousarray(sh['ao_off'], dtype=np.int32)
    xyz = np.ascontiguousarray(sh['xyz'], dtype=np.float64)
    e = np.ascontiguousarray(sh['exps'], dtype=np.float64)
    c = np.ascontiguousarray(sh['coefs'], dtype=np.float64)
    keep = (l, npr, off, aoo, xyz, e, c)
    args = [ctypes.c_int(len(l)), _p(l, ctypes.c_int), _p(npr, ctypes.c_int),
            _p(off, ctypes.c_int), _p(aoo, ctypes.c_int), _p(xyz), _p(e),
            _p(c), ctypes.c_int(int(sh['nao']))]
    return args, keep


# ------------------------------------------------------------ primitives
def boys(m, t):
    """``[F_0(t) .. F_m(t)]``."""
    out = np.zeros(m + 1)
    lib().oracle_boys(ctypes.c_int(m), ctypes.c_double(t), _p(out))
    return out


def cart2sph(l):
    nc = (l + 1) * (l + 2) // 2
    out = np.zeros((2 * l + 1, nc))
    lib().oracle_cart2sph(ctypes.c_int(l), _p(out))
    return out


def int1e(kind, ao, charges=None, atom_xyz=None):
    """kind: 'ovlp', 'kin' or 'nuc'.  ``ao`` is a shell dict."""
    k = {'ovlp': 0, 'kin': 1, 'nuc': 2}[kind]
    nao = int(ao['nao'])
    out = np.zeros((nao, nao))
    args, keep = _shell_args(ao)
    if charges is None:
        charges = np.zeros(0)
        atom_xyz = np.zeros((0, 3))
    charges = np.ascontiguousarray(charges, dtype=np.float64)
    atom_xyz = np.ascontiguousarray(atom_xyz, dtype=np.float64)
    rc = lib().oracle_int1e(ctypes.c_int(k), *args, ctypes.c_int(len(charges)),
                            _p(charges), _p(atom_xyz), _p(out))
    if rc:
        raise ValueError('oracle_int1e: unsupported angular momentum')
    return out


def int2c2e(aux):
    naux = int(aux['nao'])
    out = np.zeros((naux, naux))
    args, keep = _shell_args(aux)
    rc = lib().oracle_int2c2e(*args[:-1], ctypes.c_int(naux), _p(out))
    if rc:
        raise ValueError('oracle_int2c2e: unsupported angular momentum')
    return out


def int3c2e(ao, aux):
    """``(nao, nao, naux)`` three-centre integrals ``(ij|P)``."""
    nao, naux = int(ao['nao']), int(aux['nao'])
    out = np.zeros((nao, nao, naux))
    a1, k1 = _shell_args(ao)
    a2, k2 = _shell_args(aux)
    rc = lib().oracle_int3c2e(*a1, *a2[:-1], ctypes.c_int(naux), _p(out))
    if rc:
        raise ValueError('oracle_int3c2e: unsupported angular momentum')
    return out


# ----------------------------------------------------------- pair layout
def tril_pairs(nao):
    """``(rows, cols)`` of the packed lower triangle, ``p = i(i+1)/2 + j``."""
    rows, cols = np.tril_indices(nao)
    return rows, cols


def pair_addresses(nao):
    rows, cols = tril_pairs(nao)
    return (rows * nao + cols).astype(np.int64), np.where(rows == cols)[0]


def unpack(cderi, pair_addr, nao):
    """``(naux, npairs)`` -> ``(naux, nao, nao)``, both triangles."""
    naux = cderi.shape[0]
    rows, cols = np.divmod(np.asarray(pair_addr), nao)
    out = np.zeros((naux, nao, nao))
    out[:, rows, cols] = cderi
    out[:, cols, rows] = cderi
    return out


# ----------------------------------------------------------------- cderi
def cholesky_eri(j3c, j2c, pair_addr=None):
    """``cderi[L, p] = sum_Q (L^-1)[L,Q] (Q|p)`` over stored pairs.

    Falls back to the eigen filter (``v / sqrt(w)`` for ``w > LINEAR_DEP_THR``)
    when the metric is not positive definite (fastdfcderi.py:203-210).
    """
    nao = j3c.shape[0]
    if pair_addr is None:
        pair_addr, _ = pair_addresses(nao)
    rows, cols = np.divmod(np.asarray(pair_addr), nao)
    b = j3c[rows, cols, :].T.copy()                 # (naux, npairs)
    try:
        low = scipy.linalg.cholesky(j2c, lower=True)
        return np.ascontiguousarray(
            scipy.linalg.solve_triangular(low, b, lower=True))
    except np.linalg.LinAlgError:
        w, v = scipy.linalg.eigh(j2c)
        keep = w > LINEAR_DEP_THR
        v = v[:, keep] / np.sqrt(w[keep])
        return np.ascontiguousarray(v.T @ b)


# ------------------------------------------------------------------- J/K
def sparse_dm(dms, pair_addr, hermi):
    """``d[n,p]``: ``2 D[i,j]`` (hermi 1) or ``D[i,j] + D[j,i]`` (hermi 0)
    on the stored pairs, with the diagonal halved (fastdf.py:674-686)."""
    nao = dms.shape[-1]
    dms = np.asarray(dms).reshape(-1, nao, nao)
    rows, cols = np.divmod(np.asarray(pair_addr), nao)
    d = dms[:, rows, cols].copy()
    if hermi == 0:
        d += dms[:, cols, rows]
    else:
        d *= 2
    diag = rows == cols
    d[:, diag] *= .5
    return d


def get_j(cderi, pair_addr, dms, hermi=1):
    nao = dms.shape[-1]
    d = sparse_dm(dms, pair_addr, hermi)
    rho = d @ cderi.T                               # (n_dm, naux)
    vjp = rho @ cderi                               # (n_dm, npairs)
    rows, cols = np.divmod(np.asarray(pair_addr), nao)
    vj = np.zeros((d.shape[0], nao, nao))
    vj[:, rows, cols] = vjp
    vj[:, cols, rows] = vjp
    return vj.reshape(np.shape(dms))


def get_k(cderi, pair_addr, dms):
    """``K[n] = sum_P A_P D_n A_P`` with ``A_P`` the unpacked cderi row."""
    nao = dms.shape[-1]
    a = unpack(cderi, pair_addr, nao)
    dd = np.asarray(dms).reshape(-1, nao, nao)
    out = np.empty_like(dd)
    for n in range(dd.shape[0]):
        t = np.matmul(a, dd[n])                     # (naux, nao, nao)
        out[n] = np.einsum('Pik,Pkj->ij', t, a, optimize=True)
    return out.reshape(np.shape(dms))


def get_k_factor(cderi, pair_addr, fl, fr=None):
    """Factored exchange: ``K = sum_L (A_L F_l)(A_L F_r)^T`` (mode 0 when
    ``fr`` is None), the form the reference computes (fastdf.py:991-1005)."""
    nao = fl.shape[-2]
    a = unpack(cderi, pair_addr, nao)
    rl = np.matmul(a, fl)
    rr = rl if fr is None else np.matmul(a, fr)
    return np.einsum('Lik,Ljk->ij', rl, rr, optimize=True)


def get_jk(cderi, pair_addr, dms, hermi=1, with_j=True, with_k=True):
    vj = get_j(cderi, pair_addr, dms, hermi) if with_j else None
    vk = get_k(cderi, pair_addr, dms) if with_k else None
    return vj, vk


# --------------------------------------------------------------- DF-SCF
class DFSystem:
    """Integrals of one molecule for the oracle SCF, from shell dicts."""

    def __init__(self, ao, aux, charges, atom_xyz, nelectron, enuc):
        self.ao, self.aux = ao, aux
        self.S = int1e('ovlp', ao)
        self.T = int1e('kin', ao)
        self.V = int1e('nuc', ao, charges, atom_xyz)
        self.H = self.T + self.V
        self.nao = int(ao['nao'])
        self.j3c = int3c2e(ao, aux)
        self.j2c = int2c2e(aux)
        self.pair_addr, self.diag = pair_addresses(self.nao)
        self.cderi = cholesky_eri(self.j3c, self.j2c, self.pair_addr)
        self.nelectron = nelectron
        self.enuc = enuc


def _orth(S, thr=1e-9):
    w, v = scipy.linalg.eigh(S)
    keep = w > thr
    return v[:, keep] / np.sqrt(w[keep])


def rhf(system, conv_tol=1e-10, max_cycle=100, diis_space=8, verbose=False):
    """Closed-shell DF-RHF (core-Hamiltonian guess, DIIS).

    Returns ``dict(e_tot, dm, mo_coeff, mo_energy, cycles, converged)``.
    """
    s = system
    X = _orth(s.S)
    nocc = s.nelectron // 2

    def diag(F):
        e, c = scipy.linalg.eigh(X.T @ F @ X)
        return e, X @ c

    e_mo, C = diag(s.H)
    dm = 2 * C[:, :nocc] @ C[:, :nocc].T
    e_old = 0.
    focks, errs = [], []
    conv = False
    for cycle in range(1, max_cycle + 1):
        vj, vk = get_jk(s.cderi, s.pair_addr, dm)
        F = s.H + vj - .5 * vk
        e_tot = .5 * np.sum(dm * (s.H + F)) + s.enuc
        err = F @ dm @ s.S - s.S @ dm @ F
        err = X.T @ err @ X
        focks.append(F)
        errs.append(err)
        focks, errs = focks[-diis_space:], errs[-diis_space:]
        n = len(focks)
        if n > 1:
            B = -np.ones((n + 1, n + 1))
            B[n, n] = 0.
            for i in range(n):
                for j in range(n):
                    B[i, j] = np.sum(errs[i] * errs[j])
            rhs = np.zeros(n + 1)
            rhs[n] = -1
            coef = np.linalg.solve(B, rhs)[:n]
            F = sum(c * f for c, f in zip(coef, focks))
        e_mo, C = diag(F)
        dm = 2 * C[:, :nocc] @ C[:, :nocc].T
        de = e_tot - e_old
        e_old = e_tot
        if verbose:
            print(cycle, e_tot, de, np.abs(err).max())
        if abs(de) < conv_tol and np.abs(err).max() < 1e-6:
            conv = True
            break
    # energy of the final density
    vj, vk = get_jk(s.cderi, s.pair_addr, dm)
    F = s.H + vj - .5 * vk
    e_tot = .5 * np.sum(dm * (s.H + F)) + s.enuc
    return dict(e_tot=e_tot, dm=dm, mo_coeff=C, mo_energy=e_mo,
                cycles=cycle, converged=conv)
