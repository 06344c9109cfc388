"""FP64 symmetric eigensolver on the device: rocSOLVER's strided-batched
``syevd`` behind one service thread.

An SCF cycle diagonalises its Fock matrix once; a serving loop has many
requests doing so at the same moment from its worker threads.  Several host
threads inside the solver library at once crashed the process (a SIGSEGV in
``hipsolverDsyevd``), and one call at a time leaves the GPU mostly idle (a
``syevd`` is a long chain of small kernels).  Here every device
diagonalisation of the package is a request to one thread, the only caller of
the solver library (it also holds ``_dev.SOLVER_LOCK``, which the package's
remaining library calls take): it gathers the matrices waiting at that
moment, pads each to its size class (the next multiple of ``PAD``) and runs
one ``rocsolver_dsyevd_strided_batched`` per class on its own stream.

Determinism: each matrix of a strided batch goes through the same kernels on
its own slice, so a matrix's eigenpairs depend only on the matrix and its
size class, not on what else shares the batch; an SCF gives the same bits
alone and in a serving mix.

Padding: ``A`` (n x n) fills the leading block of an N x N matrix whose
trailing block is diagonal, ``|A|_inf + 1 + k`` for k = 0 .. N - n - 1, above
every eigenvalue of ``A``.  The blocks are exactly decoupled (every
Householder vector of the reduction is zero in the other block) and the
eigenvalues come back ascending, so ``A``'s n come first, with eigenvectors
that vanish beyond row n.
"""
import concurrent.futures
import ctypes
import os
import queue
import threading

from . import _dev

#: size classes: matrices are padded to the next multiple of this
PAD = 64

_EVECT_ORIGINAL = 211          # rocblas_evect_original
_FILL_LOWER = 122              # rocblas_fill_lower
_STATUS_OK = (0, 9, 10)        # success; size increased / unchanged (size queries)

_libs = None
_services = {}
_services_lock = threading.Lock()
STATS = {'calls': 0, 'batches': 0, 'matrices': 0}


def _load():
    global _libs
    if _libs is None:
        root = os.environ.get('ROCM_PATH', '/opt/rocm')
        rb = ctypes.CDLL(os.path.join(root, 'lib', 'librocblas.so'))
        rs = ctypes.CDLL(os.path.join(root, 'lib', 'librocsolver.so'))
        P, I32, I64, SZ = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_size_t
        for f, args in ((rb.rocblas_create_handle, [ctypes.POINTER(P)]),
                        (rb.rocblas_set_stream, [P, P]),
                        (rb.rocblas_start_device_memory_size_query, [P]),
                        (rb.rocblas_stop_device_memory_size_query, [P, ctypes.POINTER(SZ)]),
                        (rb.rocblas_set_workspace, [P, P, SZ]),
                        (rs.rocsolver_dsyevd_strided_batched,
                         [P, ctypes.c_int, ctypes.c_int, I32, P, I32, I64, P, I64, P, I64, P, I32])):
            f.argtypes = args
            f.restype = ctypes.c_int
        _libs = (rb, rs)
    return _libs


def size_class(n):
    return -(-int(n) // PAD) * PAD


class _Service:
    """The solver thread of one device."""

    def __init__(self, device):
        self.device = device
        self.q = queue.Queue()
        self.ready = threading.Event()
        self.error = None
        self.thread = threading.Thread(target=self._run, name=f'mdf-eigh-{device}', daemon=True)
        self.thread.start()
        self.ready.wait()
        if self.error is not None:
            raise self.error

    def _setup(self):
        t = _dev.torch()
        t.cuda.set_device(self.device)
        self.stream = t.cuda.Stream()
        rb, _ = _load()
        h = ctypes.c_void_p()
        self._check(rb.rocblas_create_handle(ctypes.byref(h)), 'rocblas_create_handle')
        self.handle = h
        self._check(rb.rocblas_set_stream(h, ctypes.c_void_p(self.stream.cuda_stream)), 'rocblas_set_stream')
        self.work = None
        self.work_bytes = 0
        self.sizes = {}                 # (N, batch) -> workspace bytes

    @staticmethod
    def _check(status, what):
        if status not in _STATUS_OK:
            raise RuntimeError(f'{what} returned rocblas status {status}')

    def submit(self, A):
        t = _dev.torch()
        fut = concurrent.futures.Future()
        ev = t.cuda.Event()
        ev.record(t.cuda.current_stream())
        self.q.put((A, ev, fut))
        return fut

    def _run(self):
        try:
            self._setup()
        except Exception as exc:                              # noqa: BLE001
            self.error = exc
            self.ready.set()
            return
        self.ready.set()
        while True:
            item = self.q.get()
            if item is None:
                return
            batch = [item]
            while True:
                try:
                    nxt = self.q.get_nowait()
                except queue.Empty:
                    break
                if nxt is None:
                    self.q.put(None)
                    break
                batch.append(nxt)
            groups = {}
            for it in batch:
                groups.setdefault(size_class(it[0].shape[0]), []).append(it)
            for N, group in sorted(groups.items()):
                try:
                    self._solve(N, group)
                except Exception as exc:                      # noqa: BLE001
                    for _, _, fut in group:
                        if not fut.done():
                            fut.set_exception(exc)

    def _call(self, N, buf, w, e, info, nb):
        _, rs = _load()
        return rs.rocsolver_dsyevd_strided_batched(
            self.handle, _EVECT_ORIGINAL, _FILL_LOWER, N, ctypes.c_void_p(buf.data_ptr()), N, N * N,
            ctypes.c_void_p(w.data_ptr()), N, ctypes.c_void_p(e.data_ptr()), N,
            ctypes.c_void_p(info.data_ptr()), nb)

    def _workspace(self, N, buf, w, e, info, nb):
        """Give the handle a workspace of the size this batch shape needs
        (queried once per shape), so the library allocates nothing itself."""
        t = _dev.torch()
        rb, _ = _load()
        need = self.sizes.get((N, nb))
        if need is None:
            self._check(rb.rocblas_start_device_memory_size_query(self.handle), 'size query')
            self._check(self._call(N, buf, w, e, info, nb), 'rocsolver_dsyevd_strided_batched (size query)')
            sz = ctypes.c_size_t()
            self._check(rb.rocblas_stop_device_memory_size_query(self.handle, ctypes.byref(sz)), 'size query')
            need = self.sizes[(N, nb)] = int(sz.value)
        if need > self.work_bytes:
            self.work = t.empty(max(need, 256), dtype=t.uint8, device=self.device)
            self.work_bytes = self.work.numel()
            self._check(rb.rocblas_set_workspace(self.handle, ctypes.c_void_p(self.work.data_ptr()),
                                                 self.work_bytes), 'rocblas_set_workspace')

    def _solve(self, N, group):
        t = _dev.torch()
        nb = len(group)
        dev = self.device
        with t.cuda.stream(self.stream):
            buf = t.zeros((nb, N, N), dtype=t.float64, device=dev)
            for i, (A, ev, _) in enumerate(group):
                self.stream.wait_event(ev)
                n = A.shape[0]
                buf[i, :n, :n] = A
                if n < N:
                    k = t.arange(n, N, device=dev)
                    buf[i, k, k] = A.abs().sum(dim=1).max() + 1. + (k - n).to(t.float64)
            w = t.empty((nb, N), dtype=t.float64, device=dev)
            e = t.empty((nb, N), dtype=t.float64, device=dev)
            info = t.zeros(nb, dtype=t.int32, device=dev)
            with _dev.SOLVER_LOCK:
                self._workspace(N, buf, w, e, info, nb)
                self._check(self._call(N, buf, w, e, info, nb), 'rocsolver_dsyevd_strided_batched')
            outs = []
            for i, (A, _, _) in enumerate(group):
                n = A.shape[0]
                # column j of the column-major result (row j of this view) is
                # eigenvector j; the caller gets them as columns
                outs.append((w[i, :n].clone(), buf[i, :n, :n].transpose(0, 1).contiguous()))
            infos = info.cpu()                   # waits for the batch on this stream
        STATS['batches'] += 1
        STATS['matrices'] += nb
        for i, (_, _, fut) in enumerate(group):
            if int(infos[i]) != 0:
                fut.set_exception(RuntimeError(f'syevd did not converge (info {int(infos[i])})'))
            else:
                fut.set_result(outs[i])


def _service(device):
    with _services_lock:
        s = _services.get(device)
        if s is None:
            s = _services[device] = _Service(device)
        return s


def eigh(A):
    """``(w, v)`` of the symmetric FP64 device matrix ``A`` (ascending
    eigenvalues, eigenvectors as columns), as ``torch.linalg.eigh``; callable
    from any thread, ordered after the work queued on the caller's current
    stream.  The results are complete when this returns."""
    t = _dev.torch()
    if A.dim() != 2 or A.shape[0] != A.shape[1] or A.dtype != t.float64 or not A.is_cuda:
        raise ValueError('eigsolve.eigh: a square float64 device matrix is required')
    STATS['calls'] += 1
    if A.shape[0] == 0:
        return (t.empty(0, dtype=t.float64, device=A.device),
                t.empty((0, 0), dtype=t.float64, device=A.device))
    w, v = _service(A.device.index if A.device.index is not None else t.cuda.current_device()) \
        .submit(A.contiguous()).result()
    cur = t.cuda.current_stream()
    w.record_stream(cur)
    v.record_stream(cur)
    return w, v
