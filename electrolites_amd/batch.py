"""Grouped J/K over many molecules and densities (persistent launches).

The reference contracts one molecule per ``get_jk`` call; a batch of small
molecules then leaves most of the GPU idle.  ``GroupedJK`` hands a whole batch
to ``mdf_grouped_plan_create`` / ``mdf_grouped_jk``: every item's J and K tiles
go into tables the library keeps on the device, and each phase runs as one
persistent launch over all items.  A run is launches only, so it can be
captured once in a HIP graph (``capture_graph``) and replayed.

Each item is cut and computed exactly as :func:`electrolites_amd.fastdf.get_jk`
cuts it, so results are bit-identical alone and in any batch.

The step-1 tile width is chosen by a small autotuner: the first plan for a
batch signature times every candidate width and the fastest is cached.
"""
import ctypes
import time

import numpy as np

from . import _dev
from . import _lib
from . import fastdf

P = ctypes.c_void_p


class JKItem(ctypes.Structure):
    _fields_ = [('cderi', P), ('ld', ctypes.c_long), ('naux', ctypes.c_long),
                ('npairs', ctypes.c_long), ('nao', ctypes.c_int), ('pair_map', P),
                ('n_dm', ctypes.c_int), ('dm_pairs', P), ('rho', P), ('vj_pairs', P),
                ('fl', P), ('fr', P), ('nocc', ctypes.c_int), ('vk', P), ('cderi_dtype', ctypes.c_int)]



class Entry:
    """One molecule's device buffers for a grouped run.

    ``cderi`` (naux, npairs) float64 or float32 (FP32 storage, widened on
    load; one storage type per batch); ``pair_map`` int32 (nao*nao) or None;
    ``dm_pairs`` (n_dm, npairs); ``fl``/``fr`` (n_dm, nao, nocc); outputs
    ``rho`` (n_dm, naux), ``vj_pairs`` (n_dm, npairs), ``vk`` (n_dm, nao, nao)
    are allocated here unless given.
    """

    def __init__(self, cderi, nao, n_dm, nocc, dm_pairs=None, fl=None, fr=None, pair_map=None):
        t = _dev.torch()
        dev = cderi.device
        self.cderi, self.nao, self.n_dm, self.nocc = cderi, nao, n_dm, nocc
        self.naux, self.npairs = cderi.shape
        self.pair_map = pair_map
        self.dm_pairs = dm_pairs if dm_pairs is not None else \
            t.zeros((n_dm, self.npairs), dtype=t.float64, device=dev)
        self.fl = fl if fl is not None else t.zeros((n_dm, nao, nocc), dtype=t.float64, device=dev)
        self.fr = fr
        self.rho = t.empty((n_dm, self.naux), dtype=t.float64, device=dev)
        self.vj_pairs = t.empty((n_dm, self.npairs), dtype=t.float64, device=dev)
        self.vk = t.empty((n_dm, nao, nao), dtype=t.float64, device=dev)

    def item(self):
        p = lambda x: None if x is None else x.data_ptr()
        return JKItem(self.cderi.data_ptr(), self.cderi.stride(0), self.naux, self.npairs, self.nao,
                      p(self.pair_map), self.n_dm, p(self.dm_pairs), p(self.rho), p(self.vj_pairs),
                      p(self.fl), p(self.fr), self.nocc, p(self.vk), fastdf._dtype_code(self.cderi))

    def signature(self):
        return (self.nao, self.naux, self.npairs, self.n_dm, self.nocc, self.fr is not None)


_TUNED = {}
BN_CANDIDATES = (0, 128, 320)      # 0: each item's own width (kcore::pick_bn)


class GroupedJK:
    """A device plan for one batch of :class:`Entry` objects."""

    def __init__(self, entries, with_j=True, with_k=True, precision='fp64', bn=None, tune=True,
                 work_limit=0, work_pool=None):
        self.entries = list(entries)
        self.work_limit = int(work_limit)
        # a shared grow-only workspace (an _dev.Workspace) for callers that
        # run many plans one after another on one stream; None: own buffer
        self.work_pool = work_pool
        self.with_j, self.with_k = with_j, with_k
        self.precision = precision
        self.prec = fastdf.PRECISIONS[precision]
        arr = (JKItem * max(len(self.entries), 1))(*[e.item() for e in self.entries])
        self._arr = arr
        sig = (tuple(sorted(e.signature() for e in self.entries)), with_j, with_k, precision)
        if bn is None:
            bn = _TUNED.get(sig) if tune else 0
            if bn is None:
                bn = self._autotune(sig)
        self.bn = bn
        self._plan, self.work_bytes = self._create(bn)
        self.work = None if work_pool is not None else \
            _dev.torch().empty(max(self.work_bytes, 256), dtype=_dev.torch().uint8, device=_dev.device())
        self.graph = None

    def _create(self, bn):
        plan = P()
        wb = ctypes.c_size_t()
        _lib.call('mdf_grouped_plan_create', ctypes.cast(self._arr, P), len(self.entries), int(self.with_j),
                  int(self.with_k), self.prec, bn, self.work_limit, ctypes.byref(plan), ctypes.byref(wb))
        return plan, wb.value

    def _autotune(self, sig):
        """Time each step-1 tile width once on this batch; keep the fastest."""
        t = _dev.torch()
        if not self.with_k:
            _TUNED[sig] = 0
            return 0
        best, best_t = 0, None
        for bn in BN_CANDIDATES:
            plan, wb = self._create(bn)
            try:
                work = t.empty(max(wb, 256), dtype=t.uint8, device=_dev.device())
                _lib.call('mdf_grouped_jk', plan, work.data_ptr(), work.numel(), _dev.stream_ptr())
                t.cuda.synchronize()
                t0 = time.perf_counter()
                _lib.call('mdf_grouped_jk', plan, work.data_ptr(), work.numel(), _dev.stream_ptr())
                t.cuda.synchronize()
                dt = time.perf_counter() - t0
                del work
            finally:
                _lib.call('mdf_grouped_plan_destroy', plan)
            if best_t is None or dt < best_t:
                best, best_t = bn, dt
        _TUNED[sig] = best
        return best

    def tiles(self):
        out = (ctypes.c_long * 9)()
        _lib.call('mdf_grouped_plan_info', self._plan, out, 9)
        names = ('j_pass1', 'j_reduce1', 'j_pass2', 'j_reduce2', 'k_step1', 'k_step2', 'k_finalize',
                 'k_waves', 'k_launches')
        return dict(zip(names, list(out)))

    def run(self):
        """Launch the batch on the current stream (or replay its graph)."""
        if self.graph is not None:
            self.graph.replay()
            return
        work = self.work if self.work is not None else self.work_pool.get(self.work_bytes)
        _lib.call('mdf_grouped_jk', self._plan, work.data_ptr(), work.numel(), _dev.stream_ptr())

    def capture_graph(self):
        """Capture one run in a HIP graph; later ``run()`` calls replay it."""
        t = _dev.torch()
        if self.work is None:
            raise ValueError('a captured plan needs its own workspace (work_pool=None)')
        s = t.cuda.Stream()
        s.wait_stream(t.cuda.current_stream())
        with t.cuda.stream(s):
            self.run()                      # warm once outside the capture
        t.cuda.current_stream().wait_stream(s)
        g = t.cuda.CUDAGraph()
        with t.cuda.graph(g):
            _lib.call('mdf_grouped_jk', self._plan, self.work.data_ptr(), self.work.numel(),
                      _dev.stream_ptr())
        self.graph = g
        return g

    def close(self):
        if getattr(self, '_plan', None):
            _lib.call('mdf_grouped_plan_destroy', self._plan)
            self._plan = None
        self.graph = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batched_get_jk(dfobjs, dms_list, hermi=1, precision='fp64', work_limit=0):
    """``[(vj, vk), ...]`` for each (DF object, densities) pair in one grouped
    run.  Densities follow :func:`fastdf.get_jk` (tagged or not); outputs are
    device tensors shaped like the inputs.  ``work_limit`` bounds the K
    workspace of one wave of items (bytes, 0 = the library default)."""
    t = _dev.torch()
    entries, shapes = [], []
    for dfobj, dms in zip(dfobjs, dms_list):
        cderi = fastdf._cderi_of(dfobj)
        nao = dms.shape[-1]
        lay = fastdf._layout(dfobj, cderi, nao)
        dms3 = _dev.to_device(dms).reshape(-1, nao, nao)
        fl, fr = fastdf.factorize_dm(dms, hermi)
        d = fastdf._sparse_dm(dms3, lay, hermi)
        e = Entry(cderi, nao, dms3.shape[0], fl.shape[-1], dm_pairs=d, fl=fl, fr=fr, pair_map=lay.map)
        e.layout = lay
        entries.append(e)
        shapes.append(tuple(np.shape(dms)))
    g = GroupedJK(entries, precision=precision, work_limit=work_limit)
    g.run()
    out = []
    for e, shp in zip(entries, shapes):
        vj = t.empty((e.n_dm, e.nao, e.nao), dtype=t.float64, device=e.cderi.device)
        _lib.call('mdf_unpack_pairs', _dev.ptr(e.vj_pairs), e.n_dm, e.npairs, _dev.ptr(e.layout.addr),
                  e.npairs, e.nao, _dev.ptr(vj), e.nao * e.nao, _dev.stream_ptr())
        out.append((vj.reshape(shp), e.vk.reshape(shp)))
    g.close()
    return out
