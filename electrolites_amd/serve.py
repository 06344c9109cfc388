"""Continuous-batching DF-SCF serving on one GPU.

A stream of DF-RHF energy requests (molecules of uneven size) is served by a
loop that keeps every admitted request's state in HBM -- its cderi, one-
electron matrices, density factors, DIIS history and the device buffers of
its J/K item -- and advances each of them one SCF cycle at a time:

1. admission: queued requests are admitted in arrival order while the
   resident bytes (the stored cderi + per-request matrices,
   :func:`footprint`) stay within ``hbm_budget`` and the batch has a free
   slot; an admitted request's cderi and one-electron matrices are built by
   a builder thread on a stream of its own while the loop goes on, and the
   request joins at the first step after its build finished;
2. J/K: every request whose next density is ready goes into one grouped
   launch per precision group (:class:`batch.GroupedJK`), so requests that
   arrived while others were mid-solve share launches with them;
3. absorb (:meth:`scf.RHF.absorb`: Fock, energy, DIIS, diagonalisation, next
   density, precision switch) on a worker thread and the request's own HIP
   stream (default priority), right after its J/K was launched; finished
   requests leave and free their bytes for the queue.

Schedules.  ``'async'`` (the default): no request waits for another; a step
launches the J/K of every request whose absorb finished since the last step,
so while the host diagonalises one request's Fock matrix the GPU runs the
others' J/K.  ``'lockstep'``: every active request gets its J/K in one launch,
then all absorb, then the next tick (``pipeline=True`` alternates two
groups).  Measured on the 48-request mix at 8 req/s: async 19-23 thousand
solves/hr against 13-15 thousand lockstep.

Diagonalisation.  Orthogonal dimensions up to ``host_diag_max`` (default
2000, every molecule of the mix) go to LAPACK on the host, one thread per
call, many calls side by side (:func:`scf.host_eigh`); larger ones to the
device eigensolver thread (:mod:`electrolites_amd.eigsolve`), which batches
the requests waiting at once.  Several worker threads calling hipSOLVER
``syevd`` at the same time crashed the process (SIGSEGV inside
``hipsolverDsyevd``); no worker thread calls a device solver library now.
A device ``syevd`` is a chain of small kernels (about 18 ms at 816 x 816
alone), so on this workload the host calls, running in parallel, are faster.
Requests above ``purify_min`` (default 600) skip the diagonalisation: their
occupied space comes from SP2 purification of the Fock matrix on the GPU
(:meth:`scf.RHF._purify_occ`, FP64 products only), which is faster than a
single-threaded host diagonalisation at that size.

Each request runs exactly the cycle sequence of a stand-alone
:class:`scf.RHF` with the same options, and its grouped J/K item is cut
exactly as the single-item entries cut it, so its energy is the same alone
and in any mix (bit-identical; the tests compare to 1e-12 Eh).

The reference has no serving loop (one ``get_jk`` per call,
electrolites/fastdf.py:805-1026); this is the caller-side batching built on
the grouped launch.
"""
import collections
import concurrent.futures
import math
import time

import numpy as np

from . import _dev
from . import _lib
from . import batch
from . import fastdf
from .scf import RHF, single_thread_blas

#: default resident-byte budget (of the 288 GB of HBM, leaving headroom for
#: workspaces and the framework's allocator)
HBM_BUDGET = 250e9


def footprint(mol, diis_space=8):
    """Resident device bytes of one request: the cderi over the pairs the
    build stores (the screened-in pairs, counted on the host from the
    shell-pair bounds before the build, :func:`fastdfcderi.kept_pairs`) plus
    its nao x nao matrices (S, H, X, DIIS Fock and error history, J/K
    buffers, the pair map)."""
    from .fastdfcderi import kept_pairs
    nao, naux = mol.nao, mol.naux
    npairs = int(kept_pairs(mol)[0].sum())
    return 8 * (naux * npairs + (2 * diis_space + 10) * nao * nao + 2 * npairs + naux) + 4 * nao * nao


def molecule_mix(n, seed=0, lo=3, hi=40):
    """``n`` molecules with a log-uniform number of waters in [lo, hi]; every
    second one carries a Li+ ion at the centroid (closed shell, charge +1)."""
    from .gto.cluster import water_cluster
    rng = np.random.default_rng(seed)
    mols = []
    for i in range(n):
        nw = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        mols.append(water_cluster(nw, seed=int(rng.integers(1 << 30)), li=bool(i % 2)))
    return mols


def poisson_arrivals(n, rate, seed=0):
    """Arrival times (s) of a Poisson stream of ``n`` requests at ``rate``/s."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1. / rate, size=n)) if rate > 0 else np.zeros(n)


class Request:
    """One energy request.  Filled in by the server: ``t_admit``, ``t_done``
    (seconds on the server clock), ``e_tot``, ``cycles``, ``converged``,
    ``history``."""

    def __init__(self, mol, arrival=0., rid=None, k_precision='mixed'):
        self.mol = mol
        self.arrival = float(arrival)
        self.rid = rid
        self.k_precision = k_precision
        self.bytes = footprint(mol)
        self.t_admit = self.t_done = None
        self.e_tot = None
        self.cycles = 0
        self.converged = False
        self.history = None

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival


#: ticks a grouped plan stays cached after its last use
PLAN_KEEP_TICKS = 32


class _Slot:
    """An admitted request: its SCF state and its grouped J/K item."""

    def __init__(self, req, conv_tol, max_cycle, host_diag_max=0, low_precision='fp32', purify_min=None):
        t = _dev.torch()
        self.req = req
        self.scf = RHF(req.mol, conv_tol=conv_tol, max_cycle=max_cycle,
                       k_precision=req.k_precision, host_diag_max=host_diag_max,
                       low_precision=low_precision, purify_min=purify_min)
        # the server constructs and absorbs slots inside single_thread_blas
        self.scf.concurrent_diag = True
        self.scf.build()
        self.scf.start()
        dfobj = self.scf.with_df
        self.cderi = fastdf._cderi_of(dfobj)
        nao = req.mol.nao
        self.lay = fastdf._layout(dfobj, self.cderi, nao)
        self.entry = batch.Entry(self.cderi, nao, 1, self.scf.nocc, pair_map=self.lay.map)
        self.entry.layout = self.lay
        self.nao = nao
        self._vj = t.empty((1, nao, nao), dtype=t.float64, device=self.cderi.device)
        self.stream = t.cuda.Stream()
        self.jk_done = None           # event after this slot's launched J/K
        self.ready = None             # event after this slot's last absorb
        self.future = None            # its running absorb (asynchronous schedule)

    def load_density(self):
        """Density factor and packed density of the current ``scf.dm``, the
        values :func:`fastdf.get_jk` computes for it (the factor of a density
        tagged with occupied orbitals C and occupations n, all > 0 here, is
        C sqrt(n), the same device operations), without a host round trip."""
        t = _dev.torch()
        dm = self.scf.dm
        self.entry.fl[0].copy_(dm.mo_coeff * t.sqrt(dm.mo_occ))
        e = self.entry
        _lib.call('mdf_pack_dm', _dev.ptr(dm), 1, self.nao, self.nao * self.nao,
                  _dev.ptr(self.lay.addr), e.npairs, 1, _dev.ptr(e.dm_pairs), e.npairs,
                  _dev.stream_ptr())

    def absorb(self):
        e = self.entry
        _lib.call('mdf_unpack_pairs', _dev.ptr(e.vj_pairs), 1, e.npairs, _dev.ptr(self.lay.addr),
                  e.npairs, self.nao, _dev.ptr(self._vj), self.nao * self.nao, _dev.stream_ptr())
        self.scf.absorb(self._vj[0], e.vk[0])
        return self.scf.done


class Server:
    """The serving loop.  ``submit`` queues requests; ``step`` runs one tick;
    ``serve`` drives a timed arrival stream to completion."""

    def __init__(self, hbm_budget=HBM_BUDGET, max_batch=64, conv_tol=1e-10, max_cycle=100,
                 graphs=False, workers=16, host_diag_max=2000, async_build=True,
                 low_precision='bf16x3', pipeline=False, schedule='async', purify_min=600):
        if schedule not in ('async', 'lockstep'):
            raise ValueError(f'schedule {schedule!r}: expected async or lockstep')
        self.hbm_budget = float(hbm_budget)
        self.schedule = schedule
        self.pipeline = pipeline and schedule == 'lockstep'
        if schedule == 'async':
            # an absorb holds its worker while it waits for its J/K on the
            # device, so every admitted request may need one
            workers = max(workers, max_batch)
        self._plan_tick = {}
        self._weight = [0., 0.]
        self._work = _dev.Workspace(grow=2.0)
        self.workers = workers
        self.host_diag_max = host_diag_max
        self.purify_min = purify_min
        self.low_precision = low_precision
        self._pool = concurrent.futures.ThreadPoolExecutor(workers) \
            if workers > 1 or schedule == 'async' else None
        self._builder = concurrent.futures.ThreadPoolExecutor(1) if async_build else None
        self._building = []            # (future, request) of admitted requests being built
        self.timing = collections.Counter()
        self.max_batch = max_batch
        self.conv_tol = conv_tol
        self.max_cycle = max_cycle
        self.graphs = graphs
        self.queue = collections.deque()
        self.active = []
        self.done = []
        self.resident = 0.
        self._plans = {}
        self._dead_plans = []          # plans of retired requests, closed at close()
        self._serial = 0
        self.ticks = 0
        self.t0 = time.perf_counter()
        self.trace = []               # (time, n_active, n_queued, resident bytes)

    def now(self):
        return time.perf_counter() - self.t0

    @property
    def busy(self):
        """Requests queued, being built or in the batch."""
        return bool(self.queue or self.active or self._building)

    def submit(self, req):
        if req.bytes > self.hbm_budget:
            raise ValueError(f'request {req.rid}: {req.bytes / 1e9:.1f} GB does not fit the '
                             f'{self.hbm_budget / 1e9:.1f} GB budget')
        self.queue.append(req)

    def _build(self, req):
        t = _dev.torch()
        stream = t.cuda.Stream()
        with t.cuda.stream(stream), single_thread_blas():
            slot = _Slot(req, self.conv_tol, self.max_cycle, self.host_diag_max, self.low_precision,
                         self.purify_min)
            ev = t.cuda.Event()
            ev.record(stream)
        return slot, ev

    def _admit(self):
        t = _dev.torch()
        while self.queue and len(self.active) + len(self._building) < self.max_batch and \
                self.resident + self.queue[0].bytes <= self.hbm_budget:
            req = self.queue.popleft()
            req.t_admit = self.now()
            self.resident += req.bytes
            if self._builder is None:
                with single_thread_blas():
                    self._activate(_Slot(req, self.conv_tol, self.max_cycle, self.host_diag_max,
                                         self.low_precision, self.purify_min))
            else:
                self._building.append(self._builder.submit(self._build, req))
        # requests whose build finished join the batch; when nothing is
        # active, wait for the first build rather than spin
        if self._building and not self.active:
            concurrent.futures.wait(self._building[:1])
        still = []
        for f in self._building:
            if f.done():
                slot, ev = f.result()
                t.cuda.current_stream().wait_event(ev)
                self._activate(slot)
            else:
                still.append(f)
        self._building = still

    def _activate(self, slot):
        """Join the batch, in the lighter of the two pipeline groups (work
        weighed as nao^3)."""
        slot.serial = self._serial
        self._serial += 1
        slot.weight = float(slot.nao) ** 3
        slot.group = 0 if self._weight[0] <= self._weight[1] else 1
        self._weight[slot.group] += slot.weight
        self.active.append(slot)

    def _plan(self, slots, prec):
        # keyed by the slots' serial numbers, which are never reused (an
        # object's id() is: a new request could otherwise meet a plan cached
        # for the one that held its address before, and read freed buffers)
        key = (prec,) + tuple(sorted(s.serial for s in slots))
        g = self._plans.get(key)
        if g is None:
            # plans run one after another on one stream: one shared workspace
            # (per-plan buffers would pile up in the caching allocator as the
            # membership changes, starving the libraries' own allocations)
            g = batch.GroupedJK([s.entry for s in slots], precision=prec, tune=False,
                                work_pool=None if self.graphs else self._work)
            g.uses = 0
            self._plans[key] = g
        g.uses += 1
        if self.graphs and g.graph is None and g.uses == 2:
            g.capture_graph()
        return key, g

    def _retire(self, s):
        # the plans this request was part of are never run again; their
        # references to its buffers go now (so its bytes are freed), their
        # small device tables when the server closes (freeing device memory
        # synchronises the device)
        for key in [k for k in self._plans if s.serial in k[1:]]:
            g = self._plans.pop(key)
            self._plan_tick.pop(key, None)
            g.entries = None
            g._arr = None
            g.graph = None
            g.work = None
            self._dead_plans.append(g)
        r, scf = s.req, s.scf
        r.t_done = self.now()
        r.e_tot, r.cycles, r.converged = scf.e_tot, scf.cycles, scf.converged
        r.history = scf.history
        self.resident -= r.bytes
        self.done.append(r)

    def _launch(self, slots):
        """Grouped J/K of the slots' current densities on the current stream,
        one launch per precision group; returns each slot's event, recorded
        right after its own group's launch (a request of the first group
        starts its absorb while the next group's J/K still runs)."""
        t = _dev.torch()
        main = t.cuda.current_stream()
        groups = collections.defaultdict(list)
        for s in slots:
            if s.ready is not None:
                main.wait_event(s.ready)        # its absorb wrote the density
            s.load_density()
            groups[s.scf.prec].append(s)
        # every group's plan first (a new plan uploads its tables), then the runs
        plans = []
        for prec, members in sorted(groups.items()):
            key, g = self._plan(members, prec)
            self._plan_tick[key] = self.ticks
            plans.append((g, members))
        done = {}
        for g, members in plans:
            g.run()
            ev = t.cuda.Event()
            ev.record(main)
            for s in members:
                done[s.serial] = ev
        return done

    def _absorb_all(self, slots):
        """Each slot absorbs its J/K on its own stream from a worker thread;
        returns the slots that finished."""
        t = _dev.torch()

        def absorb(s):
            with t.cuda.stream(s.stream):
                s.stream.wait_event(s.jk_done)
                finished = s.absorb()
                ev = t.cuda.Event()
                ev.record(s.stream)
            s.jk_done, s.ready = None, ev
            return finished

        with single_thread_blas():
            if self._pool is not None and len(slots) > 1:
                done = list(self._pool.map(absorb, slots))
            else:
                done = [absorb(s) for s in slots]
        return [s for s, d in zip(slots, done) if d]

    def step(self):
        """Admit, then advance the batch: :meth:`_step_async` (the default
        schedule) or one lockstep tick (:meth:`_step_lockstep`)."""
        if self.schedule == 'async':
            return self._step_async()
        return self._step_lockstep()

    def _absorb_one(self, s):
        """A worker thread's part of one cycle of one request: wait (on the
        request's stream) for its J/K, absorb it, record the event the next
        launch waits on; True when the request finished."""
        t = _dev.torch()
        with single_thread_blas():
            with t.cuda.stream(s.stream):
                s.stream.wait_event(s.jk_done)
                finished = s.absorb()
                ev = t.cuda.Event()
                ev.record(s.stream)
        s.ready = ev
        return finished

    def _step_async(self):
        """Asynchronous cycles: every request whose absorb has finished since
        the last call (its next density is being written on its stream) joins
        one grouped J/K launch per precision group at once, and its absorb is
        handed to a worker right after the launch, where it waits for that
        J/K on the device.  No request waits for another's diagonalisation:
        while the host diagonalises one request's Fock matrix, the GPU runs
        the others' J/K.  Each request still runs its stand-alone cycle
        sequence, so its result is the same alone and in the mix."""
        t0 = time.perf_counter()
        self._admit()
        t1 = time.perf_counter()
        self.timing['admit_s'] += t1 - t0
        if not self.active:
            return 0
        finished = []
        for s in self.active:
            f = s.future
            if f is not None and f.done():
                s.future = None
                if f.result():
                    finished.append(s)
        for s in finished:
            s.stream.synchronize()
            self._retire(s)
            self._weight[s.group] -= s.weight
        if finished:
            gone = set(map(id, finished))
            self.active = [s for s in self.active if id(s) not in gone]
        ready = [s for s in self.active if s.future is None]
        if ready:
            self.ticks += 1
            evs = self._launch(ready)
            for s in ready:
                s.jk_done = evs[s.serial]
                s.future = self._pool.submit(self._absorb_one, s)
            self.timing['launch_s'] += time.perf_counter() - t1
            self.trace.append((self.now(), len(self.active), len(self.queue), self.resident))
        elif not finished:
            tw = time.perf_counter()
            pending = [s.future for s in self.active if s.future is not None]
            if pending:
                concurrent.futures.wait(pending, timeout=.005,
                                        return_when=concurrent.futures.FIRST_COMPLETED)
            self.timing['wait_s'] += time.perf_counter() - tw
        return len(self.active)

    def _step_lockstep(self):
        """Admit, then one tick.  Lockstep (``pipeline=False``): every active
        request gets its J/K, then absorbs it.  Pipelined: the requests are
        in two groups; a tick launches one group's J/K and, while the GPU
        runs it, absorbs the other group's J/K from the previous tick, so the
        host-side and latency-bound parts of one group overlap the other
        group's contractions.  Either way each request follows its own
        stand-alone cycle sequence."""
        t = _dev.torch()
        t0 = time.perf_counter()
        self._admit()
        t1 = time.perf_counter()
        self.timing['admit_s'] += t1 - t0
        if not self.active:
            return 0
        self.ticks += 1
        if self.pipeline:
            g = self.ticks % 2
            launch = [s for s in self.active if s.group == g and s.jk_done is None]
            absorb = [s for s in self.active if s.group != g and s.jk_done is not None]
        else:
            launch = absorb = [s for s in self.active if s.jk_done is None]
        if launch:
            evs = self._launch(launch)
            for s in launch:
                s.jk_done = evs[s.serial]
        finished = self._absorb_all(absorb) if absorb else []
        # plans unused for a while are freed (a free synchronises the device,
        # so not every tick)
        for key in [k for k, tk in self._plan_tick.items() if tk < self.ticks - PLAN_KEEP_TICKS]:
            self._plans.pop(key).close()
            del self._plan_tick[key]
        self.timing['tick_s'] += time.perf_counter() - t1
        for s in finished:
            s.stream.synchronize()
            self._retire(s)
            self._weight[s.group] -= s.weight
        if finished:
            gone = set(map(id, finished))
            self.active = [s for s in self.active if id(s) not in gone]
        self.trace.append((self.now(), len(self.active), len(self.queue), self.resident))
        return len(self.active)

    def serve(self, requests):
        """Run a stream to completion: each request is submitted once the
        server clock passes its ``arrival``; returns :func:`summary`."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        self.t0 = time.perf_counter()
        while pending or self.busy:
            now = self.now()
            while pending and pending[0].arrival <= now:
                self.submit(pending.popleft())
            if not self.busy:
                time.sleep(max(0., pending[0].arrival - self.now()))
                continue
            self.step()
        self.close()
        return summary(self.done)

    def close(self):
        for g in self._dead_plans:
            g.close()
        self._dead_plans = []
        for g in self._plans.values():
            g.close()
        self._plans.clear()
        if self._pool is not None:
            self._pool.shutdown()
            self._pool = None
        if self._builder is not None:
            self._builder.shutdown()
            self._builder = None
        self._work.release()


def serve_one_at_a_time(requests, conv_tol=1e-10, max_cycle=100, k_precision='fp64'):
    """The plain path: each request solved alone with :class:`scf.RHF` (one
    ``get_jk`` per cycle) in arrival order, waiting for its arrival."""
    t0 = time.perf_counter()
    out = []
    for r in sorted(requests, key=lambda r: r.arrival):
        wait = r.arrival - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        r.t_admit = time.perf_counter() - t0
        mf = RHF(r.mol, conv_tol=conv_tol, max_cycle=max_cycle, k_precision=k_precision)
        r.e_tot = mf.kernel()
        r.cycles, r.converged, r.history = mf.cycles, mf.converged, mf.history
        _dev.torch().cuda.synchronize()
        r.t_done = time.perf_counter() - t0
        out.append(r)
    return summary(out)


def summary(reqs):
    """Throughput and latency of a finished stream."""
    if not reqs:
        return {'n': 0}
    lat = np.array([r.latency for r in reqs])
    t_first = min(r.arrival for r in reqs)
    t_last = max(r.t_done for r in reqs)
    span = max(t_last - t_first, 1e-9)
    return {'n': len(reqs), 'span_s': span, 'solves_per_hr': 3600. * len(reqs) / span,
            'p50_s': float(np.percentile(lat, 50)), 'p99_s': float(np.percentile(lat, 99)),
            'mean_s': float(lat.mean()), 'converged': int(sum(bool(r.converged) for r in reqs)),
            'cycles': int(sum(r.cycles for r in reqs))}
