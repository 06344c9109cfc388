"""A density-fitting object with the fields the reference's DF path reads.

``DF(mol)`` plays GPU4PySCF's ``gpu4pyscf.df.df.DF`` for this package when
GPU4PySCF is absent: ``build()`` runs the MI355X cderi build
(:func:`electrolites_amd.fastdfcderi.cholesky_eri`) and stores
``_cderi = {device_id: cderi}`` and ``_cderi_idx``; ``get_jk`` is
:func:`electrolites_amd.fastdf.get_jk` (``DF.get_jk`` -> ``df_jk.get_jk``,
gpu4pyscf/df/df.py:105-117 as the reference describes its caller).

``storage='host'`` builds the cderi into page-locked host memory (for a
tensor larger than the device); J/K then keep a prefix of its rows on the
device and stream the rest (:mod:`electrolites_amd.dfstore`).  ``loop`` is
the block iterator GPU4PySCF's other consumers use (``DF.loop``).
"""
from . import _dev
from . import fastdf
from . import fastdfcderi


class DF:
    def __init__(self, mol, auxbasis=None, verbose=0, storage='device'):
        self.mol = mol
        self.storage = storage
        self.auxbasis = auxbasis or getattr(mol, 'auxbasis', None)
        self.verbose = verbose
        self._cderi = None
        self._cderi_idx = None

    @property
    def naux(self):
        return None if self._cderi is None else next(iter(self._cderi.values())).shape[0]

    def build(self, omega=None):
        # tables of a previous tensor go with it (they are also keyed on it)
        for name in ('_mdf_layout', '_mdf_auxc', '_mdf_grad_metric', '_mdf_grad_pairs'):
            self.__dict__.pop(name, None)
        cderi, idx = fastdfcderi._cholesky_eri(self, omega=omega)
        dev = _dev.torch().cuda.current_device()
        self._cderi = {dev: cderi[0]}
        self._cderi_idx = idx
        return self

    def reset(self):
        from . import dfstore
        for c in (self._cderi or {}).values():
            dfstore.release(c)
        self.__dict__.pop('_mdf_resident', None)
        self._cderi = None
        self._cderi_idx = None
        for name in ('_mdf_layout', '_mdf_auxc', '_mdf_grad_metric', '_mdf_grad_pairs'):
            self.__dict__.pop(name, None)
        return self

    def get_jk(self, dm, hermi=1, with_j=True, with_k=True, omega=None):
        return fastdf.get_jk(self, dm, hermi, with_j, with_k, omega)

    def loop(self, blksize=None, unpack=True):
        """Blocks of aux rows as FP64 device tensors: ``(unpacked, packed)``
        with ``unpacked`` the ``(nL, nao, nao)`` symmetric matrices (or None)
        and ``packed`` the ``(nL, npairs)`` rows (``DF.loop``;
        :func:`electrolites_amd.fastdffp32.loop`)."""
        from . import fastdffp32
        return fastdffp32.loop(self, blksize, unpack)

    @property
    def nao(self):
        return self.mol.nao

    def get_j(self, dm, hermi=1):
        """J without the cderi: three-centre integrals contracted on the fly
        (:func:`electrolites_amd.fastdfj3c.get_j`; ``df_jk.get_j``)."""
        from . import fastdfj3c
        return fastdfj3c.get_j(self, dm, hermi)

    def cderi(self):
        if self._cderi is None:
            self.build()
        return next(iter(self._cderi.values()))
