"""A density-fitting object with the fields the reference's DF path reads.

``DF(mol)`` plays GPU4PySCF's ``gpu4pyscf.df.df.DF`` for this package when
GPU4PySCF is absent: ``build()`` runs the MI355X cderi build
(:func:`electrolites_amd.fastdfcderi.cholesky_eri`) and stores
``_cderi = {device_id: cderi}`` and ``_cderi_idx``; ``get_jk`` is
:func:`electrolites_amd.fastdf.get_jk` (``DF.get_jk`` -> ``df_jk.get_jk``,
gpu4pyscf/df/df.py:105-117 as the reference describes its caller).
"""
from . import _dev
from . import fastdf
from . import fastdfcderi


class DF:
    def __init__(self, mol, auxbasis=None, verbose=0):
        self.mol = mol
        self.auxbasis = auxbasis or getattr(mol, 'auxbasis', None)
        self.verbose = verbose
        self._cderi = None
        self._cderi_idx = None

    @property
    def naux(self):
        return None if self._cderi is None else next(iter(self._cderi.values())).shape[0]

    def build(self, omega=None):
        cderi, idx = fastdfcderi._cholesky_eri(self, omega=omega)
        dev = _dev.torch().cuda.current_device()
        self._cderi = {dev: cderi[0]}
        self._cderi_idx = idx
        return self

    def reset(self):
        self._cderi = None
        self._cderi_idx = None
        self.__dict__.pop('_mdf_layout', None)
        return self

    def get_jk(self, dm, hermi=1, with_j=True, with_k=True, omega=None):
        return fastdf.get_jk(self, dm, hermi, with_j, with_k, omega)

    def get_j(self, dm, hermi=1):
        """J without the cderi: three-centre integrals contracted on the fly
        (:func:`electrolites_amd.fastdfj3c.get_j`; ``df_jk.get_j``)."""
        from . import fastdfj3c
        return fastdfj3c.get_j(self, dm, hermi)

    def cderi(self):
        if self._cderi is None:
            self.build()
        return next(iter(self._cderi.values()))
