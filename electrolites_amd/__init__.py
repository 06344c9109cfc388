"""MI355X-native density-fitted two-electron path (the Electrolites DF drop-in).

When GPU4PySCF is installed, ``patch()`` installs this package's functions
over the two entry points the reference replaces (electrolites/__init__.py
:59-150, ``_TARGETS`` entries 'fastdf' and 'fastdfcderi'):

    gpu4pyscf.df.df_jk.get_jk        -> electrolites_amd.fastdf.get_jk
    gpu4pyscf.df.df._cholesky_eri    -> electrolites_amd.fastdfcderi._cholesky_eri
    gpu4pyscf.df.df_jk.get_j         -> electrolites_amd.fastdfj3c.get_j
    gpu4pyscf.df.df.DF.loop          -> electrolites_amd.fastdffp32.patched_loop
    gpu4pyscf.df.grad.{rhf,rks,uhf,uks}.Gradients.jk_energy_per_atom
                                     -> electrolites_amd.fastdfgrad.patched_jk_energy_per_atom

('fastdffp32', as the reference's entry of that name (:164), gives the
block iterator the tensors this build writes: FP32 storage widened, host
residency streamed.)
('fastdfgrad', as the reference's entry of that name (:137-143), replaces the
DF gradient's two-electron method on the four gradient classes, each of
which binds its own copy of the function.)
(the reference's 'fastdfj3c' entry patches the two Int3c2eOpt contractions
under ``df_jk.get_j``, :127-132; this build replaces the whole J-without-a-
cderi function, whose contractions are its own kernels)

Targets are checked before anything is assigned (``check_targets``, as
electrolites/__init__.py:302-322), the originals are kept for fall-through
(``fastdf._orig_get_jk``), and there is no unpatch (:19-22).

Without GPU4PySCF the same functions run on this package's own objects::

    from electrolites_amd.gto import water_cluster
    from electrolites_amd.df import DF
    mol = water_cluster(64, seed=0)
    df = DF(mol).build()                 # cderi on the GPU
    vj, vk = df.get_jk(dm)               # FP64 J/K on the GPU
"""
import importlib

__version__ = '0.1.0'

#: module -> ((target module, attribute), ...)
_TARGETS = {
    'fastdf': (('gpu4pyscf.df.df_jk', 'get_jk'),),
    'fastdfcderi': (('gpu4pyscf.df.df', '_cholesky_eri'),),
    'fastdfj3c': (('gpu4pyscf.df.df_jk', 'get_j'),),
    'fastdffp32': (('gpu4pyscf.df.df', 'DF.loop'),),
    'fastdfgrad': tuple((f'gpu4pyscf.df.grad.{k}', 'Gradients.jk_energy_per_atom')
                        for k in ('rhf', 'rks', 'uhf', 'uks')),
}
MODULES = ('fastdf', 'fastdfcderi', 'fastdfj3c', 'fastdffp32', 'fastdfgrad')
_patched = []


def check_targets(names=None):
    """``{module: None}`` when every target exists, else ``{module: reason}``."""
    out = {}
    for name in (names or MODULES):
        out[name] = None
        for mod_path, attr in _TARGETS[name]:
            try:
                obj = importlib.import_module(mod_path)
                for part in attr.split('.'):
                    obj = getattr(obj, part)
            except (ImportError, AttributeError) as exc:
                out[name] = f'cannot resolve {mod_path}.{attr}: {exc}'
                break
    return out


def _install(name):
    if name == 'fastdf':
        from . import fastdf
        mod = importlib.import_module('gpu4pyscf.df.df_jk')
        if fastdf._orig_get_jk is None:
            fastdf._orig_get_jk = mod.get_jk
        mod.get_jk = fastdf.get_jk
        # DF.get_jk looks the function up through the module at call time
        # (gpu4pyscf/df/df.py:117); rebinding the attribute covers it
    elif name == 'fastdfcderi':
        from . import fastdfcderi
        mod = importlib.import_module('gpu4pyscf.df.df')
        if getattr(fastdfcderi, '_orig_cholesky_eri', None) is None:
            fastdfcderi._orig_cholesky_eri = mod._cholesky_eri
        mod._cholesky_eri = fastdfcderi._cholesky_eri
    elif name == 'fastdfj3c':
        from . import fastdfj3c
        mod = importlib.import_module('gpu4pyscf.df.df_jk')
        if getattr(fastdfj3c, '_orig_get_j', None) is None:
            fastdfj3c._orig_get_j = mod.get_j
        mod.get_j = fastdfj3c.get_j
    elif name == 'fastdffp32':
        from . import fastdffp32
        mod = importlib.import_module('gpu4pyscf.df.df')
        if fastdffp32._orig_loop is None:
            fastdffp32._orig_loop = mod.DF.loop
        mod.DF.loop = fastdffp32.patched_loop
    elif name == 'fastdfgrad':
        from . import fastdfgrad
        for mod_path in [m for m, _ in _TARGETS['fastdfgrad']]:
            cls = importlib.import_module(mod_path).Gradients
            # each class binds the function object itself, so each is patched
            if cls not in fastdfgrad._orig_methods:
                fastdfgrad._orig_methods[cls] = cls.jk_energy_per_atom
            cls.jk_energy_per_atom = fastdfgrad.patched_jk_energy_per_atom


def patch(*names, strict=True):
    """Patch the named modules (default: all).  ``strict`` raises when a
    target is missing; otherwise missing ones are skipped."""
    names = [n for g in (names or (MODULES,)) for n in ((g,) if isinstance(g, str) else g)]
    unknown = [n for n in names if n not in _TARGETS]
    if unknown:
        raise ValueError(f'patch: {unknown} not among the patchable modules {list(MODULES)}')
    missing = {k: v for k, v in check_targets(names).items() if v}
    if missing:
        if strict:
            raise RuntimeError(f'patch(strict=True): the installed GPU4PySCF lacks targets '
                               f'of {sorted(missing)}: {missing}')
        names = [n for n in names if n not in missing]
    for name in names:
        _install(name)
        if name not in _patched:
            _patched.append(name)
    return tuple(_patched)


def patched():
    return tuple(_patched)


__all__ = ['patch', 'patched', 'check_targets', 'MODULES', '__version__']
