"""The committed generated kernel sources are what their generators write
(a kernel edited by hand, or a generator changed without regenerating,
fails here)."""
import filecmp
import importlib

import pytest


@pytest.mark.parametrize('module', ['gen_int', 'gen_int_ip'])
def test_generated_sources_match_their_generator(module, tmp_path, monkeypatch):
    gen = importlib.import_module(f'electrolites_amd.codegen.{module}')
    committed = gen.OUT
    out = tmp_path / 'gen.hip'
    monkeypatch.setattr(gen, 'OUT', str(out))
    gen.main()
    assert filecmp.cmp(str(out), committed, shallow=False), f'{committed} differs from {module}.main()'
