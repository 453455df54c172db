"""Host-side checks (CPU): the C-ABI library loads and exports every symbol
declared in include/mgw.h, and the host logic mirrors the reference."""
import ctypes
import os
import re

import numpy as np
import pytest

from tests.conftest import ROOT
from oracle import gw_oracle as O

HEADER = os.path.join(ROOT, 'include', 'mgw.h')


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    names = re.findall(r'\b(mgw_[a-z0-9_]+)\s*\(', text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from molgw_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 19
    for n in names:
        assert hasattr(lib, n), f'{n} declared in mgw.h but not exported'
    # the Python binding covers exactly the declared functions
    assert sorted(_lib.SIGNATURES) == names


def test_abi_version_and_error_string_without_gpu():
    from molgw_amd import _lib
    lib = _lib.load()
    assert lib.mgw_abi_version() == 1
    assert isinstance(lib.mgw_last_error(), bytes)
    # footprint estimator is pure host arithmetic
    b = lib.mgw_footprint_bytes(492, 1188, 492, 33, 0, 0, 0)
    assert b > 8 * 1188 * 492 * 492


def test_windowed_footprint_without_gpu():
    """mgw_footprint_bytes_nse: the whole range (nse = nstate) is
    mgw_footprint_bytes; a window needs less, never below the resident dense
    RI tensor, and more states never need less."""
    from molgw_amd import _lib
    lib = _lib.load()
    for tech, nw, prec in ((1, 32, 4), (1, 32, 0), (0, 0, 0)):   # imaginary+Pade FP16x3 / FP64, one-shot
        full = lib.mgw_footprint_bytes(768, 2688, 768, 160, tech, nw, prec)
        assert lib.mgw_footprint_bytes_nse(768, 2688, 768, 160, 768, tech, nw, prec) == full
        sizes = [lib.mgw_footprint_bytes_nse(768, 2688, 768, 160, ne, tech, nw, prec) for ne in (1, 16, 64, 768)]
        assert sizes == sorted(sizes) and sizes[1] < full
        assert sizes[0] > 8 * 2688 * 768 * 768


def test_struct_layouts_match_header():
    from molgw_amd._lib import G0W0Params, Ranges
    assert ctypes.sizeof(Ranges) == 32
    assert ctypes.sizeof(G0W0Params) == 56   # static_assert in api.hip


@pytest.mark.parametrize('occ', [
    [2, 2, 2, 0, 0, 0],
    [2, 2, 1, 0, 0],
    [2, 0, 0, 0],
])
@pytest.mark.parametrize('kw', [dict(), dict(ncoreg=1), dict(selfenergy_state_range=1),
                                dict(nvirtualg=4, ncorew=1, nvirtualw=4)])
def test_state_ranges_match_oracle(occ, kw):
    from molgw_amd.gw import selfenergy_set_state_range
    a = selfenergy_set_state_range(np.array(occ, float), **kw)
    b = O.selfenergy_set_state_range(np.array(occ, float)[:, None], **kw)
    for f in ('ncore_G', 'nvirtual_G', 'nhomo_G', 'nlumo_G', 'nsemin', 'nsemax', 'ncore_W', 'nvirtual_W'):
        assert getattr(a, f) == getattr(b, f), f


def test_imaginary_grid_matches_oracle():
    from molgw_amd.gw import imaginary_quad_grid
    for n in (1, 8, 32):
        om, wt = imaginary_quad_grid(n)
        om2, wt2 = O.imaginary_quad_grid(n)
        np.testing.assert_allclose(om, om2, rtol=1e-12)
        np.testing.assert_allclose(wt, wt2, rtol=1e-11)


def test_molecular_shapes():
    from molgw_amd.workloads import alkane_shape, water_cluster_shape
    assert water_cluster_shape(1) == (24, 84, 5)
    assert alkane_shape(8, 'cc-pVTZ') == (492, 1188, 33)
    assert alkane_shape(1) == (34, 112, 5)
    assert alkane_shape(16) == (394, 1372, 65)


def test_synthetic_system_packing_roundtrip():
    from molgw_amd.workloads import synthetic_system, dense_ri
    s = synthetic_system(9, 20, 2, seed=1)
    d = dense_ri(s)
    np.testing.assert_allclose(d, d.transpose(0, 2, 1))
    np.testing.assert_allclose(O.pack_eri_3center(d, s.index_basis), s.eri_3center)
    np.testing.assert_allclose(s.c_matrix.T @ s.c_matrix, np.eye(9), atol=1e-13)
    assert np.all(np.diff(s.energy[:2]) >= 0) and np.all(np.diff(s.energy[2:]) >= 0)
    assert s.energy[1] < s.energy[2]


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, 'molgw_amd')
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(('.py', '.hip', '.h', '.cpp', '.f90')):
                text = open(os.path.join(dirpath, f), errors='ignore').read()
                assert 'import oracle' not in text and 'from oracle' not in text, f


def test_product_does_not_use_torch():
    """The product's device memory comes from libmgw (mgw_device_alloc): a
    second HIP runtime in the process (torch's bundled one) cannot see the GPU
    once libmgw's runtime holds it."""
    paths = [os.path.join(ROOT, 'bench.py'), os.path.join(ROOT, '__graft_entry__.py')]
    for dirpath, _, files in os.walk(os.path.join(ROOT, 'molgw_amd')):
        paths += [os.path.join(dirpath, f) for f in files if f.endswith('.py')]
    for path in paths:
        text = open(path, errors='ignore').read()
        assert 'import torch' not in text, path


def test_fortran_interface_compiles(tmp_path):
    """m_mgw (the iso_c_binding module MOLGW would use) compiles and a program
    using it links against libmgw.so."""
    import subprocess
    from molgw_amd import build as b
    if not os.path.exists(b.FLANG):
        pytest.skip('amdflang not available')
    obj = tmp_path / 'm_mgw.o'
    subprocess.run([b.FLANG, '-c', os.path.join(b.FDIR, 'm_mgw.f90'), '-module-dir', str(tmp_path), '-o', str(obj)],
                   check=True)
    exe = tmp_path / 'drv'
    subprocess.run([b.FLANG, '-I', str(tmp_path), os.path.join(b.FDIR, 'mgw_fdriver.f90'), str(obj), f'-L{b.HERE}',
                    '-lmgw', f'-Wl,-rpath,{b.HERE}', '-o', str(exe)], check=True)
    assert exe.exists()


def test_poisson_stream_and_batch_are_seeded():
    from molgw_amd.serving import poisson_stream, alkane_batch
    a = poisson_stream(200, rate=5.0, seed=3)
    b = poisson_stream(200, rate=5.0, seed=3)
    assert a == b
    t = np.array([x[0] for x in a])
    assert np.all(np.diff(t) > 0)
    # mean inter-arrival 1/rate
    assert abs(np.mean(np.diff(t)) - 0.2) < 0.05
    kinds = {x[1] for x in a}
    assert kinds == {'alkane', 'water'}
    assert all(2 <= x[2] <= 24 for x in a if x[1] == 'alkane')
    assert all(1 <= x[2] <= 32 for x in a if x[1] == 'water')
    n = alkane_batch(64, seed=0)
    assert len(n) == 64 and min(n) >= 1 and max(n) <= 16 and n == alkane_batch(64, seed=0)


def test_entry_scripts_parse():
    """bench.py and __graft_entry__.py are run as scripts on the GPU box:
    they must at least compile."""
    import ast
    for name in ('bench.py', '__graft_entry__.py'):
        with open(os.path.join(ROOT, name)) as f:
            ast.parse(f.read(), name)
