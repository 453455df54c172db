"""ENERGY_QP and molgw.yaml GW sections in the reference's layouts
(src/m_io.f90:3645-3733, src/m_hamiltonian_tools.f90:532-560,
src/m_selfenergy_tools.f90:501-553)."""
import numpy as np
import yaml

from molgw_amd import formats
from molgw_amd.gw import HA_EV


def test_energy_qp_round_trip_and_status(tmp_path):
    e = np.array([-20.55512345678901, -1.3, 0.25, 1.75])
    formats.write_energy_qp(str(tmp_path / 'ENERGY_QP'), e)
    got, status = formats.read_energy_qp(4, 1, str(tmp_path))
    assert status == 0 and np.array_equal(got[:, 0], e)          # 17 significant digits: exact
    assert formats.read_energy_qp(5, 1, str(tmp_path)) == (None, 2)
    assert formats.read_energy_qp(4, 2, str(tmp_path)) == (None, 2)
    assert formats.read_energy_qp(4, 1, str(tmp_path / 'nowhere')) == (None, 1)
    e2 = np.stack([e, e + 0.1], axis=1)
    formats.write_energy_qp(str(tmp_path / 'ENERGY_QP'), e2)
    got, status = formats.read_energy_qp(4, 2, str(tmp_path))
    assert status == 0 and np.array_equal(got, e2)


def test_energy_qp_reads_fortran_list_directed_and_scissor(tmp_path):
    # list-directed output of another compiler (D exponent), lower-case name
    (tmp_path / 'energy_qp').write_text('           1\n           2\n           1  -0.5D+00\n           2   1.25D-01\n')
    got, status = formats.read_energy_qp(2, 1, str(tmp_path))
    assert status == 0 and np.array_equal(got[:, 0], [-0.5, 0.125])
    (tmp_path / 'energy_qp').write_text('1\n2\n-1 0.1\n')
    assert formats.read_energy_qp(2, 1, str(tmp_path)) == (None, -1)


def test_es18_8_matches_the_fortran_edit_descriptor():
    assert formats._es18_8(-20.5551234) == '   -2.05551234E+01'
    assert formats._es18_8(1.0) == '    1.00000000E+00'
    assert formats._es18_8(0.0) == '    0.00000000E+00'
    assert formats._es18_8(9.999999999e-3) == '    1.00000000E-02'
    assert len(formats._es18_8(-1.0e-5)) == 18


def test_yaml_sections_parse_as_molgw_py_reads_them():
    nstate, nsemin, nsemax = 10, 3, 8
    rng = np.random.default_rng(0)
    e0 = np.sort(rng.uniform(-1.0, 1.0, nstate))
    exc = rng.uniform(-0.2, 0.0, nstate)
    eqp = e0 + exc + rng.uniform(-0.05, 0.05, nstate)
    zz = rng.uniform(0.7, 0.95, nstate)
    text = (formats.yaml_qp_selfenergy('GW', e0, exc, eqp, zz, nsemin, nsemax)
            + formats.yaml_energies('GW energies', eqp, nsemin, nsemax))
    calc = yaml.safe_load(text)
    ch = calc['gw energies']['spin channel 1']
    assert sorted(ch) == list(range(nsemin, nsemax + 1))        # integer state keys
    for p in range(nsemin, nsemax + 1):
        assert abs(float(ch[p]) - eqp[p - 1] * HA_EV) <= 1e-8 * abs(eqp[p - 1] * HA_EV)
    se = calc['gw selfenergy']
    assert se['correlation']['unit'] == 'eV'
    sig = se['correlation']['spin channel 1'][5]
    assert abs(sig - (eqp[4] - e0[4] - exc[4]) * HA_EV) < 1e-6
    assert abs(se['renormalization factor']['spin channel 1'][nsemax] - zz[nsemax - 1]) < 1e-8
    # the line layout of the writer: 8 blanks, index left-adjusted in 6, ':', blank, ES18.8
    line = [ln for ln in text.splitlines() if ln.startswith('        3     :')][0]
    assert line == '        3     : ' + formats._es18_8(eqp[2] * HA_EV)


def test_screened_coulomb_round_trip_and_record_layout(tmp_path):
    from molgw_amd.formats import write_screened_coulomb, read_screened_coulomb
    rng = np.random.default_rng(3)
    pole = np.sort(rng.uniform(0.3, 3.0, 7))
    w_s = np.asfortranarray(rng.standard_normal((11, 7)))
    path = str(tmp_path / 'SCREENED_COULOMB')
    write_screened_coulomb(path, pole, w_s, 'GW')
    name, p2, w2 = read_screened_coulomb(path)
    assert name == 'GW'
    np.testing.assert_array_equal(p2, pole)
    np.testing.assert_array_equal(w2, w_s)
    # sequential unformatted layout: 4 + 7 records, each framed by its length
    raw = open(path, 'rb').read()
    assert len(raw) == (100 + 8) + 2 * (4 + 8) + (8 * 7 + 8) + 7 * (8 * 11 + 8)
    assert np.frombuffer(raw[:4], '<i4')[0] == 100 and raw[4:6] == b'GW'
    assert np.frombuffer(raw[108:112], '<i4')[0] == 4 and np.frombuffer(raw[112:116], '<i4')[0] == 11


def test_dump_gw_ingredients_layout(tmp_path):
    from molgw_amd.formats import dump_gw_ingredients
    rng = np.random.default_rng(4)
    ng, naux, npole = 5, 6, 4
    w = [np.asfortranarray(rng.standard_normal((npole, ng, ng))) for _ in range(2)]
    v = [np.asfortranarray(rng.standard_normal((naux, ng, ng))) for _ in range(2)]
    e = np.linspace(-1, 1, ng)
    pole = np.linspace(0.5, 2.0, npole)
    d = str(tmp_path / 'ingr')
    dump_gw_ingredients(d, e, pole, w, v)
    vals = np.loadtxt(tmp_path / 'ingr' / 'energies.dat')
    assert vals[0] == ng and np.array_equal(vals[1:], e)
    vals = np.loadtxt(tmp_path / 'ingr' / 'omegas.dat')
    assert vals[0] == npole and np.array_equal(vals[1:], pole)
    assert list(np.loadtxt(tmp_path / 'ingr' / 'v.dat')) == [naux, ng, ng]
    wb = np.fromfile(tmp_path / 'ingr' / 'w.bin', dtype='<f8')
    # per spin, per q: wcoeff(npole, ng) column-major
    np.testing.assert_array_equal(wb.reshape(2, ng, ng, npole)[1, 3, :, :].T, w[1][:, :, 3])
    vb = np.fromfile(tmp_path / 'ingr' / 'v.bin', dtype='<f8')
    np.testing.assert_array_equal(vb.reshape(2, ng, ng, naux)[0, 2, :, :].T, v[0][:, :, 2])
