"""The Numerics surface (src/Base/numerics/numerics.py:214-286, :610-702) and
records of it, against the reference's own record fixture
tests/data/pre_r3/pre_r3_record.json (copied to tests/golden/pre_r3/): a
record the reference wrote of its default Numerics, a NoEnvironment and a
FirstPoint with three arrays.  CPU only."""
import json
import os
import shutil

import numpy as np
import pytest

from mbptcode_amd import records
from mbptcode_amd.numerics import BSENumerics, Numerics, QuadratureNumerics
from mbptcode_amd.records import ForeignRecord

FIXTURE = os.path.join(os.path.dirname(__file__), 'golden', 'pre_r3', 'pre_r3_record.json')
OTHER_SECTIONS = {'isdf', 'scf', 'span', 'qp', 'solvent', 'walk', 'layout'}


def test_the_reference_record_is_refused_without_hold():
    with pytest.raises(TypeError, match="src.Base.numerics:IsdfNumerics.*foreign='hold'"):
        records.read_record(FIXTURE)


def test_the_reference_record_reads_with_its_numerics_mapped():
    value = records.read_record(FIXTURE, foreign='hold')
    num = value['numerics']
    assert type(num) is Numerics
    # the fixture holds the reference's default Numerics(): the modelled
    # sections must equal this package's defaults field for field
    assert num.bse == BSENumerics() and num.quadrature == QuadratureNumerics()
    assert num.bse.conv_tol == 1e-8 and num.bse.max_cycle == 150 and num.bse.nroots == 5
    assert num.quadrature.ntau_gw == 24 and num.quadrature.npade == 16
    assert num.parent is None and num.change == () and num.reason is None
    held = dict(num.held)
    assert set(held) == OTHER_SECTIONS
    for name, sec in held.items():
        assert isinstance(sec, ForeignRecord) and sec.kind == 'dataclass'
        assert sec.name.startswith('src.Base.numerics:')
    assert held['isdf'].regularization == 4e-7 and held['scf'].max_cycle == 200
    gas = value['gas']
    assert gas == ForeignRecord('src.Base.environment:NoEnvironment', 'object', {})
    point = value['first_point']
    assert point.name == 'src.gradients.excited_state:FirstPoint'
    np.testing.assert_array_equal(point.coords, np.arange(6.0).reshape(2, 3))
    assert point.result[1] == -76.25 and point.conventions == 'abc' and point.spin == 'singlet'
    assert isinstance(point.result, tuple) and point.result[2] == {'flags': ()}


def test_the_reference_record_round_trips(tmp_path):
    value = records.read_record(FIXTURE, foreign='hold')
    out = tmp_path / 'again.json'
    records.write_record(out, value)
    back = records.read_record(out, foreign='hold')
    assert back['numerics'] == value['numerics']
    assert back['gas'] == value['gas'] and back['first_point'] == value['first_point']
    # the foreign parts are written back tag for tag
    ref = json.load(open(FIXTURE))['record']
    new = json.loads(out.read_text())['record']
    assert new['gas'] == ref['gas']
    assert new['first_point']['__dataclass__'] == ref['first_point']['__dataclass__']
    fp_new, fp_ref = new['first_point']['fields'], ref['first_point']['fields']
    assert {k: v for k, v in fp_new.items() if k not in ('coords', 'charges', 'result')} == \
        {k: v for k, v in fp_ref.items() if k not in ('coords', 'charges', 'result')}
    pairs = new['numerics']['fields']['held']['__tuple__']
    held = {p['__tuple__'][0]: p['__tuple__'][1] for p in pairs}
    assert set(held) == OTHER_SECTIONS
    for name in OTHER_SECTIONS:
        assert held[name] == ref['numerics']['fields'][name], name
    # a second write gives the same settings identity
    assert back['numerics'].identity() == value['numerics'].identity()


def test_identity_and_derive():
    a, b = Numerics(), Numerics(bse={'conv_tol': 1e-8})
    assert a == b and a.identity() == b.identity()
    c = a.derive('tighter residual for a force', bse={'conv_tol': 1e-10})
    assert c.bse.conv_tol == 1e-10 and c.parent == a.identity()
    assert c.change == (('bse.conv_tol', 1e-8, 1e-10),)
    assert c.reason == 'tighter residual for a force'
    assert c.identity() != a.identity()
    # provenance is not part of the identity
    assert Numerics(bse=c.bse).identity() == c.identity()
    with pytest.raises(ValueError, match='reason'):
        a.derive('', bse={'conv_tol': 1e-9})
    with pytest.raises(ValueError, match='changes nothing'):
        a.derive('why', bse={'conv_tol': 1e-8})
    with pytest.raises(ValueError, match='no modelled section'):
        a.derive('why', isdf={'level': 'G2'})
    with pytest.raises(ValueError, match='no field'):
        a.derive('why', bse={'tolerance': 1e-9})


def test_section_validation():
    for bad in ({'solver': 'lapack'}, {'adjoint': 'implicit'}, {'preconditioner': 'ilu'},
                {'track': 'energy'}):
        with pytest.raises(ValueError):
            BSENumerics(**bad)
    BSENumerics(preconditioner='learned', track='overlap', solver='davidson')
    with pytest.raises(TypeError):
        Numerics(bse=QuadratureNumerics())


def test_to_record_and_the_reference_values_layout():
    n = Numerics(quadrature={'ntau_w': 16}).derive('finer W', quadrature={'ntau_w': 20})
    assert Numerics.from_record(n.to_record()) == n
    # the reference's to_record layout: every section as a plain mapping
    ref_layout = json.load(open(FIXTURE))['record']['numerics']['fields']
    plain = {k: v['fields'] for k, v in ref_layout.items() if isinstance(v, dict)
             and '__dataclass__' in v}
    plain.update(parent=None, change=[], reason=None)
    m = Numerics.from_record(plain)
    assert m.bse == BSENumerics() and set(dict(m.held)) == OTHER_SECTIONS


def test_records_of_numerics_round_trip(tmp_path):
    n = Numerics(bse={'nroots': 12, 'preconditioner': 'bare'}).derive(
        'more roots', bse={'nroots': 20})
    p = tmp_path / 'n.json'
    records.write_record(p, {'numerics': n, 'w': np.arange(3.0)})
    back = records.read_record(p)
    assert back['numerics'] == n and back['numerics'].identity() == n.identity()
    text = p.read_text()
    assert 'mbptcode_amd.numerics:Numerics' in text and 'mbptcode_amd.numerics:BSENumerics' in text


def test_solvers_read_the_bse_section():
    """`solve_casida_davidson` takes its settings from numerics.bse (as
    numerics.py:241-266 says every Davidson entry does); its defaults are
    the section's."""
    import inspect
    from mbptcode_amd.davidson import solve_casida_davidson
    d = {k: p.default for k, p in inspect.signature(solve_casida_davidson).parameters.items()}
    s = BSENumerics()
    assert (d['nroots'], d['conv_tol'], d['max_cycle'], d['guess_factor'], d['preconditioner']) \
        == (s.nroots, s.conv_tol, s.max_cycle, s.guess_factor, s.preconditioner)


def test_fixture_is_the_reference_file():
    """The fixture is the reference's own file, unedited (its old module
    names are what the mapping is tested on)."""
    text = open(FIXTURE).read()
    assert 'src.Base.numerics:Numerics' in text and 'src.gradients.excited_state:FirstPoint' in text
