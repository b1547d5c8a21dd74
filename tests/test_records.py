"""Records: bit-exact round trips, the atomic JSON+npz pair, refusals, hashes,
code stamps and Davidson warm starts (reference: Base/utils/record_io.py,
LR/davidson.py:94-160)."""
import json
import math
import os

import numpy as np
import pytest
import torch

from mbptcode_amd import records as R
from mbptcode_amd.davidson import davidson


def sample_tree():
    rng = np.random.default_rng(0)
    return {
        'floats': [0.1, 1.0 / 3.0, -2.5e-310, 1e308, math.pi, -0.0],
        'ints': [0, -7, 2 ** 62, 10 ** 30],
        'text': 'ω ≤ 1 meV', 'none': None, 'flags': (True, False),
        'tuple': (1, (2.5, 'x'), [3, (4,)]),
        'set': {1, 2, 3}, 'frozen': frozenset({'a', 'b'}),
        'int_keys': {1: 'one', (2, 3): [4.5], 2.5: None},
        'arrays': [rng.normal(size=(3, 4)), rng.normal(size=5).astype(np.float32),
                   np.arange(6, dtype=np.int64).reshape(2, 3), np.array([True, False]),
                   (rng.normal(size=3) + 1j * rng.normal(size=3)), np.zeros((0, 2))],
        'scalars': [np.float32(1.5), np.int64(-3), np.float64(2.0 ** -60), np.bool_(True)],
        '__dunder_key': 1,
    }


def same(a, b):
    """Equality to the bit, types included."""
    if type(a) is not type(b):
        return False
    if isinstance(a, np.ndarray):
        return a.dtype == b.dtype and a.shape == b.shape and a.tobytes() == b.tobytes()
    if isinstance(a, np.generic):
        return a.dtype == b.dtype and a.tobytes() == b.tobytes()
    if isinstance(a, float):
        return a.hex() == b.hex()
    if isinstance(a, (list, tuple)):
        return len(a) == len(b) and all(same(x, y) for x, y in zip(a, b))
    if isinstance(a, dict):
        return a.keys() == b.keys() and all(same(a[k], b[k]) for k in a)
    return a == b


def test_round_trip_is_bit_exact(tmp_path):
    value = sample_tree()
    R.write_record(tmp_path / 'res.json', value)
    back = R.read_record(tmp_path / 'res.json')
    assert same(value, back)


def test_nonfinite_floats_round_trip(tmp_path):
    value = [float('inf'), float('-inf')]
    R.write_record(tmp_path / 'f.json', value)
    assert R.read_record(tmp_path / 'f.json') == value
    R.write_record(tmp_path / 'n.json', float('nan'))
    assert math.isnan(R.read_record(tmp_path / 'n.json'))


def test_on_disk_layout_uses_the_reference_tags(tmp_path):
    R.write_record(tmp_path / 'r.json', {'x': (np.ones(2), {1: 2})})
    held = json.loads((tmp_path / 'r.json').read_text())
    assert set(held) == {'arrays', 'record'}
    assert held['arrays'].startswith('r.') and held['arrays'].endswith('.npz')
    assert len(held['arrays'].split('.')[1]) == 12
    assert held['record'] == {'x': {'__tuple__': [{'__array__': 'a0'},
                                                  {'__dict__': [[1, 2]]}]}}


def test_rewrite_keeps_one_array_file_and_spares_sibling_stems(tmp_path):
    sibling = tmp_path / 'res.x.0123456789ab.npz'
    sibling.write_bytes(b'other stem')
    for i in range(3):
        R.write_record(tmp_path / 'res.json', {'i': i, 'a': np.full(3, i)})
    mine = [p for p in tmp_path.iterdir()
            if p.name.startswith('res.') and p.name.endswith('.npz') and p != sibling]
    assert len(mine) == 1
    assert sibling.exists()
    assert R.read_record(tmp_path / 'res.json')['i'] == 2
    assert not list(tmp_path.glob('*.tmp'))


def test_a_crash_before_the_rename_leaves_the_old_record_readable(tmp_path, monkeypatch):
    path = tmp_path / 'res.json'
    R.write_record(path, {'v': np.arange(3)})
    real_replace = os.replace

    def crash_on_json(src, dst):
        if str(dst).endswith('.json'):
            raise KeyboardInterrupt('crash before the JSON rename')
        return real_replace(src, dst)
    monkeypatch.setattr(R.os, 'replace', crash_on_json)
    with pytest.raises(KeyboardInterrupt):
        R.write_record(path, {'v': np.arange(5)})
    monkeypatch.setattr(R.os, 'replace', real_replace)
    assert np.array_equal(R.read_record(path)['v'], np.arange(3))


class Foreign:
    pass


def test_refusals(tmp_path):
    with pytest.raises(TypeError):
        R.encode({'x': Foreign()})
    with pytest.raises(TypeError):
        R.encode(np.array([object()], dtype=object))
    bad = {'arrays': 'b.000000000000.npz',
           'record': {'__object__': 'os:PathLike', 'state': {}}}
    np.savez(tmp_path / 'b.000000000000.npz')
    (tmp_path / 'b.json').write_text(json.dumps(bad))
    with pytest.raises(TypeError):
        R.read_record(tmp_path / 'b.json')


def test_record_hash_follows_the_bits():
    a = sample_tree()
    b = sample_tree()
    assert R.record_hash(a) == R.record_hash(b)
    b['arrays'][0][0, 0] = np.nextafter(b['arrays'][0][0, 0], 1.0)
    assert R.record_hash(a) != R.record_hash(b)
    c = sample_tree()
    c['tuple'] = list(c['tuple'])
    assert R.record_hash(a) != R.record_hash(c)


def test_code_stamps(tmp_path):
    code = R.code_id()
    assert code and isinstance(code, str)
    assert R.same_code(code, code)
    full = 'a' * 40
    assert R.same_code(full, 'aaaaaaa')
    assert not R.same_code(full, 'aaaaaaa+dirty')
    assert not R.same_code('abc', 'abd')
    (tmp_path / 'SNAPSHOT.txt').write_text('v1.2-export extra words\n')
    assert R.code_id(tmp_path) == 'v1.2-export'
    with pytest.warns(RuntimeWarning):
        assert R.git_commit(tmp_path / 'missing') == 'unknown'


def test_warm_start_round_trips_and_restarts_a_solve(tmp_path):
    """A converged solve's roots, kept, written, read and offered as the
    guess, start the next solve of the same operator converged."""
    rng = np.random.default_rng(5)
    n, k = 400, 4
    d = np.sort(rng.uniform(0.3, 3.0, n))
    Q = rng.normal(size=(n, 6)) * 0.05
    A = np.diag(d) + Q @ Q.T
    At = torch.as_tensor(A)
    diag = torch.as_tensor(np.diag(A).copy())

    def mm(V):
        return At @ V
    x0 = torch.zeros((n, 2 * k), dtype=torch.float64)
    x0[torch.arange(2 * k), torch.arange(2 * k)] = 1.0
    first = davidson(mm, diag, k, x0, tol=1e-8)
    assert first.conv.all() and first.cycles > 2
    ws = R.DavidsonWarmStart()
    ws.keep('singlet', 20, 20, first.e, first.X.numpy())
    R.write_record(tmp_path / 'warm.json', ws.to_record())
    back = R.DavidsonWarmStart.from_record(R.read_record(tmp_path / 'warm.json'))
    assert back.guess('singlet', 20, 21) is None and back.guess('triplet', 20, 20) is None
    X, Y = back.guess('singlet', 20, 20)
    assert np.array_equal(X, first.X.numpy()) and not Y.any()
    second = davidson(mm, diag, k, torch.as_tensor(X), tol=1e-8)
    assert second.conv.all() and second.cycles == 1
    assert np.abs(second.e - first.e).max() < 1e-10
