"""Records that survive a relaunch: a JSON tree beside an .npz of its arrays
(the reference's record format, Base/utils/record_io.py:1-265).

A record holds plain values -- None, bools, numbers, strings, lists, tuples,
sets, dicts with any hashable keys, NumPy arrays and scalars -- and the
dataclasses and plain objects of this package, and reading it back gives the
same values to the bit: floats through JSON's shortest round-trip repr, arrays
through the .npz as stored, tuples, sets and non-string keys tagged.

On disk, `<stem>.json` is {"arrays": "<stem>.<token>.npz", "record": tree}, the
tree tagging what JSON cannot say with '__array__', '__scalar__', '__tuple__',
'__set__', '__frozenset__', '__dict__', '__dataclass__' and '__object__' (the
reference's tags, kept so that either code reads the other's records).

`write_record` is atomic as a pair (:41-62): the arrays go to a file named by a
fresh 12-hex-digit token, then the JSON naming it replaces the old JSON in one
rename, and only then are this stem's older token files removed, so a crash
anywhere leaves a JSON whose arrays exist.  Classes outside this package are
refused when written and when read.
"""
import dataclasses
import hashlib
import importlib
import json
import os
import pathlib
import re
import subprocess
import uuid
import warnings

import numpy as np

from .constants import GIT_PROVENANCE_TIMEOUT

#: The package whose classes a record may hold.
RECORD_PACKAGE = 'mbptcode_amd.'
#: The checkout this file belongs to (mbptcode_amd/records.py).
CHECKOUT = pathlib.Path(__file__).resolve().parents[1]

_TOKEN = re.compile(r'[0-9a-f]{12}')


def write_record(path, value):
    """Write `value` to the JSON `path` and its arrays beside it, atomically
    (record_io.py:41-62)."""
    path = pathlib.Path(path)
    tree, arrays = encode(value)
    token = uuid.uuid4().hex[:12]
    npz = path.with_name(f'{path.stem}.{token}.npz')
    part = npz.with_name(npz.name + '.tmp')
    with open(part, 'wb') as fh:
        np.savez(fh, **arrays)
    os.replace(part, npz)
    part = path.with_name(path.name + '.tmp')
    part.write_text(json.dumps({'arrays': npz.name, 'record': tree}))
    os.replace(part, path)
    # only this stem's own token files: a sibling stem '<stem>.x' keeps its
    for old in path.parent.glob(f'{path.stem}.*.npz'):
        mid = old.name[len(path.stem) + 1:-len('.npz')]
        if old.name != npz.name and _TOKEN.fullmatch(mid):
            old.unlink()


def read_record(path):
    """The value `write_record` wrote to `path` (record_io.py:65-71)."""
    path = pathlib.Path(path)
    held = json.loads(path.read_text())
    with np.load(path.with_name(held['arrays']), allow_pickle=False) as npz:
        arrays = {name: npz[name] for name in npz.files}
    return decode(held['record'], arrays)


def record_hash(value):
    """sha256 over the JSON tree (keys sorted) and each array's dtype, shape
    and bytes in key order: values that read back alike hash alike
    (record_io.py:74-83)."""
    tree, arrays = encode(value)
    digest = hashlib.sha256(json.dumps(tree, sort_keys=True).encode())
    for name in sorted(arrays):
        a = np.ascontiguousarray(arrays[name])
        digest.update(f'{name}:{a.dtype.str}:{a.shape}'.encode())
        digest.update(a.tobytes())
    return digest.hexdigest()


def _git(root, *args, binary=False):
    return subprocess.run(['git', '-C', str(root), *args], capture_output=True,
                          text=not binary, timeout=GIT_PROVENANCE_TIMEOUT,
                          check=True).stdout


def git_commit(root=None):
    """Short commit of the checkout `root` (default: this one), '+dirty' when
    its tree differs; 'unknown' with a RuntimeWarning where git cannot answer
    (record_io.py:86-108)."""
    root = CHECKOUT if root is None else root
    try:
        head = _git(root, 'rev-parse', '--short', 'HEAD').strip()
        dirty = _git(root, 'status', '--porcelain').strip()
    except (OSError, subprocess.SubprocessError) as exc:
        warnings.warn(f'no commit for the record: git at {root} failed ({exc}); '
                      'stamped unknown', RuntimeWarning)
        return 'unknown'
    return head + ('+dirty' if dirty else '')


def code_id(root=None):
    """The code a record was made with: SNAPSHOT.txt's first word in an
    exported tree, else `git_commit`, with ':' and 12 hex digits of the sha256
    of `git diff HEAD` on an edited tree (record_io.py:111-128)."""
    root = CHECKOUT if root is None else pathlib.Path(root)
    snap = pathlib.Path(root) / 'SNAPSHOT.txt'
    if snap.is_file():
        return snap.read_text().split()[0]
    code = git_commit(root)
    if not code.endswith('+dirty'):
        return code
    diff = _git(root, 'diff', 'HEAD', '--binary', binary=True)
    return f'{code}:{hashlib.sha256(diff).hexdigest()[:12]}'


def same_code(recorded, current):
    """Whether a record's code stamp names this run's code: equal, or a full
    40-digit commit that a clean short `current` abbreviates
    (record_io.py:131-140)."""
    if recorded == current:
        return True
    return (isinstance(recorded, str) and len(recorded) == 40 and '+' not in current
            and current != 'unknown' and recorded.startswith(current))


# ---------------------------------------------------------------------------
# the tree (record_io.py:143-265)
# ---------------------------------------------------------------------------

def encode(value):
    """(JSON tree, {name: array}) of `value`."""
    arrays = {}
    return _enc(value, arrays, 'record'), arrays


def decode(tree, arrays):
    """The value `encode` produced (tree, arrays) from."""
    return _dec(tree, arrays)


def _qualified(obj):
    t = type(obj)
    return f'{t.__module__}:{t.__qualname__}'


def _enc(v, arrays, at):
    if v is None or isinstance(v, (bool, str)):
        return v
    if isinstance(v, (int, float)) and not isinstance(v, np.generic):
        return v
    if isinstance(v, np.ndarray):
        if v.dtype.hasobject:
            raise TypeError(f'{at}: an object array has no record form')
        name = f'a{len(arrays)}'
        arrays[name] = v
        return {'__array__': name}
    if isinstance(v, np.generic):
        return {'__scalar__': v.dtype.str,
                'value': _enc(np.asarray(v).reshape(1), arrays, at)}
    if isinstance(v, tuple):
        return {'__tuple__': [_enc(x, arrays, f'{at}[{i}]') for i, x in enumerate(v)]}
    if isinstance(v, list):
        return [_enc(x, arrays, f'{at}[{i}]') for i, x in enumerate(v)]
    if isinstance(v, (set, frozenset)):
        tag = '__set__' if isinstance(v, set) else '__frozenset__'
        return {tag: [_enc(x, arrays, f'{at}{{}}') for x in v]}
    if isinstance(v, dict):
        if all(isinstance(key, str) and not key.startswith('__') for key in v):
            return {key: _enc(x, arrays, f'{at}.{key}') for key, x in v.items()}
        return {'__dict__': [[_enc(key, arrays, f'{at} key'), _enc(x, arrays, f'{at}[{key!r}]')]
                             for key, x in v.items()]}
    name = _qualified(v)
    if not name.startswith(RECORD_PACKAGE):
        raise TypeError(f'{at}: a {name} has no record form; records hold plain '
                        f'values and the classes of {RECORD_PACKAGE}')
    if dataclasses.is_dataclass(v):
        return {'__dataclass__': name,
                'fields': {f.name: _enc(getattr(v, f.name), arrays, f'{at}.{f.name}')
                           for f in dataclasses.fields(v)}}
    return {'__object__': name, 'state': _enc(dict(vars(v)), arrays, f'{at}.__dict__')}


def record_class(name):
    """The class a record names, from this package only."""
    if not name.startswith(RECORD_PACKAGE):
        raise TypeError(f'a record names {name}, outside {RECORD_PACKAGE}')
    module, qualname = name.split(':')
    out = importlib.import_module(module)
    for part in qualname.split('.'):
        out = getattr(out, part)
    return out


def _dec(t, arrays):
    if isinstance(t, list):
        return [_dec(x, arrays) for x in t]
    if not isinstance(t, dict):
        return t
    if '__array__' in t:
        return arrays[t['__array__']]
    if '__scalar__' in t:
        return _dec(t['value'], arrays).astype(t['__scalar__'])[0]
    if '__tuple__' in t:
        return tuple(_dec(x, arrays) for x in t['__tuple__'])
    if '__set__' in t:
        return {_dec(x, arrays) for x in t['__set__']}
    if '__frozenset__' in t:
        return frozenset(_dec(x, arrays) for x in t['__frozenset__'])
    if '__dict__' in t:
        return {_dec(key, arrays): _dec(x, arrays) for key, x in t['__dict__']}
    if '__dataclass__' in t:
        cls = record_class(t['__dataclass__'])
        values = {key: _dec(x, arrays) for key, x in t['fields'].items()}
        init = {f.name for f in dataclasses.fields(cls) if f.init}
        out = cls(**{key: x for key, x in values.items() if key in init})
        for key, x in values.items():
            if key not in init:
                object.__setattr__(out, key, x)
        return out
    if '__object__' in t:
        cls = record_class(t['__object__'])
        out = cls.__new__(cls)
        out.__dict__.update(_dec(t['state'], arrays))
        return out
    return {key: _dec(x, arrays) for key, x in t.items()}


# ---------------------------------------------------------------------------
# Davidson warm starts (LR/davidson.py:94-160)
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class DavidsonWarmStart:
    """The last converged roots of a sequence of solves on the same pair
    space, by spin: the next solve of that spin starts from them
    (`run_davidson(guess=...)`).  The reference's holder also carries the
    orbitals and geometry of a pyscf molecule to rotate the vectors into a
    moved geometry; the configs here have no molecule, so a held guess is
    offered only to a pair space of the same shape (n_occ, n_vir)."""

    held: dict = dataclasses.field(default_factory=dict)

    def keep(self, spin, nocc, nvir, omega, X, Y=None):
        X = np.asarray(X, float)
        self.held[spin] = (int(nocc), int(nvir), np.asarray(omega, float), X,
                           np.zeros_like(X) if Y is None else np.asarray(Y, float))

    def guess(self, spin, nocc, nvir):
        """(X, Y) for `run_davidson(guess=...)`, or None when nothing of this
        spin and pair-space shape is held."""
        h = self.held.get(spin)
        if h is None or (h[0], h[1]) != (int(nocc), int(nvir)):
            return None
        return h[3], h[4]

    def to_record(self):
        return {'held': dict(self.held)}

    @classmethod
    def from_record(cls, record):
        out = cls()
        for spin, (nocc, nvir, omega, X, Y) in record['held'].items():
            out.keep(spin, nocc, nvir, omega, X, Y)
        return out
