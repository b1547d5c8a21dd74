"""Results and warm starts kept on disk: a JSON tree plus one .npz of arrays.

The on-disk form is the reference's record format (Base/utils/record_io.py:
41-265), so that records move between the two codes:

* ``<stem>.json`` holds ``{"arrays": "<stem>.<token>.npz", "record": tree}``,
  the token twelve hex digits, fresh per write;
* in the tree, JSON's own values stand for themselves, and a dict whose keys
  are all strings not starting with ``__`` is a JSON object; everything else
  is a one-key-tagged object: ``__array__`` (name of an .npz entry),
  ``__scalar__`` (dtype string, with ``value``: a one-element array),
  ``__tuple__``, ``__set__``, ``__frozenset__`` (lists), ``__dict__`` (a
  list of [key, value] pairs), ``__dataclass__`` (class name, with
  ``fields``) and ``__object__`` (class name, with ``state``, its __dict__).
  Class names are ``module:qualname``.

Reading rebuilds classes of this package, and the reference's numerical
settings (``src.Base.numerics``) as the `numerics` module's sections.  Any
other class is refused -- nothing named by a file is imported from outside
the package -- unless the caller asks for ``foreign='hold'``: the class then
comes back as an inert `ForeignRecord` (its name and its decoded fields or
state), which writes back to the same tags.

Writes are atomic as a pair: the arrays land under a new token first, the
JSON naming them replaces the old JSON in one rename, and only then are this
stem's other token files removed.  A crash at any point leaves a JSON whose
arrays exist.
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

#: Class names this package writes and rebuilds start with this.
RECORD_PACKAGE = 'mbptcode_amd.'
#: The repository this file sits in (mbptcode_amd/records.py).
CHECKOUT = pathlib.Path(__file__).resolve().parents[1]

_TOKEN_FILE = re.compile(r'(?P<stem>.+)\.(?P<token>[0-9a-f]{12})\.npz')
_TAGS = ('__array__', '__scalar__', '__tuple__', '__set__', '__frozenset__', '__dict__',
         '__dataclass__', '__object__')


class ForeignRecord:
    """A class from outside this package, held as data: `name`
    ('module:qualname'), `kind` ('dataclass' or 'object') and `content`
    (its fields or its __dict__, decoded).  Nothing of the class is
    imported or run; encoding gives back the tags it was read from."""

    __slots__ = ('name', 'kind', 'content')

    def __init__(self, name, kind, content):
        if kind not in ('dataclass', 'object'):
            raise ValueError(f'kind {kind!r}: dataclass or object')
        object.__setattr__(self, 'name', str(name))
        object.__setattr__(self, 'kind', kind)
        object.__setattr__(self, 'content', dict(content))

    def __setattr__(self, key, value):
        raise AttributeError('a ForeignRecord is read-only')

    def __getattr__(self, key):
        try:
            return object.__getattribute__(self, 'content')[key]
        except KeyError:
            raise AttributeError(key) from None

    def __eq__(self, other):
        return (isinstance(other, ForeignRecord) and
                (self.name, self.kind) == (other.name, other.kind) and
                _same_value(self.content, other.content))

    __hash__ = None

    def __repr__(self):
        return f'ForeignRecord({self.name!r}, {self.kind!r}, {sorted(self.content)})'


def _same_value(a, b):
    """Equality that compares arrays elementwise (dtype and shape included)."""
    if isinstance(a, np.ndarray) or isinstance(b, np.ndarray):
        return (isinstance(a, np.ndarray) and isinstance(b, np.ndarray) and
                a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a, b))
    if isinstance(a, dict) and isinstance(b, dict):
        return a.keys() == b.keys() and all(_same_value(a[k], b[k]) for k in a)
    if isinstance(a, (list, tuple)) and isinstance(b, (list, tuple)):
        return (type(a) is type(b) and len(a) == len(b) and
                all(_same_value(x, y) for x, y in zip(a, b)))
    return a == b


# ---------------------------------------------------------------------------
# class names
# ---------------------------------------------------------------------------

def class_name(obj):
    """'module:qualname' of obj's class."""
    cls = type(obj)
    return f'{cls.__module__}:{cls.__qualname__}'


def _numerics_builder(name):
    """For a reference numerics class name (its old module path or the
    current one), the function that builds this package's counterpart from
    the record's decoded fields; else None."""
    module, _, qualname = name.partition(':')
    if module not in ('src.Base.numerics', 'src.Base.numerics.numerics'):
        return None
    from . import numerics
    return numerics.REFERENCE_BUILDERS.get(qualname)


def resolve_class(name, foreign='refuse'):
    """What a record's class name reads as: a class of this package, the
    builder of a numerics section the reference names (a function of the
    decoded fields), or (foreign='hold') None for a class kept as a
    ForeignRecord.  Otherwise a TypeError names the class."""
    if name.startswith(RECORD_PACKAGE):
        module, _, qualname = name.partition(':')
        target = importlib.import_module(module)
        for part in qualname.split('.'):
            target = getattr(target, part)
        return target
    builder = _numerics_builder(name)
    if builder is not None:
        return builder
    if foreign == 'hold':
        return None
    if name.startswith(('src.Base.numerics:', 'src.Base.numerics.numerics:')):
        raise TypeError(
            f'a record names {name}: of the reference\'s Numerics only the bse and '
            f'quadrature sections (and Numerics itself) are modelled here; read with '
            f"foreign='hold' to keep the other sections as ForeignRecord values")
    raise TypeError(f'a record names {name}, which is not a class of {RECORD_PACKAGE[:-1]}; '
                    f"read with foreign='hold' to keep it as a ForeignRecord")


# ---------------------------------------------------------------------------
# value <-> tree
# ---------------------------------------------------------------------------

class _Encoder:
    """One value's tree; its arrays are collected in `arrays` as a0, a1, ..."""

    def __init__(self):
        self.arrays = {}

    def array(self, a, where):
        if a.dtype.hasobject:
            raise TypeError(f'{where}: arrays of Python objects cannot be recorded')
        key = f'a{len(self.arrays)}'
        self.arrays[key] = a
        return {'__array__': key}

    def __call__(self, v, where='record'):
        if v is None or isinstance(v, (bool, str)):
            return v
        if isinstance(v, np.ndarray):
            return self.array(v, where)
        if isinstance(v, np.generic):
            return {'__scalar__': v.dtype.str,
                    'value': self.array(np.asarray(v).reshape(1), where)}
        if isinstance(v, (int, float)):
            return v
        if isinstance(v, list):
            return [self(x, f'{where}[{i}]') for i, x in enumerate(v)]
        if isinstance(v, tuple):
            return {'__tuple__': [self(x, f'{where}[{i}]') for i, x in enumerate(v)]}
        if isinstance(v, frozenset):
            return {'__frozenset__': [self(x, f'{where}{{}}') for x in v]}
        if isinstance(v, set):
            return {'__set__': [self(x, f'{where}{{}}') for x in v]}
        if isinstance(v, dict):
            return self.mapping(v, where)
        if isinstance(v, ForeignRecord):
            body = {k: self(x, f'{where}.{k}') for k, x in v.content.items()}
            if v.kind == 'dataclass':
                return {'__dataclass__': v.name, 'fields': body}
            return {'__object__': v.name, 'state': body}
        return self.instance(v, where)

    def mapping(self, d, where):
        if all(isinstance(k, str) and not k.startswith('__') for k in d):
            return {k: self(x, f'{where}.{k}') for k, x in d.items()}
        return {'__dict__': [[self(k, f'{where}<key>'), self(x, f'{where}[{k!r}]')]
                             for k, x in d.items()]}

    def instance(self, v, where):
        name = class_name(v)
        if not name.startswith(RECORD_PACKAGE):
            raise TypeError(f'{where}: {name} is outside {RECORD_PACKAGE[:-1]}; a record '
                            'holds plain values, arrays and this package\'s classes')
        if dataclasses.is_dataclass(v):
            return {'__dataclass__': name,
                    'fields': {f.name: self(getattr(v, f.name), f'{where}.{f.name}')
                               for f in dataclasses.fields(v)}}
        return {'__object__': name, 'state': self.mapping(dict(vars(v)), f'{where}.__dict__')}


def encode(value):
    """(tree, {name: array}) of `value`: the JSON-able tree and its arrays."""
    enc = _Encoder()
    return enc(value), enc.arrays


def _tag_of(tree):
    tags = [k for k in tree if k in _TAGS]
    return tags[0] if tags else None


def _build_dataclass(cls, fields):
    init = {f.name for f in dataclasses.fields(cls) if f.init}
    obj = cls(**{k: v for k, v in fields.items() if k in init})
    for k, v in fields.items():
        if k not in init:
            object.__setattr__(obj, k, v)
    return obj


def decode(tree, arrays, foreign='refuse'):
    """The value `encode` gave (tree, arrays) for.  foreign: 'refuse' (a
    class from outside this package raises TypeError) or 'hold' (it comes
    back as a ForeignRecord)."""
    if foreign not in ('refuse', 'hold'):
        raise ValueError(f"foreign={foreign!r}: 'refuse' or 'hold'")

    def walk(t):
        if isinstance(t, list):
            return [walk(x) for x in t]
        if not isinstance(t, dict):
            return t
        tag = _tag_of(t)
        if tag is None:
            return {k: walk(x) for k, x in t.items()}
        body = t[tag]
        if tag == '__array__':
            return arrays[body]
        if tag == '__scalar__':
            return walk(t['value']).astype(body)[0]
        if tag == '__tuple__':
            return tuple(walk(x) for x in body)
        if tag == '__set__':
            return set(walk(x) for x in body)
        if tag == '__frozenset__':
            return frozenset(walk(x) for x in body)
        if tag == '__dict__':
            return {walk(k): walk(x) for k, x in body}
        cls = resolve_class(body, foreign)
        if tag == '__dataclass__':
            fields = {k: walk(x) for k, x in t['fields'].items()}
            if cls is None:
                return ForeignRecord(body, 'dataclass', fields)
            if not isinstance(cls, type):         # a reference numerics section
                return cls(fields)
            return _build_dataclass(cls, fields)
        state = walk(t['state'])
        if cls is None:
            return ForeignRecord(body, 'object', state)
        if not isinstance(cls, type):
            raise TypeError(f'a record holds {body} as a plain object; it is a dataclass')
        obj = cls.__new__(cls)
        obj.__dict__.update(state)
        return obj

    return walk(tree)


# ---------------------------------------------------------------------------
# files
# ---------------------------------------------------------------------------

def _replace_atomically(path, write):
    """write(handle) into a temporary beside `path`, then rename it over it."""
    tmp = path.with_name(path.name + '.tmp')
    with open(tmp, 'wb') as fh:
        write(fh)
    os.replace(tmp, path)


def write_record(path, value):
    """Write `value` as `path` (JSON) and its arrays beside it; atomic as a
    pair, and this stem's older array files are removed afterwards."""
    path = pathlib.Path(path)
    tree, arrays = encode(value)
    npz = path.with_name(f'{path.stem}.{uuid.uuid4().hex[:12]}.npz')
    _replace_atomically(npz, lambda fh: np.savez(fh, **arrays))
    doc = json.dumps({'arrays': npz.name, 'record': tree}).encode()
    _replace_atomically(path, lambda fh: fh.write(doc))
    for other in path.parent.iterdir():
        m = _TOKEN_FILE.fullmatch(other.name)
        if m and m['stem'] == path.stem and other.name != npz.name:
            other.unlink()


def read_record(path, foreign='refuse'):
    """The value a record file holds (see `decode` for `foreign`).  The
    arrays are read with pickling refused."""
    path = pathlib.Path(path)
    doc = json.loads(path.read_text())
    with np.load(path.with_name(doc['arrays']), allow_pickle=False) as npz:
        arrays = {k: npz[k] for k in npz.files}
    return decode(doc['record'], arrays, foreign)


def record_hash(value):
    """sha256 hex digest of a value's record: the tree as canonical JSON
    (sorted keys, no spaces), then per array in name order its name, dtype,
    shape and C-order bytes, each part length-prefixed.  Values that read
    back alike hash alike; the digest is this package's own."""
    tree, arrays = encode(value)
    h = hashlib.sha256()

    def part(b):
        h.update(len(b).to_bytes(8, 'little'))
        h.update(b)
    part(json.dumps(tree, sort_keys=True, separators=(',', ':')).encode())
    for key in sorted(arrays):
        a = np.ascontiguousarray(arrays[key])
        part(f'{key}|{a.dtype.str}|{a.shape}'.encode())
        part(a.tobytes())
    return h.hexdigest()


# ---------------------------------------------------------------------------
# code stamps (record_io.py:86-140)
# ---------------------------------------------------------------------------

def _git(root, *args):
    """stdout bytes of `git -C root args`, or None where git cannot answer."""
    try:
        done = subprocess.run(['git', '-C', str(root), *args], capture_output=True,
                              timeout=GIT_PROVENANCE_TIMEOUT)
    except (OSError, subprocess.SubprocessError):
        return None
    return done.stdout if done.returncode == 0 else None


def git_commit(root=None):
    """The checkout's short commit, with '+dirty' when the working tree
    differs from it; 'unknown' (and a RuntimeWarning) when git gives no
    answer, since a number without its commit cannot be re-run."""
    root = CHECKOUT if root is None else root
    head = _git(root, 'rev-parse', '--short', 'HEAD')
    status = _git(root, 'status', '--porcelain') if head is not None else None
    if head is None or status is None:
        warnings.warn(f'git at {root} gave no commit for the record; stamped unknown',
                      RuntimeWarning, stacklevel=2)
        return 'unknown'
    return head.decode().strip() + ('+dirty' if status.strip() else '')


def code_id(root=None):
    """The code a record comes from: the first word of SNAPSHOT.txt in an
    exported tree, else `git_commit`, extended on an edited tree by ':' and
    the first 12 hex digits of the sha256 of `git diff HEAD`, so two edits of
    one commit are two codes."""
    root = pathlib.Path(CHECKOUT if root is None else root)
    snapshot = root / 'SNAPSHOT.txt'
    if snapshot.is_file():
        return snapshot.read_text().split()[0]
    stamp = git_commit(root)
    if not stamp.endswith('+dirty'):
        return stamp
    diff = _git(root, 'diff', 'HEAD', '--binary') or b''
    return f'{stamp}:{hashlib.sha256(diff).hexdigest()[:12]}'


def same_code(recorded, current):
    """Whether a record's code stamp is this run's code: the same stamp, or
    a full 40-digit commit of which a clean `current` is the short form."""
    if recorded == current:
        return True
    clean = current != 'unknown' and '+' not in current
    return (clean and isinstance(recorded, str) and len(recorded) == 40 and
            recorded.startswith(current))


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
