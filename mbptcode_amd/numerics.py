"""The numerical settings this path reads, decided once: `Numerics`.

The reference keeps every numerical choice of a calculation in one frozen
object (src/Base/numerics/numerics.py:610-702) with a section per stage.
The hot path here reads two of them, and those are rebuilt field for field,
with the reference's names and defaults, so that a setting means the same
in both codes and records written by either read in the other:

* `QuadratureNumerics` (numerics.py:214-238): the time and frequency grids;
* `BSENumerics` (numerics.py:241-286): the Casida eigensolver, its roots,
  residual, cycle cap, guess count and preconditioner, the (A - B) probe,
  the root's adjoint and root following.

`Numerics` holds both, plus -- read from a reference record, never used
here -- the reference's other sections as inert `ForeignRecord` values
(`held`), and the provenance of a derived object (`parent`, `change`,
`reason`).  `identity` is a digest of the choices (not of the provenance);
`derive` makes other settings for another task and writes down the parent's
identity, what changed and why.
"""
from dataclasses import asdict, dataclass, field, fields, replace
from typing import Optional

from .constants import (BSE_ADJOINTS, BSE_DAVIDSON_CONV_TOL, BSE_DAVIDSON_NROOTS,
                        BSE_FORCE_MAX_CYCLE, BSE_SOLVERS, CD_NTAU, DAVIDSON_GUESS_FACTOR,
                        DAVIDSON_PRECONDITIONER, DAVIDSON_PRECONDITIONERS, DEFAULT_NPADE,
                        DEFAULT_TAU_TARGET, GL_FREQUENCY_SCALE, ROOT_FOLLOWING)


@dataclass(frozen=True)
class QuadratureNumerics:
    """The grids (numerics.py:214-238): the self-energy's time points
    (ntau_gw), the static W's (ntau_w; None sizes them by the Laplace test at
    tau_target), the contour's frequency count, how far below the gap the
    time axis is fitted, the dRPA energy's time and frequency counts (None:
    from the excited state's), the space-time self-energy's time and chi0
    frequency counts ('auto': from the spectrum), and its Pade points on a
    Gauss-Legendre grid of scale w0_gl."""
    ntau_gw: int = CD_NTAU
    ntau_w: Optional[int] = None
    nfreq_cd: Optional[int] = None
    e_min_below_gap: Optional[float] = None
    ntau_rpa: Optional[int] = None
    nfreq_rpa: Optional[int] = None
    tau_target: float = DEFAULT_TAU_TARGET
    ntau_st: object = 'auto'
    nfreq_st: object = 'auto'
    npade: int = DEFAULT_NPADE
    w0_gl: float = GL_FREQUENCY_SCALE


@dataclass(frozen=True)
class BSENumerics:
    """The Casida eigensolver (numerics.py:241-286): solver ('auto', 'dense',
    'davidson'), nroots, residual conv_tol, max_cycle, guess_factor unit
    guesses per root, the preconditioner, the (A - B) probe of the chain
    (True, 'sign' or False; the stand-alone drivers probe only when asked),
    the root's adjoint ('explicit', 'grid') and root following (None or
    'overlap').  Values outside those sets are refused."""
    solver: str = 'auto'
    nroots: Optional[int] = BSE_DAVIDSON_NROOTS
    conv_tol: float = BSE_DAVIDSON_CONV_TOL
    max_cycle: int = BSE_FORCE_MAX_CYCLE
    guess_factor: int = DAVIDSON_GUESS_FACTOR
    preconditioner: str = DAVIDSON_PRECONDITIONER
    probe: object = True
    adjoint: str = 'explicit'
    track: Optional[str] = None

    def __post_init__(self):
        checks = (('solver', self.solver, BSE_SOLVERS), ('adjoint', self.adjoint, BSE_ADJOINTS),
                  ('preconditioner', self.preconditioner, DAVIDSON_PRECONDITIONERS))
        for name, value, allowed in checks:
            if value not in allowed:
                raise ValueError(f'BSENumerics.{name}={value!r}: one of {allowed}')
        if self.track is not None and self.track not in ROOT_FOLLOWING:
            raise ValueError(f'BSENumerics.track={self.track!r}: None or one of '
                             f'{ROOT_FOLLOWING}')


#: The sections `Numerics` models, by field name.
SECTIONS = {'quadrature': QuadratureNumerics, 'bse': BSENumerics}


@dataclass(frozen=True)
class Numerics:
    """The settings of one calculation.  A section may be given as a dict of
    its fields.  `held`: the reference's other sections, (name, value) pairs
    carried from a record unread.  parent, change, reason: where a derived
    object came from (None / () on one a driver built)."""
    quadrature: QuadratureNumerics = field(default_factory=QuadratureNumerics)
    bse: BSENumerics = field(default_factory=BSENumerics)
    held: tuple = ()
    parent: Optional[str] = None
    change: tuple = ()
    reason: Optional[str] = None

    def __post_init__(self):
        for name, cls in SECTIONS.items():
            value = getattr(self, name)
            if isinstance(value, dict):
                value = cls(**value)
                object.__setattr__(self, name, value)
            if not isinstance(value, cls):
                raise TypeError(f'Numerics.{name} must be a {cls.__name__}, '
                                f'not {type(value).__name__}')
        held = self.held.items() if isinstance(self.held, dict) else self.held
        held = tuple(sorted((str(k), v) for k, v in held))
        clash = [k for k, _ in held if k in SECTIONS]
        if clash:
            raise ValueError(f'held sections {clash} are modelled sections')
        object.__setattr__(self, 'held', held)
        object.__setattr__(self, 'change', tuple(tuple(c) for c in self.change))

    def values(self):
        """{section: {field: value}} of the modelled sections."""
        return {name: asdict(getattr(self, name)) for name in SECTIONS}

    def to_record(self):
        """The settings and provenance as plain values (`from_record`)."""
        return dict(self.values(), held=dict(self.held), parent=self.parent,
                    change=[list(c) for c in self.change], reason=self.reason)

    @classmethod
    def from_record(cls, record):
        """The object of a `to_record` mapping; a reference `to_record`
        mapping (numerics.py:643-663) reads too, its unmodelled sections
        going to `held`."""
        record = dict(record)
        provenance = {k: record.pop(k, None) for k in ('parent', 'change', 'reason')}
        held = dict(record.pop('held', {}) or {})
        sections = {}
        for name, value in record.items():
            if name in SECTIONS:
                sections[name] = SECTIONS[name](**dict(value))
            else:
                held[name] = value
        return cls(**sections, held=held, parent=provenance['parent'],
                   change=tuple(tuple(c) for c in provenance['change'] or ()),
                   reason=provenance['reason'])

    def identity(self):
        """sha256 of the settings (the modelled sections and the held ones),
        not of the provenance: equal settings, equal identity, however the
        object was made."""
        from .records import record_hash
        return record_hash({'sections': self.values(), 'held': list(self.held)})

    def derive(self, reason, **sections):
        """Settings for another task: `sections` maps a modelled section to
        {field: new value}.  The child records this object's identity as its
        parent, each changed (section.field, old, new) and `reason`.  No
        reason, an unknown section or field, or no change at all: refused."""
        if not reason or not str(reason).strip():
            raise ValueError('derive needs a reason: other settings are another '
                             'calculation, and the record must say why')
        unknown = set(sections) - set(SECTIONS)
        if unknown:
            raise ValueError(f'no modelled section {sorted(unknown)}; sections: '
                             f'{sorted(SECTIONS)}')
        updated, change = {}, []
        for name, updates in sections.items():
            current = getattr(self, name)
            names = {f.name for f in fields(current)}
            bad = set(updates) - names
            if bad:
                raise ValueError(f'{name} has no field {sorted(bad)}')
            updated[name] = replace(current, **updates)
            change.extend((f'{name}.{key}', getattr(current, key), value)
                          for key, value in sorted(updates.items())
                          if getattr(current, key) != value)
        if not change:
            raise ValueError('derive changes nothing: use the object itself')
        return replace(self, **updated, parent=self.identity(), change=tuple(change),
                       reason=str(reason))


def numerics_or_default(numerics):
    """`numerics`, or the default settings for None."""
    return Numerics() if numerics is None else numerics


def _from_reference_record(fields_):
    """A reference `Numerics` record's decoded fields as this module's
    Numerics: the modelled sections kept, every other section (a
    ForeignRecord by then) moved to `held`."""
    fields_ = dict(fields_)
    sections = {k: fields_.pop(k) for k in list(fields_) if k in SECTIONS}
    provenance = {k: fields_.pop(k, None) for k in ('parent', 'change', 'reason')}
    return Numerics(**sections, held=fields_, parent=provenance['parent'],
                    change=provenance['change'] or (), reason=provenance['reason'])


#: Reference class names (numerics.py) -> the function that builds this
#: module's counterpart from a record's decoded fields.
REFERENCE_BUILDERS = {'QuadratureNumerics': lambda f: QuadratureNumerics(**f),
                      'BSENumerics': lambda f: BSENumerics(**f),
                      'Numerics': _from_reference_record}
