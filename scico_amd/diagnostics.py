"""Iteration statistics (``scico/diagnostics.py:18-216``).

Values may be device scalars (0-d tensors) or lazy timers; they are
materialised on the host only when displayed or when ``history`` is read,
so recording statistics costs no synchronisation per iteration.
"""

from __future__ import annotations

import re
from collections import namedtuple
from typing import Optional

import torch


def _materialise(v):
    if isinstance(v, torch.Tensor):
        return v.item()
    if callable(getattr(v, "value", None)):
        return v.value()
    return v


class IterationStats:
    def __init__(self, fields: dict, ident: Optional[dict] = None, display: bool = False,
                 period: int = 1, shift_cycles: bool = True, overwrite: bool = True,
                 colsep: int = 2):
        if not isinstance(fields, dict):
            raise TypeError("Argument 'fields' must be an instance of dict.")
        self.period = period
        self.period_offset = 1 if shift_cycles else 0
        self.overwrite = overwrite
        self.colsep = colsep
        self.iterations = []
        self._raw = []
        self.fieldname, self.fieldformat, self.fieldlength, self.tuplefields = [], [], [], []
        self.headlength = 0
        fmre = re.compile(r"%(\+?-?)((?:\d+)?)(\.?)((?:\d+)?)([a-z])")
        for name in fields:
            fmt = fields[name]
            m = fmre.match(fmt)
            if not m:
                raise ValueError(f"Format string '{fmt}' could not be parsed.")
            fmflg, fmlen, fmdot, fmprc, fmtyp = m.groups()
            flen = len(fmt % 0)
            if flen < len(name):
                fmt = f"%{fmflg}{len(name)}{fmdot}{fmprc}{fmtyp}"
                flen = len(name)
            self.fieldname.append(name)
            self.fieldformat.append(fmt)
            self.fieldlength.append(flen)
            self.headlength += flen + colsep
            if ident is not None and name in ident:
                self.tuplefields.append(ident[name])
            else:
                t = re.sub(r"\W+|^(?=\d)", "_", name)
                self.tuplefields.append(t[1:] if t[0] == "_" else t)
        self.headlength -= colsep
        self.IterTuple = namedtuple("IterationStatsTuple", self.tuplefields)
        self.display = display
        self.disphdr = None
        if display:
            self.disphdr = (" " * colsep).join(
                ["%-*s" % (fl, fn) for fl, fn in zip(self.fieldlength, self.fieldname)]
            ) + "\n" + "-" * self.headlength

    def insert(self, values):
        """Record one iteration (device values stay on the device)."""
        self._raw.append(tuple(values))
        if self.display:
            vals = [_materialise(v) for v in values]
            if self.disphdr is not None:
                print(self.disphdr)
                self.disphdr = None
            end = "\n"
            if self.overwrite and (len(self._raw) - self.period_offset) % self.period != 0:
                end = "\r"
            if self.overwrite or (len(self._raw) - self.period_offset) % self.period == 0:
                print((" " * self.colsep).join(self.fieldformat) % tuple(vals), end=end)

    def _flush(self):
        if len(self.iterations) < len(self._raw):
            pending = self._raw[len(self.iterations):]
            for row in pending:
                self.iterations.append(self.IterTuple(*[_materialise(v) for v in row]))

    def end(self):
        if self.display and self.overwrite:
            print()

    def history(self, transpose: bool = False):
        self._flush()
        if transpose:
            return self.IterTuple(*[list(x) for x in zip(*self.iterations)])
        return self.iterations
