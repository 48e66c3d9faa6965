"""Device-predicate context.

Inside ``with predicated(flag):`` every kernel launched through the
library wrappers is given ``flag`` (a device int32) as its predicate and
becomes a no-op while ``*flag == 0``.  CG uses it to enqueue a fixed number
of iterations whose stopping test is evaluated on the device.
"""

from __future__ import annotations

import threading
from contextlib import contextmanager

_tls = threading.local()


def current():
    return getattr(_tls, "pred", None)


@contextmanager
def predicated(flag):
    prev = current()
    _tls.pred = flag
    try:
        yield
    finally:
        _tls.pred = prev


def resolve(pred):
    """Explicit predicate if given, else the context's."""
    return pred if pred is not None else current()
