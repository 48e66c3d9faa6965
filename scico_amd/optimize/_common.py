"""Optimizer base (``scico/optimize/_common.py:77-320``).

Iteration statistics are recorded as device values (``IterationStats``
materialises them lazily), and the ``Time`` field is the device-stream
time at the end of each iteration, so ``solve`` never synchronises
unless ``nanstop`` or ``display`` asks for host values.
"""

from __future__ import annotations

from typing import Any, Callable, Optional

import numpy as np
import torch

from ..diagnostics import IterationStats
from ..util import DeviceElapsed, Timer


def itstat_func_and_object(itstat_fields: dict, itstat_attrib: list,
                           itstat_options: Optional[dict] = None):
    """``_common.py:28-74``: build the insertion function and the stats object."""
    attribs = list(itstat_attrib)

    def itstat_func(obj):
        return tuple(eval("obj." + a, {"obj": obj}) for a in attribs)

    opts = {"fields": itstat_fields, "itstat_func": itstat_func, "display": False}
    if itstat_options:
        opts.update(itstat_options)
    func = opts.pop("itstat_func", None)
    return func, IterationStats(**opts)


class Optimizer:
    def __init__(self, **kwargs: Any):
        iter0 = kwargs.pop("iter0", 0)
        self.maxiter: int = kwargs.pop("maxiter", 100)
        self.nanstop: bool = kwargs.pop("nanstop", False)
        itstat_options = kwargs.pop("itstat_options", None)
        if kwargs:
            raise TypeError(f"Unrecognized keyword argument(s) {', '.join([k for k in kwargs])}")
        self.itnum: int = iter0
        self.timer = Timer()
        self._t_ref = None
        fields, attrib = self._itstat_default_fields()
        xf, xa = self._itstat_extra_fields()
        fields.update(xf)
        attrib.extend(xa)
        self.itstat_insert_func, self.itstat_object = itstat_func_and_object(
            fields, attrib, itstat_options
        )

    def _working_vars_finite(self) -> bool:
        raise NotImplementedError("NaN check requested but _working_vars_finite not implemented.")

    def _itstat_default_fields(self):
        fields = {"Iter": "%d", "Time": "%8.2e"}
        attrib = ["itnum", "_elapsed()"]
        if self._objective_evaluatable():
            fields.update({"Objective": "%9.3e"})
            attrib.append("objective()")
        return fields, attrib

    def _elapsed(self):
        if torch.cuda.is_available() and self._t_ref is not None:
            return DeviceElapsed(self._t_ref[0], self._t_ref[1])
        return self.timer.elapsed()

    def _objective_evaluatable(self) -> bool:
        return False

    def _itstat_extra_fields(self):
        return {}, []

    def _state_variable_names(self):
        raise NotImplementedError(f"Method _state_variables is not implemented for {type(self)}.")

    def _get_state_variables(self) -> dict:
        return {k: getattr(self, k) for k in self._state_variable_names()}

    def _set_state_variables(self, **kwargs):
        valid = self._state_variable_names()
        for k, v in kwargs.items():
            if k not in valid:
                raise RuntimeError(f"{k} is not a valid state variable for {type(self)}.")
            setattr(self, k, v)

    def save_state(self, path: str):
        """``np.savez`` of class name, itnum, history and state (``_common.py:222-239``)."""
        def host(v):
            if isinstance(v, torch.Tensor):
                return v.detach().cpu().numpy()
            if isinstance(v, list):
                return np.stack([host(e) for e in v])
            return np.asarray(v)

        from ..blockarray import BlockArray

        state = {}
        for k, v in self._get_state_variables().items():
            if isinstance(v, BlockArray):  # blocks of different shapes: one entry each
                state[f"{k}__nblocks"] = np.array(len(v))
                for i, b in enumerate(v):
                    state[f"{k}__b{i}"] = host(b)
            else:
                state[k] = host(v)
        hist = self.history(transpose=True)
        np.savez(path, opt_class=np.array(self.__class__.__name__), itnum=self.itnum,
                 history=np.array([list(c) for c in hist], dtype=np.float64) if len(hist[0]) else
                 np.zeros((len(hist), 0)),
                 history_fields=np.array(list(hist._fields)), **state)

    def load_state(self, path: str):
        """Restore a state saved by :meth:`save_state` (``_common.py:241-258``)."""
        npz = np.load(path, allow_pickle=False)
        if str(npz["opt_class"]) != self.__class__.__name__:
            raise TypeError(f"Cannot load state for {npz['opt_class']} into optimizer of type "
                            f"{self.__class__}.")
        self.itnum = int(npz["itnum"])
        hist = npz["history"]
        obj = self.itstat_object
        obj._raw = [tuple(row) for row in hist.T] if hist.size else []
        obj.iterations = []
        from ..blockarray import BlockArray

        state = {}
        for k in self._state_variable_names():
            cur = getattr(self, k)
            if f"{k}__nblocks" in npz:
                n = int(npz[f"{k}__nblocks"])
                dv = cur[0].device if isinstance(cur, list) else cur.device
                state[k] = BlockArray(torch.as_tensor(npz[f"{k}__b{i}"], device=dv)
                                      for i in range(n))
                continue
            v = npz[k]
            if isinstance(cur, list):
                state[k] = [torch.as_tensor(e, device=cur[0].device) for e in v]
            elif isinstance(cur, torch.Tensor):
                state[k] = torch.as_tensor(v, device=cur.device)
            else:
                state[k] = type(cur)(v) if np.ndim(v) == 0 else v
        self._set_state_variables(**state)
        self._after_load_state()

    def _after_load_state(self):
        pass

    def history(self, transpose: bool = False):
        return self.itstat_object.history(transpose=transpose)

    def minimizer(self):
        raise NotImplementedError(f"Method minimizer is not implemented for {type(self)}.")

    def step(self):
        raise NotImplementedError(f"Method step is not implemented for {type(self)}.")

    def solve(self, callback: Optional[Callable] = None):
        """Run ``maxiter`` iterations (``_common.py:287-320``)."""
        self.timer.start()
        if torch.cuda.is_available():
            ev = torch.cuda.Event(enable_timing=True)
            ev.record()
            self._t_ref = (ev, self.timer.elapsed())
        for self.itnum in range(self.itnum, self.itnum + self.maxiter):
            self.step()
            if self.nanstop and not self._working_vars_finite():
                raise ValueError(
                    f"NaN or Inf value encountered in working variable in iteration {self.itnum}."
                )
            self.itstat_object.insert(self.itstat_insert_func(self))
            if callback:
                self.timer.stop()
                callback(self)
                self.timer.start()
        self.timer.stop()
        self.itnum += 1
        self.itstat_object.end()
        return self.minimizer()
