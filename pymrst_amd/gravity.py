"""Process-wide gravity setting, as MRST's ``gravity`` keeps it
(pymrst/params/gravity/gravity.py:15-88): off by default; ``gravity('on')``
/ ``gravity('off')`` switch it, ``'x'``/``'y'``/``'z'`` choose the axis, a
number sets the strength along the current axis, a 2- or 3-vector sets the
vector itself, ``'reset'`` restores the default (off, 9.80665 along z).
``incomp_tpfa`` reads it when no gravity argument is given."""
from contextlib import contextmanager
from typing import Iterator

import numpy as np

from .grid import NORM_GRAVITY


class _GravitySetting:
    def __init__(self):
        self.reset()

    def reset(self):
        self.magnitude = NORM_GRAVITY
        self.axis = np.array([0.0, 0.0, 1.0])
        self.on = False

    def vector(self) -> np.ndarray:
        return self.magnitude * self.axis if self.on else np.zeros(3)

    def snapshot(self):
        return self.magnitude, self.axis.copy(), self.on

    def restore(self, snap):
        self.magnitude, self.axis, self.on = snap[0], snap[1].copy(), snap[2]


_SETTING = _GravitySetting()
_SWITCHES = {"on": True, "yes": True, "true": True, "off": False, "no": False, "false": False}


def _apply(control) -> None:
    if isinstance(control, str):
        word = control.lower()
        if word in _SWITCHES:
            _SETTING.on = _SWITCHES[word]
        elif word in ("x", "y", "z"):
            _SETTING.axis = np.eye(3)["xyz".index(word)]
        elif word == "reset":
            _SETTING.reset()
        else:
            raise ValueError(f"gravity: unknown control {control!r}")
        return
    if isinstance(control, (bool, np.bool_)):
        _SETTING.on = bool(control)
        return
    arr = np.asarray(control)
    if arr.dtype.kind not in "iuf":
        raise TypeError("gravity: a control is a string, a boolean or real numbers")
    arr = arr.astype(float).ravel()
    if arr.size == 1:
        length = float(np.linalg.norm(_SETTING.axis))
        if length == 0.0:
            raise ValueError("gravity: the direction is zero; reset it first")
        _SETTING.magnitude = float(arr[0])
        _SETTING.axis = _SETTING.axis / length
    elif arr.size in (2, 3):
        _SETTING.magnitude = 1.0
        _SETTING.axis = np.concatenate([arr, np.zeros(3 - arr.size)])
    else:
        raise ValueError("gravity: a vector has two or three components")


def gravity(*controls) -> np.ndarray:
    """Apply the controls in order and return the active gravity vector."""
    for control in controls:
        _apply(control)
    return current_gravity()


def current_gravity(griddim: int = 3) -> np.ndarray:
    if griddim not in (1, 2, 3):
        raise ValueError("griddim is 1, 2 or 3")
    return _SETTING.vector()[:griddim].copy()


@contextmanager
def isolated_gravity(*controls) -> Iterator[np.ndarray]:
    """Apply controls for the duration of a with-block, then restore."""
    snap = _SETTING.snapshot()
    try:
        yield gravity(*controls)
    finally:
        _SETTING.restore(snap)
