"""Per-report-step result store with restart, and a pickle-free wire format
for the device path's states.

:class:`ResultHandler` keeps the reference's interface and behaviour
(pymrst/modules/ad_core/utils/result_handler.py:42-318, after MRST's
``ResultHandler.m``):

* ids are MRST's 1-based step numbers and appear in the file names;
* memory and disk are independent;
* ``get_valid_ids`` reports gaps;
* ``get_restart`` returns ``(None, 1)`` on an empty store;
* reads outside the store's directory are refused.

The reference writes pickles, and unpickling runs code. Here a step is one
``.npz`` file of plain arrays, read back with ``allow_pickle=False``, so a
stored step is data, never code:

* a state dataclass becomes its fields plus a kind tag, and only kinds
  registered with :func:`register` are rebuilt;
* a device state (``DeviceOWState``, ``DeviceBOState``) is written as its
  host state;
* dicts, lists and tuples nest, with keys joined by ``/``. A list of member
  states is how an ensemble's states are stored, and a list of steps is how
  a forward trajectory is stored for an adjoint sweep.

:func:`simulate_with_output` runs a schedule one report step at a time into
a handler and resumes from what the handler already holds. This is the
restart path of MRST's packed simulation runner
(pymrst/modules/ad_core/simulators/sim_runner/packed_problem.py:14-118).
"""
from __future__ import annotations

import dataclasses
import re
from pathlib import Path
from typing import Any, Callable, Dict, Iterator, List, Optional, Sequence, Tuple

import numpy as np

_KIND = "__kind__"
_REGISTRY: Dict[str, type] = {}


def register(cls: type, name: Optional[str] = None) -> type:
    """Allow a dataclass to be rebuilt from a stored step under `name`."""
    if not dataclasses.is_dataclass(cls):
        raise TypeError(f"{cls!r} is not a dataclass")
    _REGISTRY[name or cls.__name__] = cls
    return cls


def _known(name: str) -> type:
    if name not in _REGISTRY:
        # the package's own state classes register on first need
        if name in ("OWState",):
            from .ow import OWState
            register(OWState)
        elif name in ("BOState",):
            from .bo import BOState
            register(BOState)
    if name not in _REGISTRY:
        raise ValueError(f"stored kind '{name}' is not registered (register(cls) allows it)")
    return _REGISTRY[name]


def _leaf(x) -> np.ndarray:
    if hasattr(x, "detach") and hasattr(x, "cpu"):          # a torch tensor
        x = x.detach().cpu().numpy()
    a = np.asarray(x)
    if a.dtype == object:
        raise TypeError(f"cannot store a value of type {type(x).__name__} as plain data")
    return a


def encode(obj: Any, prefix: str = "") -> Dict[str, np.ndarray]:
    """Flatten `obj` into named plain arrays (no object arrays)."""
    out: Dict[str, np.ndarray] = {}
    key = (lambda k: f"{prefix}/{k}" if prefix else str(k))
    if hasattr(obj, "to_host") and callable(obj.to_host):  # a device state
        obj = obj.to_host()
    if dataclasses.is_dataclass(obj) and not isinstance(obj, type):
        name = next((n for n, c in _REGISTRY.items() if c is type(obj)), type(obj).__name__)
        if _REGISTRY.get(name, type(obj)) is not type(obj):
            raise ValueError(f"another class is registered as '{name}'; register this one under its own name")
        out[key(_KIND)] = np.asarray(f"dataclass:{name}")
        for f in dataclasses.fields(obj):
            out.update(encode(getattr(obj, f.name), key(f.name)))
    elif isinstance(obj, dict):
        out[key(_KIND)] = np.asarray("dict")
        for k, v in obj.items():
            if not isinstance(k, str) or "/" in k or not k:
                raise ValueError(f"dict keys must be non-empty strings without '/', got {k!r}")
            out.update(encode(v, key(k)))
    elif isinstance(obj, (list, tuple)):
        out[key(_KIND)] = np.asarray(f"{type(obj).__name__}:{len(obj)}")
        for i, v in enumerate(obj):
            out.update(encode(v, key(i)))
    elif obj is None:
        out[key(_KIND)] = np.asarray("none")
    elif isinstance(obj, str):
        out[key(_KIND)] = np.asarray("str")
        out[key("value")] = np.asarray(obj)
    elif isinstance(obj, (bool, int, float, np.generic)):
        out[key(_KIND)] = np.asarray("scalar")
        out[key("value")] = np.asarray(obj)
    else:
        if not prefix:
            raise TypeError(f"cannot store a top-level {type(obj).__name__}; wrap it in a dict")
        out[prefix] = _leaf(obj)
    return out


def decode(arrays: Dict[str, np.ndarray], prefix: str = "") -> Any:
    """Rebuild what :func:`encode` flattened."""
    key = (lambda k: f"{prefix}/{k}" if prefix else str(k))
    kind_key = key(_KIND)
    if kind_key not in arrays:
        if prefix in arrays:
            return arrays[prefix]
        raise KeyError(f"stored step has no entry '{prefix}'")
    kind = str(arrays[kind_key][()])
    if kind.startswith("dataclass:"):
        cls = _known(kind.split(":", 1)[1])
        return cls(**{f.name: decode(arrays, key(f.name)) for f in dataclasses.fields(cls)})
    if kind == "dict":
        head = prefix + "/" if prefix else ""
        names = []
        for k in arrays:
            if k.startswith(head) and k != kind_key:
                name = k[len(head):].split("/", 1)[0]
                if name not in names:
                    names.append(name)
        return {name: decode(arrays, key(name)) for name in names}
    if kind.startswith(("list:", "tuple:")):
        seq_type, n = kind.split(":")
        items = [decode(arrays, key(i)) for i in range(int(n))]
        return items if seq_type == "list" else tuple(items)
    if kind == "none":
        return None
    if kind == "str":
        return str(arrays[key("value")][()])
    if kind == "scalar":
        return arrays[key("value")][()].item()
    raise ValueError(f"unknown stored kind '{kind}'")


def _time_of(state: Any) -> float:
    """The simulation time of a stored result (result_handler.py:320-330)."""
    t = state.get("time") if isinstance(state, dict) else getattr(state, "time", None)
    if t is None:
        raise KeyError("stored result has no 'time' field to search on")
    return float(t)


class _MemoryShelf:
    """Steps held as Python objects, keyed by id."""

    def __init__(self):
        self.items: Dict[int, Any] = {}

    def ids(self) -> List[int]:
        return sorted(self.items)

    def has(self, index: int) -> bool:
        return index in self.items

    def load(self, index: int) -> Any:
        return self.items[index]

    def save(self, index: int, value: Any) -> None:
        self.items[index] = value

    def remove(self, index: int) -> None:
        self.items.pop(index, None)

    def remove_all(self) -> None:
        self.items.clear()


class _DiskShelf:
    """One ``<prefix><id>.npz`` file per step in one directory.  Files are
    written beside their final name and renamed into place, so a run killed
    while writing never leaves a torn step; reads are confined to the
    directory."""

    def __init__(self, folder: Path, prefix: str, suffix: str, verbose: bool):
        self.folder, self.prefix, self.suffix, self.verbose = folder, prefix, suffix, verbose

    def file(self, index: int) -> Path:
        return self.folder / f"{self.prefix}{index}{self.suffix}"

    def _id_of(self, name: str) -> Optional[int]:
        if not (name.startswith(self.prefix) and name.endswith(self.suffix)):
            return None
        digits = name[len(self.prefix):len(name) - len(self.suffix)]
        return int(digits) if digits.isdigit() else None

    def ids(self) -> List[int]:
        if not self.folder.is_dir():
            return []
        found = (self._id_of(entry.name) for entry in self.folder.iterdir())
        return sorted(i for i in found if i is not None)

    def has(self, index: int) -> bool:
        return self.file(index).is_file()

    def _checked(self, index: int) -> Path:
        target = self.file(index).resolve()
        home = self.folder.resolve()
        if home != target.parent and home not in target.parents:
            raise ValueError(f"refusing to read {target}: outside the store directory {home}")
        return target

    def load(self, index: int) -> Any:
        target = self._checked(index)
        if not target.is_file():
            raise KeyError(f"step {index} is not stored in {self.folder}")
        with np.load(target, allow_pickle=False) as npz:
            return decode({name: npz[name] for name in npz.files})

    def save(self, index: int, value: Any) -> None:
        self.folder.mkdir(parents=True, exist_ok=True)
        final = self.file(index)
        staging = final.parent / (final.name + ".part")
        with open(staging, "wb") as handle:
            np.savez(handle, **encode(value))
        staging.replace(final)
        if self.verbose:
            print(f"stored step {index} in {final}")

    def remove(self, index: int) -> None:
        target = self.file(index)
        if target.is_file():
            target.unlink()
            if self.verbose:
                print(f"removed step {index} ({target})")

    def remove_all(self) -> None:
        for index in self.ids():
            self.remove(index)


def _camel(name: str) -> str:
    head, *rest = name.split("_")
    return head + "".join(word.capitalize() for word in rest)


class ResultHandler:
    """A store of per-step results addressed by MRST's 1-based step ids,
    held in memory, on disk, or both (the interface of the reference's
    ResultHandler, pymrst/modules/ad_core/utils/result_handler.py:42-318).

    The two places are independent shelves: a write goes to every enabled
    shelf, a read prefers memory, and the ids a store reports are those of
    its durable shelf (disk when enabled, else memory), gaps included.  A
    step on disk is one ``.npz`` of plain arrays (:func:`encode`)."""

    def __init__(self, data_directory: Optional[str | Path] = None, data_folder: str = "results",
                 data_prefix: str = "state", write_to_disk: bool = False, store_in_memory: bool = True,
                 extension: str = ".npz", cleardir: bool = False, verbose: bool = False):
        if extension != ".npz":
            raise ValueError(f"extension {extension!r} is not supported: steps are .npz files of plain arrays")
        self.data_directory = Path(data_directory) if data_directory else Path.cwd()
        self.data_folder, self.data_prefix, self.extension = data_folder, data_prefix, extension
        self.write_to_disk, self.store_in_memory = bool(write_to_disk), bool(store_in_memory)
        self.verbose = bool(verbose)
        self._memory = _MemoryShelf()
        self._disk = _DiskShelf(self.data_directory / data_folder, data_prefix, extension, self.verbose)
        if self.write_to_disk:
            self._disk.folder.mkdir(parents=True, exist_ok=True)
            if cleardir:
                self._disk.remove_all()

    @property
    def data(self) -> Dict[int, Any]:
        """The steps held in memory."""
        return self._memory.items

    def _shelves(self):
        return ([self._memory] if self.store_in_memory else []) + ([self._disk] if self.write_to_disk else [])

    def _durable(self):
        return self._disk if self.write_to_disk else self._memory

    # -- locations ---------------------------------------------------------
    def get_extension(self, wildcard: bool = False) -> str:
        return ("*" if wildcard else "") + self.extension

    def get_data_path(self, index: Optional[int] = None) -> Path:
        if index is None:
            return self._disk.folder
        if isinstance(index, bool) or not isinstance(index, (int, np.integer)):
            raise TypeError(f"a step id is one integer, not {index!r}")
        return self._disk.file(int(index))

    # -- files -------------------------------------------------------------
    def write_to_file(self, data: Any, index: int) -> "ResultHandler":
        self._disk.save(int(index), data)
        return self

    def read_from_file(self, indices: int | Sequence[int]) -> List[Any]:
        wanted = [indices] if isinstance(indices, (int, np.integer)) else list(indices)
        return [self._disk.load(int(i)) for i in wanted]

    def delete_file(self, index: int) -> "ResultHandler":
        self._disk.remove(int(index))
        return self

    # -- what is held ------------------------------------------------------
    def get_valid_ids(self) -> List[int]:
        return self._durable().ids()

    def numel_data(self) -> int:
        held = self.get_valid_ids()
        return held[-1] if held else 0

    def size_data(self, dim: Optional[int] = None):
        shape = (self.numel_data(), 1)
        return shape[dim - 1] if dim else shape

    def __len__(self) -> int:
        return len(self.get_valid_ids())

    def __contains__(self, index: int) -> bool:
        return self._durable().has(index)

    def __iter__(self) -> Iterator[Any]:
        return (self[i] for i in self.get_valid_ids())

    # -- access ------------------------------------------------------------
    def __getitem__(self, index: int) -> Any:
        for shelf in self._shelves():
            if shelf.has(index):
                return shelf.load(index)
        raise KeyError(f"step {index} is not in this store")

    def __setitem__(self, index: int, value: Any) -> None:
        shelves = self._shelves()
        if not shelves:
            raise ValueError("with neither memory nor disk enabled every stored step would be discarded; "
                             "enable store_in_memory or write_to_disk")
        if callable(getattr(value, "to_host", None)):
            value = value.to_host()
        for shelf in shelves:
            shelf.save(index, value)

    def __delitem__(self, index: int) -> None:
        self._memory.remove(index)
        if self.write_to_disk:
            self._disk.remove(index)

    def get_data(self, indices: Optional[Sequence[int]] = None) -> List[Any]:
        return [self[i] for i in (self.get_valid_ids() if indices is None else indices)]

    def reset(self) -> "ResultHandler":
        self.reset_data()
        return self

    def reset_data(self, subs=None) -> None:
        """Remove every step, or the steps `subs` names: ids, or a boolean
        mask over ids 1, 2, ... (result_handler.py:221-247)."""
        if subs is None:
            self._memory.remove_all()
            if self.write_to_disk:
                self._disk.remove_all()
            return
        sel = np.atleast_1d(np.asarray(subs))
        doomed = (np.nonzero(sel)[0] + 1) if sel.dtype == bool else sel.astype(np.int64)
        for index in map(int, doomed):
            for shelf in self._shelves():
                shelf.remove(index)

    # -- search ------------------------------------------------------------
    def _walk(self, reverse: bool):
        held = self.get_valid_ids()
        return reversed(held) if reverse else iter(held)

    def get_first_by_function(self, predicate: Callable[[Any], bool], reverse: bool = False):
        for index in self._walk(reverse):
            step = self[index]
            if predicate(step):
                return step, index
        return None, None

    def get_by_time(self, time: float, tol: float = 1e-3, reverse: bool = False):
        return self.get_first_by_function(lambda step: abs(_time_of(step) - time) < tol, reverse)

    def get_closest_by_time(self, time: float, reverse: bool = False):
        scored = [(abs(_time_of(self[i]) - time), n, i) for n, i in enumerate(self._walk(reverse))]
        if not scored:
            return None, None, float("inf")
        gap, _, index = min(scored)
        return self[index], index, gap

    def get_restart(self) -> Tuple[Any, int]:
        """(the last stored step, the id to continue at); (None, 1) when the
        store is empty."""
        last = self.numel_data()
        return (None, 1) if last == 0 else (self[last], last + 1)

    def __repr__(self) -> str:
        places = [name for name, on in (("memory", self.store_in_memory),
                                         (str(self._disk.folder), self.write_to_disk)) if on]
        return f"ResultHandler({len(self)} step(s) in {' and '.join(places) or 'no place'})"


# MRST's camel-case method names, as the reference also exposes them
for _name in ("get_extension", "get_data_path", "write_to_file", "read_from_file", "delete_file", "get_valid_ids",
              "numel_data", "size_data", "get_data", "reset_data", "get_first_by_function", "get_by_time",
              "get_closest_by_time", "get_restart"):
    setattr(ResultHandler, _camel(_name), getattr(ResultHandler, _name))
del _name


def simulate_with_output(state0, model, timesteps: Sequence[float], states: ResultHandler,
                         reports: Optional[ResultHandler] = None, simulate: Optional[Callable] = None,
                         **kwargs):
    """Run `timesteps` one report step at a time, storing step i's state as
    ``states[i]`` (1-based) and, if given, ``reports[i] = {"time",
    "dt", "iterations"}``; resume after the last step `states` already
    holds.  `simulate(state, model, [dt], **kwargs) -> (states,
    iterations)` is the schedule driver (the device ``simulate_schedule_ad``
    by default).  A model with ``run_memory()`` / ``restore_run_memory()``
    (what a run carries between steps besides the state: the oil-water
    model's previous well cdp) has it stored with each report and restored
    on resume, so the resumed run continues the interrupted one exactly;
    without a report store the resumed run starts from a fresh model's
    memory, as a restart from states alone does.  Returns every step's state
    (stored or new) and the iteration counts of the steps run now (0 for
    resumed steps)."""
    if simulate is None:
        from .ow import simulate_schedule_ad as simulate
    timesteps = np.asarray(timesteps, float)
    ids = states.get_valid_ids()
    if ids and ids != list(range(1, max(ids) + 1)):
        raise ValueError(f"the store holds steps {ids}, not a contiguous run from step 1")
    remembers = callable(getattr(model, "run_memory", None))
    if remembers and reports is not None:
        # a state whose report (and run memory) never reached the store is
        # an interrupted write: drop it and run that step again
        while ids and ids[-1] not in reports:
            del states[ids.pop()]
    last, start = states.get_restart()
    if start - 1 > len(timesteps):
        raise ValueError(f"the store holds step {start - 1}, beyond the schedule's {len(timesteps)} steps")
    state = state0 if last is None else last
    if remembers and reports is not None and start > 1:
        model.restore_run_memory(reports[start - 1].get("model_memory"))
    its = np.zeros(len(timesteps), dtype=np.int64)
    for i in range(start - 1, len(timesteps)):
        out, nit = simulate(state, model, [float(timesteps[i])], **kwargs)
        state = out[-1]
        its[i] = int(nit[0])
        # the report (with the run memory) first: a run killed between the
        # two writes leaves an orphan report, which the next step overwrites,
        # never a state without its memory
        if reports is not None:
            rep = {"time": float(timesteps[:i + 1].sum()), "dt": float(timesteps[i]), "iterations": int(its[i])}
            if remembers:
                rep["model_memory"] = model.run_memory()
            reports[i + 1] = rep
        states[i + 1] = state
    return [states[i] for i in range(1, len(timesteps) + 1)], its


__all__ = ["ResultHandler", "encode", "decode", "register", "simulate_with_output"]
