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


class ResultHandler:
    """A list-like store of per-step results, in memory and/or on disk
    (result_handler.py:42-318); steps on disk are ``.npz`` files of plain
    arrays."""

    def __init__(self, data_directory: Optional[str | Path] = None, data_folder: str = "results",
                 data_prefix: str = "state", write_to_disk: bool = False, store_in_memory: bool = True,
                 extension: str = ".npz", cleardir: bool = False, verbose: bool = False):
        if extension != ".npz":
            raise ValueError("steps are stored as .npz files of plain arrays")
        self.data_directory = Path(data_directory) if data_directory else Path.cwd()
        self.data_folder = data_folder
        self.data_prefix = data_prefix
        self.write_to_disk = bool(write_to_disk)
        self.store_in_memory = bool(store_in_memory)
        self.extension = extension
        self.verbose = bool(verbose)
        self.data: Dict[int, Any] = {}
        if self.write_to_disk:
            self.get_data_path().mkdir(parents=True, exist_ok=True)
            if cleardir:
                for path in self.get_data_path().glob(f"{self.data_prefix}*{self.extension}"):
                    path.unlink()

    # paths
    def get_extension(self, wildcard: bool = False) -> str:
        return f"*{self.extension}" if wildcard else self.extension

    def get_data_path(self, index: Optional[int] = None) -> Path:
        base = self.data_directory / self.data_folder
        if index is None:
            return base
        if not isinstance(index, (int, np.integer)) or isinstance(index, bool):
            raise TypeError(f"index must be a single integer, got {index!r}")
        return base / f"{self.data_prefix}{int(index)}{self.extension}"

    # disk
    def write_to_file(self, data: Any, index: int) -> "ResultHandler":
        path = self.get_data_path(index)
        path.parent.mkdir(parents=True, exist_ok=True)
        arrays = encode(data)
        # write beside, then rename: a run killed mid-write leaves no torn step
        tmp = path.with_name(path.name + ".part")
        with tmp.open("wb") as stream:
            np.savez(stream, **arrays)
        tmp.replace(path)
        if self.verbose:
            print(f"writing data to {path}")
        return self

    def read_from_file(self, indices: int | Sequence[int]) -> List[Any]:
        if isinstance(indices, (int, np.integer)):
            indices = [indices]
        out = []
        root = self.get_data_path().resolve()
        for index in indices:
            path = self.get_data_path(index).resolve()
            if root not in path.parents:
                raise ValueError(f"{path} is outside the handler's directory {root}")
            if not path.exists():
                raise KeyError(f"no stored result for index {index}")
            with np.load(path, allow_pickle=False) as z:
                out.append(decode({k: z[k] for k in z.files}))
        return out

    def delete_file(self, index: int) -> "ResultHandler":
        path = self.get_data_path(index)
        if path.exists():
            path.unlink()
            if self.verbose:
                print(f"deleting data at {path}")
        return self

    # what is stored
    def get_valid_ids(self) -> List[int]:
        if self.write_to_disk:
            pattern = re.compile(rf"^{re.escape(self.data_prefix)}(\d+){re.escape(self.extension)}$")
            directory = self.get_data_path()
            if not directory.is_dir():
                return []
            return sorted(int(m.group(1)) for m in (pattern.match(p.name) for p in directory.iterdir()) if m)
        return sorted(self.data)

    def numel_data(self) -> int:
        ids = self.get_valid_ids()
        return max(ids) if ids else 0

    def size_data(self, dim: Optional[int] = None):
        shape = (self.numel_data(), 1)
        return shape if dim is None else shape[dim - 1]

    def __len__(self) -> int:
        return len(self.get_valid_ids())

    def __iter__(self) -> Iterator[Any]:
        for index in self.get_valid_ids():
            yield self[index]

    # element access
    def get_data(self, indices: Optional[Sequence[int]] = None) -> List[Any]:
        return [self[i] for i in (self.get_valid_ids() if indices is None else indices)]

    def __getitem__(self, index: int) -> Any:
        if self.store_in_memory and index in self.data:
            return self.data[index]
        if self.write_to_disk:
            return self.read_from_file(index)[0]
        raise KeyError(f"no stored result for index {index}")

    def __setitem__(self, index: int, value: Any) -> None:
        if not self.store_in_memory and not self.write_to_disk:
            raise ValueError("this handler neither stores in memory nor writes to disk, so the value "
                             "would be discarded; set store_in_memory or write_to_disk")
        if hasattr(value, "to_host") and callable(value.to_host):
            value = value.to_host()
        if self.store_in_memory:
            self.data[index] = value
        if self.write_to_disk:
            self.write_to_file(value, index)

    def __delitem__(self, index: int) -> None:
        self.data.pop(index, None)
        if self.write_to_disk:
            self.delete_file(index)

    def __contains__(self, index: int) -> bool:
        return index in self.get_valid_ids()

    def reset(self) -> "ResultHandler":
        for index in list(self.get_valid_ids()):
            del self[index]
        self.data.clear()
        return self

    def reset_data(self, subs=None) -> None:
        """Drop everything, or the ids (or the 1-based mask) in `subs`
        (result_handler.py:221-247)."""
        if subs is None:
            self.data.clear()
            if self.write_to_disk:
                for path in self.get_data_path().glob(f"{self.data_prefix}{self.get_extension(True)}"):
                    path.unlink()
            return
        subs = np.atleast_1d(np.asarray(subs))
        if subs.dtype == bool:
            subs = np.flatnonzero(subs) + 1
        for index in subs.astype(int).tolist():
            if self.store_in_memory:
                self.data.pop(index, None)
            if self.write_to_disk:
                self.delete_file(index)

    # searching
    def get_first_by_function(self, predicate: Callable[[Any], bool], reverse: bool = False):
        ids = self.get_valid_ids()
        for index in (ids[::-1] if reverse else ids):
            candidate = self[index]
            if predicate(candidate):
                return candidate, index
        return None, None

    def get_by_time(self, time: float, tol: float = 1e-3, reverse: bool = False):
        return self.get_first_by_function(lambda s: abs(_time_of(s) - time) < tol, reverse)

    def get_closest_by_time(self, time: float, reverse: bool = False):
        best = (None, None, float("inf"))
        ids = self.get_valid_ids()
        for index in (ids[::-1] if reverse else ids):
            candidate = self[index]
            d = abs(_time_of(candidate) - time)
            if d < best[2]:
                best = (candidate, index, d)
        return best

    def get_restart(self) -> Tuple[Any, int]:
        """(last stored state, the step to resume at); (None, 1) when empty."""
        ids = self.get_valid_ids()
        if not ids:
            return None, 1
        last = max(ids)
        return self[last], last + 1

    def __repr__(self) -> str:
        where = (["memory"] if self.store_in_memory else []) + (
            [str(self.get_data_path())] if self.write_to_disk else [])
        return f"ResultHandler({len(self)} result(s) in {' + '.join(where) or 'nothing'})"

    # MRST's names
    getExtension = get_extension
    getDataPath = get_data_path
    writeToFile = write_to_file
    readFromFile = read_from_file
    deleteFile = delete_file
    getValidIds = get_valid_ids
    numelData = numel_data
    sizeData = size_data
    getData = get_data
    resetData = reset_data
    getFirstByFunction = get_first_by_function
    getByTime = get_by_time
    getClosestByTime = get_closest_by_time
    getRestart = get_restart


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
    last, start = states.get_restart()
    if start - 1 > len(timesteps):
        raise ValueError(f"the store holds step {start - 1}, beyond the schedule's {len(timesteps)} steps")
    state = state0 if last is None else last
    remembers = callable(getattr(model, "run_memory", None))
    if remembers and reports is not None and start > 1:
        if (start - 1) not in reports:
            raise ValueError(f"the report store lacks step {start - 1}, which the state store holds")
        model.restore_run_memory(reports[start - 1].get("model_memory"))
    its = np.zeros(len(timesteps), dtype=np.int64)
    for i in range(start - 1, len(timesteps)):
        out, nit = simulate(state, model, [float(timesteps[i])], **kwargs)
        state = out[-1]
        its[i] = int(nit[0])
        states[i + 1] = state
        if reports is not None:
            rep = {"time": float(timesteps[:i + 1].sum()), "dt": float(timesteps[i]), "iterations": int(its[i])}
            if remembers:
                rep["model_memory"] = model.run_memory()
            reports[i + 1] = rep
    return [states[i] for i in range(1, len(timesteps) + 1)], its


__all__ = ["ResultHandler", "encode", "decode", "register", "simulate_with_output"]
