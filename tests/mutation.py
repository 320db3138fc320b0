"""Planted mistakes for the oracle's checks (test infrastructure).

`planted(obj, name, (old, new), ...)` returns a context manager that
replaces the function or method `obj.name` by a copy whose source has each
`old` text replaced by `new` (each `old` must occur exactly once), compiled
in the original's module namespace.  A pinning check run inside it must
fail: that shows the check can see the mistake, i.e. it is not vacuous.
"""
import contextlib
import inspect
import textwrap


def _mutant(func, edits):
    src = textwrap.dedent(inspect.getsource(func))
    for old, new in edits:
        n = src.count(old)
        if n != 1:
            raise ValueError(f"planted edit must match once, matched {n}: {old!r}")
        src = src.replace(old, new)
    ns = {}
    exec(compile(src, inspect.getsourcefile(func) or "<planted>", "exec"), func.__globals__, ns)
    return ns[func.__name__]


@contextlib.contextmanager
def planted(obj, name, *edits):
    orig = inspect.getattr_static(obj, name)
    func = orig.__func__ if isinstance(orig, (staticmethod, classmethod)) else orig
    new = _mutant(func, edits)
    if isinstance(orig, staticmethod):
        new = staticmethod(new)
    elif isinstance(orig, classmethod):
        new = classmethod(new)
    setattr(obj, name, new)
    try:
        yield
    finally:
        setattr(obj, name, orig)


def fails(check, *args, **kwargs):
    """True when check(*args) raises AssertionError (or any numerical error a
    broken oracle can raise on the way: a failed Newton run, a singular
    matrix)."""
    try:
        check(*args, **kwargs)
    except (AssertionError, ArithmeticError, RuntimeError, ValueError, IndexError):
        return True
    return False
