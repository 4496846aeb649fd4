"""Theory registry: names and aliases -> theory packages.

Same contract as the reference's theories/registry.py: ``TheorySpec``
(:21-56), ``register_theory`` (:62-96, duplicate keys raise ValueError),
``list_theories`` (:99-113), ``get_theory_spec`` (:116-144, KeyError for an
unknown name), ``load_theory`` (:176-204, imports the package and its required
submodules), ``describe_theory`` / ``print_theory_table`` and
``discover_theories`` (:284-330, first registration of a key wins).  Lookup
is case-insensitive on stripped names.

Unlike the reference, discovery does not swallow import errors silently: a
theory that fails to import is recorded in ``DISCOVERY_ERRORS``.
"""
from __future__ import annotations

import importlib
import pkgutil
from dataclasses import dataclass
from types import ModuleType
from typing import Dict


@dataclass(frozen=True)
class TheorySpec:
    name: str
    import_path: str
    description: str = ""
    version: str = "0.0"
    aliases: tuple[str, ...] = ()
    required_submodules: tuple[str, ...] = ("kernels", "initial_config", "observables")


_REGISTRY: Dict[str, TheorySpec] = {}
DISCOVERY_ERRORS: Dict[str, str] = {}


def _key(name) -> str:
    return (name or "").strip().lower()


def register_theory(spec: TheorySpec) -> None:
    keys = [spec.name, *spec.aliases]
    for k in keys:
        key = _key(k)
        if not key:
            raise ValueError("TheorySpec names must be non-empty")
        if key in _REGISTRY:
            raise ValueError(f"Theory '{key}' is already registered")
    for k in keys:
        _REGISTRY[_key(k)] = spec


def list_theories() -> tuple[str, ...]:
    return tuple(sorted({spec.name for spec in _REGISTRY.values()}))


def get_theory_spec(name: str) -> TheorySpec:
    key = _key(name)
    if key not in _REGISTRY:
        raise KeyError(f"Unknown theory '{name}'. Available: {list_theories()}")
    return _REGISTRY[key]


def _validate_theory_package(spec: TheorySpec) -> None:
    for sub in spec.required_submodules:
        importlib.import_module(f"{spec.import_path}.{sub}")


def load_theory(name: str) -> ModuleType:
    spec = get_theory_spec(name)
    mod = importlib.import_module(spec.import_path)
    _validate_theory_package(spec)
    return mod


def _spec_lines(spec: TheorySpec) -> list[tuple[str, str]]:
    """(label, value) pairs of a spec, empty values left out."""
    rows = [("Theory", spec.name), ("Import path", spec.import_path), ("Version", spec.version),
            ("Aliases", ", ".join(spec.aliases)), ("Description", spec.description)]
    return [(k, v) for k, v in rows if v]


def describe_theory(name: str) -> None:
    """Print one registered theory, one labelled field per line."""
    for label, value in _spec_lines(get_theory_spec(name)):
        print(f"{label}: {value}")


def print_theory_table() -> None:
    """Print every registered theory once (aliases folded into their spec) as
    a table of name, version and description, sorted by name."""
    specs = sorted({id(s): s for s in _REGISTRY.values()}.values(), key=lambda s: s.name.lower())
    if not specs:
        print("No theories registered.")
        return
    cols = ("Theory", "Version", "Description")
    rows = [(s.name, s.version, s.description) for s in specs]
    widths = [max(len(r[i]) for r in [cols, *rows]) for i in range(2)]
    fmt = lambda r: "  ".join([r[0].ljust(widths[0]), r[1].ljust(widths[1]), r[2]])  # noqa: E731
    head = fmt(cols)
    print("\n".join([head, "-" * len(head), *map(fmt, rows)]))


def discover_theories(package: str = "soliton_solver_amd.theories") -> None:
    pkg = importlib.import_module(package)
    for m in pkgutil.iter_modules(pkg.__path__):
        if not m.ispkg or m.name == "__pycache__":
            continue
        path = f"{package}.{m.name}"
        try:
            sub = importlib.import_module(path)
        except Exception as exc:  # recorded, not hidden
            DISCOVERY_ERRORS[path] = repr(exc)
            continue
        spec = getattr(sub, "THEORY_SPEC", None)
        if isinstance(spec, TheorySpec):
            for k in (spec.name, *spec.aliases):
                key = _key(k)
                if key and key not in _REGISTRY:
                    _REGISTRY[key] = spec
