"""Theory plugins, discovered at import like the reference's theories/__init__.py:16-20."""
from __future__ import annotations

from .registry import (TheorySpec, describe_theory, discover_theories, get_theory_spec, list_theories,
                       load_theory, print_theory_table, register_theory)

_discovered = False
if not _discovered:
    discover_theories()
    _discovered = True


def list() -> tuple[str, ...]:  # noqa: A001 - same name as the reference's theories.list
    return list_theories()


def print_description(name: str) -> None:
    describe_theory(name)


def print_table() -> None:
    print_theory_table()
