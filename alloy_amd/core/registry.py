"""Solver capability registry: the plug-in boundary the router and the scene stepper call through.

Restates the parts of alloy/core/registry.py:34-114 the DFSPH plug-in registers into: ``SolverCapability`` (name,
factory, routes, ownership and determinism flags), idempotent ``register_solver`` (an equal capability re-registers
as a no-op, a different one under the same name is an error), the health adapter and the state adapter with its
world map.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Callable


@dataclass(frozen=True)
class SolverCapability:
    name: str
    factory: Callable[[Any, dict], Any]
    routes: frozenset = field(default_factory=frozenset)
    owns: frozenset = field(default_factory=frozenset)
    coupling: bool = False
    uniform_worlds: bool = True
    per_contact_friction: bool = False
    determinism: str = "native"
    friction_mix: str = "none"
    restitution_path: bool = False
    emitter: Any = None
    register_attributes: Callable | None = None


@dataclass
class HealthSample:
    """Per-world health inputs a solver's health adapter fills (iterations and residual)."""

    iters: Any
    residual: Any


_SOLVERS: dict[str, SolverCapability] = {}
_HEALTH: dict[type, Callable] = {}
_STATE: dict[type, tuple[Callable, Callable | None]] = {}


def register_solver(cap: SolverCapability) -> None:
    old = _SOLVERS.get(cap.name)
    if old is not None and old != cap:
        raise ValueError(f"solver capability {cap.name!r} is already registered with a different definition")
    _SOLVERS[cap.name] = cap


def get_solver(name: str) -> SolverCapability:
    return _SOLVERS[name]


def solvers_for_route(route: str) -> list[SolverCapability]:
    return [c for c in _SOLVERS.values() if route in c.routes]


def register_health_adapter(cls: type, fn: Callable) -> None:
    _HEALTH[cls] = fn


def health_adapter(cls: type) -> Callable | None:
    return _HEALTH.get(cls)


def register_state_adapter(cls: type, fn: Callable, world_map: Callable | None = None) -> None:
    _STATE[cls] = (fn, world_map)


def state_adapter(cls: type) -> tuple[Callable, Callable | None] | None:
    return _STATE.get(cls)
