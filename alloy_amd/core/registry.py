"""Solver capability registry: the plug-in boundary the router and the scene stepper call through.

Restates the parts of alloy/core/registry.py:34-215 the DFSPH plug-in registers into and its callers look up:
``SolverCapability`` (name, factory, routes, ownership and determinism flags), ``register_solver`` (an equal
capability re-registers as a no-op, a different one under the same name raises ``RegistryError`` unless
``replace``), route and name lookups, and the per-solver-class health and state adapters (looked up along the
instance's MRO).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Callable


class RegistryError(ValueError):
    """A malformed or conflicting registration, or a lookup of something never registered."""


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
    solver_cls: type | None = None

    def __post_init__(self):
        if not isinstance(self.name, str) or not self.name:
            raise RegistryError(f"SolverCapability.name must be a non-empty str, got {self.name!r}")


@dataclass
class HealthSample:
    """Per-world health inputs a solver's health adapter fills (iterations and residual)."""

    iters: Any
    residual: Any


_SOLVERS: dict[str, SolverCapability] = {}
_HEALTH: dict[type, Callable] = {}
_STATE: dict[type, Callable] = {}
_WORLD_MAPS: dict[type, Callable] = {}


def register_solver(cap: SolverCapability, *, replace: bool = False) -> None:
    """registry.py:106-113."""
    old = _SOLVERS.get(cap.name)
    if old is not None and old != cap and not replace:
        raise RegistryError(f"solver capability {cap.name!r} is already registered with different values")
    _SOLVERS[cap.name] = cap


def solver(name: str) -> SolverCapability:
    """The capability registered as ``name`` (registry.py:116-121)."""
    try:
        return _SOLVERS[name]
    except KeyError:
        raise RegistryError(f"no solver capability {name!r}; registered: {sorted(_SOLVERS)}") from None


get_solver = solver


def solvers_for_route(route: str) -> list[SolverCapability]:
    """Capabilities serving ``route``, sorted by name (registry.py:124-127)."""
    route = str(route)
    return sorted((c for c in _SOLVERS.values() if route in c.routes), key=lambda c: c.name)


def registered_solvers() -> tuple[str, ...]:
    return tuple(sorted(_SOLVERS))


def registered_routes() -> frozenset[str]:
    return frozenset(r for c in _SOLVERS.values() for r in c.routes)


def _check_cls(solver_cls: type, what: str) -> None:
    if not isinstance(solver_cls, type):
        raise RegistryError(f"{what} must be registered for a class, got {solver_cls!r}")


def _lookup(table: dict, obj: object):
    for klass in type(obj).__mro__:
        if klass in table:
            return table[klass]
    return None


def capability_for(solver_obj: object) -> SolverCapability | None:
    """The capability whose solver_cls is nearest in the instance's MRO (registry.py:140-150)."""
    by_cls = {}
    for cap in sorted(_SOLVERS.values(), key=lambda c: c.name, reverse=True):
        if cap.solver_cls is not None:
            by_cls[cap.solver_cls] = cap
    return _lookup(by_cls, solver_obj)


def register_health_adapter(solver_cls: type, fn: Callable) -> None:
    _check_cls(solver_cls, "a health adapter")
    _HEALTH[solver_cls] = fn


def health_adapter_for(solver_obj: object) -> Callable | None:
    return _lookup(_HEALTH, solver_obj)


def register_state_adapter(solver_cls: type, fn: Callable, world_map: Callable | None = None) -> None:
    """The solver's internal arrays outside State and, optionally, their item -> world maps (registry.py:191-200)."""
    _check_cls(solver_cls, "a state adapter")
    _STATE[solver_cls] = fn
    if world_map is not None:
        _WORLD_MAPS[solver_cls] = world_map
    else:
        _WORLD_MAPS.pop(solver_cls, None)


def state_arrays_for(solver_obj: object) -> dict:
    fn = _lookup(_STATE, solver_obj)
    return {} if fn is None else dict(fn(solver_obj))


def state_world_maps_for(solver_obj: object) -> dict:
    fn = _lookup(_WORLD_MAPS, solver_obj)
    return {} if fn is None else dict(fn(solver_obj))
