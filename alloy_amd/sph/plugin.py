"""Registers SolverDFSPH as the ``sph`` solver capability (alloy/solvers/sph/plugin.py:92-114).

The factory takes ``config`` (a DFSPHConfig) or config fields as kwargs plus ``capacity_per_world``; unknown kwargs
raise TypeError. The health adapter reports, per world, the last step's total pressure iterations and its mean
density error; the state adapter exposes ``alloy_state_arrays`` with its item -> world map.
"""

from __future__ import annotations

import dataclasses
import logging
from collections.abc import Mapping
from typing import Any

import numpy as np

from ..core.registry import (HealthSample, SolverCapability, register_health_adapter, register_solver,
                             register_state_adapter)
from .config import DFSPHConfig, register_custom_attributes

log = logging.getLogger("alloy.sph")

_CONFIG_FIELDS = frozenset(f.name for f in dataclasses.fields(DFSPHConfig))


def _split_kwargs(kwargs: Mapping[str, Any]) -> tuple[Any, Any, dict]:
    """(config, capacity_per_world, DFSPHConfig fields) of the factory's kwargs; a kwarg that is none of these raises
    TypeError naming it."""
    rest = dict(kwargs)
    cfg = rest.pop("config", None)
    cap = rest.pop("capacity_per_world", None)
    unknown = sorted(k for k in rest if k not in _CONFIG_FIELDS)
    if unknown:
        raise TypeError(f"sph factory: unknown kwargs {unknown}")
    return cfg, cap, rest


def _factory(view: Any, kwargs: Mapping[str, Any]):
    """Entry factory (plugin.py:36-65): ``config`` or DFSPHConfig fields, plus capacity_per_world.

    Fields given next to a ``config`` override it. A call with neither a ``config`` nor a ``world_bounds`` kwarg
    (the scene assembler's) takes its escape box from the scene's content (``escape_box.auto_box``) and keeps the
    box's padding on the solver; an explicit ``world_bounds``, None included, is used as given. Cloth absorption
    (plugin.py:46, 60-62) belongs to the cloth coupling, outside this path, so its kwarg is unknown here."""
    from . import escape_box
    from .solver import SolverDFSPH

    cfg, cap, fields = _split_kwargs(kwargs)
    pad = None
    if cfg is not None:
        if fields:
            cfg = dataclasses.replace(cfg, **fields)
    else:
        if "world_bounds" not in fields:
            box = escape_box.auto_box(view)
            bounds, pad = (None, None) if box is None else box
            fields["world_bounds"] = bounds
            log.info("sph factory: escape box %s from the scene's content (retired particles: report().escaped)",
                     bounds)
        cfg = DFSPHConfig(**fields)
    solver = SolverDFSPH(view, cfg, capacity_per_world=cap)
    if pad is not None:
        solver.escape_box_pad = pad  # an automatic box (a drive would widen it by this much)
    return solver


class _SPHEmitter:
    """Emitter protocol adapter (plugin.py:68-73): scene objects become liquid blocks."""

    def emit(self, builder: Any, obj: Any, decision: Any) -> None:
        from .build import add_fluid_block

        add_fluid_block(builder, decision.config, obj.lo, obj.counts)


def _health_adapter(solver, sample: HealthSample) -> None:
    """iters = iters_den + iters_div, residual = den_err per world (plugin.py:76-89)."""
    r = solver.report()
    sample.iters[:] = np.asarray(r.iters_density) + np.asarray(r.iters_divergence)
    sample.residual[:] = np.asarray(r.density_err)


CAPABILITY = SolverCapability(
    name="sph",
    factory=_factory,
    routes=frozenset({"liquid_sph"}),
    owns=frozenset({"particle", "body"}),
    coupling=True,
    uniform_worlds=False,
    per_contact_friction=False,
    determinism="native",
    friction_mix="none",
    restitution_path=False,
    emitter=_SPHEmitter(),
    register_attributes=register_custom_attributes,
)


def register() -> None:
    from .solver import SolverDFSPH

    register_solver(CAPABILITY)
    register_health_adapter(SolverDFSPH, _health_adapter)
    register_state_adapter(SolverDFSPH, lambda s: s.alloy_state_arrays(None), world_map=lambda s: s.alloy_state_worlds())


register()
