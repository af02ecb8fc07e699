"""Configuration, time-step helpers and the plug-in boundary (tests/solvers/sph/test_contract.py:28-136 of the
reference: capability, factory kwargs, closed forms of stable_dt / snap_down, config validation)."""

from __future__ import annotations

import dataclasses

import pytest

from alloy_amd.core.registry import get_solver, solvers_for_route
from alloy_amd.sph import plugin
from alloy_amd.sph.config import DFSPHConfig, snap_down, stable_dt
from alloy_amd.sph.scenes import dam_break


def test_defaults_and_derived():
    cfg = DFSPHConfig()
    assert cfg.h == pytest.approx(8e-3) and cfg.spacing == pytest.approx(4e-3)
    assert cfg.max_density_error == 1e-3 and cfg.pressure_solver == "cg" and cfg.neighbor_capacity == 64
    fast = DFSPHConfig.fast(particle_radius=0.005)
    assert fast.rest_jitter_speed is None and fast.cfl == 0.4 and fast.particle_radius == 0.005


@pytest.mark.parametrize("field,value,msg", [
    ("particle_radius", 0.0, "particle_radius must be > 0"),
    ("support_factor", 1.5, "support_factor"),
    ("min_iterations", 0, "min_iterations"),
    ("jacobi_omega", 0.8, "jacobi_omega"),
    ("pressure_solver", "sor", "pressure_solver"),
    ("rest_jitter_flow_ratio", 1.5, "rest_jitter_flow_ratio"),
    ("diagonal_floor", 1.0, "diagonal_floor"),
])
def test_validate_rejects(field, value, msg):
    cfg = dataclasses.replace(DFSPHConfig(), **{field: value})
    with pytest.raises(ValueError, match=msg):
        cfg.validate()


def test_stable_dt_and_snap_down_closed_forms():
    cfg = DFSPHConfig(particle_radius=1e-3)
    assert stable_dt(cfg, 0.0) == cfg.dt_max
    assert stable_dt(cfg, float("inf")) == cfg.dt_min
    assert stable_dt(cfg, 1.0) == pytest.approx(0.2 * 2e-3)
    assert stable_dt(cfg, 1e6) == cfg.dt_min
    assert snap_down(1e-3, 0.02) == pytest.approx(0.02 / 32)
    assert snap_down(0.02, 0.02) == 0.02
    assert snap_down(1e-9, 0.02) == pytest.approx(0.02 / 1024)


def test_capability_registration():
    cap = get_solver("sph")
    assert cap is plugin.CAPABILITY and "liquid_sph" in cap.routes
    assert cap in solvers_for_route("liquid_sph")
    assert cap.uniform_worlds is False and cap.coupling and cap.determinism == "native"
    plugin.register()  # idempotent


def test_factory_rejects_unknown_kwargs():
    sc = dam_break(4, 0.1)
    with pytest.raises(TypeError, match="unknown kwargs"):
        plugin._factory(sc.model, {"nonsense": 1})


def test_solver_has_no_cpu_path():
    """The product path runs only on a HIP device: a CPU model is an error, not a fallback."""
    from alloy_amd.sph.solver import SolverDFSPH

    sc = dam_break(4, 0.1)
    with pytest.raises(RuntimeError, match="HIP device"):
        SolverDFSPH(sc.model, sc.config)


@pytest.mark.parametrize("overrides", [{"surface_tension": "akinci2013"}, {"ambient_pressure": 100.0}])
def test_unbuilt_subsystems_are_refused(overrides):
    from alloy_amd.sph.solver import SolverDFSPH

    sc = dam_break(4, 0.1)
    cfg = dataclasses.replace(sc.config, **overrides)
    with pytest.raises(ValueError, match="not built"):
        SolverDFSPH(sc.model, cfg)


def test_auto_escape_box_covers_the_content_with_its_pad():
    """test_escape_box.py:46-53 of the reference: the plate's bounding sphere and the block, padded by 0.25 m; the
    ground plane is left out; an empty model has no box."""
    import numpy as np

    from alloy_amd.sph import escape_box
    from alloy_amd.sph.build import add_fluid_block, wall_box
    from alloy_amd.sph.model import GeoType, ModelBuilder

    cfg = DFSPHConfig(particle_radius=2.0e-3)
    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    b.add_shape(GeoType.PLANE, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0))
    b.begin_world()
    wall_box(b, (0.0, 0.0, 0.095), (0.05, 0.05, 0.005))
    add_fluid_block(b, cfg, (-0.008, -0.008, 0.1), (4, 4, 4))
    b.end_world()
    m = b.finalize("cpu")
    lo, hi = escape_box.auto_world_bounds(m)
    rp = float(np.linalg.norm([0.05, 0.05, 0.005]))
    assert np.allclose(lo, np.array([-rp, -rp, 0.095 - rp]) - 0.25, atol=1e-6)
    assert np.allclose(hi, np.array([rp, rp, 0.095 + rp]) + 0.25, atol=1e-6)
    assert escape_box.auto_world_bounds(ModelBuilder().finalize("cpu")) is None


@pytest.mark.parametrize("skin, ok", [(None, True), (0.0, True), (0.1, True), (-0.1, False), (1.5, False)])
def test_skin_config_range(skin, ok):
    """test_neighbors_skin.py:97-105 of the reference."""
    cfg = DFSPHConfig(reneighbor_skin=skin)
    if ok:
        cfg.validate()
    else:
        with pytest.raises(ValueError, match="reneighbor_skin"):
            cfg.validate()


def test_skin_capacity_is_the_padded_rest_fill():
    """neighbors_skin.py:146-158: 1.75 x the interior lattice neighbours within h + skin (h = 2 spacings: the 26
    interior neighbours of config.py:230-236; 32 within 2.2 spacings), at least 16 and, here, at most 64."""
    from alloy_amd.sph._abi import skin_capacity

    assert skin_capacity(2.0, 1.0, 0.0) == 46        # ceil(1.75 x 26)
    assert skin_capacity(2.0, 1.0, 0.2) == 56        # ceil(1.75 x 32)
    assert skin_capacity(2.0, 1.0, 2.0) == 64        # 4 spacings: ceil(1.75 x 250) held to 64
    assert skin_capacity(0.5, 1.0, 0.0) == 16        # no lattice neighbour within half a spacing
