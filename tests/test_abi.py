"""CPU: libcfdhip.so loads and exports every function include/cfdhip.h
declares; the ctypes table matches the header; API-level validation that
runs before any device call."""

import ctypes
import re
from pathlib import Path

import pytest

from cfd_demo_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    text = (ROOT / "include" / "cfdhip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(cfd_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    names = header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_covers_the_header():
    assert set(header_functions()) == set(_lib.SIGNATURES)


def test_solver_config_validation():
    from cfd_demo_amd import SolverConfig
    with pytest.raises(ValueError, match="Unknown interface flux"):
        SolverConfig(flux="godunov", limiter="minmod", gamma=1.4)
    with pytest.raises(ValueError, match="Unknown limiter"):
        SolverConfig(flux="rusanov", limiter="ospre", gamma=1.4)
    with pytest.raises(ValueError, match="gamma"):
        SolverConfig(flux="rusanov", limiter="minmod", gamma=1.0)
    with pytest.raises(ValueError, match="cfl"):
        SolverConfig(flux="rusanov", limiter="minmod", gamma=1.4, cfl=0.0)


def test_bad_state_shape_rejected_before_device_use():
    from cfd_demo_amd import EulerSolver, FixedState, SolverConfig, rectangle_mesh
    import numpy as np
    c = FixedState.from_primitive(1.2, 0.4, -0.3, 0.9, 1.4)
    s = EulerSolver(rectangle_mesh(2, 2), {t: c for t in ("left", "right", "bottom", "top")},
                    SolverConfig(flux="rusanov", limiter="minmod", gamma=1.4))
    with pytest.raises(ValueError, match="Expected a state of shape"):
        s.rhs(np.zeros((3, 4)))


def test_boundary_condition_errors():
    from cfd_demo_amd import EulerSolver, SlipWall, SolverConfig, rectangle_mesh
    cfg = SolverConfig(flux="rusanov", limiter="minmod", gamma=1.4)
    with pytest.raises(ValueError, match="without a boundary condition"):
        EulerSolver(rectangle_mesh(2, 2), {"left": SlipWall()}, cfg)
    with pytest.raises(ValueError, match="not in the mesh"):
        EulerSolver(rectangle_mesh(2, 2), {t: SlipWall() for t in ("left", "right", "bottom", "top", "x")}, cfg)
