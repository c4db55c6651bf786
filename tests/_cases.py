"""Build product objects (Grid, Medium, CWSource, CWRunSpec) from the golden fixtures."""

from __future__ import annotations

import numpy as np

from tests._golden import cw_case


def product_case(name: str):
    """(fixture, grid, medium, source, spec, run kwargs, solver name, oracle args)."""
    from caustica_amd import CWSource, Grid, Medium, PMLSpec
    from caustica_amd.solvers import CWRunSpec

    g, args = cw_case(name)
    m = g["meta"]
    grid = Grid(tuple(m["shape"]), m["dx"], PMLSpec(thickness=m["pml_thickness"], edge=m["pml_edge"]))
    medium = Medium(alpha=g["alpha"], rho=g["rho"], c=g["c"], beta=g["beta"])
    source = CWSource(indices=g["src_indices"].astype(np.int64), phases=g["src_phases"],
                      amplitude=m["amplitude"], f0=m["f0"], ramp_periods=m["ramp_periods"])
    spec = CWRunSpec(**m["spec"])
    kw = dict(harmonics=tuple(m["harmonics"]), record_region=args["record_region"],
              reference_point=args["reference_point"])
    return g, grid, medium, source, spec, kw, m["solver"], args
