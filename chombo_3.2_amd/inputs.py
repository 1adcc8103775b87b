"""ParmParse-style inputs for AMR Poisson solves, as the reference's driver reads them
(releasedExamples/AMRPoisson/execCell/poissonSolve.cpp with its `inputs` file).

* ``ParmParse``: ``prefix.key = value value ...`` lines; text after ``#`` is a comment;
  a later definition of a key replaces an earlier one, and ``key=value`` arguments
  given after the file replace both (ParmParse(argc, argv, NULL, file)).  ``get`` /
  ``getarr`` fail on a missing key, ``query`` returns a default, as ParmParse's do.
* ``setup_grids``: setupGrids (poissonSolve.cpp:240-374): base grids from
  domainSplit(num_cells domain, max_box_size, block_factor); the finer levels from
  ``grids.level_<l>_box_<b>_lo / _hi`` (``grids.read_in_grids = true``); refinement
  ratios ``grids.ref_ratio``; ``grids.is_periodic``; dx = domain_size[0] / num_cells[0].
  Grids made by tagging (BRMeshRefine on the gradient of the rhs, read_in_grids false)
  are not built here and are rejected.  The driver's ParseBC applies DiriBC with the
  homogeneous linear rule on every non-periodic face (poissonSolve.cpp:77-118), and
  alpha = 0, beta = 1 (:551-552).
* ``solver_params``: setupSolver (:540-586): num_smooth for pre, post and bottom,
  num_mg, max_iterations, tolerance, hang, normThresh 1e-30, then the optional
  num_pre / num_post / num_bottom; imin is AMRMultiGrid's default, 5.
* ``gaussians_rhs``: setRHS with the driver's default problem type (three Gaussian
  sources, :169-222).
* ``run``: runSolver (:588-633): ``main.iterations`` solves from a zero initial guess;
  with ``main.writePlotFiles`` (default true) the plotfile of :640-696, components phi,
  rhs and res (the composite residual) written by WriteAMRHierarchyHDF5 with dt and time
  1, as ``poissonOut.3d.hdf5`` (chombo_hdf5.py, no HDF5 library needed).

Run as a program: ``python inputs.py <inputs file> [key=value ...]``.
"""
from __future__ import annotations

import shlex
from typing import Dict, List, Optional, Sequence

import numpy as np

import hierarchy as H


class ParmParse:
    """The key/value table of an inputs file (all prefixes)."""

    def __init__(self, text: str = "", argv: Sequence[str] = ()):
        self.table: Dict[str, List[str]] = {}
        for line in text.splitlines():
            self._define(line)
        for arg in argv:
            self._define(arg)

    @classmethod
    def from_file(cls, path: str, argv: Sequence[str] = ()):
        with open(path) as f:
            return cls(f.read(), argv)

    def _define(self, line: str):
        line = line.split("#", 1)[0].strip()
        if not line:
            return
        if "=" not in line:
            raise ValueError(f"ParmParse: no '=' in {line!r}")
        key, val = line.split("=", 1)
        key = key.strip()
        if not key or any(c.isspace() for c in key):
            raise ValueError(f"ParmParse: bad key in {line!r}")
        self.table[key] = shlex.split(val)

    def contains(self, key: str) -> bool:
        return key in self.table

    def _vals(self, key: str) -> List[str]:
        if key not in self.table:
            raise KeyError(f"ParmParse: {key} not found")
        return self.table[key]

    @staticmethod
    def _conv(s: str, kind):
        if kind is bool:
            if s.lower() in ("true", "1"):
                return True
            if s.lower() in ("false", "0"):
                return False
            raise ValueError(f"ParmParse: {s!r} is not a bool")
        return kind(s)

    def get(self, key: str, kind=str):
        return self._conv(self._vals(key)[0], kind)

    def getarr(self, key: str, kind, n: int):
        v = self._vals(key)
        if len(v) < n:
            raise ValueError(f"ParmParse: {key} has {len(v)} values, {n} needed")
        return [self._conv(x, kind) for x in v[:n]]

    def query(self, key: str, kind=str, default=None):
        return self.get(key, kind) if key in self.table else default


def setup_grids(pp: ParmParse) -> H.Hierarchy:
    max_level = pp.get("grids.max_level", int)
    max_box = pp.get("grids.max_box_size", int)
    block = pp.get("grids.block_factor", int)
    ratios = pp.getarr("grids.ref_ratio", int, max_level) if max_level > 0 else []
    periodic = [v == 1 for v in pp.getarr("grids.is_periodic", int, 3)]
    ncells = pp.getarr("grids.num_cells", int, 3)
    size = pp.getarr("grids.domain_size", float, 3) if pp.contains("grids.domain_size") else [1.0, 1.0, 1.0]
    dx0 = size[0] / ncells[0]
    dom = H.box((0, 0, 0), tuple(n - 1 for n in ncells))
    levels = [H.domain_split(dom, max_box, block)]
    if max_level > 0:
        if not pp.query("grids.read_in_grids", bool, False):
            raise ValueError("grids made by tagging (BRMeshRefine, read_in_grids = false) are not supported: "
                             "give the boxes of every level (read_in_grids = true)")
        ldom = dom
        for lev in range(1, max_level + 1):
            ldom = H.box_refine(ldom, ratios[lev - 1])
            boxes = []
            for b in range(pp.get(f"grids.level_{lev}_box_count", int)):
                lo = tuple(pp.getarr(f"grids.level_{lev}_box_{b}_lo", int, 3))
                hi = tuple(pp.getarr(f"grids.level_{lev}_box_{b}_hi", int, 3))
                if any(lo[d] < ldom[0][d] or hi[d] > ldom[1][d] for d in range(3)):
                    raise ValueError("box outside of domain")
                boxes.append((lo, hi))
            levels.append(boxes)
    bc = tuple(3 if p else 1 for p in periodic)
    return H.Hierarchy(dom, levels, dx0, 0.0, 1.0, bc_type_lo=bc, bc_type_hi=bc,
                       ratios=tuple(ratios) if any(r != 2 for r in ratios) else None)


def solver_params(pp: ParmParse):
    ns = pp.get("solver.num_smooth", int)
    sp = dict(pre=ns, post=ns, bottom=ns, numMG=pp.get("solver.num_mg", int),
              iterMax=pp.get("solver.max_iterations", int), imin=5,
              eps=pp.get("solver.tolerance", float), hang=pp.get("solver.hang", float), normThresh=1.0e-30)
    sp["pre"] = pp.query("solver.num_pre", int, sp["pre"])
    sp["post"] = pp.query("solver.num_post", int, sp["post"])
    sp["bottom"] = pp.query("solver.num_bottom", int, sp["bottom"])
    ip = np.asarray([sp["pre"], sp["post"], sp["bottom"], sp["numMG"], sp["iterMax"], sp["imin"]], np.int32)
    rp = np.asarray([sp["eps"], sp["hang"], sp["normThresh"]], np.float64)
    return ip, rp


# the three sources of setRHS's `gaussians` problem: strength, scale, centre
GAUSSIANS = [(1.0, 1.0e-2, (0.25, 0.25, 0.25)), (3.0, 1.0e-2, (0.5, 0.75, 0.75)), (2.0, 1.0e-2, (0.75, 0.5, 0.5))]


def gaussians_rhs(h: H.Hierarchy) -> np.ndarray:
    """setRHS (gaussians): per cell, from 0, each source's strength * exp(-r^2 / scale)
    added in order, at loc = iv * dx + 0.5 * dx (the driver's order of operations);
    valid-packed."""
    parts = []
    for lev in range(h.nlev):
        dx = h.dx(lev)
        for (lo, hi) in h.levels[lev]:
            ax = [np.arange(lo[d], hi[d] + 1) * dx + 0.5 * dx for d in range(3)]
            x, y, z = np.meshgrid(*ax, indexing="ij")
            v = np.zeros_like(x)
            for s, w, c in GAUSSIANS:
                d0, d1, d2 = x - c[0], y - c[1], z - c[2]
                r2 = d0 * d0 + d1 * d1 + d2 * d2
                v = v + s * np.exp(-r2 / w)
            parts.append(v.ravel(order="F"))
    return np.concatenate(parts)


def run(pp: ParmParse, ctx, rhs: Optional[np.ndarray] = None, plotfile: Optional[str] = None):
    """runSolver: the hierarchy, solver and rhs the inputs describe, then
    ``main.iterations`` solves from zero, and the plotfile when ``plotfile`` is given.
    Returns (hierarchy, phi, history, info) of the last solve."""
    import amrmg

    h = setup_grids(pp)
    sp, rp = solver_params(pp)
    if rhs is None:
        rhs = gaussians_rhs(h)
    s = amrmg.AMRMultiGrid(ctx, h)
    try:
        s.setSolverParameters(sp, rp)
        s.set_field(s.RHS, rhs)
        hist, info = None, None
        for _ in range(max(1, pp.query("main.iterations", int, 1))):
            hist, info = s.solve(zero_phi=True)
        phi = s.get_field(s.PHI)
        if plotfile:
            import chombo_hdf5 as CH

            res, _ = s.computeAMRResidual()
            CH.write_amr_hierarchy(plotfile, CH.file_from_hierarchy(h, {"phi": phi, "rhs": rhs, "res": res},
                                                                    dt=1.0, time=1.0))
    finally:
        s.close()
    return h, phi, hist, info


def main(argv: Sequence[str]) -> int:
    import amrmg

    if len(argv) < 1:
        print("usage: inputs.py <inputs file> [key=value ...]")
        return 2
    pp = ParmParse.from_file(argv[0], argv[1:])
    plot = "poissonOut.3d.hdf5" if pp.query("main.writePlotFiles", bool, True) else None
    ctx = amrmg.Context(0)
    try:
        h, phi, hist, info = run(pp, ctx, plotfile=plot)
    finally:
        ctx.close()
    print(f"{h.nlev} levels, {h.total_cells()} cells: {info['iterations']} iterations, exit status "
          f"{info['exit_status']}, residual {hist[0]:.3e} -> {hist[-1]:.3e}" + (f", wrote {plot}" if plot else ""))
    return 0 if info["exit_status"] & 1 else 1


if __name__ == "__main__":
    import sys

    sys.exit(main(sys.argv[1:]))
