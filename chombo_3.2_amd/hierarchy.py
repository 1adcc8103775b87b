"""Box-layout construction for the AMR Poisson/Helmholtz hot path.

Host-side index math only (no arithmetic on field data): Chombo-style boxes,
``domainSplit`` (lib/src/BoxTools/BRMeshRefine.cpp:983-1050), the
``DisjointBoxLayout`` sort order (``Box::operator<``, lib/src/BoxTools/Box.H:2223),
and the synthetic workloads the solver is measured on:

* single level, 64^3, 32^3 boxes, Dirichlet 0, rhs = sin(2 pi x) sin(2 pi y)
  sin(2 pi z) + N(0, 1e-3)  (``single_level``)
* three AMR levels, base 256^3, refinement ratio 2, centred refined regions
  covering 25 % then 6 % of the parent domain, 64^3 maximum boxes, Gaussian-bump
  rhs (``amr_hierarchy``)

A hierarchy is handed to both the HIP library and the CPU oracle as the same
flat integer/real description (``Hierarchy.spec()``), documented in
include/amrmg.h.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

Box = Tuple[Tuple[int, int, int], Tuple[int, int, int]]


def box(lo, hi) -> Box:
    return (tuple(int(v) for v in lo), tuple(int(v) for v in hi))


def box_npts(b: Box) -> int:
    (l, h) = b
    return (h[0] - l[0] + 1) * (h[1] - l[1] + 1) * (h[2] - l[2] + 1)


def box_shape(b: Box) -> Tuple[int, int, int]:
    (l, h) = b
    return (h[0] - l[0] + 1, h[1] - l[1] + 1, h[2] - l[2] + 1)


def box_refine(b: Box, r: int) -> Box:
    (l, h) = b
    return (tuple(v * r for v in l), tuple((v + 1) * r - 1 for v in h))


def box_coarsen(b: Box, r: int) -> Box:
    (l, h) = b
    return (tuple(v // r for v in l), tuple(v // r for v in h))


def box_grow(b: Box, g: int) -> Box:
    (l, h) = b
    return (tuple(v - g for v in l), tuple(v + g for v in h))


def sort_key(b: Box):
    """Key reproducing Box::operator< : smallEnd lexicographic (dim 0 most
    significant), ties broken by bigEnd."""
    return (b[0], b[1])


def domain_split(domain: Box, max_box: int, block_factor: int) -> List[Box]:
    """Restatement of domainSplit(const Box&, Vector<Box>&, int, int, IntVect)
    (lib/src/BoxTools/BRMeshRefine.cpp:983-1050) with all directions refined."""
    if max_box == 0:
        return [domain]
    ratio = max_box // block_factor
    lo, hi = domain
    d_lo = [v // block_factor for v in lo]
    d_hi = [v // block_factor for v in hi]
    for dd in range(3):
        if d_lo[dd] * block_factor != lo[dd] or (d_hi[dd] + 1) * block_factor - 1 != hi[dd]:
            raise ValueError("domainSplit: domain not coarsenable by blocking factor")
    count = [1, 0, 0, 0]
    for i in range(3):
        c = (d_hi[i] - d_lo[i] + ratio) // ratio
        count[i + 1] = count[i] * c
    out = []
    for b in range(count[3]):
        L = list(d_lo)
        H = list(d_hi)
        for i in range(3):
            m = (b // count[i]) % (count[i + 1] // count[i])
            L[i] += m * ratio
            H[i] = min(d_hi[i], L[i] + ratio - 1)
        out.append(box_refine((tuple(L), tuple(H)), block_factor))
    return out


@dataclass
class Hierarchy:
    """An AMR hierarchy: level-0 domain, sorted boxes per level, refinement ratio
    ratios[l-1] (2 or 4) between levels l-1 and l (None: 2 everywhere)."""

    domain: Box
    levels: List[List[Box]]
    dx0: float
    alpha: float = 0.0
    beta: float = 1.0
    bc_type_lo: Tuple[int, int, int] = (1, 1, 1)  # 0 Neumann, 1 Dirichlet, 2 Dirichlet (quadratic), 3 periodic
    bc_type_hi: Tuple[int, int, int] = (1, 1, 1)
    bc_val_lo: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    bc_val_hi: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    ratios: Optional[Tuple[int, ...]] = None

    def __post_init__(self):
        self.levels = [sorted(lv, key=sort_key) for lv in self.levels]

    @property
    def nlev(self) -> int:
        return len(self.levels)

    def ratio(self, lev: int) -> int:
        """Refinement ratio between levels lev-1 and lev (lev >= 1)."""
        return 2 if self.ratios is None else int(self.ratios[lev - 1])

    def dx(self, lev: int) -> float:
        d = self.dx0
        for l in range(1, lev + 1):
            d = d / self.ratio(l)
        return d

    def level_domain(self, lev: int) -> Box:
        b = self.domain
        for l in range(1, lev + 1):
            b = box_refine(b, self.ratio(l))
        return b

    def cells(self, lev: int) -> int:
        return sum(box_npts(b) for b in self.levels[lev])

    def total_cells(self) -> int:
        return sum(self.cells(l) for l in range(self.nlev))

    def ghosted_cells(self, lev: int) -> int:
        return sum(box_npts(box_grow(b, 1)) for b in self.levels[lev])

    def level_offsets(self) -> List[int]:
        offs = [0]
        for l in range(self.nlev):
            offs.append(offs[-1] + self.cells(l))
        return offs

    def spec(self):
        """(ispec int32 array, rspec float64 array) as documented in include/amrmg.h."""
        iv = [self.nlev, *self.domain[0], *self.domain[1], *self.bc_type_lo, *self.bc_type_hi]
        iv += [len(lv) for lv in self.levels]
        for lv in self.levels:
            for b in lv:
                iv += [*b[0], *b[1]]
        rv = [self.dx0, self.alpha, self.beta, *self.bc_val_lo, *self.bc_val_hi]
        return np.asarray(iv, dtype=np.int32), np.asarray(rv, dtype=np.float64)

    # ---- field helpers (valid-packed layout: box by box, Fortran order) ----
    def cell_centres(self, lev: int, b: Box):
        dx = self.dx(lev)
        (l, h) = b
        xs = [(np.arange(l[d], h[d] + 1) + 0.5) * dx for d in range(3)]
        X, Y, Z = np.meshgrid(xs[0], xs[1], xs[2], indexing="ij")
        return X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F")

    def pack(self, fn) -> np.ndarray:
        """Evaluate fn(x, y, z, lev) on every valid cell, valid-packed."""
        parts = []
        for l in range(self.nlev):
            for b in self.levels[l]:
                x, y, z = self.cell_centres(l, b)
                parts.append(np.asarray(fn(x, y, z, l), dtype=np.float64))
        return np.concatenate(parts) if parts else np.zeros(0)

    def pack_faces(self, fn, d: int, lev: int = 0) -> np.ndarray:
        """Evaluate fn(x, y, z) on the faces normal to d of every box of level lev
        (the box grown by one face in d), per box in Fortran order: the face-packed
        layout of a FluxBox component (variable coefficients)."""
        dx = self.dx(lev)
        parts = []
        for (lo, hi) in self.levels[lev]:
            ax = []
            for q in range(3):
                n = hi[q] - lo[q] + 1 + (1 if q == d else 0)
                idx = np.arange(lo[q], lo[q] + n, dtype=np.float64)
                ax.append(idx * dx if q == d else (idx + 0.5) * dx)
            z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
            parts.append(np.asarray(fn(x, y, z), dtype=np.float64).ravel())
        return np.concatenate(parts) if parts else np.zeros(0)

    def unpack_level(self, data: np.ndarray, lev: int):
        """List of (box, 3-D array indexed [i,j,k]) for level lev of valid-packed data."""
        offs = self.level_offsets()
        o = offs[lev]
        out = []
        for b in self.levels[lev]:
            n = box_npts(b)
            out.append((b, data[o:o + n].reshape(box_shape(b), order="F")))
            o += n
        return out

    def to_global(self, data: np.ndarray, lev: int, fill=np.nan) -> np.ndarray:
        """Scatter level lev of valid-packed data into a dense array over the level domain."""
        dom = self.level_domain(lev)
        shp = box_shape(dom)
        g = np.full(shp, fill)
        for b, a in self.unpack_level(data, lev):
            sl = tuple(slice(b[0][d] - dom[0][d], b[1][d] - dom[0][d] + 1) for d in range(3))
            g[sl] = a
        return g

    def from_global(self, arrays: Sequence[np.ndarray]) -> np.ndarray:
        parts = []
        for l in range(self.nlev):
            dom = self.level_domain(l)
            for b in self.levels[l]:
                sl = tuple(slice(b[0][d] - dom[0][d], b[1][d] - dom[0][d] + 1) for d in range(3))
                parts.append(np.asarray(arrays[l][sl]).ravel(order="F"))
        return np.concatenate(parts)


def single_level(n: int = 64, max_box: int = 32, block: int = 8, alpha=0.0, beta=1.0) -> Hierarchy:
    dom = box((0, 0, 0), (n - 1, n - 1, n - 1))
    return Hierarchy(dom, [domain_split(dom, max_box, block)], 1.0 / n, alpha, beta)


def centred_region(parent_n: int, frac: float, block: int) -> Box:
    """Centred cube covering `frac` of a parent domain of side parent_n, side and
    low corner snapped to the blocking factor."""
    side = block * int(round(parent_n * frac ** (1.0 / 3.0) / block))
    side = max(block, side)
    lo = block * (((parent_n - side) // 2) // block)
    return box((lo, lo, lo), (lo + side - 1,) * 3)


def amr_hierarchy(n: int = 256, fracs=(0.25, 0.06), max_box: int = 64, block: int = 8,
                  alpha=0.0, beta=1.0) -> Hierarchy:
    """Base n^3 plus len(fracs) nested, centred refined levels (ratio 2)."""
    dom = box((0, 0, 0), (n - 1, n - 1, n - 1))
    levels = [domain_split(dom, max_box, block)]
    parent_n = n
    for f in fracs:
        reg = centred_region(parent_n, f, block)       # in parent index space
        fine = box_refine(reg, 2)
        levels.append(domain_split(fine, max_box, block))
        parent_n *= 2
    return Hierarchy(dom, levels, 1.0 / n, alpha, beta)


# ---- right-hand sides ----
def sinusoid_rhs(h: Hierarchy, seed: int, noise: float = 1e-3, k=(1, 1, 1)) -> np.ndarray:
    rng = np.random.default_rng(seed)
    base = h.pack(lambda x, y, z, l: np.sin(2 * math.pi * k[0] * x) * np.sin(2 * math.pi * k[1] * y)
                  * np.sin(2 * math.pi * k[2] * z))
    return base + noise * rng.standard_normal(base.size)


def bump_rhs(h: Hierarchy, seed: int, nbumps: int = 3) -> np.ndarray:
    """Sum of Gaussian bumps s*exp(-|x-c|^2/w) (poissonSolve.cpp:169-222 style),
    centres U(0.4,0.6)^3, widths U(5e-3, 2e-2), strengths U(1, 3)."""
    rng = np.random.default_rng(seed)
    cs = rng.uniform(0.4, 0.6, size=(nbumps, 3))
    ws = rng.uniform(5e-3, 2e-2, size=nbumps)
    ss = rng.uniform(1.0, 3.0, size=nbumps)

    def f(x, y, z, l):
        v = np.zeros_like(x)
        for c, w, s in zip(cs, ws, ss):
            r2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
            v = v + s * np.exp(-r2 / w)
        return v

    return h.pack(f)


def vc_coefficients(h: "Hierarchy", seed: int = 7, contrast: float = 4.0):
    """Smooth positive variable coefficients on every level (VCAMRPoissonOp2):
    a(x) = 1 + 0.5 sin(2 pi (x + p0)) cos(2 pi y) sin(2 pi z) at cell centres and
    b(x) = 1 + (contrast - 1) * (0.5 + 0.5 sin(2 pi (x + p1)) cos(2 pi (y + p2)) cos(pi z))
    at face centres, phases p from the seed.  Both are products of one-dimensional
    factors, evaluated per box as outer products.  Returns (a valid-packed over all
    levels, [b0, b1, b2] each face-packed level after level)."""
    rng = np.random.default_rng(seed)
    ph = rng.uniform(0, 1, size=3)
    fa = (lambda x: np.sin(2 * math.pi * (x + ph[0])), lambda y: np.cos(2 * math.pi * y),
          lambda z: np.sin(2 * math.pi * z))
    fb = (lambda x: np.sin(2 * math.pi * (x + ph[1])), lambda y: np.cos(2 * math.pi * (y + ph[2])),
          lambda z: np.cos(math.pi * z))

    def prod(fs, axes):
        return np.multiply.outer(np.multiply.outer(fs[0](axes[0]), fs[1](axes[1])), fs[2](axes[2]))

    a_parts, b_parts = [], [[], [], []]
    for lev in range(h.nlev):
        dx = h.dx(lev)
        for (lo, hi) in h.levels[lev]:
            cen = [(np.arange(lo[q], hi[q] + 1) + 0.5) * dx for q in range(3)]
            a_parts.append((1.0 + 0.5 * prod(fa, cen)).ravel(order="F"))
            for d in range(3):
                ax = list(cen)
                ax[d] = np.arange(lo[d], hi[d] + 2) * dx
                b_parts[d].append((1.0 + (contrast - 1.0) * (0.5 + 0.5 * prod(fb, ax))).ravel(order="F"))
    return np.concatenate(a_parts), [np.concatenate(p) for p in b_parts]


DEFAULT_SOLVER = dict(pre=2, post=2, bottom=2, numMG=1, iterMax=100, imin=5,
                      eps=1e-10, hang=1e-15, normThresh=1e-30)


def solver_params(**kw):
    p = dict(DEFAULT_SOLVER)
    p.update(kw)
    ip = np.asarray([p["pre"], p["post"], p["bottom"], p["numMG"], p["iterMax"], p["imin"]], np.int32)
    rp = np.asarray([p["eps"], p["hang"], p["normThresh"]], np.float64)
    return ip, rp


# ---- batched single-level problems ----
def batch_problems(count: int = 512, seed: int = 1004, n_min: int = 16, n_max: int = 128):
    """Independent single-level Helmholtz problems: edge n = 8*round(exp(U(ln n_min,
    ln n_max))/8), domainSplit with 64^3 maximum boxes (blocking factor 8),
    alpha ~ U(0,1), beta ~ U(0.5,2), homogeneous Dirichlet, rhs =
    sin(2 pi k1 x) sin(2 pi k2 y) sin(2 pi k3 z) + N(0, 1e-3) with k_d ~ U{1,2,3}.
    Returns a list of (Hierarchy, rhs)."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(count):
        n = int(8 * round(math.exp(rng.uniform(math.log(n_min), math.log(n_max))) / 8))
        n = max(8, min(n_max, n))
        alpha = float(rng.uniform(0.0, 1.0))
        beta = float(rng.uniform(0.5, 2.0))
        k = tuple(int(v) for v in rng.integers(1, 4, size=3))
        dom = box((0, 0, 0), (n - 1, n - 1, n - 1))
        h = Hierarchy(dom, [domain_split(dom, 64, 8)], 1.0 / n, alpha, beta)
        rhs = sinusoid_rhs(h, seed * 100003 + q, k=k)
        out.append((h, rhs))
    return out


# ---- serving workload ----
def bump_rhs_separable(h: Hierarchy, seed: int, nbumps: int = 3) -> np.ndarray:
    """The Gaussian-bump rhs of bump_rhs (same bump parameters for a seed), each bump
    evaluated as a product of three one-dimensional exponentials per box, which is
    fast enough to generate thousands of requests; values agree with bump_rhs to
    rounding, not bitwise."""
    rng = np.random.default_rng(seed)
    cs = rng.uniform(0.4, 0.6, size=(nbumps, 3))
    ws = rng.uniform(5e-3, 2e-2, size=nbumps)
    ss = rng.uniform(1.0, 3.0, size=nbumps)
    parts = []
    for lev in range(h.nlev):
        dx = h.dx(lev)
        for (l, hh) in h.levels[lev]:
            xs = [(np.arange(l[d], hh[d] + 1) + 0.5) * dx for d in range(3)]
            v = np.zeros((len(xs[0]), len(xs[1]), len(xs[2])))
            for c, w, sv in zip(cs, ws, ss):
                e = [np.exp(-((xs[d] - c[d]) ** 2) / w) for d in range(3)]
                v += sv * e[0][:, None, None] * e[1][None, :, None] * e[2][None, None, :]
            parts.append(v.ravel(order="F"))
    return np.concatenate(parts)


def serve_requests(count: int = 2000, seed: int = 1005, bases=(32, 64, 128, 256)):
    """Requests of 1-2-level AMR Poisson solves: base n ~ U{bases}, levels ~ U{1,2};
    the refined region covers a fraction f_r ~ U(0.05, 0.25) of the base domain's
    volume as a centred cube of side 8*round(n f_r^(1/3) / 8) (ratio 2, 64^3 maximum
    boxes, blocking factor 8).  Requests with the same hierarchy share one bump rhs
    (bump_rhs_separable, seeded by the hierarchy).  Returns a list of (Hierarchy, rhs)."""
    rng = np.random.default_rng(seed)
    cache = {}
    out = []
    for q in range(count):
        n = int(rng.choice(bases))
        nlev = int(rng.integers(1, 3))
        side = 0
        if nlev == 2:
            fr = float(rng.uniform(0.05, 0.25))
            side = max(8, 8 * int(round(n * fr ** (1.0 / 3.0) / 8)))
        key = (n, side)
        if key not in cache:
            dom = box((0, 0, 0), (n - 1, n - 1, n - 1))
            levels = [domain_split(dom, 64, 8)]
            if side:
                lo = 8 * (((n - side) // 2) // 8)
                reg = box((lo, lo, lo), (lo + side - 1,) * 3)
                levels.append(domain_split(box_refine(reg, 2), 64, 8))
            h = Hierarchy(dom, levels, 1.0 / n)
            cache[key] = (h, bump_rhs_separable(h, seed * 1009 + n * 7 + side))
        out.append(cache[key])
    return out
