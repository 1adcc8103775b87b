"""Shared test helpers: hierarchies and ghosted/valid packing."""
import numpy as np

import hierarchy as H


def ghosted_random(h, lev, rng):
    """Random ghosted-packed data for level lev (every box grown by one cell)."""
    n = h.ghosted_cells(lev)
    return rng.standard_normal(n)


def ghosted_from_fn(h, lev, fn, dx):
    parts = []
    for b in h.levels[lev]:
        g = H.box_grow(b, 1)
        (l, hh) = g
        idx = [np.arange(l[d], hh[d] + 1) for d in range(3)]
        I = np.meshgrid(*idx, indexing="ij")
        parts.append(np.asarray(fn(*[(I[d] + 0.5) * dx for d in range(3)]), dtype=np.float64).ravel(order="F"))
    return np.concatenate(parts)


def ghost_masks(h, lev):
    """For ghosted-packed data of level lev: (valid mask, face-ghost mask, cf-ghost mask)."""
    dom = h.level_domain(lev)
    boxes = h.levels[lev]
    valid, face, cf = [], [], []
    for b in boxes:
        g = H.box_grow(b, 1)
        (l, hh) = g
        idx = [np.arange(l[d], hh[d] + 1) for d in range(3)]
        I = np.meshgrid(*idx, indexing="ij")
        out = [(I[d] < b[0][d]) | (I[d] > b[1][d]) for d in range(3)]
        nout = out[0].astype(int) + out[1].astype(int) + out[2].astype(int)
        v = nout == 0
        f = nout == 1
        indom = np.ones_like(v)
        for d in range(3):
            indom &= (I[d] >= dom[0][d]) & (I[d] <= dom[1][d])
        covered = np.zeros_like(v)
        for bb in boxes:
            m = np.ones_like(v)
            for d in range(3):
                m &= (I[d] >= bb[0][d]) & (I[d] <= bb[1][d])
            covered |= m
        valid.append(v.ravel(order="F"))
        face.append(f.ravel(order="F"))
        cf.append((f & indom & ~covered).ravel(order="F"))
    return np.concatenate(valid), np.concatenate(face), np.concatenate(cf)


def ghost_dir(h, lev):
    """For ghosted-packed data of level lev: direction (0,1,2) of each face ghost, -1 otherwise."""
    parts = []
    for b in h.levels[lev]:
        g = H.box_grow(b, 1)
        (l, hh) = g
        idx = [np.arange(l[d], hh[d] + 1) for d in range(3)]
        I = np.meshgrid(*idx, indexing="ij")
        out = [(I[d] < b[0][d]) | (I[d] > b[1][d]) for d in range(3)]
        nout = out[0].astype(int) + out[1].astype(int) + out[2].astype(int)
        dirs = np.full(I[0].shape, -1)
        for d in range(3):
            dirs = np.where(out[d] & (nout == 1), d, dirs)
        parts.append(dirs.ravel(order="F"))
    return np.concatenate(parts)


def uncovered_mask(h, lev):
    """Valid-packed mask of level-lev cells not covered by level lev+1."""
    if lev + 1 >= h.nlev:
        return np.ones(h.cells(lev), dtype=bool)
    dom = h.level_domain(lev)
    cov = np.zeros(H.box_shape(dom), dtype=bool)
    sl = lambda b: tuple(slice(b[0][d] - dom[0][d], b[1][d] - dom[0][d] + 1) for d in range(3))
    for b in h.levels[lev + 1]:
        cov[sl(H.box_coarsen(b, h.ratio(lev + 1)))] = True
    return np.concatenate([(~cov[sl(b)]).ravel(order="F") for b in h.levels[lev]])


def small_single(n=32, box=16, block=8, lo=0, alpha=0.0, beta=1.0, bc=(1, 1, 1)):
    dom = H.box((lo,) * 3, (lo + n - 1,) * 3)
    return H.Hierarchy(dom, [H.domain_split(dom, box, block)], 1.0 / n, alpha, beta,
                       bc_type_lo=bc, bc_type_hi=bc)


def small_amr(n=32, box=16, block=4, alpha=0.0, beta=1.0):
    return H.amr_hierarchy(n, fracs=(0.25, 0.06), max_box=box, block=block, alpha=alpha, beta=beta)


def lshape_qcfi():
    """Hierarchy of lib/test/AMRTools/quadCFInterpTest.cpp (setHierarchy, 3-D)."""
    fine = [((0, 4, 0), (7, 12, 15)), ((8, 6, 0), (11, 9, 15)), ((4, 0, 0), (9, 3, 15)),
            ((10, 4, 0), (15, 5, 15))]
    fine = [H.box_refine(b, 2) for b in fine]
    coarse = [((0, 0, 0), (7, 15, 15)), ((8, 0, 0), (15, 7, 15)), ((8, 8, 0), (15, 15, 15))]
    return H.Hierarchy(H.box((0, 0, 0), (15, 15, 15)), [coarse, fine], 2 * 3.14159)


def lshape_amr(alpha=0.0, beta=1.0):
    """Three levels over a 32^3 base whose middle level is L-shaped in x-y, so its
    boxes do not tile a rectangle and keep per-box storage (levels 0 and 2 do tile
    one and are stored merged)."""
    dom = H.box((0, 0, 0), (31, 31, 31))
    lev1 = [((16, 16, 16), (31, 31, 47)), ((16, 32, 16), (31, 47, 47)), ((32, 16, 16), (47, 31, 47))]
    lev2 = [((40, 40, 40), (55, 55, 63)), ((40, 40, 64), (55, 55, 87))]
    return H.Hierarchy(dom, [H.domain_split(dom, 16, 8), lev1, lev2], 1.0 / 32, alpha, beta)


def mixed_ratio_amr(ratios, bc=(1, 1, 1), alpha=0.0, beta=1.0):
    """Three levels over [-8, 7]^3 (dx 1/16) with refinement ratios (r1, r2): level 1
    over coarse [-4, 7] x [-4, 3]^2 and level 2 over the last 4 coarse cells of level 1
    in x, both reaching the high x face."""
    r1, r2 = ratios
    dom = H.box((-8,) * 3, (7,) * 3)
    lev1 = H.box_refine(H.box((-4, -4, -4), (7, 3, 3)), r1)
    d1 = H.box_refine(dom, r1)
    lev2 = H.box_refine(H.box((d1[1][0] - 4 * r1 + 1, -2 * r1, -2 * r1), (d1[1][0], 2 * r1 - 1, 2 * r1 - 1)), r2)
    return H.Hierarchy(dom, [H.domain_split(dom, 8, 4), H.domain_split(lev1, 16, 4), H.domain_split(lev2, 16, 4)],
                       1.0 / 16, alpha, beta, bc_type_lo=bc, bc_type_hi=bc, ratios=tuple(ratios))


def comps_pack(h, parts):
    """nc valid-packed one-component arrays (all levels) -> one array packed box by
    box, each box component by component (the order of an FArrayBox's data)."""
    out = []
    o = 0
    for l in range(h.nlev):
        for b in h.levels[l]:
            n = H.box_npts(b)
            out.extend(p[o:o + n] for p in parts)
            o += n
    return np.concatenate(out)


def comps_unpack(h, v, nc):
    """Inverse of comps_pack: the nc one-component arrays."""
    parts = [[] for _ in range(nc)]
    o = 0
    for l in range(h.nlev):
        for b in h.levels[l]:
            n = H.box_npts(b)
            for c in range(nc):
                parts[c].append(v[o:o + n])
                o += n
    return [np.concatenate(p) for p in parts]
