"""Operator-level binding of libamrmg.so: device-resident levels, fields, operators
and copiers (the ``amrmg_level_*`` / ``amrmg_field_*`` / ``amrmg_op_*`` /
``amrmg_plan_*`` entry points of include/amrmg.h).

``AMRPoissonOp`` here has the methods of the reference's LinearOp, MGLevelOp and
AMRLevelOp virtuals (lib/src/AMRElliptic/LinearSolver.H:27-149, MultiGrid.H:88-287,
AMRMultiGrid.H:38-226) with the same names and argument meanings; ``T`` is a
``Field`` (a LevelData<FArrayBox> with one component and one ghost cell, resident in
HBM).  Solvers written against those virtuals (amrmultigrid.py) drive it unchanged.
Nothing here computes on the host: every method is one or a few C ABI calls.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import amrmg

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p
_pp = C.POINTER(C.c_void_p)
_set = False


def _lib():
    global _set
    L = amrmg.lib()
    if not _set:
        sig = {
            "amrmg_ctx_sync": [_vp],
            "amrmg_level_register": [_vp, _ip, C.c_int, _ip, C.c_int, _pp],
            "amrmg_level_destroy": [_vp],
            "amrmg_level_coarsen": [_vp, _pp],
            "amrmg_field_create": [_vp, _pp],
            "amrmg_field_create_ncomp": [_vp, C.c_int, _pp],
            "amrmg_op_set_ncomp": [_vp, C.c_int],
            "amrmg_field_destroy": [_vp],
            "amrmg_field_upload": [_vp, _dp, C.c_int],
            "amrmg_field_download": [_vp, _dp, C.c_int],
            "amrmg_field_assign": [_vp, _vp],
            "amrmg_field_incr": [_vp, _vp, C.c_double],
            "amrmg_field_axby": [_vp, _vp, _vp, C.c_double, C.c_double],
            "amrmg_field_scale": [_vp, C.c_double],
            "amrmg_field_setval": [_vp, C.c_double],
            "amrmg_field_dot": [_vp, _vp, _dp],
            "amrmg_field_dot_noalloc": [_vp, _vp, _dp],
            "amrmg_field_norm": [_vp, C.c_int, _dp],
            "amrmg_op_create": [_vp, _vp, C.c_double, C.c_double, C.c_double, C.c_double, _ip, _dp, _pp],
            "amrmg_op_create_coarser": [_vp, _pp, _pp],
            "amrmg_op_coarsened_level": [_vp, _pp],
            "amrmg_op_destroy": [_vp],
            "amrmg_op_set_alpha_beta": [_vp, C.c_double, C.c_double],
            "amrmg_op_set_relax_mode": [_vp, C.c_int],
            "amrmg_op_residual": [_vp, _vp, _vp, _vp, C.c_int],
            "amrmg_op_apply": [_vp, _vp, _vp, C.c_int],
            "amrmg_op_precond": [_vp, _vp, _vp],
            "amrmg_op_relax": [_vp, _vp, _vp, C.c_int],
            "amrmg_op_restrict_residual": [_vp, _vp, _vp, _vp],
            "amrmg_op_prolong_increment": [_vp, _vp, _vp],
            "amrmg_op_amr_residual": [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp],
            "amrmg_op_amr_operator": [_vp, _vp, _vp, _vp, _vp, C.c_int, _vp],
            "amrmg_op_amr_restrict": [_vp, _vp, _vp, _vp, _vp, _vp],
            "amrmg_op_amr_prolong": [_vp, _vp, _vp, _vp],
            "amrmg_op_amr_update_residual": [_vp, _vp, _vp, _vp],
            "amrmg_op_assign_copier": [_vp, _vp, _vp],
            "amrmg_op_zero_covered": [_vp, _vp],
            "amrmg_op_max_norm": [_vp, _vp, _dp],
            "amrmg_op_exchange": [_vp, _vp],
            "amrmg_op_bc": [_vp, _vp, C.c_int],
            "amrmg_op_set_bc_values": [_vp, amrmg.BC_VALUE_FN, _vp],
            "amrmg_op_cf_homog": [_vp, _vp],
            "amrmg_op_cf_inhomog": [_vp, _vp, _vp],
            "amrmg_op_gsrb": [_vp, _vp, _vp, C.c_int],
            "amrmg_op_reflux": [_vp, _vp, _vp, _vp, _vp],
            "amrmg_plan_copy": [_vp, _vp, _ip, C.c_int, _pp],
            "amrmg_plan_exchange": [_vp, _ip, C.c_int, _pp],
            "amrmg_plan_exec": [_vp, _vp, _vp, C.c_int],
            "amrmg_plan_destroy": [_vp],
        }
        for name, args in sig.items():
            getattr(L, name).argtypes = args
        L.amrmg_level_cells.restype = C.c_longlong
        L.amrmg_level_cells.argtypes = [_vp, C.c_int]
        _set = True
    return L


def _chk(rc):
    if rc != 0:
        raise amrmg.AMRMGError(_lib().amrmg_last_error().decode())


def _h(x):
    """Handle of an optional operand (None is the reference's undefined LevelData)."""
    return None if x is None else x.p


class Level:
    """DisjointBoxLayout + ProblemDomain on the device (amrmg_level_register)."""

    def __init__(self, ctx, boxes, domain, merge=True, _handle=None, _owned=False):
        self.ctx = ctx
        self.boxes = [tuple(map(tuple, b)) for b in boxes]
        self.domain = domain
        self._owned = _owned
        if _handle is not None:
            self.p = _handle
            return
        flat = np.asarray([v for (lo, hi) in self.boxes for v in (*lo, *hi)], dtype=np.int32)
        dom = np.asarray([*domain[0], *domain[1]], dtype=np.int32)
        p = C.c_void_p()
        _chk(_lib().amrmg_level_register(ctx.p, flat.ctypes.data_as(_ip), len(self.boxes), dom.ctypes.data_as(_ip),
                                         1 if merge else 0, C.byref(p)))
        self.p = p

    def cells(self, ghosted=False):
        return int(_lib().amrmg_level_cells(self.p, 1 if ghosted else 0))

    def coarsen(self):
        """coarsen(DisjointBoxLayout, 2) with the coarsened domain."""
        p = C.c_void_p()
        _chk(_lib().amrmg_level_coarsen(self.p, C.byref(p)))
        cb = [(tuple(v // 2 for v in lo), tuple(v // 2 for v in hi)) for (lo, hi) in self.boxes]
        dom = (tuple(v // 2 for v in self.domain[0]), tuple(v // 2 for v in self.domain[1]))
        return Level(self.ctx, cb, dom, _handle=p)

    def close(self):
        if getattr(self, "p", None) and not self._owned:
            _chk(_lib().amrmg_level_destroy(self.p))
        self.p = None


class Field:
    """LevelData<FArrayBox>(level, ncomp components, one ghost cell), zero on creation;
    host data box by box, each box component by component."""

    def __init__(self, level: Level, ncomp: int = 1):
        self.level = level
        self.ncomp = int(ncomp)
        p = C.c_void_p()
        _chk(_lib().amrmg_field_create_ncomp(level.p, self.ncomp, C.byref(p)))
        self.p = p

    def upload(self, data, ghosted=False):
        a = np.ascontiguousarray(data, dtype=np.float64)
        if a.size != self.level.cells(ghosted) * self.ncomp:
            raise ValueError("field upload: size does not match the level")
        _chk(_lib().amrmg_field_upload(self.p, a.ctypes.data_as(_dp), 1 if ghosted else 0))
        return self

    def download(self, ghosted=False):
        out = np.zeros(self.level.cells(ghosted) * self.ncomp)
        _chk(_lib().amrmg_field_download(self.p, out.ctypes.data_as(_dp), 1 if ghosted else 0))
        return out

    def close(self):
        if getattr(self, "p", None):
            _chk(_lib().amrmg_field_destroy(self.p))
        self.p = None


class Copier:
    """Copier between two levels (copyTo / copyToZero) or a level's exchange Copier."""

    def __init__(self, src: Level, dst: Level = None, items=None, exchange=False):
        p = C.c_void_p()
        it = None if items is None else np.ascontiguousarray(np.asarray(items, dtype=np.int32).reshape(-1, 8))
        ip = None if it is None else it.ctypes.data_as(_ip)
        n = 0 if it is None else it.shape[0]
        if exchange:
            _chk(_lib().amrmg_plan_exchange(src.p, ip, n, C.byref(p)))
        else:
            _chk(_lib().amrmg_plan_copy(src.p, dst.p, ip, n, C.byref(p)))
        self.p = p

    def copyTo(self, src: Field, dst: Field):
        _chk(_lib().amrmg_plan_exec(self.p, dst.p, src.p, 0))

    def copyToZero(self, dst: Field):
        _chk(_lib().amrmg_plan_exec(self.p, dst.p, None, 1))

    def close(self):
        if getattr(self, "p", None):
            _chk(_lib().amrmg_plan_destroy(self.p))
        self.p = None


class AMRPoissonOp:
    """AMRPoissonOp (alpha I + beta Laplacian) on one level, resident on the device.

    bc: (types, values), 6 each: lo x, y, z then hi x, y, z; type 0 Neumann,
    1 Dirichlet (linear rule), 2 Dirichlet (quadratic rule).  coarser: the operator of
    the next coarser AMR level or None; dxCrse > 0 turns on the homogeneous
    coarse-fine ghosts of the level's own CFRegion, and with a coarser operator
    dxCrse / dx is the refinement ratio (2 or 4)."""

    def __init__(self, level: Level, dx, dxCrse=-1.0, alpha=0.0, beta=1.0, bc=((1,) * 6, (0.0,) * 6),
                 coarser: "AMRPoissonOp" = None, _handle=None):
        self.level = level
        self.m_dx = dx
        self.m_dxCrse = dxCrse
        self.alpha, self.beta = alpha, beta
        self.bc = bc
        self.coarser = coarser
        self.finer = None
        self.mgChild = None
        if _handle is not None:
            self.p = _handle
            return
        t = np.asarray(bc[0], dtype=np.int32)
        v = np.asarray(bc[1], dtype=np.float64)
        p = C.c_void_p()
        _chk(_lib().amrmg_op_create(level.p, _h(coarser), dx, dxCrse, alpha, beta, t.ctypes.data_as(_ip),
                                    v.ctypes.data_as(_dp), C.byref(p)))
        self.p = p
        if coarser is not None:
            coarser.finer = self

    # ---- factory pieces ----
    def createCoarser(self):
        """MGnewOp at the next depth: (operator, its level)."""
        p, lp = C.c_void_p(), C.c_void_p()
        _chk(_lib().amrmg_op_create_coarser(self.p, C.byref(p), C.byref(lp)))
        cb = [(tuple(v // 2 for v in lo), tuple(v // 2 for v in hi)) for (lo, hi) in self.level.boxes]
        dom = (tuple(v // 2 for v in self.level.domain[0]), tuple(v // 2 for v in self.level.domain[1]))
        lev = Level(self.level.ctx, cb, dom, _handle=lp, _owned=True)
        op = AMRPoissonOp(lev, 2 * self.m_dx, self.m_dxCrse, self.alpha, self.beta, self.bc, _handle=p)
        self.mgChild = op
        return op

    def coarsenedLevel(self):
        """coarsen(level, refToCoarser) of an operator with a coarser AMR level
        (createCoarsened)."""
        lp = C.c_void_p()
        _chk(_lib().amrmg_op_coarsened_level(self.p, C.byref(lp)))
        r = self.refToCoarser()
        cb = [(tuple(v // r for v in lo), tuple(v // r for v in hi)) for (lo, hi) in self.level.boxes]
        dom = (tuple(v // r for v in self.level.domain[0]), tuple(v // r for v in self.level.domain[1]))
        return Level(self.level.ctx, cb, dom, _handle=lp, _owned=True)

    def setAlphaAndBeta(self, alpha, beta):
        self.alpha, self.beta = alpha, beta
        _chk(_lib().amrmg_op_set_alpha_beta(self.p, alpha, beta))

    def setRelaxMode(self, mode):
        _chk(_lib().amrmg_op_set_relax_mode(self.p, int(mode)))

    def dx(self):
        return self.m_dx

    # ---- LinearOp ----
    def create(self, like: Field) -> Field:
        return Field(like.level, like.ncomp)

    def residual(self, lhs, phi, rhs, homogeneous=False):
        _chk(_lib().amrmg_op_residual(self.p, lhs.p, phi.p, rhs.p, int(homogeneous)))

    def applyOp(self, lhs, phi, homogeneous=False):
        _chk(_lib().amrmg_op_apply(self.p, lhs.p, phi.p, int(homogeneous)))

    def preCond(self, cor, residual):
        _chk(_lib().amrmg_op_precond(self.p, cor.p, residual.p))

    def assignLocal(self, lhs, rhs):
        _chk(_lib().amrmg_field_assign(lhs.p, rhs.p))

    def dotProduct(self, a, b):
        out = C.c_double()
        _chk(_lib().amrmg_field_dot(a.p, b.p, C.byref(out)))
        return out.value

    def incr(self, lhs, x, scale):
        _chk(_lib().amrmg_field_incr(lhs.p, x.p, scale))

    def axby(self, lhs, x, y, a, b):
        _chk(_lib().amrmg_field_axby(lhs.p, x.p, y.p, a, b))

    def scale(self, lhs, s):
        _chk(_lib().amrmg_field_scale(lhs.p, s))

    def norm(self, x, ord):
        out = C.c_double()
        _chk(_lib().amrmg_field_norm(x.p, int(ord), C.byref(out)))
        return out.value

    def setToZero(self, lhs):
        _chk(_lib().amrmg_field_setval(lhs.p, 0.0))

    # ---- MGLevelOp ----
    def createCoarserField(self, fine: Field) -> Field:
        return Field(self.mgChild.level, fine.ncomp if fine is not None else 1)

    def relax(self, correction, residual, iterations):
        _chk(_lib().amrmg_op_relax(self.p, correction.p, residual.p, int(iterations)))

    def restrictResidual(self, resCoarse, phiFine, rhsFine):
        _chk(_lib().amrmg_op_restrict_residual(self.p, resCoarse.p, phiFine.p, rhsFine.p))

    def prolongIncrement(self, phiThisLevel, correctCoarse):
        _chk(_lib().amrmg_op_prolong_increment(self.p, phiThisLevel.p, correctCoarse.p))

    # ---- AMRLevelOp ----
    def refToCoarser(self):
        """dxCrse / dx for an AMR level with a coarser one (2 or 4), else 2."""
        if self.coarser is None or self.m_dxCrse <= 0:
            return 2
        return int(round(self.m_dxCrse / self.m_dx))

    def AMRResidual(self, residual, phiFine, phi, phiCoarse, rhs, homogeneousBC, finerOp):
        _chk(_lib().amrmg_op_amr_residual(self.p, residual.p, _h(phiFine), phi.p, _h(phiCoarse), rhs.p,
                                          int(homogeneousBC), _h(finerOp)))

    def AMRResidualNC(self, residual, phiFine, phi, rhs, homogeneousBC, finerOp):
        self.AMRResidual(residual, phiFine, phi, None, rhs, homogeneousBC, finerOp)

    def AMRResidualNF(self, residual, phi, phiCoarse, rhs, homogeneousBC):
        self.AMRResidual(residual, None, phi, phiCoarse, rhs, homogeneousBC, None)

    def AMROperator(self, LofPhi, phiFine, phi, phiCoarse, homogeneousBC, finerOp):
        _chk(_lib().amrmg_op_amr_operator(self.p, LofPhi.p, _h(phiFine), phi.p, _h(phiCoarse), int(homogeneousBC),
                                          _h(finerOp)))

    def AMROperatorNC(self, LofPhi, phiFine, phi, homogeneousBC, finerOp):
        self.AMROperator(LofPhi, phiFine, phi, None, homogeneousBC, finerOp)

    def AMROperatorNF(self, LofPhi, phi, phiCoarse, homogeneousBC):
        self.AMROperator(LofPhi, None, phi, phiCoarse, homogeneousBC, None)

    def AMRRestrictS(self, resCoarse, residual, correction, coarseCorrection, scratch):
        _chk(_lib().amrmg_op_amr_restrict(self.p, resCoarse.p, residual.p, correction.p, _h(coarseCorrection),
                                          scratch.p))

    def AMRProlongS(self, correction, coarseCorrection, temp):
        _chk(_lib().amrmg_op_amr_prolong(self.p, correction.p, coarseCorrection.p, temp.p))

    def AMRUpdateResidual(self, residual, correction, coarseCorrection):
        _chk(_lib().amrmg_op_amr_update_residual(self.p, residual.p, correction.p, _h(coarseCorrection)))

    def createCoarsened(self, rhs: Field, refRat=2) -> Field:
        return Field(self.coarsenedLevel(), rhs.ncomp)

    def assignCopier(self, lhs, rhs):
        _chk(_lib().amrmg_op_assign_copier(self.p, lhs.p, rhs.p))

    def zeroCovered(self, lhs):
        _chk(_lib().amrmg_op_zero_covered(self.p, lhs.p))

    def localMaxNorm(self, x):
        out = C.c_double()
        _chk(_lib().amrmg_op_max_norm(self.p, x.p, C.byref(out)))
        return out.value

    # ---- the ghost and box kernels the virtuals are made of ----
    def exchange(self, phi):
        _chk(_lib().amrmg_op_exchange(self.p, phi.p))

    def applyBC(self, phi, homogeneous):
        _chk(_lib().amrmg_op_bc(self.p, phi.p, int(homogeneous)))

    def setNComp(self, ncomp):
        """Components of the fields this operator's boundary value function serves."""
        _chk(_lib().amrmg_op_set_ncomp(self.p, int(ncomp)))

    def setBCValues(self, fn):
        """This operator's inhomogeneous BC values from fn(pos, dir, side) at the domain
        face centres (a BCValueHolder in DiriBC / NeumBC); None: the face constants."""
        cb = amrmg.bc_value_callback(fn)
        _chk(_lib().amrmg_op_set_bc_values(self.p, cb, None))
        self._bc_fn = cb

    def homogeneousCFInterp(self, phi):
        _chk(_lib().amrmg_op_cf_homog(self.p, phi.p))

    def coarseFineInterp(self, phi, phiCoarse):
        _chk(_lib().amrmg_op_cf_inhomog(self.p, phi.p, phiCoarse.p))

    def gsrbPass(self, phi, rhs, color):
        _chk(_lib().amrmg_op_gsrb(self.p, phi.p, rhs.p, int(color)))

    def reflux(self, phiFine, phi, lphi, finerOp):
        _chk(_lib().amrmg_op_reflux(self.p, phiFine.p, phi.p, lphi.p, finerOp.p))

    def close(self):
        if getattr(self, "p", None):
            _chk(_lib().amrmg_op_destroy(self.p))
            if self.coarser is not None:
                self.coarser.finer = None
        self.p = None


def sync(ctx):
    _chk(_lib().amrmg_ctx_sync(ctx.p))
