"""AMRMultiGrid, MultiGrid and BiCGStabSolver written against the operator virtuals.

The reference's solvers are templates over an operator interface
(lib/src/AMRElliptic/AMRMultiGrid.H:307-1419, MultiGrid.H:330-696,
BiCGStabSolver.H:199-527); they call LinearOp / MGLevelOp / AMRLevelOp virtuals and
never touch field data.  These classes are the same control flow in Python over any
object with those methods -- here levelops.AMRPoissonOp, whose every method is a
device operation of libamrmg.so -- so a host that keeps the reference's solver loop
and swaps only the operator gets the device path.  They also accept l_base > 0
(AMRMultiGrid::solve(phi, rhs, l_max, l_base), AMRMultiGrid.H:895-935): the coarse
levels below l_base supply coarse-fine boundary values only.

The whole-solve entry point (amrmg_solver_solve) runs the same sequence inside the
library with fused kernels; both give bitwise the same iterates.
"""
from __future__ import annotations

import math


class BiCGStabSolver:
    """BiCGStabSolver<T> (BiCGStabSolver.H:199-527): public parameters as in the
    reference, defaults from its constructor (:166-183)."""

    def __init__(self):
        self.m_op = None
        self.m_homogeneous = False
        self.m_imax = 80
        self.m_eps = 1e-6
        self.m_reps = 1e-12
        self.m_convergenceMetric = -1.0
        self.m_hang = 1e-8
        self.m_small = 1e-30
        self.m_numRestarts = 5
        self.m_normType = 2
        self.m_exitStatus = -1
        self.m_iter = -1

    def define(self, op, homogeneous):
        self.m_op = op
        self.m_homogeneous = homogeneous

    def setConvergenceMetrics(self, metric, tolerance):
        self.m_convergenceMetric = metric
        self.m_eps = tolerance

    def solve(self, phi, rhs):
        op = self.m_op
        r, r_tilde, e, p = op.create(rhs), op.create(rhs), op.create(phi), op.create(rhs)
        p_tilde, s_tilde, t, v = op.create(phi), op.create(phi), op.create(rhs), op.create(rhs)
        tmp = [r, r_tilde, e, p, p_tilde, s_tilde, t, v]
        try:
            self._solve(phi, rhs, r, r_tilde, e, p, p_tilde, s_tilde, t, v)
        finally:
            for f in tmp:
                f.close()

    def _solve(self, phi, rhs, r, r_tilde, e, p, p_tilde, s_tilde, t, v):
        op = self.m_op
        recount = 0
        op.setToZero(r)
        op.residual(r, phi, rhs, self.m_homogeneous)
        op.assignLocal(r_tilde, r)
        op.setToZero(e)
        op.setToZero(p_tilde)
        op.setToZero(s_tilde)
        self.m_iter = 0
        rho = [0.0, 0.0, 0.0, 0.0]
        norm = [op.norm(r, self.m_normType), 0.0]
        initial_norm = norm[0]
        initial_residual = norm[0]
        norm[1] = norm[0]
        alpha, beta, omega = [0.0, 0.0], [0.0, 0.0], [0.0, 0.0]
        init = True
        restarts = 0
        if self.m_convergenceMetric > 0:
            initial_norm = self.m_convergenceMetric
        while self.m_iter < self.m_imax and norm[0] > self.m_eps * norm[1] and norm[1] > 0:
            self.m_iter += 1
            norm[1] = norm[0]
            alpha[1], beta[1], omega[1] = alpha[0], beta[0], omega[0]
            rho[3], rho[2] = rho[2], rho[1]
            rho[1] = op.dotProduct(r_tilde, r)
            if rho[1] == 0.0:
                op.incr(phi, e, 1.0)
                self.m_exitStatus = 2
                return
            if init:
                op.assignLocal(p, r)
                init = False
            else:
                beta[1] = (rho[1] / rho[2]) * (alpha[1] / omega[1])
                op.scale(p, beta[1])
                op.incr(p, v, -beta[1] * omega[1])
                op.incr(p, r, 1.0)
            op.preCond(p_tilde, p)
            op.setToZero(v)
            op.applyOp(v, p_tilde, True)
            m = op.dotProduct(r_tilde, v)
            alpha[0] = rho[1] / m
            if abs(m) > self.m_small * abs(rho[1]):
                op.incr(r, v, -alpha[0])
                norm[0] = op.norm(r, self.m_normType)
                op.incr(e, p_tilde, alpha[0])
            else:
                op.setToZero(r)
                norm[0] = 0.0
            if norm[0] > self.m_eps * initial_norm and norm[0] > self.m_reps * initial_residual:
                op.preCond(s_tilde, r)
                op.setToZero(t)
                op.applyOp(t, s_tilde, True)
                omega[0] = op.dotProduct(t, r) / op.dotProduct(t, t)
                op.incr(e, s_tilde, omega[0])
                op.incr(r, t, -omega[0])
                norm[0] = op.norm(r, self.m_normType)
            if norm[0] <= self.m_eps * initial_norm or norm[0] <= self.m_reps * initial_residual:
                self.m_exitStatus = 1
                break
            if omega[0] == 0.0 or norm[0] > (1 - self.m_hang) * norm[1]:
                if recount == 0:
                    recount = 1
                else:
                    recount = 0
                    op.incr(phi, e, 1.0)
                    if restarts == self.m_numRestarts:
                        self.m_exitStatus = 3
                        return
                    op.residual(r, phi, rhs, self.m_homogeneous)
                    norm[0] = op.norm(r, self.m_normType)
                    rho[1] = rho[2] = rho[3] = 0.0
                    alpha[0] = beta[0] = omega[0] = 0.0
                    op.assignLocal(r_tilde, r)
                    op.setToZero(e)
                    restarts += 1
                    init = True
        op.incr(phi, e, 1.0)


class MultiGrid:
    """MultiGrid<T> (MultiGrid.H:330-696): V and W cycles (m_cycle > 0), F cycles
    (m_cycle < 0), a bottom solver at the deepest level."""

    def __init__(self):
        self.m_pre, self.m_post, self.m_bottom = 2, 2, 2
        self.m_cycle = 1
        self.m_numMG = 1
        self.m_op = []
        self.m_residual, self.m_correction = [], []
        self.m_bottomSolver = None
        self.m_bottomCells = 0

    def define(self, finestOp, domainCells, bottomSolver, maxDepth=-1):
        """finestOp: the AMR level's operator; coarser operators come from its
        createCoarser chain (MGnewOp), as long as the boxes stay coarsenable by
        coarsening * s_maxCoarse (s_maxCoarse = 2, AMRPoissonOp.cpp:32, 1955-1959)."""
        self.m_op = [finestOp]
        self.m_bottomCells = domainCells
        op = finestOp
        coarsening = 1
        while maxDepth < 0 or len(self.m_op) < maxDepth:
            coarsening *= 2
            if not is_coarsenable(finestOp.level.boxes, coarsening * 2):
                break
            op = op.createCoarser()
            self.m_op.append(op)
            self.m_bottomCells //= 8
        self.m_depth = len(self.m_op)
        self.setBottomSolver(bottomSolver)

    def setBottomSolver(self, solver):
        self.m_bottomSolver = solver
        if solver is not None:
            solver.define(self.m_op[-1], True)

    def init(self, e, residual):
        """MultiGrid::init (MultiGrid.H:534-548)."""
        self.close_fields()
        self.m_residual = [None] + [self.m_op[i - 1].createCoarserField(residual) for i in range(1, self.m_depth)]
        self.m_correction = [None] + [self.m_op[i - 1].createCoarserField(e) for i in range(1, self.m_depth)]

    def oneCycle(self, e, residual):
        self.cycle(0, e, residual)

    def cycle(self, depth, correction, residual):
        """MultiGrid::cycle (MultiGrid.H:630-696)."""
        op = self.m_op[depth]
        if depth == self.m_depth - 1:
            if self.m_bottomCells == 1:
                op.relax(correction, residual, 1)
            else:
                op.relax(correction, residual, self.m_bottom)
                self.m_bottomSolver.solve(correction, residual)
            return
        cycles = self.m_cycle
        if cycles < 0:
            cycles = -cycles
            op.restrictResidual(self.m_residual[depth + 1], correction, residual)
            op.setToZero(self.m_correction[depth + 1])
            self.cycle(depth + 1, self.m_correction[depth + 1], self.m_residual[depth + 1])
            op.prolongIncrement(correction, self.m_correction[depth + 1])
            op.relax(correction, residual, self.m_pre)
            for img in range(cycles):
                op.restrictResidual(self.m_residual[depth + 1], correction, residual)
                op.setToZero(self.m_correction[depth + 1])
                self.m_cycle = 1
                self.cycle(depth + 1, self.m_correction[depth + 1], self.m_residual[depth + 1])
                self.m_cycle = -cycles
                op.prolongIncrement(correction, self.m_correction[depth + 1])
            op.relax(correction, residual, self.m_post)
            return
        op.relax(correction, residual, self.m_pre)
        op.restrictResidual(self.m_residual[depth + 1], correction, residual)
        op.setToZero(self.m_correction[depth + 1])
        for img in range(cycles):
            self.cycle(depth + 1, self.m_correction[depth + 1], self.m_residual[depth + 1])
        op.prolongIncrement(correction, self.m_correction[depth + 1])
        op.relax(correction, residual, self.m_post)

    def close_fields(self):
        for f in self.m_residual + self.m_correction:
            if f is not None:
                f.close()
        self.m_residual, self.m_correction = [], []

    def close(self):
        self.close_fields()
        for op in reversed(self.m_op[1:]):
            op.close()
        self.m_op = self.m_op[:1]


class AMRMultiGrid:
    """AMRMultiGrid<T> (AMRMultiGrid.H:307-1419) over per-level AMRLevelOp objects
    (ops[l].coarser / .finer linked as AMRnewOp links them)."""

    def __init__(self):
        self.m_eps, self.m_hang, self.m_normThresh = 1e-6, 1e-15, 1e-30
        self.m_imin, self.m_iterMax, self.m_iterMin = 5, 20, -1
        self.m_pre, self.m_post, self.m_bottom, self.m_numMG = 2, 2, 2, 1
        self.m_convergenceMetric = 0.0
        self.m_bottomSolverEpsCushion = 1.0
        self.m_exitStatus = -1
        self.m_op = []
        self.m_mg = []
        self.m_bottomSolver = None
        self.m_history = []
        self.m_iterations = 0

    def define(self, ops, domainCells, bottomSolver):
        """ops: the AMR level operators; domainCells: numPts of each level's domain.
        Every level gets a MultiGrid; the one at l_base keeps its full depth, the
        others a single level (no intermediate depths: ratio 4 makes that level's
        relaxation the MultiGrid bottom branch, see relax)."""
        self.m_op = list(ops)
        self.m_bottomSolver = bottomSolver
        self.m_domainCells = list(domainCells)
        self.m_mg = [None] * len(ops)
        self.m_correction = [None] * len(ops)
        self.m_residual = [None] * len(ops)
        self.m_resC = [None] * len(ops)

    def setSolverParameters(self, pre, post, bottom, numMG, iterMax, eps, hang, normThresh):
        """AMRMultiGrid::setSolverParameters (AMRMultiGrid.H:528-553)."""
        self.m_pre, self.m_post, self.m_bottom = pre, post, bottom
        self.m_numMG = numMG
        self.m_iterMax, self.m_eps, self.m_hang, self.m_normThresh = iterMax, eps, hang, normThresh
        for mg in self.m_mg:
            if mg is not None:
                self._mg_params(mg)

    def _mg_params(self, mg):
        mg.m_pre, mg.m_post, mg.m_bottom = self.m_pre, self.m_post, self.m_bottom
        mg.m_numMG = self.m_numMG
        mg.m_cycle = self.m_numMG

    def init(self, phi, rhs, l_max, l_base):
        """AMRMultiGrid::init (AMRMultiGrid.H:1212-1249)."""
        for i in range(l_base, l_max + 1):
            op = self.m_op[i]
            self.m_correction[i] = op.create(phi[i])
            self.m_residual[i] = op.create(rhs[i])
            if i == l_base:
                mg = MultiGrid()
                self._mg_params(mg)
                mg.define(op, self.m_domainCells[i], self.m_bottomSolver)
                mg.init(phi[i], rhs[i])
                self.m_mg[i] = mg
            else:
                self.m_resC[i] = op.createCoarsened(rhs[i], op.refToCoarser())

    def revert(self, l_max, l_base):
        for i in range(l_base, l_max + 1):
            for f in (self.m_correction[i], self.m_residual[i], self.m_resC[i]):
                if f is not None:
                    f.close()
            self.m_correction[i] = self.m_residual[i] = self.m_resC[i] = None
            if self.m_mg[i] is not None:
                self.m_mg[i].close()
                self.m_mg[i] = None

    def solve(self, phi, rhs, l_max, l_base, zeroPhi=True):
        """AMRMultiGrid::solve (AMRMultiGrid.H:895-935)."""
        self.init(phi, rhs, l_max, l_base)
        try:
            lowlim = l_base - 1 if l_base > 0 else l_base
            uberResidual = [None] * len(phi)
            uberCorrection = [None] * len(phi)
            for ilev in range(lowlim, l_max + 1):
                uberCorrection[ilev] = self.m_op[ilev].create(phi[ilev])
                if ilev >= l_base:
                    uberResidual[ilev] = self.m_op[ilev].create(rhs[ilev])
            try:
                self.solveNoInitResid(phi, uberResidual, uberCorrection, rhs, l_max, l_base, zeroPhi)
            finally:
                for f in uberResidual + uberCorrection:
                    if f is not None:
                        f.close()
        finally:
            self.revert(l_max, l_base)

    def solveNoInitResid(self, phi, uberResidual, uberCorrection, rhs, l_max, l_base, zeroPhi):
        """AMRMultiGrid::solveNoInitResid (AMRMultiGrid.H:935-1124)."""
        self.m_mg[l_base].setBottomSolver(self.m_bottomSolver)
        lowlim = l_base - 1 if l_base > 0 else l_base
        for ilev in range(lowlim, l_max + 1):
            if ilev >= l_base:
                self.m_op[ilev].setToZero(uberResidual[ilev])
            self.m_op[ilev].setToZero(uberCorrection[ilev])
        if zeroPhi:
            for ilev in range(l_base, l_max + 1):
                self.m_op[ilev].setToZero(phi[ilev])
        initial = self.computeAMRResidual(uberResidual, phi, rhs, l_max, l_base, False, True)
        if self.m_convergenceMetric != 0.0:
            initial = self.m_convergenceMetric
        rnorm = initial
        norm_last = 2 * initial
        self.m_bottomSolver.setConvergenceMetrics(initial, self.m_bottomSolverEpsCushion * self.m_eps)
        it = 0
        self.m_history = [rnorm]
        goNorm = rnorm > self.m_normThresh
        goRedu = rnorm > self.m_eps * initial
        goIter = it < self.m_iterMax
        goHang = it < self.m_imin or rnorm < (1 - self.m_hang) * norm_last
        goMin = it < self.m_iterMin
        while goMin or (goIter and goRedu and goHang and goNorm):
            norm_last = rnorm
            self.AMRVCycle(uberCorrection, uberResidual, l_max, l_max, l_base)
            for ilev in range(l_base, l_max + 1):
                self.m_op[ilev].incr(phi[ilev], uberCorrection[ilev], 1.0)
                self.m_op[ilev].setToZero(uberCorrection[ilev])
            rnorm = self.computeAMRResidual(uberResidual, phi, rhs, l_max, l_base, False, True)
            it += 1
            self.m_history.append(rnorm)
            goNorm = rnorm > self.m_normThresh
            goRedu = rnorm > self.m_eps * initial
            goIter = it < self.m_iterMax
            goHang = it < self.m_imin or rnorm < (1 - self.m_hang) * norm_last
            goMin = it < self.m_iterMin
        self.m_iterations = it
        self.m_exitStatus = int(not goRedu) + int(not goIter) * 2 + int(not goHang) * 4 + int(not goNorm) * 8

    def computeAMRResidualLevel(self, resid, phi, rhs, l_max, l_base, ilev, homogeneousBC):
        """AMRMultiGrid::computeAMRResidualLevel (AMRMultiGrid.H:837-892)."""
        op = self.m_op
        if l_max != l_base:
            if ilev == l_max:
                op[l_max].AMRResidualNF(resid[l_max], phi[l_max], phi[l_max - 1], rhs[l_max], homogeneousBC)
            elif ilev == l_base:
                if l_base == 0:
                    op[l_base].AMRResidualNC(resid[l_base], phi[l_base + 1], phi[l_base], rhs[l_base], homogeneousBC,
                                             op[l_base + 1])
                else:
                    op[l_base].AMRResidual(resid[l_base], phi[l_base + 1], phi[l_base], phi[l_base - 1], rhs[l_base],
                                           homogeneousBC, op[l_base + 1])
            else:
                op[ilev].AMRResidual(resid[ilev], phi[ilev + 1], phi[ilev], phi[ilev - 1], rhs[ilev], homogeneousBC,
                                     op[ilev + 1])
        else:
            if l_base == 0:
                op[l_max].residual(resid[l_max], phi[l_max], rhs[l_max], homogeneousBC)
            else:
                op[l_max].AMRResidualNF(resid[l_max], phi[l_max], phi[l_max - 1], rhs[l_max], homogeneousBC)

    def computeAMRResidual(self, resid, phi, rhs, l_max, l_base, homogeneousBC, computeNorm):
        """AMRMultiGrid::computeAMRResidual (AMRMultiGrid.H:683-736)."""
        rnorm = 0.0
        for ilev in range(l_base, l_max + 1):
            self.computeAMRResidualLevel(resid, phi, rhs, l_max, l_base, ilev, homogeneousBC)
            if computeNorm:
                if ilev != l_max:
                    self.m_op[ilev].zeroCovered(resid[ilev])
                rnorm = max(self.m_op[ilev].localMaxNorm(resid[ilev]), rnorm)
        return rnorm

    def relax(self, correction, residual, depth, nRelax):
        """AMRMultiGrid::relax (AMRMultiGrid.H:634-666).  With ratio 4 to the coarser
        level, the level's MultiGrid cycles with m_depth 1 (init's intermediateDepth),
        which is its bottom branch (MultiGrid.H:634-646): one relaxation when its
        m_bottomCells is 1, else m_bottom, and the no-op bottom solver."""
        op = self.m_op[depth]
        if op.refToCoarser() > 2:
            nRelax = 1 if self._bottom_cells(depth) == 1 else self.m_bottom
        op.relax(correction, residual, nRelax)

    def _bottom_cells(self, depth):
        """MultiGrid::define's m_bottomCells for the level's MultiGrid: its domain over
        the multigrid levels MGnewOp would create (coarsenable by coarsening*2)."""
        cells, c = self.m_domainCells[depth], 2
        boxes = self.m_op[depth].level.boxes
        while c < (1 << 24) and is_coarsenable(boxes, c * 2):
            cells //= 8
            c *= 2
        return cells

    def AMRVCycle(self, uberCorrection, uberResidual, ilev, l_max, l_base):
        """AMRMultiGrid::AMRVCycle (AMRMultiGrid.H:1342-1419)."""
        op = self.m_op
        if ilev == l_max:
            for level in range(l_base, l_max + 1):
                op[level].assignLocal(self.m_residual[level], uberResidual[level])
                op[level].setToZero(self.m_correction[level])
        if l_max == l_base:
            self.m_mg[l_base].oneCycle(uberCorrection[ilev], uberResidual[ilev])
        elif ilev == l_base:
            self.m_mg[l_base].oneCycle(self.m_correction[ilev], self.m_residual[ilev])
            op[ilev].incr(uberCorrection[ilev], self.m_correction[ilev], 1.0)
        else:
            self.relax(self.m_correction[ilev], self.m_residual[ilev], ilev, self.m_pre)
            op[ilev].incr(uberCorrection[ilev], self.m_correction[ilev], 1.0)
            op[ilev - 1].setToZero(self.m_correction[ilev - 1])
            self.computeAMRResidualLevel(self.m_residual, uberCorrection, uberResidual, l_max, l_base, ilev - 1, True)
            op[ilev].AMRRestrictS(self.m_resC[ilev], self.m_residual[ilev], self.m_correction[ilev],
                                  self.m_correction[ilev - 1], uberCorrection[ilev])
            op[ilev - 1].assignCopier(self.m_residual[ilev - 1], self.m_resC[ilev])
            for img in range(self.m_numMG):
                self.AMRVCycle(uberCorrection, uberResidual, ilev - 1, l_max, l_base)
            op[ilev].AMRProlongS(self.m_correction[ilev], self.m_correction[ilev - 1], self.m_resC[ilev])
            op[ilev].AMRUpdateResidual(self.m_residual[ilev], self.m_correction[ilev], self.m_correction[ilev - 1])
            dCorr = uberCorrection[ilev]
            op[ilev].setToZero(dCorr)
            self.relax(dCorr, self.m_residual[ilev], ilev, self.m_post)
            op[ilev].incr(self.m_correction[ilev], dCorr, 1.0)
            op[ilev].assignLocal(uberCorrection[ilev], self.m_correction[ilev])


def is_coarsenable(boxes, r):
    return all(lo % r == 0 and (hi + 1) % r == 0 for (L, H) in boxes for lo, hi in zip(L, H))


def domain_cells(domain):
    (lo, hi) = domain
    return math.prod(h - l + 1 for l, h in zip(lo, hi))
