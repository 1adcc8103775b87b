"""ctypes binding of the CPU oracle (oracle/amr_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, by the smoke check and by the
CPU-baseline leg of bench.py.  The product never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# another build of the same source (the sanitizer build) may be named in the environment
_LIB_PATH = os.environ.get("AMRMG_ORACLE_LIB") or os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
# BCValueFunc (BCFunc.H:205-208) with a caller pointer
BC_VALUE_FN = C.CFUNCTYPE(None, _dp, _ip, _ip, _dp, C.c_void_p)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def build_native(out_path):
    """Builds the CPU-baseline timing variant (-O3 -march=native, no contraction, no
    fast-math: the parity build's arithmetic) for the CPU this runs on."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "native", f"NATIVE_OUT={out_path}"], check=True)
    return out_path


def use_library(path):
    """Load the oracle from another build of the same source (before the first call)."""
    global _LIB_PATH, _lib
    if _lib is not None and path != _LIB_PATH:
        raise RuntimeError("the oracle library is already loaded")
    _LIB_PATH = path


def set_threads(n):
    lib().oracle_set_threads(int(n))


def get_threads():
    return int(lib().oracle_get_threads())


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [_ip, _dp]
        L.oracle_create_r.restype = C.c_void_p
        L.oracle_create_r.argtypes = [_ip, _dp, _ip]
        L.oracle_amr_prolong.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_level_cells.restype = C.c_long
        L.oracle_level_cells.argtypes = [C.c_void_p, C.c_int]
        L.oracle_mg_cells.restype = C.c_long
        L.oracle_mg_cells.argtypes = [C.c_void_p, C.c_int]
        L.oracle_ghosted_cells.restype = C.c_long
        L.oracle_ghosted_cells.argtypes = [C.c_void_p, C.c_int]
        L.oracle_num_mg_depths.argtypes = [C.c_void_p]
        L.oracle_solve.argtypes = [C.c_void_p, _ip, _dp, _dp, _dp, C.c_int, _dp, C.c_int, _ip]
        L.oracle_solve_levels.argtypes = [C.c_void_p, _ip, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, _dp, C.c_int,
                                          _ip]
        L.oracle_amr_residual.restype = C.c_double
        L.oracle_amr_residual.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int]
        L.oracle_amr_operator.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
        L.oracle_level_apply.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_int]
        L.oracle_level_residual.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, C.c_int]
        L.oracle_level_relax.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_int]
        L.oracle_level_restrict.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.oracle_level_prolong.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.oracle_kernel_gsrb.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_int]
        L.oracle_homog_cf.argtypes = [C.c_void_p, C.c_int, _dp]
        L.oracle_quad_cf.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.oracle_average.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.oracle_cell_relax_count.restype = C.c_longlong
        L.oracle_bottom_solve.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_double, C.c_double, _ip]
        L.oracle_vc_set.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.oracle_face_cells.restype = C.c_long
        L.oracle_face_cells.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.oracle_vc_lambda.argtypes = [C.c_void_p, C.c_int, _dp]
        L.oracle_set_alpha_beta.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.oracle_set_bc_values.argtypes = [C.c_void_p, BC_VALUE_FN, C.c_void_p]
        L.oracle_level_norm.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, _dp]
        L.oracle_level_dot.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.oracle_set_ncomp.argtypes = [C.c_void_p, C.c_int]
        L.oracle_set_relax_mode.argtypes = [C.c_void_p, C.c_int]
        L.oracle_set_vc_average.argtypes = [C.c_void_p, C.c_int]
        L.oracle_set_bottom_params.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]
        L.oracle_amrlevel_apply.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_int]
        L.oracle_amrlevel_residual.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, C.c_int]
        L.oracle_amrlevel_relax.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_int]
        L.oracle_cf_branch_counts.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_long)]
        L.oracle_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


class Oracle:
    """One AMR hierarchy held by the CPU oracle."""

    def __init__(self, hier):
        self.hier = hier
        isp, rsp = hier.spec()
        self._isp, self._rsp = isp, rsp
        L = lib()
        ratios = getattr(hier, "ratios", None)
        if ratios is None:
            self.h = L.oracle_create(isp.ctypes.data_as(_ip), rsp.ctypes.data_as(_dp))
        else:
            self._ratios = np.ascontiguousarray(ratios, dtype=np.int32)
            self.h = L.oracle_create_r(isp.ctypes.data_as(_ip), rsp.ctypes.data_as(_dp),
                                       self._ratios.ctypes.data_as(_ip))
        if not self.h:
            raise RuntimeError(L.oracle_last_error().decode())

    def close(self):
        if self.h:
            lib().oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def mg_depths(self) -> int:
        return lib().oracle_num_mg_depths(self.h)

    def mg_cells(self, d: int) -> int:
        return lib().oracle_mg_cells(self.h, d)

    def solve(self, rhs, sparams, rparams, phi0=None, hist_cap=128, l_max=None, l_base=0, zero_phi=None):
        """AMRMultiGrid::solve(phi, rhs, l_max, l_base, zeroPhi) on valid-packed fields of
        all levels; levels outside l_base..l_max keep phi0 (level l_base-1 gives the
        coarse-fine boundary values).  zero_phi defaults to phi0 is None."""
        rhs, prhs = _d(rhs)
        phi = np.zeros_like(rhs) if phi0 is None else np.array(phi0, dtype=np.float64)
        phi, pphi = _d(phi)
        hist = np.zeros(hist_cap)
        info = np.zeros(3, dtype=np.int32)
        sp = np.ascontiguousarray(sparams, dtype=np.int32)
        rp = np.ascontiguousarray(rparams, dtype=np.float64)
        lmax = self.hier.nlev - 1 if l_max is None else int(l_max)
        zp = (phi0 is None) if zero_phi is None else bool(zero_phi)
        rc = lib().oracle_solve_levels(self.h, sp.ctypes.data_as(_ip), rp.ctypes.data_as(_dp), prhs, pphi,
                                       1 if zp else 0, lmax, int(l_base), hist.ctypes.data_as(_dp), hist_cap,
                                       info.ctypes.data_as(_ip))
        if rc != 0:
            raise RuntimeError(lib().oracle_last_error().decode())
        it = int(info[0])
        return phi, hist[:it + 1].copy(), dict(iterations=it, exit_status=int(info[1]),
                                               bottom_calls=int(info[2]))

    def amr_residual(self, phi, rhs, homog=False):
        phi, pp = _d(phi)
        rhs, pr = _d(rhs)
        res = np.zeros_like(rhs)
        n = lib().oracle_amr_residual(self.h, pp, pr, res.ctypes.data_as(_dp), int(homog))
        return res, n

    def amr_operator(self, phi, homog=False):
        phi, pp = _d(phi)
        out = np.zeros_like(phi)
        lib().oracle_amr_operator(self.h, pp, out.ctypes.data_as(_dp), int(homog))
        return out

    def level_apply(self, depth, phi, homog=False):
        phi, pp = _d(phi)
        out = np.zeros_like(phi)
        self._chk(lib().oracle_level_apply(self.h, depth, pp, out.ctypes.data_as(_dp), int(homog)))
        return out

    def level_residual(self, depth, phi, rhs, homog=False):
        phi, pp = _d(phi)
        rhs, pr = _d(rhs)
        out = np.zeros_like(phi)
        self._chk(lib().oracle_level_residual(self.h, depth, pp, pr, out.ctypes.data_as(_dp), int(homog)))
        return out

    def level_relax(self, depth, phi, rhs, iters):
        phi = np.array(phi, dtype=np.float64)
        rhs, pr = _d(rhs)
        self._chk(lib().oracle_level_relax(self.h, depth, phi.ctypes.data_as(_dp), pr, iters))
        return phi

    def level_restrict(self, depth, phi, rhs):
        phi, pp = _d(phi)
        rhs, pr = _d(rhs)
        out = np.zeros(self.mg_cells(depth + 1))
        self._chk(lib().oracle_level_restrict(self.h, depth, pp, pr, out.ctypes.data_as(_dp)))
        return out

    def level_prolong(self, depth, phi, ec):
        phi = np.array(phi, dtype=np.float64)
        ec, pe = _d(ec)
        self._chk(lib().oracle_level_prolong(self.h, depth, phi.ctypes.data_as(_dp), pe))
        return phi

    def kernel_gsrb(self, lev, phi_ghosted, rhs, color):
        phi = np.array(phi_ghosted, dtype=np.float64)
        rhs, pr = _d(rhs)
        self._chk(lib().oracle_kernel_gsrb(self.h, lev, phi.ctypes.data_as(_dp), pr, color))
        return phi

    def homog_cf(self, lev, phi_ghosted):
        phi = np.array(phi_ghosted, dtype=np.float64)
        self._chk(lib().oracle_homog_cf(self.h, lev, phi.ctypes.data_as(_dp)))
        return phi

    def quad_cf(self, lev, phi_ghosted, coarse):
        phi = np.array(phi_ghosted, dtype=np.float64)
        coarse, pc = _d(coarse)
        self._chk(lib().oracle_quad_cf(self.h, lev, phi.ctypes.data_as(_dp), pc))
        return phi

    def average(self, lev, fine):
        fine, pf = _d(fine)
        out = np.zeros(self.hier.cells(lev) // self.hier.ratio(lev) ** 3)
        self._chk(lib().oracle_average(self.h, lev, pf, out.ctypes.data_as(_dp)))
        return out

    def amr_prolong(self, lev, fine, coarse):
        """AMRProlongS's PROLONG on AMR level lev: fine + coarse (coarse valid-packed
        over the level's boxes coarsened by its ratio)."""
        out = np.array(fine, dtype=np.float64)
        coarse, pc = _d(coarse)
        self._chk(lib().oracle_amr_prolong(self.h, lev, out.ctypes.data_as(_dp), pc))
        return out

    def bottom_solve(self, corr, res, nbottom, metric, eps, iterations=False):
        """relax(nbottom) + BiCGStabSolver::solve on the deepest multigrid level; returns
        (correction, exit status) or, with iterations=True, (correction, status, iterations)."""
        corr = np.array(corr, dtype=np.float64)
        res, pr = _d(res)
        info = np.zeros(2, dtype=np.int32)
        self._chk(lib().oracle_bottom_solve(self.h, corr.ctypes.data_as(_dp), pr, nbottom, metric, eps,
                                            info.ctypes.data_as(_ip)))
        if iterations:
            return corr, int(info[0]), int(info[1])
        return corr, int(info[0])

    def set_bottom_params(self, imax=80, numRestarts=5, reps=1e-12, hang=1e-8, small=1e-30):
        """BiCGStabSolver's public parameters for every later bottom solve."""
        self._chk(lib().oracle_set_bottom_params(self.h, int(imax), int(numRestarts), float(reps), float(hang),
                                                 float(small)))

    def amrlevel_apply(self, lev, phi, homog=False):
        phi, pp = _d(phi)
        out = np.zeros_like(phi)
        self._chk(lib().oracle_amrlevel_apply(self.h, lev, pp, out.ctypes.data_as(_dp), int(homog)))
        return out

    def amrlevel_residual(self, lev, phi, rhs, homog=False):
        phi, pp = _d(phi)
        rhs, pr = _d(rhs)
        out = np.zeros_like(phi)
        self._chk(lib().oracle_amrlevel_residual(self.h, lev, pp, pr, out.ctypes.data_as(_dp), int(homog)))
        return out

    def amrlevel_relax(self, lev, phi, rhs, iters):
        phi = np.array(phi, dtype=np.float64)
        rhs, pr = _d(rhs)
        self._chk(lib().oracle_amrlevel_relax(self.h, lev, phi.ctypes.data_as(_dp), pr, iters))
        return phi

    CF_BRANCHES = ("standard", "special", "mixed_dropped", "centred", "shift_plus", "shift_minus", "drop_plus",
                   "drop_minus", "drop_zero", "mixed1", "mixed2", "mixed3", "mixed4")

    def cf_branch_counts(self, lev):
        """QuadCFStencil branch statistics of level lev (dict keyed by CF_BRANCHES)."""
        out = (C.c_long * 13)()
        self._chk(lib().oracle_cf_branch_counts(self.h, lev, out))
        return dict(zip(self.CF_BRANCHES, [int(v) for v in out]))

    def vc_set(self, a, b):
        """Variable coefficients (VCAMRPoissonOp2) on every level: a valid-packed over all
        levels, b = [b0, b1, b2], each level after level, per box its faces normal to d
        in Fortran order."""
        a, pa = _d(a)
        bs = [_d(x) for x in b]
        self._vc = (a, bs)  # keep the buffers alive for the call
        self._chk(lib().oracle_vc_set(self.h, pa, bs[0][1], bs[1][1], bs[2][1]))

    def face_cells(self, lev, d):
        return lib().oracle_face_cells(self.h, lev, d)

    def level_norm(self, depth, x, p):
        x, px = _d(x)
        out = np.zeros(1)
        self._chk(lib().oracle_level_norm(self.h, depth, px, int(p), out.ctypes.data_as(_dp)))
        return float(out[0])

    def level_dot(self, depth, a, b):
        a, pa = _d(a)
        b, pb = _d(b)
        out = np.zeros(1)
        self._chk(lib().oracle_level_dot(self.h, depth, pa, pb, out.ctypes.data_as(_dp)))
        return float(out[0])

    def set_ncomp(self, ncomp):
        """Components of the solve / AMR residual / operator / level norm and dot fields
        (valid-packed box by box, each box component by component)."""
        self._chk(lib().oracle_set_ncomp(self.h, int(ncomp)))
        self.ncomp = int(ncomp)

    def set_vc_average(self, harmonic):
        """Coefficient coarsening of the next vc_set: arithmetic (False) or harmonic."""
        self._chk(lib().oracle_set_vc_average(self.h, 1 if harmonic else 0))

    def set_relax_mode(self, mode):
        """AMRPoissonOp::s_relaxMode: 0 looseGSRB, 1 levelGSRB, 2 overlapGSRB,
        3 levelGSRBLazy, 4 levelJacobi, 5 levelMultiColor."""
        self._chk(lib().oracle_set_relax_mode(self.h, int(mode)))

    def set_alpha_beta(self, alpha, beta):
        """setAlphaAndBeta on every operator (AMRMultiGrid::getAllOperators)."""
        self._chk(lib().oracle_set_alpha_beta(self.h, float(alpha), float(beta)))

    def set_bc_values(self, fn):
        """DiriBC / NeumBC with a BCValueHolder: fn(pos, dir, side) -> value (or one
        value per component) at each domain face centre, called at every inhomogeneous
        ghost fill; None restores the constant face values."""
        if fn is None:
            cb = C.cast(None, BC_VALUE_FN)
        else:
            def _cb(pos, d, s, v, _user):
                val = fn((pos[0], pos[1], pos[2]), d[0], s[0])
                for c, x in enumerate(np.atleast_1d(val)):
                    v[c] = float(x)
            cb = BC_VALUE_FN(_cb)
        self._chk(lib().oracle_set_bc_values(self.h, cb, None))
        self._bc_fn = cb

    # ---- heat solvers over the hierarchy's AMRMultiGrid (valid-packed fields) ----
    def _acoef(self):
        return self._vc[0] if getattr(self, "_vc", None) is not None else None

    def backward_euler_step(self, phi_old, src, dt, sp, rp, zero_phi=True):
        """BackwardEuler::oneStep (BackwardEuler.cpp:145-188) with createEulerRHS
        (:196-233) and solveHelm (:236-254): rhs = a phi^n + dt src (a = 1 for
        AMRPoissonOp, aCoef for VCAMRPoissonOp2), then alpha = 1, beta = -dt and
        AMRMultiGrid::solve from phi^n (or zero)."""
        a = self._acoef()
        rhs = 0.0 + np.asarray(phi_old, dtype=np.float64)      # setVal(0); += phiOld
        if a is not None:
            rhs = rhs * a                                      # diagonalScale
        rhs = rhs + (0.0 + np.asarray(src, dtype=np.float64)) * dt   # scaleRho = dt * rho; += scaleRho
        self.set_alpha_beta(1.0, -dt)
        return self.solve(rhs, sp, rp, phi0=None if zero_phi else phi_old)

    @staticmethod
    def tga_mu():
        """AMRTGA::setMu (AMRTGA.H:368-384)."""
        eps = 1.e-12
        a = 2.0 - np.sqrt(2.0) - eps
        r = np.sqrt(a * a - 4.0 * a + 2.0)
        return (a - r) / 2.0, (a + r) / 2.0, (1.0 - a), 0.5 - a

    def tga_step(self, phi_old, src, dt, sp, rp):
        """AMRTGA::oneStep (AMRTGA.H:420-510) with applyHelm (:514-537) and solveHelm
        (:540-560).  Returns phi^{n+1} and the two solves' (history, info)."""
        mu1, mu2, mu3, mu4 = self.tga_mu()
        a = self._acoef()
        phi_old = np.asarray(phi_old, dtype=np.float64)
        srct = 0.0 + np.asarray(src, dtype=np.float64) * 1.0   # setToZero; incr(src, 1)
        if a is not None:
            srct = srct / a                                    # divideByIdentityCoef
        self.set_alpha_beta(1.0, mu4 * dt)
        rhst = self.amr_operator(srct, homog=True)             # (I + mu4 dt L) S/a
        rhst = rhst * dt                                       # scale(dt)
        self.set_alpha_beta(1.0, mu3 * dt)
        phi_new = self.amr_operator(phi_old, homog=False)      # (I + mu3 dt L) phi^n
        rhst = rhst + phi_new * 1.0                            # incr
        self.set_alpha_beta(1.0, -mu2 * dt)
        phi_new, h1, i1 = self.solve(rhst, sp, rp, phi0=phi_old)    # solveNoInit, zeroPhi false
        rhst = phi_new.copy()                                  # assignLocal
        if a is not None:
            rhst = rhst * a                                    # diagonalScale
        self.set_alpha_beta(1.0, -mu1 * dt)
        phi_new, h2, i2 = self.solve(rhst, sp, rp, phi0=phi_old)
        return phi_new, (h1, i1), (h2, i2)

    def vc_lambda(self, depth):
        out = np.zeros(self.mg_cells(depth))
        self._chk(lib().oracle_vc_lambda(self.h, depth, out.ctypes.data_as(_dp)))
        return out

    @staticmethod
    def cell_relax_count():
        return int(lib().oracle_cell_relax_count())

    @staticmethod
    def reset_counters():
        lib().oracle_reset_counters()

    @staticmethod
    def _chk(rc):
        if rc != 0:
            raise RuntimeError(lib().oracle_last_error().decode())
