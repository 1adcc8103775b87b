"""ctypes binding of libamrmg.so (C ABI declared in include/amrmg.h).

Host-side mirror of the reference's driver-level API for the elliptic path:
``AMRPoissonOpFactory`` + ``AMRMultiGrid`` become one ``AMRMultiGrid`` object
built from a ``Hierarchy`` (hierarchy.py) and device-resident fields; the level
operations of ``AMRPoissonOp`` are exposed for parity checks.  Everything runs in
the HIP library: this module only moves host buffers across the C ABI, and it
raises if the library is missing (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# another build of the same sources (the sanitizer build) may be named in the environment
LIB_PATH = os.environ.get("AMRMG_LIB") or os.path.join(_HERE, "libamrmg.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "amrmg.h")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
# amrmg_bc_value_fn: BCValueFunc (BCFunc.H:205-208) with a caller pointer
BC_VALUE_FN = C.CFUNCTYPE(None, _dp, _ip, _ip, _dp, C.c_void_p)


def bc_value_callback(fn):
    """An amrmg_bc_value_fn from fn(pos, dir, side) -> value, or a sequence of one
    value per component (pos a 3-tuple of face coordinates, side 0 low / 1 high); None
    gives a NULL pointer.  Keep the result alive while the library may call it."""
    if fn is None:
        return C.cast(None, BC_VALUE_FN)
    if isinstance(fn, BC_VALUE_FN):
        return fn

    def cb(pos, d, s, v, _user):
        val = fn((pos[0], pos[1], pos[2]), d[0], s[0])
        for c, x in enumerate(np.atleast_1d(val)):
            v[c] = float(x)
    return BC_VALUE_FN(cb)
_lib = None

KERNEL_CLASSES = ["gsrb", "residual", "apply", "restrict", "prolong", "average", "ghost", "blas",
                  "reduce", "cf_interp", "reflux", "copier"]


class AMRMGError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AMRMGError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        L.amrmg_last_error.restype = C.c_char_p
        L.amrmg_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        L.amrmg_ctx_destroy.argtypes = [C.c_void_p]
        L.amrmg_ctx_set_fused_relax.argtypes = [C.c_void_p, C.c_int]
        L.amrmg_profile_read_units.argtypes = [C.c_void_p, C.c_int, _dp]
        L.amrmg_solver_set_alpha_beta.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.amrmg_solver_set_vc_coefs.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.amrmg_level_norm.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, _dp]
        L.amrmg_solver_set_relax_mode.argtypes = [C.c_void_p, C.c_int]
        L.amrmg_solver_set_vc_average.argtypes = [C.c_void_p, C.c_int]
        L.amrmg_heat_create.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]
        L.amrmg_heat_set_source.argtypes = [C.c_void_p, _dp]
        L.amrmg_heat_step.argtypes = [C.c_void_p, C.c_double, C.c_int, _ip]
        L.amrmg_heat_destroy.argtypes = [C.c_void_p]
        L.amrmg_serve_create.argtypes = [C.c_int, C.c_int, C.c_longlong, C.c_int, C.POINTER(C.c_void_p)]
        L.amrmg_serve_set_params.argtypes = [C.c_void_p, _ip, _dp]
        L.amrmg_serve_submit.argtypes = [C.c_void_p, _ip, _dp, _dp, C.POINTER(C.c_longlong)]
        L.amrmg_serve_set_lanes.argtypes = [C.c_void_p, C.c_int, C.c_longlong]
        L.amrmg_serve_wait.argtypes = [C.c_void_p, C.c_longlong, _dp, _dp, _ip]
        L.amrmg_serve_stats.argtypes = [C.c_void_p, C.POINTER(C.c_longlong), C.c_int]
        L.amrmg_serve_destroy.argtypes = [C.c_void_p]
        L.amrmg_batch_create.argtypes = [C.c_void_p, C.c_int, _ip, _dp, C.c_int, C.POINTER(C.c_void_p)]
        L.amrmg_batch_create_lanes.argtypes = [C.c_void_p, C.c_int, _ip, _dp, C.c_int, C.c_int,
                                               C.POINTER(C.c_void_p)]
        L.amrmg_batch_lanes.argtypes = [C.c_void_p, _ip, _ip]
        L.amrmg_batch_destroy.argtypes = [C.c_void_p]
        L.amrmg_batch_set_params.argtypes = [C.c_void_p, _ip, _dp]
        L.amrmg_batch_set_rhs.argtypes = [C.c_void_p, C.c_int, _dp]
        L.amrmg_batch_get_phi.argtypes = [C.c_void_p, C.c_int, _dp]
        L.amrmg_batch_solve.argtypes = [C.c_void_p, _ip, _ip]
        L.amrmg_batch_history.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, _ip]
        L.amrmg_batch_tune_relax.argtypes = [C.c_void_p, _dp, C.c_int, _ip]
        L.amrmg_batch_set_relax_shape.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.amrmg_batch_stats.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.amrmg_batch_set_graph.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_longlong)]
        L.amrmg_batch_set_reduction_order.argtypes = [C.c_void_p, C.c_int]
        L.amrmg_solver_create.argtypes = [C.c_void_p, _ip, _dp, C.c_int, C.POINTER(C.c_void_p)]
        L.amrmg_solver_create_ratios.argtypes = [C.c_void_p, _ip, _dp, _ip, C.c_int, C.POINTER(C.c_void_p)]
        L.amrmg_solver_create_ncomp.argtypes = [C.c_void_p, _ip, _dp, _ip, C.c_int, C.c_int,
                                                C.POINTER(C.c_void_p)]
        L.amrmg_solver_destroy.argtypes = [C.c_void_p]
        L.amrmg_solver_set_params.argtypes = [C.c_void_p, _ip, _dp]
        L.amrmg_solver_set_bc_values.argtypes = [C.c_void_p, BC_VALUE_FN, C.c_void_p]
        L.amrmg_solver_level_cells.restype = C.c_longlong
        L.amrmg_solver_level_cells.argtypes = [C.c_void_p, C.c_int]
        L.amrmg_solver_num_mg_depths.argtypes = [C.c_void_p]
        L.amrmg_solver_mg_cells.restype = C.c_longlong
        L.amrmg_solver_mg_cells.argtypes = [C.c_void_p, C.c_int]
        L.amrmg_solver_device_bytes.restype = C.c_longlong
        L.amrmg_solver_device_bytes.argtypes = [C.c_void_p]
        L.amrmg_solver_tune_relax.argtypes = [C.c_void_p, _dp, C.c_int, _ip]
        L.amrmg_solver_set_relax_shape.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.amrmg_solver_set_field.argtypes = [C.c_void_p, C.c_int, _dp]
        L.amrmg_solver_get_field.argtypes = [C.c_void_p, C.c_int, _dp]
        L.amrmg_solver_solve.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, _ip]
        L.amrmg_solver_solve_levels.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, C.c_int, _ip]
        L.amrmg_solver_residual.argtypes = [C.c_void_p, C.c_int, _dp]
        L.amrmg_solver_operator.argtypes = [C.c_void_p, C.c_int]
        L.amrmg_level_apply.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_int]
        L.amrmg_level_residual.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, C.c_int]
        L.amrmg_level_relax.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_int]
        L.amrmg_level_restrict.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.amrmg_level_prolong.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.amrmg_kernel_gsrb.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_int]
        L.amrmg_homog_cf.argtypes = [C.c_void_p, C.c_int, _dp]
        L.amrmg_quad_cf.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.amrmg_average.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.amrmg_bottom_solve.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_double, C.c_double, _ip]
        L.amrmg_bottom_solve_device.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_double, C.c_double, _ip]
        L.amrmg_solver_set_bottom_params.argtypes = [C.c_void_p, _ip, _dp]
        L.amrmg_plan_check.argtypes = [_ip, _dp, C.POINTER(C.c_longlong), C.c_int]
        L.amrmg_plan_check_ratios.argtypes = [_ip, _dp, _ip, C.POINTER(C.c_longlong), C.c_int]
        L.amrmg_copy_probe.argtypes = [C.c_void_p, C.c_longlong, C.c_int, _dp]
        L.amrmg_profile_enable.argtypes = [C.c_void_p, C.c_int]
        L.amrmg_profile_reset.argtypes = [C.c_void_p]
        L.amrmg_profile_read.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.POINTER(C.c_longlong)]
        _lib = L
    return _lib


def _chk(rc):
    if rc != 0:
        raise AMRMGError(lib().amrmg_last_error().decode())


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def header_symbols():
    """Entry points declared in include/amrmg.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(amrmg_[a-z0-9_]+)\s*\(", txt)))


def plan_check(hier):
    """Build and bounds-check every device table of `hier` on the host (no device)."""
    isp, rsp = hier.spec()
    st = (C.c_longlong * 8)()
    ratios = getattr(hier, "ratios", None)
    if ratios is None:
        _chk(lib().amrmg_plan_check(isp.ctypes.data_as(_ip), rsp.ctypes.data_as(_dp), st, 8))
    else:
        r = np.ascontiguousarray(ratios, dtype=np.int32)
        _chk(lib().amrmg_plan_check_ratios(isp.ctypes.data_as(_ip), rsp.ctypes.data_as(_dp),
                                           r.ctypes.data_as(_ip), st, 8))
    keys = ["levels", "mg_depths", "ghost_patches", "cf_coarse", "cf_fine", "reflux_cells", "tiles", "cells"]
    return dict(zip(keys, list(st)))


class Context:
    """A HIP device context (one stream at default priority)."""

    def __init__(self, device: int = 0):
        p = C.c_void_p()
        _chk(lib().amrmg_ctx_create(device, C.byref(p)))
        self.p = p

    def close(self):
        if self.p:
            lib().amrmg_ctx_destroy(self.p)
            self.p = None

    def profile(self, classes=None):
        """Time kernel classes with HIP events (classes: names, None = all, [] = off)."""
        if classes is None:
            mask = -1
        else:
            mask = 0
            for c in classes:
                mask |= 1 << (KERNEL_CLASSES.index(c) if isinstance(c, str) else int(c))
        _chk(lib().amrmg_profile_enable(self.p, mask))

    def set_fused_relax(self, on: bool):
        """One-launch GSRB iterations on merged levels (True, the default) or the
        reference's four-launch sequence (False); results are bitwise the same."""
        _chk(lib().amrmg_ctx_set_fused_relax(self.p, 1 if on else 0))

    def copy_probe(self, nbytes, iters=10):
        """Average ms of one device copy of nbytes (nbytes read + nbytes written)."""
        ms = C.c_double()
        _chk(lib().amrmg_copy_probe(self.p, int(nbytes), int(iters), C.byref(ms)))
        return ms.value

    def profile_reset(self):
        _chk(lib().amrmg_profile_reset(self.p))

    def profile_read(self, cls):
        k = KERNEL_CLASSES.index(cls) if isinstance(cls, str) else int(cls)
        ms, by, n = C.c_double(), C.c_double(), C.c_longlong()
        _chk(lib().amrmg_profile_read(self.p, k, C.byref(ms), C.byref(by), C.byref(n)))
        u = C.c_double()
        _chk(lib().amrmg_profile_read_units(self.p, k, C.byref(u)))
        return dict(ms=ms.value, bytes=by.value, launches=n.value, units=u.value)


class AMRMultiGrid:
    """AMRPoissonOpFactory::define + AMRMultiGrid::define/init over a Hierarchy,
    with a BiCGStab bottom solver; fields phi/rhs live in HBM.  ncomp > 1: fields of
    that many components, valid-packed box by box, each box component by component."""

    PHI, RHS, RES = 0, 1, 2

    def __init__(self, ctx: Context, hier, precision: int = 0, ncomp: int = 1):
        self.ctx = ctx
        self.hier = hier
        self.ncomp = int(ncomp)
        isp, rsp = hier.spec()
        p = C.c_void_p()
        ratios = getattr(hier, "ratios", None)
        if self.ncomp != 1:
            r = None if ratios is None else np.ascontiguousarray(ratios, dtype=np.int32)
            _chk(lib().amrmg_solver_create_ncomp(ctx.p, isp.ctypes.data_as(_ip), rsp.ctypes.data_as(_dp),
                                                  None if r is None else r.ctypes.data_as(_ip), self.ncomp,
                                                  precision, C.byref(p)))
        elif ratios is None:
            _chk(lib().amrmg_solver_create(ctx.p, isp.ctypes.data_as(_ip), rsp.ctypes.data_as(_dp), precision,
                                            C.byref(p)))
        else:
            r = np.ascontiguousarray(ratios, dtype=np.int32)
            _chk(lib().amrmg_solver_create_ratios(ctx.p, isp.ctypes.data_as(_ip), rsp.ctypes.data_as(_dp),
                                                   r.ctypes.data_as(_ip), precision, C.byref(p)))
        self.p = p
        self.ncells = sum(lib().amrmg_solver_level_cells(p, l) for l in range(hier.nlev)) * self.ncomp

    def close(self):
        if self.p:
            lib().amrmg_solver_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def mg_depths(self):
        return lib().amrmg_solver_num_mg_depths(self.p)

    def mg_cells(self, d):
        return lib().amrmg_solver_mg_cells(self.p, d)

    def device_bytes(self):
        return lib().amrmg_solver_device_bytes(self.p)

    # (column x, column y, threads, scheme): scheme 0 keeps five LDS planes, scheme 1
    # keeps the z neighbours in registers (one x pair of cells per thread)
    RELAX_SHAPES = [(64, 8, 256, 0), (32, 8, 256, 0), (64, 4, 256, 0), (32, 16, 256, 0), (128, 4, 256, 0),
                    (64, 16, 256, 0), (32, 16, 512, 0), (64, 16, 512, 0), (64, 32, 1024, 0), (128, 8, 1024, 0),
                    (64, 13, 512, 0), (32, 13, 256, 0), (64, 29, 1024, 0), (64, 26, 1024, 1), (32, 24, 512, 1),
                    (64, 11, 512, 1), (32, 10, 256, 1), (128, 11, 1024, 1), (16, 21, 256, 1),
                    (64, 11, 256, 2), (128, 11, 512, 2), (64, 26, 512, 2), (32, 24, 256, 2),
                    (128, 26, 1024, 2), (64, 56, 1024, 2), (16, 47, 256, 2)]
    RELAX_CHUNKS = [0, 8, 16, 32]

    def tune_relax(self):
        """Autotunes the one-launch relaxation on every level that uses it; returns one
        dict per tuned level (kind 'amr' or 'mg', index, cells, chosen shape and
        z-chunk, ms per iteration, all timings [shape][chunk])."""
        cap = 64
        w = 6 + len(self.RELAX_SHAPES) * len(self.RELAX_CHUNKS)
        tab = np.zeros(cap * w)
        n = C.c_int()
        _chk(lib().amrmg_solver_tune_relax(self.p, tab.ctypes.data_as(_dp), cap, C.byref(n)))
        out = []
        for r in tab[:min(n.value, cap) * w].reshape(-1, w):
            out.append(dict(kind="amr" if r[0] == 0 else "mg", index=int(r[1]), cells=int(r[2]),
                            shape=self.RELAX_SHAPES[int(r[3])], kc=int(r[4]), ms=float(r[5]),
                            times=r[6:].reshape(len(self.RELAX_SHAPES), len(self.RELAX_CHUNKS)).tolist()))
        return out

    def set_relax_shape(self, shape: int, kc: int = 0):
        """Forces tile shape index (into RELAX_SHAPES) and z-chunk (0 = automatic)."""
        _chk(lib().amrmg_solver_set_relax_shape(self.p, int(shape), int(kc)))

    def set_alpha_beta(self, alpha, beta):
        """AMRPoissonOp::setAlphaAndBeta on every operator of the solver."""
        _chk(lib().amrmg_solver_set_alpha_beta(self.p, float(alpha), float(beta)))

    def set_bc_values(self, fn):
        """DiriBC / NeumBC with a BCValueHolder on every operator: the inhomogeneous
        value of each Dirichlet or Neumann face cell is fn(pos, dir, side) at the face
        centre (BCFunc.cpp:187-202), one value per component; None restores the
        constant face values."""
        cb = bc_value_callback(fn)
        _chk(lib().amrmg_solver_set_bc_values(self.p, cb, None))
        self._bc_fn = cb

    def set_vc_average(self, harmonic):
        """Coefficient coarsening of the next set_vc_coefs: arithmetic (False, the
        default) or harmonic (VCAMRPoissonOp2Factory's m_coefficient_average_type)."""
        _chk(lib().amrmg_solver_set_vc_average(self.p, 1 if harmonic else 0))

    def set_relax_mode(self, mode):
        """AMRPoissonOp::s_relaxMode for every operator: 0 looseGSRB, 1 levelGSRB,
        2 overlapGSRB, 3 levelGSRBLazy, 4 levelJacobi, 5 levelMultiColor."""
        _chk(lib().amrmg_solver_set_relax_mode(self.p, int(mode)))

    def set_vc_relax_shape(self, shape):
        """Tile shape of the one-launch variable-coefficient relaxation (-1 automatic, 0..3)."""
        _chk(lib().amrmg_solver_set_vc_relax_shape(self.p, int(shape)))

    def set_vc_coefs(self, a, b):
        """Variable coefficients (VCAMRPoissonOp2Factory::define on a one-level
        hierarchy): a valid-packed, b = [b0, b1, b2], each per box its faces normal
        to d (the box grown by one face in d) in Fortran order.  a=None returns to
        constant coefficients."""
        if a is None:
            _chk(lib().amrmg_solver_set_vc_coefs(self.p, None, None, None, None))
            return
        a, pa = _d(a)
        if a.size != self.ncells // self.ncomp:
            raise ValueError("aCoef size does not match the hierarchy")
        bs = [_d(x) for x in b]
        if len(bs) != 3:
            raise ValueError("bCoef needs three directions")
        _chk(lib().amrmg_solver_set_vc_coefs(self.p, pa, bs[0][1], bs[1][1], bs[2][1]))

    def setSolverParameters(self, sparams, rparams):
        sp = np.ascontiguousarray(sparams, dtype=np.int32)
        rp = np.ascontiguousarray(rparams, dtype=np.float64)
        _chk(lib().amrmg_solver_set_params(self.p, sp.ctypes.data_as(_ip), rp.ctypes.data_as(_dp)))

    def set_field(self, field, data):
        data, pd = _d(data)
        if data.size != self.ncells:
            raise ValueError("field size does not match the hierarchy")
        _chk(lib().amrmg_solver_set_field(self.p, field, pd))

    def get_field(self, field):
        out = np.zeros(self.ncells)
        _chk(lib().amrmg_solver_get_field(self.p, field, out.ctypes.data_as(_dp)))
        return out

    def solve(self, zero_phi=True, hist_cap=128, l_max=None, l_base=0):
        """AMRMultiGrid::solve on levels l_base..l_max (default all); with l_base > 0 the
        phi of level l_base-1 as set gives the coarse-fine boundary values."""
        hist = np.zeros(hist_cap)
        info = np.zeros(3, dtype=np.int32)
        lmax = self.hier.nlev - 1 if l_max is None else int(l_max)
        _chk(lib().amrmg_solver_solve_levels(self.p, lmax, int(l_base), int(zero_phi), hist.ctypes.data_as(_dp),
                                              hist_cap, info.ctypes.data_as(_ip)))
        it = int(info[0])
        return hist[:it + 1].copy(), dict(iterations=it, exit_status=int(info[1]), bottom_iters=int(info[2]))

    def computeAMRResidual(self, homogeneous=False):
        n = C.c_double()
        _chk(lib().amrmg_solver_residual(self.p, int(homogeneous), C.byref(n)))
        return self.get_field(self.RES), n.value

    def computeAMROperator(self, homogeneous=False):
        _chk(lib().amrmg_solver_operator(self.p, int(homogeneous)))
        return self.get_field(self.RES)

    # ---- level operations ----
    def level_apply(self, depth, phi, homog=False):
        phi, pp = _d(phi)
        out = np.zeros_like(phi)
        _chk(lib().amrmg_level_apply(self.p, depth, pp, out.ctypes.data_as(_dp), int(homog)))
        return out

    def level_norm(self, depth, x, p):
        """norm(BoxLayoutData, interval, p) on multigrid depth `depth` (p = 0, 1, 2)."""
        x, px = _d(x)
        out = np.zeros(1)
        _chk(lib().amrmg_level_norm(self.p, int(depth), px, int(p), out.ctypes.data_as(_dp)))
        return float(out[0])

    def level_residual(self, depth, phi, rhs, homog=False):
        phi, pp = _d(phi)
        rhs, pr = _d(rhs)
        out = np.zeros_like(phi)
        _chk(lib().amrmg_level_residual(self.p, depth, pp, pr, out.ctypes.data_as(_dp), int(homog)))
        return out

    def level_relax(self, depth, phi, rhs, iters):
        phi = np.array(phi, dtype=np.float64)
        rhs, pr = _d(rhs)
        _chk(lib().amrmg_level_relax(self.p, depth, phi.ctypes.data_as(_dp), pr, iters))
        return phi

    def level_restrict(self, depth, phi, rhs):
        phi, pp = _d(phi)
        rhs, pr = _d(rhs)
        out = np.zeros(self.mg_cells(depth + 1))
        _chk(lib().amrmg_level_restrict(self.p, depth, pp, pr, out.ctypes.data_as(_dp)))
        return out

    def level_prolong(self, depth, phi, ec):
        phi = np.array(phi, dtype=np.float64)
        ec, pe = _d(ec)
        _chk(lib().amrmg_level_prolong(self.p, depth, phi.ctypes.data_as(_dp), pe))
        return phi

    def kernel_gsrb(self, lev, phi_ghosted, rhs, color):
        phi = np.array(phi_ghosted, dtype=np.float64)
        rhs, pr = _d(rhs)
        _chk(lib().amrmg_kernel_gsrb(self.p, lev, phi.ctypes.data_as(_dp), pr, color))
        return phi

    def homog_cf(self, lev, phi_ghosted):
        phi = np.array(phi_ghosted, dtype=np.float64)
        _chk(lib().amrmg_homog_cf(self.p, lev, phi.ctypes.data_as(_dp)))
        return phi

    def quad_cf(self, lev, phi_ghosted, coarse):
        phi = np.array(phi_ghosted, dtype=np.float64)
        coarse, pc = _d(coarse)
        _chk(lib().amrmg_quad_cf(self.p, lev, phi.ctypes.data_as(_dp), pc))
        return phi

    def average(self, lev, fine):
        fine, pf = _d(fine)
        out = np.zeros(self.hier.cells(lev) // self.hier.ratio(lev) ** 3)
        _chk(lib().amrmg_average(self.p, lev, pf, out.ctypes.data_as(_dp)))
        return out

    def bottom_solve(self, corr, res, nbottom, metric, eps, device=False):
        """relax(nbottom) + BiCGStabSolver::solve on the deepest multigrid level: the
        host-driven BiCGStab (returns (correction, exit status, iterations)) or the
        device bottom solver (device=True: (correction, None, iterations))."""
        corr = np.array(corr, dtype=np.float64)
        res, pr = _d(res)
        info = np.zeros(2, dtype=np.int32)
        fn = lib().amrmg_bottom_solve_device if device else lib().amrmg_bottom_solve
        _chk(fn(self.p, corr.ctypes.data_as(_dp), pr, nbottom, metric, eps, info.ctypes.data_as(_ip)))
        if device:
            return corr, None, int(info[0])
        return corr, int(info[0]), int(info[1])

    def set_bottom_params(self, imax=80, numRestarts=5, reps=1e-12, hang=1e-8, small=1e-30):
        """BiCGStabSolver's public parameters for every bottom solve of this solver."""
        ip = np.asarray([imax, numRestarts], dtype=np.int32)
        rp = np.asarray([reps, hang, small], dtype=np.float64)
        _chk(lib().amrmg_solver_set_bottom_params(self.p, ip.ctypes.data_as(_ip), rp.ctypes.data_as(_dp)))


class HeatSolver:
    """BackwardEuler (scheme "be", BackwardEuler.cpp) or AMRTGA (scheme "tga",
    AMRTGA.H) over an AMRMultiGrid: d phi/dt = L phi + S with the solver's operator
    as L.  step() advances the solver's PHI field by dt in place."""

    SCHEMES = {"be": 0, "tga": 1}

    def __init__(self, solver: "AMRMultiGrid", scheme: str = "tga"):
        self.solver = solver
        p = C.c_void_p()
        _chk(lib().amrmg_heat_create(solver.p, self.SCHEMES[scheme], C.byref(p)))
        self.p = p

    def close(self):
        if self.p:
            lib().amrmg_heat_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_source(self, src):
        src, ps = _d(src)
        if src.size != self.solver.ncells:
            raise ValueError("source size does not match the hierarchy")
        _chk(lib().amrmg_heat_set_source(self.p, ps))

    def step(self, dt, zero_phi=False):
        """One step; returns ({iterations, exit_status} of each solve)."""
        info = np.zeros(4, dtype=np.int32)
        _chk(lib().amrmg_heat_step(self.p, float(dt), 1 if zero_phi else 0, info.ctypes.data_as(_ip)))
        return (dict(iterations=int(info[0]), exit_status=int(info[1])),
                dict(iterations=int(info[2]), exit_status=int(info[3])))


class BatchSolver:
    """Many independent one-level problems solved together, each V-cycle stage of
    all of them in one launch (amrmg_batch_*).  Per problem the arithmetic is that
    of AMRMultiGrid on it alone.  lanes=2 runs the problems with large multigrid
    bottoms in a lockstep of their own on a second stream (same results)."""

    def __init__(self, ctx: Context, hiers, precision: int = 0, lanes: int = 1):
        self.ctx = ctx
        self.hiers = list(hiers)
        isps, rsps = [], []
        for h in self.hiers:
            if h.nlev != 1:
                raise ValueError("batched problems have one level")
            i, r = h.spec()
            isps.append(np.asarray(i, np.int32))
            rsps.append(np.asarray(r, np.float64))
        isp = np.ascontiguousarray(np.concatenate(isps))
        rsp = np.ascontiguousarray(np.concatenate(rsps))
        p = C.c_void_p()
        _chk(lib().amrmg_batch_create_lanes(ctx.p, len(self.hiers), isp.ctypes.data_as(_ip),
                                             rsp.ctypes.data_as(_dp), precision, lanes, C.byref(p)))
        self.p = p
        self.n = len(self.hiers)

    def lanes(self):
        """(number of lanes, lane of each problem)"""
        n = C.c_int()
        per = np.zeros(self.n, np.int32)
        _chk(lib().amrmg_batch_lanes(self.p, C.byref(n), per.ctypes.data_as(_ip)))
        return n.value, per

    def close(self):
        if self.p:
            lib().amrmg_batch_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setSolverParameters(self, sparams, rparams):
        sp = np.ascontiguousarray(sparams, dtype=np.int32)
        rp = np.ascontiguousarray(rparams, dtype=np.float64)
        _chk(lib().amrmg_batch_set_params(self.p, sp.ctypes.data_as(_ip), rp.ctypes.data_as(_dp)))

    def set_rhs(self, q, data):
        data, pd = _d(data)
        if data.size != self.hiers[q].total_cells():
            raise ValueError("rhs size does not match the problem")
        _chk(lib().amrmg_batch_set_rhs(self.p, int(q), pd))

    def get_phi(self, q):
        out = np.zeros(self.hiers[q].total_cells())
        _chk(lib().amrmg_batch_get_phi(self.p, int(q), out.ctypes.data_as(_dp)))
        return out

    def solve(self):
        it = np.zeros(self.n, np.int32)
        ex = np.zeros(self.n, np.int32)
        _chk(lib().amrmg_batch_solve(self.p, it.ctypes.data_as(_ip), ex.ctypes.data_as(_ip)))
        return it, ex

    def history(self, q, cap=256):
        h = np.zeros(cap)
        n = C.c_int()
        _chk(lib().amrmg_batch_history(self.p, int(q), h.ctypes.data_as(_dp), cap, C.byref(n)))
        return h[:min(n.value, cap)].copy()

    def tune_relax(self):
        ms = np.zeros(64)
        n = C.c_int()
        _chk(lib().amrmg_batch_tune_relax(self.p, ms.ctypes.data_as(_dp), 64, C.byref(n)))
        return ms[:n.value].copy()

    def set_relax_shape(self, depth, shape, kc=0):
        """Forces the relaxation tile shape (index into AMRMultiGrid.RELAX_SHAPES) and
        z-chunk of one multigrid depth."""
        _chk(lib().amrmg_batch_set_relax_shape(self.p, int(depth), int(shape), int(kc)))

    def stats(self):
        cr = C.c_double()
        bi, db = C.c_longlong(), C.c_longlong()
        _chk(lib().amrmg_batch_stats(self.p, C.byref(cr), C.byref(bi), C.byref(db)))
        return dict(cell_relax=cr.value, bottom_iters=bi.value, device_bytes=db.value)

    def set_reduction_order(self, order):
        """Bottom-solve dot products and norms: 0 the reference's order (default), 1 a
        fixed tree per problem (faster; same result alone and in any batch)."""
        _chk(lib().amrmg_batch_set_reduction_order(self.p, int(order)))

    def set_graph(self, on=None):
        """V-cycles replayed from one captured HIP graph (True, the default) or
        launched directly (False); None leaves the setting.  Returns the number of
        graph replays so far."""
        n = C.c_longlong()
        _chk(lib().amrmg_batch_set_graph(self.p, -1 if on is None else int(bool(on)), C.byref(n)))
        return n.value


class ServeEngine:
    """Request loop with continuous admission (amrmg_serve_*): `slots` requests run
    at once, each on its own HIP stream, admitted FIFO as slots and the HBM budget
    allow; hierarchies stay device-resident for later requests with the same spec."""

    def __init__(self, device: int = 0, slots: int = 4, budget_bytes: int = 64 << 30, precision: int = 0,
                 sparams=None, rparams=None, max_large=None, large_cells=1 << 22):
        p = C.c_void_p()
        _chk(lib().amrmg_serve_create(int(device), int(slots), int(budget_bytes), int(precision), C.byref(p)))
        self.p = p
        self._size = {}
        if max_large is not None:
            self.set_lanes(max_large, large_cells)
        if sparams is not None:
            sp = np.ascontiguousarray(sparams, dtype=np.int32)
            rp = np.ascontiguousarray(rparams, dtype=np.float64)
            _chk(lib().amrmg_serve_set_params(self.p, sp.ctypes.data_as(_ip), rp.ctypes.data_as(_dp)))

    def set_lanes(self, max_large, large_cells=1 << 22):
        """At most max_large requests of at least large_cells cells run at once."""
        _chk(lib().amrmg_serve_set_lanes(self.p, int(max_large), C.c_longlong(int(large_cells))))

    def submit(self, hier, rhs):
        isp, rsp = hier.spec()
        isp = np.ascontiguousarray(isp, np.int32)
        rsp = np.ascontiguousarray(rsp, np.float64)
        rhs, pr = _d(rhs)
        if rhs.size != hier.total_cells():
            raise ValueError("rhs size does not match the hierarchy")
        i = C.c_longlong()
        _chk(lib().amrmg_serve_submit(self.p, isp.ctypes.data_as(_ip), rsp.ctypes.data_as(_dp), pr, C.byref(i)))
        self._size[i.value] = rhs.size
        return i.value

    def wait(self, rid, want_phi=True):
        n = self._size.get(rid)
        if n is None:
            raise AMRMGError("serve: unknown request id")
        phi = np.zeros(n) if want_phi else None
        t = np.zeros(3)
        info = np.zeros(2, np.int32)
        _chk(lib().amrmg_serve_wait(self.p, int(rid), phi.ctypes.data_as(_dp) if want_phi else None,
                                     t.ctypes.data_as(_dp), info.ctypes.data_as(_ip)))
        self._size.pop(rid, None)
        return dict(phi=phi, t_arrival=t[0], t_admit=t[1], t_done=t[2], latency=t[2] - t[0],
                    iterations=int(info[0]), exit_status=int(info[1]))

    def stats(self):
        v = (C.c_longlong * 5)()
        _chk(lib().amrmg_serve_stats(self.p, v, 5))
        return dict(max_bytes=v[0], created=v[1], reused=v[2], evicted=v[3], max_running=v[4])

    def close(self):
        if self.p:
            lib().amrmg_serve_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
