"""ctypes front-end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
import this module.  It loads oracle/liboracle.so (built by oracle/Makefile).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

dptr = C.POINTER(C.c_double)
iptr = C.POINTER(C.c_int)


def lib(path=None):
    """The oracle library: oracle/liboracle.so, or the build at `path` (the
    first call decides, e.g. a -march=native build for the CPU baseline)."""
    global _LIB
    if _LIB is None:
        path = path or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (run make -C oracle)")
        L = C.CDLL(path)
        L.orc_tables_create.restype = C.c_void_p
        L.orc_tables_create.argtypes = [C.c_int, C.c_int]
        L.orc_tables_destroy.argtypes = [C.c_void_p]
        L.orc_tables_sizes.argtypes = [C.c_void_p, iptr]
        L.orc_tables_get.restype = C.c_int
        L.orc_tables_get.argtypes = [C.c_void_p, C.c_char_p, dptr, C.c_int]
        L.orc_hier_create.restype = C.c_void_p
        L.orc_hier_create.argtypes = [C.c_int, C.c_int, iptr, iptr, iptr, dptr, dptr, iptr]
        L.orc_hier_destroy.argtypes = [C.c_void_p]
        L.orc_hier_set_conservative_reflux.argtypes = [C.c_void_p, C.c_int]
        L.orc_hier_set_cfl_uncovered.argtypes = [C.c_void_p, C.c_int]
        L.orc_hier_nslot.restype = C.c_int
        L.orc_hier_nslot.argtypes = [C.c_void_p, C.c_int]
        L.orc_hier_nfslot.restype = C.c_int
        L.orc_hier_nfslot.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_hier_valid_mask.argtypes = [C.c_void_p, C.c_int, iptr]
        L.orc_hier_face_mask.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, iptr]
        L.orc_hier_slot_centre.argtypes = [C.c_void_p, C.c_int, C.c_int, dptr]
        L.orc_project_ic.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, dptr, C.c_double, dptr]
        for nm in ("orc_pressure",):
            getattr(L, nm).restype = C.c_double
            getattr(L, nm).argtypes = [C.c_int, dptr, C.c_double]
        L.orc_pde_flux.restype = C.c_double
        L.orc_pde_flux.argtypes = [C.c_int, C.c_int, C.c_int, dptr, C.c_double]
        L.orc_pde_dflux.restype = C.c_double
        L.orc_pde_dflux.argtypes = [C.c_int, C.c_int, dptr, C.c_double]
        L.orc_cfl_lambda.restype = C.c_double
        L.orc_cfl_lambda.argtypes = [C.c_int, C.c_int, dptr, C.c_double]
        L.orc_llf.restype = C.c_double
        L.orc_llf.argtypes = [C.c_double] * 6
        L.orc_predictor.argtypes = [C.c_void_p, C.c_void_p, C.c_int, dptr, C.c_double, C.c_double, dptr]
        L.orc_ader.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(dptr), C.c_double, C.c_double, dptr]
        L.orc_average_down.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(dptr)]
        L.orc_fill_patch.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(dptr)]
        L.orc_set_dt.restype = C.c_double
        L.orc_set_dt.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(dptr), C.c_double, C.c_double]
        L.orc_evolve.restype = C.c_int
        L.orc_evolve.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(dptr), dptr, dptr, C.c_double,
                                 C.c_double, C.c_double, C.c_int]
        L.orc_lp_error.restype = C.c_double
        L.orc_lp_error.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(dptr), C.c_int, C.c_int, C.c_int,
                                   dptr, C.c_double, C.c_double]
        L.orc_totals.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(dptr), dptr]
        L.orc_limiter_tvb.restype = C.c_long
        L.orc_limiter_tvb.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(dptr), C.c_int, C.c_double, dptr,
                                      C.c_double]
        L.orc_fill_boundary.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(dptr), C.c_int]
        L.orc_ev_matrices.argtypes = [C.c_int, C.c_int, dptr, C.c_double, dptr, dptr]
        L.orc_evolve_lim.restype = C.c_int
        L.orc_evolve_lim.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(dptr), dptr, dptr, C.c_double,
                                     C.c_double, C.c_double, C.c_int, C.c_double, dptr, C.c_int,
                                     C.POINTER(C.c_long)]
        L.orc_tag_cells.restype = C.c_long
        L.orc_tag_cells.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(dptr), C.c_int, C.c_int, C.c_double,
                                    C.c_double, dptr, C.c_double, C.POINTER(C.c_ubyte)]
        L.orc_remake.restype = C.c_int
        L.orc_remake.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(dptr), C.c_void_p, C.POINTER(dptr)]
        L.orc_l2_nodes.restype = C.c_int
        L.orc_l2_nodes.argtypes = [C.c_int, dptr]
        fptr = C.POINTER(C.c_float)
        L.orc_predictor_iters.argtypes = [C.c_void_p, C.c_void_p, C.c_int, dptr, C.c_double, C.c_double,
                                          C.c_int, dptr]
        L.orc_mlp_dims.argtypes = [C.c_void_p, iptr]
        L.orc_mlp_features.argtypes = [C.c_void_p, dptr, dptr, C.c_double, dptr]
        L.orc_mlp_target.argtypes = [C.c_void_p, dptr, dptr, dptr, dptr]
        L.orc_mlp_loss_grad.restype = C.c_double
        L.orc_mlp_loss_grad.argtypes = [C.c_int] * 5 + [fptr, fptr, fptr, C.c_int, dptr, dptr]
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(dptr)


def _i(a):
    return a.ctypes.data_as(iptr)


class Tables:
    def __init__(self, P, D):
        self.P, self.D = P, D
        self.h = lib().orc_tables_create(P, D)
        if not self.h:
            raise ValueError("unsupported P/D")
        s = np.zeros(7, dtype=np.int32)
        lib().orc_tables_sizes(self.h, _i(s))
        self.Np_s, self.Np_st, self.qMp_s, self.qMp_st, self.qMp_st_bd, self.N, self.Q = map(int, s)

    def get(self, name):
        n = lib().orc_tables_get(self.h, name.encode(), None, 0)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(n)
        lib().orc_tables_get(self.h, name.encode(), _d(out), n)
        return out

    def __del__(self):
        try:
            lib().orc_tables_destroy(self.h)
        except Exception:
            pass


class Hier:
    """Boxes per level (ratio 2, one ghost cell)."""

    def __init__(self, D, levels, dom_n, prob_lo, prob_hi, periodic=None, conservative_reflux=False,
                 cfl_uncovered=False):
        self.D = D
        self.levels = [list(map(lambda b: (tuple(b[0]), tuple(b[1])), lev)) for lev in levels]
        self.nlev = len(levels)
        npatch = np.array([len(l) for l in levels], dtype=np.int32)
        flat = []
        for lev in levels:
            for lo, hi in lev:
                flat += list(lo) + list(hi)
        boxes = np.array(flat, dtype=np.int32)
        self.dom_n = np.array(dom_n, dtype=np.int32)
        self.prob_lo = np.array(prob_lo, dtype=np.float64)
        self.prob_hi = np.array(prob_hi, dtype=np.float64)
        per = np.array(periodic if periodic is not None else [1] * D, dtype=np.int32)
        self.h = lib().orc_hier_create(D, self.nlev, _i(npatch), _i(boxes), _i(self.dom_n),
                                       _d(self.prob_lo), _d(self.prob_hi), _i(per))
        if conservative_reflux:
            lib().orc_hier_set_conservative_reflux(self.h, 1)
        if cfl_uncovered:
            lib().orc_hier_set_cfl_uncovered(self.h, 1)

    def nslot(self, l):
        return lib().orc_hier_nslot(self.h, l)

    def nfslot(self, l, d):
        return lib().orc_hier_nfslot(self.h, l, d)

    def valid_mask(self, l):
        out = np.zeros(self.nslot(l), dtype=np.int32)
        lib().orc_hier_valid_mask(self.h, l, _i(out))
        return out.astype(bool)

    def slot_centre(self, l, slot):
        x = np.zeros(3)
        lib().orc_hier_slot_centre(self.h, l, slot, _d(x))
        return x[:self.D]

    def face_mask(self, l, d, kind):
        out = np.zeros(self.nfslot(l, d), dtype=np.int32)
        lib().orc_hier_face_mask(self.h, l, d, kind, _i(out))
        return out

    def __del__(self):
        try:
            lib().orc_hier_destroy(self.h)
        except Exception:
            pass


def _ptrs(Uw):
    arr = (dptr * len(Uw))()
    for i, u in enumerate(Uw):
        assert u.flags["C_CONTIGUOUS"] and u.dtype == np.float64
        arr[i] = _d(u)
    return arr


def alloc_levels(T, H):
    return [np.zeros((T.Q, T.Np_s, H.nslot(l))) for l in range(H.nlev)]


def project_ic(T, H, ic_case, params, gamma=1.4):
    prm = np.array(params, dtype=np.float64)
    Uw = alloc_levels(T, H)
    for l in range(H.nlev):
        lib().orc_project_ic(T.h, H.h, l, ic_case, _d(prm), gamma, _d(Uw[l]))
    return Uw


def predictor(T, H, lev, Uw_l, Dt, gamma=1.4):
    out = np.zeros((T.Q, T.Np_st, H.nslot(lev)))
    lib().orc_predictor(T.h, H.h, lev, _d(np.ascontiguousarray(Uw_l)), Dt, gamma, _d(out))
    return out


def ader(T, H, Uw, Dt, gamma=1.4, want_fnum=False):
    fn = None
    if want_fnum:
        tot = sum(T.Q * T.qMp_st_bd * H.nfslot(l, d) for l in range(H.nlev) for d in range(T.D))
        fn = np.zeros(tot)
    lib().orc_ader(T.h, H.h, _ptrs(Uw), Dt, gamma, _d(fn) if fn is not None else None)
    return fn


def average_down(T, H, Uw):
    lib().orc_average_down(T.h, H.h, _ptrs(Uw))


def fill_patch(T, H, Uw):
    lib().orc_fill_patch(T.h, H.h, _ptrs(Uw))


def set_dt(T, H, Uw, c_dt, gamma=1.4):
    return lib().orc_set_dt(T.h, H.h, _ptrs(Uw), c_dt, gamma)


def evolve(T, H, Uw, t, Dt, Tend, c_dt, gamma=1.4, nsteps_max=0):
    tt = np.array([t], dtype=np.float64)
    dd = np.array([Dt], dtype=np.float64)
    n = lib().orc_evolve(T.h, H.h, _ptrs(Uw), _d(tt), _d(dd), Tend, c_dt, gamma, nsteps_max)
    return n, float(tt[0]), float(dd[0])


def lp_error(T, H, Uw, q, p, ic_case, params, t, gamma=1.4):
    prm = np.array(params, dtype=np.float64)
    return lib().orc_lp_error(T.h, H.h, _ptrs(Uw), q, p, ic_case, _d(prm), gamma, t)


def limiter_tvb(T, H, Uw, lev, M, tvb_c, gamma=1.4):
    """TVB limiter of level lev in place; returns the number of cells limited."""
    c = np.ascontiguousarray(tvb_c, dtype=np.float64)
    return lib().orc_limiter_tvb(T.h, H.h, _ptrs(Uw), lev, M, _d(c), gamma)


def fill_boundary(T, H, Uw, lev):
    lib().orc_fill_boundary(T.h, H.h, _ptrs(Uw), lev)


def ev_matrices(D, d, u, gamma=1.4):
    u = np.ascontiguousarray(u, dtype=np.float64)
    Lm = np.zeros((D + 2) * (D + 2))
    Rm = np.zeros((D + 2) * (D + 2))
    lib().orc_ev_matrices(D, d, _d(u), gamma, _d(Lm), _d(Rm))
    return Lm.reshape(D + 2, D + 2), Rm.reshape(D + 2, D + 2)


def evolve_lim(T, H, Uw, t, Dt, Tend, c_dt, M, tvb_c, t_limit=1, gamma=1.4, nsteps_max=0):
    """evolve with the TVB limiter on the finest level; returns
    (steps, t, Dt, cells limited)."""
    tt = np.array([t], dtype=np.float64)
    dd = np.array([Dt], dtype=np.float64)
    c = np.ascontiguousarray(tvb_c, dtype=np.float64)
    nl = C.c_long(0)
    n = lib().orc_evolve_lim(T.h, H.h, _ptrs(Uw), _d(tt), _d(dd), Tend, c_dt, gamma, nsteps_max, M, _d(c),
                             t_limit, C.byref(nl))
    return n, float(tt[0]), float(dd[0]), nl.value


def tag_cells(T, H, Uw, lev, crit, amr_c, tvb_c=1.0, params=(5.0, 5.0), gamma=1.4):
    """Refinement tags of level lev, one per valid cell (patch order, x fastest)."""
    n = int(H.valid_mask(lev).sum())
    out = np.zeros(n, dtype=np.uint8)
    prm = np.zeros(4)
    prm[:len(params)] = params
    lib().orc_tag_cells(T.h, H.h, _ptrs(Uw), lev, crit, amr_c, tvb_c, _d(prm), gamma,
                        out.ctypes.data_as(C.POINTER(C.c_ubyte)))
    return out


def remake(T, H_old, Uw_old, H_new):
    """U_w on the layout of H_new from the state on H_old (RemakeLevel)."""
    Un = alloc_levels(T, H_new)
    rc = lib().orc_remake(T.h, H_old.h, _ptrs(Uw_old), H_new.h, _ptrs(Un))
    if rc != 0:
        raise ValueError("new layout does not nest in the one below it")
    return Un


def l2_nodes(P):
    """The 1D node set of the L2 norm's 2(P+1)-point rule, in its order."""
    out = np.zeros(2 * (P + 1))
    n = lib().orc_l2_nodes(P, _d(out))
    return out[:n]


def totals(T, H, Uw):
    out = np.zeros(T.Q)
    lib().orc_totals(T.h, H.h, _ptrs(Uw), _d(out))
    return out


def pde_flux(D, d, q, u, gamma=1.4):
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().orc_pde_flux(D, d, q, _d(u), gamma)


def pde_dflux(D, d, u, gamma=1.4):
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().orc_pde_dflux(D, d, _d(u), gamma)


def cfl_lambda(D, d, u, gamma=1.4):
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().orc_cfl_lambda(D, d, _d(u), gamma)


def llf(uL, uR, fL, fR, dfL, dfR):
    return lib().orc_llf(uL, uR, fL, fR, dfL, dfR)


def uniform_boxes(n, D, max_grid=None):
    """Chop a periodic n^D level-0 domain into boxes of edge max_grid."""
    mg = max_grid or n
    import itertools
    out = []
    # dim 0 varies fastest in patch order
    for lo_rev in itertools.product(*[range(0, n, mg)] * D):
        lo = tuple(reversed(lo_rev))
        hi = tuple(min(l + mg, n) - 1 for l in lo)
        out.append((lo, hi))
    return out


# ---- learned warm start (no counterpart in the reference) ----
def predictor_iters(T, H, lev, Uw_l, Dt, iters, gamma=1.4):
    out = np.zeros((T.Q, T.Np_st, H.nslot(lev)))
    lib().orc_predictor_iters(T.h, H.h, lev, _d(np.ascontiguousarray(Uw_l)), Dt, gamma, iters, _d(out))
    return out


def mlp_dims(T):
    o = np.zeros(5, dtype=np.int32)
    lib().orc_mlp_dims(T.h, _i(o))
    return [int(v) for v in o]  # nin, nin_p, nh, nout, nout_p


def mlp_features(T, U, tau, gamma=1.4):
    """U: [Q][Np_s] of one element, tau: [D] = Dt/dx_d."""
    nin = mlp_dims(T)[0]
    x = np.zeros(nin)
    lib().orc_mlp_features(T.h, _d(np.ascontiguousarray(U, dtype=np.float64)),
                           _d(np.ascontiguousarray(tau, dtype=np.float64)), gamma, _d(x))
    return x


def mlp_target(T, U, Hst, tau):
    nout = mlp_dims(T)[3]
    y = np.zeros(nout)
    lib().orc_mlp_target(T.h, _d(np.ascontiguousarray(U, dtype=np.float64)),
                         _d(np.ascontiguousarray(Hst, dtype=np.float64)),
                         _d(np.ascontiguousarray(tau, dtype=np.float64)), _d(y))
    return y


def mlp_loss_grad(dims, params, x, y, bf16=True, want_grad=True):
    """dims = (nin_p, nh, nout_p, nout); params flat float32; x [n, nin_p],
    y [n, nout_p] float32.  Returns (loss, grad float64 or None, ypred)."""
    nin_p, nh, nout_p, nout = dims
    params = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    n = x.shape[0]
    grad = np.zeros(params.size) if want_grad else None
    yp = np.zeros((n, nout_p))
    fptr = C.POINTER(C.c_float)
    loss = lib().orc_mlp_loss_grad(nin_p, nh, nout_p, nout, n, params.ctypes.data_as(fptr),
                                   x.ctypes.data_as(fptr), y.ctypes.data_as(fptr), 1 if bf16 else 0,
                                   _d(grad) if want_grad else None, _d(yp))
    return loss, grad, yp
