"""ctypes binding of libadg.so (the C-ABI in include/adg.h).

Loads the in-tree library built by ader-dg-amr_amd/Makefile and fails loudly
when it is missing: there is no CPU fallback for the product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libadg.so")

ADG_FP64 = 0
ADG_FP32_MFMA = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_LIB = None

# every entry point declared in include/adg.h
EXPORTS = [
    "adg_create", "adg_set_geometry", "adg_set_levels", "adg_level_nslot", "adg_level_nface", "adg_level_npred",
    "adg_upload_Uw", "adg_download_Uw", "adg_fill_ghosts", "adg_average_down", "adg_set_dt",
    "adg_set_dt_value", "adg_predictor", "adg_keep_Hw", "adg_download_Hw", "adg_face_flux",
    "adg_download_Fnum", "adg_corrector", "adg_reflux", "adg_ader", "adg_advance", "adg_step", "adg_step_graph", "adg_steps", "adg_set_graph_unroll", "adg_set_time", "adg_get_time",
    "adg_conserved_totals", "adg_sync", "adg_stream", "adg_sizes", "adg_get_table",
    "adg_project_ic", "adg_set_predictor_kernel", "adg_set_predictor_waves", "adg_set_warm_tiles", "adg_last_error", "adg_destroy",
    "adg_mlp_sizes", "adg_mlp_init", "adg_mlp_get_params", "adg_mlp_set_params", "adg_mlp_collect",
    "adg_mlp_dataset", "adg_mlp_dataset_clear", "adg_mlp_fit_normalization", "adg_mlp_train",
    "adg_mlp_loss_grad", "adg_mlp_warm_start", "adg_set_learned", "adg_learned_stats", "adg_lp_error",
    "adg_autotune", "adg_project_ic_patches", "adg_set_independent_patches",
    "adg_conserved_totals_noalloc", "adg_set_limiter", "adg_limit", "adg_limiter_stats",
    "adg_set_amr", "adg_set_tagging", "adg_tag_cells", "adg_regrid", "adg_remake", "adg_get_levels",
    "adg_set_step_count", "adg_set_conservative_reflux", "adg_set_cfl_uncovered", "adg_regrid_stats",
]


class AdgError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise AdgError(f"{LIBPATH} is not built; run __graft_entry__.build() or make -C ader-dg-amr_amd")
    L = C.CDLL(LIBPATH)
    L.adg_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_double, C.c_int, C.c_void_p]
    L.adg_set_geometry.argtypes = [C.c_void_p, _ip, _dp, _dp, _ip]
    L.adg_set_levels.argtypes = [C.c_void_p, C.c_int, _ip, _ip]
    L.adg_level_nslot.argtypes = [C.c_void_p, C.c_int]
    L.adg_level_nface.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.adg_level_npred.argtypes = [C.c_void_p, C.c_int]
    L.adg_upload_Uw.argtypes = [C.c_void_p, C.c_int, _dp]
    L.adg_download_Uw.argtypes = [C.c_void_p, C.c_int, _dp]
    L.adg_fill_ghosts.argtypes = [C.c_void_p]
    L.adg_average_down.argtypes = [C.c_void_p]
    L.adg_set_dt.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, _dp]
    L.adg_set_dt_value.argtypes = [C.c_void_p, C.c_double]
    L.adg_predictor.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip]
    L.adg_keep_Hw.argtypes = [C.c_void_p, C.c_int]
    L.adg_download_Hw.argtypes = [C.c_void_p, C.c_int, _dp]
    L.adg_face_flux.argtypes = [C.c_void_p, C.c_int]
    L.adg_download_Fnum.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
    L.adg_corrector.argtypes = [C.c_void_p, C.c_int]
    L.adg_reflux.argtypes = [C.c_void_p, C.c_int]
    L.adg_ader.argtypes = [C.c_void_p]
    L.adg_advance.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp]
    L.adg_step.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp]
    L.adg_step_graph.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, _dp]
    L.adg_steps.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, _dp]
    L.adg_set_graph_unroll.argtypes = [C.c_void_p, C.c_int]
    L.adg_set_time.argtypes = [C.c_void_p, C.c_double]
    L.adg_get_time.argtypes = [C.c_void_p, _dp, _dp]
    L.adg_conserved_totals.argtypes = [C.c_void_p, _dp]
    L.adg_conserved_totals_noalloc.argtypes = [C.c_void_p, C.c_void_p]
    L.adg_set_limiter.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, _dp, C.c_int]
    L.adg_limit.argtypes = [C.c_void_p, C.POINTER(C.c_ulonglong)]
    L.adg_limiter_stats.argtypes = [C.c_void_p, C.POINTER(C.c_ulonglong), C.c_int]
    L.adg_set_step_count.argtypes = [C.c_void_p, C.c_int]
    L.adg_set_conservative_reflux.argtypes = [C.c_void_p, C.c_int]
    L.adg_set_cfl_uncovered.argtypes = [C.c_void_p, C.c_int]
    L.adg_regrid_stats.argtypes = [C.c_void_p, _dp, C.c_int]
    L.adg_set_amr.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int]
    L.adg_set_tagging.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, C.c_int, _dp]
    L.adg_tag_cells.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_ubyte)]
    L.adg_regrid.argtypes = [C.c_void_p, C.c_double, _ip]
    L.adg_remake.argtypes = [C.c_void_p, C.c_int, _ip, _ip]
    L.adg_get_levels.argtypes = [C.c_void_p, _ip, _ip, _ip, C.c_int]
    L.adg_host_make_grids.restype = C.c_int
    L.adg_host_make_grids.argtypes = [C.c_int, _ip, _ip, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                      C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _ip, C.c_int]
    L.adg_sync.argtypes = [C.c_void_p]
    L.adg_stream.argtypes = [C.c_void_p]
    L.adg_stream.restype = C.c_void_p
    L.adg_sizes.argtypes = [C.c_void_p, _ip]
    L.adg_get_table.argtypes = [C.c_void_p, C.c_char_p, _dp, C.c_int]
    L.adg_project_ic.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
    L.adg_set_predictor_kernel.argtypes = [C.c_void_p, C.c_int]
    L.adg_set_predictor_waves.argtypes = [C.c_void_p, C.c_int]
    L.adg_set_warm_tiles.argtypes = [C.c_void_p, C.c_int]
    _fp = C.POINTER(C.c_float)
    _lp = C.POINTER(C.c_long)
    _up = C.POINTER(C.c_ulonglong)
    L.adg_mlp_sizes.argtypes = [C.c_void_p, _ip]
    L.adg_mlp_init.argtypes = [C.c_void_p, C.c_ulonglong]
    L.adg_mlp_get_params.argtypes = [C.c_void_p, _fp, C.c_long]
    L.adg_mlp_set_params.argtypes = [C.c_void_p, _fp, C.c_long]
    L.adg_mlp_collect.argtypes = [C.c_void_p, C.c_int, _lp]
    L.adg_mlp_dataset.argtypes = [C.c_void_p, C.c_long, C.c_long, _fp, _fp]
    L.adg_mlp_dataset_clear.argtypes = [C.c_void_p]
    L.adg_mlp_fit_normalization.argtypes = [C.c_void_p]
    L.adg_mlp_train.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_ulonglong, _dp]
    L.adg_mlp_loss_grad.argtypes = [C.c_void_p, C.c_int, _fp, _fp, _dp, _fp]
    L.adg_mlp_warm_start.argtypes = [C.c_void_p, C.c_int, _dp]
    L.adg_set_learned.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int]
    L.adg_learned_stats.argtypes = [C.c_void_p, _up, C.c_int]
    L.adg_lp_error.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, C.c_double, _dp]
    L.adg_autotune.argtypes = [C.c_void_p, C.c_int, _dp, _ip]
    L.adg_project_ic_patches.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp]
    L.adg_set_independent_patches.argtypes = [C.c_void_p, C.c_int]
    L.adg_last_error.argtypes = [C.c_void_p]
    L.adg_last_error.restype = C.c_char_p
    L.adg_destroy.argtypes = [C.c_void_p]
    for nm in EXPORTS:
        f = getattr(L, nm)
        if nm not in ("adg_stream", "adg_last_error", "adg_destroy"):
            f.restype = C.c_int
    _LIB = L
    return L


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _flatten_levels(levels):
    npatch = np.array([len(l) for l in levels], dtype=np.int32)
    flat = []
    for lev in levels:
        for lo, hi in lev:
            flat += list(lo) + list(hi)
    return npatch, np.array(flat, dtype=np.int32)


def make_grids(dim, dom0, levels, tags, max_level, periodic=None, blocking_factor=4, max_grid_size=None,
               grid_eff=0.7, n_error_buf=1, n_proper=1):
    """New layout from tags on the host (adg_host_make_grids): levels is the
    old layout, tags[l] an array [n][dim] of tagged cells of level l."""
    L = lib()
    per = np.array(periodic if periodic is not None else [1] * dim, dtype=np.int32)
    npatch, boxes = _flatten_levels(levels)
    nt = np.array([len(t) for t in tags], dtype=np.int32)
    tg = np.concatenate([np.asarray(t, dtype=np.int32).reshape(-1, dim) for t in tags]) if len(tags) else \
        np.zeros((0, dim), dtype=np.int32)
    tg = np.ascontiguousarray(tg, dtype=np.int32)
    if tg.size == 0:
        tg = np.zeros(dim, dtype=np.int32)
    mgs = max_grid_size or (128 if dim == 2 else 32)
    nl = C.c_int(0)
    npn = np.zeros(16, dtype=np.int32)
    args = (dim, _i(np.array(dom0, dtype=np.int32)), _i(per), max_level, blocking_factor, mgs, grid_eff,
            n_error_buf, n_proper, len(levels), _i(npatch), _i(boxes), _i(nt), _i(tg), C.byref(nl), _i(npn))
    n = L.adg_host_make_grids(*args, None, 0)
    if n < 0:
        raise AdgError(f"adg_host_make_grids failed rc={n}")
    out_b = np.zeros(max(n, 1) * 2 * dim, dtype=np.int32)
    L.adg_host_make_grids(*args, _i(out_b), n)
    out, k = [], 0
    for l in range(nl.value):
        lev = []
        for _ in range(npn[l]):
            b = out_b[k * 2 * dim:(k + 1) * 2 * dim]
            lev.append((tuple(int(v) for v in b[:dim]), tuple(int(v) for v in b[dim:])))
            k += 1
        out.append(lev)
    return out


class Context:
    """One adg_ctx: element tables, device layout of a hierarchy, fields."""

    def __init__(self, dim, p, gamma=1.4, precision=ADG_FP64):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.adg_create(C.byref(h), dim, p, gamma, precision, None)
        if rc != 0:
            raise AdgError(f"adg_create failed rc={rc}")
        self.h = h
        self.dim, self.p = dim, p
        s = np.zeros(7, dtype=np.int32)
        self.L.adg_sizes(self.h, _i(s))
        self.Np_s, self.Np_st, self.qMp_s, self.qMp_st, self.qMp_st_bd, self.N, self.Q = map(int, s)
        self.nlev = 0

    def _chk(self, rc, what):
        if rc != 0:
            raise AdgError(f"{what} failed rc={rc}: {self.L.adg_last_error(self.h).decode()}")

    def set_geometry(self, dom_n, prob_lo, prob_hi, periodic=None):
        per = np.array(periodic if periodic is not None else [1] * self.dim, dtype=np.int32)
        self._chk(self.L.adg_set_geometry(self.h, _i(np.array(dom_n, dtype=np.int32)),
                                          _d(np.array(prob_lo, dtype=np.float64)),
                                          _d(np.array(prob_hi, dtype=np.float64)), _i(per)),
                  "adg_set_geometry")

    def set_levels(self, levels):
        npatch = np.array([len(l) for l in levels], dtype=np.int32)
        flat = []
        for lev in levels:
            for lo, hi in lev:
                flat += list(lo) + list(hi)
        self._chk(self.L.adg_set_levels(self.h, len(levels), _i(npatch), _i(np.array(flat, dtype=np.int32))),
                  "adg_set_levels")
        self.nlev = len(levels)

    def nslot(self, lev):
        return self.L.adg_level_nslot(self.h, lev)

    def nface(self, lev, d):
        return self.L.adg_level_nface(self.h, lev, d)

    def npred(self, lev):
        """Slots of level lev the predictor computes (valid cells and non-copy ghosts)."""
        return self.L.adg_level_npred(self.h, lev)

    def upload_Uw(self, lev, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (self.Q, self.Np_s, self.nslot(lev))
        self._chk(self.L.adg_upload_Uw(self.h, lev, _d(U)), "adg_upload_Uw")

    def download_Uw(self, lev):
        out = np.zeros((self.Q, self.Np_s, self.nslot(lev)))
        self._chk(self.L.adg_download_Uw(self.h, lev, _d(out)), "adg_download_Uw")
        return out

    def set_predictor_kernel(self, kind):
        """0 automatic, 1 VALU (FP64), 2 MFMA quad-gather, 3 MFMA node-lane (2D; 3D p <= 2)."""
        self._chk(self.L.adg_set_predictor_kernel(self.h, kind), "adg_set_predictor_kernel")

    def set_predictor_waves(self, waves):
        """Wavefronts per workgroup of the node-lane predictor (8 or 16)."""
        self._chk(self.L.adg_set_predictor_waves(self.h, waves), "adg_set_predictor_waves")

    def set_warm_tiles(self, tiles):
        """Warm-start inference tiles per wavefront: 1, 2 (16 elements each) or 0 (8 elements)."""
        self._chk(self.L.adg_set_warm_tiles(self.h, tiles), "adg_set_warm_tiles")

    def keep_Hw(self, on=True):
        self._chk(self.L.adg_keep_Hw(self.h, 1 if on else 0), "adg_keep_Hw")

    def download_Hw(self, lev):
        out = np.zeros((self.Q, self.Np_st, self.nslot(lev)))
        self._chk(self.L.adg_download_Hw(self.h, lev, _d(out)), "adg_download_Hw")
        return out

    def download_Fnum(self, lev, d):
        out = np.zeros((self.Q, self.qMp_st_bd, self.nface(lev, d)))
        self._chk(self.L.adg_download_Fnum(self.h, lev, d, _d(out)), "adg_download_Fnum")
        return out

    def fill_ghosts(self):
        self._chk(self.L.adg_fill_ghosts(self.h), "adg_fill_ghosts")

    def average_down(self):
        self._chk(self.L.adg_average_down(self.h), "adg_average_down")

    def set_dt(self, c_dt, T, t):
        out = np.zeros(1)
        self._chk(self.L.adg_set_dt(self.h, c_dt, T, t, _d(out)), "adg_set_dt")
        return float(out[0])

    def set_dt_value(self, Dt):
        self._chk(self.L.adg_set_dt_value(self.h, Dt), "adg_set_dt_value")

    def set_time(self, t):
        self._chk(self.L.adg_set_time(self.h, t), "adg_set_time")

    def get_time(self):
        """(t, Dt) held on the device."""
        out = np.zeros(2)
        self._chk(self.L.adg_get_time(self.h, _d(out[0:1]), _d(out[1:2])), "adg_get_time")
        return float(out[0]), float(out[1])

    def predictor(self, lev=-1, mode=0):
        nf = C.c_int(0)
        self._chk(self.L.adg_predictor(self.h, lev, mode, C.byref(nf)), "adg_predictor")
        return nf.value

    def face_flux(self, lev=-1):
        self._chk(self.L.adg_face_flux(self.h, lev), "adg_face_flux")

    def corrector(self, lev=-1):
        self._chk(self.L.adg_corrector(self.h, lev), "adg_corrector")

    def reflux(self, lev=-1):
        self._chk(self.L.adg_reflux(self.h, lev), "adg_reflux")

    def ader(self):
        self._chk(self.L.adg_ader(self.h), "adg_ader")

    def advance(self, c_dt, T):
        self._chk(self.L.adg_advance(self.h, c_dt, T, None), "adg_advance")

    def step(self, c_dt, T, want_t=False):
        if want_t:
            out = np.zeros(1)
            self._chk(self.L.adg_step(self.h, c_dt, T, _d(out)), "adg_step")
            return float(out[0])
        self._chk(self.L.adg_step(self.h, c_dt, T, None), "adg_step")
        return None

    def step_graph(self, c_dt, T, nsteps):
        self._chk(self.L.adg_step_graph(self.h, c_dt, T, nsteps, None), "adg_step_graph")

    def set_graph_unroll(self, steps):
        """Steps captured per step graph (1..64)."""
        self._chk(self.L.adg_set_graph_unroll(self.h, steps), "adg_set_graph_unroll")

    def steps(self, c_dt, T, nsteps):
        """nsteps steps as stream launches, no host synchronisation between them."""
        self._chk(self.L.adg_steps(self.h, c_dt, T, nsteps, None), "adg_steps")

    def totals(self):
        out = np.zeros(self.Q)
        self._chk(self.L.adg_conserved_totals(self.h, _d(out)), "adg_conserved_totals")
        return out

    def set_limiter(self, on=True, M=0.0, tvb_c=(1.0,), t_limit=1):
        """TVB limiter on the finest level in every step (on) or none."""
        c = np.ascontiguousarray(tvb_c, dtype=np.float64)
        self._chk(self.L.adg_set_limiter(self.h, 1 if on else 0, M, c.size, _d(c), t_limit), "adg_set_limiter")

    def set_amr(self, max_level, blocking_factor=4, max_grid_size=None, grid_eff=0.7, n_error_buf=1, n_proper=1):
        mgs = max_grid_size or (128 if self.dim == 2 else 32)
        self._chk(self.L.adg_set_amr(self.h, max_level, blocking_factor, mgs, grid_eff, n_error_buf, n_proper),
                  "adg_set_amr")

    def set_tagging(self, criterion, amr_c, params=()):
        c = np.ascontiguousarray(amr_c, dtype=np.float64)
        prm = np.ascontiguousarray(params, dtype=np.float64) if len(params) else np.zeros(1)
        self._chk(self.L.adg_set_tagging(self.h, criterion, c.size, _d(c), len(params), _d(prm)), "adg_set_tagging")

    def tag_cells(self, lev):
        n = self.nvalid(lev)
        out = np.zeros(n, dtype=np.uint8)
        self._chk(self.L.adg_tag_cells(self.h, lev, out.ctypes.data_as(C.POINTER(C.c_ubyte))), "adg_tag_cells")
        return out

    def regrid(self, time=0.0):
        ch = C.c_int(0)
        self._chk(self.L.adg_regrid(self.h, time, C.byref(ch)), "adg_regrid")
        return bool(ch.value)

    def remake(self, levels):
        npatch, boxes = _flatten_levels(levels)
        self._chk(self.L.adg_remake(self.h, len(levels), _i(npatch), _i(boxes)), "adg_remake")
        self.levels = [list(l) for l in levels]

    def get_levels(self):
        nl = C.c_int(0)
        npatch = np.zeros(16, dtype=np.int32)
        n = self.L.adg_get_levels(self.h, C.byref(nl), _i(npatch), None, 0)
        if n < 0:
            raise AdgError(f"adg_get_levels failed rc={n}: {self.L.adg_last_error(self.h).decode()}")
        boxes = np.zeros(max(n, 1) * 2 * self.dim, dtype=np.int32)
        self.L.adg_get_levels(self.h, C.byref(nl), _i(npatch), _i(boxes), n)
        out, k = [], 0
        for l in range(nl.value):
            lev = []
            for _ in range(npatch[l]):
                b = boxes[k * 2 * self.dim:(k + 1) * 2 * self.dim]
                lev.append((tuple(int(v) for v in b[:self.dim]), tuple(int(v) for v in b[self.dim:])))
                k += 1
            out.append(lev)
        self.levels = out
        return out

    def nvalid(self, lev):
        return sum(int(np.prod([hi[d] - lo[d] + 1 for d in range(self.dim)])) for lo, hi in self.get_levels()[lev])

    def set_conservative_reflux(self, on=True):
        self._chk(self.L.adg_set_conservative_reflux(self.h, 1 if on else 0), "adg_set_conservative_reflux")

    def set_cfl_uncovered(self, on=True):
        self._chk(self.L.adg_set_cfl_uncovered(self.h, 1 if on else 0), "adg_set_cfl_uncovered")

    def regrid_stats(self, reset=False):
        """{"changes", "tags", "grids", "hierarchy", "setup", "transfer"}: layout
        changes and host seconds per regrid phase since the last reset."""
        o = np.zeros(6)
        self._chk(self.L.adg_regrid_stats(self.h, _d(o), 1 if reset else 0), "adg_regrid_stats")
        return dict(zip(("changes", "tags", "grids", "hierarchy", "setup", "transfer"), o.tolist()))

    def limit(self):
        """Limiter_w on the finest level now; returns the cells limited."""
        n = C.c_ulonglong(0)
        self._chk(self.L.adg_limit(self.h, C.byref(n)), "adg_limit")
        return n.value

    def limiter_stats(self, reset=False):
        n = C.c_ulonglong(0)
        self._chk(self.L.adg_limiter_stats(self.h, C.byref(n), 1 if reset else 0), "adg_limiter_stats")
        return n.value

    def totals_noalloc(self, d_out):
        """Conserved totals into device memory at address d_out (Q doubles)."""
        self._chk(self.L.adg_conserved_totals_noalloc(self.h, C.c_void_p(d_out)), "adg_conserved_totals_noalloc")

    def project_ic(self, lev, ic_case, params):
        prm = np.array(params, dtype=np.float64)
        self._chk(self.L.adg_project_ic(self.h, lev, ic_case, _d(prm)), "adg_project_ic")

    def lp_error(self, q, p, ic_case, params, t):
        """Composite Lp error of variable q against the exact solution at t."""
        prm = np.zeros(8)
        prm[:len(params)] = params
        out = np.zeros(1)
        self._chk(self.L.adg_lp_error(self.h, q, p, ic_case, _d(prm), t, _d(out)), "adg_lp_error")
        return float(out[0])

    def autotune(self, reps=10):
        """Times the predictor launch shapes and keeps the fastest; returns
        ({shape: ms}, (kind, waves))."""
        ms = np.zeros(7)
        ch = np.zeros(3, dtype=np.int32)
        self._chk(self.L.adg_autotune(self.h, reps, _d(ms), _i(ch)), "adg_autotune")
        names = ("valu", "quad8", "nodelane8", "nodelane16", "warm1", "warm2", "warm_half")
        return {n: float(v) for n, v in zip(names, ms) if v >= 0}, (int(ch[0]), int(ch[1]))

    def project_ic_patches(self, lev, ic_case, params):
        """params: [npatch][nprm] parameter sets, one per patch of the level."""
        prm = np.ascontiguousarray(params, dtype=np.float64)
        self._chk(self.L.adg_project_ic_patches(self.h, lev, ic_case, prm.shape[1], _d(prm)), "adg_project_ic_patches")

    def set_independent_patches(self, on=True):
        self._chk(self.L.adg_set_independent_patches(self.h, 1 if on else 0), "adg_set_independent_patches")

    def sync(self):
        self._chk(self.L.adg_sync(self.h), "adg_sync")

    def stream(self):
        return self.L.adg_stream(self.h)

    def table(self, name):
        n = self.L.adg_get_table(self.h, name.encode(), None, 0)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(n)
        self.L.adg_get_table(self.h, name.encode(), _d(out), n)
        return out

    def close(self):
        if getattr(self, "h", None):
            self.L.adg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- learned warm start -------------------------------------------
    def mlp_sizes(self):
        o = np.zeros(7, dtype=np.int32)
        self._chk(self.L.adg_mlp_sizes(self.h, _i(o)), "adg_mlp_sizes")
        return dict(zip(("nin", "nin_p", "nh", "nout", "nout_p", "npar", "nnorm"), map(int, o)))

    def mlp_init(self, seed=1):
        self._chk(self.L.adg_mlp_init(self.h, seed), "adg_mlp_init")

    def mlp_params(self):
        sz = self.mlp_sizes()
        out = np.zeros(sz["npar"] + sz["nnorm"], dtype=np.float32)
        self._chk(self.L.adg_mlp_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float)), out.size),
                  "adg_mlp_get_params")
        return out

    def mlp_set_params(self, par):
        par = np.ascontiguousarray(par, dtype=np.float32)
        self._chk(self.L.adg_mlp_set_params(self.h, par.ctypes.data_as(C.POINTER(C.c_float)), par.size),
                  "adg_mlp_set_params")

    def mlp_collect(self, iters):
        n = C.c_long(0)
        self._chk(self.L.adg_mlp_collect(self.h, iters, C.byref(n)), "adg_mlp_collect")
        return n.value

    def mlp_dataset(self, start, n):
        sz = self.mlp_sizes()
        x = np.zeros((n, sz["nin_p"]), dtype=np.float32)
        y = np.zeros((n, sz["nout_p"]), dtype=np.float32)
        fp = C.POINTER(C.c_float)
        self._chk(self.L.adg_mlp_dataset(self.h, start, n, x.ctypes.data_as(fp), y.ctypes.data_as(fp)),
                  "adg_mlp_dataset")
        return x, y

    def mlp_dataset_clear(self):
        self._chk(self.L.adg_mlp_dataset_clear(self.h), "adg_mlp_dataset_clear")

    def mlp_fit_normalization(self):
        self._chk(self.L.adg_mlp_fit_normalization(self.h), "adg_mlp_fit_normalization")

    def mlp_train(self, nsteps, batch=1024, lr=1e-3, seed=7):
        loss = C.c_double(0)
        self._chk(self.L.adg_mlp_train(self.h, nsteps, batch, lr, seed, C.byref(loss)), "adg_mlp_train")
        return loss.value

    def mlp_loss_grad(self, x, y):
        sz = self.mlp_sizes()
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(sz["npar"], dtype=np.float32)
        loss = C.c_double(0)
        fp = C.POINTER(C.c_float)
        self._chk(self.L.adg_mlp_loss_grad(self.h, x.shape[0], x.ctypes.data_as(fp), y.ctypes.data_as(fp),
                                           C.byref(loss), g.ctypes.data_as(fp)), "adg_mlp_loss_grad")
        return loss.value, g

    def mlp_warm_start(self, lev):
        sz = self.mlp_sizes()
        out = np.zeros((sz["nout"], self.nslot(lev)))
        self._chk(self.L.adg_mlp_warm_start(self.h, lev, _d(out)), "adg_mlp_warm_start")
        return out

    def set_learned(self, on=True, tol=1e-4, max_iters=2):
        """on: False/0 plain, True/1 learned start with early exit and
        fallback, 2 diagnostic (exactly max_iters iterations, no check)."""
        self._chk(self.L.adg_set_learned(self.h, int(on), tol, max_iters), "adg_set_learned")

    def learned_stats(self, reset=False):
        o = np.zeros(2, dtype=np.uint64)
        self._chk(self.L.adg_learned_stats(self.h, o.ctypes.data_as(C.POINTER(C.c_ulonglong)), 1 if reset else 0),
                  "adg_learned_stats")
        return int(o[0]), int(o[1])
