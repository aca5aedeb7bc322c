/*
 * adg.h -- C-ABI of the MI355X ADER-DG hot path (libadg.so).
 *
 * The reference (TrustiestMass28/ADER-DG-AMR) has no runtime ABI: its
 * boundary is the compile-time CRTP of Solver<AmrDG<P>> and
 * ModelEquation<Compressible_Euler>.  These entry points sit under that
 * interface: the CRTP methods stay on the host and their MFIter loop bodies
 * become calls below.  Each entry point names the reference member it
 * replaces (paths relative to the reference root, prefix Source/).
 *
 * Conventions
 *  - Every call returns ADG_OK (0) or a negative ADG_E_* code; nothing
 *    throws across the boundary.  adg_last_error() describes the failure.
 *  - The context owns all device memory; the caller owns box metadata.
 *  - One host thread per context.  Calls are ordered on the context's HIP
 *    stream and asynchronous, except the download / *_out readbacks and
 *    adg_sync(), which block.
 *  - Modal data layout per level: [q][n][slot] FP64, slots enumerate the
 *    grown (one ghost cell) box of every patch of the level in the order
 *    given to adg_set_levels, x fastest.
 */
#ifndef ADG_H
#define ADG_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct adg_ctx adg_ctx;

enum {
    ADG_OK = 0,
    ADG_E_ARG = -1,     /* bad argument / shape */
    ADG_E_HIP = -2,     /* HIP runtime error */
    ADG_E_STATE = -3,   /* call order violated (e.g. no levels set) */
    ADG_E_NOMEM = -4,   /* device allocation failed */
    ADG_E_UNSUP = -5    /* (dim, p) combination not compiled */
};

enum { ADG_FP64 = 0, ADG_FP32_MFMA = 1 };

/* Replaces AmrDG<P>::init (NumericalMethod/AmrDG/AmrDG.tpp:125-246: basis,
 * quadrature, element and projection matrices) together with
 * Compressible_Euler::settings (Model/CompressibleEuler/Compressible_Euler.cpp:5-48).
 * dim in {2,3}; p in {1,2,3}; precision ADG_FP64 or ADG_FP32_MFMA;
 * hip_stream may be NULL (the context creates its own stream). */
int adg_create(adg_ctx** out, int dim, int p, double gamma, int precision, void* hip_stream);

/* Domain of level 0 (AmrCore geometry, Core/Mesh.h:128-147): cells per
 * dimension, physical extent and periodicity flags. */
int adg_set_geometry(adg_ctx* ctx, const int* dom_n, const double* prob_lo,
                     const double* prob_hi, const int* periodic);

/* Replaces AmrDG::set_init_data_system / set_init_data_component
 * (AmrDG.tpp:279-334) and AmrDG::AMR_set_flux_registers
 * (AmrDG_Level.tpp:240-414): allocates the device layout of every level,
 * builds ghost-source lists, face lists, coarse-fine face masks and the
 * flux-register algebra.  boxes holds lo[dim], hi[dim] per patch, level
 * index space, ratio 2 between levels, fine boxes aligned to the ratio. */
int adg_set_levels(adg_ctx* ctx, int nlev, const int* npatch_per_lev, const int* boxes);

/* Number of slots of a level (valid + ghost cells of all its patches). */
int adg_level_nslot(adg_ctx* ctx, int lev);
/* Number of faces of direction d of a level (valid faces of all patches). */
int adg_level_nface(adg_ctx* ctx, int lev, int d);
/* Slots of level lev that the predictor computes: the valid cells and the
 * ghosts that are not plain copies (coarse-fine projections, outflow
 * mirrors).  A copy ghost's predictor result equals its source cell's bit
 * for bit (the predictor is element-local), so the face flux reads the
 * source's traces and the copy is skipped; the reference predicts every
 * ghost (AmrDG.tpp:776).  The learned start's element counts
 * (adg_learned_stats) count these slots. */
int adg_level_npred(adg_ctx* ctx, int lev);

/* Host <-> device copies of U_w for one level, layout [q][n][slot]. */
int adg_upload_Uw(adg_ctx* ctx, int lev, const double* host);
int adg_download_Uw(adg_ctx* ctx, int lev, double* host);

/* Replaces AmrDG::set_initial_condition (AmrDG.h:782-836) with the
 * pointwise Compressible_Euler::pde_IC (Compressible_Euler.h:141-368): L2
 * projection of an initial state onto the modes of every slot of a level
 * (ghost cells included), evaluated on the host and uploaded.
 *   ic_case 0: isentropic vortex, params {beta, xc, yc, u_inf, v_inf}
 *   ic_case 1: Kelvin-Helmholtz shear layer, params {amp, noise_amp, seed}
 *              (v = amp sin(4 pi x) + noise_amp * U[-1,1] per cell)
 *   ic_case 2: uniform state, params {rho, u, v, [w,] p}
 *   ic_case 3: spherical/radial shock tube, params {r0, rho_in, p_in,
 *              rho_out, p_out} centred in the domain
 *   ic_case 4: Kelvin-Helmholtz shear layer with the seeded multi-mode
 *              perturbation v = sum_{k=1..8} a_k sin(2 pi k x),
 *              a_k = amax * U[0,1), params {seed, amax}
 *   ic_case 5: smooth periodic density wave rho = 1 + amp prod_d
 *              sin(2 pi x_d / L_d), velocity (1, 0.5, 0.25), pressure p0,
 *              params {amp, p0}
 *   ic_case 6: the reference's kelvin_helmolz_instability state
 *              (Compressible_Euler.h:237-256), v = w0 sin(4 pi x) times
 *              Gaussians of width sigma about y = 0.25, 0.75, params
 *              {w0, sigma} (reference: 0.2, 0.05/sqrt(2))
 *   ic_case 7: Kelvin-Helmholtz as 1 with the noise constant on the cells
 *              of a fixed N x N grid of the unit square, params {amp,
 *              noise_amp, seed, N}: the same state at every resolution */
int adg_project_ic(adg_ctx* ctx, int lev, int ic_case, const double* params);
/* As adg_project_ic with one parameter set of nprm values per patch of the
 * level (params [npatch][nprm]); for independent patches the coordinates are
 * local to each patch, whose domain is [prob_lo, prob_lo + n dx). */
int adg_project_ic_patches(adg_ctx* ctx, int lev, int ic_case, int nprm, const double* params);
/* on != 0 before adg_set_levels: the single level's patches are independent
 * periodic domains (a batch of separate problems sharing one time step);
 * ghost cells wrap inside their own patch.  No counterpart in the reference:
 * the batched form of the hot path (every patch in one launch per stage). */
int adg_set_independent_patches(adg_ctx* ctx, int on);

/* Replaces AmrDG::AMR_FillPatch on every level (AmrDG_Level.tpp:172-224)
 * and Solver::FillBoundaryCells for periodic domains: same-level and
 * periodic copies plus L2ProjInterp::interp / amr_scatter into coarse-fine
 * ghosts (AmrDG_Interpolater.tpp:251-281,344-364). */
int adg_fill_ghosts(adg_ctx* ctx);

/* Reflux rule.  0 (default, the reference's): L2ProjInterp::reflux
 * (AmrDG_Interpolater.tpp:39-69) corrects a coarse cell only where one of
 * its faces is flagged in its own box's ownership-filtered cf_face_b_coarse
 * (AmrDG_Level.tpp:330-343), so a cell whose coarse-fine face is the low
 * face of its box is not corrected and the composite mass changes there.
 * 1: every uncovered coarse cell the flux registers reach is corrected
 * (conservative).  No counterpart in the reference (it has only rule 0).
 * With levels set, the register lists are rebuilt on the same layout. */
int adg_set_conservative_reflux(adg_ctx* ctx, int on);

/* Cells of the CFL reduction.  0 (default, the reference's set_Dt,
 * AmrDG.h:1253-1357): every valid cell of every level, covered coarse cells
 * included.  1: the uncovered cells only.  A covered coarse cell holds the
 * L2 projection of its children, which across a discontinuity can have
 * negative density at spatial node 0 where the wave speed is taken, so
 * rule 0 can collapse Dt while the composite solution is physical (3D
 * shock tubes on three levels); covered cells are overwritten by the
 * average down every step and take no part in the composite update (with
 * the conservative reflux rule).  No counterpart in the reference. */
int adg_set_cfl_uncovered(adg_ctx* ctx, int on);

/* Host time of the regrid path, cumulative since the last reset: out[0] the
 * number of layout changes (adg_remake calls that succeeded), out[1..5]
 * seconds in refinement tags (device tagging and copy back), grid
 * generation, hierarchy lists, device set-up of the new layout (uploads,
 * allocations, kernel preparation) and the transfer of U_w onto it.  The
 * reference times none of this.  reset != 0 clears the counters after
 * reading them. */
int adg_regrid_stats(adg_ctx* ctx, double* out, int reset);

/* Replaces AmrDG::AMR_average_fine_coarse (AmrDG_Level.tpp:227-238) ->
 * L2ProjInterp::average_down / amr_gather (AmrDG_Interpolater.tpp:283-398). */
int adg_average_down(adg_ctx* ctx);

/* Replaces AmrDG::set_Dt (AmrDG.h:1253-1357) and the clamp Dt=min(Dt,T-t)
 * of AmrDG::evolve (AmrDG.h:1028-1029).  The result stays on the device for
 * the next step; dt_out (nullable) receives a blocking copy. */
int adg_set_dt(adg_ctx* ctx, double c_dt, double T, double t, double* dt_out);
/* Sets the device-resident step size directly (tests, restarts). */
int adg_set_dt_value(adg_ctx* ctx, double Dt);

/* Replaces AmrDG::set_predictor + the Picard loop of AmrDG::ADER
 * (AmrDG.h:1100-1122, AmrDG.tpp:387-435,495-537,744-827) on the grown boxes
 * of every level, plus the final modal->nodal evaluation, volume flux
 * (AmrDG::flux, AmrDG.h:1167-1205) and face traces (get_H_from_H_w_bd,
 * AmrDG.tpp:437-493).  lev = -1 runs every level in one launch.
 * mode 0: exactly P Picard iterations (the reference).
 * mode 1: learned warm start (no counterpart in the reference): the network
 *   of adg_mlp_* gives the start H_0 = A0 U_w + Delta, Picard iterations run
 *   until max|H_k - H_{k-1}| <= tol rho0 (adg_set_learned, rho0 the element's
 *   mean density, so tol is relative to the state); an element that does
 *   not pass is recomputed by mode 0 (the fallback), one that passes keeps
 *   the iterate it passed with.  MFMA predictors (node-lane and
 *   quad-gather, 2D and 3D).
 * n_fallback (nullable) receives the number of elements that fell back
 * (0 in mode 0). */
int adg_predictor(adg_ctx* ctx, int lev, int mode, int* n_fallback);
/* Predictor kernel variant: 0 = automatic, 1 = VALU (FP64 only), 2 = MFMA
 * quad-gather (every Q), 3 = MFMA node-lane (2D, and 3D at p <= 2: four
 * elements per wavefront for Q = 4, three for Q = 5).  Automatic picks
 * node-lane where its fragments and tiles fit the LDS, else quad-gather,
 * else VALU; FP64 at p = 1 uses VALU.  No counterpart in the reference: a tuning knob of this
 * implementation. */
int adg_set_predictor_kernel(adg_ctx* ctx, int kind);
/* Autotuner: times the predictor launch shapes that fit this context and
 * hierarchy ({VALU} for FP64, {quad-gather, 8 waves}, {node-lane, 8 waves},
 * {node-lane, 16 waves}) over `reps` plain launches each (the plain form's
 * kernels; the learned form runs the same shapes with the learned start
 * compiled in) on the current state and keeps the fastest; with a network
 * loaded it also times the warm-start
 * kernel with one and two 16-element tiles and one 8-element tile per
 * wavefront.  ms_out[7] (nullable) gets the average launch time of each
 * candidate (-1 where it does not apply), choice_out[3] (nullable) the
 * chosen {kind, waves, warm-start tiles (1, 2, or 0 for the 8-element
 * tile)}.  The predictor kernels round differently (FP64: VALU, quad-gather
 * and node-lane agree to ~1e-15 per step, not bitwise; the node-lane shapes
 * with 8 and 16 waves give the same bits), so a run that must be
 * bit-reproducible across processes records choice_out and pins it with
 * adg_set_predictor_kernel / adg_set_predictor_waves / adg_set_warm_tiles
 * instead of re-tuning.  No counterpart in the reference. */
int adg_autotune(adg_ctx* ctx, int reps, double* ms_out, int* choice_out);
/* Wavefronts per workgroup of the node-lane predictor: 8 (default) or 16.
 * No counterpart in the reference: a launch-shape knob of this
 * implementation, set by the autotuner. */
int adg_set_predictor_waves(adg_ctx* ctx, int waves_per_block);
/* Tiles per wavefront of the warm-start inference: 1 (16 elements, default),
 * 2 (two 16-element tiles) or 0 (one 8-element tile).  No counterpart in the
 * reference: a launch-shape knob, set by the autotuner. */
int adg_set_warm_tiles(adg_ctx* ctx, int tiles);

/* Debug readback of the last predictor H_w of a level, [q][Np_st][slot]
 * (only after adg_keep_Hw(ctx,1)). */
int adg_keep_Hw(adg_ctx* ctx, int on);
int adg_download_Hw(adg_ctx* ctx, int lev, double* host);

/* Replaces AmrDG::flux_bd + AmrDG::numflux / LLF_numflux for every
 * direction (AmrDG.h:1126-1137, AmrDG.tpp:539-650,
 * AmrDG_NumericalFlux.tpp:4-41): face-centred, once per face.
 * lev = -1: every level. */
int adg_face_flux(adg_ctx* ctx, int lev);
/* Readback of Fnum for (lev, d), returned as [q][qMp_st_bd][face]. */
int adg_download_Fnum(adg_ctx* ctx, int lev, int d, double* host);

/* Replaces AmrDG::update_U_w (AmrDG.tpp:652-742) on valid cells.
 * lev = -1: every level. */
int adg_corrector(adg_ctx* ctx, int lev);

/* Replaces the FluxRegister CrseAdd/FineAdd of AmrDG::ADER
 * (AmrDG.h:1138-1153) and AmrDG::AMR_flux_correction
 * (AmrDG_Level.tpp:416-446) -> L2ProjInterp::reflux
 * (AmrDG_Interpolater.tpp:7-72). crse_lev = -1: every coarse level.
 * Coarse cells covered by the finer level are not corrected: the average
 * down that follows (AmrDG.h:1000) overwrites them, so U_w equals the
 * reference's after adg_average_down, not between the two calls. */
int adg_reflux(adg_ctx* ctx, int crse_lev);

/* Replaces AmrDG::ADER (AmrDG.h:1071-1165): predictor, face flux,
 * corrector on all levels, then flux correction. Uses the device Dt. */
int adg_ader(adg_ctx* ctx);

/* The tail of one AmrDG::evolve iteration (AmrDG.h:1017-1029): t += Dt,
 * then set_Dt on the updated U_w and the clamp Dt = min(Dt, T - t), all on
 * the device.  t_out (nullable) gets a blocking copy of the new t. */
int adg_advance(adg_ctx* ctx, double c_dt, double T, double* t_out);

/* One iteration of the AmrDG::evolve loop without regrid, limiter and
 * plots (AmrDG.h:988-1029): ADER, average_down, FillPatch, t+=Dt, set_Dt,
 * clamp to T.  t stays on the device; t_out (nullable) gets a blocking copy. */
int adg_step(adg_ctx* ctx, double c_dt, double T, double* t_out);
/* nsteps iterations of adg_step replayed from one captured hipGraph (the
 * graph is captured on first use and re-captured when c_dt, T, the
 * predictor kernel or the steps per graph change).  A graph holds
 * adg_set_graph_unroll steps (default 1); steps left over after whole
 * graphs run as the same launches on the stream.  t_out (nullable) gets a
 * blocking copy of t. */
int adg_step_graph(adg_ctx* ctx, double c_dt, double T, int nsteps, double* t_out);
/* Steps captured per step graph, 1..64.  No counterpart in the reference: a
 * launch-shape knob of this implementation. */
int adg_set_graph_unroll(adg_ctx* ctx, int steps);
/* nsteps iterations of adg_step enqueued on the library stream as separate
 * launches, with no host synchronisation between them (the same kernels in
 * the same order as adg_step_graph: bitwise the same results).  On MI355X
 * with ROCm 7 the stream launches of a Kelvin-Helmholtz step at 10 240
 * elements run about 7 us faster than a replay of the captured graph (66
 * against 73 us per step) while the host keeps ahead (about 25 us of host
 * time per step), so this is the batched step of choice there; the graph
 * saves host time where the host cannot keep ahead.  No counterpart in the
 * reference: its evolve loop (AmrDG.h:988-1029) launches nothing. */
int adg_steps(adg_ctx* ctx, double c_dt, double T, int nsteps, double* t_out);
/* Sets the device-resident time. */
int adg_set_time(adg_ctx* ctx, double t);
/* Blocking copy of the device-resident time t and step size Dt (the members
 * t, Dt of Solver, Core/Solver.h; either pointer nullable). */
int adg_get_time(adg_ctx* ctx, double* t_out, double* dt_out);

/* TVB limiter settings, the limiter arguments of
 * AmrDG::setNumericalSettings (AmrDG.tpp:50-58: limiter_type, TVB_M,
 * AMR_TVB_C, t_limit).  type 0: no limiter; type 1: "TVB".  tvb_c[ntvb]
 * holds AMR_TVB_C per level (levels past ntvb use 1.0).  With type 1 every
 * step (adg_step, adg_step_graph) runs Limiter_w on the finest level after
 * the corrector and before the average down when n % t_limit == 0, n the
 * step count since this call (AmrDG.h:990-997).  Drops a captured step
 * graph. */
int adg_set_limiter(adg_ctx* ctx, int type, double M, int ntvb, const double* tvb_c, int t_limit);
/* Sets the device step counter n of the limiter's n % t_limit (the step
 * index of AmrDG::evolve, AmrDG.h:866-875; a restart continues from
 * restart_tstep).  Every step advanced by the device adds one. */
int adg_set_step_count(adg_ctx* ctx, int n);
/* Replaces AmrDG::Limiter_w (AmrDG.h:1657-1770) with Limiter_linear_tvb
 * (AmrDG.h:1772-1855): the characteristic TVB limiting of the linear modes
 * of every valid cell of the finest level, now, with the settings of
 * adg_set_limiter (M and AMR_TVB_C; used whatever its type).
 * n_limited (nullable) receives the number of cells limited (blocking). */
int adg_limit(adg_ctx* ctx, unsigned long long* n_limited);
/* Cells limited by every limiter launch so far (n_limited_total of
 * AmrDG::evolve, AmrDG.h:1036-1046); reset != 0 clears the count after
 * reading it.  total nullable. */
int adg_limiter_stats(adg_ctx* ctx, unsigned long long* total, int reset);

/* ---- tagging-driven regrid (AmrCore::regrid(0, t) of Mesh, Core/Mesh.h:128-196,
 * at the regrid points of AmrDG::evolve, AmrDG.h:936-986) ---- */

/* AMR settings of Mesh (Core/Mesh.h:128-147): max_level (index of the finest
 * level allowed), blocking factor (the reference sets 4, Mesh.h:146),
 * max_grid_size, grid efficiency, n_error_buf, n_proper (AMReX's defaults:
 * 128 in 2D / 32 in 3D, 0.7, 1, 1). */
int adg_set_amr(adg_ctx* ctx, int max_level, int blocking_factor, int max_grid_size, double grid_eff,
                int n_error_buf, int n_proper);
/* Refinement criterion of Compressible_Euler::pde_tag_cell_refinement
 * (Compressible_Euler.h:605-682) with amr_c per level (Mesh::amr_c; n
 * values, the last repeated):
 *   0: isentropic_vortex, cell centre within amr_c^2 of (params[0], params[1])
 *      (periodic distance; the reference's centre is (5, 5));
 *   1: kelvin_helmolz_instability, the TVB troubled-cell indicator: some
 *      characteristic slope above AMR_TVB_C[lev] * amr_c[lev] * h^2
 *      (AMR_TVB_C from adg_set_limiter);
 *   2: density gradient (no counterpart in the reference): a face
 *      neighbour's cell-mean density differs by more than amr_c[lev] times
 *      the cell's own. */
int adg_set_tagging(adg_ctx* ctx, int criterion, int n, const double* amr_c, int nprm, const double* params);
/* Replaces AmrDG::AMR_tag_cell_refinement (AmrDG_TagCellRef.tpp:4-52) on
 * level lev: one byte per valid cell of the level (patch order, x fastest),
 * 1 = tagged (blocking). */
int adg_tag_cells(adg_ctx* ctx, int lev, unsigned char* tags);
/* Replaces Mesh::regrid(0, t) (AmrDG.h:948): tags levels 0..min(finest,
 * max_level-1), makes the new fine-level boxes (AmrCore::MakeNewGrids:
 * n_error_buf buffer, blocking-factor coarsening, Berger-Rigoutsos
 * clustering at grid_eff, nesting of each new level in the one below it by
 * n_proper, max_grid_size), and when they differ from the current ones
 * remakes the hierarchy as adg_remake.  *changed (nullable) = 1 when the
 * boxes changed. */
int adg_regrid(adg_ctx* ctx, double time, int* changed);
/* Replaces Mesh::RemakeLevel / MakeNewLevelFromCoarse / ClearLevel
 * (Core/Mesh.h:157-196 -> AmrDG_Level.tpp:52-224) for a given layout: level 0
 * must keep its boxes; every slot (valid and ghost) of a new level l >= 1
 * takes the old level-l valid cell at its place, else the L2 projection of
 * the new level l-1 parent (FillPatchTwoLevels / InterpFromCoarseLevel with
 * L2ProjInterp); transmissive boundary ghosts mirror; then the flux
 * registers, ghost lists and face masks are rebuilt (AMR_set_flux_registers).
 * Drops a captured step graph. */
int adg_remake(adg_ctx* ctx, int nlev, const int* npatch_per_lev, const int* boxes);
/* Current layout: *nlev, npatch[nlev], boxes (lo[dim], hi[dim] per patch,
 * at most cap patches written); returns the number of patches. */
int adg_get_levels(adg_ctx* ctx, int* nlev, int* npatch, int* boxes, int cap);

/* Conserved totals sum(vol * U_w[q][0]) over uncovered valid cells, FP64
 * (the quantity LpNorm_DG_AMR integrates over, AmrDG.h:1399-1578). */
int adg_conserved_totals(adg_ctx* ctx, double* totals);

/* The same totals, written to DEVICE memory d_totals (Q doubles) on the
 * context stream, bitwise equal to adg_conserved_totals; no counterpart in
 * the reference.  For callers that cannot allocate (code under HIP graph
 * capture, tight serving loops): it calls no hipMalloc / hipFree and no host
 * synchronisation, and keeps its per-workgroup partial sums in a static
 * device array of 5*8192 doubles per GPU, so it accepts at most
 * 256*(5*8192/Q) uncovered cells, 2 621 440 in 2D and 2 097 152 in 3D
 * (ADG_E_ARG beyond that).  Calls that overlap on
 * one GPU (several streams or contexts at once) race on that array. */
int adg_conserved_totals_noalloc(adg_ctx* ctx, double* d_totals);

/* Composite Lp error (p = 1 or 2) of conserved variable q over the uncovered
 * valid cells against the exact solution at time t, as LpNorm_DG_AMR
 * (AmrDG.h:1399-1578): (P+1)^D Gauss-Legendre points per cell, normalised
 * by the uncovered volume.  Exact solutions: ic_case 0 (isentropic vortex)
 * and 5 (density wave) translated periodically by their background
 * velocity, 2 (uniform state).  params: the initial condition's parameters,
 * 5 values read (unused ones ignored). */
int adg_lp_error(adg_ctx* ctx, int q, int p, int ic_case, const double* params, double t, double* out);

/* Blocks until all work on the context stream has finished. */
int adg_sync(adg_ctx* ctx);
/* The hipStream_t the context launches on. */
void* adg_stream(adg_ctx* ctx);
/* Basis sizes: out[0..6] = Np_s, Np_st, qMp_s, qMp_st, qMp_st_bd, N, Q. */
int adg_sizes(adg_ctx* ctx, int* out7);
/* Copies a named FP64 table of the context (same names as the reference
 * members: nodes weights V Vinv Mk_h_w_inv Mk_pred Sk_predVinv Sk_corr
 * Mkbdm Mkbdp quadmat Minv P_cf P_fc P_flux_low P_flux_high).  Returns the
 * number of doubles or a negative code. */
int adg_get_table(adg_ctx* ctx, const char* name, double* out, int cap);

/* ---- learned warm start of the predictor (no counterpart in the
 * reference).  A 3-layer GELU MLP, widths nin_p -> 128 -> 128 -> nout_p,
 * maps x = [U_w/rho0, local CFL number, Dt/dx_d] (nin = Q Np_s + 1 + D) to
 * y = (H_w* - A0 U_w) / (rho0 max_d Dt/dx_d) (nout = Q Np_st), H_w* the
 * converged predictor; BF16 MFMA products with FP32 accumulation, trained on
 * the device by its own forward / backward kernels with Adam. */
/* out[0..6] = nin, nin_p, nh, nout, nout_p, number of weights and biases,
 * number of normalisation values (mean/deviation of x and y). */
int adg_mlp_sizes(adg_ctx* ctx, int* out7);
/* Glorot-uniform weights (seeded), zero biases, identity normalisation,
 * Adam state cleared. */
int adg_mlp_init(adg_ctx* ctx, unsigned long long seed);
/* Flat parameters [W1 nh x nin_p][b1][W2 nh x nh][b2][W3 nout_p x nh][b3]
 * followed by [mu_x nin_p][sd_x nin_p][mu_y nout_p][sd_y nout_p]. */
int adg_mlp_get_params(adg_ctx* ctx, float* out, long cap);
int adg_mlp_set_params(adg_ctx* ctx, const float* in, long n);
/* Runs the predictor with `iters` Picard iterations on the current state and
 * appends (x, y) of every valid element to the device dataset; *n_total
 * (nullable) gets the dataset size.  Leaves Vol/Tr of that predictor. */
int adg_mlp_collect(adg_ctx* ctx, int iters, long* n_total);
/* Copies dataset rows [start, start + n) (raw, unstandardised) to the host;
 * x: n x nin_p, y: n x nout_p (either nullable). */
int adg_mlp_dataset(adg_ctx* ctx, long start, long n, float* x, float* y);
int adg_mlp_dataset_clear(adg_ctx* ctx);
/* Column mean / deviation of the dataset become the normalisation. */
int adg_mlp_fit_normalization(adg_ctx* ctx);
/* nsteps Adam steps (lr, beta 0.9 / 0.999) on random minibatches of `batch`
 * rows (a multiple of 16); *loss_out = mean loss of the last steps. */
int adg_mlp_train(adg_ctx* ctx, int nsteps, int batch, double lr, unsigned long long seed, double* loss_out);
/* Loss and gradient (flat weight layout) of n standardised rows x
 * (n x nin_p), y (n x nout_p), n a multiple of 16. */
int adg_mlp_loss_grad(adg_ctx* ctx, int n, const float* x, const float* y, double* loss, float* grad);
/* Warm-start increments Delta [nout][slot] of level lev (all its slots). */
int adg_mlp_warm_start(adg_ctx* ctx, int lev, double* out);
/* Predictor mode used by adg_step / adg_ader / adg_step_graph:
 * on = 1 -> mode 1 with tolerance tol and at most max_iters Picard
 * iterations from the learned start; on = 2 -> diagnostic: exactly
 * max_iters iterations from the learned start, no check, no fallback. */
int adg_set_learned(adg_ctx* ctx, int on, double tol, int max_iters);
/* out[0..1] = elements run through mode 1, elements that fell back, since
 * the last reset; reset != 0 clears them. */
int adg_learned_stats(adg_ctx* ctx, unsigned long long* out2, int reset);

const char* adg_last_error(adg_ctx* ctx);
void adg_destroy(adg_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
