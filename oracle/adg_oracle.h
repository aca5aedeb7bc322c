/*
 * adg_oracle.h -- CPU restatement of the reference AmrDG ADER-DG hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded
 * by tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of
 * bench.py, never by the product library or its host code.
 *
 * It restates, in FP64 and in the reference's loop order, the arithmetic of
 *   Source/NumericalMethod/AmrDG/AmrDG.h:53-353        (basis, GL quadrature)
 *   Source/NumericalMethod/AmrDG/AmrDG_ElementMatrix.tpp (element matrices)
 *   Source/NumericalMethod/AmrDG/AmrDG.tpp:336-827       (predictor, numflux, corrector)
 *   Source/NumericalMethod/AmrDG/AmrDG_NumericalFlux.tpp:4-41 (LLF)
 *   Source/NumericalMethod/AmrDG/AmrDG.h:1071-1357       (ADER order, set_Dt)
 *   Source/NumericalMethod/AmrDG/AmrDG_Interpolater.tpp  (scatter/gather/reflux)
 *   Source/NumericalMethod/AmrDG/AmrDG_Level.tpp:240-446 (C-F masks, flux registers)
 *   Source/Model/CompressibleEuler/Compressible_Euler.h:370-604, .cpp:5-70
 * plus the AMReX FillPatch / FluxRegister semantics those call sites rely on.
 *
 * Layout convention shared with the product's C-ABI (an interface, not an
 * algorithm): per level, modal data is [q][comp][slot]; slots enumerate the
 * grown (1 ghost cell) box of every patch of the level, patches in the order
 * given, x fastest:  slot = off[b] + ((kk*g1)+jj)*g0 + ii,  ii = i-lo0+1, ...
 * Face data of direction d is [q][m][fslot] over the valid faces of every
 * patch (n_d+1 faces along d).
 */
#ifndef ADG_ORACLE_H
#define ADG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_tables orc_tables;
typedef struct orc_hier orc_hier;

/* ---- tables (basis, quadrature, element and projection matrices) ---- */
orc_tables* orc_tables_create(int P, int D);
void        orc_tables_destroy(orc_tables*);
/* sizes: [Np_s, Np_st, qMp_s, qMp_st, qMp_st_bd, N, Q] */
void        orc_tables_sizes(const orc_tables*, int* out7);
/* Copy a named table into out (row-major as in the reference's Eigen (r,c)).
 * Returns number of doubles, or -1 for an unknown name.  Names:
 *  nodes weights val dval bd_val shifted_lo shifted_hi mi_s mi_st
 *  V Vinv Mk_h_w Mk_h_w_inv Mk_pred Sk_pred Sk_predVinv Mk_corr Sk_corr
 *  Mkbdm Mkbdp quadmat Minv P_cf P_fc P_flux_low P_flux_high
 * Per-direction tables are stacked [d][r][c]; P_cf/P_fc [child][r][c];
 * P_flux_* [d][child][r][c]. */
int         orc_tables_get(const orc_tables*, const char* name, double* out, int cap);

/* ---- hierarchy (boxes per level, ratio 2, 1 ghost cell) ---- */
/* boxes: for each level, npatch[l] entries of lo[D], hi[D] (ints, level index
 * space).  dom_n: level-0 cells per dim.  periodic: per dim flags. */
orc_hier*   orc_hier_create(int D, int nlev, const int* npatch, const int* boxes,
                            const int* dom_n, const double* prob_lo,
                            const double* prob_hi, const int* periodic);
void        orc_hier_destroy(orc_hier*);
/* 0 (default): the reference's reflux rule -- a coarse cell is corrected
 * only where one of its faces is flagged in its own box's ownership-filtered
 * cf_face_b_coarse (AmrDG_Interpolater.tpp:39-69).  1: every coarse cell the
 * registers reach is corrected (no counterpart in the reference). */
void        orc_hier_set_conservative_reflux(orc_hier*, int on);
/* 0 (default): set_Dt over every valid cell of every level, covered
 * coarse cells included (AmrDG.h:1253-1357).  1: over the uncovered cells
 * only (no counterpart in the reference). */
void        orc_hier_set_cfl_uncovered(orc_hier*, int on);
int         orc_hier_nslot(const orc_hier*, int lev);
int         orc_hier_nfslot(const orc_hier*, int lev, int d);
/* 1 if slot is a valid cell, 0 if ghost; out has nslot ints */
void        orc_hier_valid_mask(const orc_hier*, int lev, int* out);
/* flux-register face masks, restated from AMR_set_flux_registers:
 * kind 0 = b_coarse, 1 = b_fine, 2 = child idx; out has nfslot ints */
void        orc_hier_face_mask(const orc_hier*, int lev, int d, int kind, int* out);
/* cell-centre coordinate of a slot (ghosts included, not wrapped) */
void        orc_hier_slot_centre(const orc_hier*, int lev, int slot, double* x);

/* ---- initial condition: L2 projection of a pointwise IC ----
 * case 0: isentropic vortex (params: beta, xc, yc, uinf, vinf)
 * case 1: Kelvin-Helmholtz (params: amp, noise_amp, seed)  [no reference IC]
 * case 2: uniform state (params: rho, u, v, [w,] p)
 * case 3: radial shock tube (params: r0, rho_in, p_in, rho_out, p_out, centre[D])
 * case 5: smooth density wave rho = 1 + amp prod_d sin(2 pi x_d) on the unit
 *         cube, velocity (1, 0.5, 0.25), pressure p0 (params: amp, p0)
 * case 4: Kelvin-Helmholtz with v = sum_{k=1..8} a_k sin(2 pi k x),
 *         a_k = amax * r_k, r_k seeded uniform (params: seed, amax)  [no reference IC]
 * case 7: Kelvin-Helmholtz as case 1 with the noise constant on the cells of
 *         an N x N grid of the unit square (params: amp, noise_amp, seed, N)
 *         [no reference IC]
 * Fills every slot (ghosts too, as AmrDG::set_initial_condition on growntilebox). */
void        orc_project_ic(const orc_tables*, const orc_hier*, int lev, int ic_case,
                           const double* params, double gamma, double* Uw);

/* ---- pointwise model (Compressible_Euler) ---- */
double      orc_pressure(int D, const double* u, double gamma);
double      orc_pde_flux(int D, int d, int q, const double* u, double gamma);
double      orc_pde_dflux(int D, int d, const double* u, double gamma);
double      orc_cfl_lambda(int D, int d, const double* u, double gamma);
double      orc_llf(double uL, double uR, double fL, double fR, double dfL, double dfR);

/* ---- hot-path stages (Uw: array of nlev pointers to per-level data) ---- */
/* Predictor on the grown boxes of one level: P Picard iterations from the
 * set_predictor guess.  Hw_out: [q][Np_st][slot]. */
void        orc_predictor(const orc_tables*, const orc_hier*, int lev,
                          const double* Uw, double Dt, double gamma, double* Hw_out);
/* One AmrDG::ADER call (all levels, finest first, then flux correction).
 * Fnum_out (nullable): per level per d face arrays [q][qMp_st_bd][fslot],
 * packed level-major then d. */
void        orc_ader(const orc_tables*, const orc_hier*, double** Uw, double Dt,
                     double gamma, double* Fnum_out);
void        orc_average_down(const orc_tables*, const orc_hier*, double** Uw);
void        orc_fill_patch(const orc_tables*, const orc_hier*, double** Uw);
/* set_Dt: returns c_dt * min over levels and valid cells */
double      orc_set_dt(const orc_tables*, const orc_hier*, double** Uw,
                       double c_dt, double gamma);
/* Full time steps as AmrDG::evolve without regrid/limiter/plot:
 * ADER, average_down, FillPatch, t += Dt, set_Dt, clamp to T.
 * Returns the number of steps taken; *t_io is advanced; Dt_io holds the
 * step size for the next step. nsteps_max <= 0 means run to T. */
int         orc_evolve(const orc_tables*, const orc_hier*, double** Uw,
                       double* t_io, double* Dt_io, double T, double c_dt,
                       double gamma, int nsteps_max);

/* Composite Lp error of component q vs the IC case evaluated at time t
 * (periodic shift by background velocity for the vortex), over uncovered
 * cells, normalised by uncovered volume (LpNorm_DG_AMR). */
double      orc_lp_error(const orc_tables*, const orc_hier*, double** Uw, int q,
                         int p, int ic_case, const double* params, double gamma,
                         double t);
/* TVB limiter of one level (Limiter_w + Limiter_linear_tvb, AmrDG.h:1657-1855):
 * FillBoundary, then per valid cell the characteristic minmodB limiting of
 * the linear modes (threshold tvb_c[lev] * M * min(dx)^2); a troubled cell
 * keeps mode 0, gets the limited linear modes and loses every higher mode.
 * Returns the number of cells limited. */
long        orc_limiter_tvb(const orc_tables*, const orc_hier*, double** Uw, int lev,
                            double M, const double* tvb_c, double gamma);
/* MultiFab::FillBoundary of one level: same-level and periodic ghost copies. */
void        orc_fill_boundary(const orc_tables*, const orc_hier*, double** Uw, int lev);
/* pde_EV_Lmatrix / pde_EV_Rmatrix at state u (row-major Q x Q). */
void        orc_ev_matrices(int D, int d, const double* u, double gamma, double* L, double* R);
/* orc_evolve with the limiter on the finest level every t_limit steps
 * (t_limit <= 0: never); *n_limited (nullable) gets the cells limited. */
int         orc_evolve_lim(const orc_tables*, const orc_hier*, double** Uw,
                           double* t_io, double* Dt_io, double T, double c_dt,
                           double gamma, int nsteps_max, double M, const double* tvb_c,
                           int t_limit, long* n_limited);
/* Refinement tags of level lev (AMR_tag_cell_refinement + the model's
 * pde_tag_cell_refinement): crit 0 vortex geometry (prm = centre), 1 the TVB
 * troubled-cell indicator (threshold tvb_c * amr_c * min(dx)^2), 2 density
 * gradient (relative cell-mean jump above amr_c).  One byte per valid cell,
 * patch by patch, x fastest; returns the number tagged. */
long        orc_tag_cells(const orc_tables*, const orc_hier*, double** Uw, int lev, int crit,
                          double amr_c, double tvb_c, const double* prm, double gamma, unsigned char* out);
/* RemakeLevel / MakeNewLevelFromCoarse onto the layout of hn (level 0 equal):
 * old valid data where the old level had it, else the L2 projection of the
 * new coarse parent; 0, or -1 when the layout does not nest. */
int         orc_remake(const orc_tables*, const orc_hier* h_old, double** Uw_old,
                       const orc_hier* h_new, double** Uw_new);
/* The 1D nodes of the L2 norm's 2(P+1)-point rule (L2Norm_DG_AMR), in the
 * reference's order; returns their count. */
int         orc_l2_nodes(int P, double* out);
/* Conserved totals: sum over uncovered valid cells of vol * Uw[q][0] */
void        orc_totals(const orc_tables*, const orc_hier*, double** Uw, double* out);

/* ---- learned warm start (no counterpart in the reference) ---- */
/* Predictor with `iters` Picard iterations (iters = P is orc_predictor). */
void        orc_predictor_iters(const orc_tables*, const orc_hier*, int lev, const double* Uw,
                                double Dt, double gamma, int iters, double* Hw_out);
/* MLP widths [nin, nin_p, nh, nout, nout_p]: nin = Q Np_s + 1 + D inputs,
 * nout = Q Np_st outputs, _p padded to a multiple of 16, nh = 128. */
void        orc_mlp_dims(const orc_tables*, int* out5);
/* Features of one element, U [Q][Np_s], tau[d] = Dt/dx_d:
 * x = [U/rho0 (q-major), nu = sum_d tau_d (|u_d| + c) at the cell average,
 * tau_0 .. tau_{D-1}] */
void        orc_mlp_features(const orc_tables*, const double* U, const double* tau, double gamma,
                             double* x_out);
/* Target of one element: y = (H - A0 U) / (rho0 max_d tau_d), H [Q][Np_st],
 * A0 = Mk_h_w^-1 Mk_pred. */
void        orc_mlp_target(const orc_tables*, const double* U, const double* H, const double* tau,
                           double* y_out);
/* Loss (mean squared error over n rows and the first nout columns) and its
 * gradient (nullable) of the MLP with flat float params
 * [W1 nh x nin_p][b1][W2 nh x nh][b2][W3 nout_p x nh][b3]; x n x nin_p,
 * y n x nout_p.  bf16 = 1 rounds every matrix-product operand to bfloat16
 * as the MFMA kernels do; accumulation in double.  ypred (nullable) gets the
 * n x nout_p outputs. */
double      orc_mlp_loss_grad(int nin_p, int nh, int nout_p, int nout, int n, const float* params,
                              const float* x, const float* y, int bf16, double* grad, double* ypred);

#ifdef __cplusplus
}
#endif
#endif
