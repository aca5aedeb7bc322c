/*
 * adg_host.h -- host-only diagnostics of libadg.so (no device work).
 *
 * No counterpart in the reference: these expose the host AMR substrate
 * (the flattened hierarchy lists of adg_set_levels) and the MFMA lane
 * layouts the kernels assume, for tests.
 */
#ifndef ADG_HOST_H
#define ADG_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

/* Coarse-fine face masks of a box layout as AMR_set_flux_registers
 * (AmrDG_Level.tpp:240-414) defines them: kind 0 b_coarse (with the
 * ownership rule), 1 b_fine, 2 child index, for the faces of direction d of
 * level lev; kind 3: counts {slots, ghosts, faces d0..d2, register moments,
 * reflux rows, uncovered cells}.  Returns the number of ints or < 0. */
int adg_host_inspect(int dim, const int* dom_n, const int* periodic, int nlev, const int* npatch,
                     const int* boxes, int lev, int d, int kind, int* out, int cap);

/* A named element table (the names of adg_get_table) of degree P in D
 * dimensions, built on the host as adg_create builds it.  Returns the number
 * of doubles or < 0. */
int adg_host_table(int P, int D, const char* name, double* out, int cap);

/* New fine-level boxes from refinement tags, on the host (the grid
 * generation of adg_regrid, AmrCore::MakeNewGrids semantics): the old layout
 * (nlev_old, npatch_old, boxes_old as adg_set_levels takes them), the tagged
 * cells of levels 0..min(nlev_old-1, max_level-1) (ntags[l] cells of level l,
 * dim ints each, concatenated).  Writes *nlev_new, npatch_new[] (level 0
 * unchanged) and up to cap_boxes boxes; returns the number of boxes or < 0. */
int adg_host_make_grids(int dim, const int* dom0, const int* periodic, int max_level, int blocking_factor,
                        int max_grid_size, double grid_eff, int n_error_buf, int n_proper, int nlev_old,
                        const int* npatch_old, const int* boxes_old, const int* ntags, const int* tags,
                        int* nlev_new, int* npatch_new, int* boxes_new, int cap_boxes);

/* Runs one 16x16x4 MFMA on the current device and returns the 64x4
 * accumulator (prec 0: FP64, 1: FP32; mode 0 probes A rows, mode 1 B
 * columns).  Needs a GPU. */
int adg_probe_mfma_layout(int prec, int mode, double* out);

#ifdef __cplusplus
}
#endif
#endif
