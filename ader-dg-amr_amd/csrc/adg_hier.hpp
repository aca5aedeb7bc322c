// adg_hier.hpp -- host-side AMR substrate: the device layout of a patch
// hierarchy and every index list the kernels walk.
//
// The reference keeps one AMReX MultiFab per (level, variable) with one ghost
// cell and lets AMReX FillPatch / FluxRegister / average_down move data
// between boxes and levels.  Here the same semantics are flattened once, on
// the host, into plain index lists so that each stage is a single kernel
// launch over all patches of all levels:
//   - slots: every cell (valid + ghost) of every patch of every level, levels
//     concatenated, patches in order, x fastest inside a patch;
//   - faces: every valid face of direction d of every patch;
//   - ghost sources (same-level / periodic copy or coarse L2 projection);
//   - average-down groups (coarse cell <- 2^D fine cells);
//   - flux-register algebra reduced to a sparse list of signed face moments
//     per coarse cell (FluxRegister CrseAdd/FineAdd/Reflux + the reference's
//     cf_face_b_coarse ownership rule, AmrDG_Level.tpp:240-446).
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

namespace adg {

// f(i) for i in [0, n) on up to 16 host threads, items handed out one at a
// time; the first exception thrown by any item is rethrown.
template <class F>
inline void parallel_for(int n, F&& f) {
    int nt = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    nt = std::min(nt, n);
    if (nt <= 1) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int> next{0};
    std::exception_ptr err;
    std::mutex m;
    auto work = [&]() {
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= n) return;
            try {
                f(i);
            } catch (...) {
                std::lock_guard<std::mutex> g(m);
                if (!err) err = std::current_exception();
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work);
    work();
    for (std::thread& t : th) t.join();
    if (err) std::rethrow_exception(err);
}

struct LevelLayout {
    int dom[3] = {1, 1, 1};
    double dx[3] = {1, 1, 1};
    long slot_off = 0, nslot = 0;
    long face_off[3] = {0, 0, 0}, nface[3] = {0, 0, 0};
    int p_begin = 0, p_end = 0;
    long valid_begin = 0, valid_end = 0;  // range in valid_list
};

struct PatchLayout {
    int lev = 0;
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    int n[3] = {1, 1, 1}, g[3] = {1, 1, 1};
    long off = 0;        // global slot offset
    long foff[3] = {0, 0, 0};  // global face offset per direction
};

struct Hierarchy {
    int D = 2, nlev = 0;
    bool independent = false;   // patches are separate periodic domains
    int dom0[3] = {1, 1, 1};
    double plo[3] = {0, 0, 0}, phi[3] = {1, 1, 1};
    int per[3] = {1, 1, 1};
    std::vector<LevelLayout> lev;
    std::vector<PatchLayout> patch;
    long NS = 0;               // total slots
    long NF[3] = {0, 0, 0};    // total faces per direction

    std::vector<uint8_t> slot_lev;  // level of every slot
    std::vector<int> valid_list;    // global slots of valid cells, level-major
    std::vector<int> cell_face[3][2];  // per valid cell: lo/hi face id per d
    std::vector<int> face_lr[3];    // per face: left slot, right slot (2 ints)
    std::vector<uint8_t> face_lev[3];

    // ghost sources: kind & 3 = 0 copy, 1 L2 projection from coarse parent;
    // kind >> 2 = mask of mirrored directions (transmissive boundary ghosts)
    std::vector<int> g_slot, g_src;
    std::vector<uint8_t> g_kind, g_child;

    // average down, per fine level lf (index lf): coarse slot + 2^D fine slots
    std::vector<std::vector<int>> ad_coarse, ad_fine;

    // flux-register algebra
    // moments: kind 0 coarse face (Mkbd{p|m}), kind 1 fine face (P_flux_{high|low})
    std::vector<int> mom_face;     // face id (global, within its direction)
    std::vector<uint8_t> mom_d, mom_kind, mom_sel, mom_lev;  // sel: side (0 m/low,1 p/high) | child<<1
    std::vector<int> rc_slot, rc_start, rc_count;  // corrected coarse cells
    std::vector<uint8_t> rc_lev;
    std::vector<int> re_mom;       // entries: moment index
    std::vector<double> re_coef;   //          coefficient (+-1, summed)

    // uncovered valid cells (for conserved totals)
    std::vector<int> uncovered;

    // The predictor's slots: every slot but the copy ghosts (same-level or
    // periodic copies, not mirrored), in slot order, level l's in
    // [pred_off[l], pred_off[l+1]).  A copy ghost's predictor result equals
    // its source cell's bit for bit (the predictor is element-local), so the
    // face flux reads the source's traces (face_tr: face_lr with copy ghosts
    // replaced by their sources) and the copy is never predicted.
    std::vector<int> pred_slots;
    std::vector<long> pred_off;
    std::vector<int> face_tr[3];
    std::vector<int> copy_ghost, copy_src;

    // the limiter's cells (valid cells of the finest level) and their
    // same-level neighbours per direction and side, a copy ghost replaced by
    // its source cell ([2d + side][cell]); centres of the uncovered cells
    // ([cell][3])
    std::vector<int> lim_cells, lim_nbr;
    std::vector<double> unc_centre;

    // face masks for tests (global face ids per direction)
    std::vector<int8_t> bcoarse[3], bfine[3], cidx[3];

    // owner grids: per level, the slot of the valid cell at each cell of the
    // level's domain (x fastest), -1 where the level has no cell
    std::vector<std::vector<int>> grid_slot;
    std::vector<std::vector<int>> grid_patch;   // the patch of that cell, -1 elsewhere
    // build scratch: (kind, d) face -> moment id tables, -1 everywhere
    // between builds (a build resets the entries it set)
    std::vector<int> mom_tab[2][3];

    long slot_of(int p, const int* iv) const;
    long face_of(int p, int d, const int* iv) const;
    // Moves the host storage of every list of o into this hierarchy, emptied
    // (capacity kept), so that a rebuild writes into memory that is already
    // mapped instead of faulting fresh pages in; o is left empty.
    void take_storage(Hierarchy& o);
};

// Throws std::runtime_error on malformed input.
// independent != 0: one level whose patches are independent periodic
// domains (a batch of separate problems): ghost cells wrap inside their own
// patch.
// conservative_reflux != 0: every uncovered coarse cell the registers touch
// is corrected; 0 (the reference): only cells with a face flagged in their
// own patch's ownership-filtered cf_face_b_coarse (AmrDG_Interpolater.tpp:
// 39-69), which drops the correction of a cell whose coarse-fine face is the
// low face of its patch -- the reference loses mass at such faces.
// reuse (nullable): a hierarchy no longer needed whose host storage the new
// one takes over (Hierarchy::take_storage); the lists are the same with or
// without it.
Hierarchy build_hierarchy(int D, const int* dom0, const double* plo, const double* phi,
                          const int* per, int nlev, const int* npatch, const int* boxes, int independent = 0,
                          int conservative_reflux = 0, Hierarchy* reuse = nullptr);

}  // namespace adg
