// AmrDG.hpp -- C++ host side of the MI355X ADER-DG path.
//
// The reference's compile-time interface, kept with its class and member
// names so that a driver written against it reads the same:
//   Simulation<NM,EQ>          Source/Core/Simulation.h:25-200
//   Solver<NM> (CRTP)          Source/Core/Solver.h:140-200
//   AmrDG<P>                   Source/NumericalMethod/AmrDG/AmrDG.h:400-445
//   ModelEquation<EQ> (CRTP)   Source/Core/ModelEquation.h:15-130
//   Compressible_Euler         Source/Model/CompressibleEuler/Compressible_Euler.{h,cpp}
//   Mesh<NM>                   Source/Core/Mesh.h:120-200
//   BoundaryCondition<EQ,NM>   Source/Core/BoundaryCondition.h:20-100
// Every MFIter loop body of the hot path is a call into libadg.so
// (include/adg.h); the modal data stay on the GPU between calls.
//
// Differences from the reference, by design of this library:
//  - without Mesh::set_levels the fine levels start as full-domain grids
//    (InitFromScratch) and the first step regrids them from the tags, as in
//    the reference; with set_levels they start from the given boxes;
//  - the refinement criterion of a case is the reference's (vortex geometry,
//    the Kelvin-Helmholtz TVB indicator) or, for the cases the reference
//    does not tag (the noisy / multi-mode KH states, the shock tubes), the
//    library's density-gradient criterion;
//  - errors are exceptions (AdgError) on this side of the C-ABI, carrying
//    adg_last_error(); the reference has no error reporting;
//  - setLimitInitialState(true) applies the limiter once to the projected
//    initial state (a discontinuous state projected at p >= 2 has negative
//    nodal density and pressure, and the first step diverges without it);
//  - setConservativeReflux(true) corrects every coarse cell the flux
//    registers reach (the reference skips cells whose coarse-fine face is
//    the low face of their box, and loses mass there);
//  - setCflUncovered(true) takes the time step from the uncovered cells
//    only (the reference also reads covered coarse cells);
//  - plotfiles go to Simulation::setDataDir when given, else to the
//    reference's "../Results/Simulation Data/<case>" directory;
//  - the space dimension is the compile-time ADG_SPACEDIM (default 2), as
//    AMREX_SPACEDIM is in the reference.
#ifndef ADG_HOST_AMRDG_HPP
#define ADG_HOST_AMRDG_HPP

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <dirent.h>
#include <unistd.h>
#include <fstream>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "adg.h"
#include "plotfile.hpp"

#ifndef ADG_SPACEDIM
#define ADG_SPACEDIM 2
#endif

namespace adg {

constexpr int SPACEDIM = ADG_SPACEDIM;
static_assert(SPACEDIM == 2 || SPACEDIM == 3, "ADG_SPACEDIM must be 2 or 3");

class AdgError : public std::runtime_error {
  public:
    using std::runtime_error::runtime_error;
};

// Physical extent of the domain (amrex::RealBox).
struct RealBox {
    std::array<double, 3> lo{{0.0, 0.0, 0.0}};
    std::array<double, 3> hi{{1.0, 1.0, 1.0}};
};

// Cell-centred index box of one patch in its level's index space.
struct Box {
    std::array<int, 3> lo{{0, 0, 0}};
    std::array<int, 3> hi{{0, 0, 0}};
};

// Boxes of edge mg covering [lo, hi] of a level, x fastest.
inline std::vector<Box> tile_boxes(const std::array<int, 3>& lo, const std::array<int, 3>& hi, int mg) {
    std::vector<Box> out;
    const int kz0 = SPACEDIM == 3 ? lo[2] : 0, kz1 = SPACEDIM == 3 ? hi[2] : 0;
    for (int k = kz0; k <= kz1; k += mg)
        for (int j = lo[1]; j <= hi[1]; j += mg)
            for (int i = lo[0]; i <= hi[0]; i += mg) {
                Box b;
                b.lo = {{i, j, k}};
                b.hi = {{std::min(i + mg - 1, hi[0]), std::min(j + mg - 1, hi[1]),
                         SPACEDIM == 3 ? std::min(k + mg - 1, hi[2]) : 0}};
                out.push_back(b);
            }
    return out;
}

// ---------------------------------------------------------------------------
// ModelEquation<EQ>: CRTP base of the PDE model (Core/ModelEquation.h:15-130).
// The device path bakes the compressible Euler flux in with gamma as a
// parameter, so the pointwise pde_flux / pde_dflux / pde_cfl_lambda
// callbacks stay with the reference's CPU path; this base carries the model
// settings and the initial state the device projects.
template <typename EquationType>
class ModelEquation {
  public:
    ModelEquation() = default;
    virtual ~ModelEquation() = default;

    template <typename... Args>
    void settings(Args... args) {
        static_cast<EquationType*>(this)->settings(args...);
    }
    void set_pde_numeric_limits() { static_cast<EquationType*>(this)->set_pde_numeric_limits(); }

    // Initial state as the device IC case and its parameters on a domain.
    void pde_IC_case(const RealBox& rb, int* ic_case, std::vector<double>* params) const {
        static_cast<const EquationType*>(this)->pde_IC_case(rb, ic_case, params);
    }
    // Whether adg_lp_error holds the exact solution of this case at time t.
    bool pde_exact_solution(int* ic_case, std::vector<double>* params, const RealBox& rb) const {
        return static_cast<const EquationType*>(this)->pde_exact_solution(ic_case, params, rb);
    }

    int Q_model = 0;
    int Q_model_unique = 0;
    std::string model_case;
    bool flag_source_term = false;
    double gamma_adiab = 1.4;

  protected:
    double PDE_NUMERIC_LIMIT = 0.0;
};

// Compressible_Euler::settings (Compressible_Euler.cpp:5-48) and the
// initial states of pde_IC (Compressible_Euler.h:141-368) this library runs.
class Compressible_Euler : public ModelEquation<Compressible_Euler> {
  public:
    void settings(std::string euler_case) {
        model_case = euler_case;
        if (model_case == "keplerian_disc" || model_case == "double_mach_reflection" ||
            model_case == "richtmeyer_meshkov_instability" || model_case == "isentropic_vortex_static")
            throw AdgError("Compressible_Euler: case '" + model_case + "' is not part of this library");
        if (!known(model_case)) throw AdgError("Compressible_Euler: unknown case '" + model_case + "'");
        if (SPACEDIM != 2 && two_d_only(model_case))
            throw AdgError("Compressible_Euler: case '" + model_case + "' is two-dimensional");
        flag_source_term = false;
        Q_model = SPACEDIM + 2;
        Q_model_unique = Q_model;
        gamma_adiab = 1.4;
    }

    void set_pde_numeric_limits() { PDE_NUMERIC_LIMIT = 10 * std::numeric_limits<double>::epsilon(); }

    // Perturbation of the noisy / multi-mode Kelvin-Helmholtz variants.
    void setPerturbation(double amp, double noise_amp, unsigned long long seed) {
        kh_amp = amp;
        kh_noise = noise_amp;
        kh_seed = seed;
    }
    // Isentropic vortex centre, strength and background velocity.
    void setVortex(double beta, double xc, double yc, double u_inf, double v_inf) {
        vortex = {{beta, xc, yc, u_inf, v_inf}};
    }

    void pde_IC_case(const RealBox& rb, int* ic_case, std::vector<double>* prm) const {
        const std::string& c = model_case;
        if (c == "isentropic_vortex") {
            *ic_case = 0;   // Compressible_Euler.h:141-180
            *prm = {vortex[0], vortex[1], vortex[2], vortex[3], vortex[4]};
        } else if (c == "kelvin_helmolz_instability") {
            *ic_case = 6;   // Compressible_Euler.h:237-256
            *prm = {0.2, 0.05 / std::sqrt(2.0)};
        } else if (c == "kelvin_helmholtz_noise") {
            *ic_case = 1;   // v = amp sin(4 pi x) + seeded noise
            *prm = {kh_amp, kh_noise, (double)kh_seed};
        } else if (c == "kelvin_helmholtz_multimode") {
            *ic_case = 4;   // v = sum_k a_k sin(2 pi k x)
            *prm = {(double)kh_seed, kh_amp};
        } else if (c == "radial_shock_tube") {
            *ic_case = 3;   // Compressible_Euler.h:336-357, centred in the domain
            *prm = {0.4, 1.0, 1.0, 0.125, 0.125};
        } else if (c == "spherical_shock_tube") {
            *ic_case = 3;   // rho=1, p=1 inside r=0.4; rho=0.125, p=0.1 outside
            *prm = {0.4, 1.0, 1.0, 0.125, 0.1};
        } else if (c == "density_wave") {
            *ic_case = 5;
            *prm = {0.2, 1.0};
        } else {
            throw AdgError("Compressible_Euler: no initial state for '" + c + "'");
        }
        prm->resize(8, 0.0);   // the C-ABI reads up to 5 values
        (void)rb;
    }

    bool pde_exact_solution(int* ic_case, std::vector<double>* prm, const RealBox& rb) const {
        if (model_case != "isentropic_vortex" && model_case != "density_wave") return false;
        pde_IC_case(rb, ic_case, prm);
        return true;
    }

  private:
    static bool known(const std::string& c) {
        return c == "isentropic_vortex" || c == "kelvin_helmolz_instability" || c == "kelvin_helmholtz_noise" ||
               c == "kelvin_helmholtz_multimode" || c == "radial_shock_tube" || c == "spherical_shock_tube" ||
               c == "density_wave";
    }
    static bool two_d_only(const std::string& c) {
        return c == "isentropic_vortex" || c == "kelvin_helmolz_instability" || c == "kelvin_helmholtz_noise" ||
               c == "kelvin_helmholtz_multimode";
    }
    std::array<double, 5> vortex{{5.0, 5.0, 5.0, 1.0, 0.0}};   // Compressible_Euler.h:148-157
    double kh_amp = 0.01, kh_noise = 1e-3;
    unsigned long long kh_seed = 1;
};

// ---------------------------------------------------------------------------
// Mesh<NM>: geometry and the (static) level layout (Core/Mesh.h:120-200).
template <typename NumericalMethodType>
class Mesh {
  public:
    Mesh(const RealBox& rb, int max_level, const std::vector<int>& n_cell, int coord,
         const std::vector<std::array<int, 3>>& ref_ratios, const std::array<int, 3>& is_per,
         const std::vector<double>& amr_c, int dtn_regrid = 0, double dt_regrid = 0, int nghost = 1,
         int max_grid_size = 32)
        : rb_(rb), max_level_(max_level), n_cell_(n_cell), is_periodic_(is_per), amr_c_(amr_c) {
        if ((int)n_cell.size() < SPACEDIM) throw AdgError("Mesh: n_cell needs one entry per dimension");
        if (coord != 0) throw AdgError("Mesh: only Cartesian coordinates (coord = 0)");
        if (nghost != 1) throw AdgError("Mesh: the DG scheme uses exactly one ghost cell");
        if ((int)ref_ratios.size() < max_level) throw AdgError("Mesh: one refinement ratio per level");
        for (int l = 0; l < max_level; ++l)
            for (int d = 0; d < SPACEDIM; ++d)
                if (ref_ratios[l][d] != 2) throw AdgError("Mesh: refinement ratio must be 2");
        if (max_level > 0 && (int)amr_c.size() < max_level) throw AdgError("Mesh: one amr_c per coarse level");
        this->dtn_regrid = dtn_regrid;
        this->dt_regrid = dt_regrid;
        max_grid_size_ = max_grid_size;
        // AmrCore::InitFromScratch: every level 0..max_level as a full-domain
        // grid (Solver.h:566-569); the first step's regrid replaces the fine ones
        for (int l = 0; l <= max_level; ++l) {
            std::array<int, 3> lo{{0, 0, 0}};
            std::array<int, 3> hi{{(n_cell[0] << l) - 1, (n_cell[1] << l) - 1,
                                   SPACEDIM == 3 ? (n_cell[2] << l) - 1 : 0}};
            levels_.push_back(tile_boxes(lo, hi, max_grid_size));
        }
        if (max_grid_size % 4) throw AdgError("Mesh: max_grid_size must be a multiple of the blocking factor 4");
    }

    // Regrid step interval and physical-time interval (Core/Mesh.h:107-111).
    int dtn_regrid = 0;
    double dt_regrid = 0;
    int get_max_grid_size() const { return max_grid_size_; }
    // The layout after a regrid (boxes per level, level 0 unchanged).
    void set_layout(const std::vector<std::vector<Box>>& levels) { levels_ = levels; }

    // Fine levels 1..max_level, each a list of boxes in its own index space;
    // a hierarchy given this way stays as given (no regrid).
    void set_levels(const std::vector<std::vector<Box>>& fine) {
        if ((int)fine.size() != max_level_) throw AdgError("Mesh::set_levels: one box list per fine level");
        levels_.resize(1);
        for (const auto& l : fine) levels_.push_back(l);
        static_ = true;
    }
    bool is_static() const { return static_; }
    // Level 1 over the central half of the domain (the layout of the
    // reference's vortex run after its first regrid).
    void set_centre_level(int max_grid_size = 32) {
        if (max_level_ != 1) throw AdgError("Mesh::set_centre_level needs max_level = 1");
        std::array<int, 3> lo{}, hi{};
        for (int d = 0; d < 3; ++d) {
            const int n = d < SPACEDIM ? n_cell_[d] : 0;
            lo[d] = d < SPACEDIM ? n / 2 : 0;
            hi[d] = d < SPACEDIM ? n + n / 2 - 1 : 0;
        }
        levels_.resize(1);
        levels_.push_back(tile_boxes(lo, hi, max_grid_size));
        static_ = true;
    }

    int get_finest_lev() const { return (int)levels_.size() - 1; }
    int get_max_level() const { return max_level_; }
    const RealBox& get_RealBox() const { return rb_; }
    const std::vector<int>& get_n_cell() const { return n_cell_; }
    const std::array<int, 3>& get_is_periodic() const { return is_periodic_; }
    const std::vector<Box>& get_BoxArray(int l) const { return levels_.at(l); }
    const std::vector<double>& get_amr_c() const { return amr_c_; }

  private:
    RealBox rb_;
    int max_level_;
    std::vector<int> n_cell_;
    std::array<int, 3> is_periodic_;
    std::vector<double> amr_c_;
    std::vector<std::vector<Box>> levels_;
    int max_grid_size_ = 32;
    bool static_ = false;
};

// ---------------------------------------------------------------------------
// BoundaryCondition<EQ,NM> (Core/BoundaryCondition.h:20-100): per-dimension
// types 0 = Dirichlet, 1 = Neumann, 2 = periodic.  The device fills
// non-periodic ghosts transmissively (zero-gradient mirror image), which is
// the Neumann type; the reference's Dirichlet values are a stub
// (Compressible_Euler.h:58-72) and are refused.
template <typename EquationType, typename NumericalMethodType>
class BoundaryCondition {
  public:
    void settings(const std::array<int, 3>& bc_lo_type, const std::array<int, 3>& bc_hi_type,
                  const std::array<int, 3>& /*bc_lo AMReX type*/ = {}, const std::array<int, 3>& /*bc_hi*/ = {}) {
        for (int d = 0; d < SPACEDIM; ++d) {
            if (bc_lo_type[d] != bc_hi_type[d])
                throw AdgError("BoundaryCondition: lo and hi types of a dimension must agree");
            if (bc_lo_type[d] == 0) throw AdgError("BoundaryCondition: Dirichlet data is not part of this library");
            if (bc_lo_type[d] != 1 && bc_lo_type[d] != 2) throw AdgError("BoundaryCondition: unknown type");
            periodic_[d] = bc_lo_type[d] == 2 ? 1 : 0;
        }
        set_ = true;
    }
    void check(const std::array<int, 3>& is_periodic) const {
        if (!set_) return;
        for (int d = 0; d < SPACEDIM; ++d)
            if (periodic_[d] != (is_periodic[d] ? 1 : 0))
                throw AdgError("BoundaryCondition: periodic type disagrees with the mesh periodicity");
    }

  private:
    std::array<int, 3> periodic_{{1, 1, 1}};
    bool set_ = false;
};

// ---------------------------------------------------------------------------
// Solver<NM>: CRTP front of the numerical method (Core/Solver.h:140-200).
template <typename NumericalMethodType>
class Solver {
  public:
    Solver() = default;
    virtual ~Solver() = default;

    template <typename... Args>
    void settings(Args... args) {
        static_cast<NumericalMethodType*>(this)->settings(args...);
    }
    void setValidationMode(bool use_analytical_ic) { flag_analytical_ic = use_analytical_ic; }

    template <typename EquationType>
    void init(const std::shared_ptr<ModelEquation<EquationType>>& model_pde,
              std::shared_ptr<Mesh<NumericalMethodType>> mesh, int dtn_outplt, double dt_outplt,
              std::string data_dir, int restart_tstep = 0) {
        static_cast<NumericalMethodType*>(this)->init(model_pde, mesh, dtn_outplt, dt_outplt, data_dir, restart_tstep);
    }
    template <typename EquationType, typename BC>
    void evolve(const std::shared_ptr<ModelEquation<EquationType>>& model_pde, const std::shared_ptr<BC>& bdcond) {
        static_cast<NumericalMethodType*>(this)->evolve(model_pde, bdcond);
    }
    template <typename EquationType, typename BC>
    void time_integration(const std::shared_ptr<ModelEquation<EquationType>>& model_pde,
                          const std::shared_ptr<BC>& bdcond, double time) {
        static_cast<NumericalMethodType*>(this)->time_integration(model_pde, bdcond, time);
    }
    template <typename EquationType>
    void set_Dt(const std::shared_ptr<ModelEquation<EquationType>>& model_pde) {
        static_cast<NumericalMethodType*>(this)->set_Dt(model_pde);
    }
    void AMR_average_fine_coarse() { static_cast<NumericalMethodType*>(this)->AMR_average_fine_coarse(); }
    void AMR_FillPatch() { static_cast<NumericalMethodType*>(this)->AMR_FillPatch(); }

  protected:
    bool flag_analytical_ic = true;
    double T = 0.0, c_dt = 0.0, Dt = 0.0, t = 0.0;
    int n = 0;
};

// ---------------------------------------------------------------------------
// AmrDG<P>: the ADER-DG method (NumericalMethod/AmrDG/AmrDG.h:400-445,
// evolve :840-1060, ADER :1071-1165, set_Dt :1253-1357, LpNorm :1399-1578).
template <int P>
class AmrDG : public Solver<AmrDG<P>> {
    using Base = Solver<AmrDG<P>>;

  public:
    static_assert(P >= 1 && P <= 3, "the library is built for p = 1, 2, 3");
    using NumericalMethodType = AmrDG<P>;

    AmrDG() = default;
    ~AmrDG() override {
        if (ctx_) adg_destroy(ctx_);
    }
    AmrDG(const AmrDG&) = delete;
    AmrDG& operator=(const AmrDG&) = delete;

    // AmrDG.h:407-410 (AmrDG.tpp:50-58): final time, CFL factor and the
    // limiter ("TVB" with M, AMR_TVB_C per level and t_limit; "" or
    // t_limit <= 0: none).
    void settings(double T_, double c_dt_, const std::string& limiter_type = "", double TVB_M = 0.0,
                  const std::vector<double>& AMR_TVB_C = {}, int t_limit = -1) {
        if (!limiter_type.empty() && limiter_type != "TVB")
            throw AdgError("AmrDG::settings: unknown limiter '" + limiter_type + "'");
        limiter_on_ = limiter_type == "TVB" && t_limit > 0;
        TVB_M_ = TVB_M;
        AMR_TVB_C_ = AMR_TVB_C;
        t_limit_ = t_limit;
        this->T = T_;
        this->c_dt = c_dt_;
    }
    // Apply the limiter once to the projected initial state (not done by
    // the reference; needed for discontinuous initial states at p >= 2).
    void setLimitInitialState(bool on) { limit_ic_ = on; }
    // Reflux every uncovered coarse cell the flux registers reach (not the
    // reference's rule, which skips a cell whose coarse-fine face is the low
    // face of its box and so loses mass there; AmrDG_Interpolater.tpp:39-69).
    void setConservativeReflux(bool on) { conservative_reflux_ = on; }
    // Time step from the uncovered cells only (the reference's set_Dt also
    // reads covered coarse cells, whose projected data can have negative
    // nodal density across a discontinuity and collapse Dt).
    void setCflUncovered(bool on) { cfl_uncovered_ = on; }

    // Device forms: ADG_FP64 (plain) or ADG_FP32_MFMA (precision form).
    void setPrecision(int precision) { precision_ = precision; }
    // Learned warm start from a trained network (adg_mlp_get_params layout).
    void setLearnedWarmStart(const std::string& weights, double tol, int max_iters) {
        weights_ = weights;
        learned_tol_ = tol;
        learned_iters_ = max_iters;
    }
    // Steps per batch between regrid and plot points (0: one adg_step per
    // step); a batch runs as replays of the captured step graph (16 steps
    // per graph, adg_step_graph) or, with setBatchGraph(false), as stream
    // launches (adg_steps).
    void setGraphSteps(int max_batch) { graph_batch_ = max_batch; }
    void setBatchGraph(bool on) { batch_graph_ = on; }
    // Launch-shape autotuning after the hierarchy is set (0: off).
    void setAutotune(int reps) { autotune_reps_ = reps; }
    // Print the t=0 and t=T norms as the reference does.
    void setVerbose(bool v) { verbose_ = v; }

    template <typename EquationType>
    void init(const std::shared_ptr<ModelEquation<EquationType>>& model_pde, std::shared_ptr<Mesh<AmrDG<P>>> mesh,
              int dtn_outplt, double dt_outplt, const std::string& data_dir, int restart_tstep = 0) {
        mesh_ = mesh;
        dtn_outplt_ = dtn_outplt;
        dt_outplt_ = dt_outplt;
        data_dir_ = data_dir;
        restart_tstep_ = restart_tstep;
        if (ctx_) {
            adg_destroy(ctx_);
            ctx_ = nullptr;
        }
        chk(adg_create(&ctx_, SPACEDIM, P, model_pde->gamma_adiab, precision_, nullptr), "adg_create");
        chk(adg_set_graph_unroll(ctx_, 16), "adg_set_graph_unroll");   // 16 steps per captured graph
        Q = SPACEDIM + 2;
        const RealBox& rb = mesh->get_RealBox();
        int dom[3] = {0, 0, 0}, per[3] = {0, 0, 0};
        double lo[3], hi[3];
        for (int d = 0; d < SPACEDIM; ++d) {
            dom[d] = mesh->get_n_cell()[d];
            per[d] = mesh->get_is_periodic()[d];
            lo[d] = rb.lo[d];
            hi[d] = rb.hi[d];
        }
        chk(adg_set_geometry(ctx_, dom, lo, hi, per), "adg_set_geometry");
        chk(adg_set_conservative_reflux(ctx_, conservative_reflux_ ? 1 : 0), "adg_set_conservative_reflux");
        chk(adg_set_cfl_uncovered(ctx_, cfl_uncovered_ ? 1 : 0), "adg_set_cfl_uncovered");
        const int nlev = mesh->get_finest_lev() + 1;
        if (nlev != mesh->get_max_level() + 1)
            throw AdgError("AmrDG::init: give the fine boxes of every level to Mesh::set_levels");
        std::vector<int> np, boxes;
        for (int l = 0; l < nlev; ++l) {
            const auto& ba = mesh->get_BoxArray(l);
            np.push_back((int)ba.size());
            for (const Box& b : ba) {
                for (int d = 0; d < SPACEDIM; ++d) boxes.push_back(b.lo[d]);
                for (int d = 0; d < SPACEDIM; ++d) boxes.push_back(b.hi[d]);
            }
        }
        // Solver::set_initial_condition (Core/Solver.h:711-735): the analytic
        // initial state on level 0; on the finer levels the analytic state
        // too in validation mode, else the projection of the coarser level
        // (AMR_interpolate_initial_condition, AmrDG_Level.tpp:82-86); then
        // average down and FillPatch (AMR_sync_initial_condition)
        int ic = 0;
        std::vector<double> prm;
        model_pde->pde_IC_case(rb, &ic, &prm);
        if (restart_tstep > 0) {
            // Solver::init with restart (Core/Solver.h:571-591): the hierarchy
            // and U_w from the plotfiles of step restart_tstep, then FillPatch
            restart_time_ = ReadPlotFile(restart_tstep);
        } else if (this->flag_analytical_ic || nlev == 1) {
            chk(adg_set_levels(ctx_, nlev, np.data(), boxes.data()), "adg_set_levels");
            for (int l = 0; l < nlev; ++l) chk(adg_project_ic(ctx_, l, ic, prm.data()), "adg_project_ic");
        } else {
            chk(adg_set_levels(ctx_, 1, np.data(), boxes.data()), "adg_set_levels");
            chk(adg_project_ic(ctx_, 0, ic, prm.data()), "adg_project_ic");
            AMR_FillPatch();
            chk(adg_remake(ctx_, nlev, np.data(), boxes.data()), "adg_remake");
        }
        if (restart_tstep <= 0) AMR_average_fine_coarse();
        AMR_FillPatch();
        // the limiter (AmrDG::settings) and the regrid (Mesh) settings
        {
            std::vector<double> tvb = AMR_TVB_C_;
            if (tvb.empty()) tvb.assign((size_t)mesh->get_max_level() + 1, 1.0);
            chk(adg_set_limiter(ctx_, limiter_on_ ? 1 : 0, TVB_M_, (int)tvb.size(), tvb.data(),
                                t_limit_ > 0 ? t_limit_ : 1), "adg_set_limiter");
            if (limit_ic_ && restart_tstep <= 0) {
                chk(adg_limit(ctx_, nullptr), "adg_limit");
                AMR_average_fine_coarse();
                AMR_FillPatch();
            }
        }
        dynamic_ = mesh->get_max_level() > 0 && !mesh->is_static();
        if (dynamic_) {
            chk(adg_set_amr(ctx_, mesh->get_max_level(), 4, mesh->get_max_grid_size(), 0.7, 1, 1), "adg_set_amr");
            const std::string& c = model_pde->model_case;
            // Compressible_Euler::pde_tag_cell_refinement (Compressible_Euler.h:605-682)
            const int crit = c == "isentropic_vortex" ? 0 : c == "kelvin_helmolz_instability" ? 1 : 2;
            const double centre[2] = {5.0, 5.0};
            const std::vector<double>& ac = mesh->get_amr_c();
            chk(adg_set_tagging(ctx_, crit, (int)ac.size(), ac.data(), crit == 0 ? 2 : 0, centre), "adg_set_tagging");
        }
        if (autotune_reps_ > 0) chk(adg_autotune(ctx_, autotune_reps_, nullptr, nullptr), "adg_autotune");
        if (!weights_.empty()) load_network();
    }

    // AmrDG::evolve (AmrDG.h:853-1055): regrid at the first step and at the
    // regrid points (every dtn_regrid steps or dt_regrid of time,
    // AmrDG.h:936-986), ADER, the limiter on the finest level, average down,
    // FillPatch, t += Dt, set_Dt and the clamp to T.  Runs of steps between
    // regrid points run in batches (adg_steps or adg_step_graph).
    template <typename EquationType, typename BC>
    void evolve(const std::shared_ptr<ModelEquation<EquationType>>& model_pde, const std::shared_ptr<BC>& bdcond) {
        n_regrids_ = 0;
        double t_last_regrid = 0.0, t_last_plt = 0.0;
        bool first_step = true;
        if (restart_tstep_ > 0) {
            this->n = restart_tstep_;
            this->t = restart_time_;
            t_last_plt = this->t;
            t_last_regrid = this->t;
        } else {
            this->n = 0;
            this->t = 0.0;
            if (dtn_outplt_ > 0 || dt_outplt_ > 0) PlotFile(this->n, this->t);
            if (verbose_) {
                L1Norm_DG_AMR(model_pde);
                L2Norm_DG_AMR(model_pde);
            }
        }
        set_Dt(model_pde);
        chk(adg_limiter_stats(ctx_, nullptr, 1), "adg_limiter_stats");
        chk(adg_set_step_count(ctx_, this->n), "adg_set_step_count");
        const int dtn = mesh_->dtn_regrid;
        const double dtr = mesh_->dt_regrid;
        while (this->t < this->T && (max_steps_ < 0 || this->n < max_steps_ + restart_steps_offset())) {
            if (dynamic_) {
                const bool due = first_step || (dtn > 0 && this->n % dtn == 0) ||
                                 (dtr > 0 && this->t - t_last_regrid >= dtr - 1e-12);
                if (due) {
                    int changed = 0;
                    chk(adg_regrid(ctx_, this->t, &changed), "adg_regrid");
                    if (changed || first_step) ++n_regrids_;
                    if (changed) sync_mesh_layout();
                    t_last_regrid = this->t;
                }
                first_step = false;
            }
            int batch = 1;
            if (graph_batch_ > 0 && this->Dt > 0.0) {
                // stay clear of the final clamped steps and of the next regrid
                // point, which run one at a time
                double room = 0.5 * (this->T - this->t) / this->Dt;
                if (dynamic_ && dtr > 0) room = std::min(room, 0.5 * (t_last_regrid + dtr - this->t) / this->Dt);
                batch = (int)std::min<double>(graph_batch_, std::floor(room));
                if (dynamic_ && dtn > 0) batch = std::min(batch, dtn - this->n % dtn);
                // and of the plot points
                if (dtn_outplt_ > 0) batch = std::min(batch, dtn_outplt_ - this->n % dtn_outplt_);
                if (dt_outplt_ > 0)
                    batch = std::min<int>(batch, (int)std::floor(0.5 * (t_last_plt + dt_outplt_ - this->t) / this->Dt));
                if (max_steps_ >= 0) batch = std::min(batch, max_steps_ - this->n + restart_steps_offset());
            }
            if (batch >= 2) {
                if (batch_graph_) chk(adg_step_graph(ctx_, this->c_dt, this->T, batch, nullptr), "adg_step_graph");
                else chk(adg_steps(ctx_, this->c_dt, this->T, batch, nullptr), "adg_steps");
                this->n += batch;
            } else {
                time_integration(model_pde, bdcond, this->t);
                if (limiter_on_ && this->n % t_limit_ == 0) chk(adg_limit(ctx_, nullptr), "adg_limit");
                AMR_average_fine_coarse();
                AMR_FillPatch();
                chk(adg_advance(ctx_, this->c_dt, this->T, nullptr), "adg_advance");   // t += Dt; set_Dt
                this->n += 1;
            }
            chk(adg_get_time(ctx_, &this->t, &this->Dt), "adg_get_time");
            // plot points (AmrDG.h:1018-1023)
            const bool dtn_plt = dtn_outplt_ > 0 && this->n % dtn_outplt_ == 0;
            const bool dt_plt = dt_outplt_ > 0 && this->t - t_last_plt >= dt_outplt_ - 1e-12;
            if (dtn_plt) {
                PlotFile(this->n, this->t);
            } else if (dt_plt) {
                PlotFile(this->n, this->t);
                t_last_plt = this->t;
            }
        }
        if (verbose_) {
            std::printf("Total number of time steps: %d\n", this->n);
            std::printf("Total number of regrids: %d\n", n_regrids_);
            std::printf("Total number of cells limited: %llu\n", cells_limited());
            L1Norm_DG_AMR(model_pde);
            L2Norm_DG_AMR(model_pde);
        }
    }

    // PlotFile (Core/Solver.h:922-968): one plotfile per conserved variable,
    // data_dir/tstep_<n>_q_<q>_plt, components <name>_<mode>, the valid cells
    // of every level.
    void PlotFile(int tstep, double time) {
        static const char* names[5] = {"mass_density", "momentum_x", "momentum_y", "energy_density",
                                       "angular_momentum_z"};   // Compressible_Euler.h:15-24
        int sz[7];
        chk(adg_sizes(ctx_, sz), "adg_sizes");
        const int Ns = sz[0];
        const int nlev = mesh_->get_finest_lev() + 1;
        std::vector<std::vector<double>> U((size_t)nlev);
        for (int l = 0; l < nlev; ++l) U[(size_t)l] = get_U_w(l);
        const RealBox& rb = mesh_->get_RealBox();
        for (int q = 0; q < Q; ++q) {
            plt::Plotfile pf;
            pf.dim = SPACEDIM;
            pf.time = time;
            pf.step = tstep;
            for (int m = 0; m < Ns; ++m) pf.names.push_back(std::string(names[q]) + "_" + std::to_string(m));
            for (int d = 0; d < SPACEDIM; ++d) { pf.prob_lo[d] = rb.lo[d]; pf.prob_hi[d] = rb.hi[d]; }
            for (int l = 0; l < nlev; ++l) {
                plt::Level L;
                for (int d = 0; d < SPACEDIM; ++d) {
                    const int n = mesh_->get_n_cell()[d] << l;
                    L.domain_hi[d] = n - 1;
                    L.dx[d] = (rb.hi[d] - rb.lo[d]) / n;
                }
                const std::vector<double>& u = U[(size_t)l];
                const long ns = (long)u.size() / ((long)Q * Ns);
                long off = 0;
                for (const Box& b : mesh_->get_BoxArray(l)) {
                    plt::IBox ib;
                    long g[3] = {1, 1, 1};
                    for (int d = 0; d < SPACEDIM; ++d) {
                        ib.lo[d] = b.lo[d];
                        ib.hi[d] = b.hi[d];
                        g[d] = b.hi[d] - b.lo[d] + 3;
                    }
                    const long np = plt::box_cells(ib, SPACEDIM);
                    std::vector<double> v((size_t)Ns * np);
                    long k = 0;
                    for (int z = 0; z < (SPACEDIM == 3 ? b.hi[2] - b.lo[2] + 1 : 1); ++z)
                        for (int y = 0; y <= b.hi[1] - b.lo[1]; ++y)
                            for (int x = 0; x <= b.hi[0] - b.lo[0]; ++x, ++k) {
                                const long s = off + (x + 1) + g[0] * ((y + 1) + g[1] * (SPACEDIM == 3 ? z + 1 : 0));
                                for (int m = 0; m < Ns; ++m) v[(size_t)m * np + k] = u[((size_t)q * Ns + m) * ns + s];
                            }
                    off += g[0] * g[1] * g[2];
                    L.boxes.push_back(ib);
                    L.data.push_back(std::move(v));
                }
                pf.levels.push_back(std::move(L));
            }
            plt::write(data_dir_ + "/tstep_" + std::to_string(tstep) + "_q_" + std::to_string(q) + "_plt", pf);
        }
    }

    // ReadPlotFile (Core/Solver.h:599-652): the level layout (at most
    // max_level) from the q = 0 plotfile of step tstep, U_w of every level
    // from the plotfile of each q; returns the time stored there.
    double ReadPlotFile(int tstep) {
        const auto path = [&](int q) {
            return data_dir_ + "/tstep_" + std::to_string(tstep) + "_q_" + std::to_string(q) + "_plt";
        };
        plt::Plotfile meta;
        try {
            meta = plt::read(path(0));
        } catch (const std::exception& e) {
            throw AdgError(std::string("AmrDG::ReadPlotFile: ") + e.what());
        }
        const int nlev = std::min((int)meta.levels.size(), mesh_->get_max_level() + 1);
        std::vector<std::vector<Box>> lv((size_t)nlev);
        std::vector<int> np, boxes;
        for (int l = 0; l < nlev; ++l) {
            for (const plt::IBox& ib : meta.levels[(size_t)l].boxes) {
                Box b;
                for (int d = 0; d < SPACEDIM; ++d) { b.lo[d] = ib.lo[d]; b.hi[d] = ib.hi[d]; }
                lv[(size_t)l].push_back(b);
                for (int d = 0; d < SPACEDIM; ++d) boxes.push_back(b.lo[d]);
                for (int d = 0; d < SPACEDIM; ++d) boxes.push_back(b.hi[d]);
            }
            np.push_back((int)lv[(size_t)l].size());
        }
        mesh_->set_layout(lv);
        chk(adg_set_levels(ctx_, nlev, np.data(), boxes.data()), "adg_set_levels");
        int sz[7];
        chk(adg_sizes(ctx_, sz), "adg_sizes");
        const int Ns = sz[0];
        std::vector<std::vector<double>> U((size_t)nlev);
        for (int l = 0; l < nlev; ++l) {
            const int ns = adg_level_nslot(ctx_, l);
            U[(size_t)l].assign((size_t)Q * Ns * ns, 0.0);
        }
        for (int q = 0; q < Q; ++q) {
            plt::Plotfile pf;
            try {
                pf = q == 0 ? meta : plt::read(path(q));
            } catch (const std::exception& e) {
                throw AdgError(std::string("AmrDG::ReadPlotFile: ") + e.what());
            }
            for (int l = 0; l < nlev; ++l) {
                std::vector<double>& u = U[(size_t)l];
                const long ns = (long)u.size() / ((long)Q * Ns);
                long off = 0;
                const plt::Level& L = pf.levels[(size_t)l];
                for (size_t bi = 0; bi < L.boxes.size(); ++bi) {
                    const plt::IBox& b = L.boxes[bi];
                    long g[3] = {1, 1, 1};
                    for (int d = 0; d < SPACEDIM; ++d) g[d] = b.hi[d] - b.lo[d] + 3;
                    const long npt = plt::box_cells(b, SPACEDIM);
                    const std::vector<double>& v = L.data[bi];
                    long k = 0;
                    for (int z = 0; z < (SPACEDIM == 3 ? b.hi[2] - b.lo[2] + 1 : 1); ++z)
                        for (int y = 0; y <= b.hi[1] - b.lo[1]; ++y)
                            for (int x = 0; x <= b.hi[0] - b.lo[0]; ++x, ++k) {
                                const long s = off + (x + 1) + g[0] * ((y + 1) + g[1] * (SPACEDIM == 3 ? z + 1 : 0));
                                for (int m = 0; m < Ns; ++m) u[((size_t)q * Ns + m) * ns + s] = v[(size_t)m * npt + k];
                            }
                    off += g[0] * g[1] * g[2];
                }
            }
        }
        for (int l = 0; l < nlev; ++l) chk(adg_upload_Uw(ctx_, l, U[(size_t)l].data()), "adg_upload_Uw");
        return meta.time;
    }

    // Cells limited since the start of evolve (n_limited_total, AmrDG.h:864).
    unsigned long long cells_limited() {
        unsigned long long nl = 0;
        chk(adg_limiter_stats(ctx_, &nl, 0), "adg_limiter_stats");
        return nl;
    }
    int get_n_regrids() const { return n_regrids_; }

    template <typename EquationType, typename BC>
    void time_integration(const std::shared_ptr<ModelEquation<EquationType>>& model_pde,
                          const std::shared_ptr<BC>& bdcond, double time) {
        ADER(model_pde, bdcond, time);
    }

    // AmrDG::ADER (AmrDG.h:1071-1165): predictor, face flux, corrector and
    // reflux of every level.
    template <typename EquationType, typename BC>
    void ADER(const std::shared_ptr<ModelEquation<EquationType>>&, const std::shared_ptr<BC>&, double) {
        chk(adg_ader(ctx_), "adg_ader");
    }

    // AmrDG::set_Dt (AmrDG.h:1253-1357) with the clamp Dt = min(Dt, T - t).
    template <typename EquationType>
    void set_Dt(const std::shared_ptr<ModelEquation<EquationType>>&) {
        chk(adg_set_dt(ctx_, this->c_dt, this->T, this->t, &this->Dt), "adg_set_dt");
    }

    void AMR_average_fine_coarse() { chk(adg_average_down(ctx_), "adg_average_down"); }
    void AMR_FillPatch() { chk(adg_fill_ghosts(ctx_), "adg_fill_ghosts"); }

    // LpNorm_DG_AMR (AmrDG.h:1399-1578) against the exact solution at t;
    // returns one value per conserved variable (empty without one).
    template <typename EquationType>
    std::vector<double> LpNorm_DG_AMR(const std::shared_ptr<ModelEquation<EquationType>>& model_pde, int p) {
        int ic = 0;
        std::vector<double> prm;
        std::vector<double> out;
        if (!model_pde->pde_exact_solution(&ic, &prm, mesh_->get_RealBox())) return out;
        for (int q = 0; q < Q; ++q) {
            double v = 0.0;
            chk(adg_lp_error(ctx_, q, p, ic, prm.data(), this->t, &v), "adg_lp_error");
            out.push_back(v);
            if (verbose_) {
                std::printf("--multilevel--\n");
                std::printf("L%d error norm:  %.17g | DG Order:  %d | solution component: %d\n", p, v, P + 1, q);
            }
        }
        return out;
    }
    template <typename EquationType>
    std::vector<double> L1Norm_DG_AMR(const std::shared_ptr<ModelEquation<EquationType>>& m) {
        return LpNorm_DG_AMR(m, 1);
    }
    template <typename EquationType>
    std::vector<double> L2Norm_DG_AMR(const std::shared_ptr<ModelEquation<EquationType>>& m) {
        return LpNorm_DG_AMR(m, 2);
    }

    // Sum of vol * U_w[q][0] over uncovered valid cells, FP64.
    std::vector<double> conserved_totals() {
        std::vector<double> s(Q);
        chk(adg_conserved_totals(ctx_, s.data()), "adg_conserved_totals");
        return s;
    }
    // U_w of a level, [q][n][slot] (ghost slots included).
    std::vector<double> get_U_w(int lev) {
        const int ns = adg_level_nslot(ctx_, lev);
        if (ns < 0) chk(ns, "adg_level_nslot");
        int sz[7];
        chk(adg_sizes(ctx_, sz), "adg_sizes");
        std::vector<double> h((size_t)Q * sz[0] * ns);
        chk(adg_download_Uw(ctx_, lev, h.data()), "adg_download_Uw");
        return h;
    }
    unsigned long long learned_fallbacks() {
        unsigned long long st[2] = {0, 0};
        chk(adg_learned_stats(ctx_, st, 0), "adg_learned_stats");
        return st[1];
    }

    void setMaxSteps(int m) { max_steps_ = m; }
    int get_n() const { return this->n; }
    double get_t() const { return this->t; }
    double get_Dt() const { return this->Dt; }
    int get_Q() const { return Q; }
    adg_ctx* get_ctx() const { return ctx_; }

  private:
    // max_steps counts the steps of this run, also after a restart
    int restart_steps_offset() const { return restart_tstep_ > 0 ? restart_tstep_ : 0; }
    // The Mesh's boxes after a regrid, from the library's layout.
    void sync_mesh_layout() {
        int nl = 0, np[16] = {0};
        const int nb = adg_get_levels(ctx_, &nl, np, nullptr, 0);
        if (nb < 0) chk(nb, "adg_get_levels");
        std::vector<int> bx((size_t)nb * 2 * SPACEDIM);
        chk(adg_get_levels(ctx_, &nl, np, bx.data(), nb) < 0 ? ADG_E_STATE : ADG_OK, "adg_get_levels");
        std::vector<std::vector<Box>> lv((size_t)nl);
        int k = 0;
        for (int l = 0; l < nl; ++l)
            for (int i = 0; i < np[l]; ++i, ++k) {
                Box b;
                for (int d = 0; d < SPACEDIM; ++d) {
                    b.lo[d] = bx[(size_t)k * 2 * SPACEDIM + d];
                    b.hi[d] = bx[(size_t)k * 2 * SPACEDIM + SPACEDIM + d];
                }
                lv[(size_t)l].push_back(b);
            }
        mesh_->set_layout(lv);
    }
    void chk(int rc, const char* what) const {
        if (rc != ADG_OK)
            throw AdgError(std::string(what) + " failed (" + std::to_string(rc) + "): " +
                           (ctx_ ? adg_last_error(ctx_) : "no context"));
    }
    void load_network() {
        std::ifstream f(weights_, std::ios::binary | std::ios::ate);
        if (!f) throw AdgError("AmrDG: cannot open weights " + weights_);
        const std::streamsize nb = f.tellg();
        f.seekg(0);
        std::vector<float> w((size_t)nb / sizeof(float));
        f.read(reinterpret_cast<char*>(w.data()), (std::streamsize)(w.size() * sizeof(float)));
        chk(adg_mlp_set_params(ctx_, w.data(), (long)w.size()), "adg_mlp_set_params");
        chk(adg_set_learned(ctx_, 1, learned_tol_, learned_iters_), "adg_set_learned");
    }

    adg_ctx* ctx_ = nullptr;
    std::shared_ptr<Mesh<AmrDG<P>>> mesh_;
    int Q = SPACEDIM + 2;
    int precision_ = ADG_FP64;
    std::string weights_;
    double learned_tol_ = 1e-4;
    int learned_iters_ = 2;
    int graph_batch_ = 64;
    bool batch_graph_ = true;
    int autotune_reps_ = 0;
    int max_steps_ = -1;
    bool verbose_ = true;
    bool limiter_on_ = false, limit_ic_ = false, dynamic_ = false, conservative_reflux_ = false;
    bool cfl_uncovered_ = false;
    double TVB_M_ = 0.0;
    std::vector<double> AMR_TVB_C_;
    int t_limit_ = -1;
    int n_regrids_ = 0;
    int dtn_outplt_ = -1, restart_tstep_ = 0;
    double dt_outplt_ = -1.0, restart_time_ = 0.0;
    std::string data_dir_;
};

// ---------------------------------------------------------------------------
// Simulation<NM,EQ> (Core/Simulation.h:25-200).
template <typename NumericalMethodType, typename EquationType>
class Simulation {
  public:
    Simulation() = default;

    // Simulation::run (Core/Simulation.h:89-118): the output directory is
    // "../Results/Simulation Data/<case>[/<sim_run>]" unless setDataDir gave
    // one; a fresh run with output removes the plotfiles of earlier runs there
    // (tstep_*_plt directories; the reference clears the whole directory).
    void run() {
        if (!model || !solver || !mesh) throw AdgError("Simulation::run: model, numerical and geometry settings first");
        if (bdcond) bdcond->check(mesh->get_is_periodic());
        std::string dir = data_dir_;
        if (dir.empty()) {
            dir = "../Results/Simulation Data/" + model->model_case;
            if (!sim_run.empty()) dir += "/" + sim_run;
        }
        if (dtn_outplt > 0 || dt_outplt > 0 || restart_tstep > 0) {
            plt::mkdir_p(dir);
            if (restart_tstep == 0) remove_plotfiles(dir);
        }
        solver->init(std::static_pointer_cast<ModelEquation<EquationType>>(model), mesh, dtn_outplt, dt_outplt,
                     dir, restart_tstep);
        solver->evolve(std::static_pointer_cast<ModelEquation<EquationType>>(model), bdcond);
    }

    template <typename... Args>
    void setNumericalSettings(Args... args) {
        solver = std::make_shared<NumericalMethodType>();
        solver->settings(args...);
        solver->setValidationMode(validation_);
    }
    template <typename... Args>
    void setModelSettings(Args... args) {
        model = std::make_shared<EquationType>();
        model->settings(args...);
        model->set_pde_numeric_limits();
    }
    template <typename... Args>
    void setBoundaryConditions(Args... args) {
        bdcond = std::make_shared<BoundaryCondition<EquationType, NumericalMethodType>>();
        bdcond->settings(args...);
    }
    void setGeometrySettings(const RealBox& rb, int max_level, const std::vector<int>& n_cell, int coord,
                             const std::vector<std::array<int, 3>>& ref_ratios, const std::array<int, 3>& is_per,
                             const std::vector<double>& amr_c, int dtn_regrid = 0, double dt_regrid = 0,
                             int nghost = 1) {
        mesh = std::make_shared<Mesh<NumericalMethodType>>(rb, max_level, n_cell, coord, ref_ratios, is_per, amr_c,
                                                           dtn_regrid, dt_regrid, nghost);
    }
    void setIO(int n_out, double t_out, std::string run = "", int restart = 0) {
        dtn_outplt = n_out;
        dt_outplt = t_out;
        sim_run = run;
        restart_tstep = restart;
    }
    void setValidationMode(bool use_analytical_ic) {
        validation_ = use_analytical_ic;
        if (solver) solver->setValidationMode(use_analytical_ic);
    }
    int getQ() { return model ? model->Q_model : 0; }
    // Output / restart directory in place of the reference's fixed path.
    void setDataDir(const std::string& dir) { data_dir_ = dir; }

    std::shared_ptr<EquationType> model;
    std::shared_ptr<NumericalMethodType> solver;
    std::shared_ptr<Mesh<NumericalMethodType>> mesh;
    std::shared_ptr<BoundaryCondition<EquationType, NumericalMethodType>> bdcond;

  private:
    static void remove_plotfiles(const std::string& dir) {
        DIR* d = ::opendir(dir.c_str());
        if (!d) return;
        std::vector<std::string> doomed;
        while (dirent* e = ::readdir(d)) {
            const std::string n = e->d_name;
            if (n.rfind("tstep_", 0) == 0 && n.size() > 4 && n.compare(n.size() - 4, 4, "_plt") == 0)
                doomed.push_back(dir + "/" + n);
        }
        ::closedir(d);
        for (const std::string& p : doomed) remove_tree(p);
    }
    static void remove_tree(const std::string& p) {
        if (DIR* d = ::opendir(p.c_str())) {
            while (dirent* e = ::readdir(d)) {
                const std::string n = e->d_name;
                if (n != "." && n != "..") remove_tree(p + "/" + n);
            }
            ::closedir(d);
            ::rmdir(p.c_str());
        } else {
            std::remove(p.c_str());
        }
    }

    int dtn_outplt = -1;
    double dt_outplt = -1;
    std::string sim_run;
    int restart_tstep = 0;
    bool validation_ = true;
    std::string data_dir_;
};

}  // namespace adg

#endif
