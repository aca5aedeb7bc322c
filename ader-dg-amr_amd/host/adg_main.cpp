// adg_main.cpp -- driver in the shape of the reference's Source/main.cpp
// (settings block, case overrides, Simulation<AmrDG<p>,Compressible_Euler>
// construction, run), over the C++ host side in AmrDG.hpp.
//
// Usage: adg_main_2d [--case NAME] [--p 1|2|3] [--n CELLS] [--max_level L]
//                    [--T T] [--c_dt C] [--precision fp64|fp32]
//                    [--weights FILE --tol TOL --max_iters K]
//                    [--regrid] [--dtn_regrid N] [--dt_regrid T] [--amr_c C]
//                    [--limiter TVB --tvb_m M --t_limit N] [--limit_ic]
//                    [--projection_ic] [--conservative_reflux] [--cfl_uncovered]
//                    [--plot_every N] [--plot_dt T]
//                    [--data_dir DIR] [--restart N]
//                    [--graph_steps K] [--stream] [--max_steps N] [--dump FILE] [--quiet]
// Without --regrid a fine level is the static centre half of the domain; with
// --regrid the levels start full-domain and follow the tags at the case's
// regrid points (the reference's settings for the vortex and its KH case).
// Prints the reference's norm lines and, last, one JSON summary line.
// --dump writes int32 nlev, then per level int32 nslot and the FP64 U_w
// [q][n][slot] of that level, then per level int32 npatch and the boxes
// (int32 lo[dim], hi[dim] each).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "AmrDG.hpp"

namespace {

struct Options {
    std::string simulation_case = "isentropic_vortex";
    int p = 2, n = -1, max_level = -1, max_steps = -1, graph_steps = 64, max_iters = 2;
    double T = -1.0, c_dt = -1.0, tol = 1e-4;
    std::string precision = "fp64", weights, dump;
    bool quiet = false;
    bool graph = true;   // batches as step-graph replays (--stream: stream launches)
    bool regrid = false, limit_ic = false, projection_ic = false, conservative_reflux = false;
    bool cfl_uncovered = false;
    int dtn_regrid = -2, t_limit = -2;
    double dt_regrid = -2.0, amr_c = -1.0, tvb_m = -1.0;
    std::string limiter = "-";
    int plot_every = -1, restart = 0;
    double plot_dt = -1.0;
    std::string data_dir;
};

void usage() {
    std::printf(
        "usage: adg_main [--case isentropic_vortex|kelvin_helmolz_instability|kelvin_helmholtz_noise|\n"
        "                 kelvin_helmholtz_multimode|radial_shock_tube|spherical_shock_tube|density_wave]\n"
        "                [--p 1|2|3] [--n CELLS] [--max_level L] [--T T] [--c_dt C]\n"
        "                [--precision fp64|fp32] [--weights FILE --tol TOL --max_iters K]\n"
        "                [--regrid] [--dtn_regrid N] [--dt_regrid T] [--amr_c C]\n"
        "                [--limiter TVB --tvb_m M --t_limit N] [--limit_ic] [--projection_ic]\n"
        "                [--conservative_reflux] [--cfl_uncovered]\n"
        "                [--plot_every N] [--plot_dt T] [--data_dir DIR] [--restart N]\n"
        "                [--graph_steps K] [--stream] [--max_steps N] [--dump FILE] [--quiet]\n"
        "space dimension: %d\n",
        adg::SPACEDIM);
}

bool parse(int argc, char** argv, Options* o) {
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&](void) -> const char* {
            if (i + 1 >= argc) throw adg::AdgError("missing value after " + a);
            return argv[++i];
        };
        if (a == "--help" || a == "-h") return false;
        else if (a == "--case") o->simulation_case = next();
        else if (a == "--p") o->p = std::atoi(next());
        else if (a == "--n") o->n = std::atoi(next());
        else if (a == "--max_level") o->max_level = std::atoi(next());
        else if (a == "--T") o->T = std::atof(next());
        else if (a == "--c_dt") o->c_dt = std::atof(next());
        else if (a == "--precision") o->precision = next();
        else if (a == "--weights") o->weights = next();
        else if (a == "--tol") o->tol = std::atof(next());
        else if (a == "--max_iters") o->max_iters = std::atoi(next());
        else if (a == "--graph_steps") o->graph_steps = std::atoi(next());
        else if (a == "--stream") o->graph = false;
        else if (a == "--max_steps") o->max_steps = std::atoi(next());
        else if (a == "--dump") o->dump = next();
        else if (a == "--quiet") o->quiet = true;
        else if (a == "--regrid") o->regrid = true;
        else if (a == "--dtn_regrid") o->dtn_regrid = std::atoi(next());
        else if (a == "--dt_regrid") o->dt_regrid = std::atof(next());
        else if (a == "--amr_c") o->amr_c = std::atof(next());
        else if (a == "--limiter") o->limiter = next();
        else if (a == "--tvb_m") o->tvb_m = std::atof(next());
        else if (a == "--t_limit") o->t_limit = std::atoi(next());
        else if (a == "--limit_ic") o->limit_ic = true;
        else if (a == "--projection_ic") o->projection_ic = true;
        else if (a == "--conservative_reflux") o->conservative_reflux = true;
        else if (a == "--cfl_uncovered") o->cfl_uncovered = true;
        else if (a == "--plot_every") o->plot_every = std::atoi(next());
        else if (a == "--plot_dt") o->plot_dt = std::atof(next());
        else if (a == "--data_dir") o->data_dir = next();
        else if (a == "--restart") o->restart = std::atoi(next());
        else throw adg::AdgError("unknown option " + a);
    }
    return true;
}

template <int p>
int run(const Options& opt) {
    using namespace adg;
    // ------------------------------ SETTINGS (main.cpp:18-57)
    std::string simulation_case = opt.simulation_case;
    double T = 10.0, c_dt = 0.5;
    std::string limiter_type;
    double TVB_M = 0.0;
    int t_limit = -1, max_level = 0;
    int n_cell_x = 16, n_cell_y = 16, n_cell_z = 1;
    double L_lo = 0.0, L_hi = 10.0;
    std::array<int, 3> is_periodic{{1, 1, 1}};
    std::array<int, 3> bc_type{{2, 2, 2}};

    // ------------------------------ CASE OVERRIDES (main.cpp:62-111)
    if (simulation_case == "isentropic_vortex") {
        T = 10.0; c_dt = 0.9; max_level = 1;
        L_lo = 0.0; L_hi = 10.0; n_cell_x = n_cell_y = n_cell_z = 16;
    } else if (simulation_case == "kelvin_helmolz_instability" || simulation_case == "kelvin_helmholtz_noise" ||
               simulation_case == "kelvin_helmholtz_multimode") {
        T = 7.0; c_dt = 0.5; max_level = 0;
        L_lo = 0.0; L_hi = 1.0; n_cell_x = n_cell_y = n_cell_z = 64;
    } else if (simulation_case == "radial_shock_tube" || simulation_case == "spherical_shock_tube") {
        T = 0.25; c_dt = 0.5; max_level = 1;
        L_lo = 0.0; L_hi = 1.0; n_cell_x = n_cell_y = n_cell_z = 32;
        is_periodic = {{0, 0, 0}};
        bc_type = {{1, 1, 1}};
    } else if (simulation_case == "density_wave") {
        T = 1.0; c_dt = 0.5; max_level = 1;
        L_lo = 0.0; L_hi = 1.0; n_cell_x = n_cell_y = n_cell_z = 16;
    }
    // tagging-driven regrid: the reference's settings of its two cases
    // (main.cpp:62-124), the library's density-gradient tagging otherwise
    int dtn_regrid = -1;
    double dt_regrid = -1.0;
    std::vector<double> amr_c;
    bool limit_ic = opt.limit_ic;
    if (opt.regrid) {
        if (simulation_case == "isentropic_vortex") {
            max_level = 1; dtn_regrid = 1; amr_c = {1.4};
        } else if (simulation_case == "kelvin_helmolz_instability") {
            max_level = 2; dt_regrid = 0.05; amr_c = {1300.0, 2100.0};
            limiter_type = "TVB"; TVB_M = 5000.0; t_limit = 1;
        } else if (simulation_case == "radial_shock_tube" || simulation_case == "spherical_shock_tube") {
            max_level = 2; dtn_regrid = 4; amr_c = {0.2, 0.2};
            limiter_type = "TVB"; TVB_M = 0.0; t_limit = 1; limit_ic = true;
        } else {
            max_level = 2; dtn_regrid = 8; amr_c = {0.05, 0.05};
        }
    }
    if (opt.T > 0) T = opt.T;
    if (opt.c_dt > 0) c_dt = opt.c_dt;
    if (opt.n > 0) n_cell_x = n_cell_y = n_cell_z = opt.n;
    if (opt.max_level >= 0) max_level = opt.max_level;
    if (opt.dtn_regrid > -2) dtn_regrid = opt.dtn_regrid;
    if (opt.dt_regrid > -2.0) dt_regrid = opt.dt_regrid;
    if (opt.limiter != "-") limiter_type = opt.limiter;
    if (opt.tvb_m >= 0.0) TVB_M = opt.tvb_m;
    if (opt.t_limit > -2) t_limit = opt.t_limit;
    if (!opt.regrid && max_level > 1)
        throw AdgError("a static layout has at most one fine level (the centre half); use --regrid");
    if (max_level > 4) throw AdgError("max_level must be at most 4");
    if (opt.amr_c > 0.0) amr_c.assign((size_t)max_level, opt.amr_c);
    amr_c.resize((size_t)max_level, amr_c.empty() ? 1.0 : amr_c.back());

    std::vector<int> n_cell{n_cell_x, n_cell_y};
    if (SPACEDIM == 3) n_cell.push_back(n_cell_z);
    RealBox domain;
    for (int d = 0; d < 3; ++d) {
        domain.lo[d] = L_lo;
        domain.hi[d] = L_hi;
    }
    std::vector<std::array<int, 3>> amr_ratio(max_level, std::array<int, 3>{{2, 2, 2}});
    std::vector<double> amr_tvb_c((size_t)max_level + 1, 1.0);   // main.cpp:116

    // ------------------------------ SIMULATION CONSTRUCTION (main.cpp:143-151)
    Simulation<AmrDG<p>, Compressible_Euler> sim;
    sim.setModelSettings(simulation_case);
    sim.setNumericalSettings(T, c_dt, limiter_type, TVB_M, amr_tvb_c, t_limit);
    sim.setValidationMode(!opt.projection_ic);
    sim.setIO(opt.plot_every, opt.plot_dt, "", opt.restart);
    if (!opt.data_dir.empty()) sim.setDataDir(opt.data_dir);
    sim.setBoundaryConditions(bc_type, bc_type);
    sim.setGeometrySettings(domain, max_level, n_cell, 0, amr_ratio, is_periodic, amr_c,
                            opt.regrid ? dtn_regrid : -1, opt.regrid ? dt_regrid : -1.0, 1);
    if (!opt.regrid && max_level == 1) sim.mesh->set_centre_level();
    sim.solver->setLimitInitialState(limit_ic);
    sim.solver->setConservativeReflux(opt.conservative_reflux);
    sim.solver->setCflUncovered(opt.cfl_uncovered);

    sim.solver->setPrecision(opt.precision == "fp32" ? ADG_FP32_MFMA : ADG_FP64);
    if (!opt.weights.empty()) sim.solver->setLearnedWarmStart(opt.weights, opt.tol, opt.max_iters);
    sim.solver->setGraphSteps(opt.graph_steps);
    sim.solver->setBatchGraph(opt.graph);
    sim.solver->setMaxSteps(opt.max_steps);
    sim.solver->setVerbose(!opt.quiet);

    // ------------------------------ RUN (main.cpp:156-158)
    sim.run();
    auto& s = *sim.solver;
    const std::vector<double> l1 = s.L1Norm_DG_AMR(std::static_pointer_cast<ModelEquation<Compressible_Euler>>(sim.model));
    const std::vector<double> tot = s.conserved_totals();

    if (!opt.dump.empty()) {
        FILE* f = std::fopen(opt.dump.c_str(), "wb");
        if (!f) throw AdgError("cannot write " + opt.dump);
        const int nlev = sim.mesh->get_finest_lev() + 1;
        std::fwrite(&nlev, sizeof(int), 1, f);
        for (int l = 0; l < nlev; ++l) {
            const std::vector<double> u = s.get_U_w(l);
            const int ns = adg_level_nslot(s.get_ctx(), l);
            std::fwrite(&ns, sizeof(int), 1, f);
            std::fwrite(u.data(), sizeof(double), u.size(), f);
        }
        for (int l = 0; l < nlev; ++l) {
            const auto& ba = sim.mesh->get_BoxArray(l);
            const int nb = (int)ba.size();
            std::fwrite(&nb, sizeof(int), 1, f);
            for (const Box& b : ba) {
                std::fwrite(b.lo.data(), sizeof(int), SPACEDIM, f);
                std::fwrite(b.hi.data(), sizeof(int), SPACEDIM, f);
            }
        }
        std::fclose(f);
    }
    std::printf("{\"case\": \"%s\", \"p\": %d, \"dim\": %d, \"steps\": %d, \"t\": %.17g, \"Dt\": %.17g, "
                "\"regrids\": %d, \"limited\": %llu, \"levels\": %d, \"L1\": [",
                simulation_case.c_str(), p, SPACEDIM, s.get_n(), s.get_t(), s.get_Dt(), s.get_n_regrids(),
                s.cells_limited(), sim.mesh->get_finest_lev() + 1);
    for (size_t q = 0; q < l1.size(); ++q) std::printf("%s%.17g", q ? ", " : "", l1[q]);
    std::printf("], \"totals\": [");
    for (size_t q = 0; q < tot.size(); ++q) std::printf("%s%.17g", q ? ", " : "", tot[q]);
    std::printf("]}\n");
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    Options opt;
    try {
        if (!parse(argc, argv, &opt)) {
            usage();
            return 0;
        }
        if (opt.p == 1) return run<1>(opt);
        if (opt.p == 2) return run<2>(opt);
        if (opt.p == 3) return run<3>(opt);
        throw adg::AdgError("p must be 1, 2 or 3");
    } catch (const adg::AdgError& e) {
        std::fprintf(stderr, "adg_main: %s\n", e.what());
        return 2;
    }
}
