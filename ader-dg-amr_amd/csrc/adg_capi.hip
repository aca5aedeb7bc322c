// adg_capi.hip -- implementation of the C-ABI declared in include/adg.h.
//
// The context owns the element tables (one read-only device buffer), the
// flattened hierarchy lists and every field buffer.  All launches go to the
// context stream; the only host synchronisations are the explicit
// readbacks.  Time t and step size Dt live on the device (TimeState) so a
// whole step can be replayed without host round trips.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <algorithm>
#include <array>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "adg.h"
#include "adg_hier.hpp"
#include "adg_kernels.hpp"
#include "adg_mlp.hpp"
#include "adg_regrid.hpp"
#include "adg_tables.hpp"

namespace adg {
void project_ic(const ElementTables& t, const Hierarchy& H, int lev, int ic_case, const double* prm, double gamma,
                std::vector<double>& out);
void project_ic_patch(const ElementTables& t, const Hierarchy& H, int p, int ic_case, const double* prm, double gamma,
                      std::vector<double>& out);
}

// Device scratch of adg_conserved_totals_noalloc: the per-workgroup partial
// sums [Q][blocks] of the totals, kScratchDoubles doubles (320 KiB) on every
// GPU the library runs on, held until the process ends.  Every workgroup
// writes its own entries before the second stage reads them, so no call
// needs it cleared; calls that overlap on one GPU (several streams or
// contexts) race on it.  Host code never passes the name `scratch` itself to
// a kernel or copy (on the host it is only a stand-in symbol): it uses the
// device address that hipGetSymbolAddress returns for the current device,
// resolved once per device by adg_create.
constexpr long kScratchDoubles = 5L * 8192;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (nullptr on failure).
hipError_t resolve_scratch(double** out) {
    *out = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}
}  // namespace

struct adg_ctx {
    int D = 2, P = 1, Q = 4, precision = 0;
    double* scratch_dev = nullptr;   // this device's address of `scratch`
    double gamma = 1.4;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    adg::ElementTables tab;
    adg::TabOff to{};
    double* d_tab = nullptr;

    bool geo_set = false, levels_set = false;
    int dom0[3] = {1, 1, 1};
    double plo[3] = {0, 0, 0}, phi[3] = {1, 1, 1};
    int per[3] = {1, 1, 1};
    adg::Hierarchy H;
    adg::Hierarchy H_spare;   // the lists of an earlier layout, whose host storage the next build takes over
    adg::LevelGeo geo{};

    double *Uw = nullptr, *Hw_dbg = nullptr, *Mom = nullptr, *partial = nullptr;
    void *Vol = nullptr, *Tr = nullptr;     // element type of the form (double / float)
    void* Fnum[3] = {nullptr, nullptr, nullptr};
    size_t esz = sizeof(double);
    // MFMA A-operand fragments in the element type of the form:
    // [0] quad-gather predictor, [1] node-lane predictor
    adg::Fragments frags[2];
    adg::FragOff fo[2]{};
    void* d_frag[2] = {nullptr, nullptr};
    int pred_kind = 0;       // 0 auto, 1 VALU, 2 MFMA quad-gather, 3 MFMA node-lane
    int nl_wpb = 8;          // wavefronts per workgroup of the node-lane predictor (8 or 16)
    int ncu = 256;
    // resident workgroups per CU of the plain [0] and learned [1] predictor
    // kernels: quad, node-lane x8, node-lane x16
    int pred_per_cu[2][3] = {{1, 1, 1}, {1, 1, 1}};
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    double graph_cdt = -1.0, graph_T = -1.0;
    int graph_kind = -1;
    int graph_unroll = 1;      // steps captured per graph
    int graph_captured = 0;    // steps in the captured graph
    int* valid_list = nullptr;
    int* cell_face[3][2] = {};
    int* face_tr[3] = {};     // face -> (left, right) slot, copy ghosts replaced by their sources
    int* pred_slots = nullptr;
    int* cg_slot = nullptr;   // copy ghosts and their sources
    int* cg_src = nullptr;
    int *g_slot = nullptr, *g_src = nullptr;
    uint8_t *g_kind = nullptr, *g_child = nullptr, *slot_lev = nullptr;
    int* ad_coarse[adg::kMaxLev] = {};
    int* ad_fine[adg::kMaxLev] = {};
    int *mom_face = nullptr, *rc_slot = nullptr, *rc_start = nullptr, *rc_count = nullptr, *re_mom = nullptr;
    uint8_t *mom_d = nullptr, *mom_kind = nullptr, *mom_sel = nullptr, *mom_lev = nullptr, *rc_lev = nullptr;
    double* re_coef = nullptr;
    int* uncovered = nullptr;
    double* unc_centre = nullptr;   // cell centres of the uncovered cells [n][3]
    adg::TimeState* ts = nullptr;
    // TVB limiter of the finest level (Limiter_w): settings, the finest
    // level's valid slots and resolved neighbour slots, cells-limited counter
    bool lim_on = false;
    double lim_M = 0.0;
    double lim_tvb_c[adg::kMaxLev] = {};
    int lim_t_limit = 1;
    int* lim_cells = nullptr;
    int* lim_nbr = nullptr;
    long lim_n = 0;
    unsigned long long* lim_count = nullptr;
    // tagging-driven regrid: AMR settings, tagging criterion and thresholds
    adg::AmrSettings amr;
    bool amr_set = false;
    int tag_crit = 2;
    double tag_c[adg::kMaxLev] = {};
    double tag_prm[4] = {5.0, 5.0, 0.0, 0.0};
    bool keep_hw = false;
    int iters_override = 0;    // > 0: Picard iterations of the next plain predictor launches (collection)
    bool independent = false;  // adg_set_levels builds a batch of independent periodic patches
    bool conservative_reflux = false;   // reflux every cell the registers reach (not the reference's rule)
    bool cfl_uncovered = false;         // set_Dt over uncovered cells only (not the reference's rule)
    // per-step scratch (volume terms, traces, face fluxes, warm-start
    // increments): kept across layouts and grown when a layout needs more,
    // so that a regrid allocates nothing of this size
    struct Scratch { void* p = nullptr; size_t cap = 0; };
    Scratch sc_vol, sc_tr, sc_fnum[3], sc_warm;
    // host seconds spent in adg_regrid / adg_remake, cumulative: tags, grid
    // generation, hierarchy lists, device set-up (uploads, allocation),
    // data transfer onto the new layout; and the number of layout changes
    double rg_sec[5] = {0, 0, 0, 0, 0};
    long rg_changes = 0;

    // learned warm start
    struct Mlp {
        bool init = false;
        int nin = 0, nin_p = 0, nout = 0, nout_p = 0, nh = adg::kMlpHidden;
        long npar = 0, oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0;
        float *par = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr, *norm = nullptr;
        uint2* frag = nullptr;
        int f1 = 0, f2 = 0, f3 = 0, nfrag = 0;
        bool frag_ok = false;
        float *X = nullptr, *Y = nullptr;   // dataset rows (raw features / targets)
        long ndata = 0, cap = 0;
        float* work = nullptr;              // minibatch activations and gradients
        long work_rows = 0;
        double* dacc = nullptr;             // loss accumulator and column statistics
        long adam_t = 0;
    } mlp;
    float* warm = nullptr;                    // learned H_w increments [Q*Np_st][slot]
    unsigned long long* lstats = nullptr;     // [elements, fallbacks] of learned predictor launches
    bool learned = false;
    float learned_tol = 1e-2f;
    int warm_tiles = 1;           // warm-start tiles per wavefront: 1 or 2 of 16 elements, 0 = one of 8
    int warm_per_cu[3] = {0, 0, 0};
    int learned_iters = 1;
    std::string err;
    std::vector<void*> level_allocs;
};

namespace {

int fail(adg_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                              \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(ctx, ADG_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

template <class T>
int dev_upload(adg_ctx* c, T** dst, const std::vector<T>& src) {
    *dst = nullptr;
    if (src.empty()) return ADG_OK;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, src.size() * sizeof(T));
    if (e != hipSuccess) return fail(c, ADG_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->level_allocs.push_back(p);
    HIPCHK(c, hipMemcpy(p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    *dst = static_cast<T*>(p);
    return ADG_OK;
}

int dev_alloc_bytes(adg_ctx* c, void** dst, size_t bytes) {
    *dst = nullptr;
    if (bytes == 0) return ADG_OK;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(c, ADG_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->level_allocs.push_back(p);
    HIPCHK(c, hipMemsetAsync(p, 0, bytes, c->stream));
    *dst = p;
    return ADG_OK;
}
// Scratch of at least `bytes`, zeroed: the kept buffer when it is large
// enough, else a new one with a quarter more room for the next layouts.
int scratch_bytes(adg_ctx* c, adg_ctx::Scratch& sc, void** dst, size_t bytes) {
    *dst = nullptr;
    if (bytes == 0) return ADG_OK;
    if (sc.cap < bytes) {
        if (sc.p) {
            HIPCHK(c, hipStreamSynchronize(c->stream));
            (void)hipFree(sc.p);
            sc.p = nullptr;
            sc.cap = 0;
        }
        size_t cap = bytes + bytes / 4;
        hipError_t e = hipMalloc(&sc.p, cap);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            cap = bytes;
            e = hipMalloc(&sc.p, cap);
        }
        if (e != hipSuccess) {
            sc.p = nullptr;
            return fail(c, ADG_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        sc.cap = cap;
    }
    HIPCHK(c, hipMemsetAsync(sc.p, 0, bytes, c->stream));
    *dst = sc.p;
    return ADG_OK;
}

int dev_alloc(adg_ctx* c, double** dst, size_t n) {
    void* p = nullptr;
    int rc = dev_alloc_bytes(c, &p, n * sizeof(double));
    *dst = static_cast<double*>(p);
    return rc;
}

enum { PV_VALU = 0, PV_QUAD = 1, PV_NL = 2 };

// Grid of a persistent launch over nunit wavefront units (a unit: the
// elements one wavefront takes per round): no more workgroups than are
// resident at once (registers and LDS both limit it), so no workgroup waits
// for a CU, and at least one per CU while there are units to spread, so
// that a launch of less than one round reaches every CU (the deal gives
// each workgroup its share of the partial round) instead of filling a few
// CUs with whole workgroups.
inline long persistent_grid(long nunit, int wpb, int per_cu, int ncu) {
    const long nbatch = (nunit + wpb - 1) / wpb;
    return std::min<long>((long)per_cu * ncu, std::max<long>(nbatch, std::min<long>(ncu, nunit)));
}
constexpr size_t kLdsMax = 160 * 1024;

// Dynamic LDS bytes of an MFMA predictor workgroup: the fragments, plus the
// per-wave H_w re-layout buffer of the FP32 quad-gather form (8 waves) or the
// per-wave flux tile of the node-lane form (wpb waves).
size_t mfma_lds_bytes(const adg_ctx* c, int variant, int wpb = 8) {
    if (variant == PV_QUAD) {
        const size_t xch = c->esz == sizeof(double) ? 0 : (size_t)8 * c->frags[0].MB2 * 256;
        return ((size_t)c->frags[0].total * 64 + xch) * c->esz;
    }
    const bool f64 = c->esz == sizeof(double);
    const size_t tst = f64 ? adg::nl_tile_stride<double>() : adg::nl_tile_stride<float>();
    const size_t tdir = f64 ? adg::nl_tile_dirs<double>(c->D) : adg::nl_tile_dirs<float>(c->D);
    return ((size_t)c->frags[1].total * 64 + (size_t)wpb * tdir * 16 * tst) * c->esz;
}
// elements per wavefront of the node-lane predictor: 4 (Q = 4) or 3 (Q = 5)
inline int nl_elems(const adg_ctx* c) { return c->Q == 4 ? 4 : 3; }
bool nl_available(const adg_ctx* c, int wpb) {
    return (c->D == 2 || c->P <= 2) && mfma_lds_bytes(c, PV_NL, wpb) <= kLdsMax;
}
bool quad_available(const adg_ctx* c) { return mfma_lds_bytes(c, PV_QUAD) <= kLdsMax; }
// Predictor kernel of the next launch.
int pred_variant(const adg_ctx* c) {
    switch (c->pred_kind) {
        case 1: return PV_VALU;
        case 2: return PV_QUAD;
        case 3: return PV_NL;
        default: break;
    }
    if (c->precision == ADG_FP64 && c->P == 1) return PV_VALU;
    if (nl_available(c, c->nl_wpb)) return PV_NL;
    if (quad_available(c)) return PV_QUAD;
    return PV_VALU;
}

void drop_graph(adg_ctx* c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
}

void free_levels(adg_ctx* c) {
    drop_graph(c);
    for (void* p : c->level_allocs) (void)hipFree(p);
    c->level_allocs.clear();
    c->levels_set = false;
    c->Hw_dbg = nullptr;
    c->warm = nullptr;
}

bool supported(int P, int D) { return (D == 2 || D == 3) && P >= 1 && P <= 3; }

#define ADG_DISPATCH(c, MACRO)                \
    switch ((c)->P * 10 + (c)->D) {           \
        case 12: MACRO(1, 2); break;          \
        case 22: MACRO(2, 2); break;          \
        case 32: MACRO(3, 2); break;          \
        case 13: MACRO(1, 3); break;          \
        case 23: MACRO(2, 3); break;          \
        case 33: MACRO(3, 3); break;          \
        default: return fail(c, ADG_E_UNSUP, "unsupported (p, dim)"); \
    }

inline unsigned nblocks(long threads, int bs = 256) { return (unsigned)((threads + bs - 1) / bs); }

int pack_tables(adg_ctx* c) {
    const adg::ElementTables& t = c->tab;
    std::vector<double> buf;
    auto add = [&](const std::vector<double>& v) {
        int off = (int)buf.size();
        buf.insert(buf.end(), v.begin(), v.end());
        while (buf.size() % 8) buf.push_back(0.0);  // 64-byte alignment of every table
        return off;
    };
    std::vector<double> mkc(t.Np_s), minv(t.Np_s);
    for (int k = 0; k < t.Np_s; ++k) {
        mkc[k] = t.Mk_corr[(size_t)k * t.Np_s + k];
        minv[k] = t.Minv[(size_t)k * t.Np_s + k];
    }
    c->to.A0 = add(t.A0);
    c->to.B = add(t.B);
    c->to.V = add(t.V);
    c->to.Skc = add(t.Sk_corr);
    c->to.Vbd = add(t.Vbd);
    c->to.Mkbdm = add(t.Mkbdm);
    c->to.Mkbdp = add(t.Mkbdp);
    {
        std::vector<float> f32(t.Mkbdp.begin(), t.Mkbdp.end());
        std::vector<double> packed((f32.size() + 1) / 2, 0.0);
        std::memcpy(packed.data(), f32.data(), f32.size() * sizeof(float));
        c->to.Mkbdpf = add(packed);
    }
    c->to.Mkc = add(mkc);
    c->to.Minv = add(minv);
    c->to.Pcf = add(t.P_cf);
    c->to.Pfc = add(t.P_fc);
    c->to.Pfl = add(t.P_flux_low);
    c->to.Pfh = add(t.P_flux_high);
    c->to.phi0 = add(t.phi0);
    std::vector<double> sgn((size_t)t.D * t.Np_s);
    for (int d = 0; d < t.D; ++d)
        for (int k = 0; k < t.Np_s; ++k) sgn[(size_t)d * t.Np_s + k] = (t.mi_s[k][d] % 2) ? -1.0 : 1.0;
    c->to.sgn = add(sgn);
    // spatial basis at the points of the Lp norms, their weights and
    // coordinates: the N^D solution points (L1, AmrDG.h:1580-1585) and the
    // (2N)^D points of the L2 rule (AmrDG.h:1587-1655); point m has the
    // node of dimension 0 slowest
    for (int rule = 0; rule < 2; ++rule) {
        const std::vector<double>& nd = rule ? t.nodes2 : t.nodes;
        const std::vector<double>& wt = rule ? t.weights2 : t.weights;
        const int n1 = (int)nd.size(), Ns = t.Np_s, D = t.D;
        int Ms = 1;
        for (int d = 0; d < D; ++d) Ms *= n1;
        std::vector<double> vs((size_t)Ms * Ns), ws((size_t)Ms), xs((size_t)Ms * D);
        for (int m = 0; m < Ms; ++m) {
            double w = 1.0;
            int div = Ms;
            for (int d = 0; d < D; ++d) {
                div /= n1;
                const int k = (m / div) % n1;
                xs[(size_t)m * D + d] = nd[k];
                w *= wt[k];
            }
            ws[m] = w;
            for (int n = 0; n < Ns; ++n) {
                double ph = 1.0;
                for (int d = 0; d < D; ++d) ph *= std::legendre(t.mi_s[n][d], xs[(size_t)m * D + d]);
                vs[(size_t)m * Ns + n] = ph;
            }
        }
        c->to.lp[rule].Vs = add(vs);
        c->to.lp[rule].ws = add(ws);
        c->to.lp[rule].xs = add(xs);
        c->to.lp[rule].M = Ms;
    }
    c->to.total = (int)buf.size();
    void* p = nullptr;
    HIPCHK(c, hipMalloc(&p, buf.size() * sizeof(double)));
    c->d_tab = static_cast<double*>(p);
    HIPCHK(c, hipMemcpy(c->d_tab, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice));
    // MFMA A-operand fragments of both predictor layouts in the element type
    // of the form (K order / row permutation follow its accumulator rows)
    const bool fp64 = c->precision == ADG_FP64;
    for (int v = 0; v < 2; ++v) {
        adg::Fragments& f = c->frags[v];
        f = adg::build_fragments(t, fp64, v == 1);
        c->fo[v] = adg::FragOff{f.A0, f.V, f.B, f.S, f.T, f.BT, f.total, f.tail ? 1 : 0};
        HIPCHK(c, hipMalloc(&c->d_frag[v], f.data.size() * c->esz));
        if (fp64) {
            HIPCHK(c, hipMemcpy(c->d_frag[v], f.data.data(), f.data.size() * sizeof(double), hipMemcpyHostToDevice));
        } else {
            std::vector<float> ff(f.data.begin(), f.data.end());
            HIPCHK(c, hipMemcpy(c->d_frag[v], ff.data(), ff.size() * sizeof(float), hipMemcpyHostToDevice));
        }
    }
    return ADG_OK;
}

}  // namespace

extern "C" {

int adg_create(adg_ctx** out, int dim, int p, double gamma, int precision, void* hip_stream) {
    if (!out) return ADG_E_ARG;
    *out = nullptr;
    if (!supported(p, dim)) return ADG_E_UNSUP;
    if (precision != ADG_FP64 && precision != ADG_FP32_MFMA) return ADG_E_ARG;
    if (!(gamma > 1.0)) return ADG_E_ARG;
    adg_ctx* c = new adg_ctx();
    c->D = dim;
    c->P = p;
    c->Q = dim + 2;
    c->gamma = gamma;
    c->precision = precision;
    c->esz = precision == ADG_FP32_MFMA ? sizeof(float) : sizeof(double);
    try {
        c->tab = adg::build_element_tables(p, dim);
    } catch (const std::exception& e) {
        delete c;
        return ADG_E_ARG;
    }
    if (hip_stream) {
        c->stream = static_cast<hipStream_t>(hip_stream);
    } else {
        if (hipStreamCreate(&c->stream) != hipSuccess) { delete c; return ADG_E_HIP; }
        c->own_stream = true;
    }
    int rc = pack_tables(c);
    if (rc != ADG_OK) { adg_destroy(c); return rc; }
    {
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
            c->ncu = ncu;
    }
    if (precision == ADG_FP32_MFMA && !quad_available(c) && !nl_available(c, 8)) { adg_destroy(c); return ADG_E_UNSUP; }
    if (resolve_scratch(&c->scratch_dev) != hipSuccess || !c->scratch_dev) { adg_destroy(c); return ADG_E_HIP; }
    void* tsp = nullptr;
    if (hipMalloc(&tsp, sizeof(adg::TimeState)) != hipSuccess) { adg_destroy(c); return ADG_E_NOMEM; }
    c->ts = static_cast<adg::TimeState*>(tsp);
    void* lsp = nullptr;
    if (hipMalloc(&lsp, 3 * sizeof(unsigned long long)) != hipSuccess) { adg_destroy(c); return ADG_E_NOMEM; }
    c->lstats = static_cast<unsigned long long*>(lsp);
    c->lim_count = c->lstats + 2;
    const unsigned long long z3[3] = {0ull, 0ull, 0ull};
    if (hipMemcpy(c->lstats, z3, sizeof(z3), hipMemcpyHostToDevice) != hipSuccess) { adg_destroy(c); return ADG_E_HIP; }
    for (int l = 0; l < adg::kMaxLev; ++l) c->lim_tvb_c[l] = 1.0;
    adg::TimeState t0{0.0, 0.0, 0x7ff0000000000000ull, 0u, 0u};
    if (hipMemcpy(c->ts, &t0, sizeof(t0), hipMemcpyHostToDevice) != hipSuccess) { adg_destroy(c); return ADG_E_HIP; }
    *out = c;
    return ADG_OK;
}

int adg_set_geometry(adg_ctx* c, const int* dom_n, const double* prob_lo, const double* prob_hi,
                     const int* periodic) {
    if (!c || !dom_n || !prob_lo || !prob_hi || !periodic) return fail(c, ADG_E_ARG, "null argument");
    for (int d = 0; d < c->D; ++d) {
        if (dom_n[d] <= 0 || !(prob_hi[d] > prob_lo[d])) return fail(c, ADG_E_ARG, "bad domain");
        c->dom0[d] = dom_n[d];
        c->plo[d] = prob_lo[d];
        c->phi[d] = prob_hi[d];
        c->per[d] = periodic[d] ? 1 : 0;
    }
    c->geo_set = true;
    return ADG_OK;
}

static int prepare_kernels(adg_ctx* c);

int adg_set_levels(adg_ctx* c, int nlev, const int* npatch, const int* boxes) {
    if (!c) return ADG_E_ARG;
    if (!c->geo_set) return fail(c, ADG_E_STATE, "adg_set_geometry must come first");
    if (nlev < 1 || nlev > adg::kMaxLev || !npatch || !boxes) return fail(c, ADG_E_ARG, "bad level arguments");
    free_levels(c);
    const auto t_h0 = std::chrono::steady_clock::now();
    try {
        c->H = adg::build_hierarchy(c->D, c->dom0, c->plo, c->phi, c->per, nlev, npatch, boxes, c->independent ? 1 : 0,
                                    c->conservative_reflux ? 1 : 0, &c->H_spare);
    } catch (const std::exception& e) {
        return fail(c, ADG_E_ARG, e.what());
    }
    const auto t_h1 = std::chrono::steady_clock::now();
    c->rg_sec[2] += std::chrono::duration<double>(t_h1 - t_h0).count();
    adg::Hierarchy& H = c->H;
    const int D = c->D, Q = c->Q;
    const adg::ElementTables& t = c->tab;
    std::memset(&c->geo, 0, sizeof(c->geo));
    for (int l = 0; l < nlev; ++l) {
        double vol = 1.0, avg = 0.0;
        for (int d = 0; d < D; ++d) {
            c->geo.dx[l][d] = H.lev[l].dx[d];
            vol *= H.lev[l].dx[d];
            avg += H.lev[l].dx[d];
        }
        for (int d = 0; d < D; ++d) {
            double a = 1.0;
            for (int e = 0; e < D; ++e)
                if (e != d) a *= H.lev[l].dx[e];
            c->geo.area[l][d] = a;
        }
        c->geo.vol[l] = vol;
        c->geo.dx_avg[l] = avg / (double)D;
    }
    int rc;
#define UP(dst, src) if ((rc = dev_upload(c, &(dst), (src))) != ADG_OK) return rc
    UP(c->slot_lev, H.slot_lev);
    UP(c->valid_list, H.valid_list);
    for (int d = 0; d < D; ++d) {
        UP(c->cell_face[d][0], H.cell_face[d][0]);
        UP(c->cell_face[d][1], H.cell_face[d][1]);
        UP(c->face_tr[d], H.face_tr[d]);
    }
    UP(c->g_slot, H.g_slot);
    UP(c->g_src, H.g_src);
    UP(c->g_kind, H.g_kind);
    UP(c->g_child, H.g_child);
    for (int l = 1; l < nlev; ++l) {
        UP(c->ad_coarse[l], H.ad_coarse[l]);
        UP(c->ad_fine[l], H.ad_fine[l]);
    }
    UP(c->mom_face, H.mom_face);
    UP(c->mom_d, H.mom_d);
    UP(c->mom_kind, H.mom_kind);
    UP(c->mom_sel, H.mom_sel);
    UP(c->mom_lev, H.mom_lev);
    UP(c->rc_slot, H.rc_slot);
    UP(c->rc_start, H.rc_start);
    UP(c->rc_count, H.rc_count);
    UP(c->rc_lev, H.rc_lev);
    UP(c->re_mom, H.re_mom);
    UP(c->re_coef, H.re_coef);
    UP(c->uncovered, H.uncovered);
    UP(c->pred_slots, H.pred_slots);
    UP(c->cg_slot, H.copy_ghost);
    UP(c->cg_src, H.copy_src);
    c->lim_n = (long)H.lim_cells.size();
    UP(c->lim_cells, H.lim_cells);
    UP(c->lim_nbr, H.lim_nbr);
#undef UP
    const size_t NS = (size_t)H.NS;
    if ((rc = dev_alloc(c, &c->Uw, (size_t)Q * t.Np_s * NS)) != ADG_OK) return rc;
    const size_t ves = (size_t)Q * t.Np_s, tes = (size_t)2 * D * t.qMp_st_bd * Q;   // per slot
    if ((rc = scratch_bytes(c, c->sc_vol, &c->Vol, ves * NS * c->esz)) != ADG_OK) return rc;
    if ((rc = scratch_bytes(c, c->sc_tr, &c->Tr, tes * NS * c->esz)) != ADG_OK) return rc;
    for (int d = 0; d < D; ++d)
        if ((rc = scratch_bytes(c, c->sc_fnum[d], (void**)&c->Fnum[d], (size_t)Q * t.qMp_st_bd * H.NF[d] * c->esz)) !=
            ADG_OK)
            return rc;
    if ((rc = dev_alloc(c, &c->Mom, (size_t)Q * t.Np_s * H.mom_face.size())) != ADG_OK) return rc;
    if ((rc = dev_alloc(c, &c->partial, (size_t)Q * nblocks((long)H.uncovered.size()) + 1)) != ADG_OK) return rc;
    if (c->keep_hw && (rc = dev_alloc(c, &c->Hw_dbg, (size_t)Q * t.Np_st * NS)) != ADG_OK) return rc;
    if ((rc = scratch_bytes(c, c->sc_warm, (void**)&c->warm, (size_t)Q * t.Np_st * NS * sizeof(float))) != ADG_OK)
        return rc;
    if ((rc = dev_upload(c, &c->unc_centre, H.unc_centre)) != ADG_OK) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->levels_set = true;
    rc = prepare_kernels(c);
    c->rg_sec[3] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_h1).count();
    return rc;
}

// Raises the dynamic-LDS limit of the MFMA predictors once, outside any
// stream capture, and records how many of their workgroups a CU holds.
extern "C++" {
template <typename T, int P, int D>
static int prepare_pred(adg_ctx* c) {
    auto setup = [&](const void* kf, int threads, size_t lds, int& per_cu) -> int {
        per_cu = 0;
        if (lds > kLdsMax) return ADG_OK;
        HIPCHK(c, hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        int nb = 0;
        HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kf, threads, lds));
        per_cu = nb > 0 ? nb : 1;
        return ADG_OK;
    };
    int rc = ADG_OK;
    for (int l = 0; l < 2 && rc == ADG_OK; ++l) {
        int* pc = c->pred_per_cu[l];
        rc = setup(l ? adg::predictor_mfma_kernel<T, P, D, true>() : adg::predictor_mfma_kernel<T, P, D, false>(), 512,
                   mfma_lds_bytes(c, PV_QUAD), pc[0]);
        if (rc == ADG_OK && (D == 2 || P <= 2)) {
            rc = setup(adg::predictor_nl_kernel<T, P, D, 8>(l), 512, mfma_lds_bytes(c, PV_NL, 8), pc[1]);
            if (rc == ADG_OK)
                rc = setup(adg::predictor_nl_kernel<T, P, D, 16>(l), 1024, mfma_lds_bytes(c, PV_NL, 16), pc[2]);
        }
    }
    return rc;
}
}  // extern "C++"

static int prepare_kernels(adg_ctx* c) {
    int rc = ADG_OK;
    if (c->precision == ADG_FP64) {
#define L_PP(PP, DD) rc = prepare_pred<double, PP, DD>(c)
        ADG_DISPATCH(c, L_PP);
#undef L_PP
    } else {
#define L_PP(PP, DD) rc = prepare_pred<float, PP, DD>(c)
        ADG_DISPATCH(c, L_PP);
#undef L_PP
    }
    return rc;
}

int adg_level_nslot(adg_ctx* c, int lev) {
    if (!c || !c->levels_set || lev < 0 || lev >= c->H.nlev) return ADG_E_ARG;
    return (int)c->H.lev[lev].nslot;
}

int adg_level_nface(adg_ctx* c, int lev, int d) {
    if (!c || !c->levels_set || lev < 0 || lev >= c->H.nlev || d < 0 || d >= c->D) return ADG_E_ARG;
    return (int)c->H.lev[lev].nface[d];
}

int adg_level_npred(adg_ctx* c, int lev) {
    if (!c || !c->levels_set || lev < 0 || lev >= c->H.nlev) return ADG_E_ARG;
    return (int)(c->H.pred_off[lev + 1] - c->H.pred_off[lev]);
}

static int copy_level(adg_ctx* c, int lev, double* dev, int ncomp, double* host, bool up) {
    if (!c) return ADG_E_ARG;
    if (!c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    if (lev < 0 || lev >= c->H.nlev || !host || !dev) return fail(c, ADG_E_ARG, "bad level or buffer");
    const size_t ns = (size_t)c->H.lev[lev].nslot, off = (size_t)c->H.lev[lev].slot_off, NS = (size_t)c->H.NS;
    const size_t rows = (size_t)c->Q * ncomp;
    if (up)
        HIPCHK(c, hipMemcpy2DAsync(dev + off, NS * sizeof(double), host, ns * sizeof(double), ns * sizeof(double),
                                   rows, hipMemcpyHostToDevice, c->stream));
    else
        HIPCHK(c, hipMemcpy2DAsync(host, ns * sizeof(double), dev + off, NS * sizeof(double), ns * sizeof(double),
                                   rows, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ADG_OK;
}

int adg_upload_Uw(adg_ctx* c, int lev, const double* host) {
    return copy_level(c, lev, c ? c->Uw : nullptr, c ? c->tab.Np_s : 0, const_cast<double*>(host), true);
}
int adg_download_Uw(adg_ctx* c, int lev, double* host) {
    return copy_level(c, lev, c ? c->Uw : nullptr, c ? c->tab.Np_s : 0, host, false);
}

int adg_project_ic(adg_ctx* c, int lev, int ic_case, const double* params) {
    if (!c) return ADG_E_ARG;
    if (!c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    if (lev < 0 || lev >= c->H.nlev || !params || ic_case < 0 || ic_case > 7)
        return fail(c, ADG_E_ARG, "bad level, case or params");
    if (ic_case == 7 && !(params[3] >= 1.0))
        return fail(c, ADG_E_ARG, "case 7 needs a noise grid of at least one cell");
    if ((ic_case <= 1 || ic_case == 4 || ic_case == 6 || ic_case == 7) && c->D != 2)
        return fail(c, ADG_E_UNSUP, "vortex / Kelvin-Helmholtz states are 2D");
    std::vector<double> host;
    adg::project_ic(c->tab, c->H, lev, ic_case, params, c->gamma, host);
    return copy_level(c, lev, c->Uw, c->tab.Np_s, host.data(), true);
}

int adg_project_ic_patches(adg_ctx* c, int lev, int ic_case, int nprm, const double* params) {
    if (!c) return ADG_E_ARG;
    if (!c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    if (lev < 0 || lev >= c->H.nlev || !params || nprm < 1 || ic_case < 0 || ic_case > 6)
        return fail(c, ADG_E_ARG, "bad level, case or params");
    if ((ic_case <= 1 || ic_case == 4 || ic_case == 6) && c->D != 2)
        return fail(c, ADG_E_UNSUP, "vortex / Kelvin-Helmholtz states are 2D");
    const adg::LevelLayout& L = c->H.lev[lev];
    std::vector<double> host((size_t)c->Q * c->tab.Np_s * L.nslot, 0.0);
    for (int p = L.p_begin; p < L.p_end; ++p)
        adg::project_ic_patch(c->tab, c->H, p, ic_case, params + (size_t)(p - L.p_begin) * nprm, c->gamma, host);
    return copy_level(c, lev, c->Uw, c->tab.Np_s, host.data(), true);
}

int adg_set_independent_patches(adg_ctx* c, int on) {
    if (!c) return ADG_E_ARG;
    c->independent = on != 0;
    return ADG_OK;
}

int adg_set_conservative_reflux(adg_ctx* c, int on) {
    if (!c) return ADG_E_ARG;
    if ((on != 0) == c->conservative_reflux) return ADG_OK;
    c->conservative_reflux = on != 0;
    if (!c->levels_set) return ADG_OK;
    // rebuild the register lists on the current layout (the data carry over)
    std::vector<int> np((size_t)c->H.nlev), bx;
    for (int l = 0; l < c->H.nlev; ++l) np[(size_t)l] = c->H.lev[l].p_end - c->H.lev[l].p_begin;
    for (const adg::PatchLayout& P : c->H.patch) {
        for (int d = 0; d < c->D; ++d) bx.push_back(P.lo[d]);
        for (int d = 0; d < c->D; ++d) bx.push_back(P.hi[d]);
    }
    return adg_remake(c, c->H.nlev, np.data(), bx.data());
}

int adg_regrid_stats(adg_ctx* c, double* out, int reset) {
    if (!c || !out) return ADG_E_ARG;
    out[0] = (double)c->rg_changes;
    for (int k = 0; k < 5; ++k) out[1 + k] = c->rg_sec[k];
    if (reset) {
        c->rg_changes = 0;
        for (double& v : c->rg_sec) v = 0.0;
    }
    return ADG_OK;
}

int adg_set_cfl_uncovered(adg_ctx* c, int on) {
    if (!c) return ADG_E_ARG;
    if ((on != 0) != c->cfl_uncovered) drop_graph(c);
    c->cfl_uncovered = on != 0;
    return ADG_OK;
}

int adg_keep_Hw(adg_ctx* c, int on) {
    if (!c) return ADG_E_ARG;
    // a captured step graph holds the H_w pointer of its predictor launch
    if ((on != 0) != c->keep_hw) drop_graph(c);
    c->keep_hw = on != 0;
    if (c->keep_hw && c->levels_set && !c->Hw_dbg) {
        int rc = dev_alloc(c, &c->Hw_dbg, (size_t)c->Q * c->tab.Np_st * c->H.NS);
        if (rc != ADG_OK) return rc;
    }
    return ADG_OK;
}
int adg_download_Hw(adg_ctx* c, int lev, double* host) {
    if (!c) return ADG_E_ARG;
    if (!c->Hw_dbg) return fail(c, ADG_E_STATE, "adg_keep_Hw(ctx,1) was not enabled");
    return copy_level(c, lev, c->Hw_dbg, c->tab.Np_st, host, false);
}

int adg_download_Fnum(adg_ctx* c, int lev, int d, double* host) {
    if (!c || !c->levels_set || lev < 0 || lev >= c->H.nlev || d < 0 || d >= c->D || !host)
        return fail(c, ADG_E_ARG, "bad arguments");
    const size_t nf = (size_t)c->H.lev[lev].nface[d], off = (size_t)c->H.lev[lev].face_off[d];
    const size_t NF = (size_t)c->H.NF[d], rows = (size_t)c->Q * c->tab.qMp_st_bd, es = c->esz;
    std::vector<unsigned char> tmp(rows * NF * es);   // device layout [q][mb][face]
    HIPCHK(c, hipMemcpyAsync(tmp.data(), c->Fnum[d], tmp.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < rows; ++i)
        for (size_t f = 0; f < nf; ++f) {
            const size_t src = i * NF + off + f;
            host[i * nf + f] = es == sizeof(double) ? reinterpret_cast<double*>(tmp.data())[src]
                                                    : (double)reinterpret_cast<float*>(tmp.data())[src];
        }
    return ADG_OK;
}

int adg_fill_ghosts(adg_ctx* c) {
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    const long ng = (long)c->H.g_slot.size();
    if (ng == 0) return ADG_OK;
    adg::GhostArgs a{c->d_tab, c->to, c->Uw, c->g_slot, c->g_src, c->g_kind, c->g_child, c->H.NS, ng};
#define L_GH(PP, DD) hipLaunchKernelGGL((adg::k_ghosts<PP, DD>), dim3(nblocks(adg::ghost_threads(ng, c->Q, c->tab.Np_s, c->D))), dim3(256), 0, c->stream, a)
    ADG_DISPATCH(c, L_GH);
#undef L_GH
    HIPCHK(c, hipGetLastError());
    return ADG_OK;
}

int adg_average_down(adg_ctx* c) {
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    for (int l = c->H.nlev - 1; l > 0; --l) {
        const long ng = (long)c->H.ad_coarse[l].size();
        if (ng == 0) continue;
        adg::AvgArgs a{c->d_tab, c->to, c->Uw, c->ad_coarse[l], c->ad_fine[l], c->H.NS, ng};
#define L_AD(PP, DD) hipLaunchKernelGGL((adg::k_avgdown<PP, DD>), dim3(nblocks(adg::avg_threads(ng, c->Q, c->tab.Np_s, c->D))), dim3(256), 0, c->stream, a)
        ADG_DISPATCH(c, L_AD);
#undef L_AD
        HIPCHK(c, hipGetLastError());
    }
    return ADG_OK;
}

// The cells the CFL reduction visits: every valid cell (the reference's
// set_Dt, covered coarse cells included) or the uncovered ones.
static const int* cfl_cells(const adg_ctx* c) { return c->cfl_uncovered ? c->uncovered : c->valid_list; }
static long cfl_ncells(const adg_ctx* c) {
    return (long)(c->cfl_uncovered ? c->H.uncovered.size() : c->H.valid_list.size());
}

static int launch_cfl(adg_ctx* c, double c_dt, double T, int advance) {
    const long nv = cfl_ncells(c);
    const double cfl = (1.0 / (2.0 * (double)c->P + 1.0)) * (1.0 / (double)c->D);
    const int k = adg::cfl_per_thread(nv);
    adg::CflArgs a{c->d_tab, c->to, c->Uw, cfl_cells(c), c->slot_lev, c->ts, c->geo, c->H.NS, nv, c->gamma, cfl, k};
#define L_CFL(PP, DD) hipLaunchKernelGGL((adg::k_cfl<PP, DD>), dim3(nblocks(nv, 256 * k)), dim3(256), 0, c->stream, a)
    ADG_DISPATCH(c, L_CFL);
#undef L_CFL
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(adg::k_dt_finalize, dim3(1), dim3(64), 0, c->stream, c->ts, c_dt, T, advance);
    HIPCHK(c, hipGetLastError());
    return ADG_OK;
}

int adg_set_time(adg_ctx* c, double t) {
    if (!c) return ADG_E_ARG;
    HIPCHK(c, hipMemcpyAsync(&c->ts->t, &t, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ADG_OK;
}

int adg_get_time(adg_ctx* c, double* t_out, double* dt_out) {
    if (!c) return ADG_E_ARG;
    double h[2];
    HIPCHK(c, hipMemcpyAsync(h, &c->ts->t, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (t_out) *t_out = h[0];
    if (dt_out) *dt_out = h[1];
    return ADG_OK;
}

int adg_set_dt(adg_ctx* c, double c_dt, double T, double t, double* dt_out) {
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    int rc = adg_set_time(c, t);
    if (rc != ADG_OK) return rc;
    if ((rc = launch_cfl(c, c_dt, T, 0)) != ADG_OK) return rc;
    if (dt_out) {
        HIPCHK(c, hipMemcpyAsync(dt_out, &c->ts->Dt, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return ADG_OK;
}

int adg_set_dt_value(adg_ctx* c, double Dt) {
    if (!c) return ADG_E_ARG;
    if (!(Dt > 0.0)) return fail(c, ADG_E_ARG, "Dt must be positive");
    HIPCHK(c, hipMemcpyAsync(&c->ts->Dt, &Dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ADG_OK;
}

extern "C++" {
template <typename T, int P, int D>
static int launch_pred_mfma(adg_ctx* c, long begin, long n, int variant, bool learned, const int* slots) {
    const int v = variant == PV_NL ? 1 : 0;
    adg::PredMArgs<T> a{};
    a.frags = (const T*)c->d_frag[v];
    a.fo = c->fo[v];
    a.Uw = c->Uw;
    a.Vol = (T*)c->Vol;
    a.Tr = (T*)c->Tr;
    a.Hw_dbg = c->keep_hw ? c->Hw_dbg : nullptr;
    a.slot_lev = c->slot_lev;
    a.ts = c->ts;
    a.geo = c->geo;
    a.NS = c->H.NS;
    a.slot_begin = begin;
    a.nelem = n;
    a.gamma = c->gamma;
    a.iters = c->iters_override > 0 ? c->iters_override : P;
    a.warm = learned ? c->warm : nullptr;
    a.tol = c->learned_tol;
    a.learned_iters = c->learned_iters;
    a.stats = c->lstats;
    a.slots = slots;
    const int wpb = variant == PV_NL ? c->nl_wpb : 8;
    const size_t lds = mfma_lds_bytes(c, variant, wpb);
    const void* kern = variant == PV_NL ? (wpb == 16 ? adg::predictor_nl_kernel<T, P, D, 16>(learned)
                                                     : adg::predictor_nl_kernel<T, P, D, 8>(learned))
                                        : (learned ? adg::predictor_mfma_kernel<T, P, D, true>()
                                                   : adg::predictor_mfma_kernel<T, P, D, false>());
    const int per_cu = c->pred_per_cu[learned ? 1 : 0][variant == PV_NL ? (wpb == 16 ? 2 : 1) : 0];
    if (per_cu <= 0) return fail(c, ADG_E_UNSUP, "predictor workgroup does not fit a CU");
    const long unit = variant == PV_NL ? nl_elems(c) : adg::Dims<P, D>::EPM;   // elements per wavefront
    const long grid = persistent_grid((n + unit - 1) / unit, wpb, per_cu, c->ncu);
    void* kargs[] = {&a};
    HIPCHK(c, hipLaunchKernel(kern, dim3((unsigned)grid), dim3(64 * wpb), kargs, lds, c->stream));
    HIPCHK(c, hipGetLastError());
    return ADG_OK;
}
}  // extern "C++"

int adg_set_predictor_kernel(adg_ctx* c, int kind) {
    if (!c) return ADG_E_ARG;
    if (kind < 0 || kind > 3) return fail(c, ADG_E_ARG, "kind must be 0, 1, 2 or 3");
    if (kind == 1 && c->precision != ADG_FP64) return fail(c, ADG_E_UNSUP, "the VALU predictor is FP64 only");
    if (kind == 2 && !quad_available(c)) return fail(c, ADG_E_UNSUP, "MFMA fragments exceed LDS");
    if (kind == 3 && !nl_available(c, c->nl_wpb))
        return fail(c, ADG_E_UNSUP, "the node-lane predictor's fragments and tiles exceed the LDS");
    c->pred_kind = kind;
    drop_graph(c);
    return ADG_OK;
}

int adg_set_predictor_waves(adg_ctx* c, int wpb) {
    if (!c) return ADG_E_ARG;
    if (wpb != 8 && wpb != 16) return fail(c, ADG_E_ARG, "waves per block must be 8 or 16");
    if (c->pred_kind == 3 && !nl_available(c, wpb)) return fail(c, ADG_E_UNSUP, "node-lane tile exceeds LDS");
    c->nl_wpb = wpb;
    drop_graph(c);
    return ADG_OK;
}

int adg_set_warm_tiles(adg_ctx* c, int tiles) {
    if (!c) return ADG_E_ARG;
    if (tiles < 0 || tiles > 2) return fail(c, ADG_E_ARG, "warm-start tiles must be 0, 1 or 2");
    c->warm_tiles = tiles;
    drop_graph(c);
    return ADG_OK;
}

extern "C++" {
// Warm-start increments of the slots [begin, begin + n) from the packed MLP.
template <int P, int D, int NINP, int NOUTP>
static const void* warm_kernel(int nt) {
    if (nt == 0) return (const void*)adg::k_mlp_warm_half<P, D, NINP, NOUTP>;
    if (nt == 2) return (const void*)adg::k_mlp_warm2<P, D, NINP, NOUTP>;
    return (const void*)adg::k_mlp_warm1<P, D, NINP, NOUTP>;
}

// bytes of the warm-start constant block staged after the fragments
static int mlp_warm_const_pad(const adg_ctx* c) {
    return adg::warm_const_pad(2 * c->mlp.nin_p + 3 * c->mlp.nout_p + 2 * c->mlp.nh);
}

template <int P, int D>
static int launch_mlp_warm(adg_ctx* c, long begin, long n, const int* slots) {
    constexpr int Q = adg::Dims<P, D>::Q, Ns = adg::Dims<P, D>::Ns, Nst = adg::Dims<P, D>::Nst;
    constexpr int NINP = 16 * ((Q * Ns + 1 + D + 15) / 16), NOUTP = 16 * ((Q * Nst + 15) / 16);
    auto& m = c->mlp;
    if (m.nin_p != NINP || m.nout_p != NOUTP) return fail(c, ADG_E_STATE, "MLP widths do not match the tables");
    if (mlp_warm_const_pad(c) != adg::warm_const_pad(adg::warm_const_floats<NINP, NOUTP>()))
        return fail(c, ADG_E_STATE, "warm-start constant block does not match the kernel");
    adg::WarmArgs a{};
    a.stage = reinterpret_cast<const uint4*>(m.frag);
    a.f1 = m.f1;
    a.f2 = m.f2;
    a.f3 = m.f3;
    a.nfrag = m.nfrag;
    a.stage16 = (int)(((size_t)m.nfrag * 64 * sizeof(uint2) + mlp_warm_const_pad(c)) / 16);
    a.Uw = c->Uw;
    a.slot_lev = c->slot_lev;
    a.ts = c->ts;
    a.geo = c->geo;
    a.NS = c->H.NS;
    a.slot_begin = begin;
    a.nelem = n;
    a.gamma = c->gamma;
    a.warm = c->warm;
    a.slots = slots;
    const size_t lds = (size_t)m.nfrag * 64 * sizeof(uint2) + adg::warm_const_bytes<NINP, NOUTP>();
    const int nt = c->warm_tiles;
    const void* kern = warm_kernel<P, D, NINP, NOUTP>(nt);
    if (c->warm_per_cu[nt] <= 0) return fail(c, ADG_E_STATE, "warm-start kernel not configured");
    const long per_wave = nt == 0 ? 8 : 16 * nt;   // elements per wavefront; 8 wavefronts per workgroup
    const long grid = persistent_grid((n + per_wave - 1) / per_wave, 8, c->warm_per_cu[nt], c->ncu);
    void* kargs[] = {&a};
    HIPCHK(c, hipLaunchKernel(kern, dim3((unsigned)grid), dim3(512), kargs, lds, c->stream));
    HIPCHK(c, hipGetLastError());
    return ADG_OK;
}
}  // extern "C++"

static int mlp_pack(adg_ctx* c);

// Raises the LDS limit of the warm-start kernel (outside any capture).
extern "C++" {
template <int P, int D>
static int configure_mlp_warm(adg_ctx* c) {
    constexpr int Q = adg::Dims<P, D>::Q, Ns = adg::Dims<P, D>::Ns, Nst = adg::Dims<P, D>::Nst;
    constexpr int NINP = 16 * ((Q * Ns + 1 + D + 15) / 16), NOUTP = 16 * ((Q * Nst + 15) / 16);
    const size_t lds = (size_t)c->mlp.nfrag * 64 * sizeof(uint2) + adg::warm_const_bytes<NINP, NOUTP>();
    if (lds > kLdsMax) return fail(c, ADG_E_UNSUP, "network fragments exceed LDS");
    for (int nt = 0; nt <= 2; ++nt) {
        const void* kern = warm_kernel<P, D, NINP, NOUTP>(nt);
        HIPCHK(c, hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        int nb = 0;
        HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, 512, lds));
        c->warm_per_cu[nt] = nb > 0 ? nb : 1;
    }
    return ADG_OK;
}
}  // extern "C++"

int adg_autotune(adg_ctx* c, int reps, double* ms_out, int* choice_out) {
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    if (reps < 1) return fail(c, ADG_E_ARG, "reps must be >= 1");
    // candidates: {kind, waves per workgroup} of the predictor launch
    const int cand[4][2] = {{1, 8}, {2, 8}, {3, 8}, {3, 16}};
    const int kind0 = c->pred_kind, wpb0 = c->nl_wpb;
    double best = 1e300;
    int bi = -1;
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    for (int i = 0; i < 4; ++i) {
        double ms = -1.0;
        const int kind = cand[i][0], wpb = cand[i][1];
        const bool ok = (kind == 1 && c->precision == ADG_FP64) || (kind == 2 && quad_available(c)) ||
                        (kind == 3 && nl_available(c, wpb) && c->pred_per_cu[0][wpb == 16 ? 2 : 1] > 0 &&
                         c->pred_per_cu[1][wpb == 16 ? 2 : 1] > 0);
        if (ok) {
            c->pred_kind = kind;
            c->nl_wpb = wpb;
            int rc = adg_predictor(c, -1, 0, nullptr);   // warm-up launch
            if (rc != ADG_OK) { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); return rc; }
            HIPCHK(c, hipEventRecord(e0, c->stream));
            for (int r = 0; r < reps && rc == ADG_OK; ++r) rc = adg_predictor(c, -1, 0, nullptr);
            HIPCHK(c, hipEventRecord(e1, c->stream));
            HIPCHK(c, hipEventSynchronize(e1));
            float el = 0.f;
            HIPCHK(c, hipEventElapsedTime(&el, e0, e1));
            ms = (double)el / reps;
            if (ms < best) { best = ms; bi = i; }
        }
        if (ms_out) ms_out[i] = ms;
    }
    if (bi < 0) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        c->pred_kind = kind0;
        c->nl_wpb = wpb0;
        return fail(c, ADG_E_UNSUP, "no predictor candidate fits");
    }
    c->pred_kind = cand[bi][0];
    c->nl_wpb = cand[bi][1];
    // warm-start inference: one or two 16-element tiles, or one 8-element
    // tile, per wavefront
    if (ms_out) ms_out[4] = ms_out[5] = ms_out[6] = -1.0;
    if (c->mlp.init) {
        int rc = mlp_pack(c);
        double wbest = 1e300;
        for (int nt = 0; nt <= 2 && rc == ADG_OK; ++nt) {
            c->warm_tiles = nt;
#define L_MW(PP, DD) rc = launch_mlp_warm<PP, DD>(c, 0, (long)c->H.pred_slots.size(), c->pred_slots)
            ADG_DISPATCH(c, L_MW);
            HIPCHK(c, hipEventRecord(e0, c->stream));
            for (int r = 0; r < reps && rc == ADG_OK; ++r) ADG_DISPATCH(c, L_MW);
#undef L_MW
            HIPCHK(c, hipEventRecord(e1, c->stream));
            HIPCHK(c, hipEventSynchronize(e1));
            float el = 0.f;
            HIPCHK(c, hipEventElapsedTime(&el, e0, e1));
            const double ms = (double)el / reps;
            if (ms_out) ms_out[nt == 0 ? 6 : 3 + nt] = ms;
            if (ms < wbest) { wbest = ms; bi = nt; }
        }
        if (rc != ADG_OK) { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); return rc; }
        c->warm_tiles = bi;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    drop_graph(c);
    if (choice_out) { choice_out[0] = c->pred_kind; choice_out[1] = c->nl_wpb; choice_out[2] = c->warm_tiles; }
    return ADG_OK;
}

static int launch_predictor(adg_ctx* c, long begin, long n, int variant, bool learned);

int adg_predictor(adg_ctx* c, int lev, int mode, int* n_fallback) {
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    if (mode != 0 && mode != 1) return fail(c, ADG_E_ARG, "mode must be 0 (plain) or 1 (learned warm start)");
    if (lev < -1 || lev >= c->H.nlev) return fail(c, ADG_E_ARG, "bad level");
    if (n_fallback) *n_fallback = 0;
    // elements: the predictor's slot list (every slot but the copy ghosts)
    long begin = 0, n = (long)c->H.pred_slots.size();
    if (lev >= 0) { begin = c->H.pred_off[lev]; n = c->H.pred_off[lev + 1] - begin; }
    if (n == 0) return ADG_OK;
    const int variant = pred_variant(c);
    const bool learned = mode == 1;
    int rc = ADG_OK;
    if (learned) {
        if (variant == PV_VALU) return fail(c, ADG_E_UNSUP, "the learned warm start runs in the MFMA predictors");
        if (!c->mlp.init) return fail(c, ADG_E_STATE, "no network: adg_mlp_init / adg_mlp_set_params first");
        if ((rc = mlp_pack(c)) != ADG_OK) return rc;
        if (n_fallback) HIPCHK(c, hipMemsetAsync(c->lstats, 0, 2 * sizeof(unsigned long long), c->stream));
#define L_MW(PP, DD) rc = launch_mlp_warm<PP, DD>(c, begin, n, c->pred_slots)
        ADG_DISPATCH(c, L_MW);
#undef L_MW
        if (rc != ADG_OK) return rc;
    }
    if ((rc = launch_predictor(c, begin, n, variant, learned)) != ADG_OK) return rc;
    if (c->keep_hw && c->Hw_dbg && !c->H.copy_ghost.empty()) {
        // the kept H_w of copy ghosts from their sources, for inspection
        const long ng = (long)c->H.copy_ghost.size();
        hipLaunchKernelGGL(adg::k_copy_slots, dim3(nblocks(ng * c->Q * c->tab.Np_st)), dim3(256), 0, c->stream,
                           c->Hw_dbg, c->Q * c->tab.Np_st, c->H.NS, c->cg_slot, c->cg_src, ng);
        HIPCHK(c, hipGetLastError());
    }
    if (learned && n_fallback && variant != PV_VALU) {
        unsigned long long st[2];
        HIPCHK(c, hipMemcpyAsync(st, c->lstats, sizeof(st), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        *n_fallback = (int)st[1];
    }
    return ADG_OK;
}

static int launch_predictor(adg_ctx* c, long begin, long n, int variant, bool learned) {
    int rc = ADG_OK;
    if (variant != PV_VALU) {
        if (c->precision == ADG_FP64) {
#define L_PM(PP, DD) rc = launch_pred_mfma<double, PP, DD>(c, begin, n, variant, learned, c->pred_slots)
            ADG_DISPATCH(c, L_PM);
#undef L_PM
        } else {
#define L_PM(PP, DD) rc = launch_pred_mfma<float, PP, DD>(c, begin, n, variant, learned, c->pred_slots)
            ADG_DISPATCH(c, L_PM);
#undef L_PM
        }
        return rc;
    }
    adg::PredArgs a{c->d_tab, c->to, c->Uw, (double*)c->Vol, (double*)c->Tr, c->keep_hw ? c->Hw_dbg : nullptr,
                    c->slot_lev, c->ts, c->geo, c->H.NS, begin, n, c->gamma,
                    c->iters_override > 0 ? c->iters_override : c->P, c->pred_slots};
#define L_PR(PP, DD)                                                                              \
    {                                                                                             \
        constexpr int epw = adg::Dims<PP, DD>::EPW;                                               \
        long waves = (n + epw - 1) / epw;                                                         \
        hipLaunchKernelGGL((adg::k_predictor<PP, DD>), dim3(nblocks(waves * 64)), dim3(256), 0, c->stream, a); \
    }
    ADG_DISPATCH(c, L_PR);
#undef L_PR
    HIPCHK(c, hipGetLastError());
    return ADG_OK;
}

// ---------------------------------------------------------------------------
// Learned warm start: network state, dataset, training.
namespace {
uint64_t mix64_h(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
long mlp_norm_size(const adg_ctx* c) { return 2L * c->mlp.nin_p + 2L * c->mlp.nout_p; }
}  // namespace

static int mlp_alloc(adg_ctx* c) {
    auto& m = c->mlp;
    if (m.par) return ADG_OK;
    m.nin = c->Q * c->tab.Np_s + 1 + c->D;
    m.nin_p = 16 * ((m.nin + 15) / 16);
    m.nout = c->Q * c->tab.Np_st;
    m.nout_p = 16 * ((m.nout + 15) / 16);
    const long nh = m.nh;
    m.oW1 = 0;
    m.ob1 = m.oW1 + nh * m.nin_p;
    m.oW2 = m.ob1 + nh;
    m.ob2 = m.oW2 + nh * nh;
    m.oW3 = m.ob2 + nh;
    m.ob3 = m.oW3 + (long)m.nout_p * nh;
    m.npar = m.ob3 + m.nout_p;
    m.f1 = 0;
    m.f2 = m.f1 + (int)(nh / 16) * (m.nin_p / 16);
    m.f3 = m.f2 + (int)(nh / 16) * (int)(nh / 16);
    m.nfrag = m.f3 + (m.nout_p / 16) * (int)(nh / 16);
    for (float** p : {&m.par, &m.grad, &m.m, &m.v})
        HIPCHK(c, hipMalloc((void**)p, (size_t)m.npar * sizeof(float)));
    HIPCHK(c, hipMalloc((void**)&m.norm, (size_t)mlp_norm_size(c) * sizeof(float)));
    HIPCHK(c, hipMalloc((void**)&m.frag, (size_t)m.nfrag * 64 * sizeof(uint2) + mlp_warm_const_pad(c)));
    HIPCHK(c, hipMalloc((void**)&m.dacc, (size_t)(1 + 2 * (m.nin_p + m.nout_p)) * sizeof(double)));
    return ADG_OK;
}

static int mlp_fill(adg_ctx* c, float* p, long n, float v) {
    if (n <= 0) return ADG_OK;
    hipLaunchKernelGGL(adg::k_fill, dim3(nblocks(n)), dim3(256), 0, c->stream, p, n, v);
    HIPCHK(c, hipGetLastError());
    return ADG_OK;
}

static int mlp_pack(adg_ctx* c) {
    auto& m = c->mlp;
    if (m.frag_ok) return ADG_OK;
    const int nh = m.nh;
    // layers 2 and 3 at half scale: the inference kernel feeds them twice the GELU
    hipLaunchKernelGGL(adg::k_mlp_pack, dim3((nh / 16) * (m.nin_p / 16)), dim3(64), 0, c->stream, m.par + m.oW1, nh,
                       m.nin_p, 1.0f, m.frag + (long)m.f1 * 64);
    hipLaunchKernelGGL(adg::k_mlp_pack, dim3((nh / 16) * (nh / 16)), dim3(64), 0, c->stream, m.par + m.oW2, nh, nh,
                       0.5f, m.frag + (long)m.f2 * 64);
    hipLaunchKernelGGL(adg::k_mlp_pack, dim3((m.nout_p / 16) * (nh / 16)), dim3(64), 0, c->stream, m.par + m.oW3,
                       m.nout_p, nh, 0.5f, m.frag + (long)m.f3 * 64);
    const int npad = mlp_warm_const_pad(c) / 4;
    hipLaunchKernelGGL(adg::k_mlp_warm_consts, dim3((npad + 255) / 256), dim3(256), 0, c->stream, m.norm, m.par + m.ob1,
                       m.par + m.ob2, m.par + m.ob3, m.nin_p, m.nout_p, m.nout, nh, npad,
                       reinterpret_cast<float*>(m.frag + (long)m.nfrag * 64));
    HIPCHK(c, hipGetLastError());
    int rc = ADG_OK;
#define L_CW(PP, DD) rc = configure_mlp_warm<PP, DD>(c)
    ADG_DISPATCH(c, L_CW);
#undef L_CW
    if (rc != ADG_OK) return rc;
    m.frag_ok = true;
    return ADG_OK;
}

static int mlp_gemm(adg_ctx* c, const float* A, long lda, int tA, const float* B, long ldb, int tB, float* C, long ldc,
                    int M, int N, int K, int epi, const float* bias = nullptr, float* C2 = nullptr,
                    const float* aux = nullptr) {
    adg::GemmArgs g{A, lda, tA, B, ldb, tB, C, ldc, M, N, K, 1, epi, bias, C2, aux};
    if (epi == 2) g.ksplit = std::max(1, std::min(K / 16, 64));
    const long waves = (long)(M / 16) * (N / 16) * g.ksplit;
    hipLaunchKernelGGL(adg::k_gemm_bf16, dim3((unsigned)((waves * 64 + 255) / 256)), dim3(256), 0, c->stream, g);
    HIPCHK(c, hipGetLastError());
    return ADG_OK;
}

// Work buffers for minibatches of up to rows rows.
static int mlp_work(adg_ctx* c, long rows) {
    auto& m = c->mlp;
    if (m.work_rows >= rows) return ADG_OK;
    if (m.work) (void)hipFree(m.work);
    m.work = nullptr;
    const long per = 2L * m.nin_p + 3L * m.nout_p + 6L * m.nh;
    HIPCHK(c, hipMalloc((void**)&m.work, (size_t)rows * per * sizeof(float)));
    m.work_rows = rows;
    return ADG_OK;
}
struct MlpWork {
    float *Xb, *Yb, *A1, *H1, *A2, *H2, *YP, *G3, *D2, *D1;
};
static MlpWork mlp_views(adg_ctx* c) {
    auto& m = c->mlp;
    const long R = m.work_rows;
    MlpWork w;
    float* p = m.work;
    w.Xb = p; p += R * m.nin_p;
    w.Yb = p; p += R * m.nout_p;
    w.YP = p; p += R * m.nout_p;
    w.G3 = p; p += R * m.nout_p;
    w.A1 = p; p += R * m.nh;
    w.H1 = p; p += R * m.nh;
    w.A2 = p; p += R * m.nh;
    w.H2 = p; p += R * m.nh;
    w.D2 = p; p += R * m.nh;
    w.D1 = p; p += R * m.nh;
    return w;
}

// Forward and backward pass over the B rows in Xb/Yb: gradient into
// m.grad (zeroed first), summed loss added to dacc[0].
static int mlp_fwd_bwd(adg_ctx* c, long B) {
    auto& m = c->mlp;
    const MlpWork w = mlp_views(c);
    const int nh = m.nh, Bi = (int)B;
    int rc;
    float* P = m.par;
    float* G = m.grad;
    if ((rc = mlp_fill(c, G, m.npar, 0.f)) != ADG_OK) return rc;
    // forward: A1 = Xb W1^T + b1, H1 = gelu; A2 = H1 W2^T + b2, H2 = gelu; YP = H2 W3^T + b3
    if ((rc = mlp_gemm(c, w.Xb, m.nin_p, 0, P + m.oW1, m.nin_p, 1, w.A1, nh, Bi, nh, m.nin_p, 1, P + m.ob1, w.H1)))
        return rc;
    if ((rc = mlp_gemm(c, w.H1, nh, 0, P + m.oW2, nh, 1, w.A2, nh, Bi, nh, nh, 1, P + m.ob2, w.H2))) return rc;
    if ((rc = mlp_gemm(c, w.H2, nh, 0, P + m.oW3, nh, 1, w.YP, m.nout_p, Bi, m.nout_p, nh, 1, P + m.ob3, nullptr)))
        return rc;
    hipLaunchKernelGGL(adg::k_mlp_lossgrad, dim3(nblocks(B * m.nout_p)), dim3(256), 0, c->stream, w.YP, w.Yb, w.G3,
                       B, m.nout_p, m.nout, m.dacc);
    HIPCHK(c, hipGetLastError());
    auto colsum = [&](const float* Gm, int ncol, float* out) -> int {
        hipLaunchKernelGGL(adg::k_colsum, dim3((unsigned)((ncol + 255) / 256), (unsigned)((B + 255) / 256)), dim3(256),
                           0, c->stream, Gm, B, ncol, out);
        HIPCHK(c, hipGetLastError());
        return ADG_OK;
    };
    // layer 3: dW3 = G3^T H2, db3 = sum G3, D2 = (G3 W3) * gelu'(A2)
    if ((rc = mlp_gemm(c, w.G3, m.nout_p, 1, w.H2, nh, 0, G + m.oW3, nh, m.nout_p, nh, Bi, 2))) return rc;
    if ((rc = colsum(w.G3, m.nout_p, G + m.ob3))) return rc;
    if ((rc = mlp_gemm(c, w.G3, m.nout_p, 0, P + m.oW3, nh, 0, w.D2, nh, Bi, nh, m.nout_p, 3, nullptr, nullptr, w.A2)))
        return rc;
    // layer 2
    if ((rc = mlp_gemm(c, w.D2, nh, 1, w.H1, nh, 0, G + m.oW2, nh, nh, nh, Bi, 2))) return rc;
    if ((rc = colsum(w.D2, nh, G + m.ob2))) return rc;
    if ((rc = mlp_gemm(c, w.D2, nh, 0, P + m.oW2, nh, 0, w.D1, nh, Bi, nh, nh, 3, nullptr, nullptr, w.A1))) return rc;
    // layer 1
    if ((rc = mlp_gemm(c, w.D1, nh, 1, w.Xb, m.nin_p, 0, G + m.oW1, m.nin_p, nh, m.nin_p, Bi, 2))) return rc;
    return colsum(w.D1, nh, G + m.ob1);
}

extern "C" {

int adg_mlp_sizes(adg_ctx* c, int* out) {
    if (!c || !out) return ADG_E_ARG;
    int rc = mlp_alloc(c);
    if (rc != ADG_OK) return rc;
    out[0] = c->mlp.nin;
    out[1] = c->mlp.nin_p;
    out[2] = c->mlp.nh;
    out[3] = c->mlp.nout;
    out[4] = c->mlp.nout_p;
    out[5] = (int)c->mlp.npar;
    out[6] = (int)mlp_norm_size(c);
    return ADG_OK;
}

int adg_mlp_init(adg_ctx* c, unsigned long long seed) {
    if (!c) return ADG_E_ARG;
    int rc = mlp_alloc(c);
    if (rc != ADG_OK) return rc;
    auto& m = c->mlp;
    std::vector<float> par((size_t)m.npar, 0.f);
    uint64_t st = mix64_h(seed ^ 0x6d6c70ull);
    auto uni = [&]() {
        st = mix64_h(st);
        return (double)(st >> 11) * (1.0 / 9007199254740992.0);
    };
    // Glorot-uniform weights over the logical widths, zero biases; padded
    // input columns and output rows stay zero
    auto init = [&](long off, int rows, int cols, int rows_used, int cols_used) {
        const double lim = std::sqrt(6.0 / (double)(rows_used + cols_used));
        for (int i = 0; i < rows; ++i)
            for (int j = 0; j < cols; ++j)
                par[(size_t)off + (size_t)i * cols + j] =
                    (i < rows_used && j < cols_used) ? (float)((2.0 * uni() - 1.0) * lim) : 0.f;
    };
    init(m.oW1, m.nh, m.nin_p, m.nh, m.nin);
    init(m.oW2, m.nh, m.nh, m.nh, m.nh);
    init(m.oW3, m.nout_p, m.nh, m.nout, m.nh);
    HIPCHK(c, hipMemcpy(m.par, par.data(), par.size() * sizeof(float), hipMemcpyHostToDevice));
    std::vector<float> norm((size_t)mlp_norm_size(c), 0.f);
    for (int j = 0; j < m.nin_p; ++j) norm[(size_t)m.nin_p + j] = 1.f;
    for (int j = 0; j < m.nout_p; ++j) norm[(size_t)2 * m.nin_p + m.nout_p + j] = 1.f;
    HIPCHK(c, hipMemcpy(m.norm, norm.data(), norm.size() * sizeof(float), hipMemcpyHostToDevice));
    if ((rc = mlp_fill(c, m.m, m.npar, 0.f)) != ADG_OK) return rc;
    if ((rc = mlp_fill(c, m.v, m.npar, 0.f)) != ADG_OK) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    m.adam_t = 0;
    m.frag_ok = false;
    m.init = true;
    drop_graph(c);
    return ADG_OK;
}

int adg_mlp_get_params(adg_ctx* c, float* out, long cap) {
    if (!c || !out || !c->mlp.init) return fail(c, ADG_E_STATE, "no network");
    auto& m = c->mlp;
    const long tot = m.npar + mlp_norm_size(c);
    if (cap < tot) return fail(c, ADG_E_ARG, "buffer too small");
    HIPCHK(c, hipMemcpy(out, m.par, (size_t)m.npar * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(out + m.npar, m.norm, (size_t)mlp_norm_size(c) * sizeof(float), hipMemcpyDeviceToHost));
    return ADG_OK;
}

int adg_mlp_set_params(adg_ctx* c, const float* in, long n) {
    if (!c || !in) return ADG_E_ARG;
    int rc = mlp_alloc(c);
    if (rc != ADG_OK) return rc;
    auto& m = c->mlp;
    if (n != m.npar + mlp_norm_size(c)) return fail(c, ADG_E_ARG, "parameter count does not match the network");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(m.par, in, (size_t)m.npar * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(m.norm, in + m.npar, (size_t)mlp_norm_size(c) * sizeof(float), hipMemcpyHostToDevice));
    if (!m.init) {
        if ((rc = mlp_fill(c, m.m, m.npar, 0.f)) != ADG_OK) return rc;
        if ((rc = mlp_fill(c, m.v, m.npar, 0.f)) != ADG_OK) return rc;
        m.adam_t = 0;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    m.frag_ok = false;
    m.init = true;
    drop_graph(c);
    return ADG_OK;
}

int adg_mlp_collect(adg_ctx* c, int iters, long* n_total) {
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    if (iters < 1) return fail(c, ADG_E_ARG, "iters must be >= 1");
    int rc = mlp_alloc(c);
    if (rc != ADG_OK) return rc;
    auto& m = c->mlp;
    const long nv = (long)c->H.valid_list.size();
    if (m.ndata + nv > m.cap) {
        const long cap = std::max(2 * m.cap, m.ndata + nv);
        float *X = nullptr, *Y = nullptr;
        HIPCHK(c, hipMalloc((void**)&X, (size_t)cap * m.nin_p * sizeof(float)));
        HIPCHK(c, hipMalloc((void**)&Y, (size_t)cap * m.nout_p * sizeof(float)));
        if (m.ndata) {
            HIPCHK(c, hipMemcpyAsync(X, m.X, (size_t)m.ndata * m.nin_p * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(Y, m.Y, (size_t)m.ndata * m.nout_p * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (m.X) (void)hipFree(m.X);
        if (m.Y) (void)hipFree(m.Y);
        m.X = X;
        m.Y = Y;
        m.cap = cap;
    }
    // converged predictor with H_w kept
    const bool keep = c->keep_hw;
    if ((rc = adg_keep_Hw(c, 1)) != ADG_OK) return rc;
    c->iters_override = iters;
    rc = adg_predictor(c, -1, 0, nullptr);
    c->iters_override = 0;
    if (rc != ADG_OK) return rc;
    adg::CollectArgs a{c->d_tab, c->to, c->Uw, c->Hw_dbg, c->valid_list, c->slot_lev, c->ts, c->geo, c->H.NS, nv,
                       c->gamma, m.X + m.ndata * m.nin_p, m.Y + m.ndata * m.nout_p, m.nin_p, m.nout_p};
#define L_CO(PP, DD) \
    hipLaunchKernelGGL((adg::k_mlp_collect<PP, DD>), dim3((unsigned)((nv + 63) / 64)), dim3(64), 0, c->stream, a)
    ADG_DISPATCH(c, L_CO);
#undef L_CO
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->keep_hw = keep;
    m.ndata += nv;
    if (n_total) *n_total = m.ndata;
    return ADG_OK;
}

int adg_mlp_dataset(adg_ctx* c, long start, long n, float* x, float* y) {
    if (!c) return ADG_E_ARG;
    auto& m = c->mlp;
    if (start < 0 || n < 0 || start + n > m.ndata) return fail(c, ADG_E_ARG, "rows out of range");
    if (x) HIPCHK(c, hipMemcpy(x, m.X + start * m.nin_p, (size_t)n * m.nin_p * sizeof(float), hipMemcpyDeviceToHost));
    if (y) HIPCHK(c, hipMemcpy(y, m.Y + start * m.nout_p, (size_t)n * m.nout_p * sizeof(float), hipMemcpyDeviceToHost));
    return ADG_OK;
}

int adg_mlp_dataset_clear(adg_ctx* c) {
    if (!c) return ADG_E_ARG;
    c->mlp.ndata = 0;
    return ADG_OK;
}

int adg_mlp_fit_normalization(adg_ctx* c) {
    if (!c || !c->mlp.init) return fail(c, ADG_E_STATE, "no network");
    auto& m = c->mlp;
    if (m.ndata < 2) return fail(c, ADG_E_STATE, "dataset has fewer than 2 rows");
    const int nc = m.nin_p + m.nout_p;
    double* sum = m.dacc + 1;
    double* sq = sum + nc;
    HIPCHK(c, hipMemsetAsync(sum, 0, 2 * (size_t)nc * sizeof(double), c->stream));
    hipLaunchKernelGGL(adg::k_colstats, dim3((unsigned)((m.nin_p + 255) / 256), (unsigned)((m.ndata + 1023) / 1024)),
                       dim3(256), 0, c->stream, m.X, m.ndata, m.nin_p, sum, sq);
    hipLaunchKernelGGL(adg::k_colstats, dim3((unsigned)((m.nout_p + 255) / 256), (unsigned)((m.ndata + 1023) / 1024)),
                       dim3(256), 0, c->stream, m.Y, m.ndata, m.nout_p, sum + m.nin_p, sq + m.nin_p);
    HIPCHK(c, hipGetLastError());
    std::vector<double> hs(2 * (size_t)nc);
    HIPCHK(c, hipMemcpyAsync(hs.data(), sum, hs.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<float> norm((size_t)mlp_norm_size(c));
    auto stat = [&](int col, float* mu, float* sd) {
        const double mean = hs[col] / (double)m.ndata;
        const double var = std::max(0.0, hs[nc + col] / (double)m.ndata - mean * mean);
        const double dev = std::sqrt(var);
        *mu = (float)mean;
        *sd = dev > 1e-6 * std::max(1.0, std::fabs(mean)) ? (float)dev : 1.f;   // constant column
    };
    for (int j = 0; j < m.nin_p; ++j) stat(j, &norm[j], &norm[(size_t)m.nin_p + j]);
    for (int j = 0; j < m.nout_p; ++j)
        stat(m.nin_p + j, &norm[(size_t)2 * m.nin_p + j], &norm[(size_t)2 * m.nin_p + m.nout_p + j]);
    HIPCHK(c, hipMemcpy(m.norm, norm.data(), norm.size() * sizeof(float), hipMemcpyHostToDevice));
    m.frag_ok = false;
    return ADG_OK;
}

int adg_mlp_train(adg_ctx* c, int nsteps, int batch, double lr, unsigned long long seed, double* loss_out) {
    if (!c || !c->mlp.init) return fail(c, ADG_E_STATE, "no network");
    auto& m = c->mlp;
    if (m.ndata < 1) return fail(c, ADG_E_STATE, "empty dataset");
    if (nsteps < 1 || batch < 16 || batch % 16) return fail(c, ADG_E_ARG, "batch must be a positive multiple of 16");
    int rc = mlp_work(c, batch);
    if (rc != ADG_OK) return rc;
    const MlpWork w = mlp_views(c);
    const int tail = std::min(nsteps, 20);
    const float b1 = 0.9f, b2 = 0.999f;
    for (int it = 0; it < nsteps; ++it) {
        if (it == nsteps - tail) HIPCHK(c, hipMemsetAsync(m.dacc, 0, sizeof(double), c->stream));
        hipLaunchKernelGGL(adg::k_gather_batch, dim3(nblocks((long)batch * (m.nin_p + m.nout_p))), dim3(256), 0,
                           c->stream, m.X, m.Y, m.ndata, m.nin_p, m.nout_p, m.norm, (uint64_t)seed, m.adam_t,
                           (long)batch, w.Xb, w.Yb);
        HIPCHK(c, hipGetLastError());
        if ((rc = mlp_fwd_bwd(c, batch)) != ADG_OK) return rc;
        m.adam_t += 1;
        const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, (double)m.adam_t)));
        const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, (double)m.adam_t)));
        hipLaunchKernelGGL(adg::k_adam, dim3(nblocks(m.npar)), dim3(256), 0, c->stream, m.par, m.grad, m.m, m.v, m.npar,
                           (float)lr, b1, b2, 1e-8f, c1, c2);
        HIPCHK(c, hipGetLastError());
    }
    double l = 0.0;
    HIPCHK(c, hipMemcpyAsync(&l, m.dacc, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (loss_out) *loss_out = l / tail;
    m.frag_ok = false;
    drop_graph(c);
    return ADG_OK;
}

int adg_mlp_loss_grad(adg_ctx* c, int n, const float* x, const float* y, double* loss, float* grad) {
    if (!c || !c->mlp.init || !x || !y) return fail(c, ADG_E_STATE, "no network");
    if (n < 16 || n % 16) return fail(c, ADG_E_ARG, "n must be a positive multiple of 16");
    auto& m = c->mlp;
    int rc = mlp_work(c, n);
    if (rc != ADG_OK) return rc;
    const MlpWork w = mlp_views(c);
    HIPCHK(c, hipMemcpyAsync(w.Xb, x, (size_t)n * m.nin_p * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(w.Yb, y, (size_t)n * m.nout_p * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(m.dacc, 0, sizeof(double), c->stream));
    if ((rc = mlp_fwd_bwd(c, n)) != ADG_OK) return rc;
    double l = 0.0;
    HIPCHK(c, hipMemcpyAsync(&l, m.dacc, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (grad) HIPCHK(c, hipMemcpyAsync(grad, m.grad, (size_t)m.npar * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (loss) *loss = l;
    return ADG_OK;
}

int adg_mlp_warm_start(adg_ctx* c, int lev, double* host) {
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    if (!c->mlp.init) return fail(c, ADG_E_STATE, "no network");
    if (lev < 0 || lev >= c->H.nlev || !host) return fail(c, ADG_E_ARG, "bad arguments");
    int rc = mlp_pack(c);
    if (rc != ADG_OK) return rc;
    const long begin = c->H.lev[lev].slot_off, n = c->H.lev[lev].nslot;
#define L_MW(PP, DD) rc = launch_mlp_warm<PP, DD>(c, begin, n, nullptr)
    ADG_DISPATCH(c, L_MW);
#undef L_MW
    if (rc != ADG_OK) return rc;
    const long NS = c->H.NS, rows = c->mlp.nout;
    std::vector<float> tmp((size_t)rows * NS);
    HIPCHK(c, hipMemcpyAsync(tmp.data(), c->warm, tmp.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (long o = 0; o < rows; ++o)
        for (long s = 0; s < n; ++s) host[o * n + s] = tmp[(size_t)o * NS + begin + s];
    return ADG_OK;
}

int adg_set_learned(adg_ctx* c, int on, double tol, int max_iters) {
    if (!c) return ADG_E_ARG;
    if (on < 0 || on > 2) return fail(c, ADG_E_ARG, "on must be 0, 1 or 2");
    if (on == 1 && !(tol > 0.0)) return fail(c, ADG_E_ARG, "tol must be positive");
    if (on && (max_iters < 1 || max_iters > 8)) return fail(c, ADG_E_ARG, "max_iters must be 1..8");
    c->learned = on != 0;
    if (on) {
        c->learned_tol = on == 2 ? -1.f : (float)tol;
        c->learned_iters = max_iters;
    }
    drop_graph(c);
    return ADG_OK;
}

int adg_learned_stats(adg_ctx* c, unsigned long long* out, int reset) {
    if (!c) return ADG_E_ARG;
    if (out) {
        HIPCHK(c, hipMemcpyAsync(out, c->lstats, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (reset) {
        HIPCHK(c, hipMemsetAsync(c->lstats, 0, 2 * sizeof(unsigned long long), c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return ADG_OK;
}

}  // extern "C"

extern "C++" {
template <typename T>
static int launch_faceflux(adg_ctx* c) {
    adg::FaceArgs<T> a{};
    a.Tr = (const T*)c->Tr;   // the faces' trace slots: copy ghosts replaced by their sources
    long tot = 0;
    for (int d = 0; d < c->D; ++d) {
        a.Fnum[d] = (T*)c->Fnum[d];
        a.face_lr[d] = c->face_tr[d];
        a.NF[d] = c->H.NF[d];
        tot += c->H.NF[d];
    }
    a.NS = c->H.NS;
    a.gamma = c->gamma;
    tot *= c->tab.qMp_st_bd;   // one thread per (face, node)
#define L_FF(PP, DD) hipLaunchKernelGGL((adg::k_faceflux<T, PP, DD>), dim3(nblocks(tot)), dim3(256), 0, c->stream, a)
    ADG_DISPATCH(c, L_FF);
#undef L_FF
    HIPCHK(c, hipGetLastError());
    return ADG_OK;
}
}  // extern "C++"

int adg_face_flux(adg_ctx* c, int lev) {
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    (void)lev;  // faces of every level are processed in one launch
    return c->precision == ADG_FP64 ? launch_faceflux<double>(c) : launch_faceflux<float>(c);
}

extern "C++" {
template <typename T>
static int launch_corrector(adg_ctx* c, long begin, long n) {
    adg::CorrArgs<T> a{};
    a.tab = c->d_tab;
    a.to = c->to;
    a.Uw = c->Uw;
    a.Vol = (const T*)c->Vol;
    for (int d = 0; d < c->D; ++d) {
        a.Fnum[d] = (const T*)c->Fnum[d];
        a.cell_face[d][0] = c->cell_face[d][0] + begin;
        a.cell_face[d][1] = c->cell_face[d][1] + begin;
        a.NF[d] = c->H.NF[d];
    }
    a.valid_list = c->valid_list + begin;
    a.slot_lev = c->slot_lev;
    a.ts = c->ts;
    a.geo = c->geo;
    a.NS = c->H.NS;
    a.nvalid = n;
    if (c->H.NS >= (1L << 29)) return fail(c, ADG_E_UNSUP, "the corrector addresses at most 2^29 slots");
    for (int d = 0; d < c->D; ++d)
        if (c->H.NF[d] >= (1L << 29)) return fail(c, ADG_E_UNSUP, "the corrector addresses at most 2^29 faces");
#define L_CO(PP, DD) \
    hipLaunchKernelGGL((adg::k_corrector<T, PP, DD>), dim3(nblocks(adg::pad_wave(n) * c->Q)), dim3(256), 0, c->stream, a)
    ADG_DISPATCH(c, L_CO);
#undef L_CO
    HIPCHK(c, hipGetLastError());
    return ADG_OK;
}
}  // extern "C++"

int adg_corrector(adg_ctx* c, int lev) {
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    long begin = 0, n = (long)c->H.valid_list.size();
    if (lev >= 0) {
        if (lev >= c->H.nlev) return fail(c, ADG_E_ARG, "bad level");
        begin = c->H.lev[lev].valid_begin;
        n = c->H.lev[lev].valid_end - begin;
    }
    if (n == 0) return ADG_OK;
    return c->precision == ADG_FP64 ? launch_corrector<double>(c, begin, n) : launch_corrector<float>(c, begin, n);
}

extern "C++" {
template <typename T>
static int launch_moments(adg_ctx* c, long NM) {
    adg::MomArgs<T> m{};
    m.tab = c->d_tab;
    m.to = c->to;
    for (int d = 0; d < c->D; ++d) {
        m.Fnum[d] = (const T*)c->Fnum[d];
        m.NF[d] = c->H.NF[d];
    }
    m.Mom = c->Mom;
    m.mom_face = c->mom_face;
    m.mom_d = c->mom_d;
    m.mom_kind = c->mom_kind;
    m.mom_sel = c->mom_sel;
    m.mom_lev = c->mom_lev;
    m.ts = c->ts;
    m.geo = c->geo;
    m.NM = NM;
#define L_MO(PP, DD) hipLaunchKernelGGL((adg::k_moments<T, PP, DD>), dim3(nblocks(NM * c->Q * c->tab.Np_s)), dim3(256), 0, c->stream, m)
    ADG_DISPATCH(c, L_MO);
#undef L_MO
    HIPCHK(c, hipGetLastError());
    return ADG_OK;
}
}  // extern "C++"

int adg_reflux(adg_ctx* c, int crse_lev) {
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    (void)crse_lev;  // every coarse level in one launch (disjoint cells)
    const long NM = (long)c->H.mom_face.size(), nrc = (long)c->H.rc_slot.size();
    if (NM == 0 || nrc == 0) return ADG_OK;
    int rc = c->precision == ADG_FP64 ? launch_moments<double>(c, NM) : launch_moments<float>(c, NM);
    if (rc != ADG_OK) return rc;
    adg::RefluxArgs r{c->d_tab, c->to, c->Uw, c->Mom, c->rc_slot, c->rc_start, c->rc_count, c->rc_lev,
                      c->re_mom, c->re_coef, c->geo, c->H.NS, NM, nrc};
#define L_RF(PP, DD) hipLaunchKernelGGL((adg::k_reflux<PP, DD>), dim3(nblocks(nrc * c->Q * c->tab.Np_s)), dim3(256), 0, c->stream, r)
    ADG_DISPATCH(c, L_RF);
#undef L_RF
    HIPCHK(c, hipGetLastError());
    return ADG_OK;
}

extern "C++" {
template <typename T>
static adg::EpiArgs<T> epi_args(adg_ctx* c, double c_dt, double T_end) {
    adg::EpiArgs<T> a{};
    a.tab = c->d_tab;
    a.to = c->to;
    a.Uw = c->Uw;
    for (int d = 0; d < c->D; ++d) {
        a.Fnum[d] = (const T*)c->Fnum[d];
        a.NF[d] = c->H.NF[d];
    }
    a.mom_face = c->mom_face; a.mom_d = c->mom_d; a.mom_kind = c->mom_kind; a.mom_sel = c->mom_sel;
    a.mom_lev = c->mom_lev; a.rc_slot = c->rc_slot; a.rc_start = c->rc_start; a.rc_count = c->rc_count;
    a.rc_lev = c->rc_lev; a.re_mom = c->re_mom; a.re_coef = c->re_coef;
    a.nrc = (long)c->H.rc_slot.size();
    a.g_slot = c->g_slot; a.g_src = c->g_src; a.g_kind = c->g_kind; a.g_child = c->g_child;
    a.nghost = (long)c->H.g_slot.size();
    a.valid_list = cfl_cells(c);
    a.slot_lev = c->slot_lev;
    a.nvalid = cfl_ncells(c);
    a.cfl_k = adg::cfl_per_thread(a.nvalid);
    a.ncfl_blocks = nblocks(a.nvalid, 256 * a.cfl_k);
    a.cfl = (1.0 / (2.0 * (double)c->P + 1.0)) * (1.0 / (double)c->D);
    a.c_dt = c_dt;
    a.Tend = T_end;
    a.ts = c->ts;
    a.geo = c->geo;
    a.NS = c->H.NS;
    a.gamma = c->gamma;
    return a;
}

// reflux || finest average down, coarser average downs, ghosts || CFL + advance
template <typename T>
static int launch_epilogue(adg_ctx* c, double c_dt, double T_end) {
    const int L = c->H.nlev;
    adg::EpiArgs<T> a = epi_args<T>(c, c_dt, T_end);
    if (L > 1) {
        a.ad_coarse = c->ad_coarse[L - 1];
        a.ad_fine = c->ad_fine[L - 1];
        a.ngroup = (long)c->H.ad_coarse[L - 1].size();
        a.nblk_a = nblocks(adg::reflux_threads(a.nrc, c->Q, c->tab.Np_s, c->D));
        const unsigned nb = a.nblk_a + nblocks(adg::avg_threads(a.ngroup, c->Q, c->tab.Np_s, c->D));
        if (nb > 0) {
#define L_RA(PP, DD) hipLaunchKernelGGL((adg::k_reflux_avg<T, PP, DD>), dim3(nb), dim3(256), 0, c->stream, a)
            ADG_DISPATCH(c, L_RA);
#undef L_RA
            HIPCHK(c, hipGetLastError());
        }
        for (int l = L - 2; l > 0; --l) {
            const long ng = (long)c->H.ad_coarse[l].size();
            if (ng == 0) continue;
            adg::AvgArgs g{c->d_tab, c->to, c->Uw, c->ad_coarse[l], c->ad_fine[l], c->H.NS, ng};
#define L_AD(PP, DD) hipLaunchKernelGGL((adg::k_avgdown<PP, DD>), dim3(nblocks(adg::avg_threads(ng, c->Q, c->tab.Np_s, c->D))), dim3(256), 0, c->stream, g)
            ADG_DISPATCH(c, L_AD);
#undef L_AD
            HIPCHK(c, hipGetLastError());
        }
    }
    a.nblk_a = nblocks(adg::ghost_threads(a.nghost, c->Q, c->tab.Np_s, c->D));
    const unsigned nb = a.nblk_a + a.ncfl_blocks;
#define L_GC(PP, DD) hipLaunchKernelGGL((adg::k_ghost_cfl<T, PP, DD>), dim3(nb), dim3(256), 0, c->stream, a)
    ADG_DISPATCH(c, L_GC);
#undef L_GC
    HIPCHK(c, hipGetLastError());
    return ADG_OK;
}
}  // extern "C++"

static int launch_limiter(adg_ctx* c, bool every_step) {
    if (c->lim_n == 0) return ADG_OK;
    const adg::LevelLayout& fl = c->H.lev[c->H.nlev - 1];
    double h = fl.dx[0];
    for (int d = 1; d < c->D; ++d) h = std::min(h, fl.dx[d]);
    adg::LimArgs a{};
    a.Uw = c->Uw;
    a.cells = c->lim_cells;
    a.nbr = c->lim_nbr;
    a.count = c->lim_count;
    a.ts = c->ts;
    a.NS = c->H.NS;
    a.n = c->lim_n;
    a.thr = c->lim_tvb_c[c->H.nlev - 1] * c->lim_M * std::pow(h, 2.0);
    a.gamma = c->gamma;
    a.t_limit = every_step ? 1 : c->lim_t_limit;
    for (int d = 0; d < c->D; ++d)
        for (int n = 0; n < c->tab.Np_s; ++n) {
            bool lin = c->tab.mi_s[n][d] == 1, others = true;
            for (int e = 0; e < c->D; ++e)
                if (e != d && c->tab.mi_s[n][e] != 0) others = false;
            if (lin && others) { a.lin[d] = n; break; }
        }
#define L_LIM(PP, DD) hipLaunchKernelGGL((adg::k_limiter_tvb<PP, DD>), dim3(nblocks(c->lim_n)), dim3(256), 0, c->stream, a)
    ADG_DISPATCH(c, L_LIM);
#undef L_LIM
    HIPCHK(c, hipGetLastError());
    return ADG_OK;
}

static int step_launches(adg_ctx* c, double c_dt, double T_end) {
    int rc;
    if ((rc = adg_predictor(c, -1, c->learned ? 1 : 0, nullptr)) != ADG_OK) return rc;
    if ((rc = adg_face_flux(c, -1)) != ADG_OK) return rc;
    if ((rc = adg_corrector(c, -1)) != ADG_OK) return rc;
    // AmrDG::evolve: the limiter on the finest level after the ADER step,
    // before the average down (AmrDG.h:990-1000)
    if (c->lim_on && (rc = launch_limiter(c, false)) != ADG_OK) return rc;
    return c->precision == ADG_FP64 ? launch_epilogue<double>(c, c_dt, T_end) : launch_epilogue<float>(c, c_dt, T_end);
}

int adg_step_graph(adg_ctx* c, double c_dt, double T_end, int nsteps, double* t_out) {
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    if (nsteps < 0) return fail(c, ADG_E_ARG, "nsteps must be >= 0");
    const int kind = (pred_variant(c) * 32 + c->nl_wpb) * 2 + (c->learned ? 1 : 0);   // baked into the graph
    if (!c->graph_exec || c->graph_cdt != c_dt || c->graph_T != T_end || c->graph_kind != kind ||
        c->graph_captured != c->graph_unroll) {
        drop_graph(c);
        if (c->learned) {
            if (!c->mlp.init) return fail(c, ADG_E_STATE, "learned mode without a network");
            int prc = mlp_pack(c);
            if (prc != ADG_OK) return prc;
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = ADG_OK;
        for (int k = 0; k < c->graph_unroll && rc == ADG_OK; ++k) rc = step_launches(c, c_dt, T_end);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc != ADG_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return fail(c, ADG_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        c->graph = g;
        HIPCHK(c, hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
        c->graph_cdt = c_dt;
        c->graph_T = T_end;
        c->graph_kind = kind;
        c->graph_captured = c->graph_unroll;
    }
    // whole graphs, then the remaining steps as the same launches on the stream
    const int K = c->graph_captured;
    int i = 0;
    for (; i + K <= nsteps; i += K) HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
    for (; i < nsteps; ++i) {
        int rc = step_launches(c, c_dt, T_end);
        if (rc != ADG_OK) return rc;
    }
    if (t_out) {
        HIPCHK(c, hipMemcpyAsync(t_out, &c->ts->t, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return ADG_OK;
}

int adg_set_graph_unroll(adg_ctx* c, int steps) {
    if (!c) return ADG_E_ARG;
    if (steps < 1 || steps > 64) return fail(c, ADG_E_ARG, "steps per graph must be 1..64");
    c->graph_unroll = steps;
    return ADG_OK;
}

int adg_steps(adg_ctx* c, double c_dt, double T_end, int nsteps, double* t_out) {
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    if (nsteps < 0) return fail(c, ADG_E_ARG, "nsteps must be >= 0");
    for (int i = 0; i < nsteps; ++i) {
        int rc = step_launches(c, c_dt, T_end);
        if (rc != ADG_OK) return rc;
    }
    if (t_out) {
        HIPCHK(c, hipMemcpyAsync(t_out, &c->ts->t, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return ADG_OK;
}

int adg_ader(adg_ctx* c) {
    int rc;
    if ((rc = adg_predictor(c, -1, c->learned ? 1 : 0, nullptr)) != ADG_OK) return rc;
    if ((rc = adg_face_flux(c, -1)) != ADG_OK) return rc;
    if ((rc = adg_corrector(c, -1)) != ADG_OK) return rc;
    if (c->H.nlev > 1 && (rc = adg_reflux(c, -1)) != ADG_OK) return rc;
    return ADG_OK;
}

int adg_advance(adg_ctx* c, double c_dt, double T, double* t_out) {
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    int rc;
    if ((rc = launch_cfl(c, c_dt, T, 1)) != ADG_OK) return rc;
    if (t_out) {
        HIPCHK(c, hipMemcpyAsync(t_out, &c->ts->t, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return ADG_OK;
}

int adg_step(adg_ctx* c, double c_dt, double T, double* t_out) {
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    int rc;
    if ((rc = step_launches(c, c_dt, T)) != ADG_OK) return rc;
    if (t_out) {
        HIPCHK(c, hipMemcpyAsync(t_out, &c->ts->t, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return ADG_OK;
}

int adg_conserved_totals(adg_ctx* c, double* totals) {
    if (!c || !c->levels_set || !totals) return fail(c, ADG_E_STATE, "levels not set");
    const long n = (long)c->H.uncovered.size();
    const unsigned nb = nblocks(n);
    adg::TotArgs a{c->Uw, c->uncovered, c->slot_lev, c->partial, c->geo, c->H.NS, n, c->tab.Np_s, c->Q};
    hipLaunchKernelGGL(adg::k_totals, dim3(nb), dim3(256), 0, c->stream, a);
    HIPCHK(c, hipGetLastError());
    std::vector<double> h((size_t)c->Q * nb);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->partial, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int q = 0; q < c->Q; ++q) {
        double s = 0.0;
        for (unsigned b = 0; b < nb; ++b) s += h[(size_t)q * nb + b];
        totals[q] = s;
    }
    return ADG_OK;
}

int adg_set_limiter(adg_ctx* c, int type, double M, int ntvb, const double* tvb_c, int t_limit) {
    if (!c) return ADG_E_ARG;
    if (type != 0 && type != 1) return fail(c, ADG_E_ARG, "limiter type must be 0 (none) or 1 (TVB)");
    if (type == 1) {
        if (!(M >= 0.0)) return fail(c, ADG_E_ARG, "TVB M must be >= 0");
        if (t_limit < 1) return fail(c, ADG_E_ARG, "t_limit must be >= 1");
        if (ntvb < 0 || ntvb > adg::kMaxLev || (ntvb > 0 && !tvb_c)) return fail(c, ADG_E_ARG, "bad AMR_TVB_C");
        for (int l = 0; l < adg::kMaxLev; ++l) c->lim_tvb_c[l] = l < ntvb ? tvb_c[l] : 1.0;
        c->lim_M = M;
        c->lim_t_limit = t_limit;
    }
    c->lim_on = type == 1;
    drop_graph(c);
    // the step counter n of the limiter's n % t_limit starts again at 0
    const unsigned int z = 0u;
    HIPCHK(c, hipMemcpyAsync(&c->ts->nstep, &z, sizeof(z), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ADG_OK;
}

int adg_set_step_count(adg_ctx* c, int n) {
    if (!c) return ADG_E_ARG;
    if (n < 0) return fail(c, ADG_E_ARG, "step count must be >= 0");
    const unsigned int v = (unsigned int)n;
    HIPCHK(c, hipMemcpyAsync(&c->ts->nstep, &v, sizeof(v), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ADG_OK;
}

int adg_limit(adg_ctx* c, unsigned long long* n_limited) {
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    unsigned long long before = 0;
    if (n_limited) {
        HIPCHK(c, hipMemcpyAsync(&before, c->lim_count, sizeof(before), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    int rc = launch_limiter(c, true);
    if (rc != ADG_OK) return rc;
    if (n_limited) {
        unsigned long long after = 0;
        HIPCHK(c, hipMemcpyAsync(&after, c->lim_count, sizeof(after), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        *n_limited = after - before;
    }
    return ADG_OK;
}

int adg_limiter_stats(adg_ctx* c, unsigned long long* total, int reset) {
    if (!c) return ADG_E_ARG;
    if (total) {
        HIPCHK(c, hipMemcpyAsync(total, c->lim_count, sizeof(*total), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (reset) {
        const unsigned long long z = 0ull;
        HIPCHK(c, hipMemcpyAsync(c->lim_count, &z, sizeof(z), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return ADG_OK;
}

}  // extern "C"

namespace {

// Level-index cells of the valid cells of level lev, in valid-list order
// (patch by patch, x fastest).
std::vector<std::array<int, 3>> level_valid_cells(const adg::Hierarchy& H, int lev) {
    std::vector<std::array<int, 3>> out;
    const int D = H.D;
    for (int p = H.lev[lev].p_begin; p < H.lev[lev].p_end; ++p) {
        const adg::PatchLayout& P = H.patch[p];
        std::array<int, 3> iv{0, 0, 0};
        for (iv[2] = (D > 2 ? P.lo[2] : 0); iv[2] <= (D > 2 ? P.hi[2] : 0); ++iv[2])
            for (iv[1] = P.lo[1]; iv[1] <= P.hi[1]; ++iv[1])
                for (iv[0] = P.lo[0]; iv[0] <= P.hi[0]; ++iv[0]) out.push_back(iv);
    }
    return out;
}

std::vector<std::vector<adg::IBox>> level_boxes(const adg::Hierarchy& H) {
    std::vector<std::vector<adg::IBox>> out((size_t)H.nlev);
    for (const adg::PatchLayout& P : H.patch) {
        adg::IBox b;
        for (int d = 0; d < 3; ++d) { b.lo[d] = P.lo[d]; b.hi[d] = P.hi[d]; }
        out[(size_t)P.lev].push_back(b);
    }
    return out;
}

// Dense owner map of one level: level-index cell -> slot of its valid cell, or -1.
// Cell -> valid slot of one level (the hierarchy's owner grid), with the
// periodic wrap.
struct OwnerMap {
    int D = 2, dom[3] = {1, 1, 1};
    int per[3] = {1, 1, 1};
    const int* own = nullptr;
    void build(const adg::Hierarchy& H, int lev) {
        D = H.D;
        for (int d = 0; d < 3; ++d) {
            dom[d] = d < D ? H.lev[lev].dom[d] : 1;
            per[d] = d < D ? H.per[d] : 1;
        }
        own = H.grid_slot[(size_t)lev].data();
    }
    long lin(const int* iv) const {
        long k = 0;
        for (int d = D - 1; d >= 0; --d) k = k * dom[d] + iv[d];
        return k;
    }
    int find(const int* iv_in) const {
        int iv[3] = {iv_in[0], iv_in[1], iv_in[2]};
        for (int d = 0; d < D; ++d)
            if (iv[d] < 0 || iv[d] >= dom[d]) {
                if (!per[d]) return -1;
                iv[d] = ((iv[d] % dom[d]) + dom[d]) % dom[d];
            }
        return own[lin(iv)];
    }
};

int floordiv2(int a) { return a >= 0 ? a / 2 : -((-a + 1) / 2); }

}  // namespace

extern "C" {

int adg_host_make_grids(int dim, const int* dom0, const int* periodic, int max_level, int blocking_factor,
                        int max_grid_size, double grid_eff, int n_error_buf, int n_proper, int nlev_old,
                        const int* npatch_old, const int* boxes_old, const int* ntags, const int* tags,
                        int* nlev_new, int* npatch_new, int* boxes_new, int cap_boxes) {
    if ((dim != 2 && dim != 3) || !dom0 || !periodic || !npatch_old || !boxes_old || !ntags || !nlev_new ||
        !npatch_new || nlev_old < 1 || max_level < 0 || max_level >= adg::kMaxLev || blocking_factor < 2 ||
        max_grid_size < blocking_factor || !(grid_eff > 0.0 && grid_eff <= 1.0) || n_error_buf < 0 || n_proper < 0)
        return ADG_E_ARG;
    adg::AmrSettings st;
    st.max_level = max_level;
    st.blocking_factor = blocking_factor;
    st.max_grid_size = max_grid_size;
    st.grid_eff = grid_eff;
    st.n_error_buf = n_error_buf;
    st.n_proper = n_proper;
    std::vector<std::vector<adg::IBox>> old((size_t)nlev_old);
    const int* b = boxes_old;
    for (int l = 0; l < nlev_old; ++l)
        for (int k = 0; k < npatch_old[l]; ++k, b += 2 * dim) {
            adg::IBox x;
            for (int d = 0; d < dim; ++d) { x.lo[d] = b[d]; x.hi[d] = b[dim + d]; }
            old[(size_t)l].push_back(x);
        }
    const int max_crse = std::min(nlev_old - 1, max_level - 1);
    std::vector<adg::TagList> tl((size_t)std::max(max_crse + 1, 0));
    const int* t = tags;
    for (int l = 0; l <= max_crse; ++l)
        for (int k = 0; k < ntags[l]; ++k, t += dim) {
            std::array<int, 3> iv{0, 0, 0};
            for (int d = 0; d < dim; ++d) iv[d] = t[d];
            tl[(size_t)l].push_back(iv);
        }
    std::vector<std::vector<adg::IBox>> ng;
    try {
        ng = adg::make_new_grids(dim, dom0, periodic, st, old, tl);
    } catch (const std::exception&) {
        return ADG_E_ARG;
    }
    *nlev_new = 1 + (int)ng.size();
    npatch_new[0] = npatch_old[0];
    int nb = 0;
    for (int k = 0; k < npatch_old[0]; ++k, ++nb) {
        if (boxes_new && nb < cap_boxes)
            for (int d = 0; d < 2 * dim; ++d) boxes_new[(size_t)nb * 2 * dim + d] = boxes_old[(size_t)k * 2 * dim + d];
    }
    for (size_t l = 0; l < ng.size(); ++l) {
        npatch_new[l + 1] = (int)ng[l].size();
        for (const adg::IBox& x : ng[l]) {
            if (boxes_new && nb < cap_boxes)
                for (int d = 0; d < dim; ++d) {
                    boxes_new[(size_t)nb * 2 * dim + d] = x.lo[d];
                    boxes_new[(size_t)nb * 2 * dim + dim + d] = x.hi[d];
                }
            ++nb;
        }
    }
    return nb;
}

int adg_set_amr(adg_ctx* c, int max_level, int blocking_factor, int max_grid_size, double grid_eff,
                int n_error_buf, int n_proper) {
    if (!c) return ADG_E_ARG;
    if (max_level < 0 || max_level >= adg::kMaxLev) return fail(c, ADG_E_ARG, "max_level out of range");
    if (blocking_factor < 2 || blocking_factor % 2) return fail(c, ADG_E_ARG, "blocking factor must be even, >= 2");
    if (max_grid_size < blocking_factor || max_grid_size % blocking_factor)
        return fail(c, ADG_E_ARG, "max_grid_size must be a multiple of the blocking factor");
    if (!(grid_eff > 0.0 && grid_eff <= 1.0) || n_error_buf < 0 || n_proper < 0)
        return fail(c, ADG_E_ARG, "bad grid_eff, n_error_buf or n_proper");
    c->amr.max_level = max_level;
    c->amr.blocking_factor = blocking_factor;
    c->amr.max_grid_size = max_grid_size;
    c->amr.grid_eff = grid_eff;
    c->amr.n_error_buf = n_error_buf;
    c->amr.n_proper = n_proper;
    c->amr_set = true;
    return ADG_OK;
}

int adg_set_tagging(adg_ctx* c, int criterion, int n, const double* amr_c, int nprm, const double* params) {
    if (!c) return ADG_E_ARG;
    if (criterion < 0 || criterion > 2) return fail(c, ADG_E_ARG, "criterion must be 0, 1 or 2");
    if (criterion == 0 && c->D != 2) return fail(c, ADG_E_UNSUP, "the vortex criterion is two-dimensional");
    if (n < 1 || n > adg::kMaxLev || !amr_c) return fail(c, ADG_E_ARG, "amr_c must give 1..kMaxLev values");
    if (nprm < 0 || nprm > 4 || (nprm > 0 && !params)) return fail(c, ADG_E_ARG, "at most 4 parameters");
    c->tag_crit = criterion;
    for (int l = 0; l < adg::kMaxLev; ++l) c->tag_c[l] = amr_c[l < n ? l : n - 1];
    for (int i = 0; i < nprm; ++i) c->tag_prm[i] = params[i];
    return ADG_OK;
}

int adg_tag_cells(adg_ctx* c, int lev, unsigned char* tags) {
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    if (lev < 0 || lev >= c->H.nlev || !tags) return fail(c, ADG_E_ARG, "bad level or buffer");
    const adg::LevelLayout& L = c->H.lev[lev];
    const long n = L.valid_end - L.valid_begin;
    if (n == 0) return ADG_OK;
    if (c->tag_crit == 0) {
        // isentropic_vortex (Compressible_Euler.h:617-642): r <= amr_c^2
        // about (x0, y0) with periodic distance, at the cell centre
        const std::vector<std::array<int, 3>> cells = level_valid_cells(c->H, lev);
        const double Lx = c->phi[0] - c->plo[0], Ly = c->phi[1] - c->plo[1];
        for (long i = 0; i < n; ++i) {
            const double x = c->plo[0] + (cells[i][0] + 0.5) * L.dx[0];
            const double y = c->plo[1] + (cells[i][1] + 0.5) * L.dx[1];
            double dx = x - c->tag_prm[0], dy = y - c->tag_prm[1];
            if (dx > Lx / 2.0) dx -= Lx;
            if (dx < -Lx / 2.0) dx += Lx;
            if (dy > Ly / 2.0) dy -= Ly;
            if (dy < -Ly / 2.0) dy += Ly;
            const double r = std::sqrt(dx * dx + dy * dy);
            tags[i] = r <= std::pow(c->tag_c[lev], 2) ? 1 : 0;
        }
        return ADG_OK;
    }
    uint8_t* d_out = nullptr;
    HIPCHK(c, hipMalloc((void**)&d_out, (size_t)n));
    adg::TagArgs a{};
    a.Uw = c->Uw;
    a.valid_list = c->valid_list;
    for (int d = 0; d < c->D; ++d) {
        a.cell_face[d][0] = c->cell_face[d][0];
        a.cell_face[d][1] = c->cell_face[d][1];
        a.face_lr[d] = c->face_tr[d];   // a copy ghost holds its source's U_w after the ghost fill
    }
    a.out = d_out;
    a.NS = c->H.NS;
    a.vbegin = L.valid_begin;
    a.n = n;
    a.crit = c->tag_crit;
    a.gamma = c->gamma;
    if (c->tag_crit == 1) {
        double h = L.dx[0];
        for (int d = 1; d < c->D; ++d) h = std::min(h, L.dx[d]);
        a.thr = c->lim_tvb_c[lev] * c->tag_c[lev] * std::pow(h, 2.0);
    } else {
        a.thr = c->tag_c[lev];
    }
    for (int d = 0; d < c->D; ++d)
        for (int k = 0; k < c->tab.Np_s; ++k) {
            bool lin = c->tab.mi_s[k][d] == 1, others = true;
            for (int e = 0; e < c->D; ++e)
                if (e != d && c->tab.mi_s[k][e] != 0) others = false;
            if (lin && others) { a.lin[d] = k; break; }
        }
#define L_TAG(PP, DD) hipLaunchKernelGGL((adg::k_tag<PP, DD>), dim3(nblocks(n)), dim3(256), 0, c->stream, a)
    ADG_DISPATCH(c, L_TAG);
#undef L_TAG
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(tags, d_out, (size_t)n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_out);
    if (e != hipSuccess) return fail(c, ADG_E_HIP, std::string("adg_tag_cells: ") + hipGetErrorString(e));
    return ADG_OK;
}

int adg_get_levels(adg_ctx* c, int* nlev, int* npatch, int* boxes, int cap) {
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    if (!nlev || !npatch) return fail(c, ADG_E_ARG, "null argument");
    *nlev = c->H.nlev;
    const int D = c->D;
    for (int l = 0; l < c->H.nlev; ++l) npatch[l] = c->H.lev[l].p_end - c->H.lev[l].p_begin;
    int k = 0;
    for (const adg::PatchLayout& P : c->H.patch) {
        if (boxes && k < cap)
            for (int d = 0; d < D; ++d) {
                boxes[(size_t)k * 2 * D + d] = P.lo[d];
                boxes[(size_t)k * 2 * D + D + d] = P.hi[d];
            }
        ++k;
    }
    return k;
}

int adg_remake(adg_ctx* c, int nlev, const int* npatch, const int* boxes) {
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    if (nlev < 1 || nlev > adg::kMaxLev || !npatch || !boxes) return fail(c, ADG_E_ARG, "bad level arguments");
    const int D = c->D, Q = c->Q, Ns = c->tab.Np_s;
    // level 0 is not regridded
    {
        const adg::LevelLayout& L0 = c->H.lev[0];
        if (npatch[0] != L0.p_end - L0.p_begin) return fail(c, ADG_E_ARG, "level 0 boxes differ");
        for (int p = 0; p < npatch[0]; ++p)
            for (int d = 0; d < D; ++d)
                if (boxes[(size_t)p * 2 * D + d] != c->H.patch[p].lo[d] ||
                    boxes[(size_t)p * 2 * D + D + d] != c->H.patch[p].hi[d])
                    return fail(c, ADG_E_ARG, "level 0 boxes differ");
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // the old lists move out (the new build recycles the storage of the
    // layout before them, and these are kept for the next one)
    adg::Hierarchy old = std::move(c->H);
    double* old_uw = c->Uw;
    c->level_allocs.erase(std::remove(c->level_allocs.begin(), c->level_allocs.end(), (void*)old_uw),
                          c->level_allocs.end());
    int rc = adg_set_levels(c, nlev, npatch, boxes);
    if (rc != ADG_OK) { (void)hipFree(old_uw); return rc; }
    const auto t_x0 = std::chrono::steady_clock::now();
    const adg::Hierarchy& H = c->H;
    // level 0: the same slots
    HIPCHK(c, hipMemcpy2DAsync(c->Uw, (size_t)H.NS * sizeof(double), old_uw, (size_t)old.NS * sizeof(double),
                               (size_t)H.lev[0].nslot * sizeof(double), (size_t)Q * Ns, hipMemcpyDeviceToDevice,
                               c->stream));
    int* d_slot = nullptr;
    int* d_src = nullptr;
    uint8_t* d_kind = nullptr;
    uint8_t* d_child = nullptr;
    std::string err;
    for (int l = 1; l < H.nlev && err.empty(); ++l) {
        OwnerMap old_own, crse_own;
        const bool had = l < old.nlev;
        if (had) old_own.build(old, l);
        crse_own.build(H, l - 1);
        // every slot of the level (ghosts included), in slot order: each
        // patch fills its own range
        const long base = H.lev[l].slot_off, nsl = H.lev[l].nslot;
        std::vector<int> es((size_t)nsl), esrc((size_t)nsl);
        std::vector<uint8_t> ek((size_t)nsl), ech((size_t)nsl);
        std::atomic<bool> nested{true};
        adg::parallel_for(H.lev[l].p_end - H.lev[l].p_begin, [&](int i) {
            const adg::PatchLayout& P = H.patch[(size_t)(H.lev[l].p_begin + i)];
            int iv[3] = {0, 0, 0};
            long s = P.off;
            const int z0 = D > 2 ? P.lo[2] - 1 : 0, z1 = D > 2 ? P.hi[2] + 1 : 0;
            for (iv[2] = z0; iv[2] <= z1; ++iv[2])
                for (iv[1] = P.lo[1] - 1; iv[1] <= P.hi[1] + 1; ++iv[1])
                    for (iv[0] = P.lo[0] - 1; iv[0] <= P.hi[0] + 1; ++iv[0], ++s) {
                        const size_t k = (size_t)(s - base);
                        int mv[3] = {iv[0], iv[1], iv[2]}, mask = 0;
                        for (int d = 0; d < D; ++d) {
                            const int n = H.lev[l].dom[d];
                            if (H.per[d]) continue;
                            if (iv[d] < 0) { mv[d] = -1 - iv[d]; mask |= 1 << d; }
                            if (iv[d] >= n) { mv[d] = 2 * n - 1 - iv[d]; mask |= 1 << d; }
                        }
                        const int os = had ? old_own.find(mv) : -1;
                        es[k] = (int)s;
                        if (os >= 0) {
                            esrc[k] = os;
                            ek[k] = (uint8_t)(mask << 2);
                            ech[k] = 0;
                            continue;
                        }
                        int civ[3] = {0, 0, 0}, child = 0;
                        for (int d = 0; d < D; ++d) {
                            civ[d] = floordiv2(mv[d]);
                            child += (mv[d] - 2 * civ[d]) << (D - 1 - d);
                        }
                        int cs = crse_own.find(civ);
                        if (cs < 0) { nested = false; cs = 0; }   // reported below, nothing is launched
                        esrc[k] = cs;
                        ek[k] = (uint8_t)(1 | (mask << 2));
                        ech[k] = (uint8_t)child;
                    }
        });
        if (!nested) err = "new level " + std::to_string(l) + " is not nested in level " + std::to_string(l - 1);
        if (!err.empty()) break;
        const long ne = (long)es.size();
        hipError_t e = hipMalloc((void**)&d_slot, ne * sizeof(int));
        if (e == hipSuccess) e = hipMalloc((void**)&d_src, ne * sizeof(int));
        if (e == hipSuccess) e = hipMalloc((void**)&d_kind, (size_t)ne);
        if (e == hipSuccess) e = hipMalloc((void**)&d_child, (size_t)ne);
        if (e == hipSuccess) e = hipMemcpy(d_slot, es.data(), ne * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_src, esrc.data(), ne * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_kind, ek.data(), (size_t)ne, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_child, ech.data(), (size_t)ne, hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            adg::RemakeArgs a{c->d_tab, c->to, c->Uw, old_uw, d_slot, d_src, d_kind, d_child, H.NS, old.NS, ne};
#define L_RM(PP, DD) hipLaunchKernelGGL((adg::k_remake<PP, DD>), dim3(nblocks(ne * Q)), dim3(256), 0, c->stream, a)
            switch (c->P * 10 + D) {
                case 12: L_RM(1, 2); break;
                case 22: L_RM(2, 2); break;
                case 32: L_RM(3, 2); break;
                case 13: L_RM(1, 3); break;
                case 23: L_RM(2, 3); break;
                case 33: L_RM(3, 3); break;
                default: break;
            }
#undef L_RM
            e = hipGetLastError();
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        }
        if (e != hipSuccess) err = std::string("adg_remake: ") + hipGetErrorString(e);
        (void)hipFree(d_slot); (void)hipFree(d_src); (void)hipFree(d_kind); (void)hipFree(d_child);
        d_slot = d_src = nullptr;
        d_kind = d_child = nullptr;
    }
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(old_uw);
    c->H_spare = std::move(old);
    c->rg_sec[4] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_x0).count();
    if (!err.empty()) return fail(c, ADG_E_ARG, err);
    ++c->rg_changes;
    return ADG_OK;
}

int adg_regrid(adg_ctx* c, double time, int* changed) {
    (void)time;   // the criteria do not depend on time (Compressible_Euler.h:628-629)
    if (changed) *changed = 0;
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    if (!c->amr_set) return fail(c, ADG_E_STATE, "adg_set_amr must come first");
    const adg::Hierarchy& H = c->H;
    const int D = c->D;
    const int max_crse = std::min(H.nlev - 1, c->amr.max_level - 1);
    const auto t_t0 = std::chrono::steady_clock::now();
    std::vector<adg::TagList> tl((size_t)std::max(max_crse + 1, 0));
    for (int l = 0; l <= max_crse; ++l) {
        const long n = H.lev[l].valid_end - H.lev[l].valid_begin;
        std::vector<unsigned char> tg((size_t)n);
        int rc = adg_tag_cells(c, l, tg.data());
        if (rc != ADG_OK) return rc;
        const std::vector<std::array<int, 3>> cells = level_valid_cells(H, l);
        for (long i = 0; i < n; ++i)
            if (tg[(size_t)i]) tl[(size_t)l].push_back(cells[(size_t)i]);
    }
    const auto t_t1 = std::chrono::steady_clock::now();
    c->rg_sec[0] += std::chrono::duration<double>(t_t1 - t_t0).count();
    std::vector<std::vector<adg::IBox>> old = level_boxes(H), ng;
    try {
        ng = adg::make_new_grids(D, c->dom0, c->per, c->amr, old, tl);
    } catch (const std::exception& e) {
        return fail(c, ADG_E_ARG, std::string("regrid: ") + e.what());
    }
    c->rg_sec[1] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_t1).count();
    std::vector<int> np(1 + ng.size()), bx;
    np[0] = (int)old[0].size();
    for (const adg::IBox& b : old[0]) {
        for (int d = 0; d < D; ++d) bx.push_back(b.lo[d]);
        for (int d = 0; d < D; ++d) bx.push_back(b.hi[d]);
    }
    bool same = (int)ng.size() + 1 == H.nlev;
    for (size_t l = 0; l < ng.size(); ++l) {
        np[l + 1] = (int)ng[l].size();
        if (same && ng[l].size() != old[l + 1].size()) same = false;
        for (size_t k = 0; k < ng[l].size(); ++k) {
            const adg::IBox& b = ng[l][k];
            for (int d = 0; d < D; ++d) bx.push_back(b.lo[d]);
            for (int d = 0; d < D; ++d) bx.push_back(b.hi[d]);
            if (same)
                for (int d = 0; d < D; ++d)
                    if (b.lo[d] != old[l + 1][k].lo[d] || b.hi[d] != old[l + 1][k].hi[d]) same = false;
        }
    }
    if (same) return ADG_OK;
    int rc = adg_remake(c, (int)np.size(), np.data(), bx.data());
    if (rc == ADG_OK && changed) *changed = 1;
    return rc;
}

int adg_conserved_totals_noalloc(adg_ctx* c, double* d_totals) {
    if (!c || !c->levels_set) return fail(c, ADG_E_STATE, "levels not set");
    if (!d_totals) return fail(c, ADG_E_ARG, "null device output");
    const long n = (long)c->H.uncovered.size();
    const unsigned nb = nblocks(n);
    if (n == 0) return fail(c, ADG_E_ARG, "no uncovered cells");
    if ((long)c->Q * (long)nb > kScratchDoubles)
        return fail(c, ADG_E_ARG, "adg_conserved_totals_noalloc: " + std::to_string(n) +
                                      " uncovered cells exceed the static scratch (at most " +
                                      std::to_string(kScratchDoubles / c->Q * 256) + ")");
    adg::TotArgs a{c->Uw, c->uncovered, c->slot_lev, c->scratch_dev, c->geo, c->H.NS, n, c->tab.Np_s, c->Q};
    hipLaunchKernelGGL(adg::k_totals, dim3(nb), dim3(256), 0, c->stream, a);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(adg::k_totals_finish, dim3(1), dim3(256), 0, c->stream, (const double*)c->scratch_dev,
                       (int)nb, c->Q, d_totals);
    HIPCHK(c, hipGetLastError());
    return ADG_OK;
}

int adg_lp_error(adg_ctx* c, int q, int p, int ic_case, const double* params, double t, double* out) {
    if (!c || !c->levels_set || !out || !params) return fail(c, ADG_E_STATE, "levels not set");
    if (q < 0 || q >= c->Q || p < 1 || p > 2) return fail(c, ADG_E_ARG, "q out of range or p not 1 or 2");
    if (ic_case != 0 && ic_case != 2 && ic_case != 5) return fail(c, ADG_E_UNSUP, "no exact solution for this case");
    if (ic_case == 0 && c->D != 2) return fail(c, ADG_E_UNSUP, "the vortex is two-dimensional");
    const long n = (long)c->H.uncovered.size();
    const unsigned nb = nblocks(n);
    adg::LpArgs a{};
    a.tab = c->d_tab;
    a.to = c->to;
    a.Uw = c->Uw;
    a.list = c->uncovered;
    a.centre = c->unc_centre;
    a.slot_lev = c->slot_lev;
    a.partial = c->partial;
    a.geo = c->geo;
    a.NS = c->H.NS;
    a.n = n;
    a.q = q;
    a.p = p;
    a.ic = ic_case;
    for (int i = 0; i < 5; ++i) a.prm[i] = params[i];
    a.gamma = c->gamma;
    a.t = t;
    for (int d = 0; d < 3; ++d) { a.plo[d] = c->H.plo[d]; a.phi[d] = c->H.phi[d]; }
#define L_LP(PP, DD) hipLaunchKernelGGL((adg::k_lp_error<PP, DD>), dim3(nb), dim3(256), 0, c->stream, a)
    ADG_DISPATCH(c, L_LP);
#undef L_LP
    HIPCHK(c, hipGetLastError());
    std::vector<double> h(nb);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->partial, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double tot = 0.0, vamr = 0.0;
    for (unsigned b = 0; b < nb; ++b) tot += h[b];
    for (int l = 0; l < c->H.nlev; ++l) {
        long nu = 0;
        for (int s : c->H.uncovered) nu += c->H.slot_lev[s] == l;
        vamr += c->geo.vol[l] * (double)nu;
    }
    *out = std::pow(tot / vamr, 1.0 / (double)p);
    return ADG_OK;
}

int adg_sync(adg_ctx* c) {
    if (!c) return ADG_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ADG_OK;
}

void* adg_stream(adg_ctx* c) { return c ? (void*)c->stream : nullptr; }

int adg_sizes(adg_ctx* c, int* o) {
    if (!c || !o) return ADG_E_ARG;
    o[0] = c->tab.Np_s; o[1] = c->tab.Np_st; o[2] = c->tab.qMp_s; o[3] = c->tab.qMp_st;
    o[4] = c->tab.qMp_st_bd; o[5] = c->tab.N; o[6] = c->Q;
    return ADG_OK;
}

int adg_get_table(adg_ctx* c, const char* name, double* out, int cap) {
    if (!c || !name) return ADG_E_ARG;
    const std::vector<double>* v = adg::table_by_name(c->tab, name);
    if (!v) return fail(c, ADG_E_ARG, std::string("unknown table ") + name);
    if (out)
        for (size_t i = 0; i < v->size() && (int)i < cap; ++i) out[i] = (*v)[i];
    return (int)v->size();
}

// Host-only: a named element table of degree P in D dimensions, built on
// the host exactly as adg_create builds it (no device work).
int adg_host_table(int P, int D, const char* name, double* out, int cap) {
    if (!supported(P, D) || !name) return ADG_E_ARG;
    try {
        const adg::ElementTables t = adg::build_element_tables(P, D);
        const std::vector<double>* v = adg::table_by_name(t, name);
        if (!v) return ADG_E_ARG;
        if (out)
            for (size_t i = 0; i < v->size() && (int)i < cap; ++i) out[i] = (*v)[i];
        return (int)v->size();
    } catch (...) {
        return ADG_E_ARG;
    }
}

// Host-only inspection of the hierarchy lists built for a box layout (no
// device work; used by CPU tests).  kind 0: b_coarse, 1: b_fine, 2: child
// index of the faces of direction d of level lev, [face] in level order.
// kind 3: counts {NS, ghosts, faces d0, faces d1, faces d2, moments,
// reflux rows, uncovered}.  Returns the number of ints written or < 0.
int adg_host_inspect(int dim, const int* dom_n, const int* periodic, int nlev, const int* npatch, const int* boxes,
                     int lev, int d, int kind, int* out, int cap) {
    try {
        double plo[3] = {0, 0, 0}, phi[3] = {1, 1, 1};
        adg::Hierarchy H = adg::build_hierarchy(dim, dom_n, plo, phi, periodic, nlev, npatch, boxes);
        if (kind == 3) {
            const long v[8] = {H.NS, (long)H.g_slot.size(), H.NF[0], H.NF[1], H.NF[2], (long)H.mom_face.size(),
                               (long)H.rc_slot.size(), (long)H.uncovered.size()};
            for (int i = 0; i < 8 && i < cap; ++i) out[i] = (int)v[i];
            return 8;
        }
        if (lev < 0 || lev >= nlev || d < 0 || d >= dim || kind < 0 || kind > 2) return ADG_E_ARG;
        const std::vector<int8_t>& m = kind == 0 ? H.bcoarse[d] : kind == 1 ? H.bfine[d] : H.cidx[d];
        const long off = H.lev[lev].face_off[d], n = H.lev[lev].nface[d];
        for (long i = 0; i < n && i < cap; ++i) out[i] = m[off + i];
        return (int)n;
    } catch (const std::exception&) {
        return ADG_E_ARG;
    }
}

// Lane-layout probe of the MFMA builtins used by the predictor (diagnostic;
// not part of include/adg.h).  prec 0: f64, 1: f32; out receives 256 doubles.
int adg_probe_mfma_layout(int prec, int mode, double* out) {
    if (!out) return ADG_E_ARG;
    void* d = nullptr;
    if (hipMalloc(&d, 256 * sizeof(double)) != hipSuccess) return ADG_E_NOMEM;
    if (prec == 0) hipLaunchKernelGGL(adg::k_probe_mfma<double>, dim3(1), dim3(64), 0, 0, (double*)d, mode);
    else hipLaunchKernelGGL(adg::k_probe_mfma<float>, dim3(1), dim3(64), 0, 0, (float*)d, mode);
    std::vector<float> f(256);
    int rc = ADG_OK;
    if (hipDeviceSynchronize() != hipSuccess) rc = ADG_E_HIP;
    if (rc == ADG_OK) {
        if (prec == 0) rc = hipMemcpy(out, d, 256 * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess ? 0 : ADG_E_HIP;
        else {
            rc = hipMemcpy(f.data(), d, 256 * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess ? 0 : ADG_E_HIP;
            for (int i = 0; i < 256; ++i) out[i] = f[i];
        }
    }
    (void)hipFree(d);
    return rc;
}

const char* adg_last_error(adg_ctx* c) { return c ? c->err.c_str() : "null context"; }

void adg_destroy(adg_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_levels(c);
    for (adg_ctx::Scratch* sc : {&c->sc_vol, &c->sc_tr, &c->sc_fnum[0], &c->sc_fnum[1], &c->sc_fnum[2], &c->sc_warm})
        if (sc->p) (void)hipFree(sc->p);
    if (c->d_tab) (void)hipFree(c->d_tab);
    for (int v = 0; v < 2; ++v)
        if (c->d_frag[v]) (void)hipFree(c->d_frag[v]);
    if (c->ts) (void)hipFree(c->ts);
    if (c->lstats) (void)hipFree(c->lstats);
    for (void* p : {(void*)c->mlp.par, (void*)c->mlp.grad, (void*)c->mlp.m, (void*)c->mlp.v, (void*)c->mlp.norm,
                    (void*)c->mlp.frag, (void*)c->mlp.X, (void*)c->mlp.Y, (void*)c->mlp.work, (void*)c->mlp.dacc})
        if (p) (void)hipFree(p);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // extern "C"
