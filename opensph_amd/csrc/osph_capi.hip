// C ABI of the MI355X OpenSPH hot path (declared in include/osph_gpu.h): context, HBM-resident state,
// neighbour build orchestration, SymmetricSolver<3>::integrate and PredictorCorrector::step.
#include "osph_device.h"
#include "osph_internal.h"
#include "osph_gravity.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

using namespace osph;

namespace {

constexpr double PI_H = 3.14159265358979323846264338327950288;

struct ProfEvent {
    const char* name;
    hipEvent_t start, stop;
};

// Device and pinned host buffers of a batched step (osph_step_pc_batch over contexts with particles):
// the block tables of the batched launches, the member table, the members' accumulators, the batch's
// sort arrays and the captured HIP graph of the stages after the sort.  Owned by the batch's first
// member; rebuilt when the members, their sizes or the launch tuning change.
struct BatchPlan {
    std::vector<osph_ctx*> members;
    std::vector<size_t> sizes;
    Tuning tune{};
    bool mixed = false;
    size_t total = 0;
    std::vector<uint32_t> pfirst;  // host copy of the particle table's first[]
    uint32_t* tables = nullptr;
    BatchBlocks pb{}, lb{}, qb{};
    BatchMember* dev = nullptr;
    BatchMember* host = nullptr;
    unsigned long long* small = nullptr;
    unsigned long long* small_host = nullptr;
    uint64_t *keys = nullptr, *keys_sorted = nullptr;
    uint32_t *vals = nullptr, *vals_sorted = nullptr;
    void* sort_temp = nullptr;
    size_t sort_temp_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_bits = -1, graph_screened = -1, graph_plain = -1;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_bits = graph_screened = graph_plain = -1;
    }
    ~BatchPlan() {
        drop_graph();
        void* dp[] = { tables, dev, small, keys, keys_sorted, vals, vals_sorted, sort_temp };
        for (void* p : dp) (void)hipFree(p);
        if (host) (void)hipHostFree(host);
        if (small_host) (void)hipHostFree(small_host);
    }
};

}  // namespace

struct osph_ctx {
    osph_config cfg{};
    Params prm{};
    std::vector<osph_material> mats;
    osph_material* mats_dev = nullptr;
    hipStream_t stream = nullptr;
    int device = 0;

    size_t n = 0, cap = 0;
    DevState s{};
    double* dblock = nullptr;
    uint32_t* ublock = nullptr;
    // the state is kept in cell order, renewed every tune.reorder_period searches into the second block
    double* dblock2 = nullptr;
    uint32_t* ublock2 = nullptr;
    uint32_t* perm = nullptr;     // slot -> storage index
    uint32_t* perm2 = nullptr;
    uint32_t* slot_of = nullptr;  // storage index -> slot
    int since_reorder = 0;

    // neighbour structures (all sized for cap particles)
    uint32_t *keys = nullptr, *keys_sorted = nullptr, *iota = nullptr, *order = nullptr;
    const char* failed_launch = nullptr;  // name of the first timed launch that left a HIP error
    void* sort_temp = nullptr;
    size_t sort_temp_bytes = 0;
    uint32_t* cell_first = nullptr;  // [cells + 1]: first sorted position of each cell (valid rows only)
    uint32_t* row_range = nullptr;   // [2 rows]: sorted range of each row of cells
    size_t row_cap = 0;
    size_t cell_cap = 0;
    PairRec* rec = nullptr;
    PairRecF* recf = nullptr;
    Fixed fixed{};
    double4* pos = nullptr;
    float4* posf = nullptr;  // origin-relative FP32 positions of the screened list kernel
    uint32_t* group = nullptr;
    uint32_t* list = nullptr;
    uint32_t* count = nullptr;
    uint32_t kmax = 96;
    size_t list_cap = 0;
    unsigned long long* small = nullptr;     // accumulators, layout SMALL_* in osph_internal.h
    unsigned long long* small_host = nullptr;  // pinned mirror
    double* lut = nullptr;
    float* lutf = nullptr;
    Grid grid{};
    Screen screen{};
    double href = 0.0;  // class reference h of the next neighbour search (0: measure it first)
    bool neighbors_valid = false;
    // reverse-mode run (osph_adjoint_forward / _backward)
    struct {
        bool ready = false;
        int steps = 0, every = 0, saved_reorder = 0;
        bool reorder_saved = false;  // saved_reorder holds the setting to restore (a forward run began)
        std::vector<double> dt;
        double* ckpt = nullptr;  // ceil(steps / every) blocks of 13 planes of cap doubles
        size_t ckpt_blocks = 0, ckpt_cap = 0;
    } adj;
    // attached self-gravity (osph_attach_gravity): evaluated once per derivative evaluation, copied into
    // the accelerations before every pair-sum launch
    osph_gravity* grav = nullptr;
    hipEvent_t grav_ready = nullptr, grav_done = nullptr;
    const double* grav_acc[3] = { nullptr, nullptr, nullptr };

    unsigned char* staging = nullptr;
    size_t staging_cap = 0;
    double* scratch = nullptr;  // reductions (integrals)
    size_t scratch_cap = 0;
    unsigned long long* bounds_part = nullptr;  // per-block partial grid bounds
    size_t bounds_part_cap = 0;

    int prof = 0;  // 0 off, 1 every launch, 2 pair-sum launches only
    std::vector<ProfEvent> events;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, std::pair<double, uint64_t>> prof_acc;

    Tuning tune{};
    std::string err;
    BatchPlan* batch = nullptr;  // batched steps led by this context (osph_step_pc_batch)

    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
};

namespace {

#define HIPCHK(ctx, call)                                                                              \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            return (ctx)->fail(OSPH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));        \
        }                                                                                              \
    } while (0)

hipEvent_t take_event(osph_ctx* c) {
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Runs a launcher; with profiling on, brackets it with events on the context's stream.
template <typename F>
void timed(osph_ctx* c, const char* name, F&& f) {
    if (!c->prof || (c->prof == 2 && std::strncmp(name, "pairs", 5) != 0)) {
        f();
        if (!c->failed_launch && hipPeekAtLastError() != hipSuccess) c->failed_launch = name;
        return;
    }
    ProfEvent ev{ name, take_event(c), take_event(c) };
    if (ev.start) (void)hipEventRecord(ev.start, c->stream);
    f();
    if (!c->failed_launch && hipPeekAtLastError() != hipSuccess) c->failed_launch = name;
    if (ev.stop) (void)hipEventRecord(ev.stop, c->stream);
    c->events.push_back(ev);
}

void resolve_events(osph_ctx* c) {
    for (ProfEvent& e : c->events) {
        float ms = 0.f;
        if (e.start && e.stop && hipEventSynchronize(e.stop) == hipSuccess &&
            hipEventElapsedTime(&ms, e.start, e.stop) == hipSuccess) {
            auto& acc = c->prof_acc[e.name];
            acc.first += ms;
            acc.second += 1;
        }
        if (e.start) c->event_pool.push_back(e.start);
        if (e.stop) c->event_pool.push_back(e.stop);
    }
    c->events.clear();
}

Launch L(osph_ctx* c) { return Launch{ c->stream, c->tune }; }

void free_state(osph_ctx* c) {
    (void)hipFree(c->dblock);
    (void)hipFree(c->ublock);
    (void)hipFree(c->dblock2);
    (void)hipFree(c->ublock2);
    (void)hipFree(c->perm);
    (void)hipFree(c->perm2);
    (void)hipFree(c->slot_of);
    c->dblock2 = nullptr;
    c->ublock2 = nullptr;
    c->perm = c->perm2 = c->slot_of = nullptr;
    (void)hipFree(c->keys);
    (void)hipFree(c->keys_sorted);
    (void)hipFree(c->iota);
    (void)hipFree(c->order);
    (void)hipFree(c->sort_temp);
    (void)hipFree(c->rec);
    (void)hipFree(c->recf);
    (void)hipFree(c->posf);
    (void)hipFree(c->pos);
    (void)hipFree(c->group);
    (void)hipFree(c->count);
    c->dblock = nullptr;
    c->ublock = nullptr;
    c->keys = c->keys_sorted = c->iota = c->order = nullptr;
    c->sort_temp = nullptr;
    c->rec = nullptr;
    c->recf = nullptr;
    c->posf = nullptr;
    c->pos = nullptr;
    c->group = c->count = nullptr;
    c->cap = 0;
}

int ensure_staging(osph_ctx* c, size_t bytes) {
    if (bytes <= c->staging_cap) return OSPH_OK;
    (void)hipFree(c->staging);
    c->staging = nullptr;
    c->staging_cap = 0;
    HIPCHK(c, hipMalloc(&c->staging, bytes));
    c->staging_cap = bytes;
    return OSPH_OK;
}

size_t list_entries(size_t n, uint32_t kmax) {
    return (n + LIST_TILE - 1) / LIST_TILE * LIST_TILE * size_t(kmax);
}

int ensure_list(osph_ctx* c, size_t entries) {
    if (entries <= c->list_cap) return OSPH_OK;
    (void)hipFree(c->list);
    c->list = nullptr;
    c->list_cap = 0;
    HIPCHK(c, hipMalloc(&c->list, entries * sizeof(uint32_t)));
    c->list_cap = entries;
    return OSPH_OK;
}

int ensure_rows(osph_ctx* c, size_t rows) {
    if (rows <= c->row_cap) return OSPH_OK;
    (void)hipFree(c->row_range);
    c->row_range = nullptr;
    c->row_cap = 0;
    HIPCHK(c, hipMalloc(&c->row_range, 2 * rows * sizeof(uint32_t)));
    c->row_cap = rows;
    return OSPH_OK;
}

int ensure_cells(osph_ctx* c, size_t cells) {
    if (cells + 1 <= c->cell_cap) return OSPH_OK;
    (void)hipFree(c->cell_first);
    c->cell_first = nullptr;
    c->cell_cap = 0;
    HIPCHK(c, hipMalloc(&c->cell_first, (cells + 1) * sizeof(uint32_t)));
    c->cell_cap = cells + 1;
    return OSPH_OK;
}

Params make_params(const osph_config& cfg) {
    Params p{};
    p.solid_stress = cfg.solid_stress;
    p.adaptive_h = cfg.adaptive_h;
    p.sum_only_undamaged = cfg.sum_only_undamaged;
    p.has_reduce = cfg.has_reduce;
    p.has_damage = cfg.has_damage;
    p.av_alpha = cfg.av_alpha;
    p.av_beta = cfg.av_beta;
    p.h_lo = cfg.h_lo;
    p.h_hi = cfg.h_hi;
    p.criteria = cfg.criteria;
    p.courant = cfg.courant;
    p.deriv_factor = cfg.deriv_factor;
    p.div_factor = cfg.div_factor;
    p.max_dt = cfg.max_dt;
    return p;
}

// CubicSpline<3>::valueImpl / gradImpl (core/sph/kernel/Kernel.h:148-188) tabulated as in the
// LutKernel<3> constructor (:83-101).  Built on the host in strict IEEE double so that the device
// table equals the reference's table bit for bit.
void build_lut(double* values, double* grads) {
#pragma clang fp contract(off)
    const double norm = 1.0 / PI_H;
    const double rad = 2.0;
    const double qSqrToIdx = double(LUT_ENTRIES) * (1.0 / (rad * rad));
    for (int i = 0; i < LUT_ENTRIES + 1; ++i) {
        const double qSqr = double(i) / qSqrToIdx;
        const double q = std::sqrt(qSqr);
        double v = 0.0, g = 0.0;
        if (q < 1.0) {
            v = norm * (0.25 * ((2.0 - q) * (2.0 - q) * (2.0 - q)) - (1.0 - q) * (1.0 - q) * (1.0 - q));
        } else if (q < 2.0) {
            v = norm * (0.25 * ((2.0 - q) * (2.0 - q) * (2.0 - q)));
        }
        if (q == 0.0) {
            g = -3.0 * norm;
        } else if (q < 1.0) {
            g = (1.0 / q) * norm * (-0.75 * ((2.0 - q) * (2.0 - q)) + 3.0 * ((1.0 - q) * (1.0 - q)));
        } else if (q < 2.0) {
            g = (1.0 / q) * norm * (-0.75 * ((2.0 - q) * (2.0 - q)));
        }
        values[i] = v;
        grads[i] = g;
    }
}

double decode_ord(unsigned long long e) {
    unsigned long long b = (e >> 63) ? (e & ~0x8000000000000000ull) : ~e;
    double v;
    std::memcpy(&v, &b, sizeof(v));
    return v;
}

struct QDesc {
    int ncomp;
    int elem;  // 8 double, 4 uint32
    int natural;
    int planes[8];
    int off[8];
    bool u32;
};

bool describe(int q, int order, QDesc& d) {
    d = QDesc{};
    auto set = [&](int ncomp, int elem, int natural, std::initializer_list<int> pl, std::initializer_list<int> off,
                   bool u32) {
        d.ncomp = ncomp;
        d.elem = elem;
        d.natural = natural;
        d.u32 = u32;
        int k = 0;
        for (int p : pl) d.planes[k++] = p;
        k = 0;
        for (int o : off) d.off[k++] = o;
    };
    switch (q) {
    case OSPH_Q_POSITION:
        if (order == 0) set(4, 8, 32, { X, Y, Z, H }, { 0, 8, 16, 24 }, false);
        else if (order == 1) set(4, 8, 32, { VX, VY, VZ, VH }, { 0, 8, 16, 24 }, false);
        else if (order == 2) set(3, 8, 32, { AX, AY, AZ }, { 0, 8, 16 }, false);
        else return false;
        return true;
    case OSPH_Q_DEVIATORIC_STRESS:
        if (order == 0) set(5, 8, 64, { SXX, SYY, SXY, SXZ, SYZ }, { 0, 8, 16, 24, 32 }, false);
        else if (order == 1) set(5, 8, 64, { DSXX, DSYY, DSXY, DSXZ, DSYZ }, { 0, 8, 16, 24, 32 }, false);
        else return false;
        return true;
    case OSPH_Q_VELOCITY_GRADIENT:
        if (order != 0) return false;
        set(6, 8, 64, { GXX, GYY, GZZ, GXY, GXZ, GYZ }, { 0, 8, 16, 32, 40, 48 }, false);
        return true;
    default:
        break;
    }
    int plane = -1;
    bool u32 = false;
    switch (q) {
    case OSPH_Q_MASS: plane = order == 0 ? M : -1; break;
    case OSPH_Q_DENSITY: plane = order == 0 ? RHO : (order == 1 ? DRHO : -1); break;
    case OSPH_Q_ENERGY: plane = order == 0 ? U : (order == 1 ? DU : -1); break;
    case OSPH_Q_DAMAGE: plane = order == 0 ? D : (order == 1 ? DD : -1); break;
    case OSPH_Q_PRESSURE: plane = order == 0 ? P : -1; break;
    case OSPH_Q_SOUND_SPEED: plane = order == 0 ? CS : -1; break;
    case OSPH_Q_STRESS_REDUCING: plane = order == 0 ? RED : -1; break;
    case OSPH_Q_VELOCITY_DIVERGENCE: plane = order == 0 ? DIVV : -1; break;
    case OSPH_Q_EPS_MIN: plane = order == 0 ? EPSMIN : -1; break;
    case OSPH_Q_M_ZERO: plane = order == 0 ? MZERO : -1; break;
    case OSPH_Q_EXPLICIT_GROWTH: plane = order == 0 ? GROWTH : -1; break;
    case OSPH_Q_NEIGHBOR_CNT: plane = order == 0 ? NCNT : -1; u32 = true; break;
    case OSPH_Q_FLAG: plane = order == 0 ? FLAG : -1; u32 = true; break;
    case OSPH_Q_MATERIAL_ID: plane = order == 0 ? MAT : -1; u32 = true; break;
    case OSPH_Q_N_FLAWS: plane = order == 0 ? NFLAWS : -1; u32 = true; break;
    default: return false;
    }
    if (plane < 0) return false;
    set(1, u32 ? 4 : 8, u32 ? 4 : 8, { plane }, { 0 }, u32);
    return true;
}

void plane_ptrs(osph_ctx* c, const QDesc& d, void** out) {
    for (int k = 0; k < d.ncomp; ++k) {
        out[k] = d.u32 ? static_cast<void*>(c->s.u[d.planes[k]]) : static_cast<void*>(c->s.d[d.planes[k]]);
    }
}

// Adjoint plane indices (the order launch_adj_step expects): lambda of r xyz, v xyz, a xyz, rho, drho,
// u, du.
enum { LRX_ = 0, LVX_ = 3, LAX_ = 6, LRHO_ = 9, LDRHO_ = 10, LU_ = 11, LDU_ = 12 };

int sync_check(osph_ctx* c) {
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
        const std::string where = c->failed_launch ? std::string(" (first failing launch: ") + c->failed_launch + ")" : "";
        c->failed_launch = nullptr;
        return c->fail(OSPH_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(le) + where);
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OSPH_OK;
}

// Grid + sort + pack + neighbour lists for the current positions (ISymmetricFinder::build).
// FP32 screen of the neighbour test.  Coordinates are rounded once to FP32 relative to lo, so each
// carries an absolute error below eps D (D the largest extent, eps = 2^-24) and a difference below
// 2 eps D + eps |dx|.  Near the cutoff (|d| ~ R >= 2 h_min) the squared distance then has a relative
// error below 2 sqrt(3) eps (D / h_min + 1) plus a few eps of rounding, and the squared radius below
// 6 eps.  The band is twice that sum; when it exceeds 1e-2 the screen is not worth running.
Screen make_screen(const osph_ctx* c, const double lo[3], const double hi[3], double hmin) {
    Screen sc{};
    const int mode = c->tune.lists_screen;
    sc.on = mode == 1 || (mode == -1 && c->cfg.precision == OSPH_PRECISION_MIXED);
    double D = 0.0;
    for (int k = 0; k < 3; ++k) {
        sc.o[k] = lo[k];
        D = std::max(D, hi[k] - lo[k]);
    }
    const double eps = std::ldexp(1.0, -24);
    const double band = 2.0 * (2.0 * std::sqrt(3.0) * eps * (D / hmin + 1.0) + 12.0 * eps);
    if (!(band < 1e-2)) sc.on = 0;
    sc.lo = float(1.0 - band);
    sc.hi = float(1.0 + band);
    return sc;
}

// The class reference h of the first search after an upload: the smallest h (one host synchronisation).
int ensure_href(osph_ctx* c) {
    if (c->href > 0.0) return OSPH_OK;
    Launch l = L(c);
    timed(c, "init_small", [&] { launch_init_small(l, c->small); });
    timed(c, "bounds", [&] { launch_bounds_reduce(l, c->s, c->small, c->n); });
    HIPCHK(c, hipMemcpyAsync(c->small_host, c->small, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             c->stream));
    if (int e = sync_check(c)) return e;
    const double h0 = decode_ord(c->small_host[7]);
    if (!(h0 > 0.0) || !std::isfinite(h0)) return c->fail(OSPH_ERR_STATE, "smoothing lengths must be positive");
    c->href = h0;
    return OSPH_OK;
}

// PredictorCorrector's swap of the highest derivatives into the predictions storage, and the zeroing
// of the derivatives for the next evaluation (TimeStepping.cpp:331-334), as an exchange of plane
// pointers: the prediction planes take the derivatives of the last evaluation without a copy, and the
// derivative planes, whose stale values the next evaluation overwrites, take the place of the old
// predictions.  Done once before each prediction pass.
void swap_predictions(osph_ctx* c) {
    static constexpr int PAIRS[11][2] = { { AX, PAX }, { AY, PAY }, { AZ, PAZ }, { DRHO, PDRHO }, { DU, PDU },
                                          { DSXX, PDSXX }, { DSYY, PDSYY }, { DSXY, PDSXY }, { DSXZ, PDSXZ },
                                          { DSYZ, PDSYZ }, { DD, PDD } };
    for (const auto& p : PAIRS) std::swap(c->s.d[p[0]], c->s.d[p[1]]);
}

// The pass before the neighbour search (k_advance: predictions for mode 2, the material initialize for
// modes 1 and 2) with the per-class grid bounds, then the grid, the sort by cell, the cell ranges and
// the packed neighbour records (ISymmetricFinder::build).  One host synchronisation reads the class
// bounds; the first search after an upload adds one more for the class reference h.
// First half of search_grid: everything up to the class-bound readback, which is left in flight.
int search_begin(osph_ctx* c, int mode, double dt) {
    const size_t n = c->n;
    Launch l = L(c);
    if (int e = ensure_href(c)) return e;
    if (mode == 2) swap_predictions(c);
    const double href = c->href;
    unsigned long long* cacc = c->small + SMALL_CLASSES;
    const bool vel = c->cfg.precision == OSPH_PRECISION_MIXED;
    if (bounds_partials(n) > c->bounds_part_cap) {
        (void)hipFree(c->bounds_part);
        c->bounds_part = nullptr;
        c->bounds_part_cap = 0;
        HIPCHK(c, hipMalloc(&c->bounds_part, bounds_partials(n) * sizeof(unsigned long long)));
        c->bounds_part_cap = bounds_partials(n);
    }
    timed(c, mode == 2 ? "advance" : (mode == 1 ? "material_init" : "class_bounds"),
          [&] { launch_advance(l, c->s, c->mats_dev, c->prm, mode, dt, href, c->small, vel, c->bounds_part, n); });
    HIPCHK(c, hipMemcpyAsync(c->small_host + SMALL_CLASSES, cacc,
                             (SMALL_VEL + 6 - SMALL_CLASSES) * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             c->stream));
    return OSPH_OK;
}

// The plan of a neighbour search from the class (and velocity) bounds in sh (a host copy of the SMALL_*
// accumulators): the grid, the FP32 screen and, in mixed precision, the fixed-point frames; the class reference h
// of the next search; the cell and row tables sized.  *rows_out = the rows of the grid.  Host only.
int plan_search(osph_ctx* c, const unsigned long long* sh, size_t* rows_out) {
    const double href = c->href;
    double lo[3] = { INFINITY, INFINITY, INFINITY }, hi[3] = { -INFINITY, -INFINITY, -INFINITY };
    double hmin = INFINITY, hmax = 0.0;
    double clo[MAX_CLASSES][3], chi[MAX_CLASSES][3], chmax[MAX_CLASSES];
    unsigned long long ccount[MAX_CLASSES];
    int last = 0;
    for (int cl = 0; cl < MAX_CLASSES; ++cl) {
        const unsigned long long* a = sh + SMALL_CLASSES + CLASS_STRIDE * cl;
        ccount[cl] = a[8];
        if (!ccount[cl]) continue;
        last = cl;
        for (int k = 0; k < 3; ++k) {
            clo[cl][k] = decode_ord(a[k]);
            chi[cl][k] = decode_ord(a[3 + k]);
            if (!std::isfinite(clo[cl][k]) || !std::isfinite(chi[cl][k])) {
                return c->fail(OSPH_ERR_STATE, "non-finite positions");
            }
            lo[k] = std::min(lo[k], clo[cl][k]);
            hi[k] = std::max(hi[k], chi[cl][k]);
        }
        chmax[cl] = decode_ord(a[6]);
        const double chmin = decode_ord(a[7]);
        if (!(chmin > 0.0) || !std::isfinite(chmax[cl])) {
            return c->fail(OSPH_ERR_STATE, "smoothing lengths must be positive");
        }
        hmin = std::min(hmin, chmin);
        hmax = std::max(hmax, chmax[cl]);
    }
    c->href = hmin;  // class reference of the next search

    c->screen = make_screen(c, lo, hi, hmin);
    Grid g{};
    g.hmin = href;
    g.nclass = last + 1;
    uint32_t base = 0;
    for (int cl = 0; cl < g.nclass; ++cl) {
        GridClass& G = g.c[cl];
        G.base = base;
        if (ccount[cl] == 0) {
            G.empty = 1;
            continue;
        }
        double ext = 0.0;
        for (int k = 0; k < 3; ++k) {
            G.lo[k] = clo[cl][k];
            G.hi[k] = chi[cl][k];
            ext = std::max(ext, chi[cl][k] - clo[cl][k]);
        }
        G.hmax = chmax[cl];
        // cells of 2 h_max(class) / divisor, at most 250 per axis
        G.cell = std::max(2.0 * G.hmax * (1.0 + 1e-9) / double(c->tune.cell_divisor), ext / 250.0);
        G.inv_cell = 1.0 / G.cell;
        int maxdim = 1;
        for (int k = 0; k < 3; ++k) {
            G.dim[k] = int(std::floor((G.hi[k] - G.lo[k]) / G.cell)) + 1;
            maxdim = std::max(maxdim, G.dim[k]);
        }
        // 2^bits > dim: every row keeps a slot past its last cell for the end of an x-run
        int bits = 1;
        while ((1 << bits) <= maxdim) ++bits;
        G.bits = bits;
        // the last row (largest Morton code of y, z, as row_key is monotonic in each) ends the class
        const uint64_t rows_end = (uint64_t(spread2(uint32_t(G.dim[1] - 1)) | (uint64_t(spread2(uint32_t(G.dim[2] - 1))) << 1)) + 1)
                                  << bits;
        if (uint64_t(base) + rows_end > (uint64_t(1) << 32) - 2) return c->fail(OSPH_ERR_STATE, "cell count exceeds 32 bits");
        base += uint32_t(rows_end);
    }
    g.key_bits = 1;
    while (g.key_bits < 32 && (uint64_t(1) << g.key_bits) < uint64_t(base)) ++g.key_bits;
    c->grid = g;
    if (int e = ensure_cells(c, size_t(base))) return e;
    size_t rows = 0;
    for (int cl = 0; cl < g.nclass; ++cl) {
        if (!g.c[cl].empty) rows += size_t(g.c[cl].dim[1]) * size_t(g.c[cl].dim[2]);
    }
    if (int e = ensure_rows(c, rows)) return e;
    *rows_out = rows;
    if (c->cfg.precision == OSPH_PRECISION_MIXED) {
        // fixed-point frames: the smallest power-of-two steps with the position and velocity boxes within
        // 2^31 - 2 steps per axis, so that the difference of any two coordinates fits in an int32 (two
        // neighbours can differ in velocity by the whole range), with steps that stay normal FP32 numbers
        auto frame_step = [](double extent) {
            const double span = 2147483646.0;
            double st = std::exp2(-100.0);
            if (extent > 0.0) st = std::max(st, std::exp2(std::ceil(std::log2(extent / span))));
            while (extent / st > span) st *= 2.0;
            return st;
        };
        Fixed fx{};
        double ext = 0.0, vrange = 0.0;
        for (int k = 0; k < 3; ++k) {
            fx.lo[k] = lo[k];
            ext = std::max(ext, hi[k] - lo[k]);
            fx.vlo[k] = decode_ord(sh[SMALL_VEL + k]);
            const double vhi = decode_ord(sh[SMALL_VEL + 3 + k]);
            if (!std::isfinite(fx.vlo[k]) || !std::isfinite(vhi)) {
                return c->fail(OSPH_ERR_STATE, "velocities must be finite");
            }
            vrange = std::max(vrange, vhi - fx.vlo[k]);
        }
        fx.step = frame_step(ext);
        fx.inv_step = 1.0 / fx.step;
        fx.stepf = float(fx.step);
        fx.vstep = frame_step(vrange);
        fx.inv_vstep = 1.0 / fx.vstep;
        fx.vstepf = float(fx.vstep);
        c->fixed = fx;
    }
    return OSPH_OK;
}

// Whether this search moves the state into cell order (every tune.reorder_period searches).
bool reorder_due(osph_ctx* c) {
    if (c->tune.reorder_period > 0 && ++c->since_reorder >= c->tune.reorder_period) {
        c->since_reorder = 0;
        return true;
    }
    return false;
}

// The second block of state planes, which a reorder fills.
DevState other_state(const osph_ctx* c) {
    DevState dst = c->s;
    for (int k = 0; k < NDPLANES; ++k) dst.d[k] = c->dblock2 + size_t(k) * c->cap;
    for (int k = 0; k < NIPLANES; ++k) dst.u[k] = c->ublock2 + size_t(k) * c->cap;
    return dst;
}

// After a reorder into the second block: it becomes the state.
void swap_state(osph_ctx* c, const DevState& dst) {
    c->s = dst;
    std::swap(c->dblock, c->dblock2);
    std::swap(c->ublock, c->ublock2);
    std::swap(c->perm, c->perm2);
}

// Second half: waits for the class bounds, then the grid, the sort by cell, the ranges and the records.
int search_finish(osph_ctx* c) {
    const size_t n = c->n;
    Launch l = L(c);
    if (int e = sync_check(c)) return e;
    size_t rows = 0;
    if (int e = plan_search(c, c->small_host, &rows)) return e;
    const Grid g = c->grid;

    timed(c, "cell_keys", [&] { launch_cell_keys(l, c->s, g, c->keys, c->iota, c->row_range, 2 * rows, n); });
    const size_t sort_need = sort_pairs_temp_bytes(n, g.key_bits);
    if (sort_need > c->sort_temp_bytes) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        (void)hipFree(c->sort_temp);
        c->sort_temp = nullptr;
        c->sort_temp_bytes = 0;
        HIPCHK(c, hipMalloc(&c->sort_temp, sort_need));
        c->sort_temp_bytes = sort_need;
    }
    hipError_t sort_err = hipSuccess;
    timed(c, "sort", [&] {
        sort_err = sort_pairs(l, c->sort_temp, c->sort_temp_bytes, c->keys, c->keys_sorted, c->iota, c->order, n,
                              g.key_bits);
    });
    HIPCHK(c, sort_err);
    timed(c, "cell_ranges", [&] {
        launch_cell_first(l, c->keys_sorted, g, c->row_range, c->cell_first, rows, n);
    });
    if (reorder_due(c)) {
        // move the state into this search's cell order: the record gather and the pair kernels' writes
        // then run over (nearly) consecutive slots until the next reorder
        const DevState dst = other_state(c);
        timed(c, "reorder", [&] {
            launch_permute_state(l, c->s, dst, c->order, c->perm, c->perm2, c->slot_of, n);
            launch_iota(l, c->order, n);
        });
        swap_state(c, dst);
    }
    if (c->cfg.precision == OSPH_PRECISION_MIXED) {
        timed(c, "pack", [&] {
            launch_pack_mixed(l, c->s, c->prm, c->fixed, c->order, c->recf, c->pos, c->posf, c->screen, c->group, n);
        });
    } else {
        timed(c, "pack", [&] { launch_pack(l, c->s, c->prm, c->order, c->rec, c->pos, c->posf, c->screen, c->group, n); });
    }
    return OSPH_OK;
}

int search_grid(osph_ctx* c, int mode, double dt) {
    if (int e = search_begin(c, mode, dt)) return e;
    return search_finish(c);
}

uint32_t* overflow_flag(osph_ctx* c) { return reinterpret_cast<uint32_t*>(c->small + SMALL_OVERFLOW); }

int launch_lists(osph_ctx* c) {
    if (int e = ensure_list(c, list_entries(c->n, c->kmax))) return e;
    timed(c, "neighbor_lists", [&] {
        launch_neighbor_lists(L(c), c->pos, c->posf, c->screen, c->grid, c->cell_first, c->list, c->count, c->kmax,
                              overflow_flag(c), c->n);
    });
    return OSPH_OK;
}

// After an overflow of the lists (a particle with more than kmax neighbours): grow kmax to fit the
// largest count and clear the flag.  The stream must be idle.
int grow_lists(osph_ctx* c) {
    std::vector<uint32_t> cnt(c->n);
    HIPCHK(c, hipMemcpy(cnt.data(), c->count, c->n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    const uint32_t mx = *std::max_element(cnt.begin(), cnt.end());
    c->kmax = (std::max<uint32_t>(mx + 8, c->kmax * 2) + 3u) & ~3u;
    launch_clear_u32(L(c), overflow_flag(c), 1, 0u);
    return OSPH_OK;
}

bool overflowed(const osph_ctx* c) {
    uint32_t flag;
    std::memcpy(&flag, &c->small_host[SMALL_OVERFLOW], sizeof(uint32_t));
    return flag != 0;
}

// Neighbour lists for the current grid, grown and rebuilt until every list fits.
int build_lists(osph_ctx* c) {
    for (int attempt = 0; attempt < 4; ++attempt) {
        if (int e = launch_lists(c)) return e;
        HIPCHK(c, hipMemcpyAsync(&c->small_host[SMALL_OVERFLOW], overflow_flag(c), sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, c->stream));
        if (int e = sync_check(c)) return e;
        if (!overflowed(c)) {
            c->neighbors_valid = true;
            return OSPH_OK;
        }
        if (int e = grow_lists(c)) return e;
    }
    return c->fail(OSPH_ERR_STATE, "neighbour list did not fit after growing");
}

int build_neighbors(osph_ctx* c, int mode) {
    c->neighbors_valid = false;
    if (c->n == 0) {
        c->neighbors_valid = true;
        return OSPH_OK;
    }
    if (int e = search_grid(c, mode, 0.0)) return e;
    return build_lists(c);
}

// GravitySolver::loop (core/sph/solvers/GravitySolver.cpp:64-91): the attached context's self-gravity of
// the current (predicted) positions, evaluated on its own stream after this context's queued work; this
// stream then waits for it.
int gravity_eval(osph_ctx* c) {
    if (!c->grav || !c->n) return OSPH_OK;
    HIPCHK(c, hipEventRecord(c->grav_ready, c->stream));
    const double* src[5] = { c->s.d[X], c->s.d[Y], c->s.d[Z], c->s.d[H], c->s.d[M] };
    if (gravity_self_from_device(c->grav, c->n, src, c->grav_ready, c->grav_done, c->grav_acc) != 0) {
        return c->fail(OSPH_ERR_HIP, std::string("self-gravity: ") + osph_gravity_last_error(c->grav));
    }
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->grav_done, 0));
    return OSPH_OK;
}

// The pair sums; with fuse (the PC step), followed in the same kernel by the finalize, the corrector and
// the criteria (timed then as "<pairs>_finalize_correct").
void launch_pair_sums(osph_ctx* c, const FuseArgs* fuse = nullptr) {
    Launch l = L(c);
    const size_t n = c->n;
    if (c->grav && n) {
        // the pair kernels add their sums to the self-gravity (GravitySolver::loop evaluates gravity
        // into the accumulated accelerations first)
        timed(c, "gravity_copy", [&] {
            for (int k = 0; k < 3; ++k) {
                (void)hipMemcpyAsync(c->s.d[AX + k], c->grav_acc[k], n * sizeof(double), hipMemcpyDeviceToDevice,
                                     c->stream);
            }
        });
    }
    if (c->cfg.precision == OSPH_PRECISION_MIXED) {
        timed(c, fuse ? "pairs_mixed_finalize_correct" : "pairs_mixed", [&] {
            launch_pairs_mixed(l, c->s, c->prm, c->fixed, c->recf, c->group, c->order, c->list, c->count, c->kmax, n, fuse);
        });
    } else {
        timed(c, fuse ? "pairs_fp64_finalize_correct" : "pairs_fp64", [&] {
            launch_pairs_fp64(l, c->s, c->prm, c->rec, c->group, c->order, c->list, c->count, c->lut, c->kmax, n,
                              fuse);
        });
    }
}

int integrate(osph_ctx* c, osph_stats* stats) {
    const size_t n = c->n;
    Launch l = L(c);
    if (int e = build_neighbors(c, 1)) return e;
    if (int e = gravity_eval(c)) return e;
    launch_pair_sums(c);
    timed(c, "finalize", [&] { launch_finalize(l, c->s, c->mats_dev, c->prm, n); });
    if (stats) {
        std::vector<uint32_t> cnt(n);
        if (n) HIPCHK(c, hipMemcpyAsync(cnt.data(), c->s.u[NCNT], n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                         c->stream));
        if (int e = sync_check(c)) return e;
        stats->neighbors_min = n ? *std::min_element(cnt.begin(), cnt.end()) : 0;
        stats->neighbors_max = n ? *std::max_element(cnt.begin(), cnt.end()) : 0;
        double sum = 0.0;
        for (uint32_t v : cnt) sum += v;
        stats->neighbors_mean = n ? sum / double(n) : 0.0;
    }
    return OSPH_OK;
}

// MultiCriterion selection over the four per-criterion minima (TimeStepCriterion.cpp:389-419); each
// criterion is limited by the maximal step on its own.  Ids follow CriterionId.
void select_timestep(const osph_ctx* c, const unsigned long long* mins, double* dt, int* crit) {
    const int own[4] = { 3, 2, 4, 5 };  // CFL_CONDITION, DERIVATIVE, ACCELERATION, DIVERGENCE
    double best = INFINITY;
    int id = 0;
    for (int k = 0; k < 4; ++k) {
        if (!(c->cfg.criteria & (1 << k))) continue;
        double v = INFINITY;
        if (mins[k] != ~0ull) std::memcpy(&v, &mins[k], sizeof(double));
        int cid = own[k];
        if (v > c->cfg.max_dt) {
            v = c->cfg.max_dt;
            cid = 1;  // MAXIMAL_VALUE
        }
        if (v < best) {
            best = v;
            id = cid;
        }
    }
    if (dt) *dt = best;
    if (crit) *crit = id;
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    ~DevBuf() { (void)hipFree(p); }
    hipError_t alloc(size_t count) { return hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)); }
};

// Per-block partial sums of osph_integrals_noalloc: 12 per block of the integrals kernels (at most 1024
// blocks) and their totals, integrals_scratch_doubles(n) <= 1025 * 12 for every n.
constexpr size_t NOALLOC_SCRATCH_DOUBLES = 16384;  // 128 KiB per device

}  // namespace

// The static scratch of osph_integrals_noalloc.  Every GPU the library runs on holds its own copy until the
// process ends.  In host code this name stands for a host-side shadow, not a device address: kernels take
// the per-device address that bind_noalloc_scratch looks up (noalloc_scratch[device]).
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mutex;
double* noalloc_scratch[NOALLOC_MAX_DEVICES] = {};

// The device address of `device`'s copy of the scratch, looked up once per device with that device
// current (osph_create calls this for its device).
hipError_t bind_noalloc_scratch(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(noalloc_mutex);
    if (noalloc_scratch[device]) return hipSuccess;
    int current = -1;
    hipError_t e = hipGetDevice(&current);
    if (e != hipSuccess) return e;
    if (current != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (current != device) (void)hipSetDevice(current);
    if (e != hipSuccess) return e;
    noalloc_scratch[device] = static_cast<double*>(ptr);
    return hipSuccess;
}

double* noalloc_scratch_of(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(noalloc_mutex);
    return noalloc_scratch[device];
}

}  // namespace

extern "C" {

int osph_device_info(char* name, size_t cap, int* compute_units) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return OSPH_ERR_HIP;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return OSPH_ERR_HIP;
    if (name && cap) {
        if (p.name[0]) {
            std::snprintf(name, cap, "%s (%s)", p.name, p.gcnArchName);
        } else {
            std::snprintf(name, cap, "%s", p.gcnArchName);  // some runtimes report no marketing name
        }
    }
    if (compute_units) *compute_units = p.multiProcessorCount;
    return OSPH_OK;
}

int osph_stream_copy(size_t bytes, int iters, double* gbs) {
    if (bytes < 16 || iters <= 0 || !gbs) return OSPH_ERR_ARG;
    bytes &= ~size_t(15);
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = OSPH_ERR_HIP;
    if (hipMalloc(&a, bytes) == hipSuccess && hipMalloc(&b, bytes) == hipSuccess &&
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
        hipEventCreate(&e1) == hipSuccess) {
        launch_clear_u32(Launch{ st }, static_cast<uint32_t*>(a), bytes / 4, 0x3ff00000u);
        launch_stream_copy(st, a, b, bytes);  // warm-up
        (void)hipEventRecord(e0, st);
        for (int k = 0; k < iters; ++k) launch_stream_copy(st, a, b, bytes);
        (void)hipEventRecord(e1, st);
        float ms = 0.f;
        if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms > 0.f) {
            *gbs = 2.0 * double(bytes) * iters / (double(ms) * 1e-3) / 1e9;
            rc = OSPH_OK;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

int osph_kernel_lut(double* values, double* grads) {
    if (!values || !grads) return OSPH_ERR_ARG;
    build_lut(values, grads);
    return OSPH_OK;
}

int osph_create(osph_ctx** out, const osph_config* cfg, const osph_material* materials, int n_materials) {
    if (!out || !cfg || !materials || n_materials <= 0) return OSPH_ERR_ARG;
    *out = nullptr;
    if (cfg->precision != OSPH_PRECISION_FP64 && cfg->precision != OSPH_PRECISION_MIXED) return OSPH_ERR_ARG;
    osph_ctx* c = new osph_ctx();
    c->cfg = *cfg;
    c->prm = make_params(*cfg);
    c->mats.assign(materials, materials + n_materials);
    auto bail = [&](int code) {
        osph_destroy(c);
        return code;
    };
    if (hipGetDevice(&c->device) != hipSuccess) return bail(OSPH_ERR_HIP);
    if (bind_noalloc_scratch(c->device) != hipSuccess) return bail(OSPH_ERR_HIP);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail(OSPH_ERR_HIP);
    if (hipMalloc(&c->mats_dev, sizeof(osph_material) * n_materials) != hipSuccess) return bail(OSPH_ERR_HIP);
    if (hipMemcpy(c->mats_dev, materials, sizeof(osph_material) * n_materials, hipMemcpyHostToDevice) != hipSuccess)
        return bail(OSPH_ERR_HIP);
    if (hipMalloc(&c->small, SMALL_WORDS * sizeof(unsigned long long)) != hipSuccess) return bail(OSPH_ERR_HIP);
    if (hipHostMalloc(&c->small_host, SMALL_WORDS * sizeof(unsigned long long)) != hipSuccess) return bail(OSPH_ERR_HIP);
    std::vector<double> values(LUT_ENTRIES + 1), grads(LUT_ENTRIES + 1);
    build_lut(values.data(), grads.data());
    if (hipMalloc(&c->lut, grads.size() * sizeof(double)) != hipSuccess) return bail(OSPH_ERR_HIP);
    if (hipMemcpy(c->lut, grads.data(), grads.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return bail(OSPH_ERR_HIP);
    std::vector<float> gradsf(grads.begin(), grads.end());
    if (hipMalloc(&c->lutf, gradsf.size() * sizeof(float)) != hipSuccess) return bail(OSPH_ERR_HIP);
    if (hipMemcpy(c->lutf, gradsf.data(), gradsf.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return bail(OSPH_ERR_HIP);
    *out = c;
    return OSPH_OK;
}

void osph_destroy(osph_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->grav && gravity_owner(c->grav) == c) gravity_set_owner(c->grav, nullptr);
    delete c->batch;
    c->batch = nullptr;
    resolve_events(c);
    for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
    if (c->grav_ready) (void)hipEventDestroy(c->grav_ready);
    if (c->grav_done) (void)hipEventDestroy(c->grav_done);
    free_state(c);
    (void)hipFree(c->cell_first);
    (void)hipFree(c->row_range);
    (void)hipFree(c->list);
    (void)hipFree(c->staging);
    (void)hipFree(c->scratch);
    (void)hipFree(c->bounds_part);
    (void)hipFree(c->adj.ckpt);
    (void)hipFree(c->mats_dev);
    (void)hipFree(c->small);
    (void)hipFree(c->lut);
    (void)hipFree(c->lutf);
    if (c->small_host) (void)hipHostFree(c->small_host);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* osph_last_error(const osph_ctx* c) {
    return c ? c->err.c_str() : "null context";
}

int osph_resize(osph_ctx* c, size_t n) {
    if (!c) return OSPH_ERR_ARG;
    if (n > 0x7fffffffull) return c->fail(OSPH_ERR_ARG, "particle count exceeds 2^31-1");
    c->neighbors_valid = false;
    c->href = 0.0;
    if (n > c->cap) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        free_state(c);
        const size_t cap = n;
        HIPCHK(c, hipMalloc(&c->dblock, sizeof(double) * NDPLANES * cap));
        HIPCHK(c, hipMalloc(&c->ublock, sizeof(uint32_t) * NIPLANES * cap));
        HIPCHK(c, hipMalloc(&c->dblock2, sizeof(double) * NDPLANES * cap));
        HIPCHK(c, hipMalloc(&c->ublock2, sizeof(uint32_t) * NIPLANES * cap));
        HIPCHK(c, hipMalloc(&c->perm, sizeof(uint32_t) * cap));
        HIPCHK(c, hipMalloc(&c->perm2, sizeof(uint32_t) * cap));
        HIPCHK(c, hipMalloc(&c->slot_of, sizeof(uint32_t) * cap));
        HIPCHK(c, hipMalloc(&c->keys, sizeof(uint32_t) * cap));
        HIPCHK(c, hipMalloc(&c->keys_sorted, sizeof(uint32_t) * cap));
        HIPCHK(c, hipMalloc(&c->iota, sizeof(uint32_t) * cap));
        HIPCHK(c, hipMalloc(&c->order, sizeof(uint32_t) * cap));
        if (c->cfg.precision == OSPH_PRECISION_MIXED) {
            HIPCHK(c, hipMalloc(&c->recf, sizeof(PairRecF) * cap));
        } else {
            HIPCHK(c, hipMalloc(&c->rec, sizeof(PairRec) * cap));
        }
        HIPCHK(c, hipMalloc(&c->pos, sizeof(double4) * (cap + POS_SLACK)));
        HIPCHK(c, hipMalloc(&c->posf, sizeof(float4) * (cap + POS_SLACK)));
        // the slack past the last particle holds defined (zero) values
        launch_clear_u32(L(c), reinterpret_cast<uint32_t*>(c->pos + cap), 8 * POS_SLACK, 0u);
        launch_clear_u32(L(c), reinterpret_cast<uint32_t*>(c->posf + cap), 4 * POS_SLACK, 0u);
        HIPCHK(c, hipMalloc(&c->group, sizeof(uint32_t) * cap));
        HIPCHK(c, hipMalloc(&c->count, sizeof(uint32_t) * cap));
        c->sort_temp_bytes = sort_pairs_temp_bytes(cap, 32);
        HIPCHK(c, hipMalloc(&c->sort_temp, std::max<size_t>(c->sort_temp_bytes, 256)));
        c->cap = cap;
        for (int k = 0; k < NDPLANES; ++k) c->s.d[k] = c->dblock + size_t(k) * cap;
        for (int k = 0; k < NIPLANES; ++k) c->s.u[k] = c->ublock + size_t(k) * cap;
    }
    c->n = n;
    if (c->cap) {
        // every plane starts at zero (doubles: all-zero bits); slots start in storage order
        launch_clear_u32(L(c), reinterpret_cast<uint32_t*>(c->dblock), size_t(2) * NDPLANES * c->cap, 0u);
        launch_clear_u32(L(c), c->ublock, size_t(NIPLANES) * c->cap, 0u);
        launch_iota(L(c), c->perm, c->cap);
        launch_iota(L(c), c->slot_of, c->cap);
    }
    c->since_reorder = c->tune.reorder_period;  // the first search moves the state into cell order
    return sync_check(c);
}

int osph_set_quantity(osph_ctx* c, int q, int order, const void* host, size_t stride) {
    if (!c) return OSPH_ERR_ARG;
    QDesc d;
    if (!describe(q, order, d)) return c->fail(OSPH_ERR_ARG, "unknown quantity/order");
    if (stride == 0) stride = size_t(d.natural);
    if (stride < size_t(d.off[d.ncomp - 1] + d.elem)) return c->fail(OSPH_ERR_ARG, "stride smaller than element");
    const size_t n = c->n;
    if (n == 0) return OSPH_OK;
    if (!host) return c->fail(OSPH_ERR_ARG, "null host buffer");
    if (q == OSPH_Q_MATERIAL_ID) {
        // the kernels index the material table with these values: validate them here
        const unsigned char* b = static_cast<const unsigned char*>(host);
        for (size_t i = 0; i < n; ++i) {
            uint32_t v;
            std::memcpy(&v, b + i * stride, sizeof(v));
            if (v >= c->mats.size()) return c->fail(OSPH_ERR_ARG, "material id out of range");
        }
    }
    if (int e = ensure_staging(c, n * stride)) return e;
    HIPCHK(c, hipMemcpyAsync(c->staging, host, n * stride, hipMemcpyHostToDevice, c->stream));
    void* dst[8];
    plane_ptrs(c, d, dst);
    launch_stage_to_planes(L(c), c->staging, stride, d.ncomp, d.off, d.elem, dst, c->slot_of, n);
    c->neighbors_valid = false;
    c->href = 0.0;
    return sync_check(c);
}

int osph_get_quantity(osph_ctx* c, int q, int order, void* host, size_t stride) {
    if (!c) return OSPH_ERR_ARG;
    QDesc d;
    if (!describe(q, order, d)) return c->fail(OSPH_ERR_ARG, "unknown quantity/order");
    if (stride == 0) stride = size_t(d.natural);
    if (stride < size_t(d.off[d.ncomp - 1] + d.elem)) return c->fail(OSPH_ERR_ARG, "stride smaller than element");
    const size_t n = c->n;
    if (n == 0) return OSPH_OK;
    if (!host) return c->fail(OSPH_ERR_ARG, "null host buffer");
    if (int e = ensure_staging(c, n * stride)) return e;
    if ((n * stride) % 4 == 0) launch_clear_u32(L(c), reinterpret_cast<uint32_t*>(c->staging), n * stride / 4, 0u);
    void* src[8];
    plane_ptrs(c, d, src);
    launch_planes_to_stage(L(c), c->staging, stride, d.ncomp, d.off, d.elem, src, c->slot_of, n);
    HIPCHK(c, hipMemcpyAsync(host, c->staging, n * stride, hipMemcpyDeviceToHost, c->stream));
    return sync_check(c);
}

int osph_zero_derivatives(osph_ctx* c) {
    if (!c) return OSPH_ERR_ARG;
    timed(c, "zero_derivatives", [&] { launch_zero_derivatives(L(c), c->s, c->n); });
    return sync_check(c);
}

int osph_build_neighbors(osph_ctx* c) {
    if (!c) return OSPH_ERR_ARG;
    if (int e = build_neighbors(c, 0)) return e;
    return sync_check(c);
}

int osph_get_pairs(osph_ctx* c, uint32_t* pi, uint32_t* pj, size_t cap, size_t* count) {
    if (!c || !count) return OSPH_ERR_ARG;
    if (!c->neighbors_valid) return c->fail(OSPH_ERR_STATE, "neighbours not built");
    const size_t n = c->n;
    std::vector<uint32_t> cnt(n), order(n), perm(n), list(list_entries(n, c->kmax));
    if (n) {
        HIPCHK(c, hipMemcpy(cnt.data(), c->count, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(order.data(), c->order, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(perm.data(), c->perm, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < n; ++k) order[k] = perm[order[k]];  // sorted position -> storage index
        HIPCHK(c, hipMemcpy(list.data(), c->list, list.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    }
    std::vector<uint64_t> keys;
    for (size_t k = 0; k < n; ++k) {
        const uint64_t i = order[k];
        for (uint32_t m = 0; m < cnt[k] && m < c->kmax; ++m) {
            const uint64_t j = order[list[list_quad(k, m >> 2, c->kmax >> 2) * 4 + (m & 3u)]];
            if (i < j) keys.push_back((i << 32) | j);
        }
    }
    std::sort(keys.begin(), keys.end());
    *count = keys.size();
    for (size_t k = 0; k < keys.size() && k < cap; ++k) {
        if (pi) pi[k] = uint32_t(keys[k] >> 32);
        if (pj) pj[k] = uint32_t(keys[k] & 0xffffffffu);
    }
    return OSPH_OK;
}

int osph_integrate(osph_ctx* c, osph_stats* stats) {
    if (!c) return OSPH_ERR_ARG;
    if (int e = integrate(c, stats)) return e;
    return sync_check(c);
}

int osph_integrals(osph_ctx* c, double out[OSPH_INTEGRALS]) {
    if (!c || !out) return OSPH_ERR_ARG;
    std::memset(out, 0, OSPH_INTEGRALS * sizeof(double));
    const size_t n = c->n;
    if (!n) return OSPH_OK;
    const size_t need = integrals_scratch_doubles(n);
    if (need > c->scratch_cap) {
        (void)hipFree(c->scratch);
        c->scratch = nullptr;
        c->scratch_cap = 0;
        HIPCHK(c, hipMalloc(&c->scratch, need * sizeof(double)));
        c->scratch_cap = need;
    }
    timed(c, "integrals", [&] { launch_integrals(L(c), c->s, c->scratch, n); });
    double v[12];
    HIPCHK(c, hipMemcpyAsync(v, c->scratch + need - 12, sizeof(v), hipMemcpyDeviceToHost, c->stream));
    if (int e = sync_check(c)) return e;
    // TotalEnergy = kinetic + internal; CenterOfMass = sum m r / sum m (Integrals.cpp:95-180)
    out[0] = v[0] + v[1];
    out[1] = v[0];
    out[2] = v[1];
    for (int k = 0; k < 3; ++k) {
        out[3 + k] = v[2 + k];
        out[6 + k] = v[5 + k];
        out[9 + k] = v[8 + k] / v[11];
    }
    out[12] = v[11];
    return OSPH_OK;
}

int osph_integrals_noalloc(osph_ctx* c, double out[OSPH_INTEGRALS]) {
    if (!c || !out) return OSPH_ERR_ARG;
    std::memset(out, 0, OSPH_INTEGRALS * sizeof(double));
    const size_t n = c->n;
    if (!n) return OSPH_OK;
    const size_t need = integrals_scratch_doubles(n);
    if (need > NOALLOC_SCRATCH_DOUBLES) {
        return c->fail(OSPH_ERR_ARG, "osph_integrals_noalloc: " + std::to_string(need) +
                                         " partial sums exceed the static scratch of " +
                                         std::to_string(NOALLOC_SCRATCH_DOUBLES) + " doubles");
    }
    double* buf = noalloc_scratch_of(c->device);
    if (!buf) return c->fail(OSPH_ERR_STATE, "osph_integrals_noalloc: no scratch bound for this device");
    // the first kernel writes every partial the second reads: nothing needs to start at zero, and what an
    // earlier call left in the array is never read
    launch_integrals(L(c), c->s, buf, n);
    double v[12];
    HIPCHK(c, hipMemcpyAsync(v, buf + need - 12, sizeof(v), hipMemcpyDeviceToHost, c->stream));
    if (int e = sync_check(c)) return e;
    // the combination of osph_integrals (Integrals.cpp:95-180)
    out[0] = v[0] + v[1];
    out[1] = v[0];
    out[2] = v[1];
    for (int k = 0; k < 3; ++k) {
        out[3 + k] = v[2 + k];
        out[6 + k] = v[5 + k];
        out[9 + k] = v[8 + k] / v[11];
    }
    out[12] = v[11];
    return OSPH_OK;
}

int osph_timestep(osph_ctx* c, double* dt, int* crit) {
    if (!c) return OSPH_ERR_ARG;
    uint32_t* m32 = reinterpret_cast<uint32_t*>(c->small + SMALL_CRITERIA);
    launch_clear_u32(L(c), m32, 8, 0xffffffffu);
    timed(c, "criteria", [&] { launch_criteria(L(c), c->s, c->mats_dev, c->prm, c->small + SMALL_CRITERIA, c->n); });
    HIPCHK(c, hipMemcpyAsync(c->small_host + SMALL_CRITERIA, c->small + SMALL_CRITERIA,
                             4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    if (int e = sync_check(c)) return e;
    select_timestep(c, c->small_host + SMALL_CRITERIA, dt, crit);
    return OSPH_OK;
}

// One PredictorCorrector step with the state resident, in three phases so that a batch of contexts
// can interleave them:
//   begin   the fused predict pass and the class bounds (readback in flight);
//   middle  the grid, the lists, the pair sums and the fused finalize + correct + criteria (readback of
//           the criterion minima and the list overflow flag in flight);
//   end     the time step, or -- after a list overflow, when the finalize pass has left the predicted
//           state untouched -- larger lists and the second half of the step again.
int step_begin(osph_ctx* c, double dt) {
    if (!(dt > 0.0) || !std::isfinite(dt)) return c->fail(OSPH_ERR_ARG, "time step must be positive");
    c->neighbors_valid = false;
    if (c->n == 0) return OSPH_OK;
    return search_begin(c, 2, dt);
}

void launch_second_half(osph_ctx* c, double dt) {
    const bool fuse = c->tune.fuse_finalize == 1 ||
                      (c->tune.fuse_finalize == -1 && c->cfg.precision != OSPH_PRECISION_MIXED);
    if (fuse) {
        const FuseArgs fa{ c->mats_dev, dt, c->small + SMALL_CRITERIA, overflow_flag(c) };
        launch_pair_sums(c, &fa);
    } else {
        launch_pair_sums(c);
        timed(c, "finalize_correct", [&] {
            launch_finalize_correct(L(c), c->s, c->mats_dev, c->prm, dt, c->small + SMALL_CRITERIA, overflow_flag(c),
                                    c->n);
        });
    }
    (void)hipMemcpyAsync(c->small_host + SMALL_CRITERIA, c->small + SMALL_CRITERIA,
                         (SMALL_OVERFLOW + 1 - SMALL_CRITERIA) * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                         c->stream);
}

int step_middle(osph_ctx* c, double dt) {
    if (c->n == 0) return OSPH_OK;
    if (int e = search_finish(c)) return e;
    if (int e = gravity_eval(c)) return e;
    if (int e = launch_lists(c)) return e;
    launch_second_half(c, dt);
    return OSPH_OK;
}

int step_end(osph_ctx* c, double dt, double* next_dt, int* crit) {
    if (c->n == 0) {
        unsigned long long none[4] = { ~0ull, ~0ull, ~0ull, ~0ull };
        select_timestep(c, none, next_dt, crit);
        return OSPH_OK;
    }
    for (int attempt = 0; attempt < 4; ++attempt) {
        if (int e = sync_check(c)) return e;
        if (!overflowed(c)) {
            c->neighbors_valid = true;
            select_timestep(c, c->small_host + SMALL_CRITERIA, next_dt, crit);
            return OSPH_OK;
        }
        if (int e = grow_lists(c)) return e;
        if (int e = launch_lists(c)) return e;
        launch_second_half(c, dt);
    }
    return c->fail(OSPH_ERR_STATE, "neighbour list did not fit after growing");
}

int osph_step_pc(osph_ctx* c, double dt, double* next_dt, int* crit) {
    if (!c) return OSPH_ERR_ARG;
    if (int e = step_begin(c, dt)) return e;
    if (int e = step_middle(c, dt)) return e;
    return step_end(c, dt, next_dt, crit);
}

}  // extern "C"

namespace {

bool same_launch_tuning(const Tuning& a, const Tuning& b) {
    return a.pairs_block == b.pairs_block && a.pairs_lds == b.pairs_lds && a.pairs_lanes == b.pairs_lanes &&
           a.lists_block == b.lists_block && a.lists_lds == b.lists_lds && a.fuse_finalize == b.fuse_finalize &&
           a.batch_graph == b.batch_graph;
}

// The plan of a batch of members (each with particles), led by ms[0].
int make_plan(const std::vector<osph_ctx*>& ms, BatchPlan*& out) {
    osph_ctx* lead = ms[0];
    BatchPlan* old = lead->batch;
    if (old && old->members == ms && same_launch_tuning(old->tune, lead->tune)) {
        bool same = true;
        for (size_t k = 0; k < ms.size(); ++k) same = same && old->sizes[k] == ms[k]->n;
        if (same) {
            out = old;
            return OSPH_OK;
        }
    }
    HIPCHK(lead, hipStreamSynchronize(lead->stream));
    delete old;
    lead->batch = nullptr;
    BatchPlan* P = new BatchPlan();
    lead->batch = P;
    const size_t count = ms.size();
    P->members = ms;
    P->tune = lead->tune;
    P->mixed = lead->cfg.precision == OSPH_PRECISION_MIXED;
    // block tables: particles (BLOCK threads), lists (lists_block), pairs (pairs_block, lanes per particle)
    std::vector<uint32_t> host;
    auto table = [&](int block, int per, BatchBlocks& t, std::vector<uint32_t>& first) {
        first.assign(count + 1, 0u);
        for (size_t k = 0; k < count; ++k) first[k + 1] = first[k] + batch_blocks_for(ms[k]->n * size_t(per), block);
        t.total = first[count];
    };
    std::vector<uint32_t> pf, lf, qf;
    table(BLOCK, 1, P->pb, pf);
    table(P->tune.lists_block, 1, P->lb, lf);
    table(P->tune.pairs_block, P->tune.pairs_lanes, P->qb, qf);
    P->pfirst = pf;
    size_t words = 0;
    for (const auto* f : { &pf, &lf, &qf }) words += f->size() + f->back();
    std::vector<uint32_t> all;
    all.reserve(words);
    size_t offs[3][2];
    int ti = 0;
    for (const auto* f : { &pf, &lf, &qf }) {
        offs[ti][0] = all.size();
        for (size_t k = 0; k < count; ++k) {
            for (uint32_t b = (*f)[k]; b < (*f)[k + 1]; ++b) all.push_back(uint32_t(k));
        }
        offs[ti][1] = all.size();
        all.insert(all.end(), f->begin(), f->end());
        ++ti;
    }
    HIPCHK(lead, hipMalloc(&P->tables, all.size() * sizeof(uint32_t)));
    HIPCHK(lead, hipMemcpy(P->tables, all.data(), all.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    BatchBlocks* tabs[3] = { &P->pb, &P->lb, &P->qb };
    for (int t = 0; t < 3; ++t) {
        tabs[t]->member = P->tables + offs[t][0];
        tabs[t]->first = P->tables + offs[t][1];
    }
    HIPCHK(lead, hipMalloc(&P->dev, count * sizeof(BatchMember)));
    HIPCHK(lead, hipHostMalloc(&P->host, count * sizeof(BatchMember)));
    std::memset(static_cast<void*>(P->host), 0, count * sizeof(BatchMember));
    HIPCHK(lead, hipMalloc(&P->small, count * SMALL_WORDS * sizeof(unsigned long long)));
    HIPCHK(lead, hipHostMalloc(&P->small_host, count * SMALL_WORDS * sizeof(unsigned long long)));
    for (osph_ctx* c : ms) {
        P->sizes.push_back(c->n);
        P->total += c->n;
    }
    if (P->total >= (size_t(1) << 32)) return lead->fail(OSPH_ERR_ARG, "a batch must hold fewer than 2^32 particles");
    HIPCHK(lead, hipMalloc(&P->keys, P->total * sizeof(uint64_t)));
    HIPCHK(lead, hipMalloc(&P->keys_sorted, P->total * sizeof(uint64_t)));
    HIPCHK(lead, hipMalloc(&P->vals, P->total * sizeof(uint32_t)));
    HIPCHK(lead, hipMalloc(&P->vals_sorted, P->total * sizeof(uint32_t)));
    // sized once for the widest key, so that the captured graph's arguments stay valid
    P->sort_temp_bytes = std::max<size_t>(sort_pairs64_temp_bytes(P->total, 64), 256);
    HIPCHK(lead, hipMalloc(&P->sort_temp, P->sort_temp_bytes));
    out = P;
    return OSPH_OK;
}

// One predictor-corrector step of every member with one launch per stage (ms: the members with particles,
// dts: their time steps; next_dt / crit indexed like ms).  The same per-particle code and buffers as
// osph_step_pc on each member alone, so each member's result is bit-identical to stepping it alone.
int batch_step(const std::vector<osph_ctx*>& ms, const std::vector<double>& dts, double* next_dt, int* crit) {
    osph_ctx* lead = ms[0];
    BatchPlan* P = nullptr;
    if (int e = make_plan(ms, P)) return e;
    const size_t count = ms.size();
    // the members' own streams are idle (their last calls may have left work queued), and each has the
    // class reference h of its first search
    for (osph_ctx* c : ms) {
        if (hipStreamQuery(c->stream) != hipSuccess) HIPCHK(c, hipStreamSynchronize(c->stream));
        c->neighbors_valid = false;
        if (int e = ensure_href(c)) return e;
    }
    for (size_t k = 0; k < count; ++k) {
        osph_ctx* c = ms[k];
        const size_t need = size_t(P->pfirst[k + 1] - P->pfirst[k]) * bounds_values();
        if (need > c->bounds_part_cap) {
            (void)hipFree(c->bounds_part);
            c->bounds_part = nullptr;
            c->bounds_part_cap = 0;
            HIPCHK(c, hipMalloc(&c->bounds_part, need * sizeof(unsigned long long)));
            c->bounds_part_cap = need;
        }
        BatchMember& b = P->host[k];
        b.mats = c->mats_dev;
        b.prm = c->prm;
        b.dt = dts[k];
        b.href = c->href;
        b.part = c->bounds_part;
        b.n = uint32_t(c->n);
        b.small = P->small + SMALL_WORDS * k;
    }
    // the prediction planes of every member, once all allocations have succeeded
    for (size_t k = 0; k < count; ++k) {
        swap_predictions(ms[k]);
        P->host[k].s = ms[k]->s;
    }
    const Launch l{ lead->stream, P->tune };
    hipStream_t st = lead->stream;
    // predictions, material initialize and grid bounds of every member; one readback of all bounds
    HIPCHK(lead, hipMemcpyAsync(P->dev, P->host, count * sizeof(BatchMember), hipMemcpyHostToDevice, st));
    batch_advance(l, P->dev, int(count), P->pb, P->mixed);
    HIPCHK(lead, hipMemcpyAsync(P->small_host, P->small, count * SMALL_WORDS * sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, st));
    if (int e = sync_check(lead)) return e;
    // each member's grid, screen and fixed-point frames, planned as its own search plans them
    int key_bits = 1;
    bool screened = false, plain = false;
    uint32_t off = 0;
    for (size_t k = 0; k < count; ++k) {
        osph_ctx* c = ms[k];
        BatchMember& b = P->host[k];
        size_t rows = 0;
        if (int e = plan_search(c, P->small_host + SMALL_WORDS * k, &rows)) return e;
        if (int e = ensure_list(c, list_entries(c->n, c->kmax))) return e;
        int xslots = 1;
        for (int cl = 0; cl < c->grid.nclass; ++cl) {
            if (!c->grid.c[cl].empty) xslots = std::max(xslots, c->grid.c[cl].dim[0] + 1);
        }
        b.grid = c->grid;
        b.screen = c->screen;
        b.fx = c->fixed;
        b.kmax = c->kmax;
        b.rows = uint32_t(rows);
        b.xslots = xslots;
        b.reorder = reorder_due(c) ? 1 : 0;
        b.s2 = b.reorder ? other_state(c) : c->s;
        b.key_off = off;
        off += uint32_t(c->n);
        b.keys_sorted = c->keys_sorted;
        b.order = c->order;
        b.row_range = c->row_range;
        b.cell_first = c->cell_first;
        b.perm = c->perm;
        b.perm2 = c->perm2;
        b.slot_of = c->slot_of;
        b.rec = P->mixed ? static_cast<void*>(c->recf) : static_cast<void*>(c->rec);
        b.pos = c->pos;
        b.posf = c->posf;
        b.group = c->group;
        b.list = c->list;
        b.count = c->count;
        key_bits = std::max(key_bits, c->grid.key_bits);
        (b.screen.on ? screened : plain) = true;
    }
    int member_bits = 0;
    while ((size_t(1) << member_bits) < count) ++member_bits;
    HIPCHK(lead, hipMemcpyAsync(P->dev, P->host, count * sizeof(BatchMember), hipMemcpyHostToDevice, st));
    batch_cell_keys(l, P->dev, P->pb, key_bits, P->keys, P->vals);
    HIPCHK(lead, sort_pairs64(l, P->sort_temp, P->sort_temp_bytes, P->keys, P->keys_sorted, P->vals, P->vals_sorted,
                              P->total, key_bits + member_bits));
    const bool fuse = P->tune.fuse_finalize == 1 || (P->tune.fuse_finalize == -1 && !P->mixed);
    // the stages after the sort: their launch shapes and arguments are fixed for the plan (the per-step
    // values are in the member table), so they are captured once and replayed
    auto rest = [&]() {
        batch_split(l, P->dev, P->pb, key_bits, P->keys_sorted, P->vals_sorted);
        batch_cells(l, P->dev, P->pb);
        batch_permute(l, P->dev, P->pb);
        batch_pack(l, P->dev, P->pb, P->mixed);
        batch_lists(l, P->dev, P->lb, screened, plain);
        batch_pairs(l, P->dev, P->qb, P->mixed, fuse, lead->lut);
        if (!fuse) batch_finalize_correct(l, P->dev, P->pb);
    };
    if (P->tune.batch_graph) {
        if (!P->exec || P->graph_bits != key_bits || P->graph_screened != int(screened) ||
            P->graph_plain != int(plain)) {
            P->drop_graph();
            HIPCHK(lead, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            rest();
            HIPCHK(lead, hipStreamEndCapture(st, &P->graph));
            HIPCHK(lead, hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
            P->graph_bits = key_bits;
            P->graph_screened = int(screened);
            P->graph_plain = int(plain);
        }
        HIPCHK(lead, hipGraphLaunch(P->exec, st));
    } else {
        rest();
    }
    HIPCHK(lead, hipMemcpyAsync(P->small_host, P->small, count * SMALL_WORDS * sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, st));
    if (int e = sync_check(lead)) return e;
    for (size_t k = 0; k < count; ++k) {
        osph_ctx* c = ms[k];
        if (P->host[k].reorder) swap_state(c, P->host[k].s2);
        const unsigned long long* sh = P->small_host + SMALL_WORDS * k;
        uint32_t flag;
        std::memcpy(&flag, &sh[SMALL_OVERFLOW], sizeof(uint32_t));
        if (!flag) {
            select_timestep(c, sh + SMALL_CRITERIA, next_dt ? next_dt + k : nullptr, crit ? crit + k : nullptr);
            c->neighbors_valid = true;
            continue;
        }
        // a list overflowed: the member's finalize left its predicted state, which its own stream now
        // finishes with larger lists, as osph_step_pc does
        launch_init_small(L(c), c->small);
        if (int e = grow_lists(c)) return e;
        if (int e = launch_lists(c)) return e;
        launch_second_half(c, dts[k]);
        if (int e = step_end(c, dts[k], next_dt ? next_dt + k : nullptr, crit ? crit + k : nullptr)) return e;
    }
    return OSPH_OK;
}

}  // namespace

extern "C" {

int osph_step_pc_batch(osph_ctx** ctxs, int count, const double* dt, double* next_dt, int* crit) {
    if (count < 0 || (count > 0 && (!ctxs || !dt))) return OSPH_ERR_ARG;
    for (int k = 0; k < count; ++k) {
        if (!ctxs[k]) return OSPH_ERR_ARG;
        for (int j = 0; j < k; ++j) {
            if (ctxs[j] == ctxs[k]) return ctxs[k]->fail(OSPH_ERR_ARG, "a context appears twice in the batch");
        }
    }
    // one launch per stage when the members with particles share the precision and the lanes per particle
    // (which fix the order of the pair sums) and none has self-gravity attached
    std::vector<osph_ctx*> ms;
    std::vector<double> dts;
    std::vector<int> at;
    bool batched = true;
    for (int k = 0; k < count; ++k) {
        osph_ctx* c = ctxs[k];
        if (!(dt[k] > 0.0) || !std::isfinite(dt[k])) return c->fail(OSPH_ERR_ARG, "time step must be positive");
        if (!c->n) continue;
        if (!ms.empty() && (c->cfg.precision != ms[0]->cfg.precision || c->tune.pairs_lanes != ms[0]->tune.pairs_lanes)) {
            batched = false;
        }
        if (c->grav) batched = false;
        ms.push_back(c);
        dts.push_back(dt[k]);
        at.push_back(k);
    }
    if (batched && ms.size() > 1) {
        std::vector<double> nd(ms.size());
        std::vector<int> nc(ms.size());
        if (int e = batch_step(ms, dts, nd.data(), nc.data())) return e;
        for (int k = 0; k < count; ++k) {
            if (ctxs[k]->n) continue;
            if (int e = step_end(ctxs[k], dt[k], next_dt ? next_dt + k : nullptr, crit ? crit + k : nullptr)) return e;
        }
        for (size_t j = 0; j < ms.size(); ++j) {
            if (next_dt) next_dt[at[j]] = nd[j];
            if (crit) crit[at[j]] = nc[j];
        }
        return OSPH_OK;
    }
    for (int k = 0; k < count; ++k) {
        if (int e = step_begin(ctxs[k], dt[k])) return e;
    }
    for (int k = 0; k < count; ++k) {
        if (int e = step_middle(ctxs[k], dt[k])) return e;
    }
    for (int k = 0; k < count; ++k) {
        if (int e = step_end(ctxs[k], dt[k], next_dt ? next_dt + k : nullptr, crit ? crit + k : nullptr)) return e;
    }
    return OSPH_OK;
}

int osph_set_flaws(osph_ctx* c, size_t first, size_t count, double volume, double k_weibull, double m_weibull,
                    double growth_speed, uint64_t seed) {
    if (!c) return OSPH_ERR_ARG;
    if (first > c->n || count > c->n - first) return c->fail(OSPH_ERR_ARG, "particle range outside the state");
    if (!c->cfg.has_damage) return c->fail(OSPH_ERR_STATE, "the state holds no damage quantities");
    if (!(volume > 0.0) || !(k_weibull > 0.0) || !(m_weibull > 0.0) || !(growth_speed >= 0.0)) {
        return c->fail(OSPH_ERR_ARG, "volume, Weibull coefficient and exponent must be positive");
    }
    if (count == 0) return OSPH_OK;
    if (count >= (size_t(1) << 27)) return c->fail(OSPH_ERR_ARG, "a body must hold fewer than 2^27 particles");
    const uint32_t nb = uint32_t(count);
    Launch l = L(c);
    DevBuf<uint32_t> buf;
    struct StreamIdle {
        hipStream_t s;
        ~StreamIdle() { (void)hipStreamSynchronize(s); }
    } idle{ c->stream };
    HIPCHK(c, buf.alloc(3 * size_t(nb) + 2));
    uint32_t *fst = buf.p, *lst = buf.p + nb, *cnt = buf.p + 2 * size_t(nb), *acc = buf.p + 3 * size_t(nb);
    launch_clear_u32(l, fst, nb, 0xffffffffu);
    launch_clear_u32(l, lst, 2 * size_t(nb), 0u);  // last and count
    // draw flaws in batches until every particle holds one; the draw that flaws the last particle ends
    // the sequence (K = the largest first flaw), as ScalarGradyKippModel::setFlaws stops there
    uint64_t drawn = 0;
    uint32_t host[2] = { 0, 0 };
    for (;;) {
        const uint64_t batch = 4 * uint64_t(nb);
        if (drawn + batch >= 0xffffffffull) return c->fail(OSPH_ERR_STATE, "flaw count exceeds 2^32");
        flaws_first_pass(l, seed, drawn + 1, batch, nb, fst);
        drawn += batch;
        flaws_cover(l, fst, nb, acc);
        HIPCHK(c, hipMemcpyAsync(host, acc, sizeof(host), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (host[0] == 0) break;
    }
    const uint32_t K = host[1];
    flaws_count_pass(l, seed, K, nb, cnt, lst);
    const double denom = 1.0 / (std::pow(k_weibull, 1.0 / m_weibull) * std::pow(volume, 1.0 / m_weibull));
    flaws_finish(l, c->s, c->slot_of, first, nb, fst, lst, cnt, denom, 1.0 / m_weibull, growth_speed);
    return sync_check(c);
}

int osph_setup_spheres(osph_ctx* c, const osph_sphere_body* bodies, int count, const osph_collision* collision,
                       osph_body_info* info) {
    if (!c) return OSPH_ERR_ARG;
    if (count < 0 || (count > 0 && !bodies)) return c->fail(OSPH_ERR_ARG, "bodies must be given");
    constexpr double PI = 3.14159265358979323846264338327950288;  // core/math/MathUtils.h:379
    // the lattice of each body, as HexagonalPacking::generate lays it out (host: O(n^(1/3)) axis values)
    struct Plan {
        double h, dx, dy, dz;
        std::vector<double> axis[3];
        HcpLattice P{};
        size_t block0 = 0, count = 0, first = 0;
        double h_eta = 0.0, m = 0.0;
    };
    std::vector<Plan> plans(static_cast<size_t>(count));
    size_t blocks = 0, table = 0;
    for (int b = 0; b < count; ++b) {
        const osph_sphere_body& B = bodies[b];
        Plan& p = plans[size_t(b)];
        if (!(B.particle_count >= 1.0) || !(B.radius > 0.0) || !std::isfinite(B.radius) || !(B.eta > 0.0) ||
            !std::isfinite(B.center[0] + B.center[1] + B.center[2])) {
            return c->fail(OSPH_ERR_ARG, "a body needs a positive particle count, radius and eta");
        }
        if (B.material < 0 || size_t(B.material) >= c->mats.size()) {
            return c->fail(OSPH_ERR_ARG, "body material outside the material table");
        }
        // Distribution.cpp:131-138: the interparticle distance of the requested density
        const double volume = 1.3333333333333333333333 * PI * (B.radius * B.radius * B.radius);
        const double density = B.particle_count / volume;
        p.h = 1.0 / std::cbrt(density);
        p.dx = 1.1 * p.h;
        p.dy = std::sqrt(3.0) * 0.5 * p.dx;
        p.dz = std::sqrt(6.0) / 3.0 * p.dx;
        const double step[3] = { p.dx, p.dy, p.dz };
        size_t pts = 1;
        for (int a = 0; a < 3; ++a) {
            // Box(lower + 0.5 step, upper).iterateWithIndices: v = lower; v <= upper; v += step
            const double upper = B.center[a] + B.radius;
            double v = (B.center[a] - B.radius) + 0.5 * step[a];
            for (; v <= upper; v += step[a]) {
                p.axis[a].push_back(v);
                if (p.axis[a].size() > (size_t(1) << 20)) return c->fail(OSPH_ERR_ARG, "lattice too large");
            }
            pts *= std::max<size_t>(p.axis[a].size(), 1);
            if (p.axis[a].empty()) pts = 0;
        }
        if (pts > (size_t(1) << 33)) return c->fail(OSPH_ERR_ARG, "lattice too large");
        p.P.nx = p.axis[0].size();
        p.P.ny = p.axis[1].size();
        p.P.nz = p.axis[2].size();
        p.P.points = pts;
        p.P.delta_x = 0.5 * p.dx;
        p.P.delta_y = std::sqrt(3.0) / 6.0 * p.dx;
        p.P.cx = B.center[0];
        p.P.cy = B.center[1];
        p.P.cz = B.center[2];
        p.P.r2 = B.radius * B.radius;
        p.block0 = blocks;
        blocks += pts ? hcp_blocks(p.P) : 0;
        table += p.P.nx + p.P.ny + p.P.nz;
    }
    Launch l = L(c);
    DevBuf<double> axes, sums_dev;
    DevBuf<uint32_t> bcount;
    DevBuf<SerialSum> jobs_dev;
    DevBuf<unsigned long long> far;
    struct StreamIdle {  // the stream is idle before the buffers above are freed
        hipStream_t s;
        ~StreamIdle() { (void)hipStreamSynchronize(s); }
    } idle{ c->stream };
    HIPCHK(c, axes.alloc(table));
    HIPCHK(c, bcount.alloc(blocks));
    {
        std::vector<double> host;
        host.reserve(table);
        for (Plan& p : plans) {
            const size_t at = host.size();
            for (int a = 0; a < 3; ++a) host.insert(host.end(), p.axis[a].begin(), p.axis[a].end());
            p.P.xs = axes.p + at;
            p.P.ys = p.P.xs + p.P.nx;
            p.P.zs = p.P.ys + p.P.ny;
        }
        if (table) HIPCHK(c, hipMemcpy(axes.p, host.data(), table * sizeof(double), hipMemcpyHostToDevice));
    }
    for (const Plan& p : plans) {
        if (p.P.points) launch_hcp_count(l, p.P, bcount.p + p.block0);
    }
    std::vector<uint32_t> bfirst(blocks);
    if (blocks) {
        HIPCHK(c, hipMemcpyAsync(bfirst.data(), bcount.p, blocks * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    // each block's first slot: the bodies one after another, each in lattice order
    size_t total = 0;
    for (Plan& p : plans) {
        p.first = total;
        const size_t nb = p.P.points ? hcp_blocks(p.P) : 0;
        for (size_t k = 0; k < nb; ++k) {
            const uint32_t in = bfirst[p.block0 + k];
            bfirst[p.block0 + k] = uint32_t(total);
            total += in;
            if (total > 0x7fffffffull) return c->fail(OSPH_ERR_ARG, "the bodies hold more than 2^31-1 particles");
        }
        p.count = total - p.first;
        if (p.count == 0) return c->fail(OSPH_ERR_ARG, "a body holds no lattice point");
    }
    if (int e = osph_resize(c, total)) return e;
    if (blocks) {
        HIPCHK(c, hipMemcpy(bcount.p, bfirst.data(), blocks * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    for (const Plan& p : plans) {
        if (p.P.points) launch_hcp_fill(l, p.P, bcount.p + p.block0, c->s, p.h);
    }
    // left-to-right sums on the device (one thread each); `jobs` are read back as `sums`
    HIPCHK(c, jobs_dev.alloc(size_t(3) * count));
    HIPCHK(c, sums_dev.alloc(size_t(3) * count));
    std::vector<double> sums(size_t(3) * count);
    auto serial = [&](const double* weight_plane) -> int {
        std::vector<SerialSum> jobs;
        for (const Plan& p : plans) {
            for (int a = 0; a < 3; ++a) {
                jobs.push_back({ c->s.d[X + a] + p.first, weight_plane ? weight_plane + p.first : nullptr, p.count });
            }
        }
        HIPCHK(c, hipMemcpyAsync(jobs_dev.p, jobs.data(), jobs.size() * sizeof(SerialSum), hipMemcpyHostToDevice,
                                 c->stream));
        launch_serial_sums(l, jobs_dev.p, int(jobs.size()), sums_dev.p);
        HIPCHK(c, hipMemcpyAsync(sums.data(), sums_dev.p, jobs.size() * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return OSPH_OK;
    };
    // CENTER (Distribution.cpp:181-193): the lattice's centre of mass onto the domain centre; then the
    // quantities of InitialConditions::setQuantities (h = eta h, masses ~ h^3 normalised to rho0 V)
    if (int e = serial(nullptr)) return e;
    for (int b = 0; b < count; ++b) {
        const osph_sphere_body& B = bodies[b];
        Plan& p = plans[size_t(b)];
        const osph_material& mat = c->mats[size_t(B.material)];
        BodyFill f{};
        f.move = 1;
        f.fill = 1;
        for (int a = 0; a < 3; ++a) f.t[a] = B.center[a] - sums[size_t(3 * b + a)] / double(p.count);
        p.h_eta = p.h * B.eta;
        const double m0 = p.h_eta * p.h_eta * p.h_eta;
        double prelim = 0.0;
        for (size_t k = 0; k < p.count; ++k) prelim += m0;
        const double total_mass = 1.3333333333333333333333 * PI * (B.radius * B.radius * B.radius) * mat.rho0;
        p.m = m0 * (total_mass / prelim);
        f.h = p.h_eta;
        f.m = p.m;
        f.rho = mat.rho0;
        for (int a = 0; a < 3; ++a) f.v[a] = B.velocity[a];
        f.flag = B.flag;
        f.mat = uint32_t(B.material);
        launch_body_finish(l, c->s, p.first, p.count, f);
    }
    if (collision && count == 2) {
        // getBoundingSphere (ParticleJobs.cpp:499-510): centre of mass sum(m r) / sum(m), largest distance
        if (int e = serial(c->s.d[M])) return e;
        double ctr[2][3], rad[2];
        HIPCHK(c, far.alloc(2));
        launch_clear_u32(l, reinterpret_cast<uint32_t*>(far.p), 4, 0u);
        for (int b = 0; b < 2; ++b) {
            const Plan& p = plans[size_t(b)];
            double msum = 0.0;
            for (size_t k = 0; k < p.count; ++k) msum += p.m;
            for (int a = 0; a < 3; ++a) ctr[b][a] = sums[size_t(3 * b + a)] / msum;
            launch_max_dist2(l, c->s.d[X] + p.first, c->s.d[Y] + p.first, c->s.d[Z] + p.first, p.count, ctr[b],
                             far.p + b);
        }
        unsigned long long d2[2];
        HIPCHK(c, hipMemcpyAsync(d2, far.p, sizeof(d2), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int b = 0; b < 2; ++b) rad[b] = d2[b] ? std::sqrt(decode_ord_host(d2[b])) : 0.0;
        // the target's centre to the origin; the impactor to the impact angle (ParticleJobs.cpp:545-560)
        const double dist = rad[0] + rad[1];
        const double phi = collision->impact_angle_deg * (PI / 180.0);
        const double x = dist * std::cos(phi) + collision->impactor_offset_h * plans[0].h_eta;
        const double y = dist * std::sin(phi);
        BodyFill ft{}, fi{};
        ft.move = fi.move = 1;
        for (int a = 0; a < 3; ++a) ft.t[a] = -ctr[0][a];
        fi.t[0] = -ctr[1][0] + x;
        fi.t[1] = -ctr[1][1] + y;
        fi.t[2] = -ctr[1][2] + 0.0;
        launch_body_finish(l, c->s, plans[0].first, plans[0].count, ft);
        launch_body_finish(l, c->s, plans[1].first, plans[1].count, fi);
    }
    // ContinuityEquation::create: the density floor min_i m_i W(0) / h_i^3 (constant within each body)
    constexpr double W0 = 1.0 / PI;  // CubicSpline<3>::valueImpl(0) = (0.25 * 2^3 - 1) / pi
    double limit = double(1.e20f);   // LARGE
    for (const Plan& p : plans) limit = std::min(limit, p.m * W0 / (p.h_eta * p.h_eta * p.h_eta));
    if (count > 0) {
        for (osph_material& m : c->mats) m.rho_lo = std::max(limit, m.rho_lo);
        HIPCHK(c, hipMemcpyAsync(c->mats_dev, c->mats.data(), c->mats.size() * sizeof(osph_material),
                                 hipMemcpyHostToDevice, c->stream));
    }
    if (info) {
        for (int b = 0; b < count; ++b) {
            const Plan& p = plans[size_t(b)];
            const double rho0 = c->mats[size_t(bodies[b].material)].rho0;
            double vol = 0.0;
            for (size_t k = 0; k < p.count; ++k) vol += p.m / rho0;
            info[b] = { p.first, p.count, vol, p.h_eta, p.m };
        }
    }
    return sync_check(c);
}

int osph_attach_gravity(osph_ctx* c, osph_gravity* g) {
    if (!c) return OSPH_ERR_ARG;
    if (g && c->adj.ready) return c->fail(OSPH_ERR_UNSUPPORTED, "the adjoint has no self-gravity");
    if (g && gravity_owner(g) && gravity_owner(g) != c) {
        return c->fail(OSPH_ERR_ARG, "the gravity context is attached to another SPH context");
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->grav && c->grav != g && gravity_owner(c->grav) == c) gravity_set_owner(c->grav, nullptr);
    if (g) gravity_set_owner(g, c);
    if (g && !c->grav_ready) {
        HIPCHK(c, hipEventCreateWithFlags(&c->grav_ready, hipEventDisableTiming));
        HIPCHK(c, hipEventCreateWithFlags(&c->grav_done, hipEventDisableTiming));
    }
    c->grav = g;
    c->prm.add_accel = g ? 1 : 0;
    c->neighbors_valid = false;
    return OSPH_OK;
}

int osph_tune(osph_ctx* c, const char* key, long long value) {
    if (!c || !key) return OSPH_ERR_ARG;
    const std::string k(key);
    const bool block_ok = value == 64 || value == 128 || value == 256;
    const bool lds_ok = value >= 0 && value <= 65536;
    if (k == "pairs_block" && block_ok) {
        c->tune.pairs_block = int(value);
    } else if (k == "lists_block" && block_ok) {
        c->tune.lists_block = int(value);
    } else if (k == "pairs_lds" && lds_ok) {
        c->tune.pairs_lds = int(value);
    } else if (k == "lists_lds" && lds_ok) {
        c->tune.lists_lds = int(value);
    } else if (k == "pairs_lanes" && (value == 1 || value == 2 || value == 4)) {
        c->tune.pairs_lanes = int(value);
    } else if (k == "cell_divisor" && value >= 1 && value <= 4) {
        c->tune.cell_divisor = int(value);
    } else if (k == "lists_screen" && (value == -1 || value == 0 || value == 1)) {
        c->tune.lists_screen = int(value);
    } else if (k == "reorder_period" && value >= 0 && value <= 1000000) {
        c->tune.reorder_period = int(value);
    } else if (k == "fuse_finalize" && (value == -1 || value == 0 || value == 1)) {
        c->tune.fuse_finalize = int(value);
    } else if (k == "batch_graph" && (value == 0 || value == 1)) {
        c->tune.batch_graph = int(value);
    } else {
        return c->fail(OSPH_ERR_ARG, "unknown tuning key or value out of range: " + k);
    }
    return OSPH_OK;
}

int osph_autotune(osph_ctx* c, double dt, int steps, const char* const* keys, const int* n_values,
                  const long long* values, int n_keys, long long* chosen, double* best_ms, double* trial_ms,
                  int trial_cap, int* n_trials) {
    if (!c) return OSPH_ERR_ARG;
    if (!(dt > 0.0) || !std::isfinite(dt) || steps < 1) return c->fail(OSPH_ERR_ARG, "autotune needs dt > 0 and steps >= 1");
    if (c->adj.ready) return c->fail(OSPH_ERR_STATE, "autotune would change a recorded adjoint run");
    // the search space: the given keys, or the default one (the first value of each key starts the descent)
    static const char* const default_keys[] = { "pairs_lanes", "cell_divisor", "lists_block", "pairs_block",
                                                "fuse_finalize" };
    static const int default_counts[] = { 3, 3, 2, 2, 3 };
    static const long long default_values[] = { 1, 2, 4, 3, 2, 4, 256, 128, 256, 128, 1, 0, -1 };
    if (!keys) {
        keys = default_keys;
        n_values = default_counts;
        values = default_values;
        n_keys = 5;
    }
    if (n_keys < 1 || !n_values || !values) return c->fail(OSPH_ERR_ARG, "autotune needs a search space");
    std::vector<size_t> first(size_t(n_keys) + 1, 0);
    for (int k = 0; k < n_keys; ++k) {
        if (!keys[k] || n_values[k] < 1) return c->fail(OSPH_ERR_ARG, "autotune: every key needs a value");
        first[size_t(k) + 1] = first[size_t(k)] + size_t(n_values[k]);
    }
    const Tuning saved_tune = c->tune;
    std::vector<long long> best(static_cast<size_t>(n_keys));
    for (int k = 0; k < n_keys; ++k) {
        best[size_t(k)] = values[first[size_t(k)]];
        if (int e = osph_tune(c, keys[k], best[size_t(k)])) {
            c->tune = saved_tune;
            return e;
        }
    }
    if (c->n == 0) {
        if (chosen) std::copy(best.begin(), best.end(), chosen);
        if (best_ms) *best_ms = 0.0;
        if (n_trials) *n_trials = 0;
        return OSPH_OK;
    }
    // the resident state (the current block, the slot maps and the search's reference h), restored after every
    // candidate so that each candidate steps from the same state and the tuning leaves the run where it was
    DevBuf<double> dsave;
    DevBuf<uint32_t> usave;
    const size_t cap = c->cap;
    HIPCHK(c, dsave.alloc(size_t(NDPLANES) * cap));
    HIPCHK(c, usave.alloc(size_t(NIPLANES + 2) * cap));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpyAsync(dsave.p, c->dblock, sizeof(double) * NDPLANES * cap, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(usave.p, c->ublock, sizeof(uint32_t) * NIPLANES * cap, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(usave.p + size_t(NIPLANES) * cap, c->perm, sizeof(uint32_t) * cap, hipMemcpyDeviceToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(usave.p + size_t(NIPLANES + 1) * cap, c->slot_of, sizeof(uint32_t) * cap,
                             hipMemcpyDeviceToDevice, c->stream));
    const int since = c->since_reorder;
    const double href = c->href;
    // the plane order within the block (the prediction planes trade places with the derivative planes)
    std::vector<ptrdiff_t> doff(NDPLANES);
    for (int k = 0; k < NDPLANES; ++k) doff[k] = c->s.d[k] - c->dblock;
    auto restore = [&]() -> int {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpyAsync(c->dblock, dsave.p, sizeof(double) * NDPLANES * cap, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->ublock, usave.p, sizeof(uint32_t) * NIPLANES * cap, hipMemcpyDeviceToDevice,
                                 c->stream));
        HIPCHK(c, hipMemcpyAsync(c->perm, usave.p + size_t(NIPLANES) * cap, sizeof(uint32_t) * cap,
                                 hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->slot_of, usave.p + size_t(NIPLANES + 1) * cap, sizeof(uint32_t) * cap,
                                 hipMemcpyDeviceToDevice, c->stream));
        for (int k = 0; k < NDPLANES; ++k) c->s.d[k] = c->dblock + doff[k];
        for (int k = 0; k < NIPLANES; ++k) c->s.u[k] = c->ublock + size_t(k) * cap;
        c->since_reorder = since;
        c->href = href;
        c->neighbors_valid = false;
        return sync_check(c);
    };
    // mean wall time of `steps` PC steps after one untimed step (which sizes buffers for the settings)
    auto time_steps = [&](double* ms) -> int {
        double next = 0.0;
        int crit = 0;
        if (int e = osph_step_pc(c, dt, &next, &crit)) return e;
        const auto t0 = std::chrono::steady_clock::now();
        for (int k = 0; k < steps; ++k) {
            if (int e = osph_step_pc(c, dt, &next, &crit)) return e;
        }
        *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps;
        return restore();
    };
    int trials = 0;
    auto record = [&](double ms) {
        if (trial_ms && trials < trial_cap) trial_ms[trials] = ms;
        ++trials;
    };
    auto apply = [&](const std::vector<long long>& v) -> int {
        for (int k = 0; k < n_keys; ++k) {
            if (int e = osph_tune(c, keys[k], v[size_t(k)])) return e;
        }
        return OSPH_OK;
    };
    auto fail_restore = [&](int e) {
        const std::string msg = c->err;
        (void)restore();
        c->tune = saved_tune;
        c->err = msg;
        return e;
    };
    double warm = 0.0, best_t = 0.0;
    // the starting settings once untimed-to-time (the first steps on a state size its buffers), then timed
    if (int e = time_steps(&warm)) return fail_restore(e);
    if (int e = time_steps(&best_t)) return fail_restore(e);
    record(best_t);
    for (int k = 0; k < n_keys; ++k) {
        for (size_t j = first[size_t(k)] + 1; j < first[size_t(k) + 1]; ++j) {
            std::vector<long long> trial = best;
            trial[size_t(k)] = values[j];
            if (int e = apply(trial)) return fail_restore(e);
            double t = 0.0;
            if (int e = time_steps(&t)) return fail_restore(e);
            record(t);
            if (t < best_t) {
                best = trial;
                best_t = t;
            }
        }
        if (int e = apply(best)) return fail_restore(e);
    }
    if (chosen) std::copy(best.begin(), best.end(), chosen);
    if (best_ms) *best_ms = best_t;
    if (n_trials) *n_trials = trials;
    return OSPH_OK;
}

int osph_sync(osph_ctx* c) {
    if (!c) return OSPH_ERR_ARG;
    return sync_check(c);
}

int osph_profile_enable(osph_ctx* c, int enable) {
    if (!c) return OSPH_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    resolve_events(c);
    if (enable < 0 || enable > 2) return c->fail(OSPH_ERR_ARG, "profile mode must be 0, 1 or 2");
    c->prof = enable;
    c->prof_acc.clear();
    return OSPH_OK;
}

int osph_profile_read(osph_ctx* c, char* names, size_t name_cap, double* ms, uint64_t* launches, int cap,
                      int* count) {
    if (!c || !count) return OSPH_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    resolve_events(c);
    int k = 0;
    for (const auto& kv : c->prof_acc) {
        if (k < cap) {
            if (names && name_cap) std::snprintf(names + size_t(k) * name_cap, name_cap, "%s", kv.first.c_str());
            if (ms) ms[k] = kv.second.first;
            if (launches) launches[k] = kv.second.second;
        }
        ++k;
    }
    *count = k;
    return OSPH_OK;
}

// ---- reverse mode ------------------------------------------------------------------------------------

// One PC step with a given time step (the criterion's proposal is ignored).
static int step_fixed(osph_ctx* c, double dt) {
    if (int e = step_begin(c, dt)) return e;
    if (int e = step_middle(c, dt)) return e;
    return step_end(c, dt, nullptr, nullptr);
}

// The prediction of a step with time step dt and the derivatives at the predicted state (lists, pair sums
// and finalize, no correction): what the reverse step differentiates.
static int predict_and_evaluate(osph_ctx* c, double dt) {
    if (int e = step_begin(c, dt)) return e;
    if (int e = search_finish(c)) return e;
    if (int e = build_lists(c)) return e;
    launch_pair_sums(c);
    timed(c, "finalize", [&] { launch_finalize(L(c), c->s, c->mats_dev, c->prm, c->n); });
    return OSPH_OK;
}

int osph_adjoint_forward(osph_ctx* c, int steps, const double* dt, int checkpoint_every) {
    if (!c || steps < 1 || !dt || checkpoint_every < 1) return OSPH_ERR_ARG;
    if (c->cfg.solid_stress || c->cfg.has_damage || c->cfg.adaptive_h) {
        return c->fail(OSPH_ERR_UNSUPPORTED,
                       "the adjoint covers the gas terms: no solid stress, damage or adaptive smoothing length");
    }
    if (c->mats.size() > size_t(ADJ_MAX_MATERIALS)) return c->fail(OSPH_ERR_UNSUPPORTED, "too many materials");
    if (c->grav) return c->fail(OSPH_ERR_UNSUPPORTED, "the adjoint has no self-gravity: detach it first");
    for (int k = 0; k < steps; ++k) {
        if (!(dt[k] > 0.0) || !std::isfinite(dt[k])) return c->fail(OSPH_ERR_ARG, "time steps must be positive");
    }
    const size_t n = c->n;
    if (!c->adj.reorder_saved) {
        c->adj.saved_reorder = c->tune.reorder_period;
        c->adj.reorder_saved = true;
    }
    c->tune.reorder_period = 0;  // the slot order stays fixed between the checkpoints and the replays
    c->adj.ready = false;
    const size_t blocks = size_t((steps + checkpoint_every - 1) / checkpoint_every);
    if (blocks > c->adj.ckpt_blocks || c->cap > c->adj.ckpt_cap) {
        (void)hipFree(c->adj.ckpt);
        c->adj.ckpt = nullptr;
        c->adj.ckpt_blocks = 0;
        HIPCHK(c, hipMalloc(&c->adj.ckpt, blocks * 13 * std::max<size_t>(c->cap, 1) * sizeof(double)));
        c->adj.ckpt_blocks = blocks;
        c->adj.ckpt_cap = c->cap;
    }
    const size_t block_doubles = 13 * c->adj.ckpt_cap;
    for (int k = 0; k < steps; ++k) {
        if (k % checkpoint_every == 0) {
            launch_save_x(L(c), c->s, c->adj.ckpt + size_t(k / checkpoint_every) * block_doubles, c->adj.ckpt_cap, n);
        }
        if (int e = step_fixed(c, dt[k])) return e;
    }
    c->adj.steps = steps;
    c->adj.every = checkpoint_every;
    c->adj.dt.assign(dt, dt + steps);
    c->adj.ready = true;
    return sync_check(c);
}

int osph_adjoint_backward(osph_ctx* c, const double* seed_r, const double* seed_v, const double* seed_rho,
                          const double* seed_u, double* grad_r, double* grad_v, double* grad_rho, double* grad_u,
                          double* grad_eos, unsigned* clamped) {
    if (!c) return OSPH_ERR_ARG;
    if (!c->adj.ready) return c->fail(OSPH_ERR_STATE, "no forward run recorded (osph_adjoint_forward)");
    const size_t n = c->n, cap = c->adj.ckpt_cap;
    const int steps = c->adj.steps, every = c->adj.every;
    if (cap < n) return c->fail(OSPH_ERR_STATE, "the particle count changed since the forward run");
    Launch l = L(c);
    DevBuf<double> lam, gsc, seg, io, part, theta;
    DevBuf<unsigned char> rec;
    DevBuf<unsigned> clamp_count;
    // destroyed before the buffers: no queued kernel still uses them when they are freed (early returns)
    struct StreamIdle {
        hipStream_t s;
        ~StreamIdle() { (void)hipStreamSynchronize(s); }
    } idle{ c->stream };
    HIPCHK(c, lam.alloc(13 * cap));
    HIPCHK(c, gsc.alloc(9 * cap));
    HIPCHK(c, seg.alloc(size_t(every) * 13 * cap));
    HIPCHK(c, io.alloc(3 * std::max<size_t>(n, 1)));
    HIPCHK(c, part.alloc(adj_partials(n)));
    HIPCHK(c, theta.alloc(ADJ_MAX_MATERIALS * 3));
    HIPCHK(c, rec.alloc(adj_rec_bytes() * std::max<size_t>(n, 1)));
    HIPCHK(c, clamp_count.alloc(1));
    double* lp[13];
    double* gp[9];
    for (int k = 0; k < 13; ++k) lp[k] = lam.p + size_t(k) * cap;
    for (int k = 0; k < 9; ++k) gp[k] = gsc.p + size_t(k) * cap;
    launch_clear_u32(l, reinterpret_cast<uint32_t*>(theta.p), 2 * ADJ_MAX_MATERIALS * 3, 0u);
    launch_clear_u32(l, clamp_count.p, 1, 0u);
    // seeds: lambda of r and v at the final state (caller order, 3 per particle), rho and u (1 each);
    // the derivative planes start at zero
    auto seed = [&](const double* host, int nc, double* const* planes) -> int {
        if (host && n) HIPCHK(c, hipMemcpyAsync(io.p, host, n * nc * sizeof(double), hipMemcpyHostToDevice, c->stream));
        launch_scatter_planes(l, host ? io.p : nullptr, nc, planes, c->slot_of, n);
        return sync_check(c);  // io is reused
    };
    if (int e = seed(seed_r, 3, lp + LRX_)) return e;
    if (int e = seed(seed_v, 3, lp + LVX_)) return e;
    if (int e = seed(nullptr, 3, lp + LAX_)) return e;
    if (int e = seed(seed_rho, 1, lp + LRHO_)) return e;
    if (int e = seed(nullptr, 1, lp + LDRHO_)) return e;
    if (int e = seed(seed_u, 1, lp + LU_)) return e;
    if (int e = seed(nullptr, 1, lp + LDU_)) return e;

    const size_t block_doubles = 13 * cap;
    const int nblocks = (steps + every - 1) / every;
    for (int b = nblocks - 1; b >= 0; --b) {
        const int s0 = b * every, s1 = std::min(s0 + every, steps);
        // replay the segment from its checkpoint, keeping the state before each step
        launch_restore_x(l, c->s, c->adj.ckpt + size_t(b) * block_doubles, cap, n);
        launch_save_x(l, c->s, seg.p, cap, n);
        for (int k = s0; k < s1 - 1; ++k) {
            if (int e = step_fixed(c, c->adj.dt[k])) return e;
            launch_save_x(l, c->s, seg.p + size_t(k + 1 - s0) * block_doubles, cap, n);
        }
        for (int k = s1 - 1; k >= s0; --k) {
            launch_restore_x(l, c->s, seg.p + size_t(k - s0) * block_doubles, cap, n);
            if (int e = predict_and_evaluate(c, c->adj.dt[k])) return e;
            timed(c, "adjoint", [&] {
                launch_adj_step(l, c->s, lp, gp, c->order, c->list, c->count, c->kmax, c->lut, c->cfg.av_alpha,
                                c->cfg.av_beta, c->mats_dev, int(c->mats.size()), c->adj.dt[k], rec.p, part.p,
                                theta.p, clamp_count.p, n, c->cfg.precision == OSPH_PRECISION_MIXED);
            });
        }
    }
    // gradients with respect to the initial state (caller order)
    auto out = [&](double* host, int nc, double* const* planes) -> int {
        if (!host || !n) return OSPH_OK;
        launch_gather_planes(l, io.p, nc, planes, c->slot_of, n);
        HIPCHK(c, hipMemcpyAsync(host, io.p, n * nc * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        return sync_check(c);
    };
    if (int e = out(grad_r, 3, lp + LRX_)) return e;
    if (int e = out(grad_v, 3, lp + LVX_)) return e;
    if (int e = out(grad_rho, 1, lp + LRHO_)) return e;
    if (int e = out(grad_u, 1, lp + LU_)) return e;
    if (grad_eos) {
        HIPCHK(c, hipMemcpyAsync(grad_eos, theta.p, c->mats.size() * 3 * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
    }
    unsigned nclamp = 0;
    HIPCHK(c, hipMemcpyAsync(&nclamp, clamp_count.p, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    // the resident state goes back to the initial state of the forward run
    launch_restore_x(l, c->s, c->adj.ckpt, cap, n);
    c->neighbors_valid = false;
    if (int e = sync_check(c)) return e;
    if (clamped) *clamped = nclamp;
    return OSPH_OK;
}

int osph_adjoint_clear(osph_ctx* c) {
    if (!c) return OSPH_ERR_ARG;
    if (c->adj.reorder_saved) c->tune.reorder_period = c->adj.saved_reorder;
    c->adj.reorder_saved = false;
    c->adj.ready = false;
    (void)hipFree(c->adj.ckpt);
    c->adj.ckpt = nullptr;
    c->adj.ckpt_blocks = 0;
    c->adj.ckpt_cap = 0;
    return OSPH_OK;
}

}  // extern "C"
