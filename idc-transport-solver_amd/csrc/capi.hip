// C-ABI of libtsw.so: context, problem set-up, sweep dispatch and the solver drivers that run on the
// device (source iteration, Peierls GMRES with FP64 or mixed-precision sweeps).
#include <algorithm>
#include <array>
#include <climits>
#include <mutex>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "tsw_internal.hpp"

using namespace tsw;

enum class Pattern { WavefrontGraph, WavefrontEager, Staggered, StaggeredGE };

static bool is_staggered(Pattern p) { return p == Pattern::Staggered || p == Pattern::StaggeredGE; }

constexpr int kXcds = 8;  // MI355X: workgroup b runs on XCD b mod 8, each XCD with its own L2

struct tsw_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;
    std::string err;
    bool have_problem = false;
    Pattern pattern = Pattern::Staggered;
    Geom g;
    std::vector<double> h_xi, h_eta, h_w;

    // ordinates on the device, both precisions
    DBuf xi64, eta64, w64, xi32, eta32, w32;
    // staging / work arrays (grown on demand)
    DBuf psi64, psi32, q64, r64, sig_t64, sig_s64, sig_t32, sig_s32, phi64a, phi64b, phi32, red_part, red_out;
    DBuf si_state;  // source-iteration angular flux, persists between solves
    DBuf ir_state;  // FP64 iterate of the mixed-precision refinement

    // sweep timing: events recorded around every sweep
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    long long sweeps_done = 0;
    double ms_folded = 0.0;  // elapsed time of event pairs already released

    // Material table: distinct sigma_t values of the interior cells, the per-cell material index and
    // the element-matrix inverses for the current dt, in both precisions (built on first use).
    struct Materials {
        bool valid = false;
        int M = 0;
        double inv_dt = -1.0;
        DBuf idx, sig, inv64, inv32;
        bool have64 = false, have32 = false;
        DBuf key;            // device copy of the sigma_t the table was built from
        bool keyed = false;  // key holds that copy
        unsigned long long build = 0;  // counts builds and cross-section changes
    } mat;
    DBuf same_flag;  // device flags of the input change checks
    // input checks made ahead in one round trip (precheck_keys), each consumed by the same_as_key call on
    // the same key, array and length within the call that made them
    struct KeyCheck {
        const DBuf *key;
        const double *a;
        size_t n;
        bool same;
    };
    std::vector<KeyCheck> key_checks;
    // Krylov workspace kept across solves: basis, Gram-Schmidt coefficients, partial sums, the update
    // coefficients, and the solver's right-hand side, iterate and residual
    DBuf gm_v, gm_h, gm_part, gm_y, gm_vecs;
    // Cross-section degree (tsw_set_xs_degree): 0, cell means; p, Legendre coefficients per cell, used
    // through their triple-product tensors, built once per call and per coefficient array (two slots,
    // keyed by device pointer, cleared at every entry point), plus their cell means for DSA.
    int xs_deg = 0;
    struct XsCache {
        const double *key[2] = {nullptr, nullptr};
        DBuf t64[2], t32[2], mean[2];
        bool have32[2] = {false, false}, have_mean[2] = {false, false};
        int next = 0;
    } xs;
    // relative errors of an FP32-sweep GMRES product, measured per material table build `mat_build`, 1/dt
    // and DSA setting (an IDC step solves with a few different substep lengths), the latest kFp32ErrKeep
    struct Fp32Err {
        double value = -1.0;
        unsigned long long mat_build = 0;
        double inv_dt = -1.0;
        bool dsa = false;
    };
    static constexpr size_t kFp32ErrKeep = 16;
    std::vector<Fp32Err> fp32_err;
    static constexpr int kMaxMaterials = 64;

    // Inputs and work arrays of the staggered (diagonal-layout) sweep, per precision where needed:
    // cross sections and material index per quadrant, inflow traces, the source and psi per ordinate,
    // the scalar flux per quadrant, the quadrant partial moments and the refinement iterate.
    struct Diag {
        DBuf sigt64, sigt32, sigs64, sigs32, mat, tinx64, tiny64, tinx32, tiny32, q64, q32, psi64, psi32, res32,
            phid64, phid32, phiQ, ir64;
        bool inflow = false;
    } diag;

    // captured sweep graphs keyed by (precision, psi, sigma, inv_dt)
    std::map<std::tuple<int, const void *, const void *, double>, hipGraphExec_t> graphs;

    void clear_graphs() {
        for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }
    // Batch of independent problems (tsw_batch_set): one sub-context per member sharing this stream,
    // launch groups by (DG degree, compute waves, material path), side streams for the groups.
    std::vector<std::unique_ptr<tsw_ctx>> batch;
    double batch_dt = 0.0;
    struct Group {
        int p = 0, waves = 0, max_ny = 0;
        bool mat = false;
        std::vector<int> members;       // batch indices
        std::vector<int> items;         // pairs {member slot, ordinate}, heaviest first
        DBuf params64, params32, params32c, dev_items;  // params32c: FP32 correction sweeps of TSW_MIXED
        bool built64 = false, built32 = false, built32c = false, mat64 = false, mat32 = false;
        bool direct64 = false, direct32 = false;  // direct-store layout (set with the parameter table)
        size_t lds64 = 0, lds32 = 0, lds32c = 0;
        // reference-layout I/O: members swept in place on the caller's arrays; the table is rebuilt for
        // each call's arrays
        bool ref = false;
        std::vector<unsigned char> ref_host;
        DBuf ref_params;
        size_t ref_lds = 0;
        bool ref_mat = false, ref_direct = false, ref_stage = false;
        // lane-split kernel (on the caller's layout, or on the diagonal layout when a call's arrays are not
        // 16-byte aligned): members in row bands of at most `rows` rows, items {slot, band * nq + ordinate}
        bool split = false;
        int rows = 0;
        DBuf work;  // persistent launch: the item counter
        std::vector<int> band_rows;  // per member slot, fixed when the items are built
    };
    // Engine of each batch member, chosen per DG degree (tsw_batch_set): the staggered kernels through
    // layout conversions, the staggered kernels on the caller's layout, the ordinate-lane engine, or the
    // lane-split staggered kernel on the caller's layout.
    enum : char { kEngConverted = 0, kEngRef = 1, kEngLanes = 2, kEngSplit = 3 };
    std::vector<char> engine, lanes_able, split_able;
    bool batch_ref_ok = false;  // every member's array admits 32-bit byte offsets
    int batch_engine[4] = {-1, -1, -1, -1};  // per DG degree present in the batch
    std::vector<int> batch_all;              // 0 .. nprob-1
    // Many right-hand sides of this problem (tsw_sweep_apply_multi): one staggered launch over
    // (source, ordinate) items sharing the material table and element inverses; per-source inflow traces.
    struct Multi {
        std::vector<unsigned char> host;
        std::vector<int> items;
        DBuf params, dev_items, traces, stage;
        DBuf xtr64, xtr32, work;  // lane-split sweeps: exchange records per source, persistent-launch counter
        size_t xtr64_n = 0, xtr32_n = 0;
    } multi;
    // Sweeps of this problem on the caller's layout by the lane-split kernel where it is built and its
    // tables fit (tuning key ref_split: 1 default, -1 off): tsw_sweep_apply (device arrays, 16-byte aligned)
    // and tsw_sweep_apply_multi
    int ref_split_opt = 1;
    std::vector<Group> groups;
    // Batch members on the ordinate-lane engine, swept in place on the caller's arrays: one launch per DG
    // degree over (member, ordinate block, band) items, band-major so that every band's upstream band is
    // dispatched first.
    struct LaneGroup {
        int p = 0;
        std::vector<int> members;
        std::vector<int4> items;
        std::vector<unsigned char> host;  // the parameter table on its way to the device
        DBuf params, dev_items, work;     // work: the item counter of a persistent launch
    };
    std::vector<LaneGroup> lane_groups;
    // Adjoint problem: the same mesh with the negated ordinate set in reverse order (its sweeps are the
    // transposed sweeps of this problem up to the Legendre mass scaling), sharing this stream; built on
    // the first vector-Jacobian product.  Staging of the vector-Jacobian products' arrays.
    std::unique_ptr<tsw_ctx> adj;
    DBuf vjp_psi, vjp_seed, vjp_gsrc, vjp_gst, vjp_gss, vjp_y, vjp_phi;
    // Row bands of the staggered sweep: band_rows_opt 0 = automatic, -1 = one band per ordinate,
    // > 0 = rows per band.  Exchange records per precision (sentinel-filled when empty), the stall flag
    // and its pinned host mirror, read back at every synchronisation when bands ran.
    int band_rows_opt = 0;
    int prefetch_opt = 0;  // prefetch depth of the staggered sweep (0: default)
    // Autotuner: with band_rows_opt and prefetch_opt both automatic, the first staggered sweep of each
    // kind (mesh, degree, ordinates, precision, source mode, material path) times the candidate
    // (band rows, prefetch) pairs and keeps the fastest; tuned_br/tuned_pf hold the choice in force
    // (INT_MIN: none), last_br/last_pf what the last sweep used.
    bool autotune = true;
    std::map<std::array<int, 7>, std::array<int, 4>> tuned;  // (band rows, prefetch, one wave, lane split)
    int tuned_br = INT_MIN, tuned_pf = INT_MIN, tuned_w1 = INT_MIN, tuned_sl = INT_MIN;
    int last_br = 0, last_pf = 0, last_w1 = 0, last_sl = 0;
    int one_wave_opt = 0;  // 1: one wave per band of 64 rows, lane-to-lane traces (tuning key one_wave)
    // Ordinate-lane engine (sweep_lanes.hpp) on the reference layout: tuning key "lanes" (-1 off, 0
    // automatic, 1 on where built), "lanes_block" (ordinates per row group, 0 automatic) and "lanes_waves"
    // (waves per workgroup, 0 automatic); the ordinate-block table for the block size in use, the inflow
    // traces and the per-block moment partials.
    int lanes_opt = 0, lanes_J_opt = 0, lanes_W_opt = 0, lanes_pf_opt = 0;
    bool lanes_fixed_W = false;  // batch members: every member's workgroup has lanes_W_opt waves
    struct Lanes {
        DBuf blk, mom, tinx64, tiny64, tinx32, tiny32, psi_keep;
        int J = 0, nblk = 0;
    } lanes;
    int last_lanes = 0, last_lanes_J = 0, last_lanes_W = 0;
    int split_opt = 0;     // 1: 2 or 4 lanes per row on the material path (tuning key lane_split)
    // batch launch groups: -1 storer wave unless a table only fits without it, else (0, 1) direct stores
    // whenever every member has element-inverse tables (tuning key batch_direct)
    int batch_direct_opt = 0;
    int batch_ref_opt = 0;   // batch engines: 1 / -1 force / forbid the reference-layout staggered groups
    // batch engines: 1 / -1 force / forbid the lane-split groups (tuning key batch_split); their rows per
    // band (batch_split_rows, 0: the most one workgroup holds) and prefetch depth (batch_split_prefetch)
    int batch_split_opt = 0, batch_split_rows_opt = 0, batch_split_pf_opt = 0;
    int split_rows_p[4] = {0, 0, 0, 0};  // rows per band of the lane-split engine per DG degree (autotuned; 0: most)
    // lane-split groups: runs of this many consecutive items go to one XCD (tuning key batch_xcd_tile;
    // 0 or 1: dispatch order)
    int batch_xcd_tile_opt = 0;
    // lane-split groups: one persistent launch per group, workgroups taking items from a device counter
    // (tuning key batch_persistent, 1 default; 0: one workgroup per item)
    int batch_persistent_opt = 1;
    int ref_stage_opt = 0;   // 1: reference-layout sweeps stage their source loads through LDS where it fits
    // Reflecting edges (tsw_set_boundaries): mask bits 1 x-min, 2 x-max, 4 y-min, 8 y-max; the mirror map
    // [2][nq] (OrdinateSet::ReflectMap) on the device
    int reflect_mask = 0;
    bool periodic = false;  // periodic boundaries (the deck key periodic, Global::periodic)
    DBuf reflect_map_d;
    bool last_batch_ref = false;
    DBuf tune_out;
    // Source iteration replayed from a captured graph of kSiChunk iterations, with the loop test on the
    // device (SiControl); skip_flag points at the control's done flag while such a graph is captured.
    bool si_graph = true;
    int refine_steps = 1;  // FP64 refinement steps of a TSW_MIXED sweep
    bool gmres_refine = false;  // mixed GMRES by iterative refinement instead of the relaxed strategy
    double fp32_safety = 10.0;  // relaxed GMRES: FP32 products once the estimate is <= abs_tol / (fp32_safety d)
    static constexpr int kSiChunk = 8;
    DBuf si_ctl;
    SiControl *si_ctl_host = nullptr;
    double *gm_host = nullptr;  // pinned: one Arnoldi step's Hessenberg column on its way to the host
    size_t gm_host_n = 0;
    const int *skip_flag = nullptr;
    // Diffusion synthetic acceleration of the Peierls GMRES (use_dsa, PETScPeierlsSolver.cpp:151):
    // multigrid hierarchy of the diffusion operator for the current cross sections, and the apply
    // y = x + C^{-1} sigma_s x captured as a graph over staging vectors.
    bool use_dsa = false;
    int dsa_cycles = 1;
    // IDC time integrator (tsw_idc_set / tsw_idc_step): nodes on [0,1] (stage 0 = left end), the
    // integration matrix W (row 0 over [0,1], row n over [t_{n-1}, t_n]) and the stage, residual and
    // correction vectors in the reference layout, FP64.
    struct Idc {
        int num_nodes = 4, num_corr = 3, kind = TSW_IDC_LOBATTO, form = TSW_IDC_ERROR;
        bool collocation = false, left = true, right = true;
        std::vector<double> nodes, W;  // S and S*S
        DBuf stages, res, err, phi;
        // adjoint: saved forward intermediates, stage adjoints, residual adjoints, work vectors,
        // checkpoints and the states of one checkpoint segment
        DBuf saved, lam, lamR, t1, t2, t3, ckpt, seg;
    } idc;
    struct Dsa {
        std::vector<MgLevel> lv;
        std::vector<DBuf> store;
        DBuf lv_dev;
        DBuf coarse_inv;       // dense inverse of the coarsest level's operator
        int fused_from = 0;    // first level handled by the single-workgroup kernel
        size_t fused_lds = 0;  // its LDS bytes
        const double *sigma_s = nullptr;
        hipGraphExec_t graph = nullptr;
        // what the hierarchy was built for (rebuilt when any of it changes): device copies of the cross
        // sections, compared on the device
        DBuf key_st, key_ss;
        bool keyed = false;
        double key_idt = -1.0;
        int key_cycles = 0;
        ~Dsa() {
            if (graph) (void)hipGraphExecDestroy(graph);
        }
    } dsa;
    DBuf xtr64, xtr32, abort_dev;
    size_t xtr64_n = 0, xtr32_n = 0;
    int *abort_host = nullptr;
    bool bands_used = false;
    // test hook (tuning key stall_test): the next banded sweep starts with the stall flag raised, so every
    // bounded wait of its bands gives up within one poll interval and the call reports the stall
    bool stall_test = false;
    static constexpr int kSideStreams = 4;
    hipStream_t side[kSideStreams] = {};
    hipEvent_t fork_ev = nullptr, join_ev[kSideStreams] = {};

    ~tsw_ctx() {
        clear_graphs();
        for (auto e : ev_pool) (void)hipEventDestroy(e);
        for (auto &sd : side)
            if (sd) (void)hipStreamDestroy(sd);
        for (auto &e : join_ev)
            if (e) (void)hipEventDestroy(e);
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        batch.clear();
        adj.reset();
        if (abort_host) (void)hipHostFree(abort_host);
        if (si_ctl_host) (void)hipHostFree(si_ctl_host);
        if (gm_host) (void)hipHostFree(gm_host);
        if (stream && own_stream) (void)hipStreamDestroy(stream);
    }
};

// ---------------------------------------------------------------------------------------------------
// Scratch of the allocation-free sweep (tsw_sweep_apply_noalloc): one file-scope device array, of which
// every GPU holds its own copy for the life of the process.  Host code never uses the array's name as an
// address: its per-device address is looked up with hipGetSymbolAddress when the first context is
// created on that device, and that pointer is what the kernels receive.
// ---------------------------------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(128) << 20;  // 128 MiB per GPU
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxScratchDevices = 64;
constexpr int kSplitBatchPrefetch = 4;  // prefetch depth of the batch's lane-split groups
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxScratchDevices] = {};

// Looks up the address of `scratch` on `device` (the current device) once; TSW_E_HIP on failure.
static int scratch_init(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return TSW_OK;  // such devices run no allocation-free sweep
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return TSW_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return TSW_E_HIP;
    g_scratch_ptr[device] = p;
    return TSW_OK;
}

namespace {

template <typename F>
int guarded(tsw_ctx *ctx, F &&f) {
    if (!ctx) return TSW_E_ARG;
    try {
        ctx->err.clear();
        ctx->xs.key[0] = ctx->xs.key[1] = nullptr;  // cross sections may have changed since the last call
        TSW_HIP(hipSetDevice(ctx->device));
        f();
        return TSW_OK;
    } catch (const Error &e) {
        ctx->err = e.what();
        return e.code;
    } catch (const std::exception &e) {
        ctx->err = e.what();
        return TSW_E_ARG;
    }
}

void require(bool cond, int code, const std::string &msg) {
    if (!cond) throw Error(code, msg);
}

double *dbufd(DBuf &b, size_t n) { return static_cast<double *>(b.reserve(n * sizeof(double))); }

void need_problem(tsw_ctx *c) { require(c->have_problem, TSW_E_STATE, "tsw_problem_set has not been called"); }

// Doubles in one cross-section array: cell means, or (p+1)^2 Legendre coefficients per cell.
size_t xs_n(const tsw_ctx *c) { return c->g.ncell_h() * (c->xs_deg > 0 ? (size_t)c->g.N() : 1); }

// Legendre triple products with max(2p, 2) Gauss points, entries below 5e-16 set to zero
// (ComputeTPI, global.cpp:67-107).
Tpi tpi_table(int p);

// Slot of the tensor cache holding coefficient array sig (device), built on first use in this call.
int xs_slot(tsw_ctx *c, const double *sig) {
    auto &X = c->xs;
    for (int k = 0; k < 2; ++k)
        if (X.key[k] == sig) return k;
    const int k = X.next;
    X.next ^= 1;
    const Geom &g = c->g;
    launch_xs_tensor(g, sig, tpi_table(g.p), dbufd(X.t64[k], g.ncell_h() * g.N() * g.N()), c->stream);
    X.key[k] = sig;
    X.have32[k] = X.have_mean[k] = false;
    return k;
}

// Tensor of degree-p cross section sig in precision T, or null for cell means.
template <typename T>
const T *xs_tensor(tsw_ctx *c, const double *sig) {
    if (c->xs_deg == 0 || !sig) return nullptr;
    const int k = xs_slot(c, sig);
    auto &X = c->xs;
    if constexpr (sizeof(T) == 8) {
        return X.t64[k].as<double>();
    } else {
        const size_t n = c->g.ncell_h() * c->g.N() * c->g.N();
        if (!X.have32[k]) {
            launch_convert_d2f(X.t64[k].as<double>(), static_cast<float *>(X.t32[k].reserve(n * sizeof(float))), n,
                               c->stream);
            X.have32[k] = true;
        }
        return X.t32[k].as<float>();
    }
}

// Cell means of a cross-section array (the array itself for cell means).
const double *xs_means(tsw_ctx *c, const double *sig) {
    if (c->xs_deg == 0 || !sig) return sig;
    const int k = xs_slot(c, sig);
    auto &X = c->xs;
    if (!X.have_mean[k]) {
        launch_xs_means(c->g, sig, dbufd(X.mean[k], c->g.ncell_h()), c->stream);
        X.have_mean[k] = true;
    }
    return X.mean[k].as<double>();
}

// Device view of a caller array: the pointer itself for TSW_MEM_DEVICE, else a staged copy.
template <typename T>
T *stage_in(tsw_ctx *c, const T *p, size_t n, int mem, DBuf &buf) {
    require(p != nullptr, TSW_E_ARG, "null array argument");
    if (mem == TSW_MEM_DEVICE) return const_cast<T *>(p);
    require(mem == TSW_MEM_HOST, TSW_E_ARG, "mem must be TSW_MEM_HOST or TSW_MEM_DEVICE");
    T *d = static_cast<T *>(buf.reserve(n * sizeof(T)));
    TSW_HIP(hipMemcpyAsync(d, p, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return d;
}
template <typename T>
void stage_out(tsw_ctx *c, T *p, const T *d, size_t n, int mem) {
    if (mem == TSW_MEM_DEVICE) {
        if (p != d) TSW_HIP(hipMemcpyAsync(p, d, n * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
        return;
    }
    TSW_HIP(hipMemcpyAsync(p, d, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
}

void refill_exchange(tsw_ctx *c) {
    if (c->xtr64_n) launch_exchange_fill<double>(c->xtr64.as<double>(), c->xtr64_n, c->stream);
    if (c->xtr32_n) launch_exchange_fill<float>(c->xtr32.as<float>(), c->xtr32_n, c->stream);
    auto &M = c->multi;  // the per-source records of lane-split sweeps on the caller's layout
    if (M.xtr64_n) launch_exchange_fill<double>(M.xtr64.as<double>(), M.xtr64_n, c->stream);
    if (M.xtr32_n) launch_exchange_fill<float>(M.xtr32.as<float>(), M.xtr32_n, c->stream);
}

// Waits for the context's stream; reports (and clears) a row-band stall of this context or of its
// adjoint problem.
void sync(tsw_ctx *c) {
    tsw_ctx *chk[2] = {c->bands_used ? c : nullptr, (c->adj && c->adj->bands_used) ? c->adj.get() : nullptr};
    for (tsw_ctx *k : chk)
        if (k) TSW_HIP(hipMemcpyAsync(k->abort_host, k->abort_dev.ptr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    TSW_HIP(hipStreamSynchronize(c->stream));
    bool stalled = false;
    for (tsw_ctx *k : chk) {
        if (k && *k->abort_host) {
            stalled = true;
            *k->abort_host = 0;
            TSW_HIP(hipMemsetAsync(k->abort_dev.ptr, 0, sizeof(int), c->stream));
            refill_exchange(k);
            for (auto &m : k->batch) refill_exchange(m.get());  // batch members' bands report through k
        }
    }
    if (stalled) {
        TSW_HIP(hipStreamSynchronize(c->stream));
        throw Error(TSW_E_HIP, "row-band sweep stalled waiting for an upstream band (spin limit reached)");
    }
}

// Sum of the recorded (start, stop) event pairs; waits for them.
double recorded_ms(tsw_ctx *c) {
    double tot = 0.0;
    for (size_t k = 0; k + 1 < c->ev_used; k += 2) {
        TSW_HIP(hipEventSynchronize(c->ev_pool[k + 1]));
        float ms = 0.f;
        TSW_HIP(hipEventElapsedTime(&ms, c->ev_pool[k], c->ev_pool[k + 1]));
        tot += ms;
    }
    return tot;
}

hipEvent_t next_event(tsw_ctx *c) {
    if (c->ev_used >= 8192) {  // bound the pool: fold finished pairs into the running total
        c->ms_folded += recorded_ms(c);
        c->ev_used = 0;
    }
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        TSW_HIP(hipEventCreate(&e));
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

template <typename T>
SweepParams<T> params(tsw_ctx *c, T *psi, const T *sigma, double dt);
double inverse_dt(double dt) { return (dt > 0.0 && std::isfinite(dt)) ? 1.0 / dt : 0.0; }

template <>
SweepParams<double> params(tsw_ctx *c, double *psi, const double *sigma, double dt) {
    SweepParams<double> sp{psi, sigma, c->xi64.as<double>(), c->eta64.as<double>(), c->g.nx, c->g.ny, c->g.nq,
                           c->g.quad_off[1], c->g.quad_off[2], c->g.quad_off[3], c->g.dx, c->g.dy, inverse_dt(dt)};
    return sp;
}
template <>
SweepParams<float> params(tsw_ctx *c, float *psi, const float *sigma, double dt) {
    SweepParams<float> sp{psi, sigma, c->xi32.as<float>(), c->eta32.as<float>(), c->g.nx, c->g.ny, c->g.nq,
                          c->g.quad_off[1], c->g.quad_off[2], c->g.quad_off[3], (float)c->g.dx, (float)c->g.dy,
                          (float)inverse_dt(dt)};
    return sp;
}

// True when device array a (n doubles) equals the copy held in key, checked on the device (or by an
// earlier precheck_keys of this call); otherwise key receives a copy of a and the result is false.
bool same_as_key(tsw_ctx *c, const double *a, size_t n, DBuf &key, bool keyed) {
    for (size_t k = 0; k < c->key_checks.size(); ++k) {
        const auto kc = c->key_checks[k];
        if (kc.key != &key || kc.a != a || kc.n != n) continue;
        c->key_checks.erase(c->key_checks.begin() + k);
        if (kc.same) return true;
        TSW_HIP(hipMemcpyAsync(key.reserve(n * sizeof(double)), a, n * sizeof(double), hipMemcpyDeviceToDevice,
                               c->stream));
        return false;
    }
    if (keyed && key.bytes >= n * sizeof(double)) {
        int *flag = static_cast<int *>(c->same_flag.reserve(sizeof(int)));
        launch_differs(n, a, key.as<double>(), flag, c->stream);
        int h = 1;
        TSW_HIP(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        TSW_HIP(hipStreamSynchronize(c->stream));
        if (h == 0) return true;
    }
    TSW_HIP(hipMemcpyAsync(key.reserve(n * sizeof(double)), a, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return false;
}

// The checks of same_as_key for several (array, length, key, keyed) at once, with one host round trip;
// the results wait in key_checks for the same_as_key calls of the current solve (KeyCheckScope clears
// what is left when it ends).
struct KeyCheckScope {
    tsw_ctx *c;
    explicit KeyCheckScope(tsw_ctx *c_) : c(c_) { c->key_checks.clear(); }
    ~KeyCheckScope() { c->key_checks.clear(); }
};
void precheck_keys(tsw_ctx *c, const std::vector<std::tuple<const double *, size_t, DBuf *, bool>> &items) {
    std::vector<int> run;
    for (size_t k = 0; k < items.size(); ++k) {
        const auto &[a, n, key, keyed] = items[k];
        if (keyed && key->bytes >= n * sizeof(double)) run.push_back((int)k);
    }
    if (run.empty()) return;
    int *flags = static_cast<int *>(c->same_flag.reserve(sizeof(int) * items.size()));
    for (int k : run) {
        const auto &[a, n, key, keyed] = items[k];
        launch_differs(n, a, key->as<double>(), flags + k, c->stream);
    }
    std::vector<int> h(items.size(), 1);
    TSW_HIP(hipMemcpyAsync(h.data(), flags, sizeof(int) * items.size(), hipMemcpyDeviceToHost, c->stream));
    TSW_HIP(hipStreamSynchronize(c->stream));
    for (int k : run) {
        const auto &[a, n, key, keyed] = items[k];
        c->key_checks.push_back({key, a, n, h[k] == 0});
    }
}

// Builds the material table from device sigma_t (double cell array) for time step dt: reads sigma
// back, numbers its distinct interior values and uploads the per-cell index.  Leaves the table
// invalid (so sweeps eliminate per cell) when there are more than kMaxMaterials values.  When sigma
// equals the array of the last build (compared on the device) the table is kept, and the inverses too
// unless dt changed.
void prepare_materials(tsw_ctx *c, const double *sigma, double dt) {
    auto &m = c->mat;
    if (c->pattern != Pattern::Staggered) {
        m.valid = false;
        m.keyed = false;
        return;
    }
    const Geom &g = c->g;
    const bool same = same_as_key(c, sigma, g.ncell_h(), m.key, m.keyed);
    m.keyed = true;
    if (same) {
        if (m.inv_dt != inverse_dt(dt)) {
            m.inv_dt = inverse_dt(dt);
            m.have64 = m.have32 = false;
        }
        return;
    }
    m.valid = false;
    m.have64 = m.have32 = false;
    ++m.build;
    std::vector<double> h(g.ncell_h());
    TSW_HIP(hipMemcpyAsync(h.data(), sigma, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    TSW_HIP(hipStreamSynchronize(c->stream));
    std::unordered_map<double, int> ids;
    std::vector<double> vals;
    std::vector<int32_t> idx(h.size(), 0);
    for (int i = 1; i <= g.nx; ++i)
        for (int j = 1; j <= g.ny; ++j) {
            const size_t k = (size_t)i * (g.ny + 2) + j;
            auto it = ids.find(h[k]);
            if (it == ids.end()) {
                if ((int)vals.size() == tsw_ctx::kMaxMaterials) return;
                it = ids.emplace(h[k], (int)vals.size()).first;
                vals.push_back(h[k]);
            }
            idx[k] = (int32_t)it->second;
        }
    m.M = (int)vals.size();
    m.inv_dt = inverse_dt(dt);
    TSW_HIP(hipMemcpyAsync(m.idx.reserve(sizeof(int32_t) * idx.size()), idx.data(), sizeof(int32_t) * idx.size(),
                           hipMemcpyHostToDevice, c->stream));
    TSW_HIP(hipMemcpyAsync(m.sig.reserve(sizeof(double) * vals.size()), vals.data(), sizeof(double) * vals.size(),
                           hipMemcpyHostToDevice, c->stream));
    TSW_HIP(hipStreamSynchronize(c->stream));  // host vectors go out of scope
    m.valid = true;
}

// One in-place sweep of device array psi with device cell cross sections sigma (sig64: the caller's FP64
// array, whose tensor replaces sigma for degree-p cross sections).
template <typename T>
void sweep_dev(tsw_ctx *c, T *psi, const T *sigma, double dt, const double *sig64) {
    SweepParams<T> sp = params<T>(c, psi, sigma, dt);
    sp.xs = xs_tensor<T>(c, sig64);
    hipEvent_t e0 = next_event(c), e1 = next_event(c);
    TSW_HIP(hipEventRecord(e0, c->stream));
    switch (c->pattern) {
        case Pattern::WavefrontEager: launch_sweep_wavefront<T>(sp, c->g.p, c->stream); break;
        case Pattern::WavefrontGraph: {
            const auto key = std::make_tuple((int)sizeof(T), (const void *)psi, (const void *)sigma, (double)sp.inv_dt);
            auto it = c->graphs.find(key);
            if (it == c->graphs.end()) {
                hipGraph_t graph;
                TSW_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                try {
                    launch_sweep_wavefront<T>(sp, c->g.p, c->stream);
                } catch (...) {
                    (void)hipStreamEndCapture(c->stream, &graph);
                    throw;
                }
                TSW_HIP(hipStreamEndCapture(c->stream, &graph));
                hipGraphExec_t exec;
                TSW_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
                (void)hipGraphDestroy(graph);
                it = c->graphs.emplace(key, exec).first;
            }
            TSW_HIP(hipGraphLaunch(it->second, c->stream));
            break;
        }
        case Pattern::Staggered:
        case Pattern::StaggeredGE:  // meshes the staggered kernel cannot hold
            launch_sweep_wavefront<T>(sp, c->g.p, c->stream);
            break;
    }
    TSW_HIP(hipEventRecord(e1, c->stream));
    ++c->sweeps_done;
}

// Float copies of a double cell array (cross sections).
const float *as_f32(tsw_ctx *c, const double *d, size_t n, DBuf &buf) {
    float *f = static_cast<float *>(buf.reserve(n * sizeof(float)));
    launch_convert_d2f(d, f, n, c->stream);
    return f;
}

// Sweep a double device array in the requested precision.
void sweep_prec(tsw_ctx *c, double *psi, const double *sigma, double dt, int precision) {
    if (precision == TSW_FP64) {
        sweep_dev<double>(c, psi, sigma, dt, sigma);
        return;
    }
    require(precision == TSW_FP32, TSW_E_ARG, "precision must be TSW_FP64 or TSW_FP32");
    const size_t n = c->g.npsi();
    float *p32 = static_cast<float *>(c->psi32.reserve(n * sizeof(float)));
    launch_convert_d2f(psi, p32, n, c->stream);
    const float *s32 = as_f32(c, sigma, c->g.ncell_h(), c->sig_t32);
    sweep_dev<float>(c, p32, s32, dt, sigma);
    launch_convert_f2d_interior(c->g, (size_t)c->g.nq * c->g.N(), p32, psi, c->stream);  // inflow halos kept
}

double device_norm2_interior_diff(tsw_ctx *c, const double *a, const double *b) {
    const int nb = reduction_blocks();
    double *part = static_cast<double *>(c->red_part.reserve(sizeof(double) * nb * 64));
    double *out = static_cast<double *>(c->red_out.reserve(sizeof(double) * 64));
    launch_sqdist_interior(c->g, a, b, out, part, nb, c->stream);
    double h = 0.0;
    TSW_HIP(hipMemcpyAsync(&h, out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return std::sqrt(h);
}

// ---------------------------------------------------------------------------------------------------
// Restarted GMRES on phi-shaped FP64 vectors (zero halos), CGS2 orthogonalisation with batched dot
// products so each Arnoldi step costs two host round trips.
// ---------------------------------------------------------------------------------------------------
// ---------------------------------------------------------------------------------------------------
// DSA preconditioner (dsa.hip): hierarchy construction on the host from the cross sections, V-cycles,
// and the captured apply.
// ---------------------------------------------------------------------------------------------------
// One V-cycle from level l: right-hand side lv[l].b, start zero (zero_init) or lv[l].y, result in
// lv[l].y.  Fine levels: tile smoothing, the residual restricted, the coarse cycle, the prolongated
// correction smoothed; from fused_from down, one single-workgroup launch.
void dsa_vcycle(tsw_ctx *c, size_t l, bool zero_init) {
    auto &D = c->dsa;
    auto &lv = D.lv;
    const MgLevel &L = lv[l];
    hipStream_t s = c->stream;
    if ((int)l >= D.fused_from) {
        launch_mg_fused(D.lv_dev.as<MgLevel>(), (int)l, (int)lv.size(), D.fused_lds, !zero_init, s);
        return;
    }
    launch_mg_smooth(L, zero_init ? nullptr : L.y, L.x, nullptr, 0, s);  // pre-smoothing
    launch_mg_restrict_residual(L, lv[l + 1], s);
    dsa_vcycle(c, l + 1, true);
    launch_mg_smooth(L, L.x, L.y, &lv[l + 1], 1, s);  // prolongated correction, post-smoothing
}

// Builds the diffusion hierarchy for device cross sections (cell arrays) and time step dt, and captures
// the apply graph out = in + C^{-1} sigma_s in (cell means).
void dsa_build(tsw_ctx *c, const double *sigma_t, const double *sigma_s, double dt) {
    const Geom &g = c->g;
    auto &D = c->dsa;
    const size_t nh = g.ncell_h();
    const double idt = inverse_dt(dt);
    const bool same_st = same_as_key(c, sigma_t, nh, D.key_st, D.keyed);
    const bool same_ss = same_as_key(c, sigma_s, nh, D.key_ss, D.keyed);
    D.keyed = true;
    // the apply reads sigma_s from its pointer: reuse the hierarchy only for the same pointer and data
    if (D.graph && D.sigma_s == sigma_s && D.key_idt == idt && D.key_cycles == c->dsa_cycles && same_st && same_ss)
        return;
    std::vector<double> st(nh), ss(nh);
    TSW_HIP(hipMemcpyAsync(st.data(), sigma_t, sizeof(double) * nh, hipMemcpyDeviceToHost, c->stream));
    TSW_HIP(hipMemcpyAsync(ss.data(), sigma_s, sizeof(double) * nh, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    if (D.graph) {
        (void)hipGraphExecDestroy(D.graph);
        D.graph = nullptr;
    }
    struct HostLevel {
        int nx, ny;
        std::vector<double> cx, cy, sv;
    };
    std::vector<HostLevel> hl(1);
    {
        HostLevel &L = hl[0];
        L.nx = g.nx;
        L.ny = g.ny;
        L.cx.assign((size_t)(g.nx + 1) * g.ny, 0.0);
        L.cy.assign((size_t)g.nx * (g.ny + 1), 0.0);
        L.sv.assign((size_t)g.nx * g.ny, 0.0);
        auto Dc = [&](int i, int j) {  // diffusion coefficient of interior cell (i,j), 0-based
            const double t = st[(size_t)(i + 1) * (g.ny + 2) + j + 1] + idt;
            return 1.0 / (3.0 * std::max(t, 1e-12));
        };
        for (int j = 0; j < g.ny; ++j)
            for (int i = 0; i <= g.nx; ++i) {
                double rsum;
                if (i == 0) rsum = g.dx / (2 * Dc(0, j)) + 2.0;
                else if (i == g.nx) rsum = g.dx / (2 * Dc(g.nx - 1, j)) + 2.0;
                else rsum = g.dx / (2 * Dc(i - 1, j)) + g.dx / (2 * Dc(i, j));
                L.cx[(size_t)i * g.ny + j] = g.dy / rsum;
            }
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j <= g.ny; ++j) {
                double rsum;
                if (j == 0) rsum = g.dy / (2 * Dc(i, 0)) + 2.0;
                else if (j == g.ny) rsum = g.dy / (2 * Dc(i, g.ny - 1)) + 2.0;
                else rsum = g.dy / (2 * Dc(i, j - 1)) + g.dy / (2 * Dc(i, j));
                L.cy[(size_t)i * (g.ny + 1) + j] = g.dx / rsum;
            }
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j) {
                const size_t h = (size_t)(i + 1) * (g.ny + 2) + j + 1;
                L.sv[(size_t)i * g.ny + j] = std::max(st[h] + idt - ss[h], 0.0) * g.dx * g.dy;
            }
    }
    // 2x2 aggregation: coarse conductances sum the fine faces they contain, absorption sums the cells
    while ((hl.back().nx > 3 || hl.back().ny > 3) && hl.size() < 16) {
        const HostLevel &f = hl.back();
        HostLevel L;
        L.nx = (f.nx + 1) / 2;
        L.ny = (f.ny + 1) / 2;
        L.cx.assign((size_t)(L.nx + 1) * L.ny, 0.0);
        L.cy.assign((size_t)L.nx * (L.ny + 1), 0.0);
        L.sv.assign((size_t)L.nx * L.ny, 0.0);
        for (int I = 0; I <= L.nx; ++I) {
            const int fi = I == L.nx ? f.nx : 2 * I;
            for (int j = 0; j < f.ny; ++j) L.cx[(size_t)I * L.ny + j / 2] += f.cx[(size_t)fi * f.ny + j];
        }
        for (int J = 0; J <= L.ny; ++J) {
            const int fj = J == L.ny ? f.ny : 2 * J;
            for (int i = 0; i < f.nx; ++i) L.cy[(size_t)(i / 2) * (L.ny + 1) + J] += f.cy[(size_t)i * (f.ny + 1) + fj];
        }
        for (int i = 0; i < f.nx; ++i)
            for (int j = 0; j < f.ny; ++j) L.sv[(size_t)(i / 2) * L.ny + j / 2] += f.sv[(size_t)i * f.ny + j];
        hl.push_back(std::move(L));
    }
    // the coarsest level (at most 3 x 3 cells) is solved exactly with its dense inverse, the same
    // operator the smoother relaxes: diagonal sv + the four face conductances, -conductance to each
    // interior neighbour
    std::vector<double> cinv;
    {
        const HostLevel &h = hl.back();
        const int n = h.nx * h.ny;
        std::vector<double> A((size_t)n * n, 0.0);
        for (int i = 0; i < h.nx; ++i)
            for (int j = 0; j < h.ny; ++j) {
                const int c = i * h.ny + j;
                const double gw = h.cx[(size_t)i * h.ny + j], ge = h.cx[(size_t)(i + 1) * h.ny + j];
                const double gs = h.cy[(size_t)i * (h.ny + 1) + j], gn = h.cy[(size_t)i * (h.ny + 1) + j + 1];
                A[(size_t)c * n + c] = h.sv[c] + gw + ge + gs + gn;
                if (i > 0) A[(size_t)c * n + c - h.ny] = -gw;
                if (i < h.nx - 1) A[(size_t)c * n + c + h.ny] = -ge;
                if (j > 0) A[(size_t)c * n + c - 1] = -gs;
                if (j < h.ny - 1) A[(size_t)c * n + c + 1] = -gn;
            }
        // Gauss-Jordan with partial pivoting on [A | I]
        std::vector<double> I((size_t)n * n, 0.0);
        for (int k = 0; k < n; ++k) I[(size_t)k * n + k] = 1.0;
        bool ok = true;
        for (int k = 0; k < n && ok; ++k) {
            int piv = k;
            for (int r = k + 1; r < n; ++r)
                if (std::fabs(A[(size_t)r * n + k]) > std::fabs(A[(size_t)piv * n + k])) piv = r;
            if (!(std::fabs(A[(size_t)piv * n + k]) > 0.0)) {
                ok = false;
                break;
            }
            for (int col = 0; col < n; ++col) {
                std::swap(A[(size_t)k * n + col], A[(size_t)piv * n + col]);
                std::swap(I[(size_t)k * n + col], I[(size_t)piv * n + col]);
            }
            const double d = A[(size_t)k * n + k];
            for (int col = 0; col < n; ++col) {
                A[(size_t)k * n + col] /= d;
                I[(size_t)k * n + col] /= d;
            }
            for (int r = 0; r < n; ++r) {
                if (r == k) continue;
                const double f = A[(size_t)r * n + k];
                if (f == 0.0) continue;
                for (int col = 0; col < n; ++col) {
                    A[(size_t)r * n + col] -= f * A[(size_t)k * n + col];
                    I[(size_t)r * n + col] -= f * I[(size_t)k * n + col];
                }
            }
        }
        if (ok) cinv = std::move(I);
    }
    D.lv.assign(hl.size(), MgLevel{});
    D.store.clear();
    D.store.resize(hl.size() * 6);
    for (size_t l = 0; l < hl.size(); ++l) {
        const HostLevel &h = hl[l];
        MgLevel &L = D.lv[l];
        L.nx = h.nx;
        L.ny = h.ny;
        auto up = [&](DBuf &b, const std::vector<double> &v) {
            double *d = static_cast<double *>(b.reserve(sizeof(double) * v.size()));
            TSW_HIP(hipMemcpyAsync(d, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice, c->stream));
            return d;
        };
        L.cx = up(D.store[6 * l + 0], h.cx);
        L.cy = up(D.store[6 * l + 1], h.cy);
        L.sv = up(D.store[6 * l + 2], h.sv);
        const size_t n = (size_t)h.nx * h.ny;
        L.x = static_cast<double *>(D.store[6 * l + 3].reserve(sizeof(double) * n));
        L.b = static_cast<double *>(D.store[6 * l + 4].reserve(sizeof(double) * n));
        L.y = static_cast<double *>(D.store[6 * l + 5].reserve(sizeof(double) * n));
        L.inv = nullptr;
        if (l + 1 == hl.size() && !cinv.empty()) L.inv = up(D.coarse_inv, cinv);
    }
    // the small levels (at most 4096 cells each, all of their arrays within one workgroup's LDS) run
    // inside one workgroup; the coarsest level always does
    const int nl = (int)hl.size();
    require(nl <= kMgMaxLevels, TSW_E_UNSUPPORTED, "diffusion hierarchy too deep");
    const size_t lds_cap = max_lds_per_block() - 4096;
    D.fused_from = nl - 1;
    while (D.fused_from > 0 && (size_t)hl[D.fused_from - 1].nx * hl[D.fused_from - 1].ny <= 4096 &&
           mg_fused_lds_doubles(D.lv.data(), D.fused_from - 1, nl) * sizeof(double) <= lds_cap)
        --D.fused_from;
    D.fused_lds = mg_fused_lds_doubles(D.lv.data(), D.fused_from, nl) * sizeof(double);
    require(D.fused_lds <= lds_cap, TSW_E_UNSUPPORTED, "coarsest diffusion level does not fit in LDS");
    TSW_HIP(hipMemcpyAsync(D.lv_dev.reserve(sizeof(MgLevel) * D.lv.size()), D.lv.data(), sizeof(MgLevel) * D.lv.size(),
                           hipMemcpyHostToDevice, c->stream));
    sync(c);  // host vectors go out of scope
    D.sigma_s = sigma_s;
    D.key_idt = idt;
    D.key_cycles = c->dsa_cycles;
    // the graph holds the V-cycles only (level-0 right-hand side in, level-0 solution out); the
    // right-hand side and the correction are launched per apply on the caller's vectors
    hipGraph_t graph;
    TSW_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    try {
        for (int k = 0; k < c->dsa_cycles; ++k) dsa_vcycle(c, 0, k == 0);
    } catch (...) {
        (void)hipStreamEndCapture(c->stream, &graph);
        throw;
    }
    TSW_HIP(hipStreamEndCapture(c->stream, &graph));
    const hipError_t ie = hipGraphInstantiate(&D.graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    TSW_HIP(ie);
}

// y = x + C^{-1} sigma_s x on phi-layout vectors (y may equal x)
void dsa_apply(tsw_ctx *c, const double *x, double *y) {
    auto &D = c->dsa;
    const Geom &g = c->g;
    launch_dsa_rhs(g, x, D.sigma_s, D.lv[0].b, c->stream);
    TSW_HIP(hipGraphLaunch(D.graph, c->stream));
    if (y != x) TSW_HIP(hipMemcpyAsync(y, x, sizeof(double) * g.nphi(), hipMemcpyDeviceToDevice, c->stream));
    launch_dsa_add(g, D.lv[0].y, y, c->stream);
}

// Relaxed mixed-precision GMRES: the relative error d of an FP32-sweep product is measured once per
// (material table, dt, DSA setting) as ||A32 b - A64 b|| / ||b||, and FP32 products are admitted once the
// residual estimate is <= abs_tol / (kFp32ErrSafety d).  d is kept within [kFp32ErrFloor, kFp32ErrCeil]: the
// floor keeps an exactly representable problem (d = 0) from admitting FP32 products at the first step,
// the ceiling keeps a badly conditioned one from never admitting them (the final FP64 residual check
// guards the result either way).
constexpr double kFp32ErrSafety = 10.0;  // default of the tuning key fp32_safety
constexpr double kFp32ErrFloor = 1e-7;
constexpr double kFp32ErrCeil = 1e-4;

// The measured-error entry of (material table build, dt, DSA setting), created unmeasured (value < 0) when
// absent; the oldest entry goes when kFp32ErrKeep are held.
tsw_ctx::Fp32Err &fp32_err_entry(tsw_ctx *c, double dt, bool dsa) {
    const double idt = inverse_dt(dt);
    for (auto &e : c->fp32_err)
        if (e.mat_build == c->mat.build && e.inv_dt == idt && e.dsa == dsa) return e;
    if (c->fp32_err.size() >= tsw_ctx::kFp32ErrKeep) c->fp32_err.erase(c->fp32_err.begin());
    c->fp32_err.push_back({-1.0, c->mat.build, idt, dsa});
    return c->fp32_err.back();
}

struct Gmres {
    tsw_ctx *c;
    size_t n;
    int m;
    double *V;     // (m+1) x n, halos kept zero
    double *hdev;  // [2][m+2]: the two Gram-Schmidt passes' coefficients; slot j+1 of each holds
                   // ||w|| and ||w||^2 of step j
    double *hh;    // pinned host copy of hdev
    double *part;
    int nb;

    double *ydev;  // the update coefficients, negated, for one multi_axpy
    // the workspace is the context's (kept across solves, so a solve allocates nothing once warm)
    Gmres(tsw_ctx *ctx, size_t n_, int m_) : c(ctx), n(n_), m(m_) {
        V = static_cast<double *>(c->gm_v.reserve(sizeof(double) * n * (m + 1)));
        TSW_HIP(hipMemsetAsync(V, 0, sizeof(double) * n * (m + 1), c->stream));
        hdev = static_cast<double *>(c->gm_h.reserve(sizeof(double) * 2 * (m + 2)));
        nb = reduction_blocks();
        part = static_cast<double *>(c->gm_part.reserve(sizeof(double) * nb * (m + 2)));
        ydev = static_cast<double *>(c->gm_y.reserve(sizeof(double) * m));
        const size_t need = 3 * (size_t)(m + 2);
        if (c->gm_host_n < need) {
            if (c->gm_host) TSW_HIP(hipHostFree(c->gm_host));
            c->gm_host = nullptr;
            c->gm_host_n = 0;
            TSW_HIP(hipHostMalloc((void **)&c->gm_host, sizeof(double) * need, hipHostMallocDefault));
            c->gm_host_n = need;
        }
        hh = c->gm_host;
    }
    double norm(const double *y) {
        launch_multi_dot(n, 1, y, n, y, hdev, part, nb, c->stream);
        TSW_HIP(hipMemcpyAsync(hh, hdev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        sync(c);
        return std::sqrt(hh[0]);
    }

    // Solves A x = b (x holds the initial guess, updated in place). apply(x_dev, y_dev) computes y = A x
    // and leaves y's halo entries as they are (zero).  Arnoldi with CGS2: the projections, the
    // re-orthogonalisation, the norm and the normalisation stay on the device, and the Hessenberg column
    // comes back in one copy, so a step costs one host round trip.  Stops when the residual estimate
    // <= abs_tol.  Returns true when converged.
    template <typename Apply>
    bool solve(Apply &&apply, const double *b, double *x, double *r, double abs_tol, int max_its, int &its,
               double &res, bool x_zero = false) {
        return solve(apply, apply, -1.0, b, x, r, abs_tol, max_its, its, res, x_zero);
    }

    // Inexact (relaxed) form: apply_lo, a cheaper and less accurate product, replaces apply once the
    // residual estimate is <= lo_below.  A product error below ||b|| abs_tol / ||r_k|| at step k keeps the
    // attainable residual at abs_tol (the relaxation strategy of inexact Krylov methods: Bouras and
    // Fraysse 2005, Simoncini and Szyld 2003), so a fixed relative error d of apply_lo is admissible from
    // ||r_k|| <= abs_tol / d on.
    // x_zero: x is zero on entry, so the first residual is b itself and needs no product.
    // used_lo: set when a call took any apply_lo product (its residual estimate then rests on them).
    bool used_lo = false;
    template <typename Apply, typename ApplyLo>
    bool solve(Apply &&apply, ApplyLo &&apply_lo, double lo_below, const double *b, double *x, double *r,
               double abs_tol, int max_its, int &its, double &res, bool x_zero = false) {
        std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
        double *h1 = hdev, *h2 = hdev + (m + 2);
        its = 0;
        res = 0.0;
        used_lo = false;
        bool zero = x_zero;
        while (true) {
            if (zero) {
                launch_axpby(n, 1.0, b, 0.0, r, c->stream);  // r = b
                zero = false;
            } else {
                apply(x, r);  // restart residuals with the accurate product
                launch_axpby(n, 1.0, b, -1.0, r, c->stream);  // r = b - A x
            }
            const double beta = norm(r);
            res = beta;
            if (beta <= abs_tol) return true;
            if (its >= max_its) return false;
            launch_axpby(n, 1.0 / beta, r, 0.0, V, c->stream);
            std::fill(g.begin(), g.end(), 0.0);
            g[0] = beta;
            int jdone = 0;
            bool conv = false;
            for (int j = 0; j < m && its < max_its; ++j) {
                double *vn = V + (size_t)(j + 1) * n;
                if (res <= lo_below) {
                    apply_lo(V + (size_t)j * n, vn);
                    used_lo = true;
                } else {
                    apply(V + (size_t)j * n, vn);
                }
                ++its;
                launch_multi_dot(n, j + 1, V, n, vn, h1, part, nb, c->stream);
                launch_multi_axpy(n, j + 1, V, n, h1, vn, c->stream);
                launch_multi_dot(n, j + 1, V, n, vn, h2, part, nb, c->stream);
                launch_multi_axpy(n, j + 1, V, n, h2, vn, c->stream);
                launch_multi_dot(n, 1, vn, n, vn, h2 + j + 1, part, nb, c->stream);
                launch_normalize(n, h2 + j + 1, vn, h1 + j + 1, c->stream);
                TSW_HIP(hipMemcpyAsync(hh, hdev, sizeof(double) * 2 * (m + 2), hipMemcpyDeviceToHost, c->stream));
                sync(c);
                const double *hh1 = hh, *hh2 = hh + (m + 2);
                for (int i = 0; i <= j; ++i) H[i + (m + 1) * j] = hh1[i] + hh2[i];
                H[j + 1 + (m + 1) * j] = hh1[j + 1];
                for (int i = 0; i < j; ++i) {
                    const double h0 = H[i + (m + 1) * j], hx = H[i + 1 + (m + 1) * j];
                    H[i + (m + 1) * j] = cs[i] * h0 + sn[i] * hx;
                    H[i + 1 + (m + 1) * j] = -sn[i] * h0 + cs[i] * hx;
                }
                const double h0 = H[j + (m + 1) * j], hx = H[j + 1 + (m + 1) * j];
                const double den = std::sqrt(h0 * h0 + hx * hx);
                cs[j] = h0 / den;
                sn[j] = hx / den;
                H[j + (m + 1) * j] = den;
                H[j + 1 + (m + 1) * j] = 0.0;
                g[j + 1] = -sn[j] * g[j];
                g[j] = cs[j] * g[j];
                jdone = j + 1;
                res = std::fabs(g[j + 1]);
                if (res <= abs_tol) {
                    conv = true;
                    break;
                }
            }
            for (int i = jdone - 1; i >= 0; --i) {
                double s = g[i];
                for (int k = i + 1; k < jdone; ++k) s -= H[i + (m + 1) * k] * y[k];
                y[i] = s / H[i + (m + 1) * i];
            }
            if (jdone > 0) {  // x += V y in one pass (multi_axpy subtracts: it takes -y)
                double *yh = hh + 2 * (m + 2);
                for (int i = 0; i < jdone; ++i) yh[i] = -y[i];
                TSW_HIP(hipMemcpyAsync(ydev, yh, sizeof(double) * jdone, hipMemcpyHostToDevice, c->stream));
                launch_multi_axpy(n, jdone, V, n, ydev, x, c->stream);
            }
            if (conv) return true;
            if (its >= max_its) {
                sync(c);
                return false;
            }
        }
    }
};


// ---------------------------------------------------------------------------------------------------
// Staggered sweep engine: every solver of the "hip-staggered" patterns runs on the diagonal-major
// layout.  Reference-layout arrays are converted once on entry and once on exit of a call.
// ---------------------------------------------------------------------------------------------------
template <typename T>
T *dbuf(DBuf &b, size_t n) {
    return static_cast<T *>(b.reserve(n * sizeof(T)));
}

void diag_load_source(tsw_ctx *c, const double *qref, bool need64, bool need32);

// Builds the per-call inputs: materials or per-cell sigma_t, sigma_s, inflow traces and the source
// from reference-layout arrays (sigma_s, qref may be null).  need64/need32 select the precisions; src32
// false leaves out the FP32 source and inflow traces (callers whose FP32 sweeps only see residuals).
void diag_setup(tsw_ctx *c, const double *sigma_t, const double *sigma_s, const double *qref, double dt, bool need64,
                bool need32, bool src32 = true) {
    const Geom &g = c->g;
    auto &d = c->diag;
    const size_t ncell = (size_t)g.nx * g.ny;
    hipStream_t s = c->stream;
    prepare_materials(c, sigma_t, dt);
    if (c->mat.valid)
        launch_cell_to_diag<int32_t, int32_t>(g, c->mat.idx.as<int32_t>(), dbuf<int32_t>(d.mat, 4 * ncell), 1, s);
    if (need64) launch_cell_to_diag<double, double>(g, sigma_t, dbuf<double>(d.sigt64, 4 * ncell), 1, s);
    if (need32) launch_cell_to_diag<double, float>(g, sigma_t, dbuf<float>(d.sigt32, 4 * ncell), 1, s);
    if (sigma_s) {
        if (need64) launch_cell_to_diag<double, double>(g, sigma_s, dbuf<double>(d.sigs64, 4 * ncell), 1, s);
        if (need32) launch_cell_to_diag<double, float>(g, sigma_s, dbuf<float>(d.sigs32, 4 * ncell), 1, s);
    }
    diag_load_source(c, qref, need64, need32 && src32);
}

// Inflow traces and the source in the diagonal layout, from a reference-layout psi (null: no source).
void diag_load_source(tsw_ctx *c, const double *qref, bool need64, bool need32) {
    const Geom &g = c->g;
    auto &d = c->diag;
    const size_t ncell = (size_t)g.nx * g.ny, N = g.N();
    hipStream_t s = c->stream;
    d.inflow = qref != nullptr;
    if (qref) {
        const size_t nt_x = (size_t)g.nq * g.ny * (g.p + 1), nt_y = (size_t)g.nq * g.nx * (g.p + 1);
        if (need64) launch_inflow<double>(g, qref, dbuf<double>(d.tinx64, nt_x), dbuf<double>(d.tiny64, nt_y), s);
        if (need32) launch_inflow<float>(g, qref, dbuf<float>(d.tinx32, nt_x), dbuf<float>(d.tiny32, nt_y), s);
        if (need64 && need32)  // both precisions from one read of the source
            launch_psi_load_dual(g, qref, dbuf<double>(d.q64, g.nq * N * ncell), dbuf<float>(d.q32, g.nq * N * ncell), s);
        else if (need64)
            launch_psi_convert<double, double>(g, const_cast<double *>(qref), dbuf<double>(d.q64, g.nq * N * ncell), 0, s);
        else if (need32)
            launch_psi_convert<double, float>(g, const_cast<double *>(qref), dbuf<float>(d.q32, g.nq * N * ncell), 0, s);
    }
}

// Element-inverse table for precision T (null when the per-cell elimination is used).
template <typename T>
const T *diag_inverses(tsw_ctx *c, double dt) {
    auto &m = c->mat;
    if (c->pattern != Pattern::Staggered || !m.valid || m.inv_dt != inverse_dt(dt)) return nullptr;
    const size_t n = (size_t)m.M * c->g.nq * c->g.N() * c->g.N();
    bool &have = sizeof(T) == 8 ? m.have64 : m.have32;
    DBuf &buf = sizeof(T) == 8 ? m.inv64 : m.inv32;
    if (!have) {
        launch_material_inverses<T>(c->g, c->xi64.as<double>(), c->eta64.as<double>(), m.sig.as<double>(), m.M,
                                    m.inv_dt, static_cast<T *>(buf.reserve(n * sizeof(T))), c->stream);
        have = true;
    }
    return buf.as<T>();
}

// Rows per band for this context's mesh: 0 when each ordinate is one workgroup.  Automatic choice:
// one wavefront of rows per band when the mesh has more rows than that, so the ordinates' sweeps
// spread over several times more compute units than there are ordinates.
bool one_wave_for(const tsw_ctx *c) { return (c->tuned_w1 != INT_MIN ? c->tuned_w1 : c->one_wave_opt) == 1; }
bool lane_split_for(const tsw_ctx *c) { return (c->tuned_sl != INT_MIN ? c->tuned_sl : c->split_opt) == 1; }

int band_rows_for(const tsw_ctx *c) {
    const int ny = c->g.ny;
    if (one_wave_for(c)) return ny > 64 ? 64 : 0;  // the one-wave kernel takes bands of one wavefront
    int br = c->tuned_br != INT_MIN ? c->tuned_br : c->band_rows_opt;
    if (br < 0) br = ny > 448 ? 448 : 0;  // a single band cannot hold more than 448 rows
    else if (br == 0) br = ny > 64 ? 64 : 0;
    if (br >= ny) br = 0;
    return br;
}

// The context's device stall flag (allocated and cleared on first use); sync(c) reports a stall.
int *abort_flag_of(tsw_ctx *c) {
    if (!c->abort_host) {
        TSW_HIP(hipHostMalloc((void **)&c->abort_host, sizeof(int), hipHostMallocDefault));
        *c->abort_host = 0;
        c->abort_dev.reserve(sizeof(int));
        TSW_HIP(hipMemsetAsync(c->abort_dev.ptr, 0, sizeof(int), c->stream));
    }
    c->bands_used = true;
    return c->abort_dev.as<int>();
}

// Fills the band fields of sp and makes sure the exchange records exist (filled with the sentinel).
template <typename T>
void configure_bands(tsw_ctx *c, DiagSweepParams<T> &sp) {
    const Geom &g = c->g;
    const int br = band_rows_for(c);
    if (br == 0) {
        sp.nbands = 1;
        sp.band_rows = g.ny;
        return;
    }
    sp.band_rows = br;
    sp.nbands = (g.ny + br - 1) / br;
    const size_t n = (size_t)g.nq * (sp.nbands - 1) * g.nx * (g.p + 1);
    DBuf &buf = sizeof(T) == 8 ? c->xtr64 : c->xtr32;
    size_t &have = sizeof(T) == 8 ? c->xtr64_n : c->xtr32_n;
    if (have < n) {
        buf.reserve(n * sizeof(T));
        have = n;
        launch_exchange_fill<T>(buf.as<T>(), n, c->stream);
    }
    sp.xtr = buf.as<T>();
    sp.abort_flag = abort_flag_of(c);
}

// Kernel parameters of one staggered sweep: psi_out <- L^{-1}(qsrc + sigma_s phi_d / (4 pi)), inflow from
// the traces set up by diag_setup when `inflow`, else vacuum.  qsrc or phi_d may be null (not both);
// qsrc may equal psi_out.
template <typename T>
DiagSweepParams<T> make_diag_params(tsw_ctx *c, double dt, const T *qsrc, const T *phi_d, bool inflow, T *psi_out) {
    const Geom &g = c->g;
    auto &d = c->diag;
    const bool f64 = sizeof(T) == 8;
    DiagSweepParams<T> sp{};
    sp.xi = f64 ? (const T *)c->xi64.ptr : (const T *)c->xi32.ptr;
    sp.eta = f64 ? (const T *)c->eta64.ptr : (const T *)c->eta32.ptr;
    sp.nx = g.nx; sp.ny = g.ny; sp.nq = g.nq;
    sp.qo1 = g.quad_off[1]; sp.qo2 = g.quad_off[2]; sp.qo3 = g.quad_off[3];
    sp.dx = T(g.dx); sp.dy = T(g.dy); sp.inv_dt = T(inverse_dt(dt));
    sp.qsrc = qsrc;
    sp.phi = phi_d;
    sp.sig_s = phi_d ? (const T *)(f64 ? d.sigs64.ptr : d.sigs32.ptr) : nullptr;
    if (inflow) {
        require(d.inflow, TSW_E_STATE, "staggered sweep asked for inflow traces that were not set up");
        sp.tin_x = (const T *)(f64 ? d.tinx64.ptr : d.tinx32.ptr);
        sp.tin_y = (const T *)(f64 ? d.tiny64.ptr : d.tiny32.ptr);
    }
    configure_bands<T>(c, sp);
    sp.prefetch = c->tuned_pf != INT_MIN ? c->tuned_pf : c->prefetch_opt;
    sp.wave1 = one_wave_for(c) && (sp.nbands > 1 ? sp.band_rows : sp.ny) <= 64 ? 1 : 0;
    c->last_br = sp.nbands > 1 ? sp.band_rows : 0;
    c->last_pf = sp.prefetch > 0 ? sp.prefetch : (sizeof(T) == 8 ? 2 : 3);
    sp.ainv = diag_inverses<T>(c, dt);
    sp.direct = 0;
    sp.split = 0;
    if (sp.ainv) {
        sp.mat = d.mat.as<int32_t>();
        sp.M = c->mat.M;
        if (lane_split_for(c)) {  // kept where the lane-split kernel is built and its tables fit
            sp.split = 1;
            sp.wave1 = 0;
            if (diag_sweep_lds_bytes<T>(sp, g.p) == 0) sp.split = 0;
        }
        if (!sp.split && !sp.wave1 && diag_sweep_lds_bytes<T>(sp, g.p) == 0) {  // without the store staging buffer (one band)
            sp.direct = 1;
            if (diag_sweep_lds_bytes<T>(sp, g.p) == 0) {  // table too large: eliminate per cell
                sp.direct = 0;
                sp.ainv = nullptr;
            }
        }
    }
    if (!sp.ainv) sp.sig_t = (const T *)(f64 ? d.sigt64.ptr : d.sigt32.ptr);
    c->last_w1 = sp.wave1;
    c->last_sl = sp.split;
    sp.psi = psi_out;
    return sp;
}

// Autotuning of the staggered sweep: the candidates are one band per ordinate and bands of 64..256 rows
// (below ny), crossed with prefetch depths 2, 3, 4, 6; each runs twice into a scratch output (inputs
// untouched) and the second run is timed with events on the context's stream.  Every candidate gives
// the same result, so the choice only affects speed.
template <typename T>
void select_tuning(tsw_ctx *c, double dt, const T *qsrc, const T *phi_d, bool inflow) {
    c->tuned_br = c->tuned_pf = c->tuned_w1 = c->tuned_sl = INT_MIN;
    if (!c->autotune || c->band_rows_opt != 0 || c->prefetch_opt != 0 || c->one_wave_opt != 0 || c->split_opt != 0)
        return;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    TSW_HIP(hipStreamIsCapturing(c->stream, &cap));
    const Geom &g = c->g;
    const int mode = (qsrc ? 1 : 0) | (phi_d ? 2 : 0);
    const std::array<int, 7> key{g.nx, g.ny, g.p, g.nq, (int)sizeof(T), mode, c->mat.valid ? c->mat.M : 0};
    auto it = c->tuned.find(key);
    if (it == c->tuned.end()) {
        if (cap != hipStreamCaptureStatusNone) return;  // no timing inside a capture: defaults
        // (band rows, one wave, lane split): one-lane kernels over one band and bands of 64..256 rows,
        // the one-wave kernel over 64-row bands, the lane-split kernel over one band (meshes of at most
        // 128 rows) and bands of 32..128 rows
        std::vector<std::array<int, 3>> shapes{{-1, 0, 0}};
        for (int br = 64; br <= 256 && br < g.ny; br += 64) shapes.push_back({br, 0, 0});
        if (g.ny > 64) shapes.push_back({64, 1, 0});
        else shapes.push_back({-1, 1, 0});
        if (c->mat.valid && g.p >= 1) {
            if (g.ny <= 128) shapes.push_back({-1, 0, 1});
            for (int br : {32, 64, 96, 128})
                if (br < g.ny) shapes.push_back({br, 0, 1});
        }
        const int pfs[] = {2, 3, 4, 6};
        T *out = dbuf<T>(c->tune_out, (size_t)g.nq * g.N() * g.nx * g.ny);
        hipEvent_t e0, e1;
        TSW_HIP(hipEventCreate(&e0));
        TSW_HIP(hipEventCreate(&e1));
        float best = 1e30f;
        std::array<int, 4> choice{0, 0, 0, 0};
        try {
            for (const auto &sh : shapes)
                for (int pf : pfs) {
                    const int br = sh[0], w1 = sh[1], sl = sh[2];
                    c->tuned_br = br;
                    c->tuned_pf = pf;
                    c->tuned_w1 = w1;
                    c->tuned_sl = sl;
                    const DiagSweepParams<T> sp = make_diag_params<T>(c, dt, qsrc, phi_d, inflow, out);
                    if (sp.split != sl) continue;  // not built for this problem: the one-lane shapes cover it
                    launch_diag_sweep<T>(sp, g.p, c->stream);
                    TSW_HIP(hipEventRecord(e0, c->stream));
                    launch_diag_sweep<T>(sp, g.p, c->stream);
                    TSW_HIP(hipEventRecord(e1, c->stream));
                    TSW_HIP(hipEventSynchronize(e1));
                    float ms = 0.f;
                    TSW_HIP(hipEventElapsedTime(&ms, e0, e1));
                    if (ms < best) {
                        best = ms;
                        choice = {br, pf, w1, sl};
                    }
                }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            c->tuned_br = c->tuned_pf = c->tuned_w1 = c->tuned_sl = INT_MIN;
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        it = c->tuned.emplace(key, choice).first;
    }
    c->tuned_br = it->second[0];
    c->tuned_pf = it->second[1];
    c->tuned_w1 = it->second[2];
    c->tuned_sl = it->second[3];
}

// One staggered sweep on the context's stream, timed.
template <typename T>
void diag_sweep(tsw_ctx *c, double dt, const T *qsrc, const T *phi_d, bool inflow, T *psi_out) {
    select_tuning<T>(c, dt, qsrc, phi_d, inflow);
    DiagSweepParams<T> sp = make_diag_params<T>(c, dt, qsrc, phi_d, inflow, psi_out);
    sp.skip = c->skip_flag;
    if (c->stall_test && sp.nbands > 1 && !c->skip_flag) {
        c->stall_test = false;
        *c->abort_host = 1;
        TSW_HIP(hipMemcpyAsync(c->abort_dev.ptr, c->abort_host, sizeof(int), hipMemcpyHostToDevice, c->stream));
    }
    if (c->skip_flag) {  // inside a graph capture: the replay is timed as a whole
        launch_diag_sweep<T>(sp, c->g.p, c->stream);
        return;
    }
    hipEvent_t e0 = next_event(c), e1 = next_event(c);
    TSW_HIP(hipEventRecord(e0, c->stream));
    launch_diag_sweep<T>(sp, c->g.p, c->stream);
    TSW_HIP(hipEventRecord(e1, c->stream));
    ++c->sweeps_done;
}

// phi_d[Q] (precision T) from a phi-layout FP64 array
template <typename T>
T *diag_phi(tsw_ctx *c, const double *phi) {
    const Geom &g = c->g;
    auto &d = c->diag;
    const size_t n = 4 * (size_t)g.nx * g.ny * g.N();
    T *out = dbuf<T>(sizeof(T) == 8 ? d.phid64 : d.phid32, n);
    launch_cell_to_diag<double, T>(g, phi, out, g.N(), c->stream, c->skip_flag);
    return out;
}

// phi = beta xb + alpha P psi_d on the interior (xb null: xb = phi)
template <typename T>
void diag_pmv(tsw_ctx *c, const T *psi_d, double *phi, double alpha, double beta, const double *xb = nullptr) {
    const Geom &g = c->g;
    double *phiQ = dbuf<double>(c->diag.phiQ, 4 * (size_t)g.nx * g.ny * g.N());
    launch_pmv_diag<T>(g, psi_d, c->w64.as<double>(), phiQ, phi, alpha, beta, c->stream, c->skip_flag, xb);
}

// out (reference layout, interior) <- psi_d; halos of out untouched
template <typename T>
void diag_store(tsw_ctx *c, const T *psi_d, double *out_ref) {
    launch_psi_convert<double, T>(c->g, out_ref, const_cast<T *>(psi_d), 1, c->stream);
}

// True when the pattern is a staggered one and the mesh fits its kernel (else the wavefront kernels run).
bool staggered_ok(tsw_ctx *c) {
    if (!is_staggered(c->pattern) || c->xs_deg > 0) return false;  // degree-p cross sections: wavefront kernels
    DiagSweepParams<double> probe{};
    probe.nx = c->g.nx;
    probe.ny = c->g.ny;
    const int br = band_rows_for(c);
    if (br > 0) {
        probe.band_rows = br;
        probe.nbands = (c->g.ny + br - 1) / br;
    }
    return diag_sweep_lds_bytes<double>(probe, c->g.p) != 0;
}

// ---------------------------------------------------------------------------------------------------
// Ordinate-lane engine: k_lanes_sweep on the caller's reference-layout arrays (no layout conversion).
// ---------------------------------------------------------------------------------------------------
// Most ordinates of one quadrant.
int lanes_quadrant_max(const tsw_ctx *c) {
    int m = 0;
    for (int Q = 0; Q < 4; ++Q) m = std::max(m, c->g.quad_off[Q + 1] - c->g.quad_off[Q]);
    return m;
}

// Row-group width J (ordinates per row group, one block of a quadrant each) and waves per workgroup W.
// Automatic: each quadrant split into the fewest equal blocks of at most 7 ordinates (a wavefront then
// holds 9 or more rows, so a band of four waves is 36 rows or more, and a row group still reads and
// writes runs of J N values); four waves per workgroup, fewer when the mesh has fewer rows.
std::pair<int, int> lanes_shape(const tsw_ctx *c) {
    const int nQ = std::max(lanes_quadrant_max(c), 1);
    int J = c->lanes_J_opt;
    if (J <= 0) {
        const int parts = (nQ + 6) / 7;
        J = (nQ + parts - 1) / parts;
    }
    J = std::min(std::max(J, 1), 64);
    const int RW = 64 / J;
    int W = c->lanes_W_opt > 0 ? c->lanes_W_opt : 4;
    if (!c->lanes_fixed_W) W = std::min(W, std::max(1, (c->g.ny + RW - 1) / RW));
    return {J, W};
}

// True when the ordinate-lane engine runs this context's sweeps: built for DG degrees 0-2 on the
// material path of a staggered pattern, selected by the "lanes" tuning key.
bool lanes_ok(tsw_ctx *c) {
    if (c->lanes_opt <= 0 || !is_staggered(c->pattern) || c->pattern == Pattern::StaggeredGE || c->xs_deg > 0)
        return false;
    return c->g.p <= 2 && c->g.npsi() * sizeof(double) < (size_t(1) << 31);  // 32-bit byte offsets
}

// Ordinate blocks of at most J ordinates, each quadrant split into equal parts (uploaded once per J).
void lanes_blocks(tsw_ctx *c, int J) {
    auto &L = c->lanes;
    if (L.J == J && L.nblk > 0) return;
    std::vector<LaneBlock> b;
    for (int Q = 0; Q < 4; ++Q) {
        const int q0 = c->g.quad_off[Q], n = c->g.quad_off[Q + 1] - q0;
        if (n == 0) continue;
        const int parts = (n + J - 1) / J;
        for (int k = 0; k < parts; ++k) {
            const int a = q0 + (int)((long long)n * k / parts), e = q0 + (int)((long long)n * (k + 1) / parts);
            b.push_back({a, e - a, Q, 0});
        }
    }
    TSW_HIP(hipMemcpyAsync(L.blk.reserve(sizeof(LaneBlock) * b.size()), b.data(), sizeof(LaneBlock) * b.size(),
                           hipMemcpyHostToDevice, c->stream));
    TSW_HIP(hipStreamSynchronize(c->stream));  // b goes out of scope
    L.J = J;
    L.nblk = (int)b.size();
}

// Exchange records for `nbands` bands of the lane engine (shared with the staggered kernels, whose records
// are all back to the sentinel after every sweep).
template <typename T>
T *lanes_records(tsw_ctx *c, int nbands) {
    const Geom &g = c->g;
    const size_t n = (size_t)g.nq * std::max(nbands - 1, 1) * g.nx * (g.p + 1);
    DBuf &buf = sizeof(T) == 8 ? c->xtr64 : c->xtr32;
    size_t &have = sizeof(T) == 8 ? c->xtr64_n : c->xtr32_n;
    if (have < n) {
        buf.reserve(n * sizeof(T));
        have = n;
        launch_exchange_fill<T>(buf.as<T>(), n, c->stream);
    }
    if (!c->abort_host) {
        TSW_HIP(hipHostMalloc((void **)&c->abort_host, sizeof(int), hipHostMallocDefault));
        *c->abort_host = 0;
        c->abort_dev.reserve(sizeof(int));
        TSW_HIP(hipMemsetAsync(c->abort_dev.ptr, 0, sizeof(int), c->stream));
    }
    if (nbands > 1) c->bands_used = true;
    return buf.as<T>();
}

// Moment partials of the ordinate-lane engine for its current block shape.
double *lanes_mom(tsw_ctx *c) {
    lanes_blocks(c, lanes_shape(c).first);
    return dbuf<double>(c->lanes.mom, (size_t)c->lanes.nblk * c->g.nphi());
}

// Everything an ordinate-lane sweep of this context allocates or reads back (block table, element
// inverses, exchange records, trace and moment buffers), so that later sweeps only launch kernels (and can
// be captured in a graph).
template <typename T>
void lanes_prepare(tsw_ctx *c, double dt, const double *src, bool mom) {
    const Geom &g = c->g;
    const std::pair<int, int> shape = lanes_shape(c);
    lanes_blocks(c, shape.first);
    require(diag_inverses<T>(c, dt) != nullptr, TSW_E_STATE, "ordinate-lane sweep without element inverses");
    const int rows = (mom ? shape.second : std::min(shape.second, kLanesStoreWaves)) * (64 / shape.first);
    (void)lanes_records<T>(c, (g.ny + rows - 1) / rows);
    if (src) {
        auto &L = c->lanes;
        const bool f64 = sizeof(T) == 8;
        dbuf<T>(f64 ? L.tinx64 : L.tinx32, (size_t)g.nq * g.ny * (g.p + 1));
        dbuf<T>(f64 ? L.tiny64 : L.tiny32, (size_t)g.nq * g.nx * (g.p + 1));
    }
    if (mom) (void)lanes_mom(c);
}

// Parameters of one ordinate-lane sweep: psi_out <- L^{-1}(src + sigma_s phi / (4 pi)) or, with mom, the
// per-block moment partials of that solution.  src (reference layout with the inflow in its halos), phi and
// sig_s (cell arrays) may each be null (not src and phi both); the material table must be valid.  With
// `inflow` the traces of src's halos are computed here (a batch computes every member's in one launch);
// threads and lds receive the launch shape.
template <typename T>
LanesParams<T> lanes_params(tsw_ctx *c, double dt, const double *src, const double *phi, const double *sig_s,
                            double *psi_out, double *mom, bool inflow, int &threads, size_t &lds) {
    const Geom &g = c->g;
    const bool f64 = sizeof(T) == 8;
    const std::pair<int, int> shape = lanes_shape(c);
    const int J = shape.first, W = mom ? shape.second : std::min(shape.second, kLanesStoreWaves);  // + storer
    lanes_blocks(c, J);
    LanesParams<T> sp{};
    sp.xi = f64 ? (const T *)c->xi64.ptr : (const T *)c->xi32.ptr;
    sp.eta = f64 ? (const T *)c->eta64.ptr : (const T *)c->eta32.ptr;
    sp.nx = g.nx; sp.ny = g.ny; sp.nq = g.nq;
    sp.qo1 = g.quad_off[1]; sp.qo2 = g.quad_off[2]; sp.qo3 = g.quad_off[3];
    sp.dx = T(g.dx); sp.dy = T(g.dy);
    sp.blk = c->lanes.blk.as<LaneBlock>();
    sp.nblk = c->lanes.nblk;
    sp.prefetch = c->lanes_pf_opt;
    require(sp.prefetch == 0 || sp.prefetch == (f64 ? 2 : 3) || sp.prefetch == (f64 ? 4 : 6), TSW_E_ARG,
            f64 ? "ordinate-lane prefetch depth must be 2 or 4 in FP64" : "ordinate-lane prefetch depth must be 3 or 6 in FP32");
    sp.J = J;
    sp.RW = 64 / J;
    sp.src = src;
    if (src) {  // inflow traces of src's halos
        const size_t nt_x = (size_t)g.nq * g.ny * (g.p + 1), nt_y = (size_t)g.nq * g.nx * (g.p + 1);
        auto &L = c->lanes;
        T *tx = dbuf<T>(f64 ? L.tinx64 : L.tinx32, nt_x), *ty = dbuf<T>(f64 ? L.tiny64 : L.tiny32, nt_y);
        if (inflow) launch_inflow<T>(g, src, tx, ty, c->stream);
        sp.tin_x = tx;
        sp.tin_y = ty;
    }
    sp.psi = psi_out;
    require(c->mat.valid, TSW_E_STATE, "ordinate-lane sweep without a material table");
    sp.mat = c->mat.idx.as<int32_t>();
    sp.ainv = diag_inverses<T>(c, dt);
    require(sp.ainv != nullptr, TSW_E_STATE, "ordinate-lane sweep without element inverses");
    sp.M = c->mat.M;
    sp.phi = phi;
    sp.sig_s = phi ? sig_s : nullptr;
    sp.mom = mom;
    sp.w = c->w64.as<double>();
    const int rows = W * sp.RW;
    sp.band_rows = rows;
    sp.nbands = (g.ny + rows - 1) / rows;
    sp.xtr = lanes_records<T>(c, sp.nbands);
    sp.abort_flag = c->abort_dev.as<int>();
    sp.skip = c->skip_flag;
    const int N = g.N();
    auto lds_of = [&](bool table) {
        const size_t vw = 16 / sizeof(T), ps = (g.p + 1 + vw - 1) / vw * vw;  // trace slot, 16-byte rows
        size_t b = ((size_t)2 * rows * J * ps * sizeof(T) + 15) / 16 * 16;  // LanesLds: traces
        if (mom) b += (size_t)2 * rows * N * J * sizeof(double);              // moment reduction
        b += ((size_t)W * 64 * N * sizeof(T) + 15) / 16 * 16;                // per-wave source stage
        if (!mom) b += ((size_t)2 * W * 64 * N * sizeof(T) + 15) / 16 * 16;  // storer stage
        if (table) b += ((size_t)sp.M * J * N * N * sizeof(T) + 15) / 16 * 16;
        return b;
    };
    sp.inv_lds = lds_of(true) <= 96 * 1024 ? 1 : 0;
    lds = lds_of(sp.inv_lds != 0);
    require(lds <= 160 * 1024, TSW_E_UNSUPPORTED, "ordinate-lane sweep: LDS layout too large");
    threads = (W + (mom ? 0 : 1)) * 64;
    c->last_lanes = 1;
    c->last_lanes_J = J;
    c->last_lanes_W = W;
    c->last_br = sp.nbands > 1 ? sp.band_rows : 0;
    return sp;
}

// Launch of a lane sweep (or of a batch of them, all of DG degree p).
template <typename T>
void lanes_launch(int p, const LanesParams<T> &sp, int threads, size_t lds, hipStream_t s,
                  const LanesParams<T> *batch = nullptr, const int4 *items = nullptr, int nitems = 0) {
    switch (p) {
        case 0: launch_lanes_sweep_p<T, 1>(sp, threads, lds, s, batch, items, nitems); break;
        case 1: launch_lanes_sweep_p<T, 2>(sp, threads, lds, s, batch, items, nitems); break;
        case 2: launch_lanes_sweep_p<T, 3>(sp, threads, lds, s, batch, items, nitems); break;
        default: throw Error(TSW_E_UNSUPPORTED, "ordinate-lane sweep built for DG degrees 0-2");
    }
}

// One ordinate-lane sweep (lanes_params), timed.
template <typename T>
void lanes_sweep(tsw_ctx *c, double dt, const double *src, const double *phi, const double *sig_s, double *psi_out,
                 double *mom) {
    hipStream_t s = c->stream;
    int threads = 0;
    size_t lds = 0;
    const LanesParams<T> sp = lanes_params<T>(c, dt, src, phi, sig_s, psi_out, mom, true, threads, lds);
    if (c->stall_test && sp.nbands > 1 && !c->skip_flag) {
        c->stall_test = false;
        *c->abort_host = 1;
        TSW_HIP(hipMemcpyAsync(c->abort_dev.ptr, c->abort_host, sizeof(int), hipMemcpyHostToDevice, s));
    }
    if (c->skip_flag) {  // inside a graph capture: the replay is timed as a whole
        lanes_launch<T>(c->g.p, sp, threads, lds, s);
        return;
    }
    hipEvent_t e0 = next_event(c), e1 = next_event(c);
    TSW_HIP(hipEventRecord(e0, s));
    lanes_launch<T>(c->g.p, sp, threads, lds, s);
    TSW_HIP(hipEventRecord(e1, s));
    ++c->sweeps_done;
}

// In-place sweep of a reference-layout device array on the ordinate-lane engine, computing in precision T.
template <typename T>
void apply_lanes(tsw_ctx *c, double *psi_ref, const double *sigma_t, double dt) {
    prepare_materials(c, sigma_t, dt);
    require(c->mat.valid, TSW_E_UNSUPPORTED, "ordinate-lane sweep needs at most 64 distinct sigma_t values");
    lanes_sweep<T>(c, dt, psi_ref, nullptr, nullptr, psi_ref, nullptr);
}

// In-place sweep of a reference-layout device array in precision T.
template <typename T>
void apply_staggered(tsw_ctx *c, double *psi_ref, const double *sigma_t, double dt) {
    const bool f64 = sizeof(T) == 8;
    diag_setup(c, sigma_t, nullptr, psi_ref, dt, f64, !f64);
    T *q = (T *)(f64 ? c->diag.q64.ptr : c->diag.q32.ptr);
    diag_sweep<T>(c, dt, q, nullptr, true, q);
    diag_store<T>(c, q, psi_ref);
}


// TSW_MIXED sweep on the staggered engine: psi_0 = L32^{-1} Q (FP32, with the inflow), then refine_steps
// times psi += L32^{-1}(Q - (L + 1/dt) psi), the residual formed in FP64 (with the FP64 inflow) and swept in
// FP32 with zero inflow; psi_ref (reference layout) receives the FP64 result.  With one step the FP64
// iterate is never stored: the residual reads psi_0 itself and the result is written as psi_0 + d.
void mixed_refine_store(tsw_ctx *c, double dt, const float *psi0, double *psi_ref) {
    const Geom &g = c->g;
    auto &d = c->diag;
    hipStream_t s = c->stream;
    const size_t nd = (size_t)g.nq * g.N() * g.nx * g.ny;
    float *res32 = dbuf<float>(d.res32, nd);
    auto residual = [&](const auto *psi) {
        launch_residual_diag<float>(g, d.q64.as<double>(), nullptr, nullptr, d.sigt64.as<double>(), psi,
                                    d.tinx64.as<double>(), d.tiny64.as<double>(), c->xi64.as<double>(),
                                    c->eta64.as<double>(), inverse_dt(dt), res32, s);
    };
    if (c->refine_steps == 1) {
        residual(psi0);
        diag_sweep<float>(c, dt, res32, nullptr, false, res32);
        launch_psi_store_sum(g, psi_ref, psi0, res32, s);
        return;
    }
    double *psi64 = dbuf<double>(d.ir64, nd);
    launch_convert_f2d(psi0, psi64, nd, s);
    for (int k = 0; k < c->refine_steps; ++k) {
        residual(static_cast<const double *>(psi64));
        diag_sweep<float>(c, dt, res32, nullptr, false, res32);
        launch_accumulate_f2d(res32, psi64, nd, s);
    }
    diag_store<double>(c, psi64, psi_ref);
}

void apply_staggered_mixed(tsw_ctx *c, double *psi_ref, const double *sigma_t, double dt) {
    diag_setup(c, sigma_t, nullptr, psi_ref, dt, true, true);
    float *q32 = c->diag.q32.as<float>();
    diag_sweep<float>(c, dt, q32, nullptr, true, q32);
    mixed_refine_store(c, dt, q32, psi_ref);
}

// Source iteration on the staggered engine (SISolver.cpp:206-249).  Q: reference-layout source with
// inflow halos (device).  The solver's angular flux state (reference layout) seeds phi and receives
// the last iterate.  Returns iterations and the last change.
template <typename T>
void si_staggered(tsw_ctx *c, double *Q, const double *st, const double *ss, double dt, double abs_tol, int max_its,
                  int &iters, double &err) {
    const Geom &g = c->g;
    hipStream_t s = c->stream;
    const bool f64 = sizeof(T) == 8;
    diag_setup(c, st, ss, Q, dt, f64, !f64);
    double *psi_state = c->si_state.as<double>();
    double *cur = dbuf<double>(c->phi64a, g.nphi());
    double *last = dbuf<double>(c->phi64b, g.nphi());
    TSW_HIP(hipMemsetAsync(cur, 0, sizeof(double) * g.nphi(), s));
    TSW_HIP(hipMemsetAsync(last, 0, sizeof(double) * g.nphi(), s));
    launch_pmv<double, double>(g, psi_state, cur, c->w64.as<double>(), 1.0, 0.0, 0, s);  // :214
    const size_t npsi_d = (size_t)g.nq * g.N() * g.nx * g.ny;
    T *psi_d = dbuf<T>(f64 ? c->diag.psi64 : c->diag.psi32, npsi_d);
    const T *q_d = (const T *)(f64 ? c->diag.q64.ptr : c->diag.q32.ptr);
    int it = 0;
    double e = 1.79769313486231570e+308;
    while (it < max_its && e > abs_tol) {
        ++it;
        std::swap(cur, last);
        const T *phi_d = diag_phi<T>(c, last);
        diag_sweep<T>(c, dt, q_d, phi_d, true, psi_d);   // L^{-1}(Q + S phi), :226-232
        diag_pmv<T>(c, psi_d, cur, 1.0, 0.0);            // :236
        e = device_norm2_interior_diff(c, cur, last);    // :239
    }
    if (it > 0) {
        TSW_HIP(hipMemcpyAsync(psi_state, Q, sizeof(double) * g.npsi(), hipMemcpyDeviceToDevice, s));
        diag_store<T>(c, psi_d, psi_state);
        TSW_HIP(hipMemcpyAsync(Q, psi_state, sizeof(double) * g.npsi(), hipMemcpyDeviceToDevice, s));
    } else {
        TSW_HIP(hipMemcpyAsync(Q, psi_state, sizeof(double) * g.npsi(), hipMemcpyDeviceToDevice, s));
    }
    iters = it;
    err = e;
}

// Source iteration replayed from a graph: kSiChunk iterations (an even number, so the scalar-flux buffers
// alternate the same way in every replay) are captured once per call, each iteration being the
// conversion of phi, the fused sweep L^{-1}(Q + S phi), the moment contraction and the device-side loop
// test; replays continue until the test stops the loop, and iterations after that point do nothing, so
// the iterates, the count and the last change equal the eager loop's.  Timing: each replay is one
// interval of tsw_sweep_timing, counted as the iterations it ran.
template <typename T>
void si_staggered_graph(tsw_ctx *c, double *Q, const double *st, const double *ss, double dt, double abs_tol,
                        int max_its, int &iters, double &err) {
    const Geom &g = c->g;
    hipStream_t s = c->stream;
    const bool f64 = sizeof(T) == 8;
    diag_setup(c, st, ss, Q, dt, f64, !f64);
    double *psi_state = c->si_state.as<double>();
    double *bufA = dbuf<double>(c->phi64a, g.nphi());
    double *bufB = dbuf<double>(c->phi64b, g.nphi());
    TSW_HIP(hipMemsetAsync(bufA, 0, sizeof(double) * g.nphi(), s));
    TSW_HIP(hipMemsetAsync(bufB, 0, sizeof(double) * g.nphi(), s));
    launch_pmv<double, double>(g, psi_state, bufA, c->w64.as<double>(), 1.0, 0.0, 0, s);  // :214
    const size_t npsi_d = (size_t)g.nq * g.N() * g.nx * g.ny;
    T *psi_d = dbuf<T>(f64 ? c->diag.psi64 : c->diag.psi32, npsi_d);
    const T *q_d = (const T *)(f64 ? c->diag.q64.ptr : c->diag.q32.ptr);
    const int nb = reduction_blocks();
    double *part = dbuf<double>(c->red_part, (size_t)nb * 64);
    // everything the captured kernels use exists before the capture: buffers, tables, tuning
    dbuf<T>(f64 ? c->diag.phid64 : c->diag.phid32, 4 * (size_t)g.nx * g.ny * g.N());
    dbuf<double>(c->diag.phiQ, 4 * (size_t)g.nx * g.ny * g.N());
    select_tuning<T>(c, dt, q_d, diag_phi<T>(c, bufA), true);
    (void)make_diag_params<T>(c, dt, q_d, (const T *)(f64 ? c->diag.phid64.ptr : c->diag.phid32.ptr), true, psi_d);
    SiControl *ctl = static_cast<SiControl *>(c->si_ctl.reserve(sizeof(SiControl)));
    if (!c->si_ctl_host) TSW_HIP(hipHostMalloc((void **)&c->si_ctl_host, sizeof(SiControl), hipHostMallocDefault));
    SiControl init{};
    init.done = max_its <= 0 ? 1 : 0;
    init.iters = 0;
    init.max_its = max_its;
    init.tol = abs_tol;
    init.err = 1.79769313486231570e+308;
    TSW_HIP(hipMemcpyAsync(ctl, &init, sizeof(SiControl), hipMemcpyHostToDevice, s));
    TSW_HIP(hipStreamSynchronize(s));

    // iteration k (1-based) writes phi into bufB when k is odd, bufA when even (the eager loop's swap)
    hipGraph_t graph;
    c->skip_flag = &ctl->done;
    TSW_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        for (int j = 0; j < tsw_ctx::kSiChunk; ++j) {
            double *cur = (j & 1) ? bufA : bufB, *last = (j & 1) ? bufB : bufA;
            const T *phi_d = diag_phi<T>(c, last);
            diag_sweep<T>(c, dt, q_d, phi_d, true, psi_d);  // L^{-1}(Q + S phi), :226-232
            diag_pmv<T>(c, psi_d, cur, 1.0, 0.0);          // :236
            launch_si_check(g, cur, last, part, nb, ctl, s); // :239 and the loop test
        }
    } catch (...) {
        c->skip_flag = nullptr;
        (void)hipStreamEndCapture(s, &graph);
        throw;
    }
    c->skip_flag = nullptr;
    TSW_HIP(hipStreamEndCapture(s, &graph));
    hipGraphExec_t exec;
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    TSW_HIP(ie);
    try {
        SiControl &h = *c->si_ctl_host;
        h = init;
        while (!h.done) {
            const int before = h.iters;
            hipEvent_t e0 = next_event(c), e1 = next_event(c);
            TSW_HIP(hipEventRecord(e0, s));
            TSW_HIP(hipGraphLaunch(exec, s));
            TSW_HIP(hipEventRecord(e1, s));
            TSW_HIP(hipMemcpyAsync(&h, ctl, sizeof(SiControl), hipMemcpyDeviceToHost, s));
            sync(c);
            c->sweeps_done += h.iters - before;
        }
    } catch (...) {
        (void)hipGraphExecDestroy(exec);
        throw;
    }
    (void)hipGraphExecDestroy(exec);
    const SiControl &h = *c->si_ctl_host;
    if (h.iters > 0) {
        TSW_HIP(hipMemcpyAsync(psi_state, Q, sizeof(double) * g.npsi(), hipMemcpyDeviceToDevice, s));
        diag_store<T>(c, psi_d, psi_state);
    }
    TSW_HIP(hipMemcpyAsync(Q, psi_state, sizeof(double) * g.npsi(), hipMemcpyDeviceToDevice, s));
    iters = h.iters;
    err = h.err;
}

// Source iteration on the ordinate-lane engine (the material table must be valid): every iteration
// sweeps Q + S phi straight into per-block moments, so no angular flux is written until the end, where one
// more sweep of the last iteration's source writes psi (the same values that iteration's sweep computed).
// With `graph`, kSiChunk iterations are captured and replayed with the device-side loop test, as in
// si_staggered_graph.
template <typename T>
void si_lanes(tsw_ctx *c, double *Q, const double *st, const double *ss, double dt, double abs_tol, int max_its,
              bool graph, int &iters, double &err) {
    const Geom &g = c->g;
    hipStream_t s = c->stream;
    double *psi_state = c->si_state.as<double>();
    double *bufA = dbuf<double>(c->phi64a, g.nphi());
    double *bufB = dbuf<double>(c->phi64b, g.nphi());
    TSW_HIP(hipMemsetAsync(bufA, 0, sizeof(double) * g.nphi(), s));
    TSW_HIP(hipMemsetAsync(bufB, 0, sizeof(double) * g.nphi(), s));
    launch_pmv<double, double>(g, psi_state, bufA, c->w64.as<double>(), 1.0, 0.0, 0, s);  // :214
    double *mom = lanes_mom(c);
    const int nblk = c->lanes.nblk;
    lanes_prepare<T>(c, dt, Q, true);  // tables, records and trace buffers exist before any capture
    auto iteration = [&](double *cur, const double *last) {
        lanes_sweep<T>(c, dt, Q, last, ss, nullptr, mom);           // L^{-1}(Q + S phi) (:226-232) ...
        launch_mom_sum(g, nblk, mom, cur, cur, 1.0, 0.0, s, c->skip_flag);  // ... its moments (:236)
    };
    int it = 0;
    double e = 1.79769313486231570e+308;
    if (!graph) {
        double *cur = bufA, *last = bufB;
        while (it < max_its && e > abs_tol) {
            ++it;
            std::swap(cur, last);
            iteration(cur, last);
            e = device_norm2_interior_diff(c, cur, last);  // :239
        }
        if (it > 0) {
            lanes_sweep<T>(c, dt, Q, last, ss, Q, nullptr);  // psi of the last iteration
            TSW_HIP(hipMemcpyAsync(psi_state, Q, sizeof(double) * g.npsi(), hipMemcpyDeviceToDevice, s));
        } else {
            TSW_HIP(hipMemcpyAsync(Q, psi_state, sizeof(double) * g.npsi(), hipMemcpyDeviceToDevice, s));
        }
        iters = it;
        err = e;
        return;
    }
    const int nb = reduction_blocks();
    double *part = dbuf<double>(c->red_part, (size_t)nb * 64);
    SiControl *ctl = static_cast<SiControl *>(c->si_ctl.reserve(sizeof(SiControl)));
    if (!c->si_ctl_host) TSW_HIP(hipHostMalloc((void **)&c->si_ctl_host, sizeof(SiControl), hipHostMallocDefault));
    SiControl init{};
    init.done = max_its <= 0 ? 1 : 0;
    init.max_its = max_its;
    init.tol = abs_tol;
    init.err = 1.79769313486231570e+308;
    TSW_HIP(hipMemcpyAsync(ctl, &init, sizeof(SiControl), hipMemcpyHostToDevice, s));
    TSW_HIP(hipStreamSynchronize(s));
    // iteration k (1-based) writes phi into bufB when k is odd, bufA when even
    hipGraph_t gr;
    c->skip_flag = &ctl->done;
    TSW_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        for (int j = 0; j < tsw_ctx::kSiChunk; ++j) {
            double *cur = (j & 1) ? bufA : bufB, *last = (j & 1) ? bufB : bufA;
            iteration(cur, last);
            launch_si_check(g, cur, last, part, nb, ctl, s);  // :239 and the loop test
        }
    } catch (...) {
        c->skip_flag = nullptr;
        (void)hipStreamEndCapture(s, &gr);
        throw;
    }
    c->skip_flag = nullptr;
    TSW_HIP(hipStreamEndCapture(s, &gr));
    hipGraphExec_t exec;
    const hipError_t ie = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    TSW_HIP(ie);
    try {
        SiControl &h = *c->si_ctl_host;
        h = init;
        while (!h.done) {
            const int before = h.iters;
            hipEvent_t e0 = next_event(c), e1 = next_event(c);
            TSW_HIP(hipEventRecord(e0, s));
            TSW_HIP(hipGraphLaunch(exec, s));
            TSW_HIP(hipEventRecord(e1, s));
            TSW_HIP(hipMemcpyAsync(&h, ctl, sizeof(SiControl), hipMemcpyDeviceToHost, s));
            sync(c);
            c->sweeps_done += h.iters - before;
        }
    } catch (...) {
        (void)hipGraphExecDestroy(exec);
        throw;
    }
    (void)hipGraphExecDestroy(exec);
    const SiControl &h = *c->si_ctl_host;
    if (h.iters > 0) {
        const double *last = (h.iters & 1) ? bufA : bufB;  // the scalar flux the last iteration swept
        lanes_sweep<T>(c, dt, Q, last, ss, Q, nullptr);
        TSW_HIP(hipMemcpyAsync(psi_state, Q, sizeof(double) * g.npsi(), hipMemcpyDeviceToDevice, s));
    } else {
        TSW_HIP(hipMemcpyAsync(Q, psi_state, sizeof(double) * g.npsi(), hipMemcpyDeviceToDevice, s));
    }
    iters = h.iters;
    err = h.err;
}

// Peierls GMRES on the staggered engine.  FP64: sweeps in double.  FP32: an FP64 iterate psi64 (diagonal
// layout) refined with corrections whose sweeps run in float; the residual r = Q + S P psi - (L + 1/dt) psi
// is formed in FP64.  Q: reference layout with inflow halos, overwritten by psi; phi (phi layout) receives
// the scalar flux.  Returns convergence.
bool gmres_staggered(tsw_ctx *c, double *Q, const double *st, const double *ss, double dt, double abs_tol,
                     int max_its, int restart, int precision, double *phi_out, int &its_total, double &res_final) {
    const Geom &g = c->g;
    hipStream_t s = c->stream;
    const size_t n = g.nphi();
    const size_t npsi_d = (size_t)g.nq * g.N() * g.nx * g.ny;
    double *b = dbuf<double>(c->gm_vecs, 3 * n);
    double *x = b + n, *r = x + n;
    TSW_HIP(hipMemsetAsync(b, 0, sizeof(double) * 3 * n, s));
    Gmres solver(c, n, restart);
    bool converged = false;
    its_total = 0;
    // left DSA preconditioning, as PETSc applies the reference's shell preconditioner (the stop test is
    // then on the preconditioned residual, PETSc's default norm)
    const bool dsa = c->use_dsa;
    // the unchanged-input checks of the material table and the DSA hierarchy in one round trip
    KeyCheckScope checks(c);
    {
        const size_t nh = g.ncell_h();
        std::vector<std::tuple<const double *, size_t, DBuf *, bool>> items;
        if (c->pattern == Pattern::Staggered) items.emplace_back(st, nh, &c->mat.key, c->mat.keyed);
        if (dsa) {
            items.emplace_back(st, nh, &c->dsa.key_st, c->dsa.keyed);
            items.emplace_back(ss, nh, &c->dsa.key_ss, c->dsa.keyed);
        }
        precheck_keys(c, items);
    }
    if (dsa) dsa_build(c, st, ss, dt);
    if (precision == TSW_FP64) {
        diag_setup(c, st, ss, Q, dt, true, false);
        double *psi_d = dbuf<double>(c->diag.psi64, npsi_d);
        const double *q_d = c->diag.q64.as<double>();
        auto apply = [&](const double *xv, double *yv) {  // (I - P L^{-1} S) x, zero inflow (:548-555)
            const double *phi_d = diag_phi<double>(c, xv);
            diag_sweep<double>(c, dt, nullptr, phi_d, false, psi_d);
            diag_pmv<double>(c, psi_d, yv, -1.0, 1.0, xv);  // y = x - P psi (y's halos stay zero)
            if (dsa) dsa_apply(c, yv, yv);
        };
        diag_sweep<double>(c, dt, q_d, nullptr, true, psi_d);  // b = P L^{-1} Q (:461-466)
        diag_pmv<double>(c, psi_d, b, 1.0, 0.0);
        if (dsa) dsa_apply(c, b, b);
        converged = solver.solve(apply, b, x, r, abs_tol, max_its, its_total, res_final, true);
        const double *phi_d = diag_phi<double>(c, x);            // psi = L^{-1}(Q + S phi) (:504-505)
        diag_sweep<double>(c, dt, q_d, phi_d, true, psi_d);
        diag_store<double>(c, psi_d, Q);
    } else if (!c->gmres_refine) {
        // Mixed precision, relaxed: FP64 sweeps for b = P L^{-1} Q, the Krylov steps while the residual
        // estimate exceeds abs_tol / (kFp32ErrSafety d) and the final psi; FP32 sweeps for the other steps.
        // The Krylov vectors, the orthogonalisation and the DSA correction stay FP64.
        diag_setup(c, st, ss, Q, dt, true, true, false);
        double *psi_d = dbuf<double>(c->diag.psi64, npsi_d);
        float *psi32 = dbuf<float>(c->diag.psi32, npsi_d);
        const double *q_d = c->diag.q64.as<double>();
        auto apply64 = [&](const double *xv, double *yv) {
            diag_sweep<double>(c, dt, nullptr, diag_phi<double>(c, xv), false, psi_d);
            diag_pmv<double>(c, psi_d, yv, -1.0, 1.0, xv);
            if (dsa) dsa_apply(c, yv, yv);
        };
        auto apply32 = [&](const double *xv, double *yv) {
            diag_sweep<float>(c, dt, nullptr, diag_phi<float>(c, xv), false, psi32);
            diag_pmv<float>(c, psi32, yv, -1.0, 1.0, xv);
            if (dsa) dsa_apply(c, yv, yv);
        };
        diag_sweep<double>(c, dt, q_d, nullptr, true, psi_d);
        diag_pmv<double>(c, psi_d, b, 1.0, 0.0);
        if (dsa) dsa_apply(c, b, b);
        // measured FP32 product error, once per material table, dt and DSA setting
        tsw_ctx::Fp32Err &fe = fp32_err_entry(c, dt, dsa);
        if (!(fe.value >= 0.0)) {
            const double nb = solver.norm(b);
            double d = kFp32ErrFloor;
            if (nb > 0.0) {
                apply64(b, x);
                apply32(b, r);
                launch_axpby(n, -1.0, x, 1.0, r, s);  // r = A32 b - A64 b
                d = solver.norm(r) / nb;
                TSW_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
            }
            fe.value = std::min(kFp32ErrCeil, std::max(kFp32ErrFloor, d));
        }
        const double lo_below = abs_tol / (c->fp32_safety * fe.value);
        converged = solver.solve(apply64, apply32, lo_below, b, x, r, abs_tol, max_its, its_total, res_final, true);
        // psi = L^{-1}(Q + S phi) in FP64; with FP32 products in the solve, the true residual
        // b - A x = M (P psi - x) comes from the same sweep, and the solve resumes from x while it exceeds
        // abs_tol (FP64 restart residual, relaxed steps as before)
        for (int check = 0;; ++check) {
            diag_sweep<double>(c, dt, q_d, diag_phi<double>(c, x), true, psi_d);
            if (!solver.used_lo) break;
            diag_pmv<double>(c, psi_d, r, 1.0, -1.0, x);
            if (dsa) dsa_apply(c, r, r);
            res_final = solver.norm(r);
            converged = res_final <= abs_tol;
            if (converged || its_total >= max_its || check >= 3) break;
            int its = 0;
            double rr = 0.0;
            converged = solver.solve(apply64, apply32, lo_below, b, x, r, abs_tol, max_its - its_total, its, rr, false);
            its_total += its;
        }
        diag_store<double>(c, psi_d, Q);
    } else {
        diag_setup(c, st, ss, Q, dt, true, true, false);  // FP32 sweeps of residuals: no FP32 source or inflow
        double *psi64 = dbuf<double>(c->diag.ir64, npsi_d);
        float *res32 = dbuf<float>(c->diag.res32, npsi_d);
        float *psi32 = dbuf<float>(c->diag.psi32, npsi_d);
        const double *q64 = c->diag.q64.as<double>();
        TSW_HIP(hipMemsetAsync(psi64, 0, sizeof(double) * npsi_d, s));
        auto apply32 = [&](const double *xv, double *yv) {
            const float *phi_d = diag_phi<float>(c, xv);
            diag_sweep<float>(c, dt, nullptr, phi_d, false, psi32);
            diag_pmv<float>(c, psi32, yv, -1.0, 1.0, xv);
            if (dsa) dsa_apply(c, yv, yv);
        };
        const int max_refine = 30;
        for (int refine = 0;; ++refine) {
            // r = Q + S P psi - (L + 1/dt) psi (FP64), stored in float
            TSW_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
            diag_pmv<double>(c, psi64, x, 1.0, 0.0);
            const double *phi64_d = diag_phi<double>(c, x);
            launch_residual_diag<float>(g, q64, phi64_d, c->diag.sigs64.as<double>(), c->diag.sigt64.as<double>(),
                                        psi64, c->diag.tinx64.as<double>(), c->diag.tiny64.as<double>(),
                                        c->xi64.as<double>(), c->eta64.as<double>(), inverse_dt(dt), res32, s);
            // b = P L^{-1} r, the Peierls residual of the iterate
            diag_sweep<float>(c, dt, res32, nullptr, false, psi32);
            TSW_HIP(hipMemsetAsync(b, 0, sizeof(double) * n, s));
            diag_pmv<float>(c, psi32, b, 1.0, 0.0);
            if (dsa) dsa_apply(c, b, b);
            res_final = solver.norm(b);
            if (res_final <= abs_tol) {
                converged = true;
                break;
            }
            if (refine >= max_refine || its_total >= max_its) break;
            TSW_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
            const double inner_tol = std::max(0.5 * abs_tol, 1e-5 * res_final);
            int its = 0;
            double rr = 0.0;
            solver.solve(apply32, b, x, r, inner_tol, max_its - its_total, its, rr, true);
            its_total += its;
            // d = L^{-1}(r + S phi_d); psi += d
            const float *phi_d = diag_phi<float>(c, x);
            diag_sweep<float>(c, dt, res32, phi_d, false, psi32);
            launch_accumulate_f2d(psi32, psi64, npsi_d, s);
        }
        diag_store<double>(c, psi64, Q);
        TSW_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
        diag_pmv<double>(c, psi64, x, 1.0, 0.0);
    }
    if (phi_out) TSW_HIP(hipMemcpyAsync(phi_out, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    return converged;
}

// Peierls GMRES on the wavefront kernels (meshes the staggered kernel cannot hold).  Q: reference-layout
// source with inflow halos (device), overwritten by psi; pd (may be null): device phi output.
bool gmres_wavefront(tsw_ctx *ctx, double *Q, const double *st, const double *ss, double dt, double abs_tol,
                     int max_its, int restart, int precision, double *pd, int &its_total, double &res_final) {
    const Geom &g = ctx->g;
    hipStream_t s = ctx->stream;
    const size_t n = g.nphi(), npsi = g.npsi();
    const double *w = ctx->w64.as<double>();
    bool converged = false;
    its_total = 0;
    const bool dsa = ctx->use_dsa;
    if (dsa) dsa_build(ctx, xs_means(ctx, st), xs_means(ctx, ss), dt);
    const double *xt64 = xs_tensor<double>(ctx, st), *xss64 = xs_tensor<double>(ctx, ss);
    double *psi = static_cast<double *>(ctx->psi64.reserve(sizeof(double) * npsi));
    // vectors: b, x, r (phi-shaped, zero halos)
    double *b = static_cast<double *>(ctx->gm_vecs.reserve(sizeof(double) * n * 3));
    double *x = b + n, *r = x + n;
    TSW_HIP(hipMemsetAsync(b, 0, sizeof(double) * n * 3, s));
    Gmres solver(ctx, n, restart);

    // matvec y = (I - P L^{-1} S) x with zero inflow (PETScPeierlsSolver.cpp:548-555)
    auto apply = [&](const double *xv, double *yv) {
        launch_smv<double>(g, ss, xv, psi, 1.0, 0.0, 1, s, xss64);
        sweep_prec(ctx, psi, st, dt, precision);
        TSW_HIP(hipMemcpyAsync(yv, xv, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        launch_pmv<double, double>(g, psi, yv, w, -1.0, 1.0, 0, s);
        if (dsa) dsa_apply(ctx, yv, yv);
    };

    if (precision == TSW_FP64) {
        // b = P L^{-1} Q (:461-466)
        TSW_HIP(hipMemcpyAsync(psi, Q, sizeof(double) * npsi, hipMemcpyDeviceToDevice, s));
        sweep_prec(ctx, psi, st, dt, precision);
        launch_pmv<double, double>(g, psi, b, w, 1.0, 0.0, 0, s);
        if (dsa) dsa_apply(ctx, b, b);
        converged = solver.solve(apply, b, x, r, abs_tol, max_its, its_total, res_final, true);
        // source <- L^{-1}(Q + S phi) (:504-505)
        launch_smv<double>(g, ss, x, Q, 1.0, 1.0, 1, s, xss64);
        sweep_prec(ctx, Q, st, dt, precision);
    } else {
        // Mixed precision: an FP64 iterate psi64 is refined with corrections computed from FP32
        // sweeps.  The residual r = Q + S P psi - L psi is formed in FP64 with Lmv; the correction
        // solves (L - S P) d = r on the Peierls form with FP32 sweeps inside FP64 GMRES.
        const size_t ncell = g.ncell_h();
        const int N = g.N();
        double *psi64 = static_cast<double *>(ctx->ir_state.reserve(sizeof(double) * npsi));
        double *res64 = static_cast<double *>(ctx->r64.reserve(sizeof(double) * npsi));
        float *p32 = static_cast<float *>(ctx->psi32.reserve(sizeof(float) * npsi));
        float *phi32 = static_cast<float *>(ctx->phi32.reserve(sizeof(float) * n));
        const float *st32 = as_f32(ctx, st, ncell, ctx->sig_t32);
        const float *ss32 = as_f32(ctx, ss, ncell, ctx->sig_s32);
        const float *xss32 = xs_tensor<float>(ctx, ss);
        const double *xi64 = ctx->xi64.as<double>(), *eta64 = ctx->eta64.as<double>();

        // y = (I - P L^{-1} S) x with FP32 sweeps
        auto apply32 = [&](const double *xv, double *yv) {
            launch_convert_d2f(xv, phi32, n, s);
            launch_smv<float>(g, ss32, phi32, p32, 1.0f, 0.0f, 1, s, xss32);
            sweep_dev<float>(ctx, p32, st32, dt, st);
            TSW_HIP(hipMemcpyAsync(yv, xv, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            launch_pmv<float, double>(g, p32, yv, w, -1.0, 1.0, 0, s);
            if (dsa) dsa_apply(ctx, yv, yv);
        };

        // initial iterate: inflow halos from Q, zero interior
        TSW_HIP(hipMemcpyAsync(psi64, Q, sizeof(double) * npsi, hipMemcpyDeviceToDevice, s));
        launch_fill_region<double>(g, psi64, (size_t)g.nq * N, 0, 0.0, s);
        const int max_refine = 30;
        for (int refine = 0;; ++refine) {
            // r = Q + S P psi - (L + 1/dt) psi on the interior, zero halos
            TSW_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
            launch_pmv<double, double>(g, psi64, x, w, 1.0, 0.0, 0, s);
            TSW_HIP(hipMemcpyAsync(res64, Q, sizeof(double) * npsi, hipMemcpyDeviceToDevice, s));
            launch_smv<double>(g, ss, x, res64, 1.0, 1.0, 0, s, xss64);
            launch_lmv<double>(g, xi64, eta64, st, psi64, res64, -1.0, 1.0, s, xt64);
            if (inverse_dt(dt) != 0.0) launch_axpby(npsi, -inverse_dt(dt), psi64, 1.0, res64, s);
            launch_fill_region<double>(g, res64, (size_t)g.nq * N, 1, 0.0, s);
            // b = P L^{-1} r: the Peierls residual of the current iterate
            launch_convert_d2f(res64, p32, npsi, s);
            sweep_dev<float>(ctx, p32, st32, dt, st);
            TSW_HIP(hipMemsetAsync(b, 0, sizeof(double) * n, s));
            launch_pmv<float, double>(g, p32, b, w, 1.0, 0.0, 0, s);
            if (dsa) dsa_apply(ctx, b, b);
            res_final = solver.norm(b);
            if (res_final <= abs_tol) {
                converged = true;
                break;
            }
            if (refine >= max_refine || its_total >= max_its) break;
            // correction: (I - P L^{-1} S) phi_d = b, then d = L^{-1}(r + S phi_d)
            TSW_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
            const double inner_tol = std::max(0.5 * abs_tol, 1e-5 * res_final);
            int its = 0;
            double rr = 0.0;
            solver.solve(apply32, b, x, r, inner_tol, max_its - its_total, its, rr, true);
            its_total += its;
            launch_convert_d2f(res64, p32, npsi, s);
            launch_convert_d2f(x, phi32, n, s);
            launch_smv<float>(g, ss32, phi32, p32, 1.0f, 1.0f, 1, s, xss32);
            sweep_dev<float>(ctx, p32, st32, dt, st);
            launch_accumulate_f2d(p32, psi64, npsi, s);
        }
        // outputs: psi and phi = P psi
        TSW_HIP(hipMemcpyAsync(Q, psi64, sizeof(double) * npsi, hipMemcpyDeviceToDevice, s));
        TSW_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
        launch_pmv<double, double>(g, psi64, x, w, 1.0, 0.0, 0, s);
    }
    if (pd) TSW_HIP(hipMemcpyAsync(pd, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    return converged;
}

// Peierls GMRES on the ordinate-lane engine (the material table must be valid).  The same algorithm as
// gmres_staggered on the reference layouts themselves: the Arnoldi products sweep the scattering source
// of x straight into per-block moments (no angular flux is written), b = P L^{-1} Q likewise, and the final
// psi = L^{-1}(Q + S phi) is swept in place over Q.  FP32 (relaxed) as in gmres_staggered.
bool gmres_lanes(tsw_ctx *c, double *Q, const double *st, const double *ss, double dt, double abs_tol, int max_its,
                 int restart, int precision, double *phi_out, int &its_total, double &res_final) {
    const Geom &g = c->g;
    hipStream_t s = c->stream;
    const size_t n = g.nphi();
    double *b = dbuf<double>(c->gm_vecs, 3 * n);
    double *x = b + n, *r = x + n;
    TSW_HIP(hipMemsetAsync(b, 0, sizeof(double) * 3 * n, s));
    Gmres solver(c, n, restart);
    bool converged = false;
    its_total = 0;
    const bool dsa = c->use_dsa;
    if (dsa) dsa_build(c, st, ss, dt);
    double *mom = lanes_mom(c);
    const int nblk = c->lanes.nblk;
    auto apply64 = [&](const double *xv, double *yv) {  // (I - P L^{-1} S) x, zero inflow (:548-555)
        lanes_sweep<double>(c, dt, nullptr, xv, ss, nullptr, mom);
        launch_mom_sum(g, nblk, mom, xv, yv, -1.0, 1.0, s);
        if (dsa) dsa_apply(c, yv, yv);
    };
    auto apply32 = [&](const double *xv, double *yv) {
        lanes_sweep<float>(c, dt, nullptr, xv, ss, nullptr, mom);
        launch_mom_sum(g, nblk, mom, xv, yv, -1.0, 1.0, s);
        if (dsa) dsa_apply(c, yv, yv);
    };
    lanes_sweep<double>(c, dt, Q, nullptr, nullptr, nullptr, mom);  // b = P L^{-1} Q (:461-466)
    launch_mom_sum(g, nblk, mom, b, b, 1.0, 0.0, s);
    if (dsa) dsa_apply(c, b, b);
    if (precision == TSW_FP64) {
        converged = solver.solve(apply64, b, x, r, abs_tol, max_its, its_total, res_final, true);
        lanes_sweep<double>(c, dt, Q, x, ss, Q, nullptr);  // psi = L^{-1}(Q + S phi) (:504-505)
    } else {
        tsw_ctx::Fp32Err &fe = fp32_err_entry(c, dt, dsa);
        if (!(fe.value >= 0.0)) {
            const double nb = solver.norm(b);
            double d = kFp32ErrFloor;
            if (nb > 0.0) {
                apply64(b, x);
                apply32(b, r);
                launch_axpby(n, -1.0, x, 1.0, r, s);  // r = A32 b - A64 b
                d = solver.norm(r) / nb;
                TSW_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
            }
            fe.value = std::min(kFp32ErrCeil, std::max(kFp32ErrFloor, d));
        }
        const double lo_below = abs_tol / (c->fp32_safety * fe.value);
        converged = solver.solve(apply64, apply32, lo_below, b, x, r, abs_tol, max_its, its_total, res_final, true);
        // the true residual M (P psi - x) from the final FP64 psi; resume while it exceeds abs_tol
        const double *q_src = Q;
        if (solver.used_lo) {  // keep Q: the check may need further sweeps of it
            double *keep = dbuf<double>(c->lanes.psi_keep, g.npsi());
            TSW_HIP(hipMemcpyAsync(keep, Q, sizeof(double) * g.npsi(), hipMemcpyDeviceToDevice, s));
            q_src = keep;
        }
        for (int check = 0;; ++check) {
            lanes_sweep<double>(c, dt, q_src, x, ss, Q, nullptr);
            if (!solver.used_lo) break;
            launch_pmv<double, double>(g, Q, r, c->w64.as<double>(), 1.0, 0.0, 0, s);
            launch_axpby(n, -1.0, x, 1.0, r, s);
            if (dsa) dsa_apply(c, r, r);
            res_final = solver.norm(r);
            converged = res_final <= abs_tol;
            if (converged || its_total >= max_its || check >= 3) break;
            int its = 0;
            double rr = 0.0;
            converged = solver.solve(apply64, apply32, lo_below, b, x, r, abs_tol, max_its - its_total, its, rr, false);
            its_total += its;
        }
    }
    if (phi_out) TSW_HIP(hipMemcpyAsync(phi_out, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    return converged;
}

// Peierls GMRES on device arrays with the context's engine.
bool gmres_device(tsw_ctx *ctx, double *Q, const double *st, const double *ss, double dt, double abs_tol, int max_its,
                  int restart, int precision, double *pd, int &its_total, double &res_final) {
    ctx->last_lanes = 0;
    if (lanes_ok(ctx) && (precision == TSW_FP64 || !ctx->gmres_refine)) {
        prepare_materials(ctx, st, dt);
        if (ctx->mat.valid)
            return gmres_lanes(ctx, Q, st, ss, dt, abs_tol, max_its, restart, precision, pd, its_total, res_final);
    }
    if (staggered_ok(ctx))
        return gmres_staggered(ctx, Q, st, ss, dt, abs_tol, max_its, restart, precision, pd, its_total, res_final);
    return gmres_wavefront(ctx, Q, st, ss, dt, abs_tol, max_its, restart, precision, pd, its_total, res_final);
}

// In-place sweep of a reference-layout device array with the context's engine.
// TSW_MIXED on the wavefront engine: from the iterate with zero interior and Q's inflow halos,
// 1 + refine_steps times psi += L32^{-1}(Q - (L + 1/dt) psi) (FP64 residual, zero-inflow FP32 sweep).
void sweep_wavefront_mixed(tsw_ctx *ctx, double *Q, const double *st, double dt) {
    const Geom &g = ctx->g;
    hipStream_t s = ctx->stream;
    const size_t npsi = g.npsi();
    const size_t per = (size_t)g.nq * g.N();
    double *psi64 = static_cast<double *>(ctx->ir_state.reserve(sizeof(double) * npsi));
    double *res64 = static_cast<double *>(ctx->r64.reserve(sizeof(double) * npsi));
    float *p32 = static_cast<float *>(ctx->psi32.reserve(sizeof(float) * npsi));
    const float *st32 = as_f32(ctx, st, g.ncell_h(), ctx->sig_t32);
    TSW_HIP(hipMemcpyAsync(psi64, Q, sizeof(double) * npsi, hipMemcpyDeviceToDevice, s));
    launch_fill_region<double>(g, psi64, per, 0, 0.0, s);
    for (int k = 0; k <= ctx->refine_steps; ++k) {
        TSW_HIP(hipMemcpyAsync(res64, Q, sizeof(double) * npsi, hipMemcpyDeviceToDevice, s));
        launch_lmv<double>(g, ctx->xi64.as<double>(), ctx->eta64.as<double>(), st, psi64, res64, -1.0, 1.0, s,
                           xs_tensor<double>(ctx, st));
        if (inverse_dt(dt) != 0.0) launch_axpby(npsi, -inverse_dt(dt), psi64, 1.0, res64, s);
        launch_fill_region<double>(g, res64, per, 1, 0.0, s);
        launch_convert_d2f(res64, p32, npsi, s);
        sweep_dev<float>(ctx, p32, st32, dt, st);
        launch_accumulate_f2d(p32, psi64, npsi, s);
    }
    TSW_HIP(hipMemcpyAsync(Q, psi64, sizeof(double) * npsi, hipMemcpyDeviceToDevice, s));
}

// Launch shape of reference-layout sweeps: staged (coalesced) source loads with the storer wave, else
// with stores from the compute lanes (material path only), when the stages fit every member's LDS beside
// its tables; otherwise `ps` keep their shape (direct) unstaged.  Returns whether the sources are staged;
// lds and direct receive the launch's values.
template <typename T>
bool ref_stage_layout(std::vector<DiagSweepParams<T>> &ps, int p, bool mat, size_t &lds, bool &direct) {
    for (int d = 0; d <= (mat ? 1 : 0); ++d) {
        size_t need = 0;
        bool fits = true;
        for (auto &sp : ps) {
            DiagSweepParams<T> t = sp;
            t.direct = d;
            t.ref_stage = 1;
            const size_t b = diag_sweep_lds_bytes<T>(t, p);
            fits = fits && b != 0;
            need = std::max(need, b);
        }
        if (fits) {
            for (auto &sp : ps) {
                sp.direct = d;
                sp.ref_stage = 1;
            }
            lds = need;
            direct = d != 0;
            return true;
        }
    }
    for (auto &sp : ps) sp.ref_stage = 0;
    return false;
}

// Launch shape of the staggered kernel on the caller's layout for this problem (one band per ordinate,
// ny <= 448): stores from the compute lanes when the material table fits beside them, else the storer
// wave, else per-cell elimination; staged source loads with "ref_stage".  The material table, element
// inverses and per-cell arrays are set up here.
template <typename T>
DiagSweepParams<T> ref_plan(tsw_ctx *c, const double *sig, double dt, size_t &lds) {
    const Geom &g = c->g;
    const bool f64 = sizeof(T) == 8;
    diag_setup(c, sig, nullptr, nullptr, dt, f64, !f64);
    DiagSweepParams<T> base = make_diag_params<T>(c, dt, nullptr, nullptr, false, nullptr);
    base.nbands = 1;
    base.band_rows = g.ny;
    base.xtr = nullptr;
    base.wave1 = 0;
    base.split = 0;
    lds = 0;
    if (base.ainv) {
        base.direct = 1;
        lds = diag_sweep_lds_bytes<T>(base, g.p);
        if (lds == 0) {
            base.direct = 0;
            lds = diag_sweep_lds_bytes<T>(base, g.p);
        }
    }
    if (lds == 0) {  // per-cell elimination
        base.ainv = nullptr;
        base.mat = nullptr;
        base.M = 0;
        base.direct = 0;
        base.sig_t = (const T *)(f64 ? c->diag.sigt64.ptr : c->diag.sigt32.ptr);
        lds = diag_sweep_lds_bytes<T>(base, g.p);
    }
    require(lds != 0, TSW_E_UNSUPPORTED, "the mesh does not fit the staggered sweep");
    if (c->ref_stage_opt > 0) {
        std::vector<DiagSweepParams<T>> one{base};
        one[0].qref = reinterpret_cast<const double *>(1);
        bool direct = base.direct != 0;
        ref_stage_layout<T>(one, g.p, base.ainv != nullptr, lds, direct);
        base.direct = one[0].direct;
        base.ref_stage = one[0].ref_stage;
    }
    return base;
}

// One launch sweeping ordinates [q_lo, q_hi) of each array arr[r] in place with `base` (ref_plan): items
// (source, ordinate), the ordinates of one source side by side (they share cache lines); the inflow
// traces of every ordinate come first from the arrays' current halos.
template <typename T>
void ref_launch(tsw_ctx *c, const DiagSweepParams<T> &base, size_t lds, const std::vector<double *> &arr, int q_lo,
                int q_hi) {
    const Geom &g = c->g;
    const int n = (int)arr.size(), nqs = q_hi - q_lo;
    const size_t ntx = (size_t)g.nq * g.ny * (g.p + 1), nty = (size_t)g.nq * g.nx * (g.p + 1);
    T *tr = static_cast<T *>(c->multi.traces.reserve(sizeof(T) * (ntx + nty) * n));
    auto &M = c->multi;
    M.host.resize(sizeof(DiagSweepParams<T>) * n);
    DiagSweepParams<T> *ps = reinterpret_cast<DiagSweepParams<T> *>(M.host.data());
    for (int r = 0; r < n; ++r) {
        ps[r] = base;
        ps[r].qref = arr[r];
        ps[r].psi_ref = arr[r];
        ps[r].tin_x = tr + (ntx + nty) * r;
        ps[r].tin_y = tr + (ntx + nty) * r + ntx;
    }
    M.items.resize((size_t)2 * n * nqs);
    for (int r = 0; r < n; ++r)
        for (int k = 0; k < nqs; ++k) {
            M.items[2 * ((size_t)r * nqs + k)] = r;
            M.items[2 * ((size_t)r * nqs + k) + 1] = q_lo + k;
        }
    hipStream_t s = c->stream;
    TSW_HIP(hipMemcpyAsync(M.params.reserve(M.host.size()), M.host.data(), M.host.size(), hipMemcpyHostToDevice, s));
    TSW_HIP(hipMemcpyAsync(M.dev_items.reserve(sizeof(int) * M.items.size()), M.items.data(),
                           sizeof(int) * M.items.size(), hipMemcpyHostToDevice, s));
    DiagSweepParams<T> rep{};
    rep.qref = reinterpret_cast<const double *>(1);
    rep.ainv = base.ainv;
    rep.direct = base.direct;
    rep.ref_stage = base.ref_stage;
    rep.prefetch = 0;
    const DiagSweepParams<T> *dp = M.params.as<DiagSweepParams<T>>();
    launch_diag_inflow_batch<T>(dp, n, g.p, s);
    launch_diag_sweep_batch<T>(rep, dp, M.dev_items.as<int2>(), n * nqs, g.ny, lds, g.p, s);
    TSW_HIP(hipStreamSynchronize(s));  // the host tables may be rebuilt by the next launch
}

// Sweeps of n right-hand sides arr[r] of this problem (reference layout, in place, 16-byte aligned) in
// one launch of the staggered kernel on the caller's layout, the material table and element inverses
// shared by every source; counted as n sweeps.
template <typename T>
bool try_split_ref(tsw_ctx *c, const std::vector<double *> &arr, const double *sig, double dt);
template <typename T>
void multi_sweep(tsw_ctx *c, const std::vector<double *> &arr, const double *sig, double dt) {
    if (try_split_ref<T>(c, arr, sig, dt)) return;
    size_t lds = 0;
    const DiagSweepParams<T> base = ref_plan<T>(c, sig, dt, lds);
    hipEvent_t e0 = next_event(c), e1 = next_event(c);
    TSW_HIP(hipEventRecord(e0, c->stream));
    ref_launch<T>(c, base, lds, arr, 0, c->g.nq);
    TSW_HIP(hipEventRecord(e1, c->stream));
    c->sweeps_done += (long long)arr.size();
}

// Lanes per mesh row of the lane-split kernel for DG degree p (split_lanes<p + 1>).
int split_lanes_of(int p) { return p == 1 ? 2 : 4; }

// Rows per band of a problem (a batch member, or a single problem) on the lane-split kernel: the fewest
// bands of at most R rows (R: the tuning key batch_split_rows, else the batch autotuner's choice for the
// degree, else the rows one workgroup of 512 threads holds), split evenly.
int split_band_rows(const tsw_ctx *ctx, int p, int ny) {
    int R = 512 / split_lanes_of(p);
    const int want = ctx->batch_split_rows_opt > 0 ? ctx->batch_split_rows_opt : ctx->split_rows_p[p & 3];
    if (want > 0) R = std::min(R, want);
    const int nb = (ny + R - 1) / R;
    return (ny + nb - 1) / nb;
}

// Whether this context's sweeps can run the staggered kernel on the caller's layout for `psi`.
bool ref_sweep_ok(tsw_ctx *c, const double *psi) {
    return staggered_ok(c) && c->pattern == Pattern::Staggered && c->xs_deg == 0 &&
           c->g.npsi() * sizeof(double) < (size_t(1) << 32) && c->g.ny <= 448 && ((uintptr_t)psi & 15) == 0;
}

// Plan of a lane-split sweep on the caller's layout (the batch engine's kernel for one problem): false
// when it is not built for the degree, the element-inverse tables are missing or the LDS does not fit.
// Sets up the material table and inverses of sig for dt; base has the bands of split_band_rows rows.
template <typename T>
bool split_ref_plan(tsw_ctx *c, const double *sig, double dt, DiagSweepParams<T> &base, size_t &lds) {
    const Geom &g = c->g;
    if (c->ref_split_opt < 0 || g.p < 1) return false;
    // a set band_rows, prefetch, one_wave or lane_split asks for the diagonal-layout kernels
    if (c->band_rows_opt != 0 || c->prefetch_opt != 0 || c->one_wave_opt != 0 || c->split_opt != 0) return false;
    const bool f64 = sizeof(T) == 8;
    diag_setup(c, sig, nullptr, nullptr, dt, f64, !f64);
    base = make_diag_params<T>(c, dt, nullptr, nullptr, false, nullptr);
    if (!base.ainv) return false;
    const int br = split_band_rows(c, g.p, g.ny);
    lds = split_sweep_lds_bytes<T>(g.p, br, g.nx, base.M);
    if (lds == 0) return false;
    base.split = 1;
    base.direct = 0;
    base.wave1 = 0;
    base.ref_stage = 0;
    base.nbands = (g.ny + br - 1) / br;
    base.band_rows = base.nbands > 1 ? br : g.ny;
    base.xtr = nullptr;
    base.abort_flag = nullptr;
    return true;
}

// One persistent launch of the lane-split kernel sweeping every ordinate of each array arr[r] in place
// (reference layout, 16-byte aligned), with base from split_ref_plan: items {source, band * nq + ordinate},
// band-major; per-source inflow traces from the arrays' halos and per-source exchange records.  Counted
// as arr.size() sweeps.
template <typename T>
void split_ref_sweep(tsw_ctx *c, const DiagSweepParams<T> &base, size_t lds, const std::vector<double *> &arr) {
    const Geom &g = c->g;
    const int n = (int)arr.size(), nb = base.nbands;
    const size_t ntx = (size_t)g.nq * g.ny * (g.p + 1), nty = (size_t)g.nq * g.nx * (g.p + 1);
    T *tr = static_cast<T *>(c->multi.traces.reserve(sizeof(T) * (ntx + nty) * n));
    auto &M = c->multi;
    hipStream_t s = c->stream;
    const size_t per = nb > 1 ? (size_t)g.nq * (nb - 1) * g.nx * (g.p + 1) : 0;  // records of one source
    T *xtr = nullptr;
    if (per) {
        DBuf &xb = sizeof(T) == 8 ? M.xtr64 : M.xtr32;
        size_t &have = sizeof(T) == 8 ? M.xtr64_n : M.xtr32_n;
        if (have < per * n) {  // every record empty (the sentinel); a sweep leaves them empty again
            xb.reserve(sizeof(T) * per * n);
            have = per * n;
            launch_exchange_fill<T>(xb.as<T>(), have, s);
        }
        xtr = xb.as<T>();
    }
    M.host.resize(sizeof(DiagSweepParams<T>) * n);
    DiagSweepParams<T> *ps = reinterpret_cast<DiagSweepParams<T> *>(M.host.data());
    for (int r = 0; r < n; ++r) {
        ps[r] = base;
        ps[r].qref = arr[r];
        ps[r].psi_ref = arr[r];
        ps[r].tin_x = tr + (ntx + nty) * r;
        ps[r].tin_y = tr + (ntx + nty) * r + ntx;
        if (per) {
            ps[r].xtr = xtr + per * r;
            ps[r].abort_flag = abort_flag_of(c);
        }
    }
    M.items.clear();
    for (int band = 0; band < nb; ++band)
        for (int r = 0; r < n; ++r)
            for (int q = 0; q < g.nq; ++q) {
                M.items.push_back(r);
                M.items.push_back(band * g.nq + q);
            }
    TSW_HIP(hipMemcpyAsync(M.params.reserve(M.host.size()), M.host.data(), M.host.size(), hipMemcpyHostToDevice, s));
    TSW_HIP(hipMemcpyAsync(M.dev_items.reserve(sizeof(int) * M.items.size()), M.items.data(),
                           sizeof(int) * M.items.size(), hipMemcpyHostToDevice, s));
    DiagSweepParams<T> rep{};
    rep.qref = reinterpret_cast<const double *>(1);
    rep.ainv = base.ainv;
    rep.split = 1;
    rep.prefetch = c->batch_split_pf_opt > 0 ? c->batch_split_pf_opt : kSplitBatchPrefetch;
    if (c->batch_persistent_opt > 0) {
        rep.work = static_cast<int *>(M.work.reserve(sizeof(int)));
        TSW_HIP(hipMemsetAsync(rep.work, 0, sizeof(int), s));
    }
    const DiagSweepParams<T> *dp = M.params.as<DiagSweepParams<T>>();
    if (c->stall_test && per) {  // tuning key stall_test: every bounded wait of the banded sweep gives up
        c->stall_test = false;
        int *flag = abort_flag_of(c);
        *c->abort_host = 1;
        TSW_HIP(hipMemcpyAsync(flag, c->abort_host, sizeof(int), hipMemcpyHostToDevice, s));
    }
    hipEvent_t e0 = next_event(c), e1 = next_event(c);
    TSW_HIP(hipEventRecord(e0, s));
    launch_diag_inflow_batch<T>(dp, n, g.p, s);
    launch_diag_sweep_batch<T>(rep, dp, M.dev_items.as<int2>(), (int)M.items.size() / 2, base.band_rows, lds, g.p, s);
    TSW_HIP(hipEventRecord(e1, s));
    c->sweeps_done += n;
    TSW_HIP(hipStreamSynchronize(s));  // the host tables may be rebuilt by the next launch
}

// The lane-split sweep on the caller's layout of arrays arr (16-byte aligned, ref_sweep_ok) when it is
// built for this problem, in precision T; false (nothing done) otherwise.
template <typename T>
bool try_split_ref(tsw_ctx *c, const std::vector<double *> &arr, const double *sig, double dt) {
    for (double *a : arr)
        if (!ref_sweep_ok(c, a)) return false;
    DiagSweepParams<T> base;
    size_t lds = 0;
    if (!split_ref_plan<T>(c, sig, dt, base, lds)) return false;
    split_ref_sweep<T>(c, base, lds, arr);
    c->last_sl = 1;
    return true;
}

// SweepOperator::Apply with reflecting edges (SweepPatternUniCartWavefront.cpp:180-245): the quadrants
// one after the other in the order 3, 2, 1, 0 (SweepPatternUniCart.cpp:33-40), in place, and
// OrdinateFlux::ReflectBoundaries on the result after each quadrant, so a quadrant takes its reflected
// inflow from whatever the mirrored ordinates hold at that moment, as the reference does.  One sweep.
template <typename T>
void reflect_sweep(tsw_ctx *c, double *psi, const double *sig, double dt) {
    const Geom &g = c->g;
    require(ref_sweep_ok(c, psi), TSW_E_UNSUPPORTED,
            "reflecting boundaries need the hip-staggered pattern, cell-mean cross sections, ny <= 448 and a "
            "16-byte aligned psi");
    size_t lds = 0;
    const DiagSweepParams<T> base = ref_plan<T>(c, sig, dt, lds);
    const std::vector<double *> one{psi};
    hipEvent_t e0 = next_event(c), e1 = next_event(c);
    TSW_HIP(hipEventRecord(e0, c->stream));
    for (int quad : {3, 2, 1, 0}) {
        if (g.quad_off[quad + 1] > g.quad_off[quad]) ref_launch<T>(c, base, lds, one, g.quad_off[quad], g.quad_off[quad + 1]);
        launch_reflect_boundaries(g, psi, c->reflect_map_d.as<int>(), c->reflect_mask, c->stream);
    }
    TSW_HIP(hipEventRecord(e1, c->stream));
    ++c->sweeps_done;
}

void sweep_device(tsw_ctx *ctx, double *psi, const double *sig, double dt, int precision) {
    require(precision == TSW_FP64 || precision == TSW_FP32 || precision == TSW_MIXED, TSW_E_ARG, "bad precision");
    ctx->last_lanes = 0;
    if (ctx->reflect_mask) {  // TSW_MIXED: the FP64 sweep (this engine moves the same FP64 arrays in FP32)
        if (precision == TSW_FP32) reflect_sweep<float>(ctx, psi, sig, dt);
        else reflect_sweep<double>(ctx, psi, sig, dt);
        return;
    }
    if (lanes_ok(ctx) && precision != TSW_MIXED) {
        if (precision == TSW_FP64) apply_lanes<double>(ctx, psi, sig, dt);
        else apply_lanes<float>(ctx, psi, sig, dt);
    } else if (precision == TSW_FP32 ? try_split_ref<float>(ctx, {psi}, sig, dt)
                                     : try_split_ref<double>(ctx, {psi}, sig, dt)) {
        // in place on the caller's layout; TSW_MIXED: the FP64 sweep (the same FP64 arrays either way)
    } else if (staggered_ok(ctx)) {
        if (precision == TSW_FP64) apply_staggered<double>(ctx, psi, sig, dt);
        else if (precision == TSW_FP32) apply_staggered<float>(ctx, psi, sig, dt);
        else apply_staggered_mixed(ctx, psi, sig, dt);
    } else if (precision == TSW_MIXED) {
        sweep_wavefront_mixed(ctx, psi, sig, dt);
    } else {
        sweep_prec(ctx, psi, sig, dt, precision);
    }
}

// ---------------------------------------------------------------------------------------------------
// IDC time step (IDCIntegrator::Step_Nonhybrid, idc.cpp:846-940) on device arrays.
// ---------------------------------------------------------------------------------------------------
// Legendre polynomial P_n and its derivative at x
void legendre_pd(int n, double x, double &p, double &dp) {
    double p0 = 1.0, p1 = x;
    if (n == 0) {
        p = 1.0;
        dp = 0.0;
        return;
    }
    for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    p = p1;
    dp = std::fabs(1.0 - x * x) < 1e-300 ? 0.5 * n * (n + 1) * (x > 0 ? 1.0 : (n % 2 ? 1.0 : -1.0))
                                         : n * (p0 - x * p1) / (1.0 - x * x);
}

// Gauss-Legendre nodes (ascending) and weights on [-1, 1], Newton from Chebyshev guesses
void gauss_legendre(int m, std::vector<double> &x, std::vector<double> &w) {
    x.assign(m, 0.0);
    w.assign(m, 0.0);
    for (int k = 0; k < m; ++k) {
        double t = -std::cos(M_PI * (k + 0.75) / (m + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p, dp;
            legendre_pd(m, t, p, dp);
            const double dt = p / dp;
            t -= dt;
            if (std::fabs(dt) < 1e-16) break;
        }
        double p, dp;
        legendre_pd(m, t, p, dp);
        x[k] = t;
        w[k] = 2.0 / ((1.0 - t * t) * dp * dp);
    }
}

Tpi tpi_table(int p) {
    const int P = p + 1, m = std::max(2 * p, 2);
    std::vector<double> x, w;
    gauss_legendre(m, x, w);
    Tpi t{};
    for (int a = 0; a < P; ++a)
        for (int b = 0; b < P; ++b)
            for (int c = 0; c < P; ++c) {
                double v = 0.0;
                for (int i = 0; i < m; ++i) {
                    double pa, pb, pc, dp;
                    legendre_pd(a, x[i], pa, dp);
                    legendre_pd(b, x[i], pb, dp);
                    legendre_pd(c, x[i], pc, dp);
                    v += w[i] * pa * pb * pc;
                }
                t.v[a + P * (b + P * c)] = std::fabs(v) < 5e-16 ? 0.0 : v;
            }
    return t;
}

// Nodes and integration matrix (SetNodes, idc.cpp:2305-2376), from tsw_idc_nodes
void idc_tables(tsw_ctx::Idc &I) {
    int S = 0;
    require(tsw_idc_nodes(I.kind, I.num_nodes, 0, &S, nullptr, nullptr) == TSW_OK, TSW_E_ARG, "bad IDC nodes");
    I.nodes.assign(S, 0.0);
    I.W.assign((size_t)S * S, 0.0);
    require(tsw_idc_nodes(I.kind, I.num_nodes, S, &S, I.nodes.data(), I.W.data()) == TSW_OK, TSW_E_ARG,
            "bad IDC nodes");
    I.left = I.kind == TSW_IDC_LOBATTO;
    I.right = I.kind == TSW_IDC_LOBATTO || I.kind == TSW_IDC_RADAU || I.kind == TSW_IDC_CHEBYSHEV_RADAU;
}

// R = L x - S P x on the interior (ComputeResiduals, idc.cpp:1658-1667); R's halos stay zero
void idc_residual(tsw_ctx *c, const double *x, double *R, const double *st, const double *ss, double *phi) {
    const Geom &g = c->g;
    hipStream_t s = c->stream;
    launch_lmv<double>(g, c->xi64.as<double>(), c->eta64.as<double>(), st, x, R, 1.0, 0.0, s, xs_tensor<double>(c, st));
    launch_pmv<double, double>(g, x, phi, c->w64.as<double>(), 1.0, 0.0, 0, s);
    launch_smv<double>(g, ss, phi, R, -1.0, 1.0, 0, s, xs_tensor<double>(c, ss));
}

// implicit transport solve (1/dt_sub + L - S P) psi = rhs in place (rhs halos = inflow)
bool idc_solve(tsw_ctx *c, double *rhs, const double *st, const double *ss, double dt_sub, double tol, int max_its,
               int precision, int &its_total) {
    int its = 0;
    double res = 0.0;
    const bool ok = gmres_device(c, rhs, st, ss, dt_sub, tol, max_its, 30, precision, nullptr, its, res);
    its_total += its;
    return ok;
}

// One IDC step in place on psi (device, zero halos on entry and exit); src: Q with inflow halos.
bool idc_step_device(tsw_ctx *c, double *psi, const double *src, const double *st, const double *ss, double dt,
                     double tol, int max_its, int precision, int &its_total) {
    const Geom &g = c->g;
    hipStream_t s = c->stream;
    auto &I = c->idc;
    const int S = (int)I.nodes.size(), l0 = I.left ? 0 : 1;
    const size_t n = g.npsi(), K = (size_t)g.nq * g.N();
    double *stg = dbuf<double>(I.stages, (size_t)S * n);
    double *R = dbuf<double>(I.res, (size_t)S * n);
    double *e = dbuf<double>(I.err, n);
    double *phi = dbuf<double>(I.phi, g.nphi());
    TSW_HIP(hipMemsetAsync(R, 0, sizeof(double) * S * n, s));  // residual halos stay zero
    TSW_HIP(hipMemsetAsync(phi, 0, sizeof(double) * g.nphi(), s));
    auto stage = [&](int k) { return stg + (size_t)k * n; };
    bool ok = true;
    its_total = 0;
    // stage 0: the initial condition, ghost cells zero (idc.cpp:833-835, 850)
    TSW_HIP(hipMemcpyAsync(stage(0), psi, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    launch_fill_region<double>(g, stage(0), K, 1, 0.0, s);
    // prediction: implicit Euler substeps (:863-876)
    // (the sources below are each one pass over the arrays they combine, launch_lincomb)
    for (int k = 1; k < S; ++k) {
        const double h = I.nodes[k] - I.nodes[k - 1];
        LinComb L;  // stage k = stage(k-1) / (h dt) on the interior + Q (with its inflow halos)
        L.add(1.0 / (h * dt), stage(k - 1));
        L.add(1.0, src, true);
        launch_lincomb(g, K, L, stage(k), s);
        ok = idc_solve(c, stage(k), st, ss, h * dt, tol, max_its, precision, its_total) && ok;
    }
    // corrections (:882-930, ComputeSource :1686-1740)
    for (int p = 0; p < I.num_corr; ++p) {
        for (int l = l0; l < S; ++l) idc_residual(c, stage(l), R + (size_t)l * n, st, ss, phi);
        for (int k = 1; k < S; ++k) {
            const double h = I.nodes[k] - I.nodes[k - 1];
            LinComb L;
            if (I.form == TSW_IDC_ERROR) {
                // e = -stage(k)/(h dt) - sum_l W(k,l) R_l / h + stage(k-1)/(h dt) + Q, zero halos; stage(k) += e
                L.add(-1.0 / (h * dt), stage(k));
                for (int l = l0; l < S; ++l) L.add(-I.W[(size_t)k * S + l] / h, R + (size_t)l * n);
                L.add(1.0 / (h * dt), stage(k - 1));
                L.add(1.0, src);
                launch_lincomb(g, K, L, e, s);
                ok = idc_solve(c, e, st, ss, h * dt, tol, max_its, precision, its_total) && ok;
                launch_axpby(n, 1.0, e, 1.0, stage(k), s);
            } else {
                // stage(k) = sum_l (delta_kl - W(k,l)/h) R_l + stage(k-1)/(h dt) on the interior, + Q with
                // its inflow halos, then solved in place (:1711-1730, :909-911)
                for (int l = l0; l < S; ++l)
                    L.add((l == k ? 1.0 : 0.0) - I.W[(size_t)k * S + l] / h, R + (size_t)l * n);
                L.add(1.0 / (h * dt), stage(k - 1));
                L.add(1.0, src, true);
                launch_lincomb(g, K, L, stage(k), s);
                ok = idc_solve(c, stage(k), st, ss, h * dt, tol, max_its, precision, its_total) && ok;
            }
        }
    }
    // result: the last node, or the collocation formula (:933-939, ComputeCollocation_Nonhybrid :1541-1566)
    if (!I.right || I.collocation) {
        // psi = psi0 + dt Q + dt sum_l W(0,l) (S P - L) psi_l (its halos are zeroed below)
        for (int l = l0; l < S; ++l) idc_residual(c, stage(l), R + (size_t)l * n, st, ss, phi);
        LinComb L;
        L.add(1.0, stage(0));
        L.add(dt, src);
        for (int l = l0; l < S; ++l) L.add(-dt * I.W[l], R + (size_t)l * n);
        launch_lincomb(g, K, L, psi, s);
    } else {
        TSW_HIP(hipMemcpyAsync(psi, stage(S - 1), sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    }
    launch_fill_region<double>(g, psi, K, 1, 0.0, s);
    return ok;
}

// Validates and installs a mesh + ordinate set (tsw_problem_set, also used per batch member).
void set_problem(tsw_ctx *ctx, int nx, int ny, double ax, double bx, double ay, double by, int p, int nq,
                 const double *xi, const double *eta, const double *w, const int *quad_off) {
    require(nx >= 1 && ny >= 1, TSW_E_ARG, "nx and ny must be >= 1");
    require(bx > ax && by > ay, TSW_E_ARG, "empty domain");
    require(p >= 0, TSW_E_ARG, "DG degree must be >= 0");
    require(p <= 3, TSW_E_UNSUPPORTED, "DG degree " + std::to_string(p) + " not built (0..3)");
    require(nq >= 1 && xi && eta && w && quad_off, TSW_E_ARG, "bad ordinate set");
    require(quad_off[0] == 0 && quad_off[4] == nq, TSW_E_ARG, "quad_off must span [0, nq]");
    for (int k = 0; k < 4; ++k) {
        require(quad_off[k] <= quad_off[k + 1], TSW_E_ARG, "quad_off must be non-decreasing");
        for (int q = quad_off[k]; q < quad_off[k + 1]; ++q) {
            const bool xneg = (k >= 2), yneg = (k & 1);
            require(xi[q] != 0.0 && eta[q] != 0.0 && (xi[q] < 0) == xneg && (eta[q] < 0) == yneg, TSW_E_ARG,
                    "ordinate " + std::to_string(q) + " is not in quadrant " + std::to_string(k));
        }
    }
    ctx->clear_graphs();
    ctx->adj.reset();
    ctx->tuned.clear();
    if (ctx->dsa.graph) (void)hipGraphExecDestroy(ctx->dsa.graph);
    ctx->dsa.graph = nullptr;
    ctx->dsa.keyed = false;
    ctx->mat.valid = ctx->mat.keyed = ctx->mat.have64 = ctx->mat.have32 = false;
    ++ctx->mat.build;
    ctx->lanes.J = ctx->lanes.nblk = 0;
    ctx->xs_deg = 0;
    ctx->xs.key[0] = ctx->xs.key[1] = nullptr;
    Geom &g = ctx->g;
    g.nx = nx; g.ny = ny; g.p = p; g.nq = nq;
    g.ax = ax; g.bx = bx; g.ay = ay; g.by = by;
    g.dx = (bx - ax) / nx;
    g.dy = (by - ay) / ny;
    for (int k = 0; k < 5; ++k) g.quad_off[k] = quad_off[k];
    ctx->h_xi.assign(xi, xi + nq);
    ctx->h_eta.assign(eta, eta + nq);
    ctx->h_w.assign(w, w + nq);
    std::vector<float> f(nq);
    auto put64 = [&](DBuf &b, const double *src) {
        b.reserve(sizeof(double) * nq);
        TSW_HIP(hipMemcpy(b.ptr, src, sizeof(double) * nq, hipMemcpyHostToDevice));
    };
    auto put32 = [&](DBuf &b, const double *src) {
        for (int q = 0; q < nq; ++q) f[q] = (float)src[q];
        b.reserve(sizeof(float) * nq);
        TSW_HIP(hipMemcpy(b.ptr, f.data(), sizeof(float) * nq, hipMemcpyHostToDevice));
    };
    put64(ctx->xi64, xi); put64(ctx->eta64, eta); put64(ctx->w64, w);
    put32(ctx->xi32, xi); put32(ctx->eta32, eta); put32(ctx->w32, w);
    double *st = static_cast<double *>(ctx->si_state.reserve(sizeof(double) * g.npsi()));
    TSW_HIP(hipMemset(st, 0, sizeof(double) * g.npsi()));
    ctx->reflect_mask = 0;
    ctx->periodic = false;
    ctx->have_problem = true;
}

// The adjoint problem of ctx (built on first use): same mesh, ordinates -Omega in reverse order with
// their weights, so quadrant k of ctx becomes quadrant 3-k and the blocks stay ordered.
tsw_ctx *adjoint_ctx(tsw_ctx *ctx) {
    if (!ctx->adj) {
        const Geom &g = ctx->g;
        const int nq = g.nq;
        std::vector<double> xi(nq), eta(nq), w(nq);
        for (int q = 0; q < nq; ++q) {
            xi[q] = -ctx->h_xi[nq - 1 - q];
            eta[q] = -ctx->h_eta[nq - 1 - q];
            w[q] = ctx->h_w[nq - 1 - q];
        }
        int off[5];
        for (int k = 0; k < 5; ++k) off[k] = nq - g.quad_off[4 - k];
        auto a = std::make_unique<tsw_ctx>();
        a->device = ctx->device;
        a->stream = ctx->stream;
        a->own_stream = false;
        a->pattern = ctx->pattern;
        a->band_rows_opt = ctx->band_rows_opt;
        a->prefetch_opt = ctx->prefetch_opt;
        a->autotune = ctx->autotune;
        a->use_dsa = ctx->use_dsa;
        a->dsa_cycles = ctx->dsa_cycles;
        set_problem(a.get(), g.nx, g.ny, g.ax, g.bx, g.ay, g.by, g.p, nq, xi.data(), eta.data(), w.data(), off);
        ctx->adj = std::move(a);
    }
    ctx->adj->pattern = ctx->pattern;
    ctx->adj->band_rows_opt = ctx->band_rows_opt;
    ctx->adj->prefetch_opt = ctx->prefetch_opt;
    ctx->adj->autotune = ctx->autotune;
    ctx->adj->use_dsa = ctx->use_dsa;
    ctx->adj->dsa_cycles = ctx->dsa_cycles;
    ctx->adj->gmres_refine = ctx->gmres_refine;
    ctx->adj->one_wave_opt = ctx->one_wave_opt;
    ctx->adj->split_opt = ctx->split_opt;
    ctx->adj->refine_steps = ctx->refine_steps;
    return ctx->adj.get();
}

// Gradients of J = <seed, psi> once the adjoint flux zeta sits in gsrc's interior: cross-section
// gradients per cell (gst/gss may be null), then the inflow gradient added into gsrc's halos (which
// hold seed's halos: the forward map passes them through).
void vjp_finish(tsw_ctx *ctx, const double *psi, const double *phi, double *gsrc, double *gst, double *gss) {
    const Geom &g = ctx->g;
    hipStream_t s = ctx->stream;
    if (gst) TSW_HIP(hipMemsetAsync(gst, 0, sizeof(double) * g.ncell_h(), s));
    if (gss) TSW_HIP(hipMemsetAsync(gss, 0, sizeof(double) * g.ncell_h(), s));
    if (gst || gss) launch_cell_grads(g, psi, gsrc, phi, gst, gss, s);
    launch_inflow_grad(g, gsrc, gsrc, ctx->xi64.as<double>(), ctx->eta64.as<double>(), s);
}

// Vector-Jacobian product of the in-place sweep (device arrays).  zeta = L^{-T} seed is computed as
// W^{-1} R sweep_adj(R W seed): R reverses the ordinate order, W scales coefficient (d,e) by (2d+1)(2e+1),
// sweep_adj is the adjoint problem's forward sweep with vacuum inflow.
void sweep_vjp_device(tsw_ctx *ctx, const double *psi, const double *st, double dt, const double *seed, double *gsrc,
                      double *gst, int precision) {
    const Geom &g = ctx->g;
    hipStream_t s = ctx->stream;
    tsw_ctx *a = adjoint_ctx(ctx);
    double *y = dbuf<double>(ctx->vjp_y, g.npsi());
    TSW_HIP(hipMemsetAsync(y, 0, sizeof(double) * g.npsi(), s));
    launch_reverse_ordinates(g, seed, y, 1, nullptr, 0, false, s);
    sweep_device(a, y, st, dt, precision);
    if (gsrc != seed) TSW_HIP(hipMemcpyAsync(gsrc, seed, sizeof(double) * g.npsi(), hipMemcpyDeviceToDevice, s));
    launch_reverse_ordinates(g, y, gsrc, 2, nullptr, 0, false, s);
    vjp_finish(ctx, psi, nullptr, gsrc, gst, nullptr);
}

// Vector-Jacobian product of the scattering solve psi = (L - K)^{-1} Q, K psi = sigma_s P psi / (4 pi):
// zeta = L^{-T}(seed + K^T zeta), (K^T zeta)_q = w_q sigma_s sum_q' zeta_q' / (4 pi).  With chi_q = zeta_q / w_q
// and chi' = R W chi this is the adjoint problem's own scattering solve with source R W (seed_q / w_q), run
// by the same Peierls GMRES.  Returns convergence.
bool solve_vjp_device(tsw_ctx *ctx, const double *psi, const double *st, const double *ss, double dt,
                      const double *seed, double *gsrc, double *gst, double *gss, double abs_tol, int max_its,
                      int restart, int precision, int &its, double &res) {
    const Geom &g = ctx->g;
    hipStream_t s = ctx->stream;
    tsw_ctx *a = adjoint_ctx(ctx);
    const double *w = ctx->w64.as<double>();
    double *y = dbuf<double>(ctx->vjp_y, g.npsi());
    TSW_HIP(hipMemsetAsync(y, 0, sizeof(double) * g.npsi(), s));
    launch_reverse_ordinates(g, seed, y, 1, w, 2, false, s);
    const bool ok = gmres_device(a, y, st, ss, dt, abs_tol, max_its, restart, precision, nullptr, its, res);
    if (gsrc != seed) TSW_HIP(hipMemcpyAsync(gsrc, seed, sizeof(double) * g.npsi(), hipMemcpyDeviceToDevice, s));
    launch_reverse_ordinates(g, y, gsrc, 2, w, 1, true, s);
    double *phi = nullptr;
    if (gss) {
        phi = dbuf<double>(ctx->vjp_phi, g.nphi());
        TSW_HIP(hipMemsetAsync(phi, 0, sizeof(double) * g.nphi(), s));
        launch_pmv<double, double>(g, psi, phi, w, 1.0, 0.0, 0, s);
    }
    vjp_finish(ctx, psi, phi, gsrc, gst, gss);
    return ok;
}


// zeta (interior) = A^{-T} seed (interior), A = 1/dt + L - S P, by the adjoint problem's Peierls GMRES
// (the transposed solve of solve_vjp_device); zeta's halos are not written.
bool adjoint_solve(tsw_ctx *c, const double *seed, double *zeta, const double *st, const double *ss, double dt,
                   double tol, int max_its, int precision, int &its_total) {
    const Geom &g = c->g;
    hipStream_t s = c->stream;
    tsw_ctx *a = adjoint_ctx(c);
    const double *w = c->w64.as<double>();
    double *y = dbuf<double>(c->vjp_y, g.npsi());
    TSW_HIP(hipMemsetAsync(y, 0, sizeof(double) * g.npsi(), s));
    launch_reverse_ordinates(g, seed, y, 1, w, 2, false, s);
    int its = 0;
    double res = 0.0;
    const bool ok = gmres_device(a, y, st, ss, dt, tol, max_its, 30, precision, nullptr, its, res);
    its_total += its;
    launch_reverse_ordinates(g, y, zeta, 2, w, 1, true, s);
    return ok;
}

// out (interior) += alpha B^T y, B = L - S P (y's halos ignored): L^T = W^{-1} L(-Omega) W with vacuum
// inflow, evaluated with the adjoint problem's ordinates; (S P)^T by launch_kt_apply.
void bt_apply(tsw_ctx *c, const double *y, double *out, const double *st, const double *ss, double alpha) {
    const Geom &g = c->g;
    hipStream_t s = c->stream;
    tsw_ctx *a = adjoint_ctx(c);
    auto &I = c->idc;
    const size_t n = g.npsi();
    double *t1 = dbuf<double>(I.t1, n), *t2 = dbuf<double>(I.t2, n), *t3 = dbuf<double>(I.t3, n);
    TSW_HIP(hipMemsetAsync(t1, 0, sizeof(double) * n, s));
    TSW_HIP(hipMemsetAsync(t3, 0, sizeof(double) * n, s));
    launch_reverse_ordinates(g, y, t1, 1, nullptr, 0, false, s);
    launch_lmv<double>(g, a->xi64.as<double>(), a->eta64.as<double>(), st, t1, t2, 1.0, 0.0, s);
    launch_reverse_ordinates(g, t2, t3, 2, nullptr, 0, false, s);
    launch_axpby(n, alpha, t3, 1.0, out, s);
    launch_kt_apply(g, y, out, ss, c->w64.as<double>(), -alpha, s);
}

// scalar flux of x (interior, zero halos) into c->idc.phi
const double *idc_phi(tsw_ctx *c, const double *x) {
    const Geom &g = c->g;
    double *phi = dbuf<double>(c->idc.phi, g.nphi());
    TSW_HIP(hipMemsetAsync(phi, 0, sizeof(double) * g.nphi(), c->stream));
    launch_pmv<double, double>(g, x, phi, c->w64.as<double>(), 1.0, 0.0, 0, c->stream);
    return phi;
}

// Vector-Jacobian product of one IDC step (linear in psi^n for fixed cross sections): lam holds
// d J / d psi^{n+1} (interior) on entry and d J / d psi^n (interior, zero halos) on exit.  With gst / gss
// (cell arrays, accumulated; may be null) the step is first re-run from psi_n keeping its intermediates,
// and every implicit solve y = A^{-1} r adds <z, dA/dsigma y> terms (z its adjoint) and every residual
// R = B x adds <lam_R, dB/dsigma x> terms.  The reverse pass mirrors idc_step_device operation by operation.
bool idc_step_vjp_device(tsw_ctx *c, const double *psi_n, const double *src, const double *st, const double *ss,
                         double dt, double tol, int max_its, int precision, double *lam, double *gst, double *gss,
                         int &its_total) {
    const Geom &g = c->g;
    hipStream_t s = c->stream;
    auto &I = c->idc;
    const int S = (int)I.nodes.size(), P = I.num_corr, l0 = I.left ? 0 : 1;
    const size_t n = g.npsi(), K = (size_t)g.nq * g.N();
    const bool grads = gst || gss;
    const bool colloc = !I.right || I.collocation;
    bool ok = true;
    its_total = 0;
    // saved forward values: levels 0..P of all stages, then the corrections e[p][k], p = 1..P
    const size_t nlev = (size_t)(P + 1) * S, nsave = grads ? nlev + (size_t)P * S : 0;
    double *saved = grads ? dbuf<double>(I.saved, nsave * n) : nullptr;
    auto lev = [&](int p, int k) { return saved + ((size_t)p * S + k) * n; };
    auto corr = [&](int p, int k) { return saved + (nlev + (size_t)(p - 1) * S + k) * n; };
    if (grads) {
        // forward re-run (idc_step_device with every intermediate kept)
        double *R = dbuf<double>(I.res, (size_t)S * n);
        double *phi = dbuf<double>(I.phi, g.nphi());
        TSW_HIP(hipMemsetAsync(R, 0, sizeof(double) * S * n, s));
        TSW_HIP(hipMemsetAsync(phi, 0, sizeof(double) * g.nphi(), s));
        TSW_HIP(hipMemcpyAsync(lev(0, 0), psi_n, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        launch_fill_region<double>(g, lev(0, 0), K, 1, 0.0, s);
        for (int k = 1; k < S; ++k) {
            const double h = I.nodes[k] - I.nodes[k - 1];
            LinComb L;  // as idc_step_device's prediction
            L.add(1.0 / (h * dt), lev(0, k - 1));
            L.add(1.0, src, true);
            launch_lincomb(g, K, L, lev(0, k), s);
            ok = idc_solve(c, lev(0, k), st, ss, h * dt, tol, max_its, precision, its_total) && ok;
        }
        for (int p = 1; p <= P; ++p) {
            for (int l = l0; l < S; ++l) idc_residual(c, lev(p - 1, l), R + (size_t)l * n, st, ss, phi);
            TSW_HIP(hipMemcpyAsync(lev(p, 0), lev(p - 1, 0), sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            for (int k = 1; k < S; ++k) {
                const double h = I.nodes[k] - I.nodes[k - 1];
                double *e = corr(p, k);
                LinComb L;  // as idc_step_device's correction source
                L.add(-1.0 / (h * dt), lev(p - 1, k));
                for (int l = l0; l < S; ++l) L.add(-I.W[(size_t)k * S + l] / h, R + (size_t)l * n);
                L.add(1.0 / (h * dt), lev(p, k - 1));
                L.add(1.0, src);
                launch_lincomb(g, K, L, e, s);
                ok = idc_solve(c, e, st, ss, h * dt, tol, max_its, precision, its_total) && ok;
                LinComb U;  // lev(p, k) = lev(p-1, k) + e, halos included
                U.add(1.0, lev(p - 1, k), true);
                U.add(1.0, e, true);
                launch_lincomb(g, K, U, lev(p, k), s);
            }
        }
    }
    // reverse pass
    double *ls = dbuf<double>(I.lam, (size_t)S * n);     // stage adjoints (interior; halos zero)
    double *lR = dbuf<double>(I.lamR, (size_t)S * n);    // residual adjoints
    double *z = dbuf<double>(I.err, n);
    TSW_HIP(hipMemsetAsync(ls, 0, sizeof(double) * S * n, s));
    TSW_HIP(hipMemsetAsync(z, 0, sizeof(double) * n, s));
    auto lam_s = [&](int k) { return ls + (size_t)k * n; };
    launch_fill_region<double>(g, lam, K, 1, 0.0, s);
    if (colloc) {
        // out = psi0 + dt Q + dt sum_l W(0,l) (S P - L) psi_l^P
        launch_axpby(n, 1.0, lam, 1.0, lam_s(0), s);
        for (int l = l0; l < S; ++l) {
            bt_apply(c, lam, lam_s(l), st, ss, -dt * I.W[l]);
            if (grads) launch_cell_grads(g, lev(P, l), lam, idc_phi(c, lev(P, l)), gst, gss, s, dt * I.W[l], true);
        }
    } else {
        TSW_HIP(hipMemcpyAsync(lam_s(S - 1), lam, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    }
    for (int p = P; p >= 1; --p) {
        TSW_HIP(hipMemsetAsync(lR, 0, sizeof(double) * S * n, s));
        for (int k = S - 1; k >= 1; --k) {
            const double h = I.nodes[k] - I.nodes[k - 1];
            ok = adjoint_solve(c, lam_s(k), z, st, ss, h * dt, tol, max_its, precision, its_total) && ok;
            if (grads) launch_cell_grads(g, corr(p, k), z, idc_phi(c, corr(p, k)), gst, gss, s, 1.0, true);
            launch_axpby(n, 1.0 / (h * dt), z, 1.0, lam_s(k - 1), s);
            launch_axpby(n, -1.0 / (h * dt), z, 1.0, lam_s(k), s);
            for (int l = l0; l < S; ++l) launch_axpby(n, -I.W[(size_t)k * S + l] / h, z, 1.0, lR + (size_t)l * n, s);
        }
        for (int l = l0; l < S; ++l) {
            bt_apply(c, lR + (size_t)l * n, lam_s(l), st, ss, 1.0);
            if (grads)
                launch_cell_grads(g, lev(p - 1, l), lR + (size_t)l * n, idc_phi(c, lev(p - 1, l)), gst, gss, s, -1.0,
                                  true);
        }
    }
    for (int k = S - 1; k >= 1; --k) {
        const double h = I.nodes[k] - I.nodes[k - 1];
        ok = adjoint_solve(c, lam_s(k), z, st, ss, h * dt, tol, max_its, precision, its_total) && ok;
        if (grads) launch_cell_grads(g, lev(0, k), z, idc_phi(c, lev(0, k)), gst, gss, s, 1.0, true);
        launch_axpby(n, 1.0 / (h * dt), z, 1.0, lam_s(k - 1), s);
    }
    TSW_HIP(hipMemcpyAsync(lam, lam_s(0), sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    launch_fill_region<double>(g, lam, K, 1, 0.0, s);
    return ok;
}

}  // namespace

// =====================================================================================================
// C-ABI
// =====================================================================================================
extern "C" {

int tsw_create(int device, tsw_ctx **out) {
    if (!out) return TSW_E_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return TSW_E_HIP;
    tsw_ctx *c = new (std::nothrow) tsw_ctx();
    if (!c) return TSW_E_HIP;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        scratch_init(device) != TSW_OK) {
        delete c;
        return TSW_E_HIP;
    }
    *out = c;
    return TSW_OK;
}

void tsw_destroy(tsw_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
}

const char *tsw_last_error(const tsw_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int tsw_problem_set(tsw_ctx *ctx, int nx, int ny, double ax, double bx, double ay, double by, int p, int nq,
                    const double *xi, const double *eta, const double *w, const int *quad_off) {
    return guarded(ctx, [&] { set_problem(ctx, nx, ny, ax, bx, ay, by, p, nq, xi, eta, w, quad_off); });
}

int tsw_set_pattern(tsw_ctx *ctx, const char *name) {
    return guarded(ctx, [&] {
        require(name != nullptr, TSW_E_ARG, "null pattern name");
        const std::string s(name);
        if (s == "hip-wavefront") ctx->pattern = Pattern::WavefrontGraph;
        else if (s == "hip-wavefront-eager") ctx->pattern = Pattern::WavefrontEager;
        else if (s == "hip-staggered") ctx->pattern = Pattern::Staggered;
        else if (s == "hip-staggered-ge") ctx->pattern = Pattern::StaggeredGE;
        else throw Error(TSW_E_ARG, "unknown sweep_pattern '" + s + "'");
    });
}

int tsw_set_xs_degree(tsw_ctx *ctx, int degree) {
    return guarded(ctx, [&] {
        need_problem(ctx);
        require(degree == 0 || degree == ctx->g.p, TSW_E_ARG,
                "cross-section degree must be 0 or the problem's DG degree " + std::to_string(ctx->g.p));
        ctx->xs_deg = degree;
    });
}

int tsw_set_tuning(tsw_ctx *ctx, const char *key, int value) {
    return guarded(ctx, [&] {
        require(key != nullptr, TSW_E_ARG, "null tuning key");
        const std::string k(key);
        if (k == "band_rows") {
            require(value >= -1, TSW_E_ARG, "band_rows must be -1, 0 or a row count");
            ctx->band_rows_opt = value;
        } else if (k == "use_dsa") {
            require(value == 0 || value == 1, TSW_E_ARG, "use_dsa must be 0 or 1");
            ctx->use_dsa = value != 0;
        } else if (k == "dsa_cycles") {
            require(value >= 1 && value <= 16, TSW_E_ARG, "dsa_cycles must be 1..16");
            ctx->dsa_cycles = value;
        } else if (k == "si_graph") {
            require(value == 0 || value == 1, TSW_E_ARG, "si_graph must be 0 or 1");
            ctx->si_graph = value != 0;
        } else if (k == "one_wave") {
            require(value == 0 || value == 1, TSW_E_ARG, "one_wave must be 0 or 1");
            ctx->one_wave_opt = value;
        } else if (k == "lane_split") {
            require(value == 0 || value == 1, TSW_E_ARG, "lane_split must be 0 or 1");
            ctx->split_opt = value;
        } else if (k == "batch_direct") {
            require(value >= -1 && value <= 1, TSW_E_ARG, "batch_direct must be -1, 0 or 1");
            ctx->batch_direct_opt = value;
            for (auto &G : ctx->groups) G.built32 = G.built32c = G.built64 = false;
        } else if (k == "ref_stage") {
            require(value == 0 || value == 1, TSW_E_ARG, "ref_stage must be 0 or 1");
            ctx->ref_stage_opt = value;
        } else if (k == "batch_ref") {
            require(value == -1 || value == 0 || value == 1, TSW_E_ARG, "batch_ref must be -1, 0 or 1");
            ctx->batch_ref_opt = value;
        } else if (k == "batch_split") {
            require(value == -1 || value == 0 || value == 1, TSW_E_ARG, "batch_split must be -1, 0 or 1");
            ctx->batch_split_opt = value;
        } else if (k == "batch_split_rows") {
            require(value == 0 || (value >= 16 && value <= 256), TSW_E_ARG, "batch_split_rows must be 0 or 16..256");
            ctx->batch_split_rows_opt = value;
        } else if (k == "ref_split") {
            require(value == -1 || value == 1, TSW_E_ARG, "ref_split must be -1 or 1");
            ctx->ref_split_opt = value;
        } else if (k == "batch_persistent") {
            require(value == 0 || value == 1, TSW_E_ARG, "batch_persistent must be 0 or 1");
            ctx->batch_persistent_opt = value;
        } else if (k == "batch_xcd_tile") {
            require(value >= 0 && value <= 64, TSW_E_ARG, "batch_xcd_tile must be 0..64");
            ctx->batch_xcd_tile_opt = value;
        } else if (k == "batch_split_prefetch") {
            require(value == 0 || value == 2 || value == 3 || value == 4 || value == 6, TSW_E_ARG,
                    "batch_split_prefetch must be 0, 2, 3, 4 or 6");
            ctx->batch_split_pf_opt = value;
        } else if (k == "gmres_refine") {
            require(value == 0 || value == 1, TSW_E_ARG, "gmres_refine must be 0 or 1");
            ctx->gmres_refine = value != 0;
        } else if (k == "fp32_safety") {
            require(value >= 1 && value <= 1000, TSW_E_ARG, "fp32_safety must be 1..1000");
            ctx->fp32_safety = value;
        } else if (k == "refine_steps") {
            require(value >= 1 && value <= 8, TSW_E_ARG, "refine_steps must be 1..8");
            ctx->refine_steps = value;
        } else if (k == "autotune") {
            require(value == 0 || value == 1, TSW_E_ARG, "autotune must be 0 or 1");
            ctx->autotune = value != 0;
        } else if (k == "prefetch") {
            require(value == 0 || value == 2 || value == 3 || value == 4 || value == 6, TSW_E_ARG,
                    "prefetch must be 0, 2, 3, 4 or 6");
            ctx->prefetch_opt = value;
        } else if (k == "lanes") {
            require(value >= -1 && value <= 1, TSW_E_ARG, "lanes must be -1, 0 or 1");
            ctx->lanes_opt = value;
        } else if (k == "lanes_block") {
            require(value >= 0 && value <= 64, TSW_E_ARG, "lanes_block must be 0..64");
            ctx->lanes_J_opt = value;
        } else if (k == "lanes_waves") {
            require(value >= 0 && value <= 4, TSW_E_ARG, "lanes_waves must be 0..4");
            ctx->lanes_W_opt = value;
        } else if (k == "lanes_prefetch") {
            require(value == 0 || value == 2 || value == 3 || value == 4 || value == 6, TSW_E_ARG,
                    "lanes_prefetch must be 0, 2, 3, 4 or 6");
            ctx->lanes_pf_opt = value;
        } else if (k == "stall_test") {
            require(value == 0 || value == 1, TSW_E_ARG, "stall_test must be 0 or 1");
            ctx->stall_test = value != 0;
        } else {
            throw Error(TSW_E_ARG, "unknown tuning key '" + k + "'");
        }
    });
}

int tsw_get_tuning(tsw_ctx *ctx, const char *key, int *value) {
    return guarded(ctx, [&] {
        require(key != nullptr && value != nullptr, TSW_E_ARG, "null argument");
        const std::string k(key);
        if (k == "band_rows") *value = ctx->last_br;
        else if (k == "prefetch") *value = ctx->last_pf;
        else if (k == "autotune") *value = ctx->autotune ? 1 : 0;
        else if (k == "si_graph") *value = ctx->si_graph ? 1 : 0;
        else if (k == "use_dsa") *value = ctx->use_dsa ? 1 : 0;
        else if (k == "dsa_cycles") *value = ctx->dsa_cycles;
        else if (k == "refine_steps") *value = ctx->refine_steps;
        else if (k == "gmres_refine") *value = ctx->gmres_refine ? 1 : 0;
        else if (k == "fp32_safety") *value = (int)ctx->fp32_safety;
        else if (k == "one_wave") *value = ctx->last_w1;
        else if (k == "lane_split") *value = ctx->last_sl;
        else if (k == "batch_direct") *value = ctx->batch_direct_opt;
        else if (k == "batch_ref") *value = ctx->last_batch_ref ? 1 : 0;
        else if (k == "ref_stage") *value = ctx->ref_stage_opt;
        else if (k == "batch_split") *value = ctx->batch_split_opt;
        else if (k == "batch_split_rows") *value = ctx->batch_split_rows_opt;
        else if (k == "batch_xcd_tile") *value = ctx->batch_xcd_tile_opt;
        else if (k == "batch_persistent") *value = ctx->batch_persistent_opt;
        else if (k == "ref_split") *value = ctx->ref_split_opt;
        else if (k == "batch_split_prefetch")
            *value = ctx->batch_split_pf_opt > 0 ? ctx->batch_split_pf_opt : kSplitBatchPrefetch;
        else if (k.rfind("batch_engine_p", 0) == 0 && k.size() == 15 && k[14] >= '0' && k[14] <= '3')
            *value = ctx->batch_engine[k[14] - '0'];
        else if (k.rfind("batch_split_rows_p", 0) == 0 && k.size() == 19 && k[18] >= '0' && k[18] <= '3')
            *value = ctx->split_rows_p[k[18] - '0'];
        else if (k == "lanes") *value = ctx->batch.empty() ? ctx->last_lanes : (ctx->lane_groups.empty() ? 0 : 1);
        else if (k == "lanes_block") *value = ctx->last_lanes_J;
        else if (k == "lanes_waves") *value = ctx->last_lanes_W;
        else if (k == "lanes_prefetch") *value = ctx->lanes_pf_opt;
        else throw Error(TSW_E_ARG, "unknown tuning key '" + k + "'");
    });
}

int tsw_sweep_apply(tsw_ctx *ctx, double *psi_inout, const double *sigma_t, double dt, int precision, int mem) {
    return guarded(ctx, [&] {
        need_problem(ctx);
        const Geom &g = ctx->g;
        require(precision == TSW_FP64 || precision == TSW_FP32 || precision == TSW_MIXED, TSW_E_ARG,
                "bad precision");
        double *psi = stage_in(ctx, psi_inout, g.npsi(), mem, ctx->psi64);
        const double *sig = stage_in(ctx, sigma_t, xs_n(ctx), mem, ctx->sig_t64);
        sweep_device(ctx, psi, sig, dt, precision);
        if (mem == TSW_MEM_HOST) stage_out(ctx, psi_inout, psi, g.npsi(), mem);
        sync(ctx);
    });
}

int tsw_sweep_apply_multi(tsw_ctx *ctx, int n_rhs, double *const *psi_inout, const double *sigma_t, double dt,
                          int precision, int mem) {
    return guarded(ctx, [&] {
        need_problem(ctx);
        const Geom &g = ctx->g;
        require(n_rhs >= 1 && psi_inout != nullptr, TSW_E_ARG, "empty right-hand-side list");
        require(precision == TSW_FP64 || precision == TSW_FP32 || precision == TSW_MIXED, TSW_E_ARG,
                "bad precision");
        for (int r = 0; r < n_rhs; ++r) require(psi_inout[r] != nullptr, TSW_E_ARG, "null psi array");
        const double *sig = stage_in(ctx, sigma_t, xs_n(ctx), mem, ctx->sig_t64);
        std::vector<double *> dev(n_rhs);
        const size_t npsi = g.npsi(), stride = (npsi * sizeof(double) + 255) / 256 * 256;
        if (mem == TSW_MEM_HOST) {
            char *base = static_cast<char *>(ctx->multi.stage.reserve(stride * n_rhs));
            for (int r = 0; r < n_rhs; ++r) {
                dev[r] = reinterpret_cast<double *>(base + stride * r);
                TSW_HIP(hipMemcpyAsync(dev[r], psi_inout[r], npsi * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
            }
        } else {
            require(mem == TSW_MEM_DEVICE, TSW_E_ARG, "mem must be TSW_MEM_HOST or TSW_MEM_DEVICE");
            for (int r = 0; r < n_rhs; ++r) dev[r] = psi_inout[r];
        }
        bool one_launch = ctx->reflect_mask == 0;  // reflecting sweeps run one source after the other
        for (int r = 0; r < n_rhs; ++r) one_launch = one_launch && ref_sweep_ok(ctx, dev[r]);
        if (one_launch) {
            // TSW_MIXED: the FP64 sweep (this engine reads and writes the FP64 arrays in either precision)
            if (precision == TSW_FP32) multi_sweep<float>(ctx, dev, sig, dt);
            else multi_sweep<double>(ctx, dev, sig, dt);
        } else {
            for (int r = 0; r < n_rhs; ++r) sweep_device(ctx, dev[r], sig, dt, precision);
        }
        if (mem == TSW_MEM_HOST)
            for (int r = 0; r < n_rhs; ++r)
                TSW_HIP(hipMemcpyAsync(psi_inout[r], dev[r], npsi * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        sync(ctx);
    });
}

int tsw_set_boundaries(tsw_ctx *ctx, int reflect_mask, const int *reflect_map) {
    return guarded(ctx, [&] {
        need_problem(ctx);
        const int nq = ctx->g.nq;
        require(reflect_mask >= 0 && (reflect_mask & ~(15 | TSW_PERIODIC)) == 0, TSW_E_ARG,
                "reflect_mask takes TSW_REFLECT_* edges or TSW_PERIODIC");
        require(!((reflect_mask & TSW_PERIODIC) && (reflect_mask & 15)), TSW_E_ARG,
                "combined reflecting and periodic boundary conditions are not supported (init.cpp:498-510)");
        ctx->reflect_mask = 0;
        ctx->periodic = (reflect_mask & TSW_PERIODIC) != 0;
        reflect_mask &= 15;
        if (reflect_mask == 0) return;
        std::vector<int> map(2 * nq, -1);
        if (reflect_map) {
            for (int k = 0; k < 2 * nq; ++k) {
                require(reflect_map[k] >= 0 && reflect_map[k] < nq, TSW_E_ARG, "reflect_map entry out of range");
                map[k] = reflect_map[k];
            }
        } else {  // ReflectMap (OrdinateSet.cpp:1093-1140) from (xi, eta): the first ordinate within 1e-14
            const double TOL = 1.0e-14;
            const auto &xi = ctx->h_xi, &eta = ctx->h_eta;
            for (int q = 0; q < nq; ++q)
                for (int k = 0; k < q; ++k)
                    require(!(std::fabs(xi[q] - xi[k]) < TOL && std::fabs(eta[q] - eta[k]) < TOL), TSW_E_ARG,
                            "ordinates share (xi, eta): pass reflect_map");
            for (int dim = 0; dim < 2; ++dim)
                for (int q = 0; q < nq; ++q) {
                    for (int k = 0; k < nq && map[dim * nq + q] < 0; ++k) {
                        const double cx = dim == 0 ? -xi[k] : xi[k], cy = dim == 1 ? -eta[k] : eta[k];
                        if (std::fabs(xi[q] - cx) < TOL && std::fabs(eta[q] - cy) < TOL) map[dim * nq + q] = k;
                    }
                    require(map[dim * nq + q] >= 0, TSW_E_ARG,
                            "ordinate " + std::to_string(q) + " has no mirror in direction " + std::to_string(dim));
                }
        }
        TSW_HIP(hipMemcpy(ctx->reflect_map_d.reserve(sizeof(int) * map.size()), map.data(), sizeof(int) * map.size(),
                          hipMemcpyHostToDevice));
        ctx->reflect_mask = reflect_mask;
    });
}

int tsw_synchronize_halos(tsw_ctx *ctx, double *psi_inout, int mem) {
    return guarded(ctx, [&] {
        need_problem(ctx);
        const Geom &g = ctx->g;
        double *psi = stage_in(ctx, psi_inout, g.npsi(), mem, ctx->psi64);
        if (ctx->reflect_mask) launch_reflect_boundaries(g, psi, ctx->reflect_map_d.as<int>(), ctx->reflect_mask, ctx->stream);
        if (ctx->periodic) launch_periodic_halos(g, psi, ctx->stream);
        if (mem == TSW_MEM_HOST) stage_out(ctx, psi_inout, psi, g.npsi(), mem);
        sync(ctx);
    });
}

int tsw_reflect_boundaries(tsw_ctx *ctx, double *psi_inout, int mem) {
    return guarded(ctx, [&] {
        need_problem(ctx);
        const Geom &g = ctx->g;
        double *psi = stage_in(ctx, psi_inout, g.npsi(), mem, ctx->psi64);
        if (ctx->reflect_mask) launch_reflect_boundaries(g, psi, ctx->reflect_map_d.as<int>(), ctx->reflect_mask, ctx->stream);
        if (mem == TSW_MEM_HOST) stage_out(ctx, psi_inout, psi, g.npsi(), mem);
        sync(ctx);
    });
}

int tsw_sweep_apply_noalloc(tsw_ctx *ctx, double *psi_inout, const double *sigma_t, double dt, int precision) {
    return guarded(ctx, [&] {
        require(ctx->reflect_mask == 0 && !ctx->periodic, TSW_E_UNSUPPORTED, "tsw_sweep_apply_noalloc: reflecting and periodic boundaries are built for the "
                "sweeps, tsw_si_solve and the halo updates");
        need_problem(ctx);
        require(precision == TSW_FP64 || precision == TSW_FP32 || precision == TSW_MIXED, TSW_E_ARG,
                "bad precision");
        require(psi_inout != nullptr && sigma_t != nullptr, TSW_E_ARG, "null array argument");
        require(ctx->xs_deg == 0, TSW_E_UNSUPPORTED, "the allocation-free sweep takes cell-mean cross sections");
        require(ctx->device < kMaxScratchDevices && g_scratch_ptr[ctx->device] != nullptr, TSW_E_STATE,
                "no scratch address for this device");
        const Geom &g = ctx->g;
        hipStream_t s = ctx->stream;
        const size_t npsi = g.npsi(), nc = g.ncell_h(), per = (size_t)g.nq * g.N();
        auto r256 = [](size_t b) { return (b + 255) / 256 * 256; };
        // scratch carve-up: FP32 copies of psi and sigma_t; TSW_MIXED adds the FP64 iterate and residual
        const size_t b32 = r256(npsi * sizeof(float)), bs32 = r256(nc * sizeof(float)), b64 = r256(npsi * sizeof(double));
        const size_t need = precision == TSW_FP64 ? 0 : b32 + bs32 + (precision == TSW_MIXED ? 2 * b64 : 0);
        require(need <= kScratchBytes, TSW_E_ARG,
                "the allocation-free sweep needs " + std::to_string(need) + " bytes of scratch for this problem and "
                "precision; it holds " + std::to_string(kScratchBytes));
        unsigned char *base = static_cast<unsigned char *>(g_scratch_ptr[ctx->device]);
        float *p32 = reinterpret_cast<float *>(base);
        float *s32 = reinterpret_cast<float *>(base + b32);
        double *psi64 = reinterpret_cast<double *>(base + b32 + bs32);
        double *res64 = reinterpret_cast<double *>(base + b32 + bs32 + b64);
        if (precision == TSW_FP64) {
            launch_sweep_wavefront<double>(params<double>(ctx, psi_inout, sigma_t, dt), g.p, s);
        } else {
            launch_convert_d2f(sigma_t, s32, nc, s);
            const SweepParams<float> sp = params<float>(ctx, p32, s32, dt);
            if (precision == TSW_FP32) {
                launch_convert_d2f(psi_inout, p32, npsi, s);
                launch_sweep_wavefront<float>(sp, g.p, s);
                launch_convert_f2d_interior(g, per, p32, psi_inout, s);
            } else {
                // as sweep_wavefront_mixed: from the iterate with zero interior and the source's inflow halos,
                // 1 + refine_steps times psi += L32^{-1}(Q - (L + 1/dt) psi), FP64 residual, FP32 zero-inflow sweep
                TSW_HIP(hipMemcpyAsync(psi64, psi_inout, sizeof(double) * npsi, hipMemcpyDeviceToDevice, s));
                launch_fill_region<double>(g, psi64, per, 0, 0.0, s);
                for (int k = 0; k <= ctx->refine_steps; ++k) {
                    TSW_HIP(hipMemcpyAsync(res64, psi_inout, sizeof(double) * npsi, hipMemcpyDeviceToDevice, s));
                    launch_lmv<double>(g, ctx->xi64.as<double>(), ctx->eta64.as<double>(), sigma_t, psi64, res64, -1.0,
                                       1.0, s);
                    if (inverse_dt(dt) != 0.0) launch_axpby(npsi, -inverse_dt(dt), psi64, 1.0, res64, s);
                    launch_fill_region<double>(g, res64, per, 1, 0.0, s);
                    launch_convert_d2f(res64, p32, npsi, s);
                    launch_sweep_wavefront<float>(sp, g.p, s);
                    launch_accumulate_f2d(p32, psi64, npsi, s);
                }
                TSW_HIP(hipMemcpyAsync(psi_inout, psi64, sizeof(double) * npsi, hipMemcpyDeviceToDevice, s));
            }
        }
        TSW_HIP(hipStreamSynchronize(s));
    });
}

int tsw_pmv(tsw_ctx *ctx, double alpha, double beta, const double *psi, double *phi, int all_cells, int mem) {
    return guarded(ctx, [&] {
        need_problem(ctx);
        const Geom &g = ctx->g;
        const double *x = stage_in(ctx, psi, g.npsi(), mem, ctx->psi64);
        double *y = stage_in(ctx, (const double *)phi, g.nphi(), mem, ctx->phi64a);
        launch_pmv<double, double>(g, x, y, ctx->w64.as<double>(), alpha, beta, all_cells ? 1 : 0, ctx->stream);
        if (mem == TSW_MEM_HOST) stage_out(ctx, phi, y, g.nphi(), mem);
        sync(ctx);
    });
}

int tsw_smv(tsw_ctx *ctx, double alpha, double beta, const double *sigma_s, const double *phi, double *psi,
            int all_cells, int mem) {
    return guarded(ctx, [&] {
        need_problem(ctx);
        const Geom &g = ctx->g;
        const double *s = stage_in(ctx, sigma_s, xs_n(ctx), mem, ctx->sig_s64);
        const double *f = stage_in(ctx, phi, g.nphi(), mem, ctx->phi64a);
        double *y = stage_in(ctx, (const double *)psi, g.npsi(), mem, ctx->psi64);
        launch_smv<double>(g, s, f, y, alpha, beta, all_cells ? 1 : 0, ctx->stream, xs_tensor<double>(ctx, s));
        if (mem == TSW_MEM_HOST) stage_out(ctx, psi, y, g.npsi(), mem);
        sync(ctx);
    });
}

int tsw_lmv(tsw_ctx *ctx, double alpha, double beta, const double *sigma_t, const double *x, double *y, int mem) {
    return guarded(ctx, [&] {
        need_problem(ctx);
        require(x != nullptr && y != nullptr, TSW_E_ARG, "null array argument");
        const Geom &g = ctx->g;
        const double *s = stage_in(ctx, sigma_t, xs_n(ctx), mem, ctx->sig_t64);
        const double *xd = stage_in(ctx, x, g.npsi(), mem, ctx->psi64);
        double *yd = stage_in(ctx, (const double *)y, g.npsi(), mem, ctx->r64);
        if (xd == yd) {
            // y = beta y + alpha L y (Lmv_InPlace, TransportOperator.cpp:602-1200, dispatched at :1224-1228):
            // the product reads a copy of y, so every cell sees its upwind neighbours' values from before
            // the call, as the reference's in-place wavefront does
            double *xc = dbufd(ctx->psi64, g.npsi());
            TSW_HIP(hipMemcpyAsync(xc, yd, sizeof(double) * g.npsi(), hipMemcpyDeviceToDevice, ctx->stream));
            xd = xc;
        }
        launch_lmv<double>(g, ctx->xi64.as<double>(), ctx->eta64.as<double>(), s, xd, yd, alpha, beta, ctx->stream,
                           xs_tensor<double>(ctx, s));
        if (mem == TSW_MEM_HOST) stage_out(ctx, y, yd, g.npsi(), mem);
        sync(ctx);
    });
}

int tsw_si_solve(tsw_ctx *ctx, double *source_inout, const double *sigma_t, const double *sigma_s, double dt,
                 double abs_tol, int max_its, int precision, int mem, int *iters, double *err) {
    return guarded(ctx, [&] {
        need_problem(ctx);
        if (precision == TSW_MIXED) precision = TSW_FP32;  // solvers refine their FP32 sweeps themselves
        require(precision == TSW_FP64 || precision == TSW_FP32, TSW_E_ARG, "bad precision");
        const Geom &g = ctx->g;
        hipStream_t s = ctx->stream;
        double *Q = stage_in(ctx, source_inout, g.npsi(), mem, ctx->q64);
        const double *st = stage_in(ctx, sigma_t, xs_n(ctx), mem, ctx->sig_t64);
        const double *ss = stage_in(ctx, sigma_s, xs_n(ctx), mem, ctx->sig_s64);
        if (ctx->reflect_mask || ctx->periodic) {  // SISolver::Solve, reflecting or periodic (SISolver.cpp:190-254)
            double *psi = ctx->si_state.as<double>();
            double *cur = static_cast<double *>(ctx->phi64a.reserve(sizeof(double) * g.nphi()));
            double *last = static_cast<double *>(ctx->phi64b.reserve(sizeof(double) * g.nphi()));
            TSW_HIP(hipMemsetAsync(cur, 0, sizeof(double) * g.nphi(), s));
            TSW_HIP(hipMemsetAsync(last, 0, sizeof(double) * g.nphi(), s));
            const double *w = ctx->w64.as<double>();
            const int *map = ctx->reflect_map_d.as<int>();
            launch_pmv<double, double>(g, psi, cur, w, 1.0, 0.0, 0, s);  // :214
            int it = 0;
            double e = 1.79769313486231570e+308;
            while (it < max_its && e > abs_tol) {
                ++it;
                std::swap(cur, last);
                // the Smv of every iteration leaves psi's halos dirty: SynchronizeHalos reflects them, or
                // copies them across the periodic boundaries (:222); the source copy then skips those halos
                if (ctx->reflect_mask) launch_reflect_boundaries(g, psi, map, ctx->reflect_mask, s);
                if (ctx->periodic) launch_periodic_halos(g, psi, s);
                launch_copy_outside_reflect(g, Q, psi, ctx->periodic ? 15 : ctx->reflect_mask, s);  // :226
                launch_smv<double>(g, ss, last, psi, 1.0, 1.0, 0, s, xs_tensor<double>(ctx, ss));  // :229
                sweep_device(ctx, psi, st, dt, precision);                                          // :232
                launch_pmv<double, double>(g, psi, cur, w, 1.0, 0.0, 0, s);  // :236
                e = device_norm2_interior_diff(ctx, cur, last);             // :239
            }
            TSW_HIP(hipMemcpyAsync(Q, psi, sizeof(double) * g.npsi(), hipMemcpyDeviceToDevice, s));  // :249
            if (mem == TSW_MEM_HOST || Q != source_inout) stage_out(ctx, source_inout, Q, g.npsi(), mem);
            sync(ctx);
            if (iters) *iters = it;
            if (err) *err = e;
            return;
        }
        bool lanes = false;
        if (lanes_ok(ctx)) {
            prepare_materials(ctx, st, dt);
            lanes = ctx->mat.valid;
        }
        ctx->last_lanes = lanes ? 1 : 0;
        if (lanes) {
            int it = 0;
            double e = 0.0;
            if (precision == TSW_FP64) si_lanes<double>(ctx, Q, st, ss, dt, abs_tol, max_its, ctx->si_graph, it, e);
            else si_lanes<float>(ctx, Q, st, ss, dt, abs_tol, max_its, ctx->si_graph, it, e);
            if (mem == TSW_MEM_HOST || Q != source_inout) stage_out(ctx, source_inout, Q, g.npsi(), mem);
            sync(ctx);
            if (iters) *iters = it;
            if (err) *err = e;
            return;
        }
        if (staggered_ok(ctx)) {
            int it = 0;
            double e = 0.0;
            if (ctx->si_graph) {
                if (precision == TSW_FP64) si_staggered_graph<double>(ctx, Q, st, ss, dt, abs_tol, max_its, it, e);
                else si_staggered_graph<float>(ctx, Q, st, ss, dt, abs_tol, max_its, it, e);
            } else {
                if (precision == TSW_FP64) si_staggered<double>(ctx, Q, st, ss, dt, abs_tol, max_its, it, e);
                else si_staggered<float>(ctx, Q, st, ss, dt, abs_tol, max_its, it, e);
            }
            if (mem == TSW_MEM_HOST || Q != source_inout) stage_out(ctx, source_inout, Q, g.npsi(), mem);
            sync(ctx);
            if (iters) *iters = it;
            if (err) *err = e;
            return;
        }
        double *psi = ctx->si_state.as<double>();
        double *cur = static_cast<double *>(ctx->phi64a.reserve(sizeof(double) * g.nphi()));
        double *last = static_cast<double *>(ctx->phi64b.reserve(sizeof(double) * g.nphi()));
        TSW_HIP(hipMemsetAsync(cur, 0, sizeof(double) * g.nphi(), s));
        TSW_HIP(hipMemsetAsync(last, 0, sizeof(double) * g.nphi(), s));
        const double *w = ctx->w64.as<double>();
        launch_pmv<double, double>(g, psi, cur, w, 1.0, 0.0, 0, s);  // SISolver.cpp:214
        int it = 0;
        double e = 1.79769313486231570e+308;
        while (it < max_its && e > abs_tol) {
            ++it;
            std::swap(cur, last);
            TSW_HIP(hipMemcpyAsync(psi, Q, sizeof(double) * g.npsi(), hipMemcpyDeviceToDevice, s));  // :226
            launch_smv<double>(g, ss, last, psi, 1.0, 1.0, 0, s, xs_tensor<double>(ctx, ss));         // :229
            sweep_prec(ctx, psi, st, dt, precision);                                                  // :232
            launch_pmv<double, double>(g, psi, cur, w, 1.0, 0.0, 0, s);                                      // :236
            e = device_norm2_interior_diff(ctx, cur, last);                                           // :239
        }
        TSW_HIP(hipMemcpyAsync(Q, psi, sizeof(double) * g.npsi(), hipMemcpyDeviceToDevice, s));  // :249
        if (mem == TSW_MEM_HOST) stage_out(ctx, source_inout, Q, g.npsi(), mem);
        else if (Q != source_inout) stage_out(ctx, source_inout, Q, g.npsi(), mem);
        sync(ctx);
        if (iters) *iters = it;
        if (err) *err = e;
    });
}

int tsw_gmres_solve(tsw_ctx *ctx, double *source_inout, const double *sigma_t, const double *sigma_s, double dt,
                    double abs_tol, int max_its, int restart, int precision, int mem, double *phi_out, int *iters,
                    double *resid) {
    int its_total = 0;
    double res_final = 0.0;
    bool converged = false;
    const int rc = guarded(ctx, [&] {
        require(ctx->reflect_mask == 0 && !ctx->periodic, TSW_E_UNSUPPORTED, "tsw_gmres_solve: reflecting and periodic boundaries are built for the "
                "sweeps, tsw_si_solve and the halo updates");
        need_problem(ctx);
        require(restart >= 1 && max_its >= 0, TSW_E_ARG, "bad GMRES parameters");
        if (precision == TSW_MIXED) precision = TSW_FP32;  // solvers refine their FP32 sweeps themselves
        require(precision == TSW_FP64 || precision == TSW_FP32, TSW_E_ARG, "bad precision");
        const Geom &g = ctx->g;
        hipStream_t s = ctx->stream;
        const size_t n = g.nphi(), npsi = g.npsi();
        double *Q = stage_in(ctx, source_inout, npsi, mem, ctx->q64);
        const double *st = stage_in(ctx, sigma_t, xs_n(ctx), mem, ctx->sig_t64);
        const double *ss = stage_in(ctx, sigma_s, xs_n(ctx), mem, ctx->sig_s64);
        DBuf phi_dev;
        double *pd = phi_out ? (mem == TSW_MEM_DEVICE ? phi_out : static_cast<double *>(phi_dev.reserve(sizeof(double) * n))) : nullptr;
        converged = gmres_device(ctx, Q, st, ss, dt, abs_tol, max_its, restart, precision, pd, its_total, res_final);
        if (phi_out && mem == TSW_MEM_HOST)
            TSW_HIP(hipMemcpyAsync(phi_out, pd, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        if (mem == TSW_MEM_HOST) stage_out(ctx, source_inout, Q, npsi, mem);
        sync(ctx);
    });
    if (iters) *iters = its_total;
    if (resid) *resid = res_final;
    if (rc == TSW_OK && !converged) {
        ctx->err = "GMRES transport solve failed to converge";
        return TSW_E_NOCONV;
    }
    return rc;
}

int tsw_sweep_vjp(tsw_ctx *ctx, const double *psi, const double *sigma_t, double dt, const double *seed,
                  double *grad_src, double *grad_sigma_t, int precision, int mem) {
    return guarded(ctx, [&] {
        require(ctx->reflect_mask == 0 && !ctx->periodic, TSW_E_UNSUPPORTED, "tsw_sweep_vjp: reflecting and periodic boundaries are built for the "
                "sweeps, tsw_si_solve and the halo updates");
        need_problem(ctx);
        require(ctx->xs_deg == 0, TSW_E_UNSUPPORTED, "vector-Jacobian products take cell-mean cross sections "
                                                     "(tsw_set_xs_degree 0)");
        require(seed != nullptr && grad_src != nullptr, TSW_E_ARG, "seed and grad_src are required");
        require(grad_sigma_t == nullptr || psi != nullptr, TSW_E_ARG, "grad_sigma_t needs the forward psi");
        const Geom &g = ctx->g;
        const double *st = stage_in(ctx, sigma_t, xs_n(ctx), mem, ctx->sig_t64);
        const double *sd = stage_in(ctx, seed, g.npsi(), mem, ctx->vjp_seed);
        const double *pd = psi ? stage_in(ctx, psi, g.npsi(), mem, ctx->vjp_psi) : nullptr;
        double *gsrc = mem == TSW_MEM_DEVICE ? grad_src : dbuf<double>(ctx->vjp_gsrc, g.npsi());
        double *gst = grad_sigma_t ? (mem == TSW_MEM_DEVICE ? grad_sigma_t : dbuf<double>(ctx->vjp_gst, g.ncell_h()))
                                   : nullptr;
        sweep_vjp_device(ctx, pd, st, dt, sd, gsrc, gst, precision);
        if (mem == TSW_MEM_HOST) {
            stage_out(ctx, grad_src, gsrc, g.npsi(), mem);
            if (gst) stage_out(ctx, grad_sigma_t, gst, g.ncell_h(), mem);
        }
        sync(ctx);
    });
}

int tsw_solve_vjp(tsw_ctx *ctx, const double *psi, const double *sigma_t, const double *sigma_s, double dt,
                  const double *seed, double *grad_src, double *grad_sigma_t, double *grad_sigma_s, double abs_tol,
                  int max_its, int restart, int precision, int mem, int *iters, double *resid) {
    int its = 0;
    double res = 0.0;
    bool converged = false;
    const int rc = guarded(ctx, [&] {
        require(ctx->reflect_mask == 0 && !ctx->periodic, TSW_E_UNSUPPORTED, "tsw_solve_vjp: reflecting and periodic boundaries are built for the "
                "sweeps, tsw_si_solve and the halo updates");
        need_problem(ctx);
        require(ctx->xs_deg == 0, TSW_E_UNSUPPORTED, "vector-Jacobian products take cell-mean cross sections "
                                                     "(tsw_set_xs_degree 0)");
        require(seed != nullptr && grad_src != nullptr, TSW_E_ARG, "seed and grad_src are required");
        require((grad_sigma_t == nullptr && grad_sigma_s == nullptr) || psi != nullptr, TSW_E_ARG,
                "cross-section gradients need the forward psi");
        require(restart >= 1 && max_its >= 0, TSW_E_ARG, "bad GMRES parameters");
        const Geom &g = ctx->g;
        const double *st = stage_in(ctx, sigma_t, xs_n(ctx), mem, ctx->sig_t64);
        const double *ss = stage_in(ctx, sigma_s, xs_n(ctx), mem, ctx->sig_s64);
        const double *sd = stage_in(ctx, seed, g.npsi(), mem, ctx->vjp_seed);
        const double *pd = psi ? stage_in(ctx, psi, g.npsi(), mem, ctx->vjp_psi) : nullptr;
        double *gsrc = mem == TSW_MEM_DEVICE ? grad_src : dbuf<double>(ctx->vjp_gsrc, g.npsi());
        auto cellbuf = [&](double *user, DBuf &b) {
            return user ? (mem == TSW_MEM_DEVICE ? user : dbuf<double>(b, g.ncell_h())) : nullptr;
        };
        double *gst = cellbuf(grad_sigma_t, ctx->vjp_gst), *gss = cellbuf(grad_sigma_s, ctx->vjp_gss);
        converged = solve_vjp_device(ctx, pd, st, ss, dt, sd, gsrc, gst, gss, abs_tol, max_its, restart, precision,
                                     its, res);
        if (mem == TSW_MEM_HOST) {
            stage_out(ctx, grad_src, gsrc, g.npsi(), mem);
            if (gst) stage_out(ctx, grad_sigma_t, gst, g.ncell_h(), mem);
            if (gss) stage_out(ctx, grad_sigma_s, gss, g.ncell_h(), mem);
        }
        sync(ctx);
    });
    if (iters) *iters = its;
    if (resid) *resid = res;
    if (rc == TSW_OK && !converged) {
        ctx->err = "adjoint GMRES transport solve failed to converge";
        return TSW_E_NOCONV;
    }
    return rc;
}

int tsw_idc_set(tsw_ctx *ctx, int num_nodes, int num_corrections, int nodes_type, int use_collocation) {
    return guarded(ctx, [&] {
        require(nodes_type >= TSW_IDC_LOBATTO && nodes_type <= TSW_IDC_CHEBYSHEV_RADAU, TSW_E_ARG,
                "unknown IDC node type");
        require(num_nodes >= (nodes_type == TSW_IDC_LOBATTO ? 2 : 1) && num_nodes <= 12, TSW_E_ARG,
                "IDC node count out of range");
        require(num_corrections >= 0 && num_corrections <= 32, TSW_E_ARG, "IDC correction count out of range");
        auto &I = ctx->idc;
        I.num_nodes = num_nodes;
        I.num_corr = num_corrections;
        I.kind = nodes_type;
        I.collocation = use_collocation != 0;
        idc_tables(I);
    });
}

int tsw_idc_set_type(tsw_ctx *ctx, int idc_type) {
    return guarded(ctx, [&] {
        require(idc_type == TSW_IDC_ERROR || idc_type == TSW_IDC_UPDATE, TSW_E_ARG, "unknown IDC type");
        ctx->idc.form = idc_type;
    });
}

int tsw_idc_step(tsw_ctx *ctx, double *psi_inout, const double *source, const double *sigma_t,
                 const double *sigma_s, double dt, double abs_tol, int max_its, int precision, int mem) {
    bool ok = true;
    const int rc = guarded(ctx, [&] {
        require(ctx->reflect_mask == 0 && !ctx->periodic, TSW_E_UNSUPPORTED, "tsw_idc_step: reflecting and periodic boundaries are built for the "
                "sweeps, tsw_si_solve and the halo updates");
        need_problem(ctx);
        require(dt > 0.0 && std::isfinite(dt), TSW_E_ARG, "IDC needs a finite time step");
        if (precision == TSW_MIXED) precision = TSW_FP32;  // solvers refine their FP32 sweeps themselves
        require(precision == TSW_FP64 || precision == TSW_FP32, TSW_E_ARG, "bad precision");
        if (ctx->idc.nodes.empty()) idc_tables(ctx->idc);
        const Geom &g = ctx->g;
        double *psi = stage_in(ctx, psi_inout, g.npsi(), mem, ctx->q64);
        const double *src = stage_in(ctx, source, g.npsi(), mem, ctx->vjp_seed);
        const double *st = stage_in(ctx, sigma_t, xs_n(ctx), mem, ctx->sig_t64);
        const double *ss = stage_in(ctx, sigma_s, xs_n(ctx), mem, ctx->sig_s64);
        int its = 0;
        ok = idc_step_device(ctx, psi, src, st, ss, dt, abs_tol, max_its, precision, its);
        if (mem == TSW_MEM_HOST) stage_out(ctx, psi_inout, psi, g.npsi(), mem);
        sync(ctx);
    });
    if (rc == TSW_OK && !ok) {
        ctx->err = "an implicit solve of the IDC step failed to converge";
        return TSW_E_NOCONV;
    }
    return rc;
}

int tsw_idc_run(tsw_ctx *ctx, double *psi_inout, const double *source, const double *sigma_t, const double *sigma_s,
                double dt, int nsteps, double abs_tol, int max_its, int precision, int mem) {
    bool ok = true;
    const int rc = guarded(ctx, [&] {
        require(ctx->reflect_mask == 0 && !ctx->periodic, TSW_E_UNSUPPORTED, "tsw_idc_run: reflecting and periodic boundaries are built for the "
                "sweeps, tsw_si_solve and the halo updates");
        need_problem(ctx);
        require(dt > 0.0 && std::isfinite(dt) && nsteps >= 0, TSW_E_ARG, "IDC needs a finite time step and nsteps >= 0");
        if (precision == TSW_MIXED) precision = TSW_FP32;  // solvers refine their FP32 sweeps themselves
        require(precision == TSW_FP64 || precision == TSW_FP32, TSW_E_ARG, "bad precision");
        if (ctx->idc.nodes.empty()) idc_tables(ctx->idc);
        const Geom &g = ctx->g;
        double *psi = stage_in(ctx, psi_inout, g.npsi(), mem, ctx->q64);
        const double *src = stage_in(ctx, source, g.npsi(), mem, ctx->vjp_seed);
        const double *st = stage_in(ctx, sigma_t, xs_n(ctx), mem, ctx->sig_t64);
        const double *ss = stage_in(ctx, sigma_s, xs_n(ctx), mem, ctx->sig_s64);
        for (int k = 0; k < nsteps && ok; ++k) {
            int its = 0;
            ok = idc_step_device(ctx, psi, src, st, ss, dt, abs_tol, max_its, precision, its);
        }
        if (mem == TSW_MEM_HOST) stage_out(ctx, psi_inout, psi, g.npsi(), mem);
        sync(ctx);
    });
    if (rc == TSW_OK && !ok) {
        ctx->err = "an implicit solve of an IDC step failed to converge";
        return TSW_E_NOCONV;
    }
    return rc;
}

int tsw_idc_adjoint(tsw_ctx *ctx, const double *psi0, const double *source, const double *sigma_t,
                    const double *sigma_s, double dt, int nsteps, int checkpoint_every, const double *seed,
                    double *grad_psi0, double *grad_sigma_t, double *grad_sigma_s, double abs_tol, int max_its,
                    int precision, int mem) {
    bool ok = true;
    const int rc = guarded(ctx, [&] {
        require(ctx->reflect_mask == 0 && !ctx->periodic, TSW_E_UNSUPPORTED, "tsw_idc_adjoint: reflecting and periodic boundaries are built for the "
                "sweeps, tsw_si_solve and the halo updates");
        need_problem(ctx);
        require(ctx->xs_deg == 0, TSW_E_UNSUPPORTED, "vector-Jacobian products take cell-mean cross sections "
                                                     "(tsw_set_xs_degree 0)");
        require(dt > 0.0 && std::isfinite(dt) && nsteps >= 0, TSW_E_ARG, "IDC needs a finite time step and nsteps >= 0");
        require(checkpoint_every >= 1, TSW_E_ARG, "checkpoint_every must be >= 1");
        require(seed && grad_psi0, TSW_E_ARG, "seed and grad_psi0 are required");
        const bool grads = grad_sigma_t || grad_sigma_s;
        require(!grads || psi0, TSW_E_ARG, "cross-section gradients need psi0");
        if (precision == TSW_MIXED) precision = TSW_FP32;  // solvers refine their FP32 sweeps themselves
        require(precision == TSW_FP64 || precision == TSW_FP32, TSW_E_ARG, "bad precision");
        if (ctx->idc.nodes.empty()) idc_tables(ctx->idc);
        const Geom &g = ctx->g;
        hipStream_t s = ctx->stream;
        const size_t n = g.npsi();
        const double *src = stage_in(ctx, source, n, mem, ctx->vjp_seed);
        const double *st = stage_in(ctx, sigma_t, xs_n(ctx), mem, ctx->sig_t64);
        const double *ss = stage_in(ctx, sigma_s, xs_n(ctx), mem, ctx->sig_s64);
        const double *p0 = psi0 ? stage_in(ctx, psi0, n, mem, ctx->vjp_psi) : nullptr;
        double *lam = mem == TSW_MEM_DEVICE ? grad_psi0 : dbuf<double>(ctx->vjp_gsrc, n);
        if (lam != seed) {
            const double *sd = stage_in(ctx, seed, n, mem, ctx->q64);
            TSW_HIP(hipMemcpyAsync(lam, sd, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        }
        auto cellbuf = [&](double *user, DBuf &b) {
            return user ? (mem == TSW_MEM_DEVICE ? user : dbuf<double>(b, g.ncell_h())) : nullptr;
        };
        double *gst = cellbuf(grad_sigma_t, ctx->vjp_gst), *gss = cellbuf(grad_sigma_s, ctx->vjp_gss);
        if (gst) TSW_HIP(hipMemsetAsync(gst, 0, sizeof(double) * g.ncell_h(), s));
        if (gss) TSW_HIP(hipMemsetAsync(gss, 0, sizeof(double) * g.ncell_h(), s));
        auto &I = ctx->idc;
        int its = 0;
        if (grads) {
            // forward pass keeping a checkpoint every `checkpoint_every` steps
            const int C = checkpoint_every, nck = (nsteps + C - 1) / C;
            double *ck = dbuf<double>(I.ckpt, (size_t)std::max(nck, 1) * n);
            double *seg = dbuf<double>(I.seg, (size_t)C * n);
            double *cur = dbuf<double>(ctx->vjp_y, n);  // idc steps use their own buffers, not vjp_y
            TSW_HIP(hipMemcpyAsync(cur, p0, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            for (int k = 0; k < nsteps && ok; ++k) {
                if (k % C == 0) TSW_HIP(hipMemcpyAsync(ck + (size_t)(k / C) * n, cur, sizeof(double) * n,
                                                       hipMemcpyDeviceToDevice, s));
                ok = idc_step_device(ctx, cur, src, st, ss, dt, abs_tol, max_its, precision, its);
            }
            // backward pass: per segment, recompute its states from the checkpoint, then step back
            for (int sg = nck - 1; sg >= 0 && ok; --sg) {
                const int k0 = sg * C, k1 = std::min(k0 + C, nsteps);
                TSW_HIP(hipMemcpyAsync(seg, ck + (size_t)sg * n, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
                for (int k = k0 + 1; k < k1 && ok; ++k) {
                    double *dst = seg + (size_t)(k - k0) * n;
                    TSW_HIP(hipMemcpyAsync(dst, dst - n, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
                    ok = idc_step_device(ctx, dst, src, st, ss, dt, abs_tol, max_its, precision, its);
                }
                for (int k = k1 - 1; k >= k0 && ok; --k)
                    ok = idc_step_vjp_device(ctx, seg + (size_t)(k - k0) * n, src, st, ss, dt, abs_tol, max_its,
                                             precision, lam, gst, gss, its);
            }
        } else {
            // the step is linear in psi: its transpose needs no forward states
            for (int k = nsteps - 1; k >= 0 && ok; --k)
                ok = idc_step_vjp_device(ctx, nullptr, src, st, ss, dt, abs_tol, max_its, precision, lam, nullptr,
                                         nullptr, its);
        }
        if (mem == TSW_MEM_HOST) {
            stage_out(ctx, grad_psi0, lam, n, mem);
            if (gst) stage_out(ctx, grad_sigma_t, gst, g.ncell_h(), mem);
            if (gss) stage_out(ctx, grad_sigma_s, gss, g.ncell_h(), mem);
        }
        sync(ctx);
    });
    if (rc == TSW_OK && !ok) {
        ctx->err = "an implicit solve of the IDC adjoint failed to converge";
        return TSW_E_NOCONV;
    }
    return rc;
}

}  // extern "C"

namespace {
void batch_choose_engines(tsw_ctx *ctx);
void build_batch_groups(tsw_ctx *ctx, const std::vector<int> &members);

}  // namespace

extern "C" {

int tsw_batch_set(tsw_ctx *ctx, int nprob, const tsw_problem *probs, const double *const *sigma_t, double dt,
                  int mem) {
    return guarded(ctx, [&] {
        require(nprob >= 1 && probs && sigma_t, TSW_E_ARG, "empty batch");
        ctx->batch.clear();
        ctx->groups.clear();
        ctx->lane_groups.clear();
        ctx->engine.assign(nprob, tsw_ctx::kEngConverted);
        ctx->lanes_able.assign(nprob, 0);
        ctx->split_able.assign(nprob, 0);
        for (int &r : ctx->split_rows_p) r = 0;
        ctx->batch_all.resize(nprob);
        for (int b = 0; b < nprob; ++b) ctx->batch_all[b] = b;
        ctx->batch_dt = dt;
        for (int b = 0; b < nprob; ++b) {
            const tsw_problem &pr = probs[b];
            std::unique_ptr<tsw_ctx> sub(new tsw_ctx());
            sub->device = ctx->device;
            sub->stream = ctx->stream;
            sub->own_stream = false;
            sub->pattern = Pattern::Staggered;
            sub->band_rows_opt = -1;
            set_problem(sub.get(), pr.nx, pr.ny, pr.ax, pr.bx, pr.ay, pr.by, pr.p, pr.nq, pr.xi, pr.eta, pr.w,
                        pr.quad_off);
            require(staggered_ok(sub.get()), TSW_E_UNSUPPORTED,
                    "batch member " + std::to_string(b) + " does not fit the staggered sweep (ny <= 448)");
            const Geom &g = sub->g;
            double *sig = static_cast<double *>(sub->sig_t64.reserve(sizeof(double) * g.ncell_h()));
            TSW_HIP(hipMemcpyAsync(sig, sigma_t[b], sizeof(double) * g.ncell_h(),
                                   mem == TSW_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
            diag_setup(sub.get(), sig, nullptr, nullptr, dt, true, true);
            // the ordinate-lane engine where it is built for the member (three compute waves and a storer)
            sub->lanes_opt = 1;
            sub->lanes_J_opt = ctx->lanes_J_opt;
            sub->lanes_W_opt = ctx->lanes_W_opt > 0 ? std::min(ctx->lanes_W_opt, kLanesStoreWaves) : kLanesStoreWaves;
            sub->lanes_fixed_W = true;
            sub->lanes_pf_opt = ctx->lanes_pf_opt;
            ctx->lanes_able[b] = ctx->lanes_opt >= 0 && lanes_ok(sub.get()) && sub->mat.valid &&
                                 diag_inverses<double>(sub.get(), dt) && diag_inverses<float>(sub.get(), dt);
            // the lane-split kernel where it is built for the degree and the member's tables fit its LDS
            const int sbr = g.p >= 1 ? split_band_rows(ctx, g.p, g.ny) : 0;
            ctx->split_able[b] = ctx->batch_split_opt >= 0 && g.p >= 1 && sub->mat.valid &&
                                 diag_inverses<double>(sub.get(), dt) && diag_inverses<float>(sub.get(), dt) &&
                                 split_sweep_lds_bytes<double>(g.p, sbr, g.nx, sub->mat.M) != 0 &&
                                 split_sweep_lds_bytes<float>(g.p, sbr, g.nx, sub->mat.M) != 0;
            ctx->batch.push_back(std::move(sub));
        }
        ctx->batch_ref_ok = true;  // 32-bit byte offsets into every member's array
        for (auto &sub : ctx->batch)
            ctx->batch_ref_ok = ctx->batch_ref_ok && sub->g.npsi() * sizeof(double) < (size_t(1) << 32);
        for (int k = 0; k < tsw_ctx::kSideStreams; ++k) {
            if (!ctx->side[k]) TSW_HIP(hipStreamCreateWithFlags(&ctx->side[k], hipStreamNonBlocking));
            if (!ctx->join_ev[k]) TSW_HIP(hipEventCreateWithFlags(&ctx->join_ev[k], hipEventDisableTiming));
        }
        if (!ctx->fork_ev) TSW_HIP(hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming));
        batch_choose_engines(ctx);
        build_batch_groups(ctx, ctx->batch_all);
        sync(ctx);
    });
}

}  // extern "C"

namespace {
// Builds (once) the device parameter table of a launch group for precision T: the sweep of each member's
// source in place with its inflow, or (corr, FP32 only) the zero-inflow sweep of its residual buffer.
// With `ref`, the table sweeps each member's reference-layout array ref[b] in place (inflow traces into the
// member's trace buffers by launch_diag_inflow_batch) and is rebuilt and uploaded for every call, without
// a host wait; G.ref_* receive the launch choices.
template <typename T>
void batch_group_params(tsw_ctx *ctx, tsw_ctx::Group &G, bool corr, const std::vector<double *> *ref = nullptr) {
    bool &built = sizeof(T) == 8 ? G.built64 : (corr ? G.built32c : G.built32);
    if (built && !ref) return;
    std::vector<DiagSweepParams<T>> ps;
    size_t lds = 0;
    for (int b : G.members) {
        tsw_ctx *sub = ctx->batch[b].get();
        const Geom &g = sub->g;
        const size_t nd = (size_t)g.nq * g.N() * g.nx * g.ny;
        DiagSweepParams<T> sp;
        if (ref) {
            sp = make_diag_params<T>(sub, ctx->batch_dt, nullptr, nullptr, false, nullptr);
            sp.qref = (*ref)[b];
            sp.psi_ref = (*ref)[b];
            auto &d = sub->diag;
            const bool f64 = sizeof(T) == 8;
            sp.tin_x = dbuf<T>(f64 ? d.tinx64 : d.tinx32, (size_t)g.nq * g.ny * (g.p + 1));
            sp.tin_y = dbuf<T>(f64 ? d.tiny64 : d.tiny32, (size_t)g.nq * g.nx * (g.p + 1));
        } else {
            T *q = corr ? (T *)dbuf<float>(sub->diag.res32, nd)
                        : (T *)(sizeof(T) == 8 ? sub->diag.q64.ptr : sub->diag.q32.ptr);
            sp = make_diag_params<T>(sub, ctx->batch_dt, q, nullptr, !corr, q);
        }
        if (!G.split) lds = std::max(lds, diag_sweep_lds_bytes<T>(sp, g.p));
        ps.push_back(sp);
    }
    if (G.split) {  // lane-split kernel: material path, row bands of split_band_rows rows
        for (size_t k = 0; k < ps.size(); ++k) {
            tsw_ctx *sub = ctx->batch[G.members[k]].get();
            const Geom &g = sub->g;
            DiagSweepParams<T> &sp = ps[k];
            require(sp.ainv != nullptr, TSW_E_STATE, "lane-split batch member without element inverses");
            sp.split = 1;
            sp.direct = 0;
            sp.wave1 = 0;
            const int br = G.band_rows[k];
            sp.nbands = (g.ny + br - 1) / br;
            sp.band_rows = sp.nbands > 1 ? br : g.ny;
            if (sp.nbands > 1) {
                sp.xtr = lanes_records<T>(sub, sp.nbands);
                sp.abort_flag = abort_flag_of(ctx);  // a stall is reported (and the records refilled) by sync(ctx)
            }
            const size_t b = split_sweep_lds_bytes<T>(g.p, G.rows, g.nx, sp.M);
            require(b != 0, TSW_E_UNSUPPORTED, "lane-split batch group does not fit one CU's LDS");
            lds = std::max(lds, b);
        }
        if (ref) {
            G.ref_direct = false;
            G.ref_stage = false;
            G.ref_mat = true;
        } else {
            (sizeof(T) == 8 ? G.direct64 : G.direct32) = false;
            (sizeof(T) == 8 ? G.mat64 : G.mat32) = true;
        }
    }
    // one launch serves the group: if any member needs the direct-store layout every member uses it;
    // members whose tables do not fit fall back to elimination, and a mixed group eliminates everywhere
    bool any_direct = false;
    for (auto &sp : ps) any_direct = any_direct || (sp.ainv && sp.direct);
    if (!G.split && !any_direct && ctx->batch_direct_opt != -1) {
        // Without the storer wave and its staging rows a workgroup is one wave and 2*rows*NS values
        // smaller, so more of them share a CU: the batched sweep is limited by how many item sweeps run
        // at once (on the 256-problem batch the FP32 kernels take 6.5 ms direct, 8.6 ms with the storer).
        bool all_tables = true;
        for (auto &sp : ps) all_tables = all_tables && sp.ainv != nullptr;
        any_direct = all_tables;
    }
    if (!G.split && any_direct) {
        lds = 0;
        for (size_t k = 0; k < ps.size(); ++k) {
            tsw_ctx *sub = ctx->batch[G.members[k]].get();
            if (!ps[k].ainv) continue;
            ps[k].direct = 1;
            const size_t b = diag_sweep_lds_bytes<T>(ps[k], sub->g.p);
            if (b == 0) ps[k].ainv = nullptr;
            lds = std::max(lds, b);
        }
    }
    bool all_mat = G.mat;
    for (auto &sp : ps) all_mat = all_mat && sp.ainv != nullptr;
    if (!ref && !G.split) {
        (sizeof(T) == 8 ? G.direct64 : G.direct32) = all_mat && any_direct;
        (sizeof(T) == 8 ? G.mat64 : G.mat32) = all_mat;
    }
    if (!all_mat) {
        lds = 0;
        for (size_t k = 0; k < ps.size(); ++k) {
            tsw_ctx *sub = ctx->batch[G.members[k]].get();
            ps[k].ainv = nullptr;
            ps[k].mat = nullptr;
            ps[k].M = 0;
            ps[k].direct = 0;
            ps[k].sig_t = (const T *)(sizeof(T) == 8 ? sub->diag.sigt64.ptr : sub->diag.sigt32.ptr);
            lds = std::max(lds, diag_sweep_lds_bytes<T>(ps[k], sub->g.p));
        }
    }
    require(lds != 0, TSW_E_UNSUPPORTED, "batch group does not fit the staggered sweep");
    if (ref) {
        if (!G.split) {
            G.ref_direct = all_mat && any_direct;
            G.ref_stage = ctx->ref_stage_opt > 0 && ref_stage_layout<T>(ps, G.p, all_mat, lds, G.ref_direct);
        }
        G.ref_host.resize(sizeof(DiagSweepParams<T>) * ps.size());
        std::memcpy(G.ref_host.data(), ps.data(), G.ref_host.size());
        TSW_HIP(hipMemcpyAsync(G.ref_params.reserve(G.ref_host.size()), G.ref_host.data(), G.ref_host.size(),
                               hipMemcpyHostToDevice, ctx->stream));
        G.ref_lds = lds;
        G.ref_mat = all_mat;
        return;
    }
    DBuf &buf = sizeof(T) == 8 ? G.params64 : (corr ? G.params32c : G.params32);
    TSW_HIP(hipMemcpyAsync(buf.reserve(sizeof(DiagSweepParams<T>) * ps.size()), ps.data(),
                           sizeof(DiagSweepParams<T>) * ps.size(), hipMemcpyHostToDevice, ctx->stream));
    TSW_HIP(hipStreamSynchronize(ctx->stream));  // ps goes out of scope
    (sizeof(T) == 8 ? G.lds64 : (corr ? G.lds32c : G.lds32)) = lds;
    built = true;
}

// The lane groups' parameter tables for one batched sweep of lane members b in place over arr[b] (their
// inflow from arr[b]'s halos), uploaded on the context's stream; threads / lds per group.
template <typename T>
void lanes_batch_params(tsw_ctx *ctx, const std::vector<double *> &arr, std::vector<int> &threads,
                        std::vector<size_t> &lds) {
    threads.assign(ctx->lane_groups.size(), 0);
    lds.assign(ctx->lane_groups.size(), 0);
    for (size_t k = 0; k < ctx->lane_groups.size(); ++k) {
        auto &LG = ctx->lane_groups[k];
        LG.host.resize(sizeof(LanesParams<T>) * LG.members.size());
        LanesParams<T> *ps = reinterpret_cast<LanesParams<T> *>(LG.host.data());
        for (size_t slot = 0; slot < LG.members.size(); ++slot) {
            const int b = LG.members[slot];
            int th = 0;
            size_t l = 0;
            ps[slot] = lanes_params<T>(ctx->batch[b].get(), ctx->batch_dt, arr[b], nullptr, nullptr, arr[b], nullptr,
                                       false, th, l);
            if (ps[slot].nbands > 1) ps[slot].abort_flag = abort_flag_of(ctx);  // reported by sync(ctx)
            require(threads[k] == 0 || threads[k] == th, TSW_E_STATE, "lane batch members of unequal workgroups");
            threads[k] = th;
            lds[k] = std::max(lds[k], l);
        }
        TSW_HIP(hipMemcpyAsync(LG.params.reserve(LG.host.size()), LG.host.data(), LG.host.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    }
}

// The inflow traces of lane group k's members from their arrays' halos (before the side streams fork).
template <typename T>
void lanes_batch_inflow(tsw_ctx *ctx, size_t k, hipStream_t s) {
    auto &LG = ctx->lane_groups[k];
    const LanesParams<T> *dp = LG.params.as<LanesParams<T>>();
    const int n = (int)LG.members.size();
    switch (LG.p) {
        case 0: launch_lanes_inflow_batch_p<T, 1>(dp, n, s); break;
        case 1: launch_lanes_inflow_batch_p<T, 2>(dp, n, s); break;
        case 2: launch_lanes_inflow_batch_p<T, 3>(dp, n, s); break;
        default: throw Error(TSW_E_UNSUPPORTED, "ordinate-lane sweep built for DG degrees 0-2");
    }
}

template <typename T>
void lanes_batch_launch(tsw_ctx *ctx, size_t k, int threads, size_t lds, hipStream_t s) {
    auto &LG = ctx->lane_groups[k];
    const LanesParams<T> *dp = LG.params.as<LanesParams<T>>();
    LanesParams<T> rep = reinterpret_cast<const LanesParams<T> *>(LG.host.data())[0];  // the mode
    // persistent launch: the counter was zeroed on the context's stream before the fork
    rep.work = ctx->batch_persistent_opt > 0 && LG.work.ptr ? LG.work.as<int>() : nullptr;
    lanes_launch<T>(LG.p, rep, threads, lds, s, dp, LG.dev_items.as<int4>(), (int)LG.items.size());
}

// Launch groups of the members `members` by their engines: staggered groups by (DG degree, compute waves,
// material path, layout), work items {member slot, ordinate} heaviest first; lane groups by DG degree, items
// {member slot, ordinate block, band} band-major, heaviest members first within a band.
void build_batch_groups(tsw_ctx *ctx, const std::vector<int> &members) {
    ctx->groups.clear();
    ctx->lane_groups.clear();
    std::map<std::tuple<int, int, bool, int>, int> gid;
    std::map<int, int> lid;
    for (int b : members) {
        tsw_ctx *sub = ctx->batch[b].get();
        if (ctx->engine[b] == tsw_ctx::kEngLanes) {
            auto it = lid.find(sub->g.p);
            if (it == lid.end()) {
                it = lid.emplace(sub->g.p, (int)ctx->lane_groups.size()).first;
                ctx->lane_groups.emplace_back();
                ctx->lane_groups.back().p = sub->g.p;
                ctx->lane_groups.back().work.reserve(sizeof(int));
            }
            ctx->lane_groups[it->second].members.push_back(b);
            continue;
        }
        const int e = ctx->engine[b];
        const bool split = e == tsw_ctx::kEngSplit;
        // staggered groups hold members of equal waves per workgroup, split groups of equal row slots
        const int rpw = split ? 64 / split_lanes_of(sub->g.p) : 64;
        const int br = split ? split_band_rows(ctx, sub->g.p, sub->g.ny) : sub->g.ny;
        const auto key = std::make_tuple(sub->g.p, (br + rpw - 1) / rpw, sub->mat.valid, e);
        auto it = gid.find(key);
        if (it == gid.end()) {
            it = gid.emplace(key, (int)ctx->groups.size()).first;
            ctx->groups.emplace_back();
            auto &G = ctx->groups.back();
            G.p = sub->g.p;
            G.waves = (br + rpw - 1) / rpw;
            G.mat = sub->mat.valid;
            G.ref = e == tsw_ctx::kEngRef || split;
            G.split = split;
            if (split) G.work.reserve(sizeof(int));
        }
        auto &G = ctx->groups[it->second];
        G.members.push_back(b);
        G.band_rows.push_back(br);
        G.max_ny = std::max(G.max_ny, sub->g.ny);
        G.rows = std::max(G.rows, br);
    }
    for (auto &G : ctx->groups) {
        // (-work, slot, band * nq + ordinate), band-major: every band's upstream band is dispatched first
        std::vector<std::tuple<int, long long, int, int>> w;
        for (int slot = 0; slot < (int)G.members.size(); ++slot) {
            const Geom &g = ctx->batch[G.members[slot]]->g;
            const int br = G.band_rows[slot];
            const int nb = (g.ny + br - 1) / br;
            for (int band = 0; band < nb; ++band)
                for (int q = 0; q < g.nq; ++q)
                    w.emplace_back(band, -(long long)(g.nx + g.ny) * g.ny, slot, band * g.nq + q);
        }
        std::sort(w.begin(), w.end());
        const int T = G.split ? ctx->batch_xcd_tile_opt : 0;
        if (T <= 1) {
            for (auto &t : w) {
                G.items.push_back(std::get<2>(t));
                G.items.push_back(std::get<3>(t));
            }
        } else {
            // Within each band's run of items (so the order stays band-major), runs of T consecutive items
            // (neighbouring ordinates of one member, whose cells share cache lines in the caller's layout)
            // go to one XCD and the runs go round the XCDs: workgroup k of the run (padded to a multiple
            // of kXcds * T with empty items) takes item ((m / T) * kXcds + k % kXcds) * T + m % T, m = k / kXcds.
            for (size_t s0 = 0; s0 < w.size();) {
                size_t s1 = s0;
                while (s1 < w.size() && std::get<0>(w[s1]) == std::get<0>(w[s0])) ++s1;
                const size_t L = s1 - s0, C = (size_t)kXcds * T, Lp = (L + C - 1) / C * C;
                for (size_t k = 0; k < Lp; ++k) {
                    const size_t m = k / kXcds, r = k % kXcds, j = ((m / T) * kXcds + r) * T + m % T;
                    G.items.push_back(j < L ? std::get<2>(w[s0 + j]) : -1);
                    G.items.push_back(j < L ? std::get<3>(w[s0 + j]) : 0);
                }
                s0 = s1;
            }
        }
        TSW_HIP(hipMemcpyAsync(G.dev_items.reserve(sizeof(int) * G.items.size()), G.items.data(),
                               sizeof(int) * G.items.size(), hipMemcpyHostToDevice, ctx->stream));
    }
    for (auto &LG : ctx->lane_groups) {
        std::vector<std::tuple<long long, int>> order;  // (-work, slot)
        std::vector<int> nblk(LG.members.size()), nbands(LG.members.size());
        int max_bands = 0;
        for (int slot = 0; slot < (int)LG.members.size(); ++slot) {
            tsw_ctx *sub = ctx->batch[LG.members[slot]].get();
            const std::pair<int, int> shape = lanes_shape(sub);
            lanes_blocks(sub, shape.first);
            const int rows = shape.second * (64 / shape.first);
            nblk[slot] = sub->lanes.nblk;
            nbands[slot] = (sub->g.ny + rows - 1) / rows;
            max_bands = std::max(max_bands, nbands[slot]);
            order.emplace_back(-(long long)(sub->g.nx + sub->g.ny) * sub->g.nq * sub->g.N(), slot);
        }
        std::sort(order.begin(), order.end());
        for (int band = 0; band < max_bands; ++band)
            for (auto &o : order) {
                const int slot = std::get<1>(o);
                if (band < nbands[slot])
                    for (int bi = 0; bi < nblk[slot]; ++bi) LG.items.push_back(int4{slot, bi, band, 0});
            }
        TSW_HIP(hipMemcpyAsync(LG.dev_items.reserve(sizeof(int4) * LG.items.size()), LG.items.data(),
                               sizeof(int4) * LG.items.size(), hipMemcpyHostToDevice, ctx->stream));
    }
    TSW_HIP(hipStreamSynchronize(ctx->stream));  // the item vectors may be rebuilt by the next call
}

// Launch of one staggered group in precision T, whose parameter table batch_group_params has built and
// uploaded on the context's stream before the side streams forked from it: `ref` sweeps the members'
// arrays dev[b] in place (their inflow traces computed before the fork, batch_group_inflow).
// The inflow traces of a group on the caller's layout from its members' halos, on the context's stream
// before the side streams fork (so no side stream waits on them behind another stream's sweeps).
template <typename T>
void batch_group_inflow(tsw_ctx *ctx, tsw_ctx::Group &G) {
    launch_diag_inflow_batch<T>(G.ref_params.as<DiagSweepParams<T>>(), (int)G.members.size(), G.p, ctx->stream);
}

template <typename T>
void batch_group_launch(tsw_ctx *ctx, tsw_ctx::Group &G, bool corr, const std::vector<double *> *ref, hipStream_t sd) {
    DiagSweepParams<T> rep{};  // representative parameters for the host-side dispatch
    const DiagSweepParams<T> *probs;
    size_t lds;
    if (ref) {
        probs = G.ref_params.as<DiagSweepParams<T>>();
        rep.qref = reinterpret_cast<const double *>(1);
        rep.ainv = G.ref_mat ? reinterpret_cast<const T *>(1) : nullptr;
        rep.direct = G.ref_direct ? 1 : 0;
        rep.ref_stage = G.ref_stage ? 1 : 0;
        lds = G.ref_lds;
    } else {
        probs = (const DiagSweepParams<T> *)(sizeof(T) == 8 ? G.params64.ptr : (corr ? G.params32c.ptr : G.params32.ptr));
        rep.qsrc = reinterpret_cast<const T *>(1);
        rep.ainv = (sizeof(T) == 8 ? G.mat64 : G.mat32) ? reinterpret_cast<const T *>(1) : nullptr;
        rep.direct = (sizeof(T) == 8 ? G.direct64 : G.direct32) ? 1 : 0;
        lds = sizeof(T) == 8 ? G.lds64 : (corr ? G.lds32c : G.lds32);
    }
    rep.prefetch = ctx->prefetch_opt;
    if (G.split) {
        rep.split = 1;
        rep.prefetch = ctx->batch_split_pf_opt > 0 ? ctx->batch_split_pf_opt : kSplitBatchPrefetch;
        // persistent launch: the counter was zeroed on the context's stream before the fork
        if (ctx->batch_persistent_opt > 0 && G.work.ptr) rep.work = G.work.as<int>();
    }
    launch_diag_sweep_batch<T>(rep, probs, G.dev_items.as<int2>(), (int)G.items.size() / 2,
                               G.split ? G.rows : G.max_ny, lds, G.p, sd);
}

// One batched sweep over the current groups, all on the side streams, timed as one sweep.  Converted
// groups sweep their diagonal-layout buffers in FP64 (conv64) or FP32 (corr: the FP32 correction sweeps of
// TSW_MIXED, converted groups only); groups on the caller's layout (lane groups, and staggered groups
// with `ref_call`) sweep dev[b] in place in FP64 (inplace64) or FP32.
void batch_sweep(tsw_ctx *ctx, bool conv64, bool inplace64, bool corr, const std::vector<double *> &dev, bool ref_call) {
    std::vector<int> lthreads;
    std::vector<size_t> llds;
    const bool lanes = !corr && !ctx->lane_groups.empty();
    if (lanes && inplace64) lanes_batch_params<double>(ctx, dev, lthreads, llds);
    else if (lanes) lanes_batch_params<float>(ctx, dev, lthreads, llds);
    std::vector<std::pair<tsw_ctx::Group *, bool>> diag;  // (group, on the caller's layout)
    for (auto &G : ctx->groups) {
        const bool in_place = G.ref && ref_call;
        if (!corr || !in_place) diag.emplace_back(&G, in_place);
    }
    if (diag.empty() && !lanes) return;
    // every parameter table is on the device (uploads on ctx->stream) before the fork below
    for (auto &gd : diag) {
        if (gd.second) {
            if (inplace64) batch_group_params<double>(ctx, *gd.first, false, &dev);
            else batch_group_params<float>(ctx, *gd.first, false, &dev);
        } else if (conv64 && !corr) {
            batch_group_params<double>(ctx, *gd.first, false);
        } else {
            batch_group_params<float>(ctx, *gd.first, corr);
        }
    }
    if (ctx->stall_test) {  // tuning key stall_test: every bounded wait of a banded lane-split group gives up
        bool banded = false;
        for (auto &gd : diag)
            for (size_t m = 0; gd.first->split && m < gd.first->members.size(); ++m)
                banded = banded || gd.first->band_rows[m] < ctx->batch[gd.first->members[m]]->g.ny;
        if (banded) {
            ctx->stall_test = false;
            int *flag = abort_flag_of(ctx);
            *ctx->abort_host = 1;
            TSW_HIP(hipMemcpyAsync(flag, ctx->abort_host, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        }
    }
    hipEvent_t e0 = next_event(ctx), e1 = next_event(ctx);
    TSW_HIP(hipEventRecord(e0, ctx->stream));
    // every inflow trace of the sweeps on the caller's layout, before the fork
    if (lanes)
        for (size_t j = 0; j < ctx->lane_groups.size(); ++j) {
            if (inplace64) lanes_batch_inflow<double>(ctx, j, ctx->stream);
            else lanes_batch_inflow<float>(ctx, j, ctx->stream);
        }
    for (auto &gd : diag)
        if (gd.second) {
            if (inplace64) batch_group_inflow<double>(ctx, *gd.first);
            else batch_group_inflow<float>(ctx, *gd.first);
        }
    if (ctx->batch_persistent_opt > 0) {  // the persistent launches' item counters start at zero
        if (lanes)
            for (auto &LG : ctx->lane_groups)
                if (LG.work.ptr) TSW_HIP(hipMemsetAsync(LG.work.ptr, 0, sizeof(int), ctx->stream));
        for (auto &gd : diag)
            if (gd.first->split && gd.first->work.ptr)
                TSW_HIP(hipMemsetAsync(gd.first->work.ptr, 0, sizeof(int), ctx->stream));
    }
    TSW_HIP(hipEventRecord(ctx->fork_ev, ctx->stream));
    int k = 0;
    if (lanes)
        for (size_t j = 0; j < ctx->lane_groups.size(); ++j, ++k) {
            hipStream_t sd = ctx->side[k % tsw_ctx::kSideStreams];
            TSW_HIP(hipStreamWaitEvent(sd, ctx->fork_ev, 0));
            if (inplace64) lanes_batch_launch<double>(ctx, j, lthreads[j], llds[j], sd);
            else lanes_batch_launch<float>(ctx, j, lthreads[j], llds[j], sd);
        }
    for (auto &gd : diag) {
        hipStream_t sd = ctx->side[k % tsw_ctx::kSideStreams];
        TSW_HIP(hipStreamWaitEvent(sd, ctx->fork_ev, 0));
        if (gd.second) {
            if (inplace64) batch_group_launch<double>(ctx, *gd.first, false, &dev, sd);
            else batch_group_launch<float>(ctx, *gd.first, false, &dev, sd);
        } else if (conv64 && !corr) {
            batch_group_launch<double>(ctx, *gd.first, false, nullptr, sd);
        } else {
            batch_group_launch<float>(ctx, *gd.first, corr, nullptr, sd);
        }
        ++k;
    }
    for (int j = 0; j < std::min<int>(k, tsw_ctx::kSideStreams); ++j) {
        TSW_HIP(hipEventRecord(ctx->join_ev[j], ctx->side[j]));
        TSW_HIP(hipStreamWaitEvent(ctx->stream, ctx->join_ev[j], 0));
    }
    TSW_HIP(hipEventRecord(e1, ctx->stream));
    ++ctx->sweeps_done;
}

// One batched sweep of the members `members` over their device arrays dev[b] (reference layout, in place),
// by the current groups.  Members on the caller's layout (lane members; reference-layout members when
// every array is 16-byte aligned, else they go through the conversions for this call) take the FP64 sweep
// under TSW_MIXED: those sweeps read and write FP64 in either precision and their FP32 arithmetic is not
// faster, so a refinement would only add a residual and a second sweep.  Converted members: inflow traces
// and the source to the diagonal layout, the sweep (FP32 with the FP64 refinement of mixed_refine_store
// under TSW_MIXED), the result back.
void batch_apply(tsw_ctx *ctx, const std::vector<int> &members, std::vector<double *> &dev, int precision) {
    bool aligned = true;
    for (int b : members) aligned = aligned && ((uintptr_t)dev[b] & 15) == 0;
    bool any_ref = false;
    for (auto &G : ctx->groups) any_ref = any_ref || G.ref;
    const bool ref_call = aligned && any_ref;
    ctx->last_batch_ref = ref_call;
    auto converted = [&](int b) {
        const char e = ctx->engine[b];
        return e == tsw_ctx::kEngConverted || ((e == tsw_ctx::kEngRef || e == tsw_ctx::kEngSplit) && !ref_call);
    };
    const bool conv64 = precision == TSW_FP64, mixed = precision == TSW_MIXED;
    const bool inplace64 = precision != TSW_FP32;
    // per-member layout conversions spread over the side streams, joined before the next phase
    auto on_side_streams = [&](auto &&fn) {
        bool any = false;
        for (int b : members) any = any || converted(b);
        if (!any) return;
        TSW_HIP(hipEventRecord(ctx->fork_ev, ctx->stream));
        for (auto sd : ctx->side) TSW_HIP(hipStreamWaitEvent(sd, ctx->fork_ev, 0));
        int k = 0;
        for (int b : members) {
            if (!converted(b)) continue;
            tsw_ctx *sub = ctx->batch[b].get();
            sub->stream = ctx->side[k++ % tsw_ctx::kSideStreams];
            try {
                fn(b, sub);
            } catch (...) {
                sub->stream = ctx->stream;
                throw;
            }
            sub->stream = ctx->stream;
        }
        for (int j = 0; j < tsw_ctx::kSideStreams; ++j) {
            TSW_HIP(hipEventRecord(ctx->join_ev[j], ctx->side[j]));
            TSW_HIP(hipStreamWaitEvent(ctx->stream, ctx->join_ev[j], 0));
        }
    };
    auto nd_of = [](const tsw_ctx *sub) { return (size_t)sub->g.nq * sub->g.N() * sub->g.nx * sub->g.ny; };
    on_side_streams([&](int b, tsw_ctx *sub) { diag_load_source(sub, dev[b], conv64 || mixed, !conv64); });
    batch_sweep(ctx, conv64, inplace64, false, dev, ref_call);
    const bool fused_store = mixed && ctx->refine_steps == 1;
    auto residual = [&](tsw_ctx *sub, const auto *psi) {
        auto &d = sub->diag;
        launch_residual_diag<float>(sub->g, d.q64.as<double>(), nullptr, nullptr, d.sigt64.as<double>(), psi,
                                    d.tinx64.as<double>(), d.tiny64.as<double>(), sub->xi64.as<double>(),
                                    sub->eta64.as<double>(), inverse_dt(ctx->batch_dt),
                                    dbuf<float>(d.res32, nd_of(sub)), sub->stream);
    };
    if (fused_store) {
        on_side_streams([&](int, tsw_ctx *sub) { residual(sub, sub->diag.q32.as<float>()); });
        batch_sweep(ctx, false, inplace64, true, dev, ref_call);
    } else if (mixed) {
        on_side_streams([&](int, tsw_ctx *sub) {
            launch_convert_f2d(sub->diag.q32.as<float>(), dbuf<double>(sub->diag.ir64, nd_of(sub)), nd_of(sub),
                               sub->stream);
        });
        for (int k = 0; k < ctx->refine_steps; ++k) {
            on_side_streams([&](int, tsw_ctx *sub) {
                residual(sub, static_cast<const double *>(sub->diag.ir64.as<double>()));
            });
            batch_sweep(ctx, false, inplace64, true, dev, ref_call);
            on_side_streams([&](int, tsw_ctx *sub) {
                launch_accumulate_f2d(sub->diag.res32.as<float>(), sub->diag.ir64.as<double>(), nd_of(sub), sub->stream);
            });
        }
    }
    on_side_streams([&](int b, tsw_ctx *sub) {
        if (conv64) diag_store<double>(sub, sub->diag.q64.as<double>(), dev[b]);
        else if (fused_store)
            launch_psi_store_sum(sub->g, dev[b], sub->diag.q32.as<float>(), sub->diag.res32.as<float>(), sub->stream);
        else if (mixed) diag_store<double>(sub, sub->diag.ir64.as<double>(), dev[b]);
        else diag_store<float>(sub, sub->diag.q32.as<float>(), dev[b]);
    });
}

// Autotuning of the batch engines, per DG degree: each candidate (staggered kernels through the layout
// conversions or on the caller's layout, one ordinate per wavefront; the ordinate-lane engine, J ordinates
// per wavefront) sweeps the degree's members, or the largest ones that fit a scratch budget (4 GiB, at most
// a quarter of the free device memory), twice
// in FP64 on zero-source scratch arrays, and the second sweep is timed.  Every engine gives the same bits,
// so the choice only affects speed.  Without the autotuner (or with an engine forced by the tuning keys
// "lanes" and "batch_ref") the lane engine takes the degrees it is built for and the staggered kernels on
// the caller's layout the others.
void batch_choose_engines(tsw_ctx *ctx) {
    const int nprob = (int)ctx->batch.size();
    for (int p = 0; p < 4; ++p) ctx->batch_engine[p] = -1;
    for (int p = 0; p < 4; ++p) {
        std::vector<int> mem;
        bool all_lanes = true, all_split = ctx->batch_ref_ok;
        for (int b = 0; b < nprob; ++b)
            if (ctx->batch[b]->g.p == p) {
                mem.push_back(b);
                all_lanes = all_lanes && ctx->lanes_able[b];
                all_split = all_split && ctx->split_able[b];
            }
        if (mem.empty()) continue;
        std::vector<char> cand;
        if (ctx->batch_split_opt > 0 && all_split) cand = {tsw_ctx::kEngSplit};
        else if (ctx->lanes_opt > 0 && all_lanes) cand = {tsw_ctx::kEngLanes};
        else if (ctx->batch_ref_opt > 0 && ctx->batch_ref_ok) cand = {tsw_ctx::kEngRef};
        else {
            if (ctx->batch_split_opt == 0 && all_split) cand.push_back(tsw_ctx::kEngSplit);
            if (ctx->lanes_opt == 0 && all_lanes) cand.push_back(tsw_ctx::kEngLanes);
            if (ctx->batch_ref_opt == 0 && ctx->batch_ref_ok) cand.push_back(tsw_ctx::kEngRef);
            if (ctx->batch_ref_opt <= 0 && ctx->lanes_opt <= 0 && ctx->batch_split_opt <= 0)
                cand.push_back(tsw_ctx::kEngConverted);
            if (cand.empty()) cand.push_back(tsw_ctx::kEngConverted);
        }
        // the lane-split engine is tried with the most rows one workgroup holds and with 64-row bands
        // (smaller workgroups, more of them per CU where the tables leave room), unless rows are set
        std::vector<std::pair<char, int>> tries;
        for (char e : cand) {
            tries.emplace_back(e, 0);
            if (e == tsw_ctx::kEngSplit && ctx->batch_split_rows_opt == 0) tries.emplace_back(e, 64);
        }
        char best = cand[0];
        int best_rows = 0;
        ctx->split_rows_p[p] = 0;
        if (tries.size() > 1 && ctx->autotune) {
            // the heaviest members up to the scratch budget
            std::sort(mem.begin(), mem.end(),
                      [&](int a, int b) { return ctx->batch[a]->g.npsi() > ctx->batch[b]->g.npsi(); });
            std::vector<int> sub_members;
            size_t bytes = 0, free_b = 0, total_b = 0;
            TSW_HIP(hipMemGetInfo(&free_b, &total_b));
            const size_t budget = std::min(size_t(4) << 30, free_b / 4);  // leave room for the engines' buffers
            for (int b : mem) {
                const size_t nb = ctx->batch[b]->g.npsi() * sizeof(double);
                if (!sub_members.empty() && bytes + nb > budget) continue;
                sub_members.push_back(b);
                bytes += nb;
            }
            DBuf scratch;
            char *base = static_cast<char *>(scratch.reserve(bytes + 256 * sub_members.size()));
            TSW_HIP(hipMemsetAsync(base, 0, bytes + 256 * sub_members.size(), ctx->stream));
            std::vector<double *> dev(nprob, nullptr);
            size_t off = 0;
            for (int b : sub_members) {
                dev[b] = reinterpret_cast<double *>(base + off);
                off += (ctx->batch[b]->g.npsi() * sizeof(double) + 255) / 256 * 256;
            }
            double best_ms = 0.0;
            for (size_t t = 0; t < tries.size(); ++t) {
                const char e = tries[t].first;
                ctx->split_rows_p[p] = tries[t].second;
                for (int b : sub_members) ctx->engine[b] = e;
                build_batch_groups(ctx, sub_members);
                batch_apply(ctx, sub_members, dev, TSW_FP64);  // first touch of tables and buffers
                hipEvent_t t0 = next_event(ctx), t1 = next_event(ctx);
                TSW_HIP(hipEventRecord(t0, ctx->stream));
                batch_apply(ctx, sub_members, dev, TSW_FP64);
                TSW_HIP(hipEventRecord(t1, ctx->stream));
                TSW_HIP(hipEventSynchronize(t1));
                float ms = 0.f;
                TSW_HIP(hipEventElapsedTime(&ms, t0, t1));
                if (t == 0 || ms < best_ms) {
                    best_ms = ms;
                    best = e;
                    best_rows = tries[t].second;
                }
            }
            ctx->split_rows_p[p] = best_rows;
            sync(ctx);
            ctx->ev_used = 0;  // the tuning sweeps are not counted
            ctx->sweeps_done = 0;
            ctx->ms_folded = 0.0;
        }
        for (int b : mem) {
            ctx->engine[b] = best;
            if (best != tsw_ctx::kEngConverted) {  // diagonal-layout buffers of the tuning sweeps
                auto &d = ctx->batch[b]->diag;
                d.q64.release();
                d.q32.release();
                d.res32.release();
                d.ir64.release();
            }
        }
        ctx->batch_engine[p] = best;
    }
    ctx->groups.clear();
    ctx->lane_groups.clear();
}
}  // namespace

extern "C" {

int tsw_batch_sweep(tsw_ctx *ctx, double *const *psi, int precision, int mem) {
    return guarded(ctx, [&] {
        require(!ctx->batch.empty(), TSW_E_STATE, "tsw_batch_set has not been called");
        require(psi != nullptr, TSW_E_ARG, "null psi array list");
        require(precision == TSW_FP64 || precision == TSW_FP32 || precision == TSW_MIXED, TSW_E_ARG,
                "bad precision");
        std::vector<double *> dev(ctx->batch.size());
        for (size_t b = 0; b < ctx->batch.size(); ++b) {
            require(psi[b] != nullptr, TSW_E_ARG, "null psi array");
            tsw_ctx *sub = ctx->batch[b].get();
            dev[b] = stage_in(sub, psi[b], sub->g.npsi(), mem, sub->psi64);
        }
        batch_apply(ctx, ctx->batch_all, dev, precision);
        if (mem == TSW_MEM_HOST)
            for (size_t b = 0; b < ctx->batch.size(); ++b) {
                tsw_ctx *sub = ctx->batch[b].get();
                stage_out(sub, psi[b], dev[b], sub->g.npsi(), mem);
            }
        sync(ctx);
    });
}

int tsw_sweep_timing(tsw_ctx *ctx, int reset, long long *n_sweeps, double *total_ms) {
    return guarded(ctx, [&] {
        sync(ctx);
        const double tot = ctx->ms_folded + recorded_ms(ctx);
        if (n_sweeps) *n_sweeps = ctx->sweeps_done;
        if (total_ms) *total_ms = tot;
        if (reset) {
            ctx->ev_used = 0;
            ctx->sweeps_done = 0;
            ctx->ms_folded = 0.0;
        }
    });
}

}  // extern "C"
