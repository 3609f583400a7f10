// voce_serve.cpp -- FE^2-style request server over resident RVEs.
//
// Every registered id is a context whose state lives on the device for the
// server's lifetime; its per-increment workspace (spectra, x-pass sums,
// block partials, pre-update stress) is taken from a pool only while one of
// its increments is in flight.  One scheduling round (voce_server_step):
//   1. admit queued requests in arrival order: the id must be idle, a slot
//      free and the workspace must fit in the pool (the first request that
//      does not fit stops admission, so large requests are not starved);
//   2. start the admitted increments (new velocity gradient, begin kernel);
//   3. launch one chunk of FFT iterations over everything in flight, grouped
//      by grid size and precision, through block maps into one descriptor
//      table (rows: ids, then audit shadows);
//   4. poll the in-flight rows and close out the finished ones (skew pass,
//      close-out, increment finaliser), copy their macro state to the host.
// A member's arithmetic does not depend on what else is in flight: its
// blocks, its partial sums and its launch choices are the same alone or in
// a mix.  An audited request also runs from a copy of the id's state in FP64
// (a shadow row aliasing the id's constant tables); the shadow's macro
// stress is compared with the served one and then discarded.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <deque>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/voce_batch.h"
#include "voce_host.h"

namespace {

double mono_now()
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// first-fit allocator over one device buffer
struct Pool {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    std::map<size_t, size_t> free_;   // offset -> length
    void init(char *b, size_t c)
    {
        base = b;
        cap = c;
        used = 0;
        free_.clear();
        if (c) free_[0] = c;
    }
    bool alloc(size_t n, size_t *off)
    {
        n = al256(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const size_t rest = it->second - n, o2 = it->first + n;
                free_.erase(it);
                if (rest) free_[o2] = rest;
                used += n;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t n)
    {
        n = al256(n);
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Request {
    long tag;
    int id;
    int audit;
    double L[9];
    double t_submit;
};

// per-increment workspace of one row, carved from the pool
struct Ws {
    size_t off = 0, len = 0;
    bool held = false;
};

// FP64 copy of a row's evolving state, for an audit
struct Shadow {
    int row = -1;            // table row
    int src = -1;            // the audited id
    char *arena = nullptr;   // own device memory
    size_t bytes = 0;
    long V = 0;
    int nx = 0, ny = 0, nz = 0, two = 0;
    bool busy = false;
};

struct Inflight {
    int row;
    long seq;                // admission sequence number (pairs a row with its shadow)
    Request req;
    double t_admit;
    int shadow;              // index into shadows, or -1
    int launched = 0;        // FFT iterations launched since admission
};

struct PendingAudit {       // a finished served row waiting for its shadow (or vice versa)
    voce_result res;
    bool main_done = false, shadow_done = false;
    MacroState shadow_ms;
};

}  // namespace

struct voce_server {
    voce_server_opts o;
    hipStream_t stream = nullptr;
    int rows = 0, max_shadow = 0;
    RveDev *htab = nullptr;                 // pinned host mirror of the table
    RveDev *dtab = nullptr;
    std::vector<voce_ctx *> ctx;            // by id
    std::vector<Ws> ws;                     // by row
    std::vector<int> busy;                  // by row
    std::vector<int> pred;                  // by id: iterations its last increment took
    Pool pool;
    char *pool_mem = nullptr;
    std::vector<Shadow> shadows;
    std::deque<Request> queue;
    std::vector<Inflight> inflight;
    std::vector<voce_result> done;
    std::map<long, PendingAudit> audits;    // by admission sequence number (tags need not be unique)
    long next_seq = 0;
    int fail_shadow_alloc = 0;              // test hook: fail the n-th next shadow allocation (0: never)
    std::map<std::tuple<int, int, int>, voce_tune_choice> tuned;
    // device scratch: maps, poll indices and flags; pinned staging
    int2 *d_maps = nullptr;
    size_t maps_cap = 0;
    int2 *h_maps = nullptr;
    int *d_idx = nullptr, *d_flags = nullptr, *h_idx = nullptr, *h_flags = nullptr;
    double *h_vg = nullptr;                 // 9 doubles per row
    int *h_one = nullptr;                   // constants for small H2D writes
    MacroState *h_ms = nullptr;             // per row
    size_t idx_cap = 0;
    std::vector<LaunchSet> iter_sets;
    bool set_changed = true;
    long resident_bytes = 0, steps = 0, iters = 0, n_audits = 0, n_done = 0;
};

#define SCK(x)                                                                        \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            fprintf(stderr, "voce server: %s failed: %s\n", #x, hipGetErrorString(e_)); \
            return VOCE_ERR_DEVICE;                                                   \
        }                                                                             \
    } while (0)

extern "C" void voce_server_destroy(voce_server *s)
{
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (voce_ctx *c : s->ctx)
        if (c) voce_ctx_destroy(c);
    for (Shadow &sh : s->shadows)
        if (sh.arena) (void)hipFree(sh.arena);
    void *dev[] = {s->dtab, s->pool_mem, s->d_maps, s->d_idx, s->d_flags};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    void *host[] = {s->htab, s->h_maps, s->h_idx, s->h_flags, s->h_vg, s->h_one, s->h_ms};
    for (void *p : host)
        if (p) (void)hipHostFree(p);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

extern "C" int voce_server_create(const voce_server_opts *opts, voce_server **out)
{
    if (!out || !opts) return VOCE_ERR_ARG;
    *out = nullptr;
    if (opts->max_rves <= 0 || (opts->precision != VOCE_FP64 && opts->precision != VOCE_MIXED)) return VOCE_ERR_ARG;
    voce_server *s = new voce_server();
    s->o = *opts;
    if (s->o.chunk <= 0) s->o.chunk = 16;
    s->max_shadow = 16;
    s->rows = s->o.max_rves + s->max_shadow;
    int rc = 0;
#define FAIL(code)              \
    do {                        \
        rc = (code);            \
        voce_server_destroy(s); \
        return rc;              \
    } while (0)
    if (hipStreamCreate(&s->stream) != hipSuccess) FAIL(VOCE_ERR_DEVICE);
    if (hipHostMalloc((void **)&s->htab, s->rows * sizeof(RveDev), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&s->h_vg, s->rows * 9 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&s->h_one, 16 * sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&s->h_ms, s->rows * sizeof(MacroState), hipHostMallocDefault) != hipSuccess)
        FAIL(VOCE_ERR_NOMEM);
    memset(s->htab, 0, s->rows * sizeof(RveDev));
    for (int i = 0; i < 16; ++i) s->h_one[i] = (i == 1) ? 1 : 0;
    if (hipMalloc((void **)&s->dtab, s->rows * sizeof(RveDev)) != hipSuccess) FAIL(VOCE_ERR_NOMEM);
    s->ctx.assign(s->o.max_rves, nullptr);
    s->pred.assign(s->o.max_rves, 0);
    s->ws.assign(s->rows, Ws());
    s->busy.assign(s->rows, 0);
    s->shadows.resize(s->max_shadow);
    for (int i = 0; i < s->max_shadow; ++i) s->shadows[i].row = s->o.max_rves + i;
    size_t wsb = s->o.workspace_bytes;
    if (wsb == 0) {
        // default pool: an eighth of the free memory, at most 16 GiB (several
        // hundred 64^3 increments in flight), leaving the rest for resident ids
        size_t fr = 0, tot = 0;
        (void)hipMemGetInfo(&fr, &tot);
        wsb = fr / 8;
        if (wsb > ((size_t)16 << 30)) wsb = (size_t)16 << 30;
    }
    wsb = al256(wsb);
    if (hipMalloc((void **)&s->pool_mem, wsb) != hipSuccess) FAIL(VOCE_ERR_NOMEM);
    s->pool.init(s->pool_mem, wsb);
#undef FAIL
    *out = s;
    return VOCE_OK;
}

static int upload_row(voce_server *s, int row)
{
    SCK(hipMemcpyAsync(s->dtab + row, s->htab + row, sizeof(RveDev), hipMemcpyHostToDevice, s->stream));
    return 0;
}

static bool take_ws(voce_server *s, int row, RveDev &d)
{
    Ws &w = s->ws[row];
    const size_t n = voce_ws_bytes(d);
    size_t off = 0;
    if (!s->pool.alloc(n, &off)) return false;
    w.off = off;
    w.len = n;
    w.held = true;
    voce_ws_carve(d, s->pool.base + off);
    return true;
}

static void drop_ws(voce_server *s, int row, RveDev &d)
{
    Ws &w = s->ws[row];
    if (w.held) s->pool.release(w.off, w.len);
    w.held = false;
    d.P6 = nullptr;
    d.KCV = nullptr;
    d.spec = nullptr;
    d.usum = nullptr;
    d.part = nullptr;
}

// throughput-tuned launch choices per (grid, precision), kept for the process
static std::mutex g_srv_tune_mu;
static std::map<std::tuple<int, int, int, int>, voce_tune_choice> g_srv_tune;

extern "C" int voce_server_add_rve(voce_server *s, int id, const voce_setup *su)
{
    if (!s || !su || id < 0 || id >= s->o.max_rves) return VOCE_ERR_ARG;
    if (s->ctx[id]) return VOCE_ERR_ARG;
    voce_setup cp = *su;
    cp.precision = s->o.precision;
    voce_build_opts bo;
    memset(&bo, 0, sizeof(bo));
    bo.stream = s->stream;
    bo.on_tab = s->htab + id;
    bo.no_workspace = 1;
    voce_ctx *c = nullptr;
    int rc = voce_ctx_build(&cp, &bo, &c);
    if (rc) return rc;
    s->ctx[id] = c;
    s->pred[id] = su->iter_max > 0 ? su->iter_max : 1;
    s->resident_bytes += (long)c->arena_bytes;
    if (s->o.tune) {
        const auto key = std::make_tuple(c->dims[0], c->dims[1], c->dims[2]);
        auto it = s->tuned.find(key);
        voce_tune_choice ch;
        const auto gkey = std::make_tuple(c->dims[0], c->dims[1], c->dims[2], s->o.precision);
        if (it == s->tuned.end()) {
            std::lock_guard<std::mutex> lk(g_srv_tune_mu);
            auto git = g_srv_tune.find(gkey);
            if (git != g_srv_tune.end()) it = s->tuned.emplace(key, git->second).first;
        }
        if (it == s->tuned.end()) {
            // A round launches the kernels over many RVEs at once, so the
            // choice is timed for throughput: on a temporary batch of copies
            // of this RVE, about 2 M voxels in all (at most 64 copies).  The
            // choice is kept for the process, so servers in one process run
            // a grid size with the same kernels (and the same rounding).
            const long per = c->V > 0 ? c->V : 1;
            int k = (int)((2L << 20) / per);
            k = k < 1 ? 1 : (k > 64 ? 64 : k);
            std::vector<voce_setup> cps((size_t)k, cp);
            voce_batch *tb = nullptr;
            rc = voce_batch_create(cps.data(), k, 1, &tb);
            if (rc == 0) {
                rc = voce_get_tuning(voce_batch_member(tb, 0), &ch);
                voce_batch_destroy(tb);
            } else {
                // no room for the copies: one RVE's latency, on a lent workspace
                if (!take_ws(s, id, c->d)) return VOCE_ERR_NOMEM;
                if ((rc = voce_sync_tab(c))) return rc;
                rc = voce_autotune(c, &ch);
                drop_ws(s, id, c->d);
            }
            if (rc) return rc;
            s->tuned[key] = ch;
            std::lock_guard<std::mutex> lk(g_srv_tune_mu);
            g_srv_tune.emplace(gkey, ch);
        } else {
            ch = it->second;
        }
        if ((rc = voce_apply_tuning(c, &ch))) return rc;
    }
    if ((rc = voce_sync_tab(c))) return rc;   // mirrors into htab[id]
    rc = upload_row(s, id);
    if (rc) return rc;
    SCK(hipStreamSynchronize(s->stream));
    return 0;
}

extern "C" int voce_server_submit(voce_server *s, int id, const double L[9], int audit, long tag)
{
    if (!s || !L || id < 0 || id >= s->o.max_rves || !s->ctx[id]) return VOCE_ERR_ARG;
    // a workspace larger than the whole pool would never be admitted
    if (voce_ws_bytes(s->ctx[id]->d) > s->pool.cap) return VOCE_ERR_NOMEM;
    Request r;
    r.tag = tag;
    r.id = id;
    r.audit = audit ? 1 : 0;
    memcpy(r.L, L, sizeof(r.L));
    r.t_submit = mono_now();
    s->queue.push_back(r);
    return 0;
}

// FP64 shadow of row `src` in shadow slot k: own evolving state and
// workspace, the source's constant tables, a copy of its state
// bytes of a shadow arena for id `src`: evolving arrays, P6, spectra, sums,
// partials, MacroState, logs, twiddles (FP64)
static size_t shadow_bytes(voce_server *s, int src)
{
    voce_ctx *c = s->ctx[src];
    const RveDev &d0 = c->d;
    const long V = d0.V;
    const size_t db = sizeof(double), cb = 16;
    const int two = d0.CRSSB != d0.CRSS;
    RveDev d = d0;
    d.fft_f32 = 0;
    voce_spec_layout(&d);
    const size_t nspec = d.nspec;
    size_t need = 0;
    const size_t sz_state[11] = {6 * V * db, 6 * V * db, 5 * V * db, 5 * V * db, 12 * V * db, two ? 12 * V * db : 0,
                                 9 * V * db, 3 * V * db, 3 * V * db, V * db, 12 * V * db};
    for (size_t b : sz_state) need += al256(b);
    need += al256(6 * V * db) + al256(VOCE_KCV_COMP * voce_kcv_voxels(V) * db) + al256(6 * nspec * cb) + al256(3 * nspec * cb) +
            al256((size_t)d0.nblk_vox * 16 * db);
    need += al256(sizeof(MacroState)) + al256(3 * (size_t)c->log_cap * db);
    need += al256(d0.nx * cb) + al256(d0.ny * cb) + al256(d0.nz * cb);
    return need;
}

// device memory of shadow slot k, large enough for id `src` (the only step
// of an audit's set-up that can run out of memory; done before the request
// takes anything else)
static int shadow_reserve(voce_server *s, int k, int src)
{
    Shadow &sh = s->shadows[k];
    const size_t need = shadow_bytes(s, src);
    if (s->fail_shadow_alloc > 0 && --s->fail_shadow_alloc == 0) return VOCE_ERR_NOMEM;
    if (sh.arena && sh.bytes < need) {
        (void)hipFree(sh.arena);
        sh.arena = nullptr;
        sh.bytes = 0;
    }
    if (!sh.arena) {
        if (hipMalloc((void **)&sh.arena, need) != hipSuccess) {
            sh.arena = nullptr;
            return VOCE_ERR_NOMEM;
        }
        sh.bytes = need;
    }
    return 0;
}

// FP64 shadow of row `src` in shadow slot k (reserved by shadow_reserve):
// own evolving state and workspace, the source's constant tables, a copy of
// its state
static int shadow_start(voce_server *s, int k, int src, const double L[9])
{
    Shadow &sh = s->shadows[k];
    voce_ctx *c = s->ctx[src];
    const RveDev &d0 = c->d;
    const long V = d0.V;
    const size_t db = sizeof(double), cb = 16;
    const int two = d0.CRSSB != d0.CRSS;
    // layout: evolving arrays, P6, spectra, sums, partials, MacroState, logs, twiddles
    RveDev d = d0;   // constant tables (ORI, PH, SLOC, phase) stay aliased
    d.fft_f32 = 0;
    voce_spec_layout(&d);   // FP64 spectrum layout
    const size_t nspec = d.nspec;
    const size_t sz_state[11] = {6 * V * db, 6 * V * db, 5 * V * db, 5 * V * db, 12 * V * db, two ? 12 * V * db : 0,
                                 9 * V * db, 3 * V * db, 3 * V * db, V * db, 12 * V * db};
    if (!sh.arena || sh.bytes < shadow_bytes(s, src)) return VOCE_ERR_ARG;
    sh.src = src;
    sh.V = V;
    // no diagnostics: the shadow must not write into the id's arrays (the
    // fluctuation sums FLU are updated whenever the pointer is set)
    d.GAMDOT = d.TAUX = d.SVM = d.EVM = d.WORKP = d.RHOR = d.RHOC = d.FLU = nullptr;
    char *p = sh.arena;
    double *dst[11];
    const double *srcp[11] = {d0.E6, d0.S6, d0.EP5, d0.ED5, d0.CRSS, two ? d0.CRSSB : nullptr, d0.R, d0.WG, d0.WF,
                              d0.GAC, d0.GTOT};
    for (int i = 0; i < 11; ++i) {
        dst[i] = sz_state[i] ? (double *)p : nullptr;
        p += al256(sz_state[i]);
        if (sz_state[i]) SCK(hipMemcpyAsync(dst[i], srcp[i], sz_state[i], hipMemcpyDeviceToDevice, s->stream));
    }
    d.E6 = dst[0]; d.S6 = dst[1]; d.EP5 = dst[2]; d.ED5 = dst[3]; d.CRSS = dst[4];
    d.CRSSB = two ? dst[5] : dst[4];
    d.R = dst[6]; d.WG = dst[7]; d.WF = dst[8]; d.GAC = dst[9]; d.GTOT = dst[10];
    d.P6 = (double *)p;
    p += al256(6 * V * db);
    d.KCV = (double *)p;
    p += al256(VOCE_KCV_COMP * voce_kcv_voxels(V) * db);
    d.spec = p;
    p += al256(6 * nspec * cb);
    d.usum = p;
    p += al256(3 * nspec * cb);
    d.part = (double *)p;
    p += al256((size_t)d0.nblk_vox * 16 * db);
    d.ms = (MacroState *)p;
    p += al256(sizeof(MacroState));
    double *logs = (double *)p;
    p += al256(3 * (size_t)c->log_cap * db);
    d.tw_x = p;
    p += al256(d0.nx * cb);
    d.tw_y = p;
    p += al256(d0.ny * cb);
    d.tw_z = p;
    // FP64 launch choices (defaults; audits are not tuned)
    d.plv = voce_default_plv(&d);
    d.xv = 0;
    d.tk = voce_xgreen_tile(d.nx, 0);
    d.pl_r8 = d.x_r8 = 1;
    {
        // FP64 twiddles e^{-2 pi i k/n}, as the contexts compute them
        for (int a = 0; a < 3; ++a) {
            const int n = a == 0 ? d.nx : (a == 1 ? d.ny : d.nz);
            std::vector<double> tw(2 * n);
            for (int k2 = 0; k2 < n; ++k2) {
                const long double ang = -2.0L * 3.14159265358979323846264338327950288L * k2 / n;
                tw[2 * k2] = (double)cosl(ang);
                tw[2 * k2 + 1] = (double)sinl(ang);
            }
            void *t = a == 0 ? d.tw_x : (a == 1 ? d.tw_y : d.tw_z);
            SCK(hipMemcpy(t, tw.data(), 2 * n * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    sh.nx = d.nx;
    sh.ny = d.ny;
    sh.nz = d.nz;
    sh.two = two;
    sh.busy = true;
    // macro state: the source's, with the shadow's logs, no diagnostics, L pending
    SCK(hipMemcpyAsync(d.ms, d0.ms, sizeof(MacroState), hipMemcpyDeviceToDevice, s->stream));
    MacroState *stage = s->h_ms + sh.row;
    stage->logs = logs;
    SCK(hipMemcpyAsync((char *)d.ms + offsetof(MacroState, logs), &stage->logs, sizeof(double *), hipMemcpyHostToDevice,
                       s->stream));
    SCK(hipMemcpyAsync((char *)d.ms + offsetof(MacroState, diag), s->h_one + 0, sizeof(int), hipMemcpyHostToDevice,
                       s->stream));
    memcpy(s->h_vg + 9 * sh.row, L, 9 * sizeof(double));
    SCK(hipMemcpyAsync((char *)d.ms + offsetof(MacroState, vg_next), s->h_vg + 9 * sh.row, 9 * sizeof(double),
                       hipMemcpyHostToDevice, s->stream));
    SCK(hipMemcpyAsync((char *)d.ms + offsetof(MacroState, vg_pending), s->h_one + 1, sizeof(int), hipMemcpyHostToDevice,
                       s->stream));
    if (!voce_plane_lds(&d, d.plv) || !voce_xgreen_lds(&d, d.xv, d.tk)) return VOCE_ERR_UNSUPPORTED_GRID;
    s->htab[sh.row] = d;
    return upload_row(s, sh.row);
}

// launch sets over a list of rows, one per (grid, precision)
static int build_sets(voce_server *s, const std::vector<int> &rows, std::vector<LaunchSet> &sets, int2 *h, int2 *dmaps,
                      size_t cap, size_t *used)
{
    // one launch set per grid, precision and kernel variants (an audit's
    // FP64 shadow may run other variants than a tuned row of the same grid)
    std::map<std::tuple<int, int, int, int, int, int, int>, std::vector<int>> g;
    for (int r : rows) {
        const RveDev &d = s->htab[r];
        g[std::make_tuple(d.nx, d.ny, d.nz, d.fft_f32, d.plv, d.xv, d.tk)].push_back(r);
    }
    sets.clear();
    size_t off = 0;
    for (auto &kv : g) {
        LaunchSet L;
        memset(&L, 0, sizeof(L));
        L.tab = s->dtab;
        const RveDev &d0 = s->htab[kv.second[0]];
        L.plv = d0.plv;
        L.p2_n = d0.ny;
        L.p2_nx = d0.nx;
        L.sp_ny = d0.ny;
        L.sp_nz = d0.nz;
        L.p2_tk = d0.tk;
        L.xv = d0.xv;
        L.fft_f32 = d0.fft_f32;
        L.uniform_phase = 1;
        size_t need = 0;
        for (int r : kv.second) {
            const RveDev &d = s->htab[r];
            if (d.plv != d0.plv || d.xv != d0.xv) return VOCE_ERR_ARG;
            need += (size_t)d.nx * 15 + (d.ny * d.nzh + d.tk - 1) / d.tk + d.nblk_vox + 2;
        }
        if (off + need > cap) return VOCE_ERR_NOMEM;
        auto put = [&](int count_per, int kind, const int2 **mp, int *np) {
            (void)count_per;
            *mp = dmaps + off;
            int n = 0;
            for (int r : kv.second) {
                const RveDev &d = s->htab[r];
                int m = 0;
                if (kind == 5) m = d.nx * 6 + 1;   // forward planes + finaliser block
                else if (kind == 0) m = d.nx * 6;
                else if (kind == 1) m = d.nx * 3;
                else if (kind == 2) m = (d.ny * d.nzh + d.tk - 1) / d.tk;
                else if (kind == 3) m = d.nblk_vox;
                else m = 1;
                for (int k = 0; k < m; ++k) h[off + n++] = make_int2(r, k);
            }
            *np = n;
            off += n;
        };
        put(0, 5, &L.map_pfwd, &L.n_pfwd);
        put(0, 0, &L.map_plane6, &L.n_plane6);
        put(0, 1, &L.map_plane3, &L.n_plane3);
        put(0, 2, &L.map_x, &L.n_x);
        put(0, 3, &L.map_vox, &L.n_vox);
        put(0, 4, &L.map_one, &L.n_one);
        for (int r : kv.second) {
            const RveDev &d = s->htab[r];
            L.uniform_phase = L.uniform_phase && d.uniform_phase;
            L.two_crss = L.two_crss || d.two_crss;
            const size_t a = voce_plane_lds(&d, d.plv), b = voce_xgreen_lds(&d, d.xv, d.tk);
            if (!a || !b) return VOCE_ERR_UNSUPPORTED_GRID;
            if (a > L.pl_lds) L.pl_lds = a;
            if (b > L.x_lds) L.x_lds = b;
        }
        sets.push_back(L);
    }
    *used = off;
    return 0;
}

static int ensure_scratch(voce_server *s, size_t nmaps, size_t nidx)
{
    if (nmaps > s->maps_cap) {
        size_t cap = s->maps_cap ? s->maps_cap : 65536;
        while (cap < nmaps) cap *= 2;
        SCK(hipStreamSynchronize(s->stream));
        if (s->d_maps) (void)hipFree(s->d_maps);
        if (s->h_maps) (void)hipHostFree(s->h_maps);
        s->d_maps = nullptr;
        s->h_maps = nullptr;
        // three regions: begin, iteration, close-out maps
        if (hipMalloc((void **)&s->d_maps, 3 * cap * sizeof(int2)) != hipSuccess) return VOCE_ERR_NOMEM;
        if (hipHostMalloc((void **)&s->h_maps, 3 * cap * sizeof(int2), hipHostMallocDefault) != hipSuccess)
            return VOCE_ERR_NOMEM;
        s->maps_cap = cap;
        s->set_changed = true;
    }
    if (nidx > s->idx_cap) {
        size_t cap = s->idx_cap ? s->idx_cap : 1024;
        while (cap < nidx) cap *= 2;
        SCK(hipStreamSynchronize(s->stream));
        void *ps[] = {s->d_idx, s->d_flags};
        for (void *p : ps)
            if (p) (void)hipFree(p);
        void *hs[] = {s->h_idx, s->h_flags};
        for (void *p : hs)
            if (p) (void)hipHostFree(p);
        if (hipMalloc((void **)&s->d_idx, cap * sizeof(int)) != hipSuccess ||
            hipMalloc((void **)&s->d_flags, 2 * cap * sizeof(int)) != hipSuccess ||
            hipHostMalloc((void **)&s->h_idx, cap * sizeof(int), hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&s->h_flags, 2 * cap * sizeof(int), hipHostMallocDefault) != hipSuccess)
            return VOCE_ERR_NOMEM;
        s->idx_cap = cap;
    }
    return 0;
}

static size_t map_need(voce_server *s, const std::vector<int> &rows)
{
    size_t n = 0;
    for (int r : rows) {
        const RveDev &d = s->htab[r];
        n += (size_t)d.nx * 15 + (d.ny * d.nzh + d.tk - 1) / d.tk + d.nblk_vox + 2;
    }
    return n;
}

static void fill_result(voce_result &res, const MacroState &M)
{
    memset(&res.rep, 0, sizeof(res.rep));
    res.rep.istep = M.istep;
    res.rep.iter = M.iter;
    res.rep.iter_total = M.iter_total;
    res.rep.e_vm = M.e_vm; res.rep.evmp = M.evmp; res.rep.d_vm = M.d_vm; res.rep.dvmp = M.dvmp;
    res.rep.s_vm = M.s_vm; res.rep.s_vm1 = M.s_vm1;
    memcpy(res.rep.sig_avg, M.SigAvg, 72);
    memcpy(res.rep.sig_avg1, M.SigAvg1, 72);
    memcpy(res.rep.disgrad_avg, M.DisGradAvg_actual, 72);
    memcpy(res.rep.velgrad_avg, M.VelGradAvg, 72);
    memcpy(res.rep.eps_avg, M.EpsAvg, 72);
    memcpy(res.rep.edot_avg, M.EdotAvg, 72);
}

static double sig_rel(const double *a, const double *b)
{
    double m = 0.0, r = 0.0;
    for (int i = 0; i < 9; ++i) {
        m = fmax(m, fabs(b[i]));
        r = fmax(r, fabs(a[i] - b[i]));
    }
    return m > 0 ? r / m : r;
}

static int server_step(voce_server *s);

extern "C" int voce_server_step(voce_server *s, int *inflight)
{
    if (!s) return VOCE_ERR_ARG;
    const int rc = server_step(s);
    if (inflight) *inflight = (int)s->inflight.size();
    return rc;
}

// 0 or an error code
static int server_step(voce_server *s)
{
    int rc = 0;
    ++s->steps;
    // 1. admission.  A failure part way (no memory for an audit's shadow, a
    // device error) puts the request back where it was, releases what it had
    // taken, and still starts the increments admitted before it, so the
    // server stays consistent and the request can be retried.
    std::vector<int> started;
    int adm_rc = 0;
    {
        std::deque<Request> keep;
        bool blocked = false;
        while (!s->queue.empty()) {
            Request r = s->queue.front();
            s->queue.pop_front();
            const int row = r.id;
            voce_ctx *c = s->ctx[row];
            bool ok = !blocked && !s->busy[row] &&
                      (s->o.max_inflight <= 0 || (int)s->inflight.size() < s->o.max_inflight);
            int sk = -1;
            if (ok && r.audit) {
                for (int k = 0; k < s->max_shadow; ++k)
                    if (!s->shadows[k].busy && !s->busy[s->shadows[k].row]) {
                        sk = k;
                        break;
                    }
                if (sk < 0) ok = false;
            }
            bool took_ws = false;
            auto fail = [&](int code) {
                if (took_ws) drop_ws(s, row, c->d);
                s->htab[row] = c->d;
                s->busy[row] = 0;
                keep.push_back(r);   // back in arrival order, ahead of the rest
                while (!s->queue.empty()) {
                    keep.push_back(s->queue.front());
                    s->queue.pop_front();
                }
                adm_rc = code;
            };
            if (ok && r.audit) {
                const int e = shadow_reserve(s, sk, row);
                if (e) {
                    fail(e);
                    break;
                }
            }
            if (ok && !take_ws(s, row, c->d)) {
                ok = false;
                blocked = true;   // keep arrival order under memory pressure
            }
            if (!ok) {
                keep.push_back(r);
                continue;
            }
            took_ws = true;
            // the id's next increment: velocity gradient, workspace pointers
            memcpy(s->h_vg + 9 * row, r.L, 9 * sizeof(double));
            if (hipMemcpyAsync((char *)c->d.ms + offsetof(MacroState, vg_next), s->h_vg + 9 * row, 9 * sizeof(double),
                               hipMemcpyHostToDevice, s->stream) != hipSuccess ||
                hipMemcpyAsync((char *)c->d.ms + offsetof(MacroState, vg_pending), s->h_one + 1, sizeof(int),
                               hipMemcpyHostToDevice, s->stream) != hipSuccess) {
                fail(VOCE_ERR_DEVICE);
                break;
            }
            s->htab[row] = c->d;
            if (upload_row(s, row)) {
                fail(VOCE_ERR_DEVICE);
                break;
            }
            Inflight f;
            f.row = row;
            f.seq = s->next_seq++;
            f.req = r;
            f.t_admit = mono_now();
            f.shadow = -1;
            if (r.audit) {
                // the shadow copies the state before the increment starts
                const int e = shadow_start(s, sk, row, r.L);
                if (e) {
                    fail(e);
                    break;
                }
                s->busy[s->shadows[sk].row] = 1;
                f.shadow = sk;
                Inflight fs;
                fs.row = s->shadows[sk].row;
                fs.seq = f.seq;
                fs.req = r;
                fs.t_admit = f.t_admit;
                fs.shadow = sk;
                s->inflight.push_back(fs);
                started.push_back(fs.row);
                PendingAudit pa;
                memset(&pa.res, 0, sizeof(pa.res));
                s->audits[f.seq] = pa;
                ++s->n_audits;
            }
            s->busy[row] = 1;
            started.push_back(row);
            s->inflight.push_back(f);
            s->set_changed = true;
        }
        s->queue.swap(keep);
    }
    if (adm_rc) {
        // start what was admitted before the failure, then report it
        if (!started.empty()) {
            std::vector<int> rows;
            for (auto &f : s->inflight) rows.push_back(f.row);
            if (ensure_scratch(s, map_need(s, rows), rows.size())) return adm_rc;
            int2 *hb = s->h_maps;
            for (size_t i = 0; i < started.size(); ++i) hb[i] = make_int2(started[i], 0);
            LaunchSet B;
            memset(&B, 0, sizeof(B));
            B.tab = s->dtab;
            B.map_one = s->d_maps;
            B.n_one = (int)started.size();
            if (hipMemcpyAsync(s->d_maps, hb, started.size() * sizeof(int2), hipMemcpyHostToDevice, s->stream) ==
                hipSuccess)
                (void)voce_launch_begin_increment(&B, s->stream);
            (void)hipStreamSynchronize(s->stream);
        }
        return adm_rc;
    }
    if (s->inflight.empty()) return 0;
    std::vector<int> rows;
    for (auto &f : s->inflight) rows.push_back(f.row);
    if ((rc = ensure_scratch(s, map_need(s, rows), rows.size()))) return rc;
    const size_t cap = s->maps_cap;
    // 2. start the admitted increments
    if (!started.empty()) {
        int2 *hb = s->h_maps;
        for (size_t i = 0; i < started.size(); ++i) hb[i] = make_int2(started[i], 0);
        SCK(hipMemcpyAsync(s->d_maps, hb, started.size() * sizeof(int2), hipMemcpyHostToDevice, s->stream));
        LaunchSet B;
        memset(&B, 0, sizeof(B));
        B.tab = s->dtab;
        B.map_one = s->d_maps;
        B.n_one = (int)started.size();
        if (voce_launch_begin_increment(&B, s->stream)) return VOCE_ERR_DEVICE;
    }
    // 3. one chunk of iterations over everything in flight
    if (s->set_changed) {
        size_t used = 0;
        if ((rc = build_sets(s, rows, s->iter_sets, s->h_maps + cap, s->d_maps + cap, cap, &used))) return rc;
        SCK(hipMemcpyAsync(s->d_maps + cap, s->h_maps + cap, used * sizeof(int2), hipMemcpyHostToDevice, s->stream));
        s->set_changed = false;
    }
    // chunk: up to the nearest predicted end among the rows in flight (each
    // id's last increment), between 2 and the configured maximum; a short
    // chunk lets finished rows leave and queued requests join sooner
    int chunk = s->o.chunk;
    for (const Inflight &f : s->inflight) {
        const int rem = s->pred[f.req.id] - f.launched;
        if (rem < chunk) chunk = rem;
    }
    if (chunk < 2) chunk = 2;
    if (chunk > s->o.chunk) chunk = s->o.chunk;
    for (int it = 0; it < chunk; ++it)
        for (LaunchSet &L : s->iter_sets)
            if (voce_launch_iteration(&L, s->stream, nullptr)) return VOCE_ERR_DEVICE;
    for (LaunchSet &L : s->iter_sets)   // the chunk's last Newton step
        if (voce_launch_finalize(&L, s->stream)) return VOCE_ERR_DEVICE;
    for (Inflight &f : s->inflight) f.launched += chunk;
    s->iters += chunk;
    // 4. poll and close out the finished rows
    for (size_t i = 0; i < rows.size(); ++i) s->h_idx[i] = rows[i];
    SCK(hipMemcpyAsync(s->d_idx, s->h_idx, rows.size() * sizeof(int), hipMemcpyHostToDevice, s->stream));
    if (voce_launch_poll(s->dtab, s->d_idx, (int)rows.size(), s->d_flags, s->stream)) return VOCE_ERR_DEVICE;
    SCK(hipMemcpyAsync(s->h_flags, s->d_flags, 2 * rows.size() * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    SCK(hipStreamSynchronize(s->stream));
    std::vector<int> fin;
    for (size_t i = 0; i < rows.size(); ++i) {
        if (s->h_flags[2 * i + 1]) return s->h_flags[2 * i + 1];
        if (!s->h_flags[2 * i]) fin.push_back((int)i);
    }
    if (fin.empty()) return 0;
    std::vector<int> frows;
    for (int i : fin) frows.push_back(rows[i]);
    std::vector<LaunchSet> csets;
    size_t used = 0;
    if ((rc = build_sets(s, frows, csets, s->h_maps + 2 * cap, s->d_maps + 2 * cap, cap, &used))) return rc;
    SCK(hipMemcpyAsync(s->d_maps + 2 * cap, s->h_maps + 2 * cap, used * sizeof(int2), hipMemcpyHostToDevice, s->stream));
    for (LaunchSet &L : csets)
        if (voce_launch_closeout(&L, s->stream, 1)) return VOCE_ERR_DEVICE;
    for (int r : frows)
        SCK(hipMemcpyAsync(s->h_ms + r, s->htab[r].ms, sizeof(MacroState), hipMemcpyDeviceToHost, s->stream));
    SCK(hipStreamSynchronize(s->stream));
    const double t_done = mono_now();
    std::vector<Inflight> still;
    std::vector<bool> is_fin(rows.size(), false);
    for (int i : fin) is_fin[i] = true;
    for (size_t i = 0; i < s->inflight.size(); ++i) {
        Inflight &f = s->inflight[i];
        if (!is_fin[i]) {
            still.push_back(f);
            continue;
        }
        const MacroState &M = s->h_ms[f.row];
        const bool is_shadow = f.row >= s->o.max_rves;
        if (is_shadow) {
            Shadow &sh = s->shadows[f.shadow];
            sh.busy = false;
            s->busy[f.row] = 0;
            s->htab[f.row].ms = nullptr;
            PendingAudit &pa = s->audits[f.seq];
            pa.shadow_ms = M;
            pa.shadow_done = true;
        } else {
            voce_ctx *c = s->ctx[f.row];
            drop_ws(s, f.row, c->d);
            s->htab[f.row] = c->d;
            if ((rc = upload_row(s, f.row))) return rc;
            s->busy[f.row] = 0;
            c->n_steps_done = M.istep;
            s->pred[f.req.id] = M.iter > 0 ? M.iter : 1;
            voce_result res;
            memset(&res, 0, sizeof(res));
            res.tag = f.req.tag;
            res.id = f.req.id;
            res.status = M.error;
            res.audited = f.req.audit;
            fill_result(res, M);
            res.t_submit = f.req.t_submit;
            res.t_admit = f.t_admit;
            res.t_done = t_done;
            if (f.req.audit) {
                PendingAudit &pa = s->audits[f.seq];
                pa.res = res;
                pa.main_done = true;
            } else {
                s->done.push_back(res);
                ++s->n_done;
            }
        }
        if (f.req.audit) {
            auto it = s->audits.find(f.seq);
            if (it != s->audits.end() && it->second.main_done && it->second.shadow_done) {
                voce_result res = it->second.res;
                res.audit_sig_rel = sig_rel(res.rep.sig_avg, it->second.shadow_ms.SigAvg);
                res.t_done = t_done;
                s->done.push_back(res);
                ++s->n_done;
                s->audits.erase(it);
            }
        }
    }
    s->inflight.swap(still);
    s->set_changed = true;
    return 0;
}

extern "C" int voce_server_poll(voce_server *s, voce_result *out, int max)
{
    if (!s || (!out && max > 0)) return VOCE_ERR_ARG;
    int n = 0;
    while (n < max && n < (int)s->done.size()) {
        out[n] = s->done[n];
        ++n;
    }
    s->done.erase(s->done.begin(), s->done.begin() + n);
    return n;
}

extern "C" int voce_server_info(const voce_server *s, long out[10])
{
    if (!s || !out) return VOCE_ERR_ARG;
    long reg = 0;
    for (voce_ctx *c : s->ctx) reg += c != nullptr;
    out[0] = reg;
    out[1] = s->resident_bytes;
    out[2] = (long)s->pool.cap;
    out[3] = (long)s->pool.used;
    out[4] = (long)s->inflight.size();
    out[5] = (long)s->queue.size();
    out[6] = s->n_done;
    out[7] = s->n_audits;
    out[8] = s->steps;
    out[9] = s->iters;
    return 0;
}

extern "C" int voce_server_inject_fault(voce_server *s, int kind, int after)
{
    if (!s || kind != VOCE_FAULT_SHADOW_ALLOC || after < 0) return VOCE_ERR_ARG;
    s->fail_shadow_alloc = after;
    return 0;
}

extern "C" voce_ctx *voce_server_rve(voce_server *s, int id)
{
    if (!s || id < 0 || id >= s->o.max_rves) return nullptr;
    return s->ctx[id];
}
