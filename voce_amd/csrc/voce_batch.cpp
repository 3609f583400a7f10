// voce_batch.cpp -- many RVEs advanced together on one GPU.
//
// Each member is a full context (its own device state) whose MacroState
// sits in one batch-owned array, so the per-increment read-back is one copy.
// Members are grouped by grid size; a group's kernels cover all of its
// members through block maps (block -> member, local block) over one
// device descriptor table.  An iteration is the five kernels of every
// group; chunks of iterations are captured once into HIP graphs and
// replayed, with a one-block poll kernel between chunks.  Converged members
// return at the top of every kernel, so the rest keep iterating.
#include <hip/hip_runtime.h>

#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/voce_batch.h"
#include "voce_host.h"

namespace {

struct Group {
    int nx, ny, nz;
    std::vector<int> members;
    LaunchSet L;
    int2 *dmap = nullptr;   // device block maps (batch-owned), reused when they fit
    size_t dmap_cap = 0;
};

}  // namespace

struct voce_batch {
    int n = 0;
    int precision = 0;
    hipStream_t stream = nullptr;
    std::vector<voce_ctx *> m;
    std::vector<RveDev> htab;         // host mirror of the descriptor table
    RveDev *dtab = nullptr;
    MacroState *ms_all = nullptr;     // device, n
    double *logs_all = nullptr;       // device, 3 x iter_max per member
    std::vector<size_t> log_off;
    size_t log_total = 0;
    MacroState *h_ms = nullptr;       // pinned, n
    double *h_logs = nullptr;         // pinned
    int *d_cnt = nullptr, *h_cnt = nullptr;
    int2 *map_all = nullptr;          // (i, 0) for every member
    std::vector<Group> groups;
    std::vector<void *> allocs;       // batch-owned device buffers
    std::map<int, hipGraphExec_t> graphs;
    int pred = 0, iter_max_any = 0;
    long replays = 0, iters_launched = 0, last_max_iter = 0, dev_bytes = 0;
};

static int balloc(voce_batch *b, void **p, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    if (hipMalloc(p, bytes) != hipSuccess) return VOCE_ERR_NOMEM;
    b->allocs.push_back(*p);
    return 0;
}

// block maps of one group, uploaded into one buffer
static int build_group_maps(voce_batch *b, Group &g)
{
    std::vector<int2> pf, p6, p3, x, vox, one;
    bool uni = true, two = false;
    size_t pl_lds = 0, x_lds = 0;
    for (int i : g.members) {
        const RveDev &d = b->htab[i];
        for (int k = 0; k <= d.nx * 6; ++k) pf.push_back(make_int2(i, k));   // + finaliser block
        for (int k = 0; k < d.nx * 6; ++k) p6.push_back(make_int2(i, k));
        for (int k = 0; k < d.nx * 3; ++k) p3.push_back(make_int2(i, k));
        const int nxb = (d.ny * d.nzh + d.tk - 1) / d.tk;
        for (int k = 0; k < nxb; ++k) x.push_back(make_int2(i, k));
        for (int k = 0; k < d.nblk_vox; ++k) vox.push_back(make_int2(i, k));
        one.push_back(make_int2(i, 0));
        uni = uni && d.uniform_phase;
        two = two || d.two_crss;
        const size_t a = voce_plane_lds(&d, d.plv), c = voce_xgreen_lds(&d, d.xv, d.tk);
        if (!a || !c) return VOCE_ERR_UNSUPPORTED_GRID;
        if (a > pl_lds) pl_lds = a;
        if (c > x_lds) x_lds = c;
        if (d.plv != b->htab[g.members[0]].plv || d.xv != b->htab[g.members[0]].xv) return VOCE_ERR_ARG;
    }
    const size_t tot = pf.size() + p6.size() + p3.size() + x.size() + vox.size() + one.size();
    if (tot > g.dmap_cap) {
        int rc = balloc(b, (void **)&g.dmap, tot * sizeof(int2));
        if (rc) return rc;
        g.dmap_cap = tot;
    }
    int2 *dm = g.dmap;
    std::vector<int2> all;
    all.reserve(tot);
    LaunchSet &L = g.L;
    memset(&L, 0, sizeof(L));
    L.tab = b->dtab;
    size_t off = 0;
    auto put = [&](std::vector<int2> &v, const int2 **mp, int *np) {
        *mp = dm + off;
        *np = (int)v.size();
        all.insert(all.end(), v.begin(), v.end());
        off += v.size();
    };
    put(pf, &L.map_pfwd, &L.n_pfwd);
    put(p6, &L.map_plane6, &L.n_plane6);
    put(p3, &L.map_plane3, &L.n_plane3);
    put(x, &L.map_x, &L.n_x);
    put(vox, &L.map_vox, &L.n_vox);
    put(one, &L.map_one, &L.n_one);
    const RveDev &d0 = b->htab[g.members[0]];
    L.plv = d0.plv;
    L.p2_n = d0.ny;
    L.p2_nx = d0.nx;
    L.sp_ny = d0.ny;
    L.sp_nz = d0.nz;
    L.p2_tk = d0.tk;
    L.xv = d0.xv;
    L.pl_lds = pl_lds;
    L.x_lds = x_lds;
    L.fft_f32 = d0.fft_f32;
    L.uniform_phase = uni;
    L.two_crss = two;
    if (hipMemcpyAsync(dm, all.data(), tot * sizeof(int2), hipMemcpyHostToDevice, b->stream) != hipSuccess)
        return VOCE_ERR_DEVICE;
    return hipStreamSynchronize(b->stream) == hipSuccess ? 0 : VOCE_ERR_DEVICE;
}

static int upload_tab(voce_batch *b)
{
    if (hipMemcpyAsync(b->dtab, b->htab.data(), b->n * sizeof(RveDev), hipMemcpyHostToDevice, b->stream) != hipSuccess)
        return VOCE_ERR_DEVICE;
    return hipStreamSynchronize(b->stream) == hipSuccess ? 0 : VOCE_ERR_DEVICE;
}

static int group_apply(voce_batch *b, Group &g, const voce_tune_choice &ch)
{
    int rc;
    for (int i : g.members)
        if ((rc = voce_apply_tuning(b->m[i], &ch))) return rc;
    if ((rc = upload_tab(b))) return rc;
    return build_group_maps(b, g);
}

// Throughput tuning of one group: every FFT launch choice is timed on the
// whole group (all of its members' blocks in one launch, as an iteration
// runs them).  A full GPU can favour other shapes than the latency of a
// single RVE, which is what voce_autotune measures.  The probes overwrite
// the spectra (scratch) and update E6, which is saved and restored.
// group-tuned choices per (grid, precision, members), kept for the process so
// that equal batches run the same kernels (and round the same way)
static std::mutex g_grp_tune_mu;
static std::map<std::tuple<int, int, int, int, int>, voce_tune_choice> g_grp_tune;

static int tune_group_timed(voce_batch *b, Group &g);

static int tune_group(voce_batch *b, Group &g)
{
    const auto key = std::make_tuple(g.nx, g.ny, g.nz, b->precision, (int)g.members.size());
    {
        std::lock_guard<std::mutex> lk(g_grp_tune_mu);
        auto it = g_grp_tune.find(key);
        if (it != g_grp_tune.end()) return group_apply(b, g, it->second);
    }
    const int rc = tune_group_timed(b, g);
    if (rc) return rc;
    voce_tune_choice ch;
    if (voce_get_tuning(b->m[g.members[0]], &ch) == 0) {
        std::lock_guard<std::mutex> lk(g_grp_tune_mu);
        g_grp_tune.emplace(key, ch);
    }
    return 0;
}

static int tune_group_timed(voce_batch *b, Group &g)
{
    voce_ctx *c0 = b->m[g.members[0]];
    const voce_tune_choice keep = {c0->d.plv, c0->d.xv, c0->d.tk, c0->d.pl_r8, c0->d.x_r8, 0.0, 0.0};
    size_t tot = 0;
    for (int i : g.members) tot += 6 * (size_t)b->m[i]->V;
    double *save = nullptr;
    if (hipMalloc((void **)&save, tot * sizeof(double)) != hipSuccess) return VOCE_ERR_NOMEM;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = VOCE_ERR_DEVICE;
    const size_t act_off = offsetof(MacroState, active);
    static const int flag_on = 1, flag_off = 0;
    auto set_active = [&](const int *flag) {
        for (int i : g.members)
            if (hipMemcpyAsync((char *)(b->ms_all + i) + act_off, flag, sizeof(int), hipMemcpyHostToDevice, b->stream) !=
                hipSuccess)
                return VOCE_ERR_DEVICE;
        return hipStreamSynchronize(b->stream) == hipSuccess ? 0 : VOCE_ERR_DEVICE;
    };
    {
        size_t off = 0;
        for (int i : g.members) {
            const size_t n6 = 6 * (size_t)b->m[i]->V;
            if (!rc && hipMemcpyAsync(save + off, b->m[i]->d.E6, n6 * sizeof(double), hipMemcpyDeviceToDevice,
                                      b->stream) != hipSuccess)
                rc = VOCE_ERR_DEVICE;
            off += n6;
        }
    }
    if (!rc) rc = set_active(&flag_on);
    auto probe = [&](int which, double *us) -> int {
        const int reps = 3;
        for (int i = 0; i < 2; ++i)
            if (voce_launch_fft_probe(&g.L, b->stream, which)) return VOCE_ERR_DEVICE;
        if (hipEventRecord(e0, b->stream) != hipSuccess) return VOCE_ERR_DEVICE;
        for (int i = 0; i < reps; ++i)
            if (voce_launch_fft_probe(&g.L, b->stream, which)) return VOCE_ERR_DEVICE;
        if (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return VOCE_ERR_DEVICE;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return VOCE_ERR_DEVICE;
        *us = 1e3 * ms / reps;
        return 0;
    };
    auto fits = [&](const voce_tune_choice &ch) {
        RveDev d = c0->d;
        d.tk = ch.tk;
        return voce_plane_lds(&d, ch.plv) > 0 && voce_xgreen_lds(&d, ch.xv, ch.tk) > 0;
    };
    voce_tune_choice best = keep, cand;
    double best_pl = 1e300, best_x = 1e300;
    for (int plv = 0; plv < VOCE_N_PLV && !rc; ++plv)
        for (int r8 = 1; r8 >= ((plv == 4 || plv == 6) ? 1 : 0) && !rc; --r8) {   // plv 4 and 6 have fixed radices
            cand = keep;
            cand.plv = plv;
            cand.pl_r8 = r8;
            if (!fits(cand)) continue;
            if ((rc = group_apply(b, g, cand))) break;
            double a = 0, bb = 0;
            if ((rc = probe(0, &a)) || (rc = probe(2, &bb))) break;
            if (a + bb < best_pl) {
                best_pl = a + bb;
                best.plv = plv;
                best.pl_r8 = r8;
            }
        }
    const int tks[5] = {1, 2, 4, 8, 16};
    for (int xv = 0; xv < VOCE_N_XV && !rc; ++xv)
        for (int ti = 0; ti < 5 && !rc; ++ti)
            for (int r8 = 1; r8 >= (xv >= 3 ? 1 : 0) && !rc; --r8) {   // xv 3 and 4 have fixed radices
                cand = keep;
                cand.xv = xv;
                cand.tk = tks[ti];
                cand.x_r8 = r8;
                if (!fits(cand)) continue;
                if ((rc = group_apply(b, g, cand))) break;
                double a = 0;
                if ((rc = probe(1, &a))) break;
                if (a < best_x) {
                    best_x = a;
                    best.xv = xv;
                    best.tk = tks[ti];
                    best.x_r8 = r8;
                }
            }
    if (!rc && (best_pl >= 1e300 || best_x >= 1e300)) rc = VOCE_ERR_UNSUPPORTED_GRID;
    best.us_plane = best_pl;
    best.us_x = best_x;
    // restore E6 and the flags whatever happened, then settle on a choice
    {
        size_t off = 0;
        for (int i : g.members) {
            const size_t n6 = 6 * (size_t)b->m[i]->V;
            if (hipMemcpyAsync(b->m[i]->d.E6, save + off, n6 * sizeof(double), hipMemcpyDeviceToDevice, b->stream) !=
                    hipSuccess && !rc)
                rc = VOCE_ERR_DEVICE;
            off += n6;
        }
    }
    const int rc2 = set_active(&flag_off);
    if (!rc) rc = rc2;
    const int rc3 = group_apply(b, g, rc ? keep : best);
    if (!rc) rc = rc3;
    (void)hipFree(save);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

static void free_graphs(voce_batch *b)
{
    for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
}

extern "C" void voce_batch_destroy(voce_batch *b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    free_graphs(b);
    for (voce_ctx *c : b->m) voce_ctx_destroy(c);
    for (void *p : b->allocs) (void)hipFree(p);
    if (b->h_ms) (void)hipHostFree(b->h_ms);
    if (b->h_logs) (void)hipHostFree(b->h_logs);
    if (b->h_cnt) (void)hipHostFree(b->h_cnt);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

extern "C" int voce_batch_create(const voce_setup *setups, int n, int tune, voce_batch **out)
{
    if (!out) return VOCE_ERR_ARG;
    *out = nullptr;
    if (!setups || n <= 0) return VOCE_ERR_ARG;
    for (int i = 1; i < n; ++i)
        if (setups[i].precision != setups[0].precision) return VOCE_ERR_ARG;
    voce_batch *b = new voce_batch();
    b->n = n;
    b->precision = setups[0].precision;
    size_t free0 = 0, total0 = 0, free1 = 0, total1 = 0;
    (void)hipMemGetInfo(&free0, &total0);
    int rc = 0;
#define FAIL(code)              \
    do {                        \
        rc = (code);            \
        voce_batch_destroy(b);  \
        return rc;              \
    } while (0)
    if (hipStreamCreate(&b->stream) != hipSuccess) FAIL(VOCE_ERR_DEVICE);
    b->log_off.resize(n);
    for (int i = 0; i < n; ++i) {
        b->log_off[i] = b->log_total;
        const int cap = setups[i].iter_max > 0 ? setups[i].iter_max : 1;
        b->log_total += 3 * (size_t)cap;
        if (setups[i].iter_max > b->iter_max_any) b->iter_max_any = setups[i].iter_max;
    }
    if ((rc = balloc(b, (void **)&b->ms_all, n * sizeof(MacroState)))) FAIL(rc);
    if ((rc = balloc(b, (void **)&b->logs_all, b->log_total * sizeof(double)))) FAIL(rc);
    if ((rc = balloc(b, (void **)&b->dtab, n * sizeof(RveDev)))) FAIL(rc);
    if ((rc = balloc(b, (void **)&b->d_cnt, 4 * sizeof(int)))) FAIL(rc);
    if ((rc = balloc(b, (void **)&b->map_all, n * sizeof(int2)))) FAIL(rc);
    {
        std::vector<int2> ma(n);
        for (int i = 0; i < n; ++i) ma[i] = make_int2(i, 0);
        if (hipMemcpyAsync(b->map_all, ma.data(), n * sizeof(int2), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
            hipStreamSynchronize(b->stream) != hipSuccess)
            FAIL(VOCE_ERR_DEVICE);
    }
    if (hipHostMalloc((void **)&b->h_ms, n * sizeof(MacroState), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&b->h_logs, b->log_total * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&b->h_cnt, 4 * sizeof(int), hipHostMallocDefault) != hipSuccess)
        FAIL(VOCE_ERR_NOMEM);
    b->htab.resize(n);   // members keep pointers into it: no reallocation after this
    b->m.assign(n, nullptr);
    for (int i = 0; i < n; ++i) {
        voce_build_opts o;
        memset(&o, 0, sizeof(o));
        o.stream = b->stream;
        o.ms_dev = b->ms_all + i;
        o.logs_dev = b->logs_all + b->log_off[i];
        o.on_tab = &b->htab[i];
        if ((rc = voce_ctx_build(&setups[i], &o, &b->m[i]))) FAIL(rc);
    }
    // groups by grid size, in first-appearance order
    std::map<std::tuple<int, int, int>, int> gi;
    for (int i = 0; i < n; ++i) {
        const auto k = std::make_tuple(setups[i].dims[0], setups[i].dims[1], setups[i].dims[2]);
        auto it = gi.find(k);
        if (it == gi.end()) {
            gi[k] = (int)b->groups.size();
            Group g;
            g.nx = setups[i].dims[0];
            g.ny = setups[i].dims[1];
            g.nz = setups[i].dims[2];
            b->groups.push_back(g);
            it = gi.find(k);
        }
        b->groups[it->second].members.push_back(i);
    }
    if ((rc = upload_tab(b))) FAIL(rc);
    for (Group &g : b->groups)
        if ((rc = build_group_maps(b, g))) FAIL(rc);
    if (tune)
        for (Group &g : b->groups)
            if ((rc = tune_group(b, g))) FAIL(rc);
    b->pred = b->iter_max_any < 8 ? b->iter_max_any : 8;
    if (b->pred < 1) b->pred = 1;
    (void)hipMemGetInfo(&free1, &total1);
    b->dev_bytes = free0 > free1 ? (long)(free0 - free1) : 0;
#undef FAIL
    *out = b;
    return VOCE_OK;
}

// graph of k iterations of every group
static int graph_for(voce_batch *b, int k, hipGraphExec_t *ex)
{
    auto it = b->graphs.find(k);
    if (it != b->graphs.end()) {
        *ex = it->second;
        return 0;
    }
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return VOCE_ERR_DEVICE;
    int rc = 0;
    for (int i = 0; i < k && !rc; ++i)
        for (Group &gr : b->groups)
            if (voce_launch_iteration(&gr.L, b->stream, nullptr)) {
                rc = VOCE_ERR_DEVICE;
                break;
            }
    // the chunk's last Newton step is finalised by a launch of its own
    for (Group &gr : b->groups)
        if (!rc && voce_launch_finalize(&gr.L, b->stream)) rc = VOCE_ERR_DEVICE;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return VOCE_ERR_DEVICE;
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return VOCE_ERR_DEVICE;
    b->graphs[k] = x;
    *ex = x;
    return 0;
}

extern "C" int voce_batch_increment(voce_batch *b, voce_incr_report *reps)
{
    if (!b) return VOCE_ERR_ARG;
    int rc = upload_tab(b);   // members may have changed (diagnostics switched on)
    if (rc) return rc;
    LaunchSet all;
    memset(&all, 0, sizeof(all));
    // begin_increment over every member: one block each
    all.tab = b->dtab;
    all.n_one = b->n;
    all.map_one = b->map_all;
    if (voce_launch_begin_increment(&all, b->stream)) return VOCE_ERR_DEVICE;
    int launched = 0;
    // the previous increment's largest iteration count, then chunks of 4, 8
    // and 16 (one captured graph per length)
    int chunk = b->pred, next_chunk = 4;
    while (launched < b->iter_max_any) {
        hipGraphExec_t ex;
        if ((rc = graph_for(b, chunk, &ex))) return rc;
        if (hipGraphLaunch(ex, b->stream) != hipSuccess) return VOCE_ERR_DEVICE;
        ++b->replays;
        launched += chunk;
        b->iters_launched += chunk;
        if (voce_launch_count_active(b->dtab, b->n, b->d_cnt, b->stream)) return VOCE_ERR_DEVICE;
        if (hipMemcpyAsync(b->h_cnt, b->d_cnt, 2 * sizeof(int), hipMemcpyDeviceToHost, b->stream) != hipSuccess)
            return VOCE_ERR_DEVICE;
        if (hipStreamSynchronize(b->stream) != hipSuccess) return VOCE_ERR_DEVICE;
        if (b->h_cnt[1]) return b->h_cnt[1];
        if (b->h_cnt[0] == 0) break;
        chunk = next_chunk;
        next_chunk = next_chunk < 16 ? 2 * next_chunk : 16;
    }
    for (Group &g : b->groups)
        if (voce_launch_closeout(&g.L, b->stream, 1)) return VOCE_ERR_DEVICE;
    if (hipMemcpyAsync(b->h_ms, b->ms_all, b->n * sizeof(MacroState), hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
        hipMemcpyAsync(b->h_logs, b->logs_all, b->log_total * sizeof(double), hipMemcpyDeviceToHost, b->stream) !=
            hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return VOCE_ERR_DEVICE;
    int mx = 0;
    for (int i = 0; i < b->n; ++i) {
        const MacroState &M = b->h_ms[i];
        if (M.error) return M.error;
        if (M.iter > mx) mx = M.iter;
        voce_ctx *c = b->m[i];
        c->n_steps_done = M.istep;
        c->pred_iters = M.iter > 0 ? M.iter : 1;
        if (reps) {
            const int nl = M.iter < c->log_cap ? M.iter : c->log_cap;
            const double *lg = b->h_logs + b->log_off[i];
            c->log_e.assign(lg, lg + nl);
            c->log_s.assign(lg + c->log_cap, lg + c->log_cap + nl);
            c->log_v.assign(lg + 2 * c->log_cap, lg + 2 * c->log_cap + nl);
            voce_fill_report(c, M, &reps[i]);
        }
    }
    b->last_max_iter = mx;
    b->pred = mx > 0 ? mx : 1;
    return VOCE_OK;
}

extern "C" int voce_batch_size(const voce_batch *b) { return b ? b->n : 0; }

extern "C" voce_ctx *voce_batch_member(voce_batch *b, int i)
{
    if (!b || i < 0 || i >= b->n) return nullptr;
    return b->m[i];
}

extern "C" int voce_batch_info(const voce_batch *b, long out[8])
{
    if (!b || !out) return VOCE_ERR_ARG;
    long vox = 0;
    for (voce_ctx *c : b->m) vox += c->V;
    out[0] = b->n;
    out[1] = (long)b->groups.size();
    out[2] = vox;
    out[3] = b->replays;
    out[4] = (long)b->graphs.size();
    out[5] = b->iters_launched;
    out[6] = b->last_max_iter;
    out[7] = b->dev_bytes;
    return 0;
}
