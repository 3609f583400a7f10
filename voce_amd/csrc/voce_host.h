// voce_host.h -- host-side context object shared by the single-RVE C ABI
// (voce_ctx.cpp), the batch (voce_batch.cpp) and the request server
// (voce_serve.cpp).  C++ only; not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/voce.h"
#include "../../include/voce_prof.h"
#include "voce_internal.h"

struct voce_ctx {
    int dims[3];
    long V;
    int precision;
    int n_steps_done;
    int diag;
    int pred_iters;
    int update_flag, creep, iter_max;
    hipStream_t stream;
    bool own_stream;       // false: the stream belongs to a batch or server
    RveDev d;              // host copy of the descriptor
    RveDev *dtab;          // device copy (a one-entry RVE table)
    RveDev *on_tab;        // non-null: a host mirror slot of a batch table kept in step with d
    LaunchSet L;           // launch sequence over that table
    MacroState *h_ms;      // pinned mirror of the device MacroState
    int *h_ctl;            // pinned: istep, iter, iter_total, active, error
    RveDev *h_tab;         // pinned staging for a descriptor written to dtab
    int device;            // HIP device the context was created on
    int two_crss_any;
    int log_cap;
    double *logs_dev;      // 3 x log_cap: Err_e, Err_s, s_vm per iteration
    std::vector<int> grain, phase_id;
    std::vector<double> log_e, log_s, log_v;
    std::vector<void *> allocs;
    size_t arena_bytes;    // device bytes of the context's arena
    // kernel timing (include/voce_prof.h)
    int prof;
    std::vector<hipEvent_t> ev;    // 8 per enqueued iteration (start/end per kernel)
    double prof_ms[VOCE_PROF_NK];
    long prof_n[VOCE_PROF_NK];
};

// How a context is placed when a batch or server owns it.
struct voce_build_opts {
    hipStream_t stream;    // shared stream (null: the context creates its own)
    MacroState *ms_dev;    // MacroState slot in a caller-owned array (null: own)
    double *logs_dev;      // with ms_dev: 3 x iter_max doubles for the logs
    RveDev *on_tab;        // host mirror slot of the caller's descriptor table
    int no_workspace;      // 1: no spectra / partial-sum buffers (assigned per request)
};

int voce_ctx_build(const voce_setup *su, const voce_build_opts *opt, voce_ctx **out);
int voce_dev_alloc(voce_ctx *c, void **p, size_t bytes);
int voce_sync_tab(voce_ctx *c);
int voce_read_logs(voce_ctx *c, int n);
int voce_run_increment(voce_ctx *c, voce_incr_report *rep, int timing);
size_t voce_ws_bytes(const RveDev &d);       // per-increment workspace bytes of one RVE
void voce_ws_carve(RveDev &d, char *p);      // point d's workspace fields into p
void voce_fill_report(voce_ctx *c, const MacroState &M, voce_incr_report *rep);
