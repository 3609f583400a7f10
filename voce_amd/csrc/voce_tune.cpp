// voce_tune.cpp -- launch tuning of the FFT kernels per grid size.
//
// The plane kernels (z r2c + y pass, and the inverse) and the x-pass kernel
// each come in launch-shape variants (threads per block, largest tile);
// the x pass also takes a tile width (columns per block) and both take a
// radix set (8/4/2 or 4/2).  The tuner times every valid combination on
// the context's own buffers and keeps the fastest.  Results are cached per
// (nx, ny, nz, precision) for the process and, with VOCE_TUNE_FILE set,
// in a text file, one line per grid size:
//   nx ny nz precision plv xv tk pl_r8 x_r8 us_plane us_x
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>

#include "../../include/voce_batch.h"
#include "voce_host.h"

namespace {

using Key = std::tuple<int, int, int, int>;
std::mutex g_mu;
std::map<Key, voce_tune_choice> g_cache;
bool g_file_read = false;

void read_file_locked()
{
    if (g_file_read) return;
    g_file_read = true;
    const char *fn = getenv("VOCE_TUNE_FILE");
    if (!fn || !*fn) return;
    FILE *f = fopen(fn, "r");
    if (!f) return;
    int nx, ny, nz, pr;
    voce_tune_choice c;
    while (fscanf(f, "%d %d %d %d %d %d %d %d %d %lf %lf", &nx, &ny, &nz, &pr, &c.plv, &c.xv, &c.tk, &c.pl_r8,
                  &c.x_r8, &c.us_plane, &c.us_x) == 11)
        g_cache[Key(nx, ny, nz, pr)] = c;
    fclose(f);
}

void append_file(const Key &k, const voce_tune_choice &c)
{
    const char *fn = getenv("VOCE_TUNE_FILE");
    if (!fn || !*fn) return;
    FILE *f = fopen(fn, "a");
    if (!f) return;
    fprintf(f, "%d %d %d %d %d %d %d %d %d %.3f %.3f\n", std::get<0>(k), std::get<1>(k), std::get<2>(k), std::get<3>(k),
            c.plv, c.xv, c.tk, c.pl_r8, c.x_r8, c.us_plane, c.us_x);
    fclose(f);
}

bool choice_fits(const RveDev &d, const voce_tune_choice &c)
{
    return voce_plane_lds(&d, c.plv) > 0 && voce_xgreen_lds(&d, c.xv, c.tk) > 0 && (c.pl_r8 == 0 || c.pl_r8 == 1) &&
           (c.x_r8 == 0 || c.x_r8 == 1);
}

// average microseconds of `which` over `reps` launches after 2 warm-ups
int time_probe1(voce_ctx *c, int which, int reps, hipEvent_t e0, hipEvent_t e1, double *us)
{
    for (int i = 0; i < 2; ++i)
        if (voce_launch_fft_probe(&c->L, c->stream, which)) return VOCE_ERR_DEVICE;
    if (hipEventRecord(e0, c->stream) != hipSuccess) return VOCE_ERR_DEVICE;
    for (int i = 0; i < reps; ++i)
        if (voce_launch_fft_probe(&c->L, c->stream, which)) return VOCE_ERR_DEVICE;
    if (hipEventRecord(e1, c->stream) != hipSuccess) return VOCE_ERR_DEVICE;
    if (hipEventSynchronize(e1) != hipSuccess) return VOCE_ERR_DEVICE;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return VOCE_ERR_DEVICE;
    *us = 1e3 * ms / reps;
    return 0;
}

// the smaller of two such averages: one disturbed measurement does not
// decide the choice
int time_probe(voce_ctx *c, int which, int reps, hipEvent_t e0, hipEvent_t e1, double *us)
{
    double a = 0, b = 0;
    int rc = time_probe1(c, which, reps, e0, e1, &a);
    if (!rc) rc = time_probe1(c, which, reps, e0, e1, &b);
    *us = a < b ? a : b;
    return rc;
}

int apply(voce_ctx *c, const voce_tune_choice &ch)
{
    c->d.plv = ch.plv;
    c->d.xv = ch.xv;
    c->d.tk = ch.tk;
    c->d.pl_r8 = ch.pl_r8;
    c->d.x_r8 = ch.x_r8;
    return voce_sync_tab(c);
}

int measure(voce_ctx *c, voce_tune_choice *best)
{
    RveDev &d = c->d;
    const long V = c->V;
    const size_t eb = 6 * (size_t)V * sizeof(double);
    const voce_tune_choice keep = {d.plv, d.xv, d.tk, d.pl_r8, d.x_r8, 0.0, 0.0};
    double *e6_save = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    const size_t act_off = offsetof(MacroState, active);
    int one = 1, zero = 0;
    if (hipMalloc((void **)&e6_save, eb) != hipSuccess) return VOCE_ERR_NOMEM;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = VOCE_ERR_DEVICE;
    // the probes run the iteration kernels' FFT stages: they overwrite the
    // spectra (scratch between increments) and update E6, which is saved
    if (!rc && hipMemcpyAsync(e6_save, d.E6, eb, hipMemcpyDeviceToDevice, c->stream) != hipSuccess) rc = VOCE_ERR_DEVICE;
    if (!rc && hipMemcpyAsync((char *)d.ms + act_off, &one, sizeof(int), hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = VOCE_ERR_DEVICE;
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = VOCE_ERR_DEVICE;
    const int reps = V >= 1000000 ? 3 : 6;
    double best_pl = 1e300, best_x = 1e300;
    voce_tune_choice cand = keep, b = keep;
    // plane kernels: variant x radix set (x-pass choice held at its current value)
    for (int plv = 0; plv < VOCE_N_PLV && !rc; ++plv)
        for (int r8 = 1; r8 >= ((plv == 4 || plv == 6) ? 1 : 0) && !rc; --r8) {   // plv 4 and 6 have fixed radices
            cand = keep;
            cand.plv = plv;
            cand.pl_r8 = r8;
            if (!choice_fits(d, cand)) continue;
            if ((rc = apply(c, cand))) break;
            double a = 0, bb = 0;
            if ((rc = time_probe(c, 0, reps, e0, e1, &a))) break;
            if ((rc = time_probe(c, 2, reps, e0, e1, &bb))) break;
            if (a + bb < best_pl) {
                best_pl = a + bb;
                b.plv = plv;
                b.pl_r8 = r8;
            }
        }
    // x pass: variant x tile x radix set
    const int tks[5] = {1, 2, 4, 8, 16};
    for (int xv = 0; xv < VOCE_N_XV && !rc; ++xv)
        for (int ti = 0; ti < 5 && !rc; ++ti)
            for (int r8 = 1; r8 >= (xv >= 3 ? 1 : 0) && !rc; --r8) {   // xv 3 and 4 have fixed radices
                cand = keep;
                cand.xv = xv;
                cand.tk = tks[ti];
                cand.x_r8 = r8;
                if (!choice_fits(d, cand)) continue;
                if ((rc = apply(c, cand))) break;
                double a = 0;
                if ((rc = time_probe(c, 1, reps, e0, e1, &a))) break;
                if (a < best_x) {
                    best_x = a;
                    b.xv = xv;
                    b.tk = tks[ti];
                    b.x_r8 = r8;
                }
            }
    b.us_plane = best_pl;
    b.us_x = best_x;
    // restore the state
    if (hipMemcpyAsync(d.E6, e6_save, eb, hipMemcpyDeviceToDevice, c->stream) != hipSuccess && !rc) rc = VOCE_ERR_DEVICE;
    if (hipMemcpyAsync((char *)d.ms + act_off, &zero, sizeof(int), hipMemcpyHostToDevice, c->stream) != hipSuccess && !rc)
        rc = VOCE_ERR_DEVICE;
    if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = VOCE_ERR_DEVICE;
    (void)hipFree(e6_save);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) {
        (void)apply(c, keep);
        return rc;
    }
    if (best_pl >= 1e300 || best_x >= 1e300) {
        (void)apply(c, keep);
        return VOCE_ERR_UNSUPPORTED_GRID;
    }
    *best = b;
    return 0;
}

}  // namespace

extern "C" int voce_apply_tuning(voce_ctx *c, const voce_tune_choice *ch)
{
    if (!c || !ch) return VOCE_ERR_ARG;
    if (!choice_fits(c->d, *ch)) return VOCE_ERR_ARG;
    return apply(c, *ch);
}

extern "C" int voce_get_tuning(const voce_ctx *c, voce_tune_choice *out)
{
    if (!c || !out) return VOCE_ERR_ARG;
    *out = voce_tune_choice{c->d.plv, c->d.xv, c->d.tk, c->d.pl_r8, c->d.x_r8, 0.0, 0.0};
    return 0;
}

extern "C" int voce_autotune(voce_ctx *c, voce_tune_choice *out)
{
    if (!c) return VOCE_ERR_ARG;
    const Key k(c->dims[0], c->dims[1], c->dims[2], c->precision);
    voce_tune_choice ch;
    bool have = false;
    {
        std::lock_guard<std::mutex> g(g_mu);
        read_file_locked();
        auto it = g_cache.find(k);
        if (it != g_cache.end() && choice_fits(c->d, it->second)) {
            ch = it->second;
            have = true;
        }
    }
    if (!have) {
        int rc = measure(c, &ch);
        if (rc) return rc;
        std::lock_guard<std::mutex> g(g_mu);
        g_cache[k] = ch;
        append_file(k, ch);
    }
    int rc = apply(c, ch);
    if (rc) return rc;
    if (out) *out = ch;
    return 0;
}
