// voce_gen.cpp -- synthetic microstructures.  No counterpart in the
// reference, whose testf.c:97-130 only writes lamellae.
//
// Periodic Voronoi grains: seeds uniform in [0,N)^3 from SplitMix64, voxel
// centres at (i+0.5, j+0.5, k+0.5), minimum-image distance, ties to the
// lowest seed id; then one Bunge triple per grain, uniform on SO(3), rounded
// as a mesh file prints it ("%.6f").  The voxel loop runs on host threads.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <thread>
#include <vector>

#include "../../include/voce_batch.h"

namespace {

// SplitMix64 (Steele, Lea, Flood 2014) and its 53-bit uniform double
uint64_t sm64_next(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double sm64_uniform(uint64_t *s) { return (double)(sm64_next(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

extern "C" int voce_voronoi(const int dims[3], int n_grains, uint64_t seed, int *grain, double *euler_by_grain)
{
    const int nx = dims ? dims[0] : 0, ny = dims ? dims[1] : 0, nz = dims ? dims[2] : 0;
    if (nx <= 0 || ny <= 0 || nz <= 0 || n_grains <= 0 || !grain || !euler_by_grain) return VOCE_ERR_ARG;
    const double L[3] = {(double)nx, (double)ny, (double)nz};
    const long V = (long)nx * ny * nz;
    std::vector<double> sd(3 * (size_t)n_grains);
    uint64_t st = seed;
    for (int g = 0; g < n_grains; ++g) {
        sd[3 * g] = sm64_uniform(&st) * nx;
        sd[3 * g + 1] = sm64_uniform(&st) * ny;
        sd[3 * g + 2] = sm64_uniform(&st) * nz;
    }
    for (int g = 0; g < n_grains; ++g) {
        double e[3];
        e[0] = 360.0 * sm64_uniform(&st);
        e[1] = acos(2.0 * sm64_uniform(&st) - 1.0) * 180.0 / 3.141592653589793;
        e[2] = 360.0 * sm64_uniform(&st);
        for (int k = 0; k < 3; ++k) {
            char buf[64];
            snprintf(buf, sizeof(buf), "%.6f", e[k]);
            euler_by_grain[3 * g + k] = strtod(buf, nullptr);
        }
    }
    auto work = [&](long v0, long v1) {
        for (long v = v0; v < v1; ++v) {
            const double p[3] = {(double)(v / ((long)ny * nz)) + 0.5, (double)((v / nz) % ny) + 0.5,
                                 (double)(v % nz) + 0.5};
            double best = INFINITY;
            int arg = 0;
            for (int q = 0; q < n_grains; ++q) {
                double d[3];
                for (int k = 0; k < 3; ++k) {
                    d[k] = p[k] - sd[3 * q + k];
                    d[k] -= L[k] * nearbyint(d[k] / L[k]);   // minimum image, ties to even
                }
                const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                if (r2 < best) {   // strict: ties stay with the lower seed id
                    best = r2;
                    arg = q;
                }
            }
            grain[v] = arg + 1;
        }
    };
    unsigned nt = std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    if (nt > 16) nt = 16;
    if ((long)nt * 4096 > V) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) th.emplace_back(work, V * t / nt, V * (t + 1) / nt);
    for (auto &x : th) x.join();
    return 0;
}
