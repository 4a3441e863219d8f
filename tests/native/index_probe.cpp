// Host build of the kernels' lattice index helpers (heatflow-mpm_amd/csrc/hfm_device.h), driven with
// points inside, on and outside the box and with huge and non-finite coordinates.  Built with the
// address and undefined-behaviour sanitizers by tests/test_index_host.py: every lattice read must stay
// inside its exactly sized heap array and no conversion or integer operation may be undefined.
// Prints one line per sample: the three velocity components and the cell sample (or "nan").
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <vector>

#include "../../heatflow-mpm_amd/csrc/hfm_device.h"

int main() {
    const int n = 5;             // odd and even lattice lengths (n, n + 1)
    const float h = 0.01f, dt = 2e-3f;
    const long N = (long)n * n * n, F = (long)(n + 1) * n * n;
    // exactly sized heap arrays, so any read outside them is reported
    float* u = (float*)malloc(sizeof(float) * F);
    float* v = (float*)malloc(sizeof(float) * F);
    float* w = (float*)malloc(sizeof(float) * F);
    float* T = (float*)malloc(sizeof(float) * N);
    for (long q = 0; q < F; ++q) {
        u[q] = 0.1f * (float)(q % 7);
        v[q] = -0.05f * (float)(q % 5);
        w[q] = 0.02f * (float)(q % 3);
    }
    for (long q = 0; q < N; ++q) T[q] = 20.0f + (float)q;
    const float inf = std::numeric_limits<float>::infinity(), nan = std::numeric_limits<float>::quiet_NaN();
    const float big[] = {0.0f, 0.5f * h, h, 2.5f * h, (float)n * h, -h, 1e4f, -1e4f, 3e7f, -3e7f, 2.2e9f * h,
                         -2.2e9f * h, 1e30f, -1e30f, 3.4e38f, -3.4e38f, inf, -inf, nan};
    const int nb = (int)(sizeof(big) / sizeof(big[0]));
    // pos / h by division, then by the corrected product of the reciprocal (the kernels' fast form)
    const hfm::CellSize forms[2] = {hfm::CellSize(h), hfm::CellSize(h, 1.0f / h, 1)};
    for (const hfm::CellSize& hs : forms)
        for (int a = 0; a < nb; ++a)
            for (int b = 0; b < nb; b += 3)
                for (int c = 0; c < nb; c += 5) {
                    float x = big[a], y = big[b], z = big[c];
                    float vx, vy, vz;
                    hfm::sample_velocity(u, v, w, x, y, z, hs, n, vx, vy, vz);
                    float bx = x, by = y, bz = z;
                    hfm::backtrace(u, v, w, bx, by, bz, hs, n, dt);
                    const float t = hfm::sample(T, bx, by, bz, 0.5f, 0.5f, 0.5f, hs, n, n, n);
                    std::printf("%g %g %g %g\n", (double)vx, (double)vy, (double)vz, (double)t);
                }
    // velocities that are themselves huge or non-finite (a diverged pressure projection)
    const float vel[] = {1e30f, -1e30f, inf, -inf, nan};
    for (float s : vel) {
        for (long q = 0; q < F; ++q) {
            u[q] = s;
            v[q] = -s;
            w[q] = s;
        }
        for (int a = 0; a < nb; a += 2) {
            float x = big[a], y = 0.5f * h, z = 0.5f * h;
            hfm::backtrace(u, v, w, x, y, z, h, n, dt);
            std::printf("%g\n", (double)hfm::sample(T, x, y, z, 0.5f, 0.5f, 0.5f, h, n, n, n));
        }
    }
    free(u);
    free(v);
    free(w);
    free(T);
    std::printf("done\n");
    return 0;
}
