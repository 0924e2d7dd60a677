// Exercises include/pm_universe.hpp the way the reference's demo drives its
// Universe class (main.cu:52-145): construct, clear, add galaxies (one with a
// black hole), generate frames on the stepping thread, pop them, stop.
// Checks the addGalaxy initial-condition law (cuh:1071-1168) on the state read
// back before any step, and the frame format.  Exit status 0 = pass.
#include <cmath>
#include <cstdio>
#include <thread>
#include <vector>

#include "pm_universe.hpp"

static int fails = 0;
#define CHECK(c, ...)                             \
    do {                                          \
        if (!(c)) {                               \
            std::printf("FAIL: " __VA_ARGS__);    \
            std::printf("\n");                    \
            ++fails;                              \
        }                                         \
    } while (0)

int main() {
    pm::UniverseConstants k;
    k.N = 256;
    k.dt = 1.0f;
    k.seed = 7;
    const int total = 20000;
    const int dev[1] = {0};
    const float perf[1] = {1.0f};
    pm::Universe<2> u(total, dev, perf, false, k);
    CHECK(u.getLatticeSize() == 256, "lattice size");

    u.clear();
    const int n1 = 12000, n2 = total - n1;
    const float R = 0.2f, w = 1.0f, mpp = 1.0f, vcx = 0.001f;
    u.addGalaxy(n1, 0.35f, 0.5f, w, mpp, R, vcx, 0.0f, true);
    u.addGalaxy(n2, 0.65f, 0.5f, w, mpp, 0.15f, -vcx, 0.0f, false);

    std::vector<float> x, y, xo, yo;
    u.readState(x, y, xo, yo);
    const std::vector<float>& m = u.masses();
    const float N = 256.0f, dt = k.dt, Rc = N * R;
    const float totalMass = mpp * n1 + n1 * mpp / 1000.0f;
    double worst_r = 0, worst_v = 0;
    for (int i = 0; i < n1 - 1; ++i) {
        const float cx = 0.35f * N, cy = 0.5f * N;
        const float dx = x[i] - cx, dy = y[i] - cy, r = std::sqrt(dx * dx + dy * dy);
        // radius in [0.1 R, R]
        worst_r = std::fmax(worst_r, std::fmax(0.1f * Rc - r, r - Rc));
        // velocity relative to the centre of mass: tangential, |v| = r * (w/dt) * sqrt(4.4e-6 M r / R^2)
        const float vx = (x[i] - xo[i]) / dt - (vcx / dt) * N, vy = (y[i] - yo[i]) / dt;
        const float want = r * (w / dt) * std::sqrt(0.0000044f * totalMass * r / (Rc * Rc));
        const float got = std::sqrt(vx * vx + vy * vy);
        worst_v = std::fmax(worst_v, std::fabs(got - want) / want);
        CHECK(std::fabs(vx * dx + vy * dy) <= 1e-2f * got * r + 1e-6f, "velocity of particle %d not tangential", i);
        CHECK(m[i] == mpp / total, "mass of particle %d", i);
    }
    CHECK(worst_r <= 1e-3, "radius outside [0.1R, R] by %g", worst_r);
    CHECK(worst_v <= 2e-3, "tangential speed law off by %g", worst_v);
    // black hole: last particle of galaxy 1 at the centre with 1/1000 of the galaxy's mass
    const int bh = n1 - 1;
    CHECK(std::fabs(x[bh] - 0.35f * N) < 1e-3f && std::fabs(y[bh] - 0.5f * N) < 1e-3f, "black hole position");
    CHECK(std::fabs(m[bh] - (n1 * mpp / 1000.0f) / total) <= 1e-6f * m[bh], "black hole mass");

    u.nBodyStartGeneratingFrames();
    int got = 0;
    float first_max = 0.f;
    for (int spin = 0; spin < 20000 && got < 5; ++spin) {
        bool ready = false;
        std::vector<float>& f = u.popFrame(ready);
        if (ready) {
            CHECK(f.size() == (size_t)(N * N), "frame size");
            float lo = 1e30f, hi = -1e30f;
            for (float v : f) { lo = std::fmin(lo, v); hi = std::fmax(hi, v); }
            CHECK(lo >= 0.0f && hi <= 1.8f, "frame range [%g, %g]", lo, hi);
            if (got == 0) first_max = hi;
            ++got;
        } else {
            std::this_thread::sleep_for(std::chrono::milliseconds(1));
        }
    }
    u.nBodyStop();
    CHECK(got == 5, "got %d frames", got);
    CHECK(first_max == 1.8f, "first frame max %g", first_max);
    std::printf("%s (%d frames)\n", fails ? "universe FAILED" : "universe ok", got);
    return fails ? 1 : 0;
}
