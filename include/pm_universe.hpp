// pm_universe.hpp -- header-only drop-in for the reference's simulator class
// `template<int NUM_TIME_STEPS_PER_RENDER> struct Universe` (CosmosCuda.cuh:
// 830-1428), built on the C ABI of pm.h.  Same public members, argument
// meanings and host-side state:
//
//   Universe(particles, devices, performances, lowAccuracy)   cuh:881-1035
//   clear()                                                   cuh:1036-1069
//   addGalaxy(n, cx, cy, angularVelocity, massPerParticle,
//             normalizedRadius, comVx, comVy, blackHole)      cuh:1071-1168
//   getLatticeSize()                                          cuh:1310-1312
//   nBodyStartGeneratingFrames() / popFrame(ready) /
//   nBodyStop()                                               cuh:1318-1395
//
// Differences a caller sees:
//   * one device: `devices` and `performances` have one entry (the reference
//     splits particles over several GPUs, cuh:882-921);
//   * the constants of the reference's `namespace Constants` (cuh:29-74) are a
//     runtime `UniverseConstants` argument (defaults: N 2048, dt 50, G 1,
//     G(0) 3, 40 buffered frames, blur 3);
//   * random numbers come from a seeded generator (the reference seeds from
//     std::random_device), so runs repeat;
//   * errors throw std::runtime_error with pm_last_error's text instead of
//     exit() (cuh:19-27);
//   * addGalaxy while frames are being generated is serialised with the
//     stepping thread (the reference races, cuh:1151-1155 vs 1337-1341).
// The position-Verlet state (x, xOld) of the reference is kept on the host and
// converted by pm_load_verlet / pm_read_verlet.
#ifndef COSMOS_PM_UNIVERSE_HPP
#define COSMOS_PM_UNIVERSE_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "pm.h"

namespace pm {

struct UniverseConstants {
    int N = 2048;                 // lattice size (power of two, 64..8192)
    float dt = 50.0f;             // time step
    float gravityMultiplier = 1.0f;
    double greenOrigin = 3.0;     // G(0) of the 1/r Green's function (cuh:113-116)
    int maxFramesBuffered = 40;   // frame ring (cuh:54)
    int blurR = 3;                // render filter size (cuh:55)
    std::uint64_t seed = 1;       // initial-condition generator seed
};

template <int NUM_TIME_STEPS_PER_RENDER>
class Universe {
public:
    Universe(int particles, const int (&cudaDevices)[1], const float (&devicePerformances)[1], bool lowAccuracy,
             const UniverseConstants& constants = UniverseConstants())
        : c_(constants), rng_(constants.seed) {
        (void)devicePerformances;
        if (particles <= 0) throw std::runtime_error("Universe: particles must be positive");
        total_ = particles;
        x_.resize(particles);
        y_.resize(particles);
        xOld_.resize(particles);
        yOld_.resize(particles);
        m_.resize(particles);
        pm_config cfg;
        pm_config_default(&cfg);
        cfg.n_lattice = c_.N;
        cfg.n_particles = particles;
        cfg.mode = lowAccuracy ? PM_NGP : PM_CIC;       // accuracy = !lowAccuracy (cuh:890)
        cfg.dt = c_.dt;
        cfg.gravity = c_.gravityMultiplier;
        cfg.green_origin = c_.greenOrigin;
        cfg.wrap_mode = PM_WRAP_REFERENCE;             // reproduce cuh:288-303
        cfg.device = cudaDevices[0];
        check(pm_create(&cfg, &h_), nullptr);
        // uniform in [0.01 N, 0.99 N)^2 at rest, mass 1/particles (cuh:931-952)
        std::uniform_real_distribution<float> u(0.0f, 1.0f);
        const float n = (float)c_.N;
        for (int i = 0; i < particles; ++i) {
            x_[i] = 0.01f * n + 0.98f * u(rng_) * n;
            y_[i] = 0.01f * n + 0.98f * u(rng_) * n;
            xOld_[i] = x_[i];
            yOld_[i] = y_[i];
            m_[i] = 1.0f / (float)particles;
        }
        particleCounter_ = particles;                  // as cuh:967: the next addGalaxy first wraps to 0
        upload();
    }

    Universe(const Universe&) = delete;
    Universe& operator=(const Universe&) = delete;

    ~Universe() {
        if (running_) nBodyStop();
        if (h_) pm_destroy(h_);
    }

    // Moves all particles out of the simulated area (parked at -10, never
    // deposited or moved; cuh:1036-1069).  Masses are kept.
    void clear() {
        std::lock_guard<std::mutex> lk(sim_);
        for (int i = 0; i < total_; ++i) x_[i] = y_[i] = xOld_[i] = yOld_[i] = -10.0f;
        upload();
        particleCounter_ = 0;
    }

    // Adds a galaxy of n particles (cuh:1071-1168): radius uniform in
    // [0.1 R, R] with R = N * normalizedRadius, uniform angle, tangential
    // speed sqrt(4.4e-6 * totalMass * r / R^2) * angularVelocity / dt plus the
    // centre-of-mass velocity, masses massPerParticle / particles, and with
    // blackHole the last particle at the centre carrying 1/1000 of the
    // galaxy's mass.  Other particles keep their current state.
    void addGalaxy(int n, float normalizedCenterX = 0.5f, float normalizedCenterY = 0.5f,
                   float angularVelocity = 1.0f, float massPerParticle = 1.0f, float normalizedRadius = 0.3f,
                   float centerOfMassVelocityX = 0.0f, float centerOfMassVelocityY = 0.0f, bool blackHole = false) {
        std::lock_guard<std::mutex> lk(sim_);
        download();
        const float N = (float)c_.N, dt = c_.dt;
        const int maxCount = std::min(particleCounter_ + n, total_);
        const int count = maxCount - particleCounter_;
        angularVelocity /= dt;
        const float radiusSquared = (N * normalizedRadius) * (N * normalizedRadius);
        const float totalMass =
            massPerParticle * (float)count + (blackHole ? (float)count * massPerParticle / 1000.0f : 0.0f);
        centerOfMassVelocityX /= dt;
        centerOfMassVelocityY /= dt;
        std::uniform_real_distribution<float> u(0.0f, 1.0f);
        const float cx = normalizedCenterX * N, cy = normalizedCenterY * N;
        for (int i = particleCounter_; i < maxCount; ++i) {
            const float r = (0.1f + 0.9f * u(rng_)) * (N * normalizedRadius);
            const float a = (float)(3.14159265358979323846 * 2.0 * u(rng_));
            const float speed = std::sqrt(0.0000044f * totalMass * r / radiusSquared);
            x_[i] = r * std::cos(a) + cx;
            y_[i] = r * std::sin(a) + cy;
            const float vecX = x_[i] - cx, vecY = y_[i] - cy;
            xOld_[i] = -dt * ((vecY * angularVelocity) * speed + centerOfMassVelocityX * N) + x_[i];
            yOld_[i] = -dt * ((-vecX * angularVelocity) * speed + centerOfMassVelocityY * N) + y_[i];
            m_[i] = massPerParticle / (float)total_;
            if (blackHole && i == maxCount - 1) {
                x_[i] = cx;
                y_[i] = cy;
                xOld_[i] = x_[i] - centerOfMassVelocityX * N * dt;
                yOld_[i] = y_[i] - centerOfMassVelocityY * N * dt;
                m_[i] = ((float)count * massPerParticle / 1000.0f) / (float)total_;
            }
        }
        particleCounter_ = maxCount;
        upload();
        if (particleCounter_ >= total_) particleCounter_ = 0;
    }

    int getLatticeSize() const { return c_.N; }

    // Starts the stepping thread: NUM_TIME_STEPS_PER_RENDER steps, then one
    // density frame into the ring (dropped when the ring is full), until
    // nBodyStop (cuh:1318-1361).
    void nBodyStartGeneratingFrames() {
        if (running_) return;
        {
            std::lock_guard<std::mutex> lk(frames_mu_);
            frames_.assign((size_t)c_.maxFramesBuffered, std::vector<float>((size_t)c_.N * c_.N));
            pushCtr_ = popCtr_ = 0;
            working_ = true;
        }
        running_ = true;
        worker_ = std::thread([this] { loop(); });
    }

    // Returns the oldest buffered frame (N*N floats, index i + j*N) and sets
    // ready; ready = false when none is buffered (cuh:1363-1380).
    std::vector<float>& popFrame(bool& ready) {
        std::lock_guard<std::mutex> lk(frames_mu_);
        if (result_.size() != (size_t)c_.N * c_.N) result_.resize((size_t)c_.N * c_.N);
        if (popCtr_ < pushCtr_) {
            result_.swap(frames_[(size_t)(popCtr_ % c_.maxFramesBuffered)]);
            ++popCtr_;
            ready = true;
        } else {
            ready = false;
        }
        return result_;
    }

    // Stops the stepping thread after its current batch (cuh:1381-1395).
    void nBodyStop() {
        {
            std::lock_guard<std::mutex> lk(frames_mu_);
            working_ = false;
        }
        if (worker_.joinable()) worker_.join();
        running_ = false;
        if (!error_.empty()) throw std::runtime_error(error_);
    }

    // Additions: the handle (for pm_* calls between batches) and a blocking
    // read of the current Verlet state into the host arrays.
    pm_handle handle() const { return h_; }
    void readState(std::vector<float>& x, std::vector<float>& y, std::vector<float>& xOld,
                   std::vector<float>& yOld) {
        std::lock_guard<std::mutex> lk(sim_);
        download();
        x = x_;
        y = y_;
        xOld = xOld_;
        yOld = yOld_;
    }
    const std::vector<float>& masses() const { return m_; }

private:
    static void check(pm_status st, pm_handle h) {
        if (st != PM_OK) throw std::runtime_error(std::string("pm: ") + pm_last_error(h));
    }
    void upload() {
        check(pm_load_verlet(h_, x_.data(), y_.data(), xOld_.data(), yOld_.data(), m_.data(), (int64_t)total_), h_);
    }
    void download() {
        check(pm_read_verlet(h_, x_.data(), y_.data(), xOld_.data(), yOld_.data()), h_);
    }
    void loop() {
        std::vector<float> frame((size_t)c_.N * c_.N);
        bool first = true;
        for (;;) {
            {
                std::lock_guard<std::mutex> lk(sim_);
                pm_status st = pm_step(h_, NUM_TIME_STEPS_PER_RENDER);
                if (st == PM_OK) st = pm_frame(h_, c_.blurR, first ? 1 : 0, frame.data());
                if (st != PM_OK) {
                    error_ = std::string("pm: ") + pm_last_error(h_);
                    return;
                }
                first = false;
            }
            std::lock_guard<std::mutex> lk(frames_mu_);
            if (pushCtr_ < popCtr_ + c_.maxFramesBuffered - 1) {      // cuh:1302-1307
                frames_[(size_t)(pushCtr_ % c_.maxFramesBuffered)].swap(frame);
                frame.resize((size_t)c_.N * c_.N);
                ++pushCtr_;
            }
            if (!working_) return;
        }
    }

    UniverseConstants c_;
    pm_handle h_ = nullptr;
    std::mt19937 rng_;
    int total_ = 0, particleCounter_ = 0;
    std::vector<float> x_, y_, xOld_, yOld_, m_;
    std::mutex sim_;                      // the handle and the host arrays
    std::mutex frames_mu_;                // the frame ring
    std::vector<std::vector<float>> frames_;
    std::vector<float> result_;
    long long pushCtr_ = 0, popCtr_ = 0;
    bool working_ = false, running_ = false;
    std::thread worker_;
    std::string error_;
};

}  // namespace pm

#endif  // COSMOS_PM_UNIVERSE_HPP
