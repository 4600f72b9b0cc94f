// fsim_simulator.hpp — C++ facade over the C ABI (include/fsim.h) with the
// reference's Simulator interface.
//
// Mirrors genericfsim::simulator::Simulator (src/Simulator/simulator/
// simulator.h:16-99), its SimulatorConfig (:30-39), the obstacle hierarchy
// (src/Simulator/simulator/macGrid/obstacles.hpp:7-65), the MacGrid public
// knobs (macGrid.h:168-182) and the HashedParticles accessors the manager
// uses (hashedParticles.h:17-179): same names, same argument meaning.
//
// Differences a caller sees:
//   * the step runs on the GPU; the host copies of particles and grid fields
//     are caches, refreshed when read after a step (forEach, getParticleAt,
//     cellType...) and pushed before the next step when written;
//   * errors: the reference has no error returns (simulate is void); here a
//     non-OK status of the C ABI is thrown as std::runtime_error carrying
//     fsim_last_error(), which the reference's main() already catches;
//   * vectors are fsim::dvec3 (x, y, z members, like glm::dvec3).
#pragma once

#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fsim.h"

namespace fsim {

struct dvec3 {
    double x = 0, y = 0, z = 0;
    dvec3() = default;
    dvec3(double x_, double y_, double z_) : x(x_), y(y_), z(z_) {}
};
struct ivec3 {
    int x = 0, y = 0, z = 0;
};

// Throws std::runtime_error for a non-OK status.
void check(int status, const fsim_ctx* ctx, const char* what);

namespace simulator {
class Simulator;
}

namespace obstacle {   // obstacles.hpp:7-65

struct Obstacle {
    dvec3 pos, prevPos, speed;
    explicit Obstacle(const dvec3& p = dvec3()) : pos(p), prevPos(p) {}
    virtual ~Obstacle() = default;
    void setNewPos(const dvec3& p) {
        prevPos = pos;
        pos = p;
    }
    void calculateSpeed(double dt) {
        speed = dvec3((pos.x - prevPos.x) / dt, (pos.y - prevPos.y) / dt, (pos.z - prevPos.z) / dt);
    }
    virtual Obstacle* clone() = 0;
    virtual fsim_obstacle toC() const = 0;
};

struct RectengularObstacle : Obstacle {
    const dvec3 size;
    RectengularObstacle(const dvec3& s, const dvec3& p = dvec3()) : Obstacle(p), size(s) {}
    Obstacle* clone() override { return new RectengularObstacle(*this); }
    fsim_obstacle toC() const override;
};

struct SphericalObstacle : Obstacle {
    const double r;
    SphericalObstacle(double r_, const dvec3& p = dvec3()) : Obstacle(p), r(r_) {}
    Obstacle* clone() override { return new SphericalObstacle(*this); }
    fsim_obstacle toC() const override;
};

struct SphericalParticleSource : SphericalObstacle {
    const double particleSpawnRate;
    const double particleSpawnSpeed;
    double lastSpawnFraction = 0;
    SphericalParticleSource(double r_, double rate, double speed_, const dvec3& p = dvec3())
        : SphericalObstacle(r_, p), particleSpawnRate(rate), particleSpawnSpeed(speed_) {}
    Obstacle* clone() override { return new SphericalParticleSource(*this); }
    fsim_obstacle toC() const override;
};

struct SphericalParticleSink : SphericalObstacle {
    SphericalParticleSink(double r_, const dvec3& p = dvec3()) : SphericalObstacle(r_, p) {}
    Obstacle* clone() override { return new SphericalParticleSink(*this); }
    fsim_obstacle toC() const override;
};

}  // namespace obstacle

namespace particles {   // particle.h:8-12, hashedParticles.h

struct Particle {
    dvec3 pos, v;
    dvec3 c[3];   // c[axis] = gradient of v_axis (APIC)
};

class HashedParticles {
public:
    HashedParticles(std::vector<Particle> initial, double particleR);
    // hashedParticles.cpp:23-42: num particles at uniform positions in
    // [dimensions/2 + cellD, dimensions - 1.1 r - cellD] per axis (z = the
    // given z when zConst), at rest.  The positions come from a SplitMix64
    // counter RNG keyed by `seed` in place of std::rand (random.h:13-28).
    HashedParticles(int num, double r, dvec3 dimensions, dvec3 cellD, bool zConst, double z, uint64_t seed = 0);

    // forEach may write (the next step uploads the particles again);
    // forEachConst only reads (extension for snapshot readers).
    void forEach(bool parallel, const std::function<void(Particle&)>& fn);
    void forEachConst(const std::function<void(const Particle&)>& fn);
    Particle& getParticleAt(size_t idx);
    size_t getParticleNum();
    // hashedParticles.cpp:132-151: drops particles from the end, or adds
    // particles at rest at uniform positions in [1.1 r + cellD,
    // dimensions - 1.1 r - cellD] per axis (z = the fixed z when zConst)
    void setParticleNum(size_t n);
    // hashedParticles.cpp:340-353: new grid spacing and size; every particle
    // is clamped into [1.1 r + cellD, dimensions - 1.1 r - cellD]
    void updateGridParams(const dvec3& cellD, const dvec3& dimensions);
    double getParticleR() const { return r_; }
    void setParticleR(double r);
    void addParticles(const std::vector<Particle>& ps);

private:
    friend class fsim::simulator::Simulator;
    std::vector<Particle> host_;
    double r_;
    dvec3 dims_{1, 1, 1}, cellD_{0, 0, 0};   // set by the constructor above or by the Simulator's grid
    bool grid_known_ = false;
    bool zConst_ = false;
    double z_ = 0;
    uint64_t seed_ = 0, draws_ = 0;
    double uniform(double lo, double hi);   // getDoubleInRange (random.h:24-26)
    void setGrid(const dvec3& cellD, const dvec3& dimensions);
    bool host_stale_ = false;      // the device holds newer particles
    bool device_stale_ = true;     // the host was written after the last upload
    bool r_dirty_ = true;
    std::function<void()> refresh_;
    void ensureHost();
};

}  // namespace particles

namespace macgrid {   // macGrid.h

enum class CellType { WATER = FSIM_WATER, AIR = FSIM_AIR, SOLID = FSIM_SOLID };

class MacGrid {
public:
    // macGrid.cpp:9-13: cellD = 1 / resolution on every axis, gridSize =
    // targetDimensions / cellD (truncated), dimensions = gridSize * cellD.
    // 2D: z is 3 cells of the target depth / 3, and dimensions.z = the depth.
    MacGrid(const dvec3& targetDimensions, double resolution, bool twoD = false);
    virtual ~MacGrid() = default;

    const dvec3 cellD;
    const dvec3 cellDInv;
    const ivec3 gridSize;
    const dvec3 dimensions;

    bool isTopOfContainerSolid = false;
    double pressureK = 2.0, averagePressure = 2.0;
    int incompressibilityMaxIterationCount = 80;
    bool pressureEnabled = true;
    double fluidDensity = 1.0;
    double residualTolerance = 1e-6;
    const bool twoD;

    // read-only views of the last step's grid (flat order (x*Ny + y)*Nz + z)
    CellType cellType(int x, int y, int z);
    const std::vector<float>& faceVelocities(int axis);   // Face::v2 after the step
    const std::vector<double>& pressure();

private:
    friend class fsim::simulator::Simulator;
    std::function<void(int, void*, size_t)> fetch_;
    std::vector<uint8_t> label_;
    std::vector<float> v2_[3];
    std::vector<double> p_;
    bool stale_[5] = {true, true, true, true, true};
};

// The two solver plugins the manager chooses between by
// SimulationConfig::GridSolverType (simulationManager.cpp:16-19):
// BridsonSolverGrid solves with MIC(0)-PCG (bridsonSolverGrid.h:10-43),
// BasicMacGrid with red-black SOR sweeps (basicMacGrid.h:7-16).
class BridsonSolverGrid : public MacGrid {
public:
    BridsonSolverGrid(const dvec3& targetDimensions, float resolution, bool twoD, double fluidDensity_ = 1.0)
        : MacGrid(targetDimensions, resolution, twoD) {
        fluidDensity = fluidDensity_;
    }
};
class BasicMacGrid : public MacGrid {
public:
    using MacGrid::MacGrid;
};

}  // namespace macgrid

namespace simulator {

class Simulator {
public:
    enum class P2G2PType { PIC, FLIP, APIC };

    struct SimulatorConfig {   // simulator.h:30-39, same defaults
        P2G2PType transferType = P2G2PType::FLIP;
        float flipRatio = 0.99f;
        float gravity = 150.0f;
        bool gravityEnabled = true, pushParticlesApartEnabled = true;
        bool pushApartEnabled = true;
        bool particleSpawningEnabled = false;
        bool particleDespawningEnabled = false;
        bool stopParticles = false;
    };

    Simulator(SimulatorConfig config, std::shared_ptr<particles::HashedParticles> hashedParticles,
              std::shared_ptr<macgrid::MacGrid> macGrid);
    ~Simulator();
    Simulator(const Simulator&) = delete;
    Simulator& operator=(const Simulator&) = delete;

    void setNewMacGrid(std::shared_ptr<macgrid::MacGrid> macGrid);
    void setNewHashedParticles(std::shared_ptr<particles::HashedParticles> particles);

    // One simulation iteration of dt seconds on the GPU (simulator.cpp:51-104).
    void simulate(double dt);

    // Obstacle speed, prevPos and pos are updated externally (simulator.h:71-74).
    std::vector<std::unique_ptr<obstacle::Obstacle>> obstacles;

    // Sliding averages (factor 0.9) of each stage's duration in microseconds,
    // keyed like the reference's stepDuration (simulator.cpp:31-41).
    std::map<std::string, long long> getStepDuration() const { return stepDuration; }

    // Extensions: the last step's solver statistics and the raw context.
    const fsim_step_stats& lastStats() const { return stats_; }
    fsim_ctx* context() { return ctx_; }
    // Pressure-solve options with no reference counterpart.  A BasicMacGrid
    // grid always solves with FSIM_SOLVER_BASIC_SOR.
    int stopRule = FSIM_STOP_ABS_INF;
    int precond = FSIM_PRECOND_MIC0;
    uint64_t seed = 0;

    SimulatorConfig config;

private:
    std::shared_ptr<particles::HashedParticles> hashedParticles;
    std::shared_ptr<macgrid::MacGrid> macGrid;
    std::map<std::string, long long> stepDuration;
    fsim_ctx* ctx_ = nullptr;
    fsim_step_stats stats_{};

    void createContext();
    fsim_params makeParams() const;
    void pushParams();
    void pushParticlesIfDirty();
    void pullParticles();
};

}  // namespace simulator

}  // namespace fsim
