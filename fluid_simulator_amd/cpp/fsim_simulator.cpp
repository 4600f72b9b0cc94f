// fsim_simulator.cpp — the C++ facade's forwarding to the C ABI.
#include "fsim_simulator.hpp"

#include <algorithm>
#include <cmath>

namespace fsim {

void check(int status, const fsim_ctx* ctx, const char* what) {
    if (status == FSIM_OK) return;
    std::string msg = std::string(what) + ": " + fsim_status_string(status);
    if (ctx) {
        const char* e = fsim_last_error(ctx);
        if (e && *e) msg += std::string(" (") + e + ")";
    }
    throw std::runtime_error(msg);
}

namespace obstacle {

static fsim_obstacle base_obstacle(const Obstacle& o, int kind) {
    fsim_obstacle c{};
    c.kind = kind;
    c.pos[0] = o.pos.x;
    c.pos[1] = o.pos.y;
    c.pos[2] = o.pos.z;
    c.speed[0] = o.speed.x;
    c.speed[1] = o.speed.y;
    c.speed[2] = o.speed.z;
    return c;
}
fsim_obstacle RectengularObstacle::toC() const {
    fsim_obstacle c = base_obstacle(*this, FSIM_OBS_BOX);
    c.size[0] = size.x;
    c.size[1] = size.y;
    c.size[2] = size.z;
    return c;
}
fsim_obstacle SphericalObstacle::toC() const {
    fsim_obstacle c = base_obstacle(*this, FSIM_OBS_SPHERE);
    c.r = r;
    return c;
}
fsim_obstacle SphericalParticleSource::toC() const {
    fsim_obstacle c = base_obstacle(*this, FSIM_OBS_SOURCE);
    c.r = r;
    c.spawn_rate = particleSpawnRate;
    c.spawn_speed = particleSpawnSpeed;
    c.last_spawn_fraction = lastSpawnFraction;
    return c;
}
fsim_obstacle SphericalParticleSink::toC() const {
    fsim_obstacle c = base_obstacle(*this, FSIM_OBS_SINK);
    c.r = r;
    return c;
}

}  // namespace obstacle

namespace particles {

HashedParticles::HashedParticles(std::vector<Particle> initial, double particleR)
    : host_(std::move(initial)), r_(particleR) {}

HashedParticles::HashedParticles(int num, double r, dvec3 dimensions, dvec3 cellD, bool zConst, double z,
                                 uint64_t seed)
    : r_(r), dims_(dimensions), cellD_(cellD), grid_known_(true), zConst_(zConst), z_(z), seed_(seed) {
    host_.reserve((size_t)std::max(num, 0));
    for (int p = 0; p < num; p++) {
        Particle q;
        q.pos.x = uniform(dims_.x * 0.5 + cellD_.x, dims_.x - 1.1 * r_ - cellD_.x);
        q.pos.y = uniform(dims_.y * 0.5 + cellD_.y, dims_.y - 1.1 * r_ - cellD_.y);
        q.pos.z = zConst_ ? z_ : uniform(dims_.z * 0.5 + cellD_.z, dims_.z - 1.1 * r_ - cellD_.z);
        host_.push_back(q);
    }
}

// SplitMix64 (Steele, Lea, Flood 2014), the n-th output for the seed; the
// top 53 bits as a double in [0, 1)
static double splitmix_unit(uint64_t seed, uint64_t n) {
    uint64_t z = seed + (n + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

double HashedParticles::uniform(double lo, double hi) { return lo + (hi - lo) * splitmix_unit(seed_, draws_++); }

void HashedParticles::setGrid(const dvec3& cellD, const dvec3& dimensions) {
    cellD_ = cellD;
    dims_ = dimensions;
    grid_known_ = true;
}

void HashedParticles::ensureHost() {
    if (host_stale_ && refresh_) refresh_();
    host_stale_ = false;
}
void HashedParticles::forEach(bool, const std::function<void(Particle&)>& fn) {
    ensureHost();
    for (Particle& p : host_) fn(p);
    device_stale_ = true;
}
void HashedParticles::forEachConst(const std::function<void(const Particle&)>& fn) {
    ensureHost();
    for (const Particle& p : host_) fn(p);
}
Particle& HashedParticles::getParticleAt(size_t idx) {
    ensureHost();
    device_stale_ = true;   // the caller may write through the reference
    return host_.at(idx);
}
size_t HashedParticles::getParticleNum() {
    ensureHost();
    return host_.size();
}
void HashedParticles::setParticleNum(size_t n) {
    ensureHost();
    if (n > host_.size() && !grid_known_)
        throw std::runtime_error("setParticleNum: the grid of the particles is not known yet");
    if (n < host_.size()) host_.resize(n);
    while (host_.size() < n) {
        Particle q;
        q.pos.x = uniform(1.1 * r_ + cellD_.x, dims_.x - 1.1 * r_ - cellD_.x);
        q.pos.y = uniform(1.1 * r_ + cellD_.y, dims_.y - 1.1 * r_ - cellD_.y);
        q.pos.z = zConst_ ? z_ : uniform(1.1 * r_ + cellD_.z, dims_.z - 1.1 * r_ - cellD_.z);
        host_.push_back(q);
    }
    device_stale_ = true;
}

void HashedParticles::updateGridParams(const dvec3& cellD, const dvec3& dimensions) {
    ensureHost();
    setGrid(cellD, dimensions);
    auto clampv = [](double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); };
    for (Particle& q : host_) {
        q.pos.x = clampv(q.pos.x, 1.1 * r_ + cellD.x, dimensions.x - 1.1 * r_ - cellD.x);
        q.pos.y = clampv(q.pos.y, 1.1 * r_ + cellD.y, dimensions.y - 1.1 * r_ - cellD.y);
        q.pos.z = zConst_ ? z_ : clampv(q.pos.z, 1.1 * r_ + cellD.z, dimensions.z - 1.1 * r_ - cellD.z);
    }
    device_stale_ = true;
}
void HashedParticles::setParticleR(double r) {
    r_ = r;
    r_dirty_ = true;
}
void HashedParticles::addParticles(const std::vector<Particle>& ps) {
    ensureHost();
    host_.insert(host_.end(), ps.begin(), ps.end());
    device_stale_ = true;
}

}  // namespace particles

namespace macgrid {

MacGrid::MacGrid(const dvec3& targetDimensions, double resolution, bool twoD_)
    : cellD(1 / resolution, 1 / resolution, twoD_ ? targetDimensions.z / 3 : 1 / resolution),
      cellDInv(1.0 / cellD.x, 1.0 / cellD.y, 1.0 / cellD.z),
      gridSize{(int)(targetDimensions.x / cellD.x), (int)(targetDimensions.y / cellD.y),
               twoD_ ? 3 : (int)(targetDimensions.z / cellD.z)},
      dimensions(gridSize.x * cellD.x, gridSize.y * cellD.y, twoD_ ? targetDimensions.z : gridSize.z * cellD.z),
      twoD(twoD_) {}

CellType MacGrid::cellType(int x, int y, int z) {
    const size_t nc = (size_t)gridSize.x * gridSize.y * gridSize.z;
    if (stale_[0]) {
        label_.resize(nc);
        if (fetch_) fetch_(FSIM_FIELD_LABEL, label_.data(), nc);
        stale_[0] = false;
    }
    return (CellType)label_.at(((size_t)x * gridSize.y + y) * gridSize.z + z);
}
const std::vector<float>& MacGrid::faceVelocities(int axis) {
    const size_t nc = (size_t)gridSize.x * gridSize.y * gridSize.z;
    if (stale_[1 + axis]) {
        v2_[axis].resize(nc);
        if (fetch_) fetch_(FSIM_FIELD_U2 + axis, v2_[axis].data(), 4 * nc);
        stale_[1 + axis] = false;
    }
    return v2_[axis];
}
const std::vector<double>& MacGrid::pressure() {
    const size_t nc = (size_t)gridSize.x * gridSize.y * gridSize.z;
    if (stale_[4]) {
        p_.resize(nc);
        if (fetch_) fetch_(FSIM_FIELD_PRESSURE, p_.data(), 8 * nc);
        stale_[4] = false;
    }
    return p_;
}

}  // namespace macgrid

namespace simulator {

static const char* const kTimerKeys[FSIM_NUM_TIMERS] = {
    "SimulateParticles", "PushParticlesApart", "PushParticlesOutOfObstacles", "P2GTransfer",
    "IncompressibilityPrep", "Incompressibility", "VelocityExtrapolation", "G2PTransfer"};

Simulator::Simulator(SimulatorConfig cfg, std::shared_ptr<particles::HashedParticles> hp,
                     std::shared_ptr<macgrid::MacGrid> mg)
    : config(cfg), hashedParticles(std::move(hp)), macGrid(std::move(mg)) {
    for (const char* k : kTimerKeys) stepDuration[k] = 0;
    if (!hashedParticles->grid_known_) hashedParticles->setGrid(macGrid->cellD, macGrid->dimensions);
    createContext();
}

Simulator::~Simulator() {
    if (hashedParticles) hashedParticles->refresh_ = nullptr;
    if (macGrid) macGrid->fetch_ = nullptr;
    if (ctx_) fsim_destroy(ctx_);
}

void Simulator::createContext() {
    if (ctx_) {
        fsim_destroy(ctx_);
        ctx_ = nullptr;
    }
    fsim_grid_desc g{};
    g.nx = macGrid->gridSize.x;
    g.ny = macGrid->gridSize.y;
    g.nz = macGrid->gridSize.z;
    g.cell_d = macGrid->cellD.x;
    g.particle_r = hashedParticles->getParticleR();
    g.two_d = macGrid->twoD ? 1 : 0;
    g.cell_d_z = macGrid->twoD ? macGrid->cellD.z : 0.0;
    const fsim_params p = makeParams();
    check(fsim_create(&g, &p, &ctx_), ctx_, "fsim_create");
    hashedParticles->device_stale_ = true;
    hashedParticles->r_dirty_ = false;
    fsim_ctx* ctx = ctx_;
    macGrid->fetch_ = [ctx](int field, void* dst, size_t bytes) {
        check(fsim_get_grid(ctx, field, dst, bytes), ctx, "fsim_get_grid");
    };
    hashedParticles->refresh_ = [this]() { pullParticles(); };
}

fsim_params Simulator::makeParams() const {
    fsim_params p{};
    p.transfer = config.transferType == P2G2PType::PIC ? FSIM_PIC
                 : config.transferType == P2G2PType::FLIP ? FSIM_FLIP
                                                           : FSIM_APIC;
    p.flip_ratio = config.flipRatio;
    p.gravity = config.gravity;
    p.gravity_enabled = config.gravityEnabled;
    p.push_apart_enabled = config.pushApartEnabled;
    p.spawning_enabled = config.particleSpawningEnabled;
    p.despawning_enabled = config.particleDespawningEnabled;
    p.stop_particles = config.stopParticles;
    p.top_solid = macGrid->isTopOfContainerSolid;
    p.pressure_k = macGrid->pressureK;
    p.average_pressure = macGrid->averagePressure;
    p.max_iterations = macGrid->incompressibilityMaxIterationCount;
    p.pressure_enabled = macGrid->pressureEnabled;
    p.fluid_density = macGrid->fluidDensity;
    p.residual_tolerance = macGrid->residualTolerance;
    p.stop_rule = stopRule;
    p.precond = dynamic_cast<const macgrid::BasicMacGrid*>(macGrid.get()) ? FSIM_SOLVER_BASIC_SOR : precond;
    p.seed = seed;
    return p;
}

void Simulator::pushParams() {
    const fsim_params p = makeParams();
    check(fsim_set_params(ctx_, &p), ctx_, "fsim_set_params");
}

void Simulator::pushParticlesIfDirty() {
    particles::HashedParticles& hp = *hashedParticles;
    if (hp.r_dirty_) {
        check(fsim_set_particle_radius(ctx_, hp.r_), ctx_, "fsim_set_particle_radius");
        hp.r_dirty_ = false;
    }
    if (!hp.device_stale_) return;
    const size_t n = hp.host_.size();
    std::vector<float> pos(3 * n), vel(3 * n), c(9 * n);
    for (size_t i = 0; i < n; i++) {
        const particles::Particle& q = hp.host_[i];
        const double pv[3] = {q.pos.x, q.pos.y, q.pos.z}, vv[3] = {q.v.x, q.v.y, q.v.z};
        for (int a = 0; a < 3; a++) {
            pos[3 * i + a] = (float)pv[a];
            vel[3 * i + a] = (float)vv[a];
            const double cv[3] = {q.c[a].x, q.c[a].y, q.c[a].z};
            for (int b = 0; b < 3; b++) c[9 * i + 3 * a + b] = (float)cv[b];
        }
    }
    check(fsim_set_particles(ctx_, (int64_t)n, pos.data(), vel.data(), c.data()), ctx_, "fsim_set_particles");
    hp.device_stale_ = false;
}

void Simulator::pullParticles() {
    int64_t n = 0;
    check(fsim_num_particles(ctx_, &n), ctx_, "fsim_num_particles");
    std::vector<float> pos(3 * n), vel(3 * n), c(9 * n);
    int64_t got = 0;
    check(fsim_get_particles(ctx_, n, pos.data(), vel.data(), c.data(), &got), ctx_, "fsim_get_particles");
    particles::HashedParticles& hp = *hashedParticles;
    hp.host_.resize((size_t)got);
    for (int64_t i = 0; i < got; i++) {
        particles::Particle& q = hp.host_[i];
        q.pos = dvec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        q.v = dvec3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
        for (int a = 0; a < 3; a++) q.c[a] = dvec3(c[9 * i + 3 * a], c[9 * i + 3 * a + 1], c[9 * i + 3 * a + 2]);
    }
    hp.device_stale_ = false;
}

void Simulator::setNewMacGrid(std::shared_ptr<macgrid::MacGrid> mg) {
    macGrid->fetch_ = nullptr;
    macGrid = std::move(mg);
    createContext();
}

void Simulator::setNewHashedParticles(std::shared_ptr<particles::HashedParticles> hp) {
    hashedParticles->refresh_ = nullptr;
    hashedParticles = std::move(hp);
    if (!hashedParticles->grid_known_) hashedParticles->setGrid(macGrid->cellD, macGrid->dimensions);
    hashedParticles->device_stale_ = true;
    hashedParticles->r_dirty_ = true;
    hashedParticles->refresh_ = [this]() { pullParticles(); };
}

void Simulator::simulate(double dt) {
    pushParams();
    pushParticlesIfDirty();
    std::vector<fsim_obstacle> obs;
    obs.reserve(obstacles.size());
    for (const auto& o : obstacles) obs.push_back(o->toC());
    check(fsim_set_obstacles(ctx_, (int)obs.size(), obs.data()), ctx_, "fsim_set_obstacles");
    check(fsim_step(ctx_, dt, &stats_), ctx_, "fsim_step");
    // the sources' fractional spawn carry is state of the obstacle objects
    std::vector<fsim_obstacle> back(obs.size());
    int nb = 0;
    check(fsim_get_obstacles(ctx_, (int)back.size(), back.data(), &nb), ctx_, "fsim_get_obstacles");
    for (int i = 0; i < nb && i < (int)obstacles.size(); i++)
        if (auto* src = dynamic_cast<obstacle::SphericalParticleSource*>(obstacles[i].get()))
            src->lastSpawnFraction = back[i].last_spawn_fraction;
    // stepDuration: sliding average of whole microseconds, stored as long long
    // after every update (simulator.cpp:52-101)
    constexpr double kSlidingAvg = 0.9;
    for (int k = 0; k < FSIM_NUM_TIMERS; k++) {
        long long& d = stepDuration[kTimerKeys[k]];
        const long long us = (long long)(stats_.timer_ms[k] * 1000.0);
        d = (long long)(d * kSlidingAvg + us * (1.0 - kSlidingAvg));
    }
    // the solver's iteration count, set raw after every step (simulator.cpp:91)
    stepDuration["Incompressibility it count"] = stats_.pcg_iterations;
    hashedParticles->host_stale_ = true;
    for (bool& s : macGrid->stale_) s = true;
}

}  // namespace simulator

}  // namespace fsim
