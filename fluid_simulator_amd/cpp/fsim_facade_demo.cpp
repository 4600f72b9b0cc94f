// fsim_facade_demo.cpp — drives the C++ facade the way the reference's
// SimulationManager drives Simulator: build a grid and a dam-break particle
// set, step, read particles and timings back.  Prints one JSON line.
//
// usage: fsim_facade_demo <n> <steps> <transfer 0|1|2> [out.bin]
//   out.bin receives the final particle positions (float32, 3 per particle)
//        fsim_facade_demo particles <num> <grown> <shrunk>
//   builds HashedParticles(num, ...) on a 32^3 unit grid, grows it to
//   <grown> and shrinks it to <shrunk> with setParticleNum (host only, no
//   device needed) and prints the ranges it saw
//        fsim_facade_demo 2d <n> <steps> <transfer> <num> <init.bin> <out.bin>
//   the 2D mode: MacGrid(1 x 1 x 0.1, n, twoD), HashedParticles(num, ...,
//   zConst at the depth / 2) with push-apart on; writes the initial and the
//   final positions (float32, 3 per particle)
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <string>
#include <vector>

#include "fsim_simulator.hpp"

using namespace fsim;

static int particles_demo(int num, int grown, int shrunk) {
    const double dx = 1.0 / 32, r = dx / 4;
    particles::HashedParticles hp(num, r, dvec3(1, 1, 1), dvec3(dx, dx, dx), false, 0.0, 7);
    double lo0 = 1e9, hi0 = -1e9, lo1 = 1e9, hi1 = -1e9, vmax = 0;
    const size_t n0 = hp.getParticleNum();
    std::vector<double> first;
    for (size_t i = 0; i < n0; i++) {
        const auto& q = hp.getParticleAt(i);
        for (double v : {q.pos.x, q.pos.y, q.pos.z}) lo0 = v < lo0 ? v : lo0, hi0 = v > hi0 ? v : hi0;
        first.push_back(q.pos.x);
    }
    hp.setParticleNum((size_t)grown);
    const size_t n1 = hp.getParticleNum();
    double mean = 0;
    for (size_t i = n0; i < n1; i++) {
        const auto& q = hp.getParticleAt(i);
        for (double v : {q.pos.x, q.pos.y, q.pos.z}) lo1 = v < lo1 ? v : lo1, hi1 = v > hi1 ? v : hi1, mean += v;
        for (double v : {q.v.x, q.v.y, q.v.z}) vmax = v * v > vmax ? v * v : vmax;
    }
    mean /= 3.0 * (double)(n1 - n0 > 0 ? n1 - n0 : 1);
    hp.setParticleNum((size_t)shrunk);
    const size_t n2 = hp.getParticleNum();
    bool kept = true;
    for (size_t i = 0; i < n2 && i < first.size(); i++) kept = kept && hp.getParticleAt(i).pos.x == first[i];
    std::printf("{\"n0\": %zu, \"lo0\": %.9f, \"hi0\": %.9f, \"n1\": %zu, \"lo1\": %.9f, \"hi1\": %.9f, "
                "\"mean1\": %.6f, \"vmax1\": %.3g, \"n2\": %zu, \"kept\": %s}\n",
                n0, lo0, hi0, n1, lo1, hi1, mean, vmax, n2, kept ? "true" : "false");
    return 0;
}

static void write_positions(particles::HashedParticles& hp, const char* path) {
    std::FILE* f = std::fopen(path, "wb");
    if (!f) throw std::runtime_error("cannot open output file");
    hp.forEachConst([&](const particles::Particle& p) {
        const float v[3] = {(float)p.pos.x, (float)p.pos.y, (float)p.pos.z};
        std::fwrite(v, sizeof(float), 3, f);
    });
    std::fclose(f);
}

// The reference's SimulationManager in 2D (simulationManager.cpp:10-27):
// the grid gets twoD, the particles zConst with z = the target depth / 2.
static int demo_2d(int n, int steps, int transfer, int num, const char* init, const char* out) {
    try {
        const dvec3 target(1, 1, 0.1);
        auto grid = std::make_shared<macgrid::MacGrid>(target, (double)n, true);
        grid->pressureEnabled = false;
        grid->incompressibilityMaxIterationCount = 10000;
        const double r = grid->cellD.x / 4;
        auto parts = std::make_shared<particles::HashedParticles>(num, r, grid->dimensions, grid->cellD, true,
                                                                  target.z / 2, 11);
        write_positions(*parts, init);
        simulator::Simulator::SimulatorConfig cfg;
        cfg.transferType = (simulator::Simulator::P2G2PType)transfer;
        cfg.flipRatio = 0.95f;
        cfg.gravity = -9.81f;
        cfg.pushApartEnabled = true;
        simulator::Simulator sim(cfg, parts, grid);
        sim.stopRule = FSIM_STOP_REL_L2;
        for (int s = 0; s < steps; s++) sim.simulate(1.0 / 240.0);
        write_positions(*parts, out);
        int water = 0;
        for (int x = 0; x < grid->gridSize.x; x++)
            for (int y = 0; y < grid->gridSize.y; y++)
                for (int z = 0; z < grid->gridSize.z; z++) water += grid->cellType(x, y, z) == macgrid::CellType::WATER;
        std::printf("{\"particles\": %zu, \"grid\": [%d, %d, %d], \"cell_d\": [%.17g, %.17g, %.17g], "
                    "\"dims\": [%.17g, %.17g, %.17g], \"water_cells\": %d, \"pcg_iterations\": %d}\n",
                    parts->getParticleNum(), grid->gridSize.x, grid->gridSize.y, grid->gridSize.z, grid->cellD.x,
                    grid->cellD.y, grid->cellD.z, grid->dimensions.x, grid->dimensions.y, grid->dimensions.z, water,
                    sim.lastStats().pcg_iterations);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc > 7 && std::string(argv[1]) == "2d")
        return demo_2d(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]), argv[6],
                       argv[7]);
    if (argc > 4 && std::string(argv[1]) == "particles")
        return particles_demo(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]));
    const int n = argc > 1 ? std::atoi(argv[1]) : 32;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 3;
    const int transfer = argc > 3 ? std::atoi(argv[3]) : 1;
    const char* out = argc > 4 ? argv[4] : nullptr;
    try {
        auto grid = std::make_shared<macgrid::MacGrid>(dvec3(1, 1, 1), n);
        grid->pressureEnabled = false;
        grid->incompressibilityMaxIterationCount = 10000;
        const double r = grid->cellD.x / 4;
        // dam break: block x < 0.4, y < 0.6 (the C ABI's scene generator, seed 0)
        fsim_grid_desc g{n, n, n, grid->cellD.x, r, 0};
        const double block[6] = {0, 0, 0, 0.4, 0.6, 1.0};
        int64_t np = 0;
        check(fsim_scene_dam_break(&g, 1, block, 0, nullptr, 0, 0, nullptr, &np), nullptr, "scene");
        std::vector<float> pos(3 * np);
        check(fsim_scene_dam_break(&g, 1, block, 0, nullptr, 0, np, pos.data(), &np), nullptr, "scene");
        std::vector<particles::Particle> ps((size_t)np);
        for (int64_t i = 0; i < np; i++) ps[i].pos = dvec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        auto parts = std::make_shared<particles::HashedParticles>(std::move(ps), r);

        simulator::Simulator::SimulatorConfig cfg;   // the 3D preset: no push-apart
        cfg.transferType = (simulator::Simulator::P2G2PType)transfer;
        cfg.flipRatio = 0.95f;
        cfg.gravity = -9.81f;
        cfg.pushApartEnabled = false;
        cfg.pushParticlesApartEnabled = false;
        simulator::Simulator sim(cfg, parts, grid);
        sim.stopRule = FSIM_STOP_REL_L2;
        for (int s = 0; s < steps; s++) sim.simulate(1.0 / 240.0);

        double ymin = 1e9, ymax = -1e9;
        parts->forEachConst([&](const particles::Particle& p) {
            ymin = p.pos.y < ymin ? p.pos.y : ymin;
            ymax = p.pos.y > ymax ? p.pos.y : ymax;
        });
        int water = 0;
        for (int x = 0; x < n; x++)
            for (int y = 0; y < n; y++)
                for (int z = 0; z < n; z++) water += grid->cellType(x, y, z) == macgrid::CellType::WATER;
        if (out) write_positions(*parts, out);
        const auto d = sim.getStepDuration();
        std::printf("{\"particles\": %zu, \"water_cells\": %d, \"pcg_iterations\": %d, \"ymin\": %.6f, \"ymax\": %.6f, "
                    "\"Incompressibility_us\": %lld, \"P2GTransfer_us\": %lld, \"it_count\": %lld, "
                    "\"step_duration_keys\": %zu}\n",
                    parts->getParticleNum(), water, sim.lastStats().pcg_iterations, ymin, ymax,
                    d.at("Incompressibility"), d.at("P2GTransfer"), d.at("Incompressibility it count"), d.size());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    return 0;
}
