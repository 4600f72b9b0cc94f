// fsim_internal.hpp — device state and launch interfaces of libfsim_hip.so.
//
// Device layout (one context, one GPU):
//   * grid fields are dense arrays over all Nc = Nx*Ny*Nz cells in the
//     reference's flat order idx = (x*Ny + y)*Nz + z (z fastest);
//   * face velocities are FP32 SoA: u,v,w (P2G result after boundary
//     conditions, the reference's Face::v) and u2,v2,w2 (Face::v2);
//   * labels are uint8 fsim_cell_type; nbinfo (uint16) caches, per WATER cell,
//     which of the 6 neighbours are WATER and how many are not SOLID;
//   * pressure-solve vectors are dense FP32 (b, r, z, s, q, y) with FP64
//     pressure x and FP64 MIC(0) diagonal; entries outside WATER cells are 0;
//     they, nbinfo and the packed W use the rotated layout of ridx() below;
//   * particles are FP32 SoA (pos, vel, APIC C) plus the int32 particle id
//     (its index in the reference's particle vector), kept sorted by cell
//     after every BIN stage.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fsim.h"

namespace fsim {

constexpr int kMaxObstacles = 64;
constexpr int kBlock = 256;

// nbinfo bit layout
constexpr uint16_t NB_FLUID = 1u << 0;      // the cell itself is an interior WATER cell
constexpr uint16_t NB_MX = 1u << 1;         // -x neighbour is WATER (fluid)
constexpr uint16_t NB_MY = 1u << 2;
constexpr uint16_t NB_MZ = 1u << 3;
constexpr uint16_t NB_PX = 1u << 4;         // +x neighbour is WATER (fluid)
constexpr uint16_t NB_PY = 1u << 5;
constexpr uint16_t NB_PZ = 1u << 6;
constexpr int NB_NONSOLID_SHIFT = 7;        // bits 7..9: number of non-SOLID neighbours
// bits 10..15: for each axis a, whether the -a neighbour's +((a+1)%3) and
// +((a+2)%3) neighbours are WATER (the MIC(0) factor's "other" terms)
constexpr int NB_UP_SHIFT = 10;

// Persistent wavefront kernels (MIC(0) factor and triangular sweeps): a tile
// of WF_TY x WF_TZ (y, z) columns per workgroup, WF_NW compute waves of WF_R
// rows each plus two communication waves.
#ifndef FSIM_WF_R
#define FSIM_WF_R 2
#endif
#ifndef FSIM_WF_NW
#define FSIM_WF_NW 4
#endif
constexpr int WF_R = FSIM_WF_R;
constexpr int WF_NW = FSIM_WF_NW;
constexpr int WF_TY = WF_R * WF_NW;
constexpr int WF_TZ = 64;
constexpr int WF_THREADS = (WF_NW + 4) * WF_TZ;   // compute waves, 2 receiving and 2 sending edge waves
#ifndef FSIM_WF_PREFETCH
#define FSIM_WF_PREFETCH 8
#endif
constexpr int WF_PREFETCH = FSIM_WF_PREFETCH;   // steps of per-cell inputs loaded ahead (and the loop unroll)
// Steps of one tile: all rows of compute wave w are at x-index t - w - tz at
// step t, so the last cell (wave WF_NW-1, lane 63, x-index nxi-1) is at step
// nxi + WF_NW + WF_TZ - 3; rounded up to whole blocks of WF_PREFETCH steps
// (the extra steps are inactive everywhere) so that no path through the step
// loop skips a prefetch.
__host__ __device__ inline int wave_steps(int nxi) {
    return (nxi + WF_NW + WF_TZ - 2 + WF_PREFETCH - 1) / WF_PREFETCH * WF_PREFETCH;
}
// the wavefront kernels address cells with 32-bit byte offsets into arrays
// of up to 8-byte words; larger grids use the plane-by-plane kernels
constexpr int64_t kWaveMaxCells = ((int64_t(1) << 31) - 1) / 8;

// Solver-vector layout ("rotated"): cell (x, y, z) of the pressure solve is
// stored at plane (x + ((z - 1) mod 64)) mod nx of a dense nx*ny*nz array,
// row y, column z.  The map is a bijection on every (y, z) column, and the
// 64 cells of a z-tile [1 + 64k, 64 + 64k] that one wavefront step visits
// (x = c - ((z - 1) mod 64)) share one plane, so a step reads and writes 64
// contiguous words.  Stencil neighbours stay contiguous chunks too.
__host__ __device__ inline int rot_of(int z) { return (z - 1) & (WF_TZ - 1); }
__host__ __device__ inline int rot_plane(int x, int z, int nx) {
    const int p = x + rot_of(z);
    return p < nx ? p : p % nx;
}
__host__ __device__ inline int unrot_x(int p, int z, int nx) {
    const int x = p - rot_of(z);
    return x >= 0 ? x : (x % nx + nx) % nx;
}
__host__ __device__ inline int64_t ridx(int x, int y, int z, int nx, int ny, int nz) {
    return ((int64_t)rot_plane(x, z, nx) * ny + y) * nz + z;
}

// Dimensions and scale of a pressure solve; stencil neighbours in the rotated layout.
struct SolveArgs {
    int nx, ny, nz;
    int off63;          // 63 mod nx: the plane shift of a z step across a 64-wide z-tile edge
    int64_t nc;
    float scale_f;      // dt / (rho dx^2)
    double scale;
};

// flat index of the grid fields (labels, velocities, density)
__device__ inline int64_t sidx(const SolveArgs& S, int x, int y, int z) { return ((int64_t)x * S.ny + y) * S.nz + z; }
// index of the solver vectors, nbinfo and W (rotated layout)
__device__ inline int64_t rsidx(const SolveArgs& S, int x, int y, int z) { return ridx(x, y, z, S.nx, S.ny, S.nz); }
// Rotated indices of the six neighbours of the interior cell stored at
// plane p, row y, column z (rotated index i): +-y are +-nz; +-x are the
// next/previous plane; +-z move one column and one plane, except across a
// 64-wide z-tile edge where rot(z) wraps 63 <-> 0 and the plane moves by 63.
struct RNeighbours {
    int64_t px, mx, py, my, pz, mz;
};
__device__ inline RNeighbours rneighbours(const SolveArgs& S, int64_t i, int p, int y, int z) {
    const int64_t plane = (int64_t)S.ny * S.nz;
    const int64_t yz = (int64_t)y * S.nz + z;
    const int rz = rot_of(z);
    const int pn = p + 1 == S.nx ? 0 : p + 1;
    const int pp = p == 0 ? S.nx - 1 : p - 1;
    RNeighbours n;
    n.px = pn * plane + yz;
    n.mx = pp * plane + yz;
    n.py = i + S.nz;
    n.my = i - S.nz;
    const int o = S.off63;
    n.pz = (rz == WF_TZ - 1 ? (p >= o ? p - o : p - o + S.nx) : pn) * plane + yz + 1;
    n.mz = (rz == 0 ? (p + o >= S.nx ? p + o - S.nx : p + o) : pp) * plane + yz - 1;
    return n;
}
// rotated indices of the +a and -a neighbours of interior cell (x, y, z)
__device__ inline int64_t rnb_plus(const SolveArgs& S, int x, int y, int z, int a) {
    return rsidx(S, x + (a == 0), y + (a == 1), z + (a == 2));
}
__device__ inline int64_t rnb_minus(const SolveArgs& S, int x, int y, int z, int a) {
    return rsidx(S, x - (a == 0), y - (a == 1), z - (a == 2));
}

struct WaveArgs {
    int nx, ny, nz;
    int TJ, TK;                   // tiles along y and z
    int SP;                       // boundary-packet rows (steps) per tile
    const uint16_t* nbinfo;
    const float* in;
    float* out;
    float ms;                     // -scale in FP32
    double scale;                 // dt / (rho dx^2)
    double* P64;
    float* P32;
    uint64_t* ybnd;               // packets across y tile edges: [tile J*TK+K][step][64][words]
    uint64_t* zbnd;               // packets across z tile edges: [tile J*TK+K][step][WF_TY][words]
    unsigned long long* ticket;
    unsigned long long ticket_base;
    uint32_t epoch;
    int* err;
    const uint8_t* colw;          // [y * nz + z]: 1 if the (y, z) column holds a WATER cell
    int send_delay;               // test option: s_sleep(127) periods of the sending waves per step
};

struct GridDims {
    int nx, ny, nz;
    int64_t nc;
    double cd, cdinv;     // cellD.x == cellD.y
    double cdz, cdzinv;   // cellD.z (cd except in 2D)
    double dims[3];
    double pr;
    int two_d;
};

// Device-resident scalars of the pressure solve (one cache line each group).
struct SolveScalars {
    double sigma, alpha, beta, sq;   // PCG scalars
    double rr, rmax, bb;             // recursive residual ||r||^2, max|r|, ||b||^2
    double tr_ss, tr_max;            // true residual ||b-Ax||^2, max|b-Ax|
    double zr;                       // z.r
    double zq;                       // z.q (flexible PCG)
    double nfluid;                   // number of WATER cells of the solve
    int nan_flag;
    int pad;
};

// Kernel groups timed with HIP events when profiling is on.
enum KernelId {
    KID_SPAWN = 0, KID_MOVE, KID_BIN_KEYS, KID_BIN_SCAN, KID_BIN_SCATTER, KID_BIN_SEGSORT, KID_BIN_GATHER,
    KID_P2G, KID_MARK, KID_OBSTACLES, KID_BORDERS, KID_POST_P2G, KID_RHS, KID_MIC_FACTOR, KID_SWEEP_FWD,
    KID_SWEEP_BWD, KID_SPMV, KID_UPDATE, KID_DOT, KID_TRUE_RES, KID_PRESSURE, KID_EXTRAPOLATE, KID_G2P,
    KID_LEARNED, KID_SOR, KID_PUSH_APART, KID_MG, KID_COUNT
};

struct Particles {
    int64_t n = 0, cap = 0;
    float* base = nullptr;   // one allocation, 16 float planes of cap entries
    float* pos[3] = {};
    float* vel[3] = {};
    float* c[9] = {};
    int32_t* pid = nullptr;
};

}  // namespace fsim

namespace fsim {
struct LearnedNet;
struct MgState;
constexpr int kMgMaxLevels = 8;
constexpr int kMgCoarseMaxSide = 16;   // the network's level: the first whose largest side is <= this
}

struct fsim_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t xstream = nullptr;    // PCG: x += alpha s beside the r update and the preconditioner
    fsim::GridDims g{};
    fsim_params prm{};
    std::string err;

    // particles (double-buffered for the binning gather)
    fsim::Particles P, Palt;
    uint32_t* key = nullptr;     // cell key per particle (Nc = removed)
    int32_t* rank = nullptr;     // slot within cell from the histogram atomics
    int32_t* perm = nullptr;     // sorted slot -> source slot
    int64_t pwork_cap = 0;
    uint8_t* removed = nullptr;  // sink removal flags (per particle)
    bool c_valid = true;         // APIC matrices carried (false after a PIC/FLIP bin)
    int64_t scan_tmp_cap = 0;

    // grid
    float *u = nullptr, *v = nullptr, *w = nullptr;
    float *u2 = nullptr, *v2 = nullptr, *w2 = nullptr;
    uint8_t* label = nullptr;
    uint8_t* valid = nullptr;
    float* density = nullptr;
    int32_t* cell_count = nullptr;   // Nc + 1 (last = removed particles)
    int32_t* cell_start = nullptr;   // Nc + 2
    int32_t* scan_tmp = nullptr;
    uint16_t* nbinfo = nullptr;
    uint8_t* colw = nullptr;          // WATER columns (y, z) of the current labels, for the wavefront kernels
    uint32_t* chunk_flags = nullptr;  // solver-vector chunks (64 or 256 cells) holding a WATER cell, per solve

    // pressure solve
    float *b = nullptr, *r = nullptr, *z = nullptr, *s = nullptr, *q = nullptr, *y = nullptr;
    float* s2 = nullptr;              // second search-direction buffer (swapped with s)
    float* P32 = nullptr;
    double* P64 = nullptr;
    double* x = nullptr;
    double* sor64 = nullptr;          // BasicMacGrid SOR: FP64 faces u, v, w (allocated on first use)
    double* partials = nullptr;       // block partial sums (FP64)
    unsigned* red_count = nullptr;    // blocks finished in the running reduction (0 between launches)
    hipEvent_t res_ev = nullptr;      // PCG: the iteration's scalars are on the host
    hipEvent_t dir_ev = nullptr;      // PCG: the direction step (s, alpha) is done
    hipEvent_t x_ev = nullptr;        // PCG: the x update on xstream is done
    int last_mic_iterations = 0;      // MIC(0)-PCG iterations of the last solve that used it
    int learned_since_mic = 0;        // learned solves since that MIC(0) solve
    int learned_refresh = 50;         // learned solves between MIC(0) solves that renew the count (0: never)
    fsim::LearnedNet* learned = nullptr;   // learned preconditioner (network parameters and buffers)
    fsim::MgState* mg = nullptr;          // multigrid hierarchy of the learned preconditioner
    int mg_coarse = 0;                    // coarsest-level solve: FSIM_MG_COARSE_*
    int reliable_drop = 3;                // PCG: true residual after a 10^drop drop of the recursive one (0: at the tolerance)
    int mg_graph = 1;                     // replay the multigrid V-cycle as a HIP graph
    int p2g_rows = 0;                     // 1: the one-row-per-block P2G gather instead of the tiled one
    int wave_send_delay = 0;              // test option: sending-wave pause per step (s_sleep 127 units)
    int learned_gen = 0;                  // bumped when the network's per-size buffers are reallocated
    bool mg_harvesting = false;           // a solve run by fsim_learned_harvest
    int partials_cap = 0;
    fsim::SolveScalars* scal = nullptr;   // device
    fsim::SolveScalars* scal_host = nullptr;  // pinned host mirror
    double last_dt = 0.0;
    // wavefront sweeps
    fsim::WaveArgs wave{};
    uint64_t* ybnd = nullptr;
    uint64_t* zbnd = nullptr;
    unsigned long long* wave_ticket = nullptr;   // 3 counters: factor, fwd, bwd
    unsigned long long wave_ticket_base[3] = {0, 0, 0};
    uint32_t wave_epoch = 0;
    int* wave_err = nullptr;
    int sweep_mode = 0;                          // 0 wavefront, 1 one launch per plane
    int learned_smoothing = 2;                   // red-black sweeps after the network on the coarsest level
    int learned_loss = 0;                        // training loss: 0 masked MSE, 1 energy norm
    int learned_conv_lds = 1;                    // 1: LDS-staged convolutions where they fit

    // renderer snapshot staging (grown on demand)
    fsim_gfx_particle* gfx_staging = nullptr;
    int64_t gfx_cap = 0;

    // obstacles
    std::vector<fsim_obstacle> obs;
    fsim_obstacle* obs_dev = nullptr;
    uint64_t spawn_calls = 0;

    // timing
    hipEvent_t ev[2 * FSIM_NUM_TIMERS + 4] = {};
    fsim_step_stats stats{};
    // per-kernel profiling
    bool prof_on = false;
    std::vector<hipEvent_t> prof_pool;
    size_t prof_next = 0;
    std::vector<int> prof_pending;   // triples: kernel id, start event, end event
    double prof_ms[fsim::KID_COUNT] = {};
    int64_t prof_n[fsim::KID_COUNT] = {};
};

namespace fsim {

// ---- error plumbing ------------------------------------------------------
int set_err(fsim_ctx* ctx, int code, const std::string& msg);
#define FSIM_HIP(ctx, expr)                                                              \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return ::fsim::set_err((ctx), FSIM_ERR_HIP,                                  \
                                   std::string(#expr) + ": " + hipGetErrorString(e_));   \
    } while (0)
#define FSIM_TRY(expr)              \
    do {                            \
        int rc_ = (expr);           \
        if (rc_ != FSIM_OK) return rc_; \
    } while (0)

inline unsigned nblocks(int64_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

// ---- profiling (fsim_capi.hip) ----------------------------------------------
int set_profiling(fsim_ctx* ctx, bool on);
int prof_event(fsim_ctx* ctx);          // records a pooled event, returns its index (-1 if off)
void prof_close(fsim_ctx* ctx, int kid, int start);
int prof_collect(fsim_ctx* ctx);        // after a stream synchronisation

// Times the enclosed launches as kernel group `kid` when profiling is on.
struct ProfRange {
    fsim_ctx* c;
    int kid, start;
    ProfRange(fsim_ctx* ctx, int k) : c(ctx), kid(k), start(prof_event(ctx)) {}
    ~ProfRange() {
        if (start >= 0) prof_close(c, kid, start);
    }
};
#define FSIM_CAT2(a, b) a##b
#define FSIM_CAT(a, b) FSIM_CAT2(a, b)
#define FSIM_PROF(ctx, kid) ::fsim::ProfRange FSIM_CAT(prof_range_, __LINE__)((ctx), (kid))

// ---- particle stages (fsim_particles.hip) ---------------------------------
int particles_reserve(fsim_ctx* ctx, int64_t n);
int stage_spawn(fsim_ctx* ctx, double dt);
// flags: 1 = advect, 2 = push out, 4 = stop particles
int stage_move(fsim_ctx* ctx, double dt, int flags);
int stage_bin(fsim_ctx* ctx);
int stage_push_apart(fsim_ctx* ctx);
int upload_obstacles(fsim_ctx* ctx);

// ---- grid stages (fsim_grid.hip) ------------------------------------------
int stage_p2g(fsim_ctx* ctx);
int stage_mark(fsim_ctx* ctx);
int stage_obstacles(fsim_ctx* ctx);
int stage_borders(fsim_ctx* ctx);
int stage_post_p2g(fsim_ctx* ctx, double grav_inc);
int stage_extrapolate(fsim_ctx* ctx);
int stage_g2p(fsim_ctx* ctx);
int fill_u8(fsim_ctx* ctx, uint8_t* p, uint8_t v, int64_t n);
int fill_f32(fsim_ctx* ctx, float* p, float v, int64_t n);
int fill_f64(fsim_ctx* ctx, double* p, double v, int64_t n);
int fill_i32(fsim_ctx* ctx, int32_t* p, int32_t v, int64_t n);

// ---- pressure solve (fsim_solver.hip) -------------------------------------
int stage_solve(fsim_ctx* ctx, double dt);
// kind 0 = MIC(0) factor, 1 = forward sweep, 2 = backward sweep
int wave_launch(fsim_ctx* ctx, int kind, const float* in, float* out, double scale);
int wave_init(fsim_ctx* ctx);
int solver_apply(fsim_ctx* ctx, int op, double dt, const double* in, double* out);

// ---- renderer snapshot (fsim_gfx.hip) --------------------------------------
int gfx_snapshot(fsim_ctx* ctx, fsim_gfx_particle* host_out);
int gfx_snapshot_noalloc(fsim_ctx* ctx, fsim_gfx_particle* host_out);
int scratch_init(fsim_ctx* ctx);
int64_t scratch_particles();
void gfx_free(fsim_ctx* ctx);

// ---- multigrid of the learned preconditioner (fsim_mg.hip) ------------------
int mg_levels_for(int nx, int ny, int nz);
int mg_setup(fsim_ctx* ctx);                                        // coarse operators of the current labels
int mg_apply(fsim_ctx* ctx, const float* r, float* z, double scale);  // z = V-cycle(r), rotated layout
int mg_info(fsim_ctx* ctx, int* levels, int* dims);
int mg_harvest_begin(fsim_ctx* ctx, int cap);
int mg_harvest_end(fsim_ctx* ctx, int cap, float* b, float* op, float* x, int* n_out);
void mg_free(fsim_ctx* ctx);

// ---- learned preconditioner (fsim_learned.hip) ------------------------------
// the network on the coarsest multigrid level (flat, operator arrays): x = out * max|b| * norm
int learned_apply_level(fsim_ctx* ctx, int nx, int ny, int nz, const float* b, const float* diag, const float* ox,
                        const float* oy, const float* oz, double norm, float* x);
// the network on the solve's own grid (rotated layout): z = out * max|r| * binv
int learned_prepare(fsim_ctx* ctx, int nx, int ny, int nz);
int learned_apply_level0(fsim_ctx* ctx, const float* r, float* z, float binv);
int64_t learned_num_params();
int learned_init(fsim_ctx* ctx, uint64_t seed);
int learned_set_params(fsim_ctx* ctx, const float* params);
int learned_get_params(fsim_ctx* ctx, float* params);
bool learned_ready(const fsim_ctx* ctx);
void label_operator(int nx, int ny, int nz, const uint8_t* label, float* op);   // op[4][nc] of a label field
int learned_forward_host(fsim_ctx* ctx, const float* b, const float* op, float* out);
int learned_grad_host(fsim_ctx* ctx, const float* b, const float* op, const double* p, float* grad, double* loss);
int learned_train_batch(fsim_ctx* ctx, int n, const float* b, const float* op, const double* p, float lr,
                        double* mean_loss);
void learned_free(fsim_ctx* ctx);

}  // namespace fsim
