// fsim_capi.hip — the extern "C" boundary of libfsim_hip.so (include/fsim.h).
//
// fsim_step is Simulator::simulate(double dt) (simulator.cpp:51-100) as a
// sequence of HIP kernels on the context's stream:
//   spawn -> advect + sink removal + push-out + stop (fused per particle)
//   -> bin by cell -> P2G (+ fluid marking and particle density) -> obstacles
//   -> borders -> v2 + gravity -> pressure solve -> extrapolation -> G2P.
// Stage GPU times are HIP-event intervals grouped by the reference's
// stepDuration keys (simulator.cpp:31-41).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "fsim_internal.hpp"

namespace fsim {

int set_err(fsim_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int stage_p2g_fused(fsim_ctx* ctx);

static const char* const kKernelNames[KID_COUNT] = {
    "spawn", "move", "bin_keys", "bin_scan", "bin_scatter", "bin_segsort", "bin_gather", "p2g", "mark",
    "obstacles", "borders", "post_p2g", "rhs", "mic_factor", "sweep_fwd", "sweep_bwd", "spmv", "update", "dot",
    "true_residual", "pressure", "extrapolate", "g2p", "learned", "basic_sor", "push_apart", "mg"};

int set_profiling(fsim_ctx* ctx, bool on) {
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    FSIM_TRY(prof_collect(ctx));
    ctx->prof_on = on;
    return FSIM_OK;
}

int prof_event(fsim_ctx* ctx) {
    if (!ctx->prof_on) return -1;
    if (ctx->prof_next == ctx->prof_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        ctx->prof_pool.push_back(e);
    }
    const int idx = (int)ctx->prof_next++;
    if (hipEventRecord(ctx->prof_pool[idx], ctx->stream) != hipSuccess) return -1;
    return idx;
}

void prof_close(fsim_ctx* ctx, int kid, int start) {
    const int end = prof_event(ctx);
    if (end < 0) return;
    ctx->prof_pending.push_back(kid);
    ctx->prof_pending.push_back(start);
    ctx->prof_pending.push_back(end);
}

int prof_collect(fsim_ctx* ctx) {
    for (size_t k = 0; k + 2 < ctx->prof_pending.size(); k += 3) {
        float ms = 0.f;
        FSIM_HIP(ctx, hipEventElapsedTime(&ms, ctx->prof_pool[ctx->prof_pending[k + 1]],
                                          ctx->prof_pool[ctx->prof_pending[k + 2]]));
        ctx->prof_ms[ctx->prof_pending[k]] += ms;
        ctx->prof_n[ctx->prof_pending[k]] += 1;
    }
    ctx->prof_pending.clear();
    ctx->prof_next = 0;
    return FSIM_OK;
}

}  // namespace fsim

using namespace fsim;

static int check_grid(const fsim_grid_desc* g) {
    if (!g) return FSIM_ERR_INVALID;
    const bool two_d = g->two_d != 0;
    if (g->nx < 4 || g->ny < 4 || (two_d ? g->nz != 3 : g->nz < 4) || !(g->cell_d > 0.0) || !(g->particle_r >= 0.0))
        return FSIM_ERR_INVALID;
    // 2D: the z cell size is the target depth / 3 (macGrid.cpp:10); 3D: cellD.z == cellD.x
    if (two_d ? !(g->cell_d_z > 0.0) : (g->cell_d_z != 0.0 && g->cell_d_z != g->cell_d)) return FSIM_ERR_INVALID;
    if ((int64_t)g->nx * g->ny * g->nz >= (int64_t)INT32_MAX - 8) return FSIM_ERR_INVALID;
    return FSIM_OK;
}

// two_d: the learned multigrid preconditioner (no reference counterpart) is
// built for 3D grids only; the 2D mode takes MIC(0)-PCG or red-black SOR.
static int check_params(const fsim_params* p, bool two_d) {
    if (!p) return FSIM_ERR_INVALID;
    if (p->transfer < FSIM_PIC || p->transfer > FSIM_APIC) return FSIM_ERR_INVALID;
    if (p->max_iterations < 0 || !(p->fluid_density > 0.0)) return FSIM_ERR_INVALID;
    if (p->stop_rule != FSIM_STOP_ABS_INF && p->stop_rule != FSIM_STOP_REL_L2) return FSIM_ERR_INVALID;
    if (p->precond != FSIM_PRECOND_MIC0 && p->precond != FSIM_PRECOND_LEARNED && p->precond != FSIM_SOLVER_BASIC_SOR)
        return FSIM_ERR_INVALID;
    if (two_d && p->precond == FSIM_PRECOND_LEARNED) return FSIM_ERR_INVALID;
    return FSIM_OK;
}

template <typename T>
static int dalloc(fsim_ctx* ctx, T** p, int64_t n) {
    if (hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (size_t)std::max<int64_t>(n, 1)) != hipSuccess)
        return set_err(ctx, FSIM_ERR_OOM, "device allocation failed");
    return FSIM_OK;
}

static void free_all(fsim_ctx* ctx) {
    learned_free(ctx);
    mg_free(ctx);
    gfx_free(ctx);
    void* ptrs[] = {ctx->P.base, ctx->Palt.base, ctx->key, ctx->rank, ctx->perm, ctx->removed, ctx->u, ctx->v, ctx->w,
                    ctx->u2, ctx->v2, ctx->w2, ctx->label, ctx->valid, ctx->density, ctx->cell_count, ctx->cell_start,
                    ctx->scan_tmp, ctx->nbinfo, ctx->chunk_flags, ctx->b, ctx->r, ctx->z, ctx->s, ctx->q, ctx->y, ctx->P32, ctx->P64,
                    ctx->x, ctx->partials, ctx->scal, ctx->obs_dev, ctx->ybnd, ctx->zbnd, ctx->s2, ctx->red_count,
                    ctx->wave_ticket, ctx->wave_err, ctx->sor64, ctx->colw};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (ctx->scal_host) (void)hipHostFree(ctx->scal_host);
    if (ctx->res_ev) (void)hipEventDestroy(ctx->res_ev);
    if (ctx->dir_ev) (void)hipEventDestroy(ctx->dir_ev);
    if (ctx->x_ev) (void)hipEventDestroy(ctx->x_ev);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : ctx->prof_pool) (void)hipEventDestroy(e);
    if (ctx->xstream) (void)hipStreamDestroy(ctx->xstream);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
}

extern "C" {

int fsim_abi_version(void) { return FSIM_ABI_VERSION; }

const char* fsim_status_string(int s) {
    switch (s) {
    case FSIM_OK: return "ok";
    case FSIM_ERR_INVALID: return "invalid argument";
    case FSIM_ERR_HIP: return "HIP runtime error";
    case FSIM_ERR_UNSUPPORTED: return "unsupported feature";
    case FSIM_ERR_OOM: return "out of device memory";
    case FSIM_ERR_NUMERIC: return "numerical failure";
    case FSIM_ERR_TIMEOUT: return "device wait timed out";
    case FSIM_ERR_NO_DEVICE: return "no HIP device";
    default: return "unknown status";
    }
}

const char* fsim_last_error(const fsim_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int fsim_create(const fsim_grid_desc* grid, const fsim_params* params, fsim_ctx** out) {
    if (!out) return FSIM_ERR_INVALID;
    *out = nullptr;
    FSIM_TRY(check_grid(grid));
    FSIM_TRY(check_params(params, grid->two_d != 0));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return FSIM_ERR_NO_DEVICE;
    fsim_ctx* ctx = new (std::nothrow) fsim_ctx();
    if (!ctx) return FSIM_ERR_OOM;
    auto fail = [&](int rc) {
        free_all(ctx);
        delete ctx;
        return rc;
    };
    if (hipGetDevice(&ctx->device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->xstream, hipStreamNonBlocking) != hipSuccess)
        return fail(FSIM_ERR_HIP);
    GridDims& g = ctx->g;
    g.nx = grid->nx;
    g.ny = grid->ny;
    g.nz = grid->nz;
    g.nc = (int64_t)g.nx * g.ny * g.nz;
    g.cd = grid->cell_d;
    g.cdinv = 1.0 / grid->cell_d;
    g.two_d = grid->two_d ? 1 : 0;
    g.cdz = g.two_d ? grid->cell_d_z : grid->cell_d;
    g.cdzinv = 1.0 / g.cdz;
    g.dims[0] = g.nx * g.cd;
    g.dims[1] = g.ny * g.cd;
    g.dims[2] = g.nz * g.cdz;
    g.pr = grid->particle_r;
    ctx->prm = *params;
    const int64_t nc = g.nc;
    int rc = FSIM_OK;
    if ((rc = dalloc(ctx, &ctx->u, nc)) || (rc = dalloc(ctx, &ctx->v, nc)) || (rc = dalloc(ctx, &ctx->w, nc)) ||
        (rc = dalloc(ctx, &ctx->u2, nc)) || (rc = dalloc(ctx, &ctx->v2, nc)) || (rc = dalloc(ctx, &ctx->w2, nc)) ||
        (rc = dalloc(ctx, &ctx->label, nc)) || (rc = dalloc(ctx, &ctx->valid, nc)) ||
        (rc = dalloc(ctx, &ctx->density, nc)) || (rc = dalloc(ctx, &ctx->cell_count, nc + 1)) ||
        (rc = dalloc(ctx, &ctx->cell_start, nc + 2)) || (rc = dalloc(ctx, &ctx->nbinfo, nc)) ||
        (rc = dalloc(ctx, &ctx->chunk_flags, nc / 64 + 1)) ||
        (rc = dalloc(ctx, &ctx->b, nc)) || (rc = dalloc(ctx, &ctx->r, nc)) || (rc = dalloc(ctx, &ctx->z, nc)) ||
        (rc = dalloc(ctx, &ctx->s, nc)) || (rc = dalloc(ctx, &ctx->s2, nc)) || (rc = dalloc(ctx, &ctx->q, nc)) ||
        (rc = dalloc(ctx, &ctx->y, nc)) || (rc = dalloc(ctx, &ctx->red_count, 1)) ||
        (rc = dalloc(ctx, &ctx->P32, nc)) || (rc = dalloc(ctx, &ctx->P64, nc)) || (rc = dalloc(ctx, &ctx->x, nc)) ||
        (rc = dalloc(ctx, &ctx->scal, 1)) || (rc = dalloc(ctx, &ctx->obs_dev, kMaxObstacles)))
        return fail(rc);
    ctx->scan_tmp_cap = (nc + 2) / 4096 + 2;
    if ((rc = dalloc(ctx, &ctx->scan_tmp, ctx->scan_tmp_cap))) return fail(rc);
    ctx->partials_cap = 2 * (int)nblocks(nc) + 16;
    if ((rc = dalloc(ctx, &ctx->partials, ctx->partials_cap))) return fail(rc);
    if (hipHostMalloc(reinterpret_cast<void**>(&ctx->scal_host), sizeof(SolveScalars), hipHostMallocDefault) != hipSuccess)
        return fail(FSIM_ERR_OOM);
    for (auto& e : ctx->ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(FSIM_ERR_HIP);
    if (hipEventCreateWithFlags(&ctx->res_ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->dir_ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->x_ev, hipEventDisableTiming) != hipSuccess)
        return fail(FSIM_ERR_HIP);
    // initial grid state: zero faces, AIR labels (MacGrid ctor, macGrid.cpp:9-17)
    float* f32s[] = {ctx->u, ctx->v, ctx->w, ctx->u2, ctx->v2, ctx->w2, ctx->density, ctx->b, ctx->r, ctx->z, ctx->s, ctx->s2,
                     ctx->q, ctx->y, ctx->P32};
    for (float* p : f32s)
        if (fill_f32(ctx, p, 0.0f, nc)) return fail(FSIM_ERR_HIP);
    if (fill_u8(ctx, ctx->label, FSIM_AIR, nc) || fill_f64(ctx, ctx->x, 0.0, nc) || fill_f64(ctx, ctx->P64, 0.0, nc) ||
        fill_i32(ctx, ctx->cell_count, 0, nc + 1) || fill_i32(ctx, ctx->cell_start, 0, nc + 2))
        return fail(FSIM_ERR_HIP);
    if (hipMemsetAsync(ctx->nbinfo, 0, sizeof(uint16_t) * nc, ctx->stream) != hipSuccess ||
        hipMemsetAsync(ctx->red_count, 0, sizeof(unsigned), ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return fail(FSIM_ERR_HIP);
    if (particles_reserve(ctx, 1024) != FSIM_OK) return fail(FSIM_ERR_OOM);
    if ((rc = wave_init(ctx)) != FSIM_OK) return fail(rc);
    if ((rc = scratch_init(ctx)) != FSIM_OK) return fail(rc);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(FSIM_ERR_HIP);
    *out = ctx;
    return FSIM_OK;
}

int fsim_destroy(fsim_ctx* ctx) {
    if (!ctx) return FSIM_OK;
    (void)hipStreamSynchronize(ctx->stream);
    free_all(ctx);
    delete ctx;
    return FSIM_OK;
}

int fsim_set_params(fsim_ctx* ctx, const fsim_params* params) {
    if (!ctx) return FSIM_ERR_INVALID;
    int rc = check_params(params, ctx->g.two_d != 0);
    if (rc) return set_err(ctx, rc, "invalid parameters");
    ctx->prm = *params;
    return FSIM_OK;
}

int fsim_get_params(const fsim_ctx* ctx, fsim_params* out) {
    if (!ctx || !out) return FSIM_ERR_INVALID;
    *out = ctx->prm;
    return FSIM_OK;
}

int fsim_set_particle_radius(fsim_ctx* ctx, double r) {
    if (!ctx || !(r >= 0.0)) return FSIM_ERR_INVALID;
    ctx->g.pr = r;
    return FSIM_OK;
}

int fsim_set_particles(fsim_ctx* ctx, int64_t n, const float* pos, const float* vel, const float* c) {
    if (!ctx || n < 0 || (n > 0 && !pos)) return FSIM_ERR_INVALID;
    FSIM_TRY(particles_reserve(ctx, std::max<int64_t>(n, 1)));   // also grows the scan buffer
    std::vector<float> plane((size_t)n);
    auto put = [&](float* dst) -> int {
        FSIM_HIP(ctx, hipMemcpy(dst, plane.data(), sizeof(float) * n, hipMemcpyHostToDevice));
        return FSIM_OK;
    };
    for (int a = 0; a < 3; a++) {
        for (int64_t i = 0; i < n; i++) plane[i] = pos[3 * i + a];
        FSIM_TRY(put(ctx->P.pos[a]));
        for (int64_t i = 0; i < n; i++) plane[i] = vel ? vel[3 * i + a] : 0.0f;
        FSIM_TRY(put(ctx->P.vel[a]));
    }
    for (int k = 0; k < 9; k++) {
        for (int64_t i = 0; i < n; i++) plane[i] = c ? c[9 * i + k] : 0.0f;
        FSIM_TRY(put(ctx->P.c[k]));
    }
    std::vector<int32_t> ids((size_t)n);
    for (int64_t i = 0; i < n; i++) ids[i] = (int32_t)i;
    if (n) FSIM_HIP(ctx, hipMemcpy(ctx->P.pid, ids.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    FSIM_HIP(ctx, hipMemset(ctx->removed, 0, ctx->P.cap));
    ctx->P.n = n;
    ctx->c_valid = true;
    return FSIM_OK;
}

int fsim_num_particles(const fsim_ctx* ctx, int64_t* n_out) {
    if (!ctx || !n_out) return FSIM_ERR_INVALID;
    *n_out = ctx->P.n;
    return FSIM_OK;
}

int fsim_get_slot_ids(fsim_ctx* ctx, int64_t capacity, int32_t* ids, int64_t* n_out) {
    if (!ctx) return FSIM_ERR_INVALID;
    if (n_out) *n_out = ctx->P.n;
    if (!ids) return FSIM_OK;
    if (capacity < ctx->P.n) return set_err(ctx, FSIM_ERR_INVALID, "capacity smaller than the particle count");
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->P.n) FSIM_HIP(ctx, hipMemcpy(ids, ctx->P.pid, sizeof(int32_t) * ctx->P.n, hipMemcpyDeviceToHost));
    return FSIM_OK;
}

int fsim_get_particles(fsim_ctx* ctx, int64_t capacity, float* pos, float* vel, float* c, int64_t* n_out) {
    if (!ctx) return FSIM_ERR_INVALID;
    const int64_t n = ctx->P.n;
    if (n_out) *n_out = n;
    if (!pos && !vel && !c) return FSIM_OK;
    if (capacity < n) return set_err(ctx, FSIM_ERR_INVALID, "capacity smaller than the particle count");
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<int32_t> ids((size_t)n);
    std::vector<float> plane((size_t)n);
    if (n) FSIM_HIP(ctx, hipMemcpy(ids.data(), ctx->P.pid, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++)
        if (ids[i] < 0 || ids[i] >= n) return set_err(ctx, FSIM_ERR_INVALID, "corrupt particle ids");
    auto get = [&](const float* src, float* dst, int stride, int off) -> int {
        if (!n) return FSIM_OK;
        FSIM_HIP(ctx, hipMemcpy(plane.data(), src, sizeof(float) * n, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; i++) dst[(int64_t)stride * ids[i] + off] = plane[i];
        return FSIM_OK;
    };
    for (int a = 0; a < 3; a++) {
        if (pos) FSIM_TRY(get(ctx->P.pos[a], pos, 3, a));
        if (vel) FSIM_TRY(get(ctx->P.vel[a], vel, 3, a));
    }
    if (c) {
        if (ctx->c_valid) {
            for (int k = 0; k < 9; k++) FSIM_TRY(get(ctx->P.c[k], c, 9, k));
        } else {
            std::memset(c, 0, sizeof(float) * 9 * n);
        }
    }
    return FSIM_OK;
}

int fsim_set_obstacles(fsim_ctx* ctx, int n, const fsim_obstacle* obs) {
    if (!ctx || n < 0 || n > kMaxObstacles || (n > 0 && !obs)) return FSIM_ERR_INVALID;
    for (int i = 0; i < n; i++)
        if (obs[i].kind < FSIM_OBS_BOX || obs[i].kind > FSIM_OBS_SINK)
            return set_err(ctx, FSIM_ERR_INVALID, "unknown obstacle kind");
    ctx->obs.assign(obs, obs + n);
    return FSIM_OK;
}

int fsim_get_obstacles(const fsim_ctx* ctx, int capacity, fsim_obstacle* out, int* n_out) {
    if (!ctx) return FSIM_ERR_INVALID;
    if (n_out) *n_out = (int)ctx->obs.size();
    if (!out) return FSIM_OK;
    if (capacity < (int)ctx->obs.size()) return FSIM_ERR_INVALID;
    std::copy(ctx->obs.begin(), ctx->obs.end(), out);
    return FSIM_OK;
}

int fsim_get_stats(const fsim_ctx* ctx, fsim_step_stats* out) {
    if (!ctx || !out) return FSIM_ERR_INVALID;
    *out = ctx->stats;
    return FSIM_OK;
}

static int run_stage(fsim_ctx* ctx, int stage, double dt) {
    switch (stage) {
    case FSIM_STAGE_SPAWN: return stage_spawn(ctx, dt);
    case FSIM_STAGE_ADVECT: {
        FSIM_TRY(upload_obstacles(ctx));
        FSIM_TRY(stage_move(ctx, dt, 1));
        // removeParticles runs inside advectParticles (simulator.cpp:250): drop
        // the particles a sink took now (the binning compacts them)
        bool sink = false;
        for (const fsim_obstacle& o : ctx->obs) sink = sink || o.kind == FSIM_OBS_SINK;
        return (sink && ctx->prm.despawning_enabled) ? stage_bin(ctx) : FSIM_OK;
    }
    case FSIM_STAGE_PUSH_OUT: FSIM_TRY(upload_obstacles(ctx)); return stage_move(ctx, dt, 2);
    case FSIM_STAGE_STOP: return stage_move(ctx, dt, 4);
    case FSIM_STAGE_BIN: return stage_bin(ctx);
    case FSIM_STAGE_P2G: FSIM_TRY(stage_bin(ctx)); return stage_p2g(ctx);
    case FSIM_STAGE_MARK: FSIM_TRY(stage_bin(ctx)); return stage_mark(ctx);
    case FSIM_STAGE_OBSTACLES: return stage_obstacles(ctx);
    case FSIM_STAGE_BORDERS: return stage_borders(ctx);
    case FSIM_STAGE_POST_P2G:
        return stage_post_p2g(ctx, ctx->prm.gravity_enabled ? (double)ctx->prm.gravity * dt : 0.0);
    case FSIM_STAGE_SOLVE: return stage_solve(ctx, dt);
    case FSIM_STAGE_EXTRAPOLATE: return stage_extrapolate(ctx);
    case FSIM_STAGE_G2P: return stage_g2p(ctx);
    case FSIM_STAGE_PUSH_APART: return stage_push_apart(ctx);
    default: return set_err(ctx, FSIM_ERR_INVALID, "unknown stage");
    }
}

int fsim_run_stage(fsim_ctx* ctx, int stage, double dt) {
    if (!ctx) return FSIM_ERR_INVALID;
    FSIM_TRY(run_stage(ctx, stage, dt));
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return prof_collect(ctx);
}

int fsim_step(fsim_ctx* ctx, double dt, fsim_step_stats* stats) {
    if (!ctx) return FSIM_ERR_INVALID;
    if (!(dt > 0.0)) return set_err(ctx, FSIM_ERR_INVALID, "dt must be positive");
    auto t0 = std::chrono::steady_clock::now();
    hipEvent_t* ev = ctx->ev;
    const bool apart = ctx->prm.push_apart_enabled != 0;
    FSIM_HIP(ctx, hipEventRecord(ev[0], ctx->stream));
    FSIM_TRY(upload_obstacles(ctx));
    if (ctx->prm.spawning_enabled) FSIM_TRY(stage_spawn(ctx, dt));
    if (apart) {   // advect | push apart | push out + stop (simulator.cpp:57-74)
        FSIM_TRY(stage_move(ctx, dt, 1));
        FSIM_HIP(ctx, hipEventRecord(ev[7], ctx->stream));
        FSIM_TRY(stage_push_apart(ctx));
        FSIM_HIP(ctx, hipEventRecord(ev[8], ctx->stream));
        FSIM_TRY(stage_move(ctx, dt, 2 | 4));
    } else {
        FSIM_TRY(stage_move(ctx, dt, 1 | 2 | 4));   // advect and push-out fused per particle
    }
    FSIM_HIP(ctx, hipEventRecord(ev[1], ctx->stream));   // SimulateParticles (+ push-out when fused)
    FSIM_TRY(stage_bin(ctx));
    FSIM_TRY(stage_p2g_fused(ctx));   // reset + P2G + fluid marking + particle density
    FSIM_HIP(ctx, hipEventRecord(ev[2], ctx->stream));   // P2GTransfer
    FSIM_TRY(stage_obstacles(ctx));
    FSIM_TRY(stage_borders(ctx));
    FSIM_TRY(stage_post_p2g(ctx, ctx->prm.gravity_enabled ? (double)ctx->prm.gravity * dt : 0.0));
    FSIM_HIP(ctx, hipEventRecord(ev[3], ctx->stream));   // IncompressibilityPrep
    FSIM_TRY(stage_solve(ctx, dt));
    FSIM_HIP(ctx, hipEventRecord(ev[4], ctx->stream));   // Incompressibility
    FSIM_TRY(stage_extrapolate(ctx));
    FSIM_HIP(ctx, hipEventRecord(ev[5], ctx->stream));   // VelocityExtrapolation
    FSIM_TRY(stage_g2p(ctx));
    FSIM_HIP(ctx, hipEventRecord(ev[6], ctx->stream));   // G2PTransfer
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    fsim_step_stats& st = ctx->stats;
    float ms[6];
    for (int k = 0; k < 6; k++) FSIM_HIP(ctx, hipEventElapsedTime(&ms[k], ev[k], ev[k + 1]));
    std::memset(st.timer_ms, 0, sizeof(st.timer_ms));
    st.timer_ms[FSIM_T_SIMULATE_PARTICLES] = ms[0];
    if (apart) {
        float a[3];
        FSIM_HIP(ctx, hipEventElapsedTime(&a[0], ev[0], ev[7]));
        FSIM_HIP(ctx, hipEventElapsedTime(&a[1], ev[7], ev[8]));
        FSIM_HIP(ctx, hipEventElapsedTime(&a[2], ev[8], ev[1]));
        st.timer_ms[FSIM_T_SIMULATE_PARTICLES] = a[0];
        st.timer_ms[FSIM_T_PUSH_APART] = a[1];
        st.timer_ms[FSIM_T_PUSH_OUT] = a[2];
    }
    st.timer_ms[FSIM_T_P2G] = ms[1];
    st.timer_ms[FSIM_T_INCOMPRESSIBILITY_PREP] = ms[2];
    st.timer_ms[FSIM_T_INCOMPRESSIBILITY] = ms[3];
    st.timer_ms[FSIM_T_EXTRAPOLATION] = ms[4];
    st.timer_ms[FSIM_T_G2P] = ms[5];
    st.solve_ms = ms[3];
    st.num_particles = ctx->P.n;
    st.step_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
    return prof_collect(ctx);
}

int fsim_get_grid(fsim_ctx* ctx, int field, void* dst, size_t bytes) {
    if (!ctx || !dst) return FSIM_ERR_INVALID;
    const int64_t nc = ctx->g.nc;
    const void* src = nullptr;
    size_t need = 0;
    switch (field) {
    case FSIM_FIELD_U: src = ctx->u; need = 4 * nc; break;
    case FSIM_FIELD_V: src = ctx->v; need = 4 * nc; break;
    case FSIM_FIELD_W: src = ctx->w; need = 4 * nc; break;
    case FSIM_FIELD_U2: src = ctx->u2; need = 4 * nc; break;
    case FSIM_FIELD_V2: src = ctx->v2; need = 4 * nc; break;
    case FSIM_FIELD_W2: src = ctx->w2; need = 4 * nc; break;
    case FSIM_FIELD_LABEL: src = ctx->label; need = nc; break;
    case FSIM_FIELD_DENSITY: src = ctx->density; need = 4 * nc; break;
    case FSIM_FIELD_PRESSURE: src = ctx->x; need = 8 * nc; break;
    case FSIM_FIELD_RHS: src = ctx->b; need = 4 * nc; break;
    case FSIM_FIELD_PRECON: src = ctx->P64; need = 8 * nc; break;
    case FSIM_FIELD_CELL_COUNT: src = ctx->cell_count; need = 4 * nc; break;
    case FSIM_FIELD_CELL_START: src = ctx->cell_start; need = 4 * (nc + 1); break;
    default: return set_err(ctx, FSIM_ERR_INVALID, "unknown field");
    }
    if (bytes < need) return set_err(ctx, FSIM_ERR_INVALID, "destination too small");
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const bool rotated = field == FSIM_FIELD_PRESSURE || field == FSIM_FIELD_RHS || field == FSIM_FIELD_PRECON;
    if (!rotated) {
        FSIM_HIP(ctx, hipMemcpy(dst, src, need, hipMemcpyDeviceToHost));
        return FSIM_OK;
    }
    // solver vectors are stored in the rotated layout: gather into flat order
    const size_t elem = need / nc;
    std::vector<unsigned char> tmp(need);
    FSIM_HIP(ctx, hipMemcpy(tmp.data(), src, need, hipMemcpyDeviceToHost));
    unsigned char* out = static_cast<unsigned char*>(dst);
    const GridDims& g = ctx->g;
    for (int x = 0; x < g.nx; x++)
        for (int y = 0; y < g.ny; y++)
            for (int z = 0; z < g.nz; z++) {
                const int64_t i = ((int64_t)x * g.ny + y) * g.nz + z;
                std::memcpy(out + i * elem, tmp.data() + ridx(x, y, z, g.nx, g.ny, g.nz) * elem, elem);
            }
    return FSIM_OK;
}

int fsim_set_grid(fsim_ctx* ctx, int field, const void* src, size_t bytes) {
    if (!ctx || !src) return FSIM_ERR_INVALID;
    const int64_t nc = ctx->g.nc;
    void* dst = nullptr;
    size_t need = 0;
    switch (field) {
    case FSIM_FIELD_U: dst = ctx->u; need = 4 * nc; break;
    case FSIM_FIELD_V: dst = ctx->v; need = 4 * nc; break;
    case FSIM_FIELD_W: dst = ctx->w; need = 4 * nc; break;
    case FSIM_FIELD_U2: dst = ctx->u2; need = 4 * nc; break;
    case FSIM_FIELD_V2: dst = ctx->v2; need = 4 * nc; break;
    case FSIM_FIELD_W2: dst = ctx->w2; need = 4 * nc; break;
    case FSIM_FIELD_LABEL: dst = ctx->label; need = nc; break;
    case FSIM_FIELD_DENSITY: dst = ctx->density; need = 4 * nc; break;
    default: return set_err(ctx, FSIM_ERR_INVALID, "field is read-only or unknown");
    }
    if (bytes < need) return set_err(ctx, FSIM_ERR_INVALID, "source too small");
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    FSIM_HIP(ctx, hipMemcpy(dst, src, need, hipMemcpyHostToDevice));
    return FSIM_OK;
}

int fsim_gfx_snapshot(fsim_ctx* ctx, int64_t capacity, fsim_gfx_particle* out, int64_t* n_out) {
    if (!ctx) return FSIM_ERR_INVALID;
    if (n_out) *n_out = ctx->P.n;
    if (!out) return FSIM_OK;
    if (capacity < ctx->P.n) return set_err(ctx, FSIM_ERR_INVALID, "capacity smaller than the particle count");
    return gfx_snapshot(ctx, out);
}

int fsim_gfx_snapshot_noalloc(fsim_ctx* ctx, int64_t capacity, fsim_gfx_particle* out, int64_t* n_out) {
    if (!ctx) return FSIM_ERR_INVALID;
    if (n_out) *n_out = ctx->P.n;
    if (!out) return FSIM_OK;
    if (capacity < ctx->P.n) return set_err(ctx, FSIM_ERR_INVALID, "capacity smaller than the particle count");
    return gfx_snapshot_noalloc(ctx, out);
}

int64_t fsim_gfx_snapshot_noalloc_max(void) { return scratch_particles(); }

int fsim_num_kernels(void) { return KID_COUNT; }

const char* fsim_kernel_name(int kid) { return (kid >= 0 && kid < KID_COUNT) ? kKernelNames[kid] : "unknown"; }

int fsim_get_kernel_time(fsim_ctx* ctx, int kid, double* total_ms, int64_t* launches) {
    if (!ctx || kid < 0 || kid >= KID_COUNT) return FSIM_ERR_INVALID;
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    FSIM_TRY(prof_collect(ctx));
    if (total_ms) *total_ms = ctx->prof_ms[kid];
    if (launches) *launches = ctx->prof_n[kid];
    return FSIM_OK;
}

int fsim_reset_kernel_times(fsim_ctx* ctx) {
    if (!ctx) return FSIM_ERR_INVALID;
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    FSIM_TRY(prof_collect(ctx));
    std::fill(std::begin(ctx->prof_ms), std::end(ctx->prof_ms), 0.0);
    std::fill(std::begin(ctx->prof_n), std::end(ctx->prof_n), 0);
    return FSIM_OK;
}

int fsim_set_option(fsim_ctx* ctx, int option, int value) {
    if (!ctx) return FSIM_ERR_INVALID;
    switch (option) {
    case FSIM_OPT_SWEEP_MODE:
        if (value != 0 && value != 1) return set_err(ctx, FSIM_ERR_INVALID, "sweep mode is 0 or 1");
        if (value == 0 && ctx->g.nc > kWaveMaxCells)
            return set_err(ctx, FSIM_ERR_INVALID, "grid too large for the wavefront sweeps (32-bit offsets)");
        ctx->sweep_mode = value;
        return FSIM_OK;
    case FSIM_OPT_PROFILE:
        return set_profiling(ctx, value != 0);
    case FSIM_OPT_LEARNED_SMOOTHING:
        if (value < 0 || value > 16) return set_err(ctx, FSIM_ERR_INVALID, "learned smoothing steps are 0 .. 16");
        ctx->learned_smoothing = value;
        return FSIM_OK;
    case FSIM_OPT_LEARNED_LOSS:
        if (value != 0 && value != 1) return set_err(ctx, FSIM_ERR_INVALID, "learned loss is 0 (MSE) or 1 (energy)");
        ctx->learned_loss = value;
        return FSIM_OK;
    case FSIM_OPT_LEARNED_CONV:
        if (value != 0 && value != 1) return set_err(ctx, FSIM_ERR_INVALID, "learned convolution kernel is 0 or 1");
        ctx->learned_conv_lds = value;
        return FSIM_OK;
    case FSIM_OPT_LEARNED_REFRESH:
        if (value < 0) return set_err(ctx, FSIM_ERR_INVALID, "learned refresh period must be >= 0");
        ctx->learned_refresh = value;
        return FSIM_OK;
    case FSIM_OPT_LEARNED_MIC_COUNT:
        if (value <= 0) return set_err(ctx, FSIM_ERR_INVALID, "the MIC(0) iteration count must be positive");
        ctx->last_mic_iterations = value;
        ctx->learned_since_mic = 0;
        return FSIM_OK;
    case FSIM_OPT_P2G_KERNEL:
        if (value != 0 && value != 1) return set_err(ctx, FSIM_ERR_INVALID, "P2G kernel option is 0 or 1");
        ctx->p2g_rows = value;
        return FSIM_OK;
    case FSIM_OPT_WAVE_SEND_DELAY:
        if (value < 0 || value > 64) return set_err(ctx, FSIM_ERR_INVALID, "sending-wave delay is 0 .. 64");
        ctx->wave_send_delay = value;
        return FSIM_OK;
    case FSIM_OPT_MG_GRAPH:
        if (value != 0 && value != 1) return set_err(ctx, FSIM_ERR_INVALID, "multigrid graph option is 0 or 1");
        ctx->mg_graph = value;
        return FSIM_OK;
    case FSIM_OPT_RELIABLE_DROP:
        if (value < 0 || value > 15) return set_err(ctx, FSIM_ERR_INVALID, "reliable-update drop is 0 .. 15");
        ctx->reliable_drop = value;
        return FSIM_OK;
    case FSIM_OPT_MG_COARSE:
        if (value < FSIM_MG_COARSE_NETWORK || value > FSIM_MG_COARSE_ACCURATE)
            return set_err(ctx, FSIM_ERR_INVALID, "unknown coarse-level mode");
        ctx->mg_coarse = value;
        return FSIM_OK;
    default: return set_err(ctx, FSIM_ERR_INVALID, "unknown option");
    }
}

int fsim_solver_apply(fsim_ctx* ctx, int op, double dt, const double* in, double* out) {
    if (!ctx || !out || (op != FSIM_OP_FACTOR_MIC0 && !in) || !(dt > 0.0)) return FSIM_ERR_INVALID;
    return solver_apply(ctx, op, dt, in, out);
}

int64_t fsim_learned_num_params(void) { return learned_num_params(); }

int fsim_learned_init(fsim_ctx* ctx, uint64_t seed) {
    if (!ctx) return FSIM_ERR_INVALID;
    return learned_init(ctx, seed);
}

int fsim_learned_set_params(fsim_ctx* ctx, const float* params) {
    if (!ctx || !params) return FSIM_ERR_INVALID;
    for (int64_t i = 0; i < learned_num_params(); i++)
        if (!std::isfinite(params[i])) return set_err(ctx, FSIM_ERR_INVALID, "non-finite network parameter");
    return learned_set_params(ctx, params);
}

int fsim_learned_get_params(fsim_ctx* ctx, float* params) {
    if (!ctx || !params) return FSIM_ERR_INVALID;
    return learned_get_params(ctx, params);
}

// the operator of a label field, or of op, as flat float arrays [4][Nc]
static std::vector<float> label_op(const fsim_ctx* ctx, const uint8_t* label, int n = 1) {
    const int64_t nc = ctx->g.nc;
    std::vector<float> op((size_t)(4 * nc * n));
    for (int k = 0; k < n; k++) label_operator(ctx->g.nx, ctx->g.ny, ctx->g.nz, label + k * nc, op.data() + 4 * k * nc);
    return op;
}

int fsim_learned_forward(fsim_ctx* ctx, const float* b, const uint8_t* label, float* out) {
    if (!ctx || !b || !label || !out) return FSIM_ERR_INVALID;
    return learned_forward_host(ctx, b, label_op(ctx, label).data(), out);
}

int fsim_learned_forward_op(fsim_ctx* ctx, const float* b, const float* op, float* out) {
    if (!ctx || !b || !op || !out) return FSIM_ERR_INVALID;
    return learned_forward_host(ctx, b, op, out);
}

int fsim_learned_grad(fsim_ctx* ctx, const float* b, const uint8_t* label, const double* p, float* grad,
                      double* loss) {
    if (!ctx || !b || !label || !p || !grad || !loss) return FSIM_ERR_INVALID;
    return learned_grad_host(ctx, b, label_op(ctx, label).data(), p, grad, loss);
}

int fsim_learned_grad_op(fsim_ctx* ctx, const float* b, const float* op, const double* p, float* grad, double* loss) {
    if (!ctx || !b || !op || !p || !grad || !loss) return FSIM_ERR_INVALID;
    return learned_grad_host(ctx, b, op, p, grad, loss);
}

int fsim_learned_train_batch(fsim_ctx* ctx, int n, const float* b, const uint8_t* label, const double* p, float lr,
                             double* mean_loss) {
    if (!ctx || !b || !label || !p || !mean_loss || n <= 0 || !(lr > 0.f)) return FSIM_ERR_INVALID;
    return learned_train_batch(ctx, n, b, label_op(ctx, label, n).data(), p, lr, mean_loss);
}

int fsim_learned_train_batch_op(fsim_ctx* ctx, int n, const float* b, const float* op, const double* p, float lr,
                                double* mean_loss) {
    if (!ctx || !b || !op || !p || !mean_loss || n <= 0 || !(lr > 0.f)) return FSIM_ERR_INVALID;
    return learned_train_batch(ctx, n, b, op, p, lr, mean_loss);
}

int fsim_mg_info(fsim_ctx* ctx, int* levels, int dims[3]) {
    if (!ctx) return FSIM_ERR_INVALID;
    return mg_info(ctx, levels, dims);
}

int fsim_learned_harvest(fsim_ctx* ctx, double dt, int capacity, float* b, float* op, float* x, int* n_out) {
    if (!ctx || capacity < 0 || !n_out || (capacity > 0 && (!b || !op || !x)) || !(dt > 0.0)) return FSIM_ERR_INVALID;
    *n_out = 0;
    FSIM_TRY(mg_harvest_begin(ctx, std::max(capacity, 1)));
    const int precond = ctx->prm.precond, coarse = ctx->mg_coarse;
    ctx->prm.precond = FSIM_PRECOND_LEARNED;
    ctx->mg_coarse = FSIM_MG_COARSE_ACCURATE;
    ctx->mg_harvesting = true;
    int rc = stage_solve(ctx, dt);
    if (rc == FSIM_OK) rc = hipStreamSynchronize(ctx->stream) == hipSuccess ? FSIM_OK : FSIM_ERR_HIP;
    ctx->prm.precond = precond;
    ctx->mg_coarse = coarse;
    ctx->mg_harvesting = false;
    if (rc != FSIM_OK) return rc;
    if (capacity == 0) return FSIM_OK;
    return mg_harvest_end(ctx, capacity, b, op, x, n_out);
}

}  // extern "C"
