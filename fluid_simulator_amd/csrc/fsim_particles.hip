// fsim_particles.hip — particle-side stages of the step on gfx950.
//
//  * stage_move:  advectParticles (simulator.cpp:134-251), removal of
//                 particles that hit a sink, pushParticlesOutOfObstacles
//                 (simulator.cpp:253-312) and stopParticles
//                 (simulator.cpp:71-74), fused per particle.  The particle
//                 state is FP32 but the collision logic runs in FP64 with
//                 contraction off, so every branch (wall hit, back-step
//                 count, push axis) is decided exactly as an FP64 evaluation
//                 of the same expression order would decide it.
//  * stage_spawn: spawnParticles (simulator.cpp:106-125) with a SplitMix64
//                 counter RNG in place of std::rand (random.h:13-15).
//  * stage_bin:   counting sort of the particles by grid cell: key +
//                 histogram (integer atomics), exclusive scan, scatter,
//                 per-cell sort by particle id, payload gather.  The result
//                 (counts, offsets, permutation) is the serial stable
//                 counting sort by (cell, particle id), bit for bit.
#include <hip/hip_runtime.h>

#include <climits>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "fsim_internal.hpp"

namespace fsim {

// ---- SplitMix64 counter RNG (Steele, Lea, Flood 2014) ----------------------
__host__ __device__ inline uint64_t splitmix64(uint64_t seed, uint64_t n) {
    uint64_t z = seed + (n + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__host__ __device__ inline double u01d(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }
__host__ __device__ inline float u01f(uint64_t x) { return (float)(x >> 40) * (1.0f / 16777216.0f); }
constexpr uint64_t kSpawnSalt = 0x5350415741ULL;

struct MoveArgs {
    float* pos[3];
    float* vel[3];
    uint8_t* removed;
    const fsim_obstacle* obs;
    int nobs;
    int64_t n;
    double dt, pr;
    double cd[3], cdinv[3];   // cellD, cellDInv per axis (z differs in 2D)
    double dims[3];
    int ncell[3];
    int flags;
    int despawn;
    int two_d;
};

__host__ __device__ inline bool is_sphere(int k) { return k == FSIM_OBS_SPHERE || k == FSIM_OBS_SOURCE || k == FSIM_OBS_SINK; }

__host__ __device__ inline bool in_rect(const double p[3], double pr, const double rp[3], const double rs[3]) {
#pragma clang fp contract(off)
    return fabs(p[0] - rp[0]) < rs[0] * 0.5 + pr && fabs(p[1] - rp[1]) < rs[1] * 0.5 + pr &&
           fabs(p[2] - rp[2]) < rs[2] * 0.5 + pr;
}

__host__ __device__ inline double dot3(const double a[3], const double b[3]) {
#pragma clang fp contract(off)
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// advectParticles per-particle body, simulator.cpp:146-249.  Returns true if
// the particle entered a despawning sink.
__host__ __device__ bool advect_one(const MoveArgs& A, double pos[3], double v[3], const double lo[3], const double hi[3]) {
#pragma clang fp contract(off)
    const double wall_rest = 0.3, sphere_rest = 1.0, rect_rest = 0.2;
    const double dt = A.dt, pr = A.pr;
    bool removed = false;
    double t = 0;
    int run = 0;
    const int max_run = 200;
    while (run < max_run) {
        run++;
        double min_t = 1e6;
        int min_axis = 0;
        for (int a = 0; a < 3; a++) {
            double comp = v[a];
            double tmp = 1e6;
            if (comp > 1e-6) tmp = (hi[a] - pos[a]) / comp;
            else if (comp < -1e-6) tmp = (pos[a] - lo[a]) / -comp;
            if (tmp < min_t) { min_t = tmp; min_axis = a; }
        }
        if (min_t <= (dt - t)) {
            for (int a = 0; a < 3; a++) pos[a] += v[a] * min_t * 0.999;
            v[min_axis] *= -wall_rest;
            t += 0.999 * min_t;
            continue;
        }
        bool collision = false;
        for (int a = 0; a < 3; a++) pos[a] += v[a] * (dt - t);
        for (int oi = 0; oi < A.nobs; oi++) {
            const fsim_obstacle& o = A.obs[oi];
            if (is_sphere(o.kind)) {
                double r = o.r + pr;
                double d3[3] = {o.pos[0] - pos[0], o.pos[1] - pos[1], o.pos[2] - pos[2]};
                double d = sqrt(dot3(d3, d3)) - r;
                if (d < 0) {
                    if (o.kind == FSIM_OBS_SINK && A.despawn) {
                        removed = true;
                        collision = false;
                        break;
                    }
                    double bt = 0;
                    for (;;) {
                        double q[3];
                        for (int a = 0; a < 3; a++) q[a] = (pos[a] - v[a] * bt) - (o.pos[a] - o.speed[a] * bt);
                        if (!(sqrt(dot3(q, q)) < r && bt < t + (double)0.001f)) break;
                        bt += 0.0002;
                    }
                    if (bt > dt - t) continue;
                    bt += 0.0002;
                    double nrm[3], relv[3];
                    for (int a = 0; a < 3; a++) pos[a] -= bt * v[a];
                    for (int a = 0; a < 3; a++) nrm[a] = pos[a] - (o.pos[a] - o.speed[a] * bt);
                    double inv = 1.0 / sqrt(dot3(nrm, nrm));
                    for (int a = 0; a < 3; a++) nrm[a] *= inv;
                    for (int a = 0; a < 3; a++) relv[a] = v[a] - o.speed[a];
                    double ssn = -dot3(nrm, relv);
                    if (ssn <= 0.0f) continue;
                    for (int a = 0; a < 3; a++) {
                        double mirror = -relv[a] / ssn;
                        v[a] = (nrm[a] - mirror) * 2.0 + mirror;
                        v[a] *= ssn * sphere_rest;
                        v[a] += o.speed[a];
                    }
                    t += (dt - t) - bt;
                    collision = true;
                    break;
                }
            }
            if (o.kind == FSIM_OBS_BOX) {
                if (in_rect(pos, pr, o.pos, o.size)) {
                    double bt = 0;
                    for (;;) {
                        double pp[3], op[3];
                        for (int a = 0; a < 3; a++) {
                            pp[a] = pos[a] - v[a] * bt;
                            op[a] = o.pos[a] - o.speed[a] * bt;
                        }
                        if (!(in_rect(pp, pr, op, o.size) && bt < dt - t)) break;
                        bt += 0.0002;
                    }
                    if (bt > dt - t) continue;
                    bt += 0.0002;
                    double ptmp[3];
                    for (int a = 0; a < 3; a++) pos[a] -= bt * v[a];
                    for (int a = 0; a < 3; a++) ptmp[a] = o.pos[a] - o.speed[a] * bt;
                    for (int a = 0; a < 3; a++) {
                        if (pos[a] >= ptmp[a] + o.size[a] * 0.5 + pr) {
                            v[a] = -v[a] * rect_rest + o.speed[a];
                            break;
                        } else if (pos[a] <= ptmp[a] - o.size[a] * 0.5 - pr) {
                            v[a] = -v[a] * rect_rest - o.speed[a];
                            break;
                        }
                    }
                    t += (dt - t) - bt;
                    collision = true;
                    break;
                }
            }
        }
        if (!collision) run = max_run;
    }
    for (int a = 0; a < 3; a++) pos[a] = pos[a] < lo[a] ? lo[a] : (pos[a] > hi[a] ? hi[a] : pos[a]);
    return removed;
}

// MacGrid::getMinMaxRect, macGrid.h:139-145, expanded by the particle radius
// (simulator.cpp:281-285).
__host__ __device__ inline void rect_bounds(const MoveArgs& A, const fsim_obstacle& o, double lo[3], double hi[3]) {
#pragma clang fp contract(off)
    for (int a = 0; a < 3; a++) {
        double c = o.pos[a] * A.cdinv[a];
        double d = o.size[a] * A.cdinv[a] * 0.5;
        int mn = (int)fmax(round(c - d), 1.0);
        int mx = (int)fmin(round(c + d), A.ncell[a] - 1.0);
        lo[a] = (double)mn * A.cd[a] - A.pr;
        hi[a] = (double)mx * A.cd[a] + A.pr;
    }
}

// pushParticlesOutOfObstacles per particle, simulator.cpp:263-310.  The
// reference runs one particle pass per obstacle; passes are independent per
// particle, so looping the obstacles inside one pass is the same map.
__host__ __device__ void push_out_one(const MoveArgs& A, double pos[3], const double plo[3], const double phi[3]) {
#pragma clang fp contract(off)
    for (int oi = 0; oi < A.nobs; oi++) {
        const fsim_obstacle& o = A.obs[oi];
        if (is_sphere(o.kind)) {
            double r = o.r + A.pr, r2 = r * r;
            double o2p[3] = {pos[0] - o.pos[0], pos[1] - o.pos[1], pos[2] - o.pos[2]};
            double d2 = dot3(o2p, o2p);
            if (d2 < r2) {
                double d = sqrt(d2);
                for (int a = 0; a < 3; a++) {
                    double val = o.pos[a] + o2p[a] / d * r;
                    pos[a] = val < plo[a] ? plo[a] : (val > phi[a] ? phi[a] : val);
                }
            }
        }
        if (o.kind == FSIM_OBS_BOX) {
            double st[3], en[3];
            rect_bounds(A, o, st, en);
            if (pos[0] > en[0] || pos[0] < st[0] || pos[1] > en[1] || pos[1] < st[1] || pos[2] > en[2] ||
                pos[2] < st[2])
                continue;
            int axis = 0;
            double amount = 1e6;
            for (int a = 0; a < 3; a++) {
                if (pos[a] > o.pos[a]) {
                    double tmp = en[a] - pos[a];
                    if (tmp < fabs(amount)) { amount = tmp; axis = a; }
                } else {
                    double tmp = st[a] - pos[a];
                    if (-tmp < fabs(amount)) { amount = tmp; axis = a; }
                }
            }
            double val = pos[axis] + amount;
            pos[axis] = val < plo[axis] ? plo[axis] : (val > phi[axis] ? phi[axis] : val);
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_move(MoveArgs A) {
#pragma clang fp contract(off)
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    double pos[3], v[3];
    for (int a = 0; a < 3; a++) {
        pos[a] = A.pos[a][i];
        v[a] = A.vel[a][i];
    }
    bool removed = false;
    if (A.flags & 1) {
        double lo[3], hi[3];
        // gridLow = cellD + (r, r, 2D ? 0 : r) * 1.01 (simulator.cpp:140)
        for (int a = 0; a < 3; a++) {
            lo[a] = A.cd[a] + (a == 2 && A.two_d ? 0.0 : A.pr) * 1.01;
            hi[a] = A.dims[a] - lo[a];
        }
        removed = advect_one(A, pos, v, lo, hi);
        if (removed) A.removed[i] = 1;
    }
    if ((A.flags & 2) && !removed) {
        double plo[3], phi[3];
        for (int a = 0; a < 3; a++) {   // particleLow, simulator.cpp:257
            plo[a] = A.cd[a] + A.pr * 1.01;
            phi[a] = A.dims[a] - plo[a];
        }
        push_out_one(A, pos, plo, phi);
    }
    if (A.flags & 4) v[0] = v[1] = v[2] = 0.0;
    for (int a = 0; a < 3; a++) {
        A.pos[a][i] = (float)pos[a];
        A.vel[a][i] = (float)v[a];
    }
}

int upload_obstacles(fsim_ctx* ctx) {
    if (ctx->obs.empty()) return FSIM_OK;
    FSIM_HIP(ctx, hipMemcpyAsync(ctx->obs_dev, ctx->obs.data(), sizeof(fsim_obstacle) * ctx->obs.size(),
                                 hipMemcpyHostToDevice, ctx->stream));
    return FSIM_OK;
}

int stage_move(fsim_ctx* ctx, double dt, int flags) {
    if (ctx->P.n == 0) return FSIM_OK;
    FSIM_PROF(ctx, KID_MOVE);
    if (!ctx->prm.stop_particles) flags &= ~4;
    MoveArgs A{};
    for (int a = 0; a < 3; a++) {
        A.pos[a] = ctx->P.pos[a];
        A.vel[a] = ctx->P.vel[a];
        A.dims[a] = ctx->g.dims[a];
    }
    A.removed = ctx->removed;
    A.obs = ctx->obs_dev;
    A.nobs = (int)ctx->obs.size();
    A.n = ctx->P.n;
    A.dt = dt;
    A.pr = ctx->g.pr;
    for (int a = 0; a < 3; a++) {
        A.cd[a] = a == 2 ? ctx->g.cdz : ctx->g.cd;
        A.cdinv[a] = a == 2 ? ctx->g.cdzinv : ctx->g.cdinv;
    }
    A.two_d = ctx->g.two_d;
    A.ncell[0] = ctx->g.nx;
    A.ncell[1] = ctx->g.ny;
    A.ncell[2] = ctx->g.nz;
    A.flags = flags;
    A.despawn = ctx->prm.despawning_enabled;
    hipLaunchKernelGGL(k_move, dim3(nblocks(A.n)), dim3(kBlock), 0, ctx->stream, A);
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

// ---- pushParticlesApart (hashedParticles.cpp:64-107) -----------------------------
// Deterministic Jacobi form of the reference's racy parallel loop: every
// particle sums, from the positions before the pass, its own offset
// (p_i - p_j) (D - |p_i - p_j|) / |p_i - p_j| * 0.5 for each other particle
// closer than the diameter D = 2r and not closer than 1e-4 (:86-95), visiting
// them through the cell binning in (cell, id) order (all lie within
// R = ceil(D / cellD) cells), and is clamped once (:96-101).  An isolated pair
// ends exactly D apart, as in the reference.  FP64 sums in that fixed order;
// the result goes to the alternate position planes.
struct ApartArgs {
    const float* pos[3];
    float* out[3];
    const int32_t* start;
    int64_t n;
    int nx, ny, nz;
    int R[3];            // search reach in cells per axis
    double cdinv[3], D, D2;
    double lo[3], hi[3];
    int two_d;           // 2D: z is set to the depth / 2 (hashedParticles.cpp:96-101)
    double zc;
};

__global__ void __launch_bounds__(kBlock) k_push_apart(ApartArgs A) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const double pi[3] = {(double)A.pos[0][i], (double)A.pos[1][i], (double)A.pos[2][i]};
    const int n3[3] = {A.nx, A.ny, A.nz};
    int c[3];
#pragma unroll
    for (int a = 0; a < 3; a++) c[a] = min(max((int)(pi[a] * A.cdinv[a]), 0), n3[a] - 1);   // the bin key's cell
    double q[3] = {pi[0], pi[1], pi[2]};
    for (int x = max(c[0] - A.R[0], 0); x <= min(c[0] + A.R[0], A.nx - 1); x++)
        for (int y = max(c[1] - A.R[1], 0); y <= min(c[1] + A.R[1], A.ny - 1); y++) {
            const int64_t row = ((int64_t)x * A.ny + y) * A.nz;
            const int z0 = max(c[2] - A.R[2], 0), z1 = min(c[2] + A.R[2], A.nz - 1);
            // the cells z0..z1 of one row are one contiguous slot range
            const int b = A.start[row + z0], e = A.start[row + z1 + 1];
            for (int j = b; j < e; j++) {
                if (j == (int)i) continue;
                const double d[3] = {pi[0] - (double)A.pos[0][j], pi[1] - (double)A.pos[1][j],
                                     pi[2] - (double)A.pos[2][j]};
                const double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                if (d2 > A.D2 || d2 < 1e-8) continue;
                const double dist = sqrt(d2);
                const double tmp = (A.D - dist) / dist;
#pragma unroll
                for (int a = 0; a < 3; a++) q[a] += d[a] * tmp * 0.5;
            }
        }
#pragma unroll
    for (int a = 0; a < 3; a++) A.out[a][i] = (float)(q[a] < A.lo[a] ? A.lo[a] : (q[a] > A.hi[a] ? A.hi[a] : q[a]));
    if (A.two_d) A.out[2][i] = (float)A.zc;
}

int stage_push_apart(fsim_ctx* ctx) {
    FSIM_TRY(stage_bin(ctx));   // sorted by (cell, id); sink removals compacted
    const int64_t n = ctx->P.n;
    if (n == 0) return FSIM_OK;
    FSIM_PROF(ctx, KID_PUSH_APART);
    ApartArgs A{};
    for (int a = 0; a < 3; a++) {
        A.pos[a] = ctx->P.pos[a];
        A.out[a] = ctx->Palt.pos[a];
        A.lo[a] = (a == 2 ? ctx->g.cdz : ctx->g.cd) + ctx->g.pr * 1.01;
        A.hi[a] = ctx->g.dims[a] - A.lo[a];
    }
    A.two_d = ctx->g.two_d;
    A.zc = ctx->g.dims[2] / 2;
    A.start = ctx->cell_start;
    A.n = n;
    A.nx = ctx->g.nx;
    A.ny = ctx->g.ny;
    A.nz = ctx->g.nz;
    A.D = ctx->g.pr * 2.0;
    A.D2 = A.D * A.D;
    for (int a = 0; a < 3; a++) {
        A.cdinv[a] = a == 2 ? ctx->g.cdzinv : ctx->g.cdinv;
        A.R[a] = (int)std::ceil(A.D * A.cdinv[a]);
    }
    hipLaunchKernelGGL(k_push_apart, dim3(nblocks(n)), dim3(kBlock), 0, ctx->stream, A);
    FSIM_HIP(ctx, hipGetLastError());
    for (int a = 0; a < 3; a++)
        FSIM_HIP(ctx, hipMemcpyAsync(ctx->P.pos[a], ctx->Palt.pos[a], sizeof(float) * (size_t)n,
                                     hipMemcpyDeviceToDevice, ctx->stream));
    return FSIM_OK;
}

// ---- spawning -----------------------------------------------------------------
struct SpawnArgs {
    float* pos[3];
    float* vel[3];
    float* c[9];
    int32_t* pid;
    uint8_t* removed;
    int64_t base;       // first new slot
    int64_t pid_base;   // first new particle id
    int count;
    double opos[3];
    double r;
    double speed;
    uint64_t seed;
    uint64_t ctr_base;
};

__global__ void __launch_bounds__(kBlock) k_spawn(SpawnArgs A) {
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.count) return;
    uint64_t ctr = A.ctr_base + 2ull * (uint64_t)i;
    double theta = u01d(splitmix64(A.seed, ctr)) * (2.0 * M_PI);
    double phi = u01d(splitmix64(A.seed, ctr + 1)) * M_PI;
    double r = A.r;
    double nrm[3] = {r * sin(phi) * cos(theta), r * sin(phi) * sin(theta), r * cos(phi)};
    double inv = 1.0 / sqrt(dot3(nrm, nrm));
    int64_t j = A.base + i;
    for (int a = 0; a < 3; a++) {
        A.pos[a][j] = (float)(A.opos[a] + nrm[a]);
        A.vel[a][j] = (float)(A.speed * (nrm[a] * inv));
    }
    for (int k = 0; k < 9; k++) A.c[k][j] = 0.0f;
    A.pid[j] = (int32_t)(A.pid_base + i);
    A.removed[j] = 0;
}

int stage_spawn(fsim_ctx* ctx, double dt) {
    // counts on the host in FP64, exactly as simulator.cpp:112-114
    std::vector<int> nsp(ctx->obs.size(), 0);
    int64_t total = 0;
    for (size_t o = 0; o < ctx->obs.size(); o++) {
        fsim_obstacle& ob = ctx->obs[o];
        if (ob.kind != FSIM_OBS_SOURCE) continue;
        double num_d = ob.spawn_rate * dt + ob.last_spawn_fraction;
        int num = (int)num_d;
        ob.last_spawn_fraction = num_d - num;
        nsp[o] = num;
        total += num;
    }
    FSIM_TRY(upload_obstacles(ctx));
    FSIM_PROF(ctx, KID_SPAWN);
    if (total > 0) {
        FSIM_TRY(particles_reserve(ctx, ctx->P.n + total));
        int64_t base = ctx->P.n;
        for (size_t o = 0; o < ctx->obs.size(); o++) {
            if (!nsp[o]) continue;
            const fsim_obstacle& ob = ctx->obs[o];
            SpawnArgs A{};
            for (int a = 0; a < 3; a++) {
                A.pos[a] = ctx->P.pos[a];
                A.vel[a] = ctx->P.vel[a];
                A.opos[a] = ob.pos[a];
            }
            for (int k = 0; k < 9; k++) A.c[k] = ctx->P.c[k];
            A.pid = ctx->P.pid;
            A.removed = ctx->removed;
            A.base = base;
            A.pid_base = base;
            A.count = nsp[o];
            A.r = ob.r + ctx->g.pr;
            A.speed = ob.spawn_speed;
            A.seed = ctx->prm.seed ^ kSpawnSalt;
            A.ctr_base = ((ctx->spawn_calls << 6) + (uint64_t)o) << 32;
            hipLaunchKernelGGL(k_spawn, dim3(nblocks(A.count)), dim3(kBlock), 0, ctx->stream, A);
            FSIM_HIP(ctx, hipGetLastError());
            base += nsp[o];
        }
        ctx->P.n = base;
    }
    ctx->spawn_calls++;
    return FSIM_OK;
}

// ---- particle storage -------------------------------------------------------
static int alloc_planes(fsim_ctx* ctx, Particles& P, int64_t cap) {
    float* base = nullptr;
    if (hipMalloc(&base, sizeof(float) * (size_t)cap * 16) != hipSuccess)
        return set_err(ctx, FSIM_ERR_OOM, "particle buffer allocation failed");
    P.base = base;
    P.cap = cap;
    for (int a = 0; a < 3; a++) {
        P.pos[a] = base + (size_t)cap * a;
        P.vel[a] = base + (size_t)cap * (3 + a);
    }
    for (int k = 0; k < 9; k++) P.c[k] = base + (size_t)cap * (6 + k);
    P.pid = reinterpret_cast<int32_t*>(base + (size_t)cap * 15);
    return FSIM_OK;
}

int particles_reserve(fsim_ctx* ctx, int64_t n) {
    if (n <= ctx->P.cap) return FSIM_OK;
    if (n > INT32_MAX) return set_err(ctx, FSIM_ERR_INVALID, "particle count exceeds int32 range");
    const int64_t cap = std::max<int64_t>(n + n / 8, 1024);
    // the id-renumbering scan runs over up to cap particles
    const int64_t need_tmp = (cap + 2) / 4096 + 2;
    // every new buffer is owned here until all of them exist: a failure frees
    // the ones already made and leaves the context as it was
    struct Fresh {
        Particles NP, NA;
        uint32_t* key = nullptr;
        int32_t *rank = nullptr, *perm = nullptr, *tmp = nullptr;
        uint8_t* removed = nullptr;
        bool keep = false;
        ~Fresh() {
            if (keep) return;
            void* ptrs[] = {NP.base, NA.base, key, rank, perm, tmp, removed};
            for (void* p : ptrs)
                if (p) (void)hipFree(p);
        }
    } f;
    FSIM_TRY(alloc_planes(ctx, f.NP, cap));
    FSIM_TRY(alloc_planes(ctx, f.NA, cap));
    if (hipMalloc(&f.key, sizeof(uint32_t) * (cap + 1)) != hipSuccess ||
        hipMalloc(&f.rank, sizeof(int32_t) * cap) != hipSuccess ||
        hipMalloc(&f.perm, sizeof(int32_t) * cap) != hipSuccess || hipMalloc(&f.removed, cap) != hipSuccess ||
        (ctx->scan_tmp_cap < need_tmp && hipMalloc(&f.tmp, sizeof(int32_t) * need_tmp) != hipSuccess))
        return set_err(ctx, FSIM_ERR_OOM, "particle buffer allocation failed");
    if (ctx->P.n > 0) {
        for (int k = 0; k < 16; k++)
            FSIM_HIP(ctx, hipMemcpyAsync(f.NP.base + (size_t)cap * k, ctx->P.base + (size_t)ctx->P.cap * k,
                                         sizeof(float) * ctx->P.n, hipMemcpyDeviceToDevice, ctx->stream));
    }
    FSIM_HIP(ctx, hipMemsetAsync(f.removed, 0, cap, ctx->stream));
    if (ctx->P.n > 0)
        FSIM_HIP(ctx, hipMemcpyAsync(f.removed, ctx->removed, ctx->P.n, hipMemcpyDeviceToDevice, ctx->stream));
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    void* old[] = {ctx->P.base, ctx->Palt.base, ctx->key, ctx->rank, ctx->perm, ctx->removed};
    for (void* p : old)
        if (p) (void)hipFree(p);
    f.NP.n = ctx->P.n;
    ctx->P = f.NP;
    ctx->Palt = f.NA;
    ctx->key = f.key;
    ctx->rank = f.rank;
    ctx->perm = f.perm;
    ctx->removed = f.removed;
    ctx->pwork_cap = cap;
    if (f.tmp) {
        (void)hipFree(ctx->scan_tmp);
        ctx->scan_tmp = f.tmp;
        ctx->scan_tmp_cap = need_tmp;
    }
    f.keep = true;
    return FSIM_OK;
}

// ---- binning: counting sort by grid cell --------------------------------------
// Each wave takes kKeyChunks consecutive 64-particle chunks: the loads of all
// of them are issued first and their run counts go out as atomics together,
// so a wave has several memory round trips in flight instead of one.
constexpr int kKeyChunks = 4;
static unsigned key_blocks(int64_t n) {
    return (unsigned)((n + (int64_t)kBlock * kKeyChunks - 1) / ((int64_t)kBlock * kKeyChunks));
}
__global__ void __launch_bounds__(kBlock) k_bin_keys(const float* __restrict__ px, const float* __restrict__ py,
                                                     const float* __restrict__ pz, const uint8_t* __restrict__ removed,
                                                     uint32_t* __restrict__ key, int32_t* __restrict__ rank,
                                                     int32_t* __restrict__ count, int64_t n, double cdinv,
                                                     double cdzinv, int nx, int ny, int nz, uint32_t nc) {
    const int lane = threadIdx.x & 63;
    const int64_t first = ((int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * (64 * kKeyChunks) + lane;
    float x[kKeyChunks], y[kKeyChunks], z[kKeyChunks];
    uint8_t rm[kKeyChunks];
#pragma unroll
    for (int c = 0; c < kKeyChunks; c++) {
        const int64_t i = first + 64 * c;
        x[c] = y[c] = z[c] = 0.f;
        rm[c] = 0;
        if (i < n) {
            rm[c] = removed[i];
            x[c] = px[i];
            y[c] = py[i];
            z[c] = pz[i];
        }
    }
    uint32_t k[kKeyChunks];
#pragma unroll
    for (int c = 0; c < kKeyChunks; c++) {
        const int64_t i = first + 64 * c;
        k[c] = 0xffffffffu;   // lanes past n: a key of their own, no count
        if (i < n) {
            if (rm[c]) {
                k[c] = nc;
            } else {
                // ivec3(pos * cellDInv), simulator.cpp:358-359, evaluated in FP64
                // positions outside the grid (never produced by the step, whose
                // advection clamps them) are clamped into it so no key leaves [0, Nc)
                const int cx = min(max((int)((double)x[c] * cdinv), 0), nx - 1);
                const int cy = min(max((int)((double)y[c] * cdinv), 0), ny - 1);
                const int cz = min(max((int)((double)z[c] * cdzinv), 0), nz - 1);
                k[c] = (uint32_t)(((int64_t)cx * ny + cy) * nz + cz);
            }
            key[i] = k[c];
        }
    }
    // The particles arrive sorted by cell (the previous bin), so neighbouring
    // lanes share keys: each run of equal keys in a chunk takes one atomicAdd
    // of its length and its lanes number themselves from the returned base.
    // The rank inside a cell is provisional (the segment sort fixes the order).
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    int s[kKeyChunks], base[kKeyChunks];
#pragma unroll
    for (int c = 0; c < kKeyChunks; c++) {
        const uint32_t kprev = (uint32_t)__shfl_up((int)k[c], 1, 64);
        const bool head = lane == 0 || kprev != k[c];
        const uint64_t heads = __builtin_amdgcn_ballot_w64(head);
        s[c] = 63 - __clzll(heads & upto);   // this lane's run head
        const uint64_t later = heads & ~upto;
        const int e = later ? __ffsll((long long)later) - 1 : 64;   // next run head
        base[c] = 0;
        if (head && first + 64 * c < n) base[c] = atomicAdd(&count[k[c]], e - s[c]);
    }
#pragma unroll
    for (int c = 0; c < kKeyChunks; c++) {
        const int b = __shfl(base[c], s[c], 64);
        const int64_t i = first + 64 * c;
        if (i < n) rank[i] = b + (lane - s[c]);
    }
}

// kKeyChunks particles per thread (a wave's chunks of 64 consecutive ones),
// each step's loads issued for all of them before the next step's
__global__ void __launch_bounds__(kBlock) k_bin_scatter(const uint32_t* __restrict__ key, const int32_t* __restrict__ rank,
                                                        const int32_t* __restrict__ start, int32_t* __restrict__ perm,
                                                        int64_t n) {
    const int64_t first =
        ((int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * (64 * kKeyChunks) + (threadIdx.x & 63);
    uint32_t k[kKeyChunks];
    int32_t r[kKeyChunks], st[kKeyChunks];
#pragma unroll
    for (int c = 0; c < kKeyChunks; c++) {
        const int64_t i = first + 64 * c;
        k[c] = 0;
        r[c] = 0;
        if (i < n) {
            k[c] = key[i];
            r[c] = rank[i];
        }
    }
#pragma unroll
    for (int c = 0; c < kKeyChunks; c++) st[c] = first + 64 * c < n ? start[k[c]] : 0;
#pragma unroll
    for (int c = 0; c < kKeyChunks; c++) {
        const int64_t i = first + 64 * c;
        if (i < n) perm[st[c] + r[c]] = (int32_t)i;
    }
}

// Canonical order inside a cell: ascending particle id.  A cell of up to
// kSegRegs particles (a resting fluid holds ~8) is read once into registers,
// sorted by a bitonic network (compile-time indices, padding keys INT_MAX last)
// and written back only if its order changed; larger cells use an insertion
// sort in place.  Particle ids are unique, so both give the same order.
constexpr int kSegRegs = 16;
__global__ void __launch_bounds__(kBlock) k_bin_segsort(const int32_t* __restrict__ start, int32_t* __restrict__ perm,
                                                        const int32_t* __restrict__ pid, int64_t nc) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    int b = start[c], e = start[c + 1];
    if (e - b < 2) return;
    if (e - b <= kSegRegs) {
        const int n = e - b;
        int32_t key[kSegRegs], val[kSegRegs];
#pragma unroll
        for (int i = 0; i < kSegRegs; i++) {
            val[i] = i < n ? perm[b + i] : 0;
            key[i] = i < n ? pid[val[i]] : INT_MAX;
        }
        bool sorted = true;
#pragma unroll
        for (int i = 0; i + 1 < kSegRegs; i++) sorted = sorted && key[i] <= key[i + 1];
        if (sorted) return;
#pragma unroll
        for (int k = 2; k <= kSegRegs; k <<= 1)
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
                for (int i = 0; i < kSegRegs; i++) {
                    const int l = i ^ j;
                    if (l > i) {
                        const bool up = (i & k) == 0;
                        const bool swap = up ? key[i] > key[l] : key[i] < key[l];
                        const int32_t ki = key[i], vi = val[i];
                        key[i] = swap ? key[l] : ki;
                        val[i] = swap ? val[l] : vi;
                        key[l] = swap ? ki : key[l];
                        val[l] = swap ? vi : val[l];
                    }
                }
#pragma unroll
        for (int i = 0; i < kSegRegs; i++)
            if (i < n) perm[b + i] = val[i];
        return;
    }
    for (int i = b + 1; i < e; i++) {
        int32_t v = perm[i];
        int32_t kv = pid[v];
        int j = i - 1;
        while (j >= b && pid[perm[j]] > kv) {
            perm[j + 1] = perm[j];
            j--;
        }
        perm[j + 1] = v;
    }
}

struct GatherArgs {
    const float* src[16];
    float* dst[16];
    int nplanes;
    const int32_t* perm;
    int64_t n;
};

__global__ void __launch_bounds__(kBlock) k_bin_gather(GatherArgs A) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.n) return;
    int32_t s = A.perm[j];
#pragma unroll 4
    for (int k = 0; k < A.nplanes; k++) A.dst[k][j] = A.src[k][s];
}

// pid renumbering after sink removals: a removed particle's slot in the
// reference's vector is erased (hashedParticles.cpp:158-172), so every later
// id shifts down by the number of removed ids before it.
__global__ void __launch_bounds__(kBlock) k_mark_removed_ids(const int32_t* __restrict__ pid, const uint8_t* __restrict__ removed,
                                                             int32_t* __restrict__ flag, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[pid[i]] = removed[i] ? 1 : 0;
}
__global__ void __launch_bounds__(kBlock) k_renumber(int32_t* __restrict__ pid, const int32_t* __restrict__ before, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pid[i] -= before[pid[i]];
}

// ---- exclusive scan of int32 (n values -> n+1 offsets) -------------------------
constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;

__device__ inline int32_t block_exclusive_scan(int32_t v, int32_t* lds_waves, int32_t* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds_waves[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t acc = 0;
        for (int w = 0; w < kBlock / 64; w++) {
            int32_t t = lds_waves[w];
            lds_waves[w] = acc;
            acc += t;
        }
        lds_waves[kBlock / 64] = acc;
    }
    __syncthreads();
    int32_t res = x - v + lds_waves[wave];
    *total = lds_waves[kBlock / 64];
    __syncthreads();
    return res;
}

__global__ void __launch_bounds__(kBlock) k_scan_tiles(const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                       int32_t* __restrict__ tile_sums, int64_t n) {
    __shared__ int32_t lds[kBlock / 64 + 1];
    int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int32_t vals[kScanItems];
    int32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int64_t i = base + k;
        vals[k] = i < n ? in[i] : 0;
        sum += vals[k];
    }
    int32_t total;
    int32_t off = block_exclusive_scan(sum, lds, &total);
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        int64_t i = base + k;
        if (i < n) out[i] = off;
        off += vals[k];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kBlock) k_scan_tile_sums(int32_t* __restrict__ tile_sums, int ntiles,
                                                           int32_t* __restrict__ out_total) {
    __shared__ int32_t lds[kBlock / 64 + 1];
    int32_t carry = 0;
    for (int b = 0; b < ntiles; b += kBlock) {
        int i = b + threadIdx.x;
        int32_t v = i < ntiles ? tile_sums[i] : 0;
        int32_t total;
        int32_t ex = block_exclusive_scan(v, lds, &total);
        if (i < ntiles) tile_sums[i] = ex + carry;
        carry += total;
    }
    if (threadIdx.x == 0) *out_total = carry;
}

__global__ void __launch_bounds__(kBlock) k_scan_add(int32_t* __restrict__ out, const int32_t* __restrict__ tile_sums, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] += tile_sums[i / kScanTile];
}

// out[0..n] = exclusive scan of in[0..n-1]; tmp holds >= ceil(n/4096) ints.
static int exclusive_scan(fsim_ctx* ctx, const int32_t* in, int32_t* out, int32_t* tmp, int64_t n) {
    int ntiles = (int)((n + kScanTile - 1) / kScanTile);
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(kBlock), 0, ctx->stream, in, out, tmp, n);
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(kBlock), 0, ctx->stream, tmp, ntiles, out + n);
    hipLaunchKernelGGL(k_scan_add, dim3(nblocks(n)), dim3(kBlock), 0, ctx->stream, out, tmp, n);
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

int stage_bin(fsim_ctx* ctx) {
    const int64_t nc = ctx->g.nc;
    const int64_t n = ctx->P.n;
    // counts of Nc cells + 1 bucket for removed particles
    FSIM_TRY(fill_i32(ctx, ctx->cell_count, 0, nc + 1));
    if (n > 0) {
        FSIM_PROF(ctx, KID_BIN_KEYS);
        hipLaunchKernelGGL(k_bin_keys, dim3(key_blocks(n)), dim3(kBlock), 0, ctx->stream, ctx->P.pos[0], ctx->P.pos[1],
                           ctx->P.pos[2], ctx->removed, ctx->key, ctx->rank, ctx->cell_count, n, ctx->g.cdinv,
                           ctx->g.cdzinv, ctx->g.nx, ctx->g.ny, ctx->g.nz, (uint32_t)nc);
        FSIM_HIP(ctx, hipGetLastError());
    }
    {
        FSIM_PROF(ctx, KID_BIN_SCAN);
        FSIM_TRY(exclusive_scan(ctx, ctx->cell_count, ctx->cell_start, ctx->scan_tmp, nc + 1));
    }
    if (n == 0) return FSIM_OK;
    {
    FSIM_PROF(ctx, KID_BIN_SCATTER);
    hipLaunchKernelGGL(k_bin_scatter, dim3(key_blocks(n)), dim3(kBlock), 0, ctx->stream, ctx->key, ctx->rank,
                       ctx->cell_start, ctx->perm, n);
    }
    {
    FSIM_PROF(ctx, KID_BIN_SEGSORT);
    hipLaunchKernelGGL(k_bin_segsort, dim3(nblocks(nc)), dim3(kBlock), 0, ctx->stream, ctx->cell_start, ctx->perm,
                       ctx->P.pid, nc);
    }
    FSIM_HIP(ctx, hipGetLastError());
    int64_t n_new = n;
    if (ctx->prm.despawning_enabled) {
        int32_t kept = 0;
        FSIM_HIP(ctx, hipMemcpyAsync(&kept, ctx->cell_start + nc, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        n_new = kept;
        if (n_new < n) {
            // flags over the id space, scanned, then ids shifted down
            int32_t* flag = ctx->rank;  // rank is free after the scatter
            hipLaunchKernelGGL(k_mark_removed_ids, dim3(nblocks(n)), dim3(kBlock), 0, ctx->stream, ctx->P.pid,
                               ctx->removed, flag, n);
            FSIM_TRY(exclusive_scan(ctx, flag, reinterpret_cast<int32_t*>(ctx->key), ctx->scan_tmp, n));
            hipLaunchKernelGGL(k_renumber, dim3(nblocks(n)), dim3(kBlock), 0, ctx->stream, ctx->P.pid,
                               reinterpret_cast<const int32_t*>(ctx->key), n);
            FSIM_HIP(ctx, hipGetLastError());
        }
    }
    GatherArgs G{};
    const bool apic = ctx->prm.transfer == FSIM_APIC;
    int k = 0;
    for (int a = 0; a < 3; a++) { G.src[k] = ctx->P.pos[a]; G.dst[k++] = ctx->Palt.pos[a]; }
    for (int a = 0; a < 3; a++) { G.src[k] = ctx->P.vel[a]; G.dst[k++] = ctx->Palt.vel[a]; }
    if (apic)
        for (int c = 0; c < 9; c++) { G.src[k] = ctx->P.c[c]; G.dst[k++] = ctx->Palt.c[c]; }
    G.src[k] = reinterpret_cast<const float*>(ctx->P.pid);
    G.dst[k++] = reinterpret_cast<float*>(ctx->Palt.pid);
    G.nplanes = k;
    G.perm = ctx->perm;
    G.n = n_new;
    if (n_new > 0) {
        FSIM_PROF(ctx, KID_BIN_GATHER);
        hipLaunchKernelGGL(k_bin_gather, dim3(nblocks(n_new)), dim3(kBlock), 0, ctx->stream, G);
        FSIM_HIP(ctx, hipGetLastError());
    }
    // PIC/FLIP never read or write the APIC matrices, so they are not carried
    // through the gather; they read as zero until APIC mode re-initialises them.
    if (!apic) ctx->c_valid = false;
    std::swap(ctx->P, ctx->Palt);
    ctx->P.n = n_new;
    ctx->Palt.n = 0;
    if (n_new < n || ctx->prm.despawning_enabled)
        FSIM_HIP(ctx, hipMemsetAsync(ctx->removed, 0, ctx->P.cap, ctx->stream));
    return FSIM_OK;
}

}  // namespace fsim

// ---- synthetic scene generator (host only) -------------------------------------
using namespace fsim;

// Voxelisation rule of MacGrid::addObstacle for spheres and boxes
// (macGrid.cpp:235-291), used to keep scene particles out of solid cells.
static bool scene_cell_solid(const fsim_grid_desc* g, int x, int y, int z, int nobs, const fsim_obstacle* obs) {
    const int n[3] = {g->nx, g->ny, g->nz};
    const double inv = 1.0 / g->cell_d;
    const double invd[3] = {inv, inv, 1.0 / (g->two_d ? g->cell_d_z : g->cell_d)};
    const int p[3] = {x, y, z};
    for (int i = 0; i < nobs; i++) {
        const fsim_obstacle& o = obs[i];
        double ctr[3] = {o.pos[0] * invd[0], o.pos[1] * invd[1], o.pos[2] * invd[2]};
        if (o.kind == FSIM_OBS_SPHERE || o.kind == FSIM_OBS_SOURCE) {
            double r = inv * o.r, r2 = r * r;
            bool in = true;
            for (int a = 0; a < 3; a++) {
                int mn = (int)std::fmax(ctr[a] - r, 1.0), mx = (int)std::fmin(ctr[a] + r, n[a] - 2.0);
                if (p[a] < mn || p[a] > mx) in = false;
            }
            double c2o[3] = {x + 0.5 - ctr[0], y + 0.5 - ctr[1], z + 0.5 - ctr[2]};
            if (in && c2o[0] * c2o[0] + c2o[1] * c2o[1] + c2o[2] * c2o[2] < r2) return true;
        } else if (o.kind == FSIM_OBS_BOX) {
            bool in = true;
            for (int a = 0; a < 3; a++) {
                double d = o.size[a] * invd[a] * 0.5;
                int mn = (int)std::fmax(std::round(ctr[a] - d), 1.0), mx = (int)std::fmin(std::round(ctr[a] + d), n[a] - 1.0);
                if (p[a] < mn || p[a] >= mx) in = false;
            }
            if (in) return true;
        }
    }
    return false;
}

extern "C" int fsim_scene_dam_break(const fsim_grid_desc* g, int n_blocks, const double* blk, int n_obstacles,
                                    const fsim_obstacle* obs, uint64_t seed, int64_t capacity, float* pos,
                                    int64_t* n_out) {
    const bool two_d = g && g->two_d;
    if (!g || !n_out || n_blocks < 0 || (n_blocks > 0 && !blk) || g->nx < 4 || g->ny < 4 ||
        (two_d ? g->nz != 3 || !(g->cell_d_z > 0.0) : g->nz < 4))
        return FSIM_ERR_INVALID;
    const double cd = g->cell_d;
    const double cdd[3] = {cd, cd, two_d ? g->cell_d_z : cd};
    const int per = two_d ? 4 : 8;   // 2D: 2 x 2 sub-cells in (x, y), all at the mid-depth z
    int64_t np = 0;
    for (int x = 1; x < g->nx - 1; x++)
        for (int y = 1; y < g->ny - 1; y++)
            for (int z = 1; z < g->nz - 1; z++) {
                const double c[3] = {(x + 0.5) * cdd[0], (y + 0.5) * cdd[1], (z + 0.5) * cdd[2]};
                bool in = false;
                for (int b = 0; b < n_blocks && !in; b++) {
                    const double* lo = blk + 6 * b;
                    const double* hi = lo + 3;
                    in = c[0] >= lo[0] && c[0] < hi[0] && c[1] >= lo[1] && c[1] < hi[1] && c[2] >= lo[2] && c[2] < hi[2];
                }
                if (!in || scene_cell_solid(g, x, y, z, n_obstacles, obs)) continue;
                if (pos) {
                    if (np + per > capacity) return FSIM_ERR_INVALID;
                    const int ci[3] = {x, y, z};
                    for (int s = 0; s < per; s++) {
                        const int sub[3] = {two_d ? (s >> 1) & 1 : (s >> 2) & 1, two_d ? s & 1 : (s >> 1) & 1, s & 1};
                        int64_t id = np + s;
                        for (int a = 0; a < 3; a++) {
                            float u = u01f(splitmix64(seed, 3ull * (uint64_t)id + (uint64_t)a));
                            pos[3 * id + a] = two_d && a == 2 ? (float)(1.5 * cdd[2])
                                                              : (float)((ci[a] + ((double)sub[a] + (double)u) * 0.5) * cdd[a]);
                        }
                    }
                }
                np += per;
            }
    *n_out = np;
    return FSIM_OK;
}
