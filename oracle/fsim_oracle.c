/*
 * fsim_oracle.c — FP64 CPU oracle of the reference's Simulator step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP product
 * (libfsim_hip.so) and the CPU baseline timed beside it; nothing in the
 * product links, loads or calls it.
 *
 * Clean-room restatement of dkiovics/fluid_simulator's step hot path.  Each
 * function cites the reference file:line it follows (paths relative to the
 * reference repository).  The data structures keep the reference's cost
 * structure: AoS FP64 particles (particle.h:8-12), AoS cells with three faces
 * each (macGridCell.h:8-54), atomic double scatter in P2G, OpenMP
 * `parallel for` over x for cell loops, and the serial MIC(0) factor/sweeps.
 *
 * Parity pins (tests/test_oracle_*.py): partition of unity of the tent
 * weights, uniform/affine field reproduction of P2G/G2P, FLIP identity at
 * zero grid change, brute-force voxelisation of obstacles and borders,
 * brute-force histogram for the binning, dense (L L^T)^-1 for MIC(0) apply,
 * a scipy direct solve for PCG, post-projection divergence, closed-form wall
 * bounces and clamp invariants for advection, count/radius/speed invariants
 * for spawning, and hand-worked cases (tests/handworked.py) for the
 * averaging of extrapolateVelocities, the top-row gravity wrap, the density
 * splat and the density term of the RHS, and the sphere (static and moving,
 * with the back-step cap) and box (moving, low-x face) collision responses.
 * Parity unpinned: whole-step trajectories, which compose those stages (the
 * reference cannot be built here: glm is absent from the image).
 */
#include "fsim_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define FO_PI 3.14159265358979323846
#define FO_MAX_OBS 64

/* ---- reference data structures ------------------------------------------ */

/* MacGridCell::Face, macGridCell.h:9-20 (atomic v and weight sum, v2, pos) */
typedef struct { double v, v2, wsum; double pos[3]; } fo_face;
/* MacGridCell, macGridCell.h:22-54 */
typedef struct { fo_face f[3]; double pos[3]; int type; int id; double avg; } fo_cell;
/* Particle, particle.h:8-12 */
typedef struct { double pos[3], v[3], c[3][3]; } fo_particle;

struct fo_sim {
    int n[3];
    int64_t nc;
    double cd, cdinv;          /* cellD.x (== .y): the solver's scales (bridsonSolverGrid.cpp:41,81,296) */
    double cdv[3], cdinvv[3];  /* cellD, cellDInv per axis: z is the depth / 3 in 2D (macGrid.cpp:10) */
    int two_d;                 /* MacGrid::twoD */
    double zc;                 /* 2D: the z particles keep (HashedParticles z = depth / 2) */
    double dims[3];
    double pr;
    fsim_params prm;
    int nthreads;
    fo_cell* cells;
    fo_particle* p;
    int64_t np, cap;
    int64_t* fluid; /* fluidCellPositions as flat indices (macGrid.h:189) */
    int64_t nfluid;
    fsim_obstacle obs[FO_MAX_OBS];
    int nobs;
    /* Bridson solver state (bridsonSolverGrid.h:19-31) */
    double (*arow)[4]; /* nonSolidNeighbours, xWater, yWater, zWater */
    double* precon;
    double* pressure;
    double* rhs;
    uint64_t spawn_calls;
    fsim_step_stats stats;
};

static inline int64_t cidx(const fo_sim* s, int x, int y, int z) {
    return ((int64_t)x * s->n[1] + y) * s->n[2] + z;
}

static int fo_threads(const fo_sim* s) { return s->nthreads > 1 ? s->nthreads : 1; }

/* ---- util/random: SplitMix64 counter RNG (replaces random.h:13-37) ------- */
uint64_t fo_splitmix64(uint64_t seed, uint64_t n) {
    uint64_t z = seed + (n + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static inline double u01d(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

#define FO_SPAWN_SALT 0x5350415741ULL

/* ---- util/interpolation.h:7-35 ------------------------------------------ */
static inline double tri_w(const double c[3], const double p[3], const double inv[3]) {
    double x = 1.0 - fabs((p[0] - c[0]) * inv[0]);
    double y = 1.0 - fabs((p[1] - c[1]) * inv[1]);
    double z = 1.0 - fabs((p[2] - c[2]) * inv[2]);
    return x * y * z;
}
static inline void tri_grad(const double c[3], const double p[3], const double inv[3], double g[3]) {
    double vx = (p[0] - c[0]) * inv[0], vy = (p[1] - c[1]) * inv[1], vz = (p[2] - c[2]) * inv[2];
    double xa = 1.0 - fabs(vx), ya = 1.0 - fabs(vy), za = 1.0 - fabs(vz);
    double xs = vx > 0.0 ? -1.0 : 1.0, ys = vy > 0.0 ? -1.0 : 1.0, zs = vz > 0.0 ? -1.0 : 1.0;
    g[0] = xs * ya * za * inv[0];
    g[1] = ys * xa * za * inv[1];
    g[2] = zs * xa * ya * inv[2];
}

static inline double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* Corner order of getFacesAround / getCellsAround, macGrid.cpp:150-201 */
static const int CORNER[8][3] = {{1, 1, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0},
                                 {1, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 0, 0}};

/* MacGrid::getFacesAround, macGrid.cpp:142-181: base cell of the 8 faces of
 * each axis. */
static inline void faces_base(const fo_sim* s, const double pos[3], int base[3][3]) {
    for (int a = 0; a < 3; a++) {
        for (int d = 0; d < 3; d++) {
            double g = pos[d] * s->cdinvv[d] - (d == a ? 0.0 : 0.5);
            base[a][d] = (int)g;
        }
        base[a][a] -= 1;
    }
}

/* ---- construction: MacGrid ctor + initNewGrid, macGrid.cpp:9-33 --------- */
static void init_cells(fo_sim* s) {
    for (int x = 0; x < s->n[0]; x++)
        for (int y = 0; y < s->n[1]; y++)
            for (int z = 0; z < s->n[2]; z++) {
                fo_cell* c = &s->cells[cidx(s, x, y, z)];
                memset(c, 0, sizeof(*c));
                const double cx = s->cdv[0], cy = s->cdv[1], cz = s->cdv[2];
                c->f[0].pos[0] = (x + 1) * cx; c->f[0].pos[1] = (y + 0.5) * cy; c->f[0].pos[2] = (z + 0.5) * cz;
                c->f[1].pos[0] = (x + 0.5) * cx; c->f[1].pos[1] = (y + 1) * cy; c->f[1].pos[2] = (z + 0.5) * cz;
                c->f[2].pos[0] = (x + 0.5) * cx; c->f[2].pos[1] = (y + 0.5) * cy; c->f[2].pos[2] = (z + 1) * cz;
                c->pos[0] = (x + 0.5) * cx; c->pos[1] = (y + 0.5) * cy; c->pos[2] = (z + 0.5) * cz;
                c->type = FSIM_AIR;
            }
}

fo_sim* fo_create(const fsim_grid_desc* g, const fsim_params* prm) {
    if (!g || !prm || g->nx < 4 || g->ny < 4) return NULL;
    if (g->two_d ? (g->nz != 3 || !(g->cell_d_z > 0.0)) : (g->nz < 4 || (g->cell_d_z != 0.0 && g->cell_d_z != g->cell_d)))
        return NULL;
    fo_sim* s = (fo_sim*)calloc(1, sizeof(fo_sim));
    s->n[0] = g->nx; s->n[1] = g->ny; s->n[2] = g->nz;
    s->nc = (int64_t)g->nx * g->ny * g->nz;
    s->cd = g->cell_d;
    s->cdinv = 1.0 / g->cell_d;
    s->two_d = g->two_d ? 1 : 0;
    for (int a = 0; a < 3; a++) {
        s->cdv[a] = a == 2 && s->two_d ? g->cell_d_z : g->cell_d;
        s->cdinvv[a] = 1.0 / s->cdv[a];
        s->dims[a] = s->n[a] * s->cdv[a];
    }
    s->zc = s->dims[2] / 2;
    s->pr = g->particle_r;
    s->prm = *prm;
    s->nthreads = 1;
    s->cells = (fo_cell*)malloc(sizeof(fo_cell) * s->nc);
    s->fluid = (int64_t*)malloc(sizeof(int64_t) * s->nc);
    s->arow = malloc(sizeof(double[4]) * s->nc);
    s->precon = (double*)calloc(s->nc, sizeof(double));
    s->pressure = (double*)calloc(s->nc, sizeof(double));
    s->rhs = (double*)calloc(s->nc, sizeof(double));
    if (!s->cells || !s->fluid || !s->arow) { fo_destroy(s); return NULL; }
    init_cells(s);
    return s;
}

void fo_destroy(fo_sim* s) {
    if (!s) return;
    free(s->cells); free(s->p); free(s->fluid); free(s->arow);
    free(s->precon); free(s->pressure); free(s->rhs);
    free(s);
}

void fo_set_threads(fo_sim* s, int n) { s->nthreads = n; }
void fo_set_params(fo_sim* s, const fsim_params* prm) { s->prm = *prm; }
void fo_set_particle_radius(fo_sim* s, double r) { s->pr = r; }

static int reserve(fo_sim* s, int64_t n) {
    if (n <= s->cap) return 0;
    int64_t cap = n > 2 * s->cap ? n : 2 * s->cap;
    fo_particle* p = (fo_particle*)realloc(s->p, sizeof(fo_particle) * (size_t)cap);
    if (!p) return -1;
    s->p = p;
    s->cap = cap;
    return 0;
}

int fo_set_particles(fo_sim* s, int64_t n, const double* pos, const double* vel, const double* c) {
    if (reserve(s, n)) return -1;
    for (int64_t i = 0; i < n; i++) {
        fo_particle* q = &s->p[i];
        for (int a = 0; a < 3; a++) {
            q->pos[a] = pos[3 * i + a];
            q->v[a] = vel ? vel[3 * i + a] : 0.0;
            for (int b = 0; b < 3; b++) q->c[a][b] = c ? c[9 * i + 3 * a + b] : 0.0;
        }
    }
    s->np = n;
    return 0;
}

int64_t fo_num_particles(const fo_sim* s) { return s->np; }

void fo_get_particles(const fo_sim* s, double* pos, double* vel, double* c) {
    for (int64_t i = 0; i < s->np; i++) {
        const fo_particle* q = &s->p[i];
        for (int a = 0; a < 3; a++) {
            if (pos) pos[3 * i + a] = q->pos[a];
            if (vel) vel[3 * i + a] = q->v[a];
            if (c) for (int b = 0; b < 3; b++) c[9 * i + 3 * a + b] = q->c[a][b];
        }
    }
}

void fo_set_obstacles(fo_sim* s, int n, const fsim_obstacle* o) {
    s->nobs = n > FO_MAX_OBS ? FO_MAX_OBS : n;
    if (s->nobs) memcpy(s->obs, o, sizeof(fsim_obstacle) * s->nobs);
}
int fo_get_obstacles(const fo_sim* s, fsim_obstacle* out) {
    if (out && s->nobs) memcpy(out, s->obs, sizeof(fsim_obstacle) * s->nobs);
    return s->nobs;
}

static inline int is_sphere(int kind) {
    return kind == FSIM_OBS_SPHERE || kind == FSIM_OBS_SOURCE || kind == FSIM_OBS_SINK;
}

/* ---- Simulator::spawnParticles, simulator.cpp:106-125 ------------------- */
static void fo_spawn(fo_sim* s, double dt) {
    int64_t total = 0;
    int nsp[FO_MAX_OBS];
    for (int o = 0; o < s->nobs; o++) {
        nsp[o] = 0;
        fsim_obstacle* ob = &s->obs[o];
        if (ob->kind != FSIM_OBS_SOURCE) continue;
        double num_d = ob->spawn_rate * dt + ob->last_spawn_fraction;
        int num = (int)num_d;
        ob->last_spawn_fraction = num_d - num;
        nsp[o] = num;
        total += num;
    }
    reserve(s, s->np + total);
    int64_t w = s->np;
    for (int o = 0; o < s->nobs; o++) {
        const fsim_obstacle* ob = &s->obs[o];
        if (ob->kind != FSIM_OBS_SOURCE) continue;
        double r = ob->r + s->pr;
        uint64_t seed = s->prm.seed ^ FO_SPAWN_SALT;
        for (int i = 0; i < nsp[o]; i++) {
            uint64_t ctr = (((s->spawn_calls << 6) + (uint64_t)o) << 32) + 2ull * (uint64_t)i;
            double theta = u01d(fo_splitmix64(seed, ctr)) * (2.0 * FO_PI);
            double phi = u01d(fo_splitmix64(seed, ctr + 1)) * FO_PI;
            double nrm[3] = {r * sin(phi) * cos(theta), r * sin(phi) * sin(theta), r * cos(phi)};
            double inv = 1.0 / sqrt(dot3(nrm, nrm));
            fo_particle* q = &s->p[w++];
            memset(q, 0, sizeof(*q));
            for (int a = 0; a < 3; a++) {
                q->pos[a] = ob->pos[a] + nrm[a];
                q->v[a] = ob->spawn_speed * (nrm[a] * inv);
            }
        }
    }
    s->np = w;
    s->spawn_calls++;
}

/* isParticleInRectangle, simulator.cpp:127-132 */
static inline int in_rect(const double p[3], double pr, const double rp[3], const double rs[3]) {
    return fabs(p[0] - rp[0]) < rs[0] * 0.5 + pr && fabs(p[1] - rp[1]) < rs[1] * 0.5 + pr &&
           fabs(p[2] - rp[2]) < rs[2] * 0.5 + pr;
}

/* HashedParticles::removeParticles, hashedParticles.cpp:158-172 (stable) */
static void remove_marked(fo_sim* s, const unsigned char* rm) {
    int64_t w = 0;
    for (int64_t i = 0; i < s->np; i++) {
        if (rm[i]) continue;
        if (w != i) s->p[w] = s->p[i];
        w++;
    }
    s->np = w;
}

/* ---- Simulator::advectParticles, simulator.cpp:134-251 ------------------ */
static void advect_one(const fo_sim* s, fo_particle* P, double dt, const double lo[3], const double hi[3],
                       unsigned char* removed) {
    const double wall_rest = 0.3, sphere_rest = 1.0, rect_rest = 0.2;
    const double pr = s->pr;
    double t = 0;
    int run = 0;
    const int max_run = 200;
    while (run < max_run) {
        run++;
        double min_t = 1e6;
        int min_axis = 0;
        for (int a = 0; a < 3; a++) {
            double comp = P->v[a];
            double tmp = 1e6;
            if (comp > 1e-6) tmp = (hi[a] - P->pos[a]) / comp;
            else if (comp < -1e-6) tmp = (P->pos[a] - lo[a]) / -comp;
            if (tmp < min_t) { min_t = tmp; min_axis = a; }
        }
        if (min_t <= (dt - t)) {
            for (int a = 0; a < 3; a++) P->pos[a] += P->v[a] * min_t * 0.999;
            P->v[min_axis] *= -wall_rest;
            t += 0.999 * min_t;
            continue;
        }
        int collision = 0;
        for (int a = 0; a < 3; a++) P->pos[a] += P->v[a] * (dt - t);
        for (int oi = 0; oi < s->nobs; oi++) {
            const fsim_obstacle* o = &s->obs[oi];
            if (is_sphere(o->kind)) {
                int sink = o->kind == FSIM_OBS_SINK;
                double r = o->r + pr;
                double d3[3] = {o->pos[0] - P->pos[0], o->pos[1] - P->pos[1], o->pos[2] - P->pos[2]};
                double d = sqrt(dot3(d3, d3)) - r;
                if (d < 0) {
                    if (sink && s->prm.despawning_enabled) {
                        *removed = 1;
                        collision = 0;
                        break;
                    }
                    double bt = 0;
                    for (;;) {
                        double q[3];
                        for (int a = 0; a < 3; a++)
                            q[a] = (P->pos[a] - P->v[a] * bt) - (o->pos[a] - o->speed[a] * bt);
                        if (!(sqrt(dot3(q, q)) < r && bt < t + (double)0.001f)) break;
                        bt += 0.0002;
                    }
                    if (bt > dt - t) continue;
                    bt += 0.0002;
                    double ptmp[3], nrm[3], relv[3];
                    for (int a = 0; a < 3; a++) P->pos[a] -= bt * P->v[a];
                    for (int a = 0; a < 3; a++) ptmp[a] = o->pos[a] - o->speed[a] * bt;
                    for (int a = 0; a < 3; a++) nrm[a] = P->pos[a] - ptmp[a];
                    double inv = 1.0 / sqrt(dot3(nrm, nrm));
                    for (int a = 0; a < 3; a++) nrm[a] *= inv;
                    for (int a = 0; a < 3; a++) relv[a] = P->v[a] - o->speed[a];
                    double ssn = -dot3(nrm, relv);
                    if (ssn <= 0.0f) continue;
                    for (int a = 0; a < 3; a++) {
                        double mirror = -relv[a] / ssn;
                        P->v[a] = (nrm[a] - mirror) * 2.0 + mirror;
                        P->v[a] *= ssn * sphere_rest;
                        P->v[a] += o->speed[a];
                    }
                    t += (dt - t) - bt;
                    collision = 1;
                    break;
                }
            }
            if (o->kind == FSIM_OBS_BOX) {
                if (in_rect(P->pos, pr, o->pos, o->size)) {
                    double bt = 0;
                    for (;;) {
                        double pp[3], op[3];
                        for (int a = 0; a < 3; a++) {
                            pp[a] = P->pos[a] - P->v[a] * bt;
                            op[a] = o->pos[a] - o->speed[a] * bt;
                        }
                        if (!(in_rect(pp, pr, op, o->size) && bt < dt - t)) break;
                        bt += 0.0002;
                    }
                    if (bt > dt - t) continue;
                    bt += 0.0002;
                    double ptmp[3];
                    for (int a = 0; a < 3; a++) P->pos[a] -= bt * P->v[a];
                    for (int a = 0; a < 3; a++) ptmp[a] = o->pos[a] - o->speed[a] * bt;
                    for (int a = 0; a < 3; a++) {
                        if (P->pos[a] >= ptmp[a] + o->size[a] * 0.5 + pr) {
                            P->v[a] = -P->v[a] * rect_rest + o->speed[a];
                            break;
                        } else if (P->pos[a] <= ptmp[a] - o->size[a] * 0.5 - pr) {
                            P->v[a] = -P->v[a] * rect_rest - o->speed[a];
                            break;
                        }
                    }
                    t += (dt - t) - bt;
                    collision = 1;
                    break;
                }
            }
        }
        if (!collision) run = max_run;
    }
    for (int a = 0; a < 3; a++) P->pos[a] = P->pos[a] < lo[a] ? lo[a] : (P->pos[a] > hi[a] ? hi[a] : P->pos[a]);
}

static void fo_advect(fo_sim* s, double dt) {
    double lo[3], hi[3];
    for (int a = 0; a < 3; a++) {   /* gridLow = cellD + (r, r, twoD ? 0 : r) * 1.01, simulator.cpp:140 */
        lo[a] = s->cdv[a] + (a == 2 && s->two_d ? 0.0 : s->pr) * 1.01;
        hi[a] = s->dims[a] - lo[a];
    }
    unsigned char* rm = (unsigned char*)calloc(s->np ? s->np : 1, 1);
    int64_t np = s->np;
#pragma omp parallel for num_threads(fo_threads(s)) schedule(static)
    for (int64_t i = 0; i < np; i++) advect_one(s, &s->p[i], dt, lo, hi, &rm[i]);
    remove_marked(s, rm);
    free(rm);
}

/* MacGrid::getMinMaxRect, macGrid.h:139-145 */
static void min_max_rect(const fo_sim* s, const double pos[3], const double size[3], double lo[3], double hi[3]) {
    for (int a = 0; a < 3; a++) {
        double c = pos[a] * s->cdinvv[a];
        double d = size[a] * s->cdinvv[a] * 0.5;
        int mn = (int)fmax(round(c - d), 1.0);
        int mx = (int)fmin(round(c + d), s->n[a] - 1.0);
        lo[a] = (double)mn * s->cdv[a];
        hi[a] = (double)mx * s->cdv[a];
    }
}

/* ---- Simulator::pushParticlesOutOfObstacles, simulator.cpp:253-312 ------ */
static void fo_push_out(fo_sim* s) {
    const double pr = s->pr;
    double plo[3], phi[3];
    for (int a = 0; a < 3; a++) {
        plo[a] = s->cdv[a] + pr * 1.01;
        phi[a] = s->dims[a] - plo[a];
    }
    int64_t np = s->np;
    for (int oi = 0; oi < s->nobs; oi++) {
        const fsim_obstacle* o = &s->obs[oi];
        if (is_sphere(o->kind)) {
            double r = o->r + pr, r2 = r * r;
#pragma omp parallel for num_threads(fo_threads(s)) schedule(static)
            for (int64_t i = 0; i < np; i++) {
                fo_particle* P = &s->p[i];
                double o2p[3] = {P->pos[0] - o->pos[0], P->pos[1] - o->pos[1], P->pos[2] - o->pos[2]};
                double d2 = dot3(o2p, o2p);
                if (d2 >= r2) continue;
                double d = sqrt(d2);
                for (int a = 0; a < 3; a++) {
                    double v = o->pos[a] + o2p[a] / d * r;
                    P->pos[a] = v < plo[a] ? plo[a] : (v > phi[a] ? phi[a] : v);
                }
            }
        }
        if (o->kind == FSIM_OBS_BOX) {
            double st[3], en[3];
            min_max_rect(s, o->pos, o->size, st, en);
            for (int a = 0; a < 3; a++) { st[a] -= pr; en[a] += pr; }
#pragma omp parallel for num_threads(fo_threads(s)) schedule(static)
            for (int64_t i = 0; i < np; i++) {
                fo_particle* P = &s->p[i];
                if (P->pos[0] > en[0] || P->pos[0] < st[0] || P->pos[1] > en[1] || P->pos[1] < st[1] ||
                    P->pos[2] > en[2] || P->pos[2] < st[2])
                    continue;
                int axis = 0;
                double amount = 1e6;
                for (int a = 0; a < 3; a++) {
                    if (P->pos[a] > o->pos[a]) {
                        double tmp = en[a] - P->pos[a];
                        if (tmp < fabs(amount)) { amount = tmp; axis = a; }
                    } else {
                        double tmp = st[a] - P->pos[a];
                        if (-tmp < fabs(amount)) { amount = tmp; axis = a; }
                    }
                }
                double v = P->pos[axis] + amount;
                P->pos[axis] = v < plo[axis] ? plo[axis] : (v > phi[axis] ? phi[axis] : v);
            }
        }
    }
}

/* ---- HashedParticles::pushParticlesApart, hashedParticles.cpp:64-107 -----
 * For every pair closer than the particle diameter D = 2r (and not closer
 * than 1e-4, :88), each particle moves by half the overlap along the line
 * between them, offset = (p_i - p_j) (D - |p_i - p_j|) / |p_i - p_j| * 0.5
 * (:91-95), and is clamped to [cellD + 1.01 r, L - cellD - 1.01 r] (:96-101).
 * The reference runs this as a parallel loop whose iterations move both
 * particles of a pair while other iterations read them (a data race, so its
 * result depends on the schedule).  This is its deterministic Jacobi form:
 * every particle sums its own offsets from all pairs, computed from the
 * positions before the pass (an isolated pair ends exactly D apart, as in
 * the reference), visiting the other particles in (grid cell, id) order
 * through the cell binning (all pairs within D lie within ceil(D / cellD)
 * cells), then is clamped once. */
static void fo_push_apart(fo_sim* s) {
    const int64_t np = s->np;
    if (np == 0) return;
    const double D = s->pr * 2.0, D2 = D * D;
    double lo[3], hi[3];
    for (int a = 0; a < 3; a++) {   /* particleLow, hashedParticles.cpp:67-68 */
        lo[a] = s->cdv[a] + s->pr * 1.01;
        hi[a] = s->dims[a] - lo[a];
    }
    int32_t* count = (int32_t*)malloc(sizeof(int32_t) * (size_t)s->nc);
    int32_t* start = (int32_t*)malloc(sizeof(int32_t) * (size_t)(s->nc + 1));
    int64_t* perm = (int64_t*)malloc(sizeof(int64_t) * (size_t)np);
    fo_bin(s, count, start, perm);
    int R[3];
    for (int a = 0; a < 3; a++) R[a] = (int)ceil(D * s->cdinvv[a]);
    double* out = (double*)malloc(sizeof(double) * 3 * (size_t)np);
#pragma omp parallel for num_threads(fo_threads(s)) schedule(static)
    for (int64_t i = 0; i < np; i++) {
        const double* pi = s->p[i].pos;
        int c[3];
        for (int a = 0; a < 3; a++) c[a] = (int)(pi[a] * s->cdinvv[a]);
        double q[3] = {pi[0], pi[1], pi[2]};
        for (int x = c[0] - R[0]; x <= c[0] + R[0]; x++) {
            if (x < 0 || x >= s->n[0]) continue;
            for (int y = c[1] - R[1]; y <= c[1] + R[1]; y++) {
                if (y < 0 || y >= s->n[1]) continue;
                for (int z = c[2] - R[2]; z <= c[2] + R[2]; z++) {
                    if (z < 0 || z >= s->n[2]) continue;
                    const int64_t cell = cidx(s, x, y, z);
                    for (int32_t k = start[cell]; k < start[cell + 1]; k++) {
                        const int64_t j = perm[k];
                        if (j == i) continue;
                        const double* pj = s->p[j].pos;
                        const double d[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
                        const double d2 = dot3(d, d);
                        if (d2 > D2 || d2 < 1e-8) continue;
                        const double dist = sqrt(d2);
                        const double tmp = (D - dist) / dist;
                        for (int a = 0; a < 3; a++) q[a] += d[a] * tmp * 0.5;
                    }
                }
            }
        }
        for (int a = 0; a < 3; a++) out[3 * i + a] = q[a] < lo[a] ? lo[a] : (q[a] > hi[a] ? hi[a] : q[a]);
        if (s->two_d) out[3 * i + 2] = s->zc;   /* 2D: z = depth / 2 (hashedParticles.cpp:96-101) */
    }
    for (int64_t i = 0; i < np; i++)
        for (int a = 0; a < 3; a++) s->p[i].pos[a] = out[3 * i + a];
    free(out);
    free(perm);
    free(start);
    free(count);
}

/* ---- stopParticles branch, simulator.cpp:71-74 --------------------------- */
static void fo_stop(fo_sim* s) {
    for (int64_t i = 0; i < s->np; i++) s->p[i].v[0] = s->p[i].v[1] = s->p[i].v[2] = 0.0;
}

/* ---- MacGrid::resetGridValues, macGrid.cpp:221-229 ----------------------- */
static void fo_reset(fo_sim* s) {
#pragma omp parallel for num_threads(fo_threads(s)) schedule(static)
    for (int x = 0; x < s->n[0]; x++)
        for (int y = 0; y < s->n[1]; y++)
            for (int z = 0; z < s->n[2]; z++) {
                fo_cell* c = &s->cells[cidx(s, x, y, z)];
                for (int a = 0; a < 3; a++) c->f[a].v = c->f[a].v2 = c->f[a].wsum = 0.0;
                c->avg = 0.0;
                c->type = FSIM_AIR;
            }
}

/* ---- Simulator::p2gTransfer, simulator.cpp:314-352 ----------------------- */
static void fo_p2g(fo_sim* s) {
    fo_reset(s);
    const int apic = s->prm.transfer == FSIM_APIC;
    int64_t np = s->np;
#pragma omp parallel for num_threads(fo_threads(s)) schedule(static)
    for (int64_t i = 0; i < np; i++) {
        const fo_particle* P = &s->p[i];
        int base[3][3];
        faces_base(s, P->pos, base);
        for (int a = 0; a < 3; a++)
            for (int k = 0; k < 8; k++) {
                fo_face* f = &s->cells[cidx(s, base[a][0] + CORNER[k][0], base[a][1] + CORNER[k][1],
                                            base[a][2] + CORNER[k][2])].f[a];
                double w = tri_w(f->pos, P->pos, s->cdinvv);
                double add;
                if (!apic) {
                    add = P->v[a] * w;
                } else {
                    double p2f[3] = {f->pos[0] - P->pos[0], f->pos[1] - P->pos[1], f->pos[2] - P->pos[2]};
                    add = (P->v[a] + dot3(P->c[a], p2f)) * w;
                }
#pragma omp atomic
                f->v += add;
#pragma omp atomic
                f->wsum += w;
            }
    }
#pragma omp parallel for num_threads(fo_threads(s)) schedule(static)
    for (int x = 0; x < s->n[0]; x++)
        for (int y = 0; y < s->n[1]; y++)
            for (int z = 0; z < s->n[2]; z++) {
                fo_cell* c = &s->cells[cidx(s, x, y, z)];
                for (int a = 0; a < 3; a++) {
                    double w = c->f[a].wsum;
                    c->f[a].v = w > 1e-6 ? c->f[a].v / w : 0.0;
                }
            }
}

/* ---- markFluidCellsAndCalculateParticleDensities, simulator.cpp:354-368 --- */
static void fo_mark(fo_sim* s) {
    int64_t np = s->np;
#pragma omp parallel for num_threads(fo_threads(s)) schedule(static)
    for (int64_t i = 0; i < np; i++) {
        const fo_particle* P = &s->p[i];
        int cx = (int)(P->pos[0] * s->cdinvv[0]), cy = (int)(P->pos[1] * s->cdinvv[1]), cz = (int)(P->pos[2] * s->cdinvv[2]);
        s->cells[cidx(s, cx, cy, cz)].type = FSIM_WATER;
        int b[3];
        for (int d = 0; d < 3; d++) b[d] = (int)(P->pos[d] * s->cdinvv[d] - 0.5);
        for (int k = 0; k < 8; k++) {
            fo_cell* c = &s->cells[cidx(s, b[0] + CORNER[k][0], b[1] + CORNER[k][1], b[2] + CORNER[k][2])];
            double w = tri_w(c->pos, P->pos, s->cdinvv);
#pragma omp atomic
            c->avg += w;
        }
    }
}

/* ---- MacGrid::addObstacle, macGrid.cpp:231-292 (serial, order-dependent) -- */
static void solidify(fo_sim* s, int x, int y, int z, const double speed[3]) {
    fo_cell* c = &s->cells[cidx(s, x, y, z)];
    c->type = FSIM_SOLID;
    const int off[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int a = 0; a < 3; a++) {
        fo_cell* np_ = &s->cells[cidx(s, x + off[a][0], y + off[a][1], z + off[a][2])];
        fo_cell* nm = &s->cells[cidx(s, x - off[a][0], y - off[a][1], z - off[a][2])];
        if (np_->type == FSIM_WATER) c->f[a].v = speed[a];
        if (nm->type == FSIM_WATER) nm->f[a].v = speed[a];
    }
}

static void fo_add_obstacle(fo_sim* s, const fsim_obstacle* o) {
    double ctr[3];
    for (int a = 0; a < 3; a++) ctr[a] = o->pos[a] * s->cdinvv[a];
    if ((o->kind == FSIM_OBS_SPHERE || o->kind == FSIM_OBS_SOURCE)) {
        double r = s->cdinv * o->r, r2 = r * r;   /* cellDInv.x * r */
        int mn[3], mx[3];
        for (int a = 0; a < 3; a++) {
            mn[a] = (int)fmax(ctr[a] - r, 1.0);
            mx[a] = (int)fmin(ctr[a] + r, s->n[a] - 2.0);
        }
        for (int x = mn[0]; x <= mx[0]; x++)
            for (int y = mn[1]; y <= mx[1]; y++)
                for (int z = mn[2]; z <= mx[2]; z++) {
                    double c2o[3] = {x + 0.5 - ctr[0], y + 0.5 - ctr[1], z + 0.5 - ctr[2]};
                    if (dot3(c2o, c2o) < r2) solidify(s, x, y, z, o->speed);
                }
    }
    if (o->kind == FSIM_OBS_BOX) {
        int mn[3], mx[3];
        for (int a = 0; a < 3; a++) {
            double d = o->size[a] * s->cdinvv[a] * 0.5;
            mn[a] = (int)fmax(round(ctr[a] - d), 1.0);
            mx[a] = (int)fmin(round(ctr[a] + d), s->n[a] - 1.0);
        }
        for (int x = mn[0]; x < mx[0]; x++)
            for (int y = mn[1]; y < mx[1]; y++)
                for (int z = mn[2]; z < mx[2]; z++) solidify(s, x, y, z, o->speed);
    }
}

static void fo_obstacles(fo_sim* s) {
    for (int i = 0; i < s->nobs; i++) fo_add_obstacle(s, &s->obs[i]);
}

/* ---- MacGrid::restoreBorderingSolidCellsAndSpeeds, macGrid.cpp:75-140 ---- */
static void fo_borders(fo_sim* s) {
    const int X = s->n[0], Y = s->n[1], Z = s->n[2];
#pragma omp parallel for num_threads(fo_threads(s)) schedule(static)
    for (int y = 0; y < Y; y++)
        for (int z = 0; z < Z; z++) {
            s->cells[cidx(s, 0, y, z)].type = s->cells[cidx(s, X - 1, y, z)].type = FSIM_SOLID;
            if (s->cells[cidx(s, 1, y, z)].type == FSIM_WATER) s->cells[cidx(s, 0, y, z)].f[0].v = 0;
            if (s->cells[cidx(s, X - 2, y, z)].type == FSIM_WATER) s->cells[cidx(s, X - 2, y, z)].f[0].v = 0;
        }
#pragma omp parallel for num_threads(fo_threads(s)) schedule(static)
    for (int x = 0; x < X; x++) {
        for (int z = 0; z < Z; z++) {
            s->cells[cidx(s, x, 0, z)].type = FSIM_SOLID;
            if (s->cells[cidx(s, x, 1, z)].type == FSIM_WATER) s->cells[cidx(s, x, 0, z)].f[1].v = 0;
            if (s->prm.top_solid) {
                s->cells[cidx(s, x, Y - 1, z)].type = FSIM_SOLID;
                if (s->cells[cidx(s, x, Y - 2, z)].type == FSIM_WATER) s->cells[cidx(s, x, Y - 2, z)].f[1].v = 0;
            }
        }
        for (int y = 0; y < Y; y++) {
            s->cells[cidx(s, x, y, 0)].type = s->cells[cidx(s, x, y, Z - 1)].type = FSIM_SOLID;
            if (s->cells[cidx(s, x, y, 1)].type == FSIM_WATER) s->cells[cidx(s, x, y, 0)].f[2].v = 0;
            if (s->cells[cidx(s, x, y, Z - 2)].type == FSIM_WATER) s->cells[cidx(s, x, y, Z - 2)].f[2].v = 0;
        }
    }
}

/* fluidCellPositions rebuild, macGrid.cpp:212-218 (serial, lexicographic) */
static void build_fluid_list(fo_sim* s) {
    s->nfluid = 0;
    for (int x = 1; x < s->n[0] - 1; x++)
        for (int y = 1; y < s->n[1] - 1; y++)
            for (int z = 1; z < s->n[2] - 1; z++) {
                fo_cell* c = &s->cells[cidx(s, x, y, z)];
                if (c->type == FSIM_WATER) {
                    s->fluid[s->nfluid] = cidx(s, x, y, z);
                    c->id = (int)s->nfluid;
                    s->nfluid++;
                }
            }
}

/* ---- MacGrid::postP2GUpdate, macGrid.cpp:204-219 ------------------------- */
static void fo_post_p2g(fo_sim* s, double grav_inc) {
    const int64_t nc = s->nc, stride_y = s->n[2];
#pragma omp parallel for num_threads(fo_threads(s)) schedule(static)
    for (int x = 0; x < s->n[0]; x++)
        for (int y = 0; y < s->n[1]; y++)
            for (int z = 0; z < s->n[2]; z++) {
                int64_t i = cidx(s, x, y, z);
                fo_cell* c = &s->cells[i];
                for (int a = 0; a < 3; a++) c->f[a].v2 = c->f[a].v;
                /* cell<1,1>(pos) is the flat index +Nz: for the top row it is
                 * cell(x+1, 0, z), a SOLID border cell (no gravity there).  An
                 * index past the end (x = Nx-1, y = Ny-1) is only reached when
                 * c itself is not SOLID, which the border pass rules out; it is
                 * treated as SOLID here. */
                int64_t nb = i + stride_y;
                if (c->type != FSIM_SOLID && nb < nc && s->cells[nb].type != FSIM_SOLID) c->f[1].v2 += grav_inc;
            }
    build_fluid_list(s);
}

/* ---- BridsonSolverGrid, bridsonSolverGrid.cpp --------------------------- */
static inline const fo_cell* nbc(const fo_sim* s, int64_t i, int a, int sgn) {
    int64_t st = a == 0 ? (int64_t)s->n[1] * s->n[2] : (a == 1 ? s->n[2] : 1);
    return &s->cells[i + sgn * st];
}

/* calculateAMatrix, bridsonSolverGrid.cpp:40-77 */
static void calc_a(fo_sim* s, double dt) {
    const double scale = dt / (s->prm.fluid_density * s->cd * s->cd);
    int64_t nf = s->nfluid;
#pragma omp parallel for num_threads(fo_threads(s)) schedule(static)
    for (int64_t p = 0; p < nf; p++) {
        double* row = s->arow[p];
        row[0] = row[1] = row[2] = row[3] = 0;
        int64_t i = s->fluid[p];
        for (int a = 0; a < 3; a++) {
            int tp = nbc(s, i, a, +1)->type;
            if (tp == FSIM_WATER) { row[0] += scale; row[1 + a] = -scale; }
            else if (tp == FSIM_AIR) row[0] += scale;
            if (nbc(s, i, a, -1)->type != FSIM_SOLID) row[0] += scale;
        }
    }
}

/* calculateRHS, bridsonSolverGrid.cpp:79-89 */
static void calc_rhs(fo_sim* s, double* r) {
    const double scale = 1.0 / s->cd;
    int64_t nf = s->nfluid;
#pragma omp parallel for num_threads(fo_threads(s)) schedule(static)
    for (int64_t p = 0; p < nf; p++) {
        int64_t i = s->fluid[p];
        const fo_cell* c = &s->cells[i];
        r[p] = -scale * (c->f[0].v2 + c->f[1].v2 + c->f[2].v2 - nbc(s, i, 0, -1)->f[0].v2 - nbc(s, i, 1, -1)->f[1].v2 -
                         nbc(s, i, 2, -1)->f[2].v2) +
               (s->prm.pressure_enabled ? (c->avg - s->prm.average_pressure) * s->prm.pressure_k : 0.0);
    }
}

/* calculatePreconditioner (MIC(0), tau 0.97, sigma 0.25), bridsonSolverGrid.cpp:91-124, .h:26-27 */
static void calc_precon(fo_sim* s) {
    const double tau = 0.97, sigma = 0.25;
    for (int64_t p = 0; p < s->nfluid; p++) {
        int64_t i = s->fluid[p];
        double eneg = 0, eneg_tau = 0;
        for (int a = 0; a < 3; a++) {
            const fo_cell* n = nbc(s, i, a, -1);
            if (n->type == FSIM_WATER) {
                const double* A = s->arow[n->id];
                double ap = A[1 + a] * s->precon[n->id];
                eneg += ap * ap;
                eneg_tau += ap * (A[1 + (a + 1) % 3] + A[1 + (a + 2) % 3]) * s->precon[n->id];
            }
        }
        double diag = s->arow[p][0];
        double e = diag - eneg - eneg_tau * tau;
        if (e < sigma * diag) e = diag < 1e-6 ? 1.0 : diag;
        s->precon[p] = 1.0 / sqrt(e);
    }
}

/* The reference sums the two "other" off-diagonals as (y + z), (x + z),
 * (x + y) for the -x, -y, -z neighbour.  (a+1)%3,(a+2)%3 gives (y,z),(z,x),
 * (x,y); the -y case differs only in the order of an exact sum of two values
 * that are each 0 or -scale, so the result is identical. */

/* applyPreconditioner, bridsonSolverGrid.cpp:126-163 */
static void apply_precon(const fo_sim* s, const double* r, double* q, double* z) {
    for (int64_t p = 0; p < s->nfluid; p++) {
        int64_t i = s->fluid[p];
        double qneg = 0;
        for (int a = 0; a < 3; a++) {
            const fo_cell* n = nbc(s, i, a, -1);
            if (n->type == FSIM_WATER) qneg += s->arow[n->id][1 + a] * q[n->id] * s->precon[n->id];
        }
        q[p] = (r[p] - qneg) * s->precon[p];
    }
    for (int64_t p = s->nfluid - 1; p >= 0; p--) {
        int64_t i = s->fluid[p];
        double tneg = 0;
        for (int a = 0; a < 3; a++) {
            const fo_cell* n = nbc(s, i, a, +1);
            if (n->type == FSIM_WATER) tneg += s->arow[p][1 + a] * z[n->id];
        }
        z[p] = (q[p] - tneg * s->precon[p]) * s->precon[p];
    }
}

/* applyAMatrix, bridsonSolverGrid.cpp:165-198 */
static void apply_a(const fo_sim* s, const double* v, double* out) {
    int64_t nf = s->nfluid;
#pragma omp parallel for num_threads(fo_threads(s)) schedule(static)
    for (int64_t p = 0; p < nf; p++) {
        int64_t i = s->fluid[p];
        const double* A = s->arow[p];
        double val = A[0] * v[p];
        for (int a = 0; a < 3; a++) {
            const fo_cell* n = nbc(s, i, a, +1);
            if (n->type == FSIM_WATER) val += A[1 + a] * v[n->id];
        }
        for (int a = 0; a < 3; a++) {
            const fo_cell* n = nbc(s, i, a, -1);
            if (n->type == FSIM_WATER) val += s->arow[n->id][1 + a] * v[n->id];
        }
        out[p] = val;
    }
}

/* dotProduct / multAdd / multSelfAndAdd, bridsonSolverGrid.cpp:200-242 */
static double dotv(const fo_sim* s, const double* a, const double* b, int64_t n) {
    double r = 0.0;
#pragma omp parallel for num_threads(fo_threads(s)) reduction(+ : r) schedule(static)
    for (int64_t i = 0; i < n; i++) r += a[i] * b[i];
    return r;
}

/* applyPressureToVelocities, bridsonSolverGrid.cpp:295-327 */
static void apply_pressure(fo_sim* s, double dt, const double* pres) {
    const double scale = dt / (s->prm.fluid_density * s->cd);
    int64_t nf = s->nfluid;
#pragma omp parallel for num_threads(fo_threads(s)) schedule(static)
    for (int64_t p = 0; p < nf; p++) {
        int64_t i = s->fluid[p];
        fo_cell* c = &s->cells[i];
        for (int a = 0; a < 3; a++) {
            const fo_cell* n = nbc(s, i, a, +1);
            if (n->type != FSIM_SOLID) {
                if (n->type == FSIM_AIR) c->f[a].v2 += scale * pres[p];
                else c->f[a].v2 += scale * (pres[p] - pres[n->id]);
            }
        }
        for (int a = 0; a < 3; a++) {
            fo_cell* n = (fo_cell*)nbc(s, i, a, -1);
            if (n->type == FSIM_AIR) n->f[a].v2 -= scale * pres[p];
        }
    }
}

/* FP64 true residual b - A p for the stats (no reference counterpart). */
static void true_residual(fo_sim* s, const double* b, const double* pres, double* tmp, double* abs_inf, double* l2) {
    apply_a(s, pres, tmp);
    double mx = 0, ss = 0;
    for (int64_t i = 0; i < s->nfluid; i++) {
        double r = b[i] - tmp[i];
        ss += r * r;
        if (fabs(r) > mx) mx = fabs(r);
    }
    *abs_inf = mx;
    *l2 = sqrt(ss);
}

/* BridsonSolverGrid::solveIncompressibility, bridsonSolverGrid.cpp:244-293.
 * stop_rule ABS_INF is the reference's max|r| < tol; REL_L2 (extension)
 * tests ||r||_2 <= tol * ||b||_2 at the same point of the loop. */
/* ---- BasicMacGrid::solveIncompressibility, basicMacGrid.cpp:5-102 --------
 * The release build's parallel branch (paralellDefine.h:5-7): per iteration
 * one pass over the cells with x + y + z odd (z from 1 + (x+y)%2, :20), then
 * one over the even ones (z from 2 - (x+y)%2, :62).  A WATER cell with s > 0
 * non-SOLID neighbours moves its 6 faces by d = (-u - v - w + u(-x) + v(-y) +
 * w(-z) [+ (avgPNum - averagePressure) pressureK]) * 1.98 / s, each face whose
 * other cell is not SOLID (:31-55).  Cells of one colour share no face, so a
 * pass's result does not depend on its order. */
static void sor_pass(fo_sim* s, int odd) {
    const double omega = 1.98; /* overRelaxation, :13 */
    const double term_k = s->prm.pressure_k, avgp = s->prm.average_pressure;
    const int pe = s->prm.pressure_enabled;
#pragma omp parallel for num_threads(fo_threads(s)) schedule(static)
    for (int x = 1; x < s->n[0] - 1; x++)
        for (int y = 1; y < s->n[1] - 1; y++)
            for (int z = odd ? 1 + ((x + y) % 2) : 2 - ((x + y) % 2); z < s->n[2] - 1; z += 2) {
                const int64_t i = cidx(s, x, y, z);
                fo_cell* c = &s->cells[i];
                if (c->type != FSIM_WATER) continue;
                fo_cell* xn = (fo_cell*)nbc(s, i, 0, -1);
                fo_cell* yn = (fo_cell*)nbc(s, i, 1, -1);
                fo_cell* zn = (fo_cell*)nbc(s, i, 2, -1);
                const int s1 = nbc(s, i, 2, +1)->type != FSIM_SOLID, s2 = zn->type != FSIM_SOLID;
                const int s3 = nbc(s, i, 1, +1)->type != FSIM_SOLID, s4 = yn->type != FSIM_SOLID;
                const int s5 = nbc(s, i, 0, +1)->type != FSIM_SOLID, s6 = xn->type != FSIM_SOLID;
                const int sum = s1 + s2 + s3 + s4 + s5 + s6;
                if (sum == 0) continue;
                double d = -c->f[0].v2 - c->f[1].v2 - c->f[2].v2 + xn->f[0].v2 + yn->f[1].v2 + zn->f[2].v2 +
                           (pe ? (c->avg - avgp) * term_k : 0.0);
                d = d * omega / sum;
                if (s1) c->f[2].v2 += d;
                if (s2) zn->f[2].v2 -= d;
                if (s3) c->f[1].v2 += d;
                if (s4) yn->f[1].v2 -= d;
                if (s5) c->f[0].v2 += d;
                if (s6) xn->f[0].v2 -= d;
            }
}

static double rhs_norm(fo_sim* s, double* r) {
    calc_rhs(s, r);
    double ss = 0;
    for (int64_t p = 0; p < s->nfluid; p++) ss += r[p] * r[p];
    return sqrt(ss);
}

static int fo_basic_sor(fo_sim* s) {
    const int64_t nf = s->nfluid;
    double* r = (double*)malloc((size_t)(nf ? nf : 1) * sizeof(double));
    const double b0 = rhs_norm(s, r);
    memset(s->pressure, 0, sizeof(double) * s->nc);
    const int iters = s->prm.max_iterations;
    for (int n = 0; n < iters; n++) {
        sor_pass(s, 1);
        sor_pass(s, 0);
    }
    /* the right-hand side left on the projected faces */
    const double b1 = rhs_norm(s, r);
    memset(s->rhs, 0, sizeof(double) * s->nc);
    for (int64_t p = 0; p < nf; p++) s->rhs[s->fluid[p]] = r[p];
    free(r);
    s->stats.early_exit = 0;
    s->stats.solve_nan = 0;
    s->stats.b_norm_l2 = b0;
    s->stats.residual_rel_l2 = b0 > 0 ? b1 / b0 : 0.0; /* the right-hand side left over, relative */
    s->stats.residual_abs_inf = 0;
    s->stats.pcg_iterations = iters; /* the reference returns the iteration count, :10 */
    s->stats.num_fluid_cells = nf;
    return iters;
}

static int fo_solve(fo_sim* s, double dt) {
    if (s->prm.precond == FSIM_SOLVER_BASIC_SOR) return fo_basic_sor(s);
    const int64_t nf = s->nfluid;
    size_t sz = (size_t)(nf ? nf : 1) * sizeof(double);
    double* pres = (double*)calloc(1, sz);
    double* z = (double*)calloc(1, sz);
    double* q = (double*)calloc(1, sz);
    double* r = (double*)malloc(sz);
    double* sv = (double*)malloc(sz);
    double* b = (double*)malloc(sz);
    calc_rhs(s, r);
    memcpy(b, r, sz);
    memset(s->rhs, 0, sizeof(double) * s->nc);
    memset(s->pressure, 0, sizeof(double) * s->nc);
    for (int64_t p = 0; p < nf; p++) s->rhs[s->fluid[p]] = r[p];

    double total = 0;
    for (int64_t i = 0; i < nf; i++) total += r[i] * r[i];
    s->stats.b_norm_l2 = sqrt(total);
    s->stats.early_exit = 0;
    s->stats.solve_nan = 0;
    s->stats.residual_abs_inf = 0;
    s->stats.residual_rel_l2 = 0;
    int it = 0;
    if (total < 1e-7) {
        s->stats.early_exit = 1;
        s->stats.pcg_iterations = 0;
        {
            double mx = 0;
            for (int64_t i = 0; i < nf; i++) if (fabs(r[i]) > mx) mx = fabs(r[i]);
            s->stats.residual_abs_inf = mx;
            s->stats.residual_rel_l2 = 1.0;
        }
        goto done;
    }
    calc_a(s, dt);
    calc_precon(s);
    apply_precon(s, r, q, z);
    memcpy(sv, z, sz);
    double sigma = dotv(s, z, r, nf);
    const double bnorm = sqrt(total);
    for (; it < s->prm.max_iterations; it++) {
        apply_a(s, sv, z);
        double alpha = sigma / dotv(s, sv, z, nf);
        if (alpha != alpha) { s->stats.solve_nan = 1; break; }
        for (int64_t i = 0; i < nf; i++) pres[i] += sv[i] * alpha;
        for (int64_t i = 0; i < nf; i++) r[i] += z[i] * -alpha;
        if (s->prm.stop_rule == FSIM_STOP_REL_L2) {
            double ss = 0;
            for (int64_t i = 0; i < nf; i++) ss += r[i] * r[i];
            if (sqrt(ss) <= s->prm.residual_tolerance * bnorm) break;
        } else {
            double mx = 0.0;
            for (int64_t i = 0; i < nf; i++) if (fabs(r[i]) > mx) mx = fabs(r[i]);
            if (mx < s->prm.residual_tolerance) break;
        }
        apply_precon(s, r, q, z);
        double sigma_new = dotv(s, z, r, nf);
        if (sigma != sigma) { s->stats.solve_nan = 1; break; }
        double beta = sigma_new / sigma;
        for (int64_t i = 0; i < nf; i++) sv[i] = sv[i] * beta + z[i];
        sigma = sigma_new;
    }
    apply_pressure(s, dt, pres);
    {
        double ai, l2;
        true_residual(s, b, pres, q, &ai, &l2);
        s->stats.residual_abs_inf = ai;
        s->stats.residual_rel_l2 = bnorm > 0 ? l2 / bnorm : 0.0;
    }
    s->stats.pcg_iterations = it;
    for (int64_t p = 0; p < nf; p++) s->pressure[s->fluid[p]] = pres[p];
done:
    s->stats.num_fluid_cells = nf;
    free(pres); free(z); free(q); free(r); free(sv); free(b);
    return it;
}

/* ---- MacGrid::extrapolateVelocities, macGrid.cpp:294-336 ----------------- */
static void fo_extrapolate(fo_sim* s) {
    const int X = s->n[0], Y = s->n[1], Z = s->n[2];
    unsigned char* valid = (unsigned char*)malloc((size_t)s->nc);
    memset(valid, 100, (size_t)s->nc);
    for (int64_t p = 0; p < s->nfluid; p++) valid[s->fluid[p]] = 0;
    for (int it = 0; it < 2; it++) {
#pragma omp parallel for num_threads(fo_threads(s)) schedule(static)
        for (int x = 0; x < X; x++)
            for (int y = 0; y < Y; y++)
                for (int z = 0; z < Z; z++) {
                    int64_t i = cidx(s, x, y, z);
                    if (valid[i] <= it) continue;
                    int cnt = 0;
                    double vs[3] = {0, 0, 0};
                    int pos[3] = {x, y, z};
                    for (int a = 0; a < 3; a++)
                        for (int off = 0; off < 2; off++) {
                            int q[3] = {pos[0], pos[1], pos[2]};
                            q[a] += off * 2 - 1;
                            if (q[0] < 0 || q[0] >= X || q[1] < 0 || q[1] >= Y || q[2] < 0 || q[2] >= Z) continue;
                            int64_t j = cidx(s, q[0], q[1], q[2]);
                            if (valid[j] > it) continue;
                            vs[0] += s->cells[j].f[0].v2;
                            vs[1] += s->cells[j].f[1].v2;
                            vs[2] += s->cells[j].f[2].v2;
                            cnt++;
                        }
                    if (cnt > 0) {
                        fo_cell* c = &s->cells[i];
                        if (x + 1 < X && s->cells[cidx(s, x + 1, y, z)].type != FSIM_WATER) c->f[0].v2 = vs[0] / cnt;
                        if (y + 1 < Y && s->cells[cidx(s, x, y + 1, z)].type != FSIM_WATER) c->f[1].v2 = vs[1] / cnt;
                        if (z + 1 < Z && s->cells[cidx(s, x, y, z + 1)].type != FSIM_WATER) c->f[2].v2 = vs[2] / cnt;
                        valid[i] = (unsigned char)(it + 1);
                    }
                }
    }
    free(valid);
}

/* ---- Simulator::g2pTransfer, simulator.cpp:375-418 ----------------------- */
static void fo_g2p(fo_sim* s) {
    const int tt = s->prm.transfer;
    const float fr = s->prm.flip_ratio;
    int64_t np = s->np;
#pragma omp parallel for num_threads(fo_threads(s)) schedule(static)
    for (int64_t i = 0; i < np; i++) {
        fo_particle* P = &s->p[i];
        int base[3][3];
        faces_base(s, P->pos, base);
        for (int a = 0; a < 3; a++) {
            if (a == 2 && s->two_d) {   /* 2D: v.z = 0, c[2] kept (simulator.cpp:382-385) */
                P->v[2] = 0.0;
                break;
            }
            double pic = 0, flip = 0, cv[3] = {0, 0, 0};
            for (int k = 0; k < 8; k++) {
                const fo_face* f = &s->cells[cidx(s, base[a][0] + CORNER[k][0], base[a][1] + CORNER[k][1],
                                                  base[a][2] + CORNER[k][2])].f[a];
                double w = tri_w(f->pos, P->pos, s->cdinvv);
                pic += f->v2 * w;
                if (tt == FSIM_FLIP) flip += (f->v2 - f->v) * w;
                if (tt == FSIM_APIC) {
                    double g[3];
                    tri_grad(f->pos, P->pos, s->cdinvv, g);
                    for (int d = 0; d < 3; d++) cv[d] += g[d] * f->v2;
                }
            }
            if (tt == FSIM_PIC) P->v[a] = pic;
            else if (tt == FSIM_FLIP) P->v[a] = pic * (double)(1 - fr) + (flip + P->v[a]) * (double)fr;
            else {
                P->v[a] = pic;
                for (int d = 0; d < 3; d++) P->c[a][d] = cv[d];
            }
        }
    }
}

/* Serial stable counting sort by grid cell (the binning the GPU path uses in
 * place of the reference's 2r intersection hash, hashedParticles.cpp:195-244). */
void fo_bin(const fo_sim* s, int32_t* count, int32_t* start, int64_t* perm) {
    memset(count, 0, sizeof(int32_t) * (size_t)s->nc);
    int64_t* key = (int64_t*)malloc(sizeof(int64_t) * (size_t)(s->np ? s->np : 1));
    for (int64_t i = 0; i < s->np; i++) {
        const double* p = s->p[i].pos;
        key[i] = cidx(s, (int)(p[0] * s->cdinvv[0]), (int)(p[1] * s->cdinvv[1]), (int)(p[2] * s->cdinvv[2]));
        count[key[i]]++;
    }
    int64_t acc = 0;
    for (int64_t c = 0; c < s->nc; c++) { start[c] = (int32_t)acc; acc += count[c]; }
    start[s->nc] = (int32_t)acc;
    int32_t* fill = (int32_t*)calloc((size_t)s->nc, sizeof(int32_t));
    for (int64_t i = 0; i < s->np; i++) perm[start[key[i]] + fill[key[i]]++] = i;
    free(fill);
    free(key);
}

/* ---- SimulationManager's gfx snapshot, simulationManager.cpp:223-231 ---- */
void fo_gfx_snapshot(const fo_sim* s, double* out) {
    for (int64_t i = 0; i < s->np; i++) {
        const fo_particle* P = &s->p[i];
        /* getCellsAround, macGrid.cpp:188-202 */
        int base[3];
        for (int d = 0; d < 3; d++) base[d] = (int)(P->pos[d] * s->cdinvv[d] - 0.5);
        float density = 0; /* the reference's float accumulator (:226) */
        for (int k = 0; k < 8; k++) {
            const int x = base[0] + CORNER[k][0], y = base[1] + CORNER[k][1], z = base[2] + CORNER[k][2];
            const fo_cell* c = &s->cells[cidx(s, x, y, z)];
            density += tri_w(P->pos, c->pos, s->cdinvv) * c->avg; /* trilinearInterpoll(p.pos, c.cell.pos) */
        }
        out[5 * i + 0] = (float)P->pos[0];
        out[5 * i + 1] = (float)P->pos[1];
        out[5 * i + 2] = (float)P->pos[2];
        out[5 * i + 3] = (float)sqrt(dot3(P->v, P->v)); /* glm::length(p.v) */
        out[5 * i + 4] = density;
    }
}

int fo_run_stage(fo_sim* s, int stage, double dt) {
    switch (stage) {
    case FSIM_STAGE_SPAWN: fo_spawn(s, dt); return 0;
    case FSIM_STAGE_ADVECT: fo_advect(s, dt); return 0;
    case FSIM_STAGE_PUSH_OUT: fo_push_out(s); return 0;
    case FSIM_STAGE_STOP: fo_stop(s); return 0;
    case FSIM_STAGE_BIN: return 0; /* the oracle keeps reference (unsorted) order */
    case FSIM_STAGE_P2G: fo_p2g(s); return 0;
    case FSIM_STAGE_MARK: fo_mark(s); return 0;
    case FSIM_STAGE_OBSTACLES: fo_obstacles(s); return 0;
    case FSIM_STAGE_BORDERS: fo_borders(s); return 0;
    case FSIM_STAGE_POST_P2G:
        fo_post_p2g(s, s->prm.gravity_enabled ? (double)s->prm.gravity * dt : 0.0);
        return 0;
    case FSIM_STAGE_SOLVE: fo_solve(s, dt); return 0;
    case FSIM_STAGE_EXTRAPOLATE: fo_extrapolate(s); return 0;
    case FSIM_STAGE_G2P: fo_g2p(s); return 0;
    case FSIM_STAGE_PUSH_APART: fo_push_apart(s); return 0;
    default: return -1;
    }
}

/* ---- Simulator::simulate, simulator.cpp:51-100 --------------------------- */
void fo_step(fo_sim* s, double dt, fsim_step_stats* out) {
    double t0 = omp_get_wtime(), t;
    memset(s->stats.timer_ms, 0, sizeof(s->stats.timer_ms));
    t = omp_get_wtime();
    if (s->prm.spawning_enabled) fo_spawn(s, dt);
    fo_advect(s, dt);
    s->stats.timer_ms[FSIM_T_SIMULATE_PARTICLES] = (omp_get_wtime() - t) * 1e3;
    t = omp_get_wtime();
    if (s->prm.push_apart_enabled) fo_push_apart(s);
    s->stats.timer_ms[FSIM_T_PUSH_APART] = (omp_get_wtime() - t) * 1e3;
    t = omp_get_wtime();
    fo_push_out(s);
    s->stats.timer_ms[FSIM_T_PUSH_OUT] = (omp_get_wtime() - t) * 1e3;
    if (s->prm.stop_particles) fo_stop(s);
    t = omp_get_wtime();
    fo_p2g(s);
    s->stats.timer_ms[FSIM_T_P2G] = (omp_get_wtime() - t) * 1e3;
    t = omp_get_wtime();
    fo_mark(s);
    fo_obstacles(s);
    fo_borders(s);
    fo_post_p2g(s, s->prm.gravity_enabled ? (double)s->prm.gravity * dt : 0.0);
    s->stats.timer_ms[FSIM_T_INCOMPRESSIBILITY_PREP] = (omp_get_wtime() - t) * 1e3;
    t = omp_get_wtime();
    fo_solve(s, dt);
    s->stats.timer_ms[FSIM_T_INCOMPRESSIBILITY] = (omp_get_wtime() - t) * 1e3;
    s->stats.solve_ms = s->stats.timer_ms[FSIM_T_INCOMPRESSIBILITY];
    t = omp_get_wtime();
    fo_extrapolate(s);
    s->stats.timer_ms[FSIM_T_EXTRAPOLATION] = (omp_get_wtime() - t) * 1e3;
    t = omp_get_wtime();
    fo_g2p(s);
    s->stats.timer_ms[FSIM_T_G2P] = (omp_get_wtime() - t) * 1e3;
    s->stats.step_ms = (omp_get_wtime() - t0) * 1e3;
    s->stats.num_particles = s->np;
    s->stats.num_fluid_cells = s->nfluid;
    if (out) *out = s->stats;
}

void fo_get_stats(const fo_sim* s, fsim_step_stats* out) { *out = s->stats; }

int fo_get_grid(const fo_sim* s, int field, void* dst) {
    const int64_t nc = s->nc;
    if (field >= FSIM_FIELD_U && field <= FSIM_FIELD_W2) {
        double* d = (double*)dst;
        int a = (field - FSIM_FIELD_U) % 3;
        int two = field >= FSIM_FIELD_U2;
        for (int64_t i = 0; i < nc; i++) d[i] = two ? s->cells[i].f[a].v2 : s->cells[i].f[a].v;
        return 0;
    }
    switch (field) {
    case FSIM_FIELD_LABEL:
        for (int64_t i = 0; i < nc; i++) ((unsigned char*)dst)[i] = (unsigned char)s->cells[i].type;
        return 0;
    case FSIM_FIELD_DENSITY:
        for (int64_t i = 0; i < nc; i++) ((double*)dst)[i] = s->cells[i].avg;
        return 0;
    case FSIM_FIELD_PRESSURE: memcpy(dst, s->pressure, sizeof(double) * nc); return 0;
    case FSIM_FIELD_RHS: memcpy(dst, s->rhs, sizeof(double) * nc); return 0;
    case FSIM_FIELD_PRECON: {
        double* d = (double*)dst;
        memset(d, 0, sizeof(double) * nc);
        for (int64_t p = 0; p < s->nfluid; p++) d[s->fluid[p]] = s->precon[p];
        return 0;
    }
    default: return -1;
    }
}

int fo_set_grid(fo_sim* s, int field, const void* src) {
    const int64_t nc = s->nc;
    if (field >= FSIM_FIELD_U && field <= FSIM_FIELD_W2) {
        const double* d = (const double*)src;
        int a = (field - FSIM_FIELD_U) % 3;
        int two = field >= FSIM_FIELD_U2;
        for (int64_t i = 0; i < nc; i++) {
            if (two) s->cells[i].f[a].v2 = d[i];
            else s->cells[i].f[a].v = d[i];
        }
        return 0;
    }
    switch (field) {
    case FSIM_FIELD_LABEL:
        for (int64_t i = 0; i < nc; i++) s->cells[i].type = ((const unsigned char*)src)[i];
        build_fluid_list(s);
        return 0;
    case FSIM_FIELD_DENSITY:
        for (int64_t i = 0; i < nc; i++) s->cells[i].avg = ((const double*)src)[i];
        return 0;
    default: return -1;
    }
}

/* Pressure-solve operators on dense vectors (test hook). */
int fo_solver_apply(fo_sim* s, int op, double dt, const double* in, double* out) {
    build_fluid_list(s);
    const int64_t nf = s->nfluid;
    calc_a(s, dt);
    calc_precon(s);
    size_t sz = (size_t)(nf ? nf : 1) * sizeof(double);
    double* a = (double*)malloc(sz);
    double* b = (double*)malloc(sz);
    double* q = (double*)malloc(sz);
    for (int64_t p = 0; p < nf; p++) a[p] = in ? in[s->fluid[p]] : 0.0;
    memset(out, 0, sizeof(double) * s->nc);
    int rc = 0;
    if (op == FSIM_OP_APPLY_A) apply_a(s, a, b);
    else if (op == FSIM_OP_APPLY_MIC0) apply_precon(s, a, q, b);
    else if (op == FSIM_OP_FACTOR_MIC0) memcpy(b, s->precon, sz);
    else rc = -1;
    if (!rc) for (int64_t p = 0; p < nf; p++) out[s->fluid[p]] = b[p];
    free(a); free(b); free(q);
    return rc;
}
