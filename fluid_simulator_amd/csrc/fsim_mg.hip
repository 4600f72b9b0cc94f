// fsim_mg.hip — the learned multigrid preconditioner of flexible PCG (no
// counterpart in the reference; MIC(0)-PCG stays the fallback and the
// solve its results are measured against).
//
// z = M(r) is one V-cycle on a hierarchy of aggregation levels whose
// coarsest level is solved by the learned network:
//   * level 0 is the solve's own operator A = s L0 (s = dt / (rho dx^2),
//     L0 the 7-point matrix of calculateAMatrix, bridsonSolverGrid.cpp:40-77:
//     diagonal = non-SOLID neighbours, -1 towards WATER neighbours), read
//     from nbinfo in the solver's rotated layout;
//   * level l+1 aggregates 2x2x2 cells of level l: P is piecewise constant,
//     R = P^T, and L_{l+1} = P^T L_l P (Galerkin).  Aggregates of a 7-point
//     operator couple only face neighbours, so each level stores four FP32
//     arrays: diagonal and the couplings to the +x, +y, +z neighbour;
//   * smoothing: red-black Gauss-Seidel over the parity of x + y + z (one
//     colour has no coupling inside itself, so a half-sweep is a parallel
//     pass with no races), two sweeps before and two after the coarse
//     correction, which is added with the over-correction factor 1.6 that
//     piecewise-constant aggregation needs;
//   * the coarsest level (the first whose largest side is <= 16 cells) is
//     solved by the network from (b / max|b|, the level's operator
//     normalised by 4^-level), followed by a few red-black sweeps; for
//     comparisons it can instead be left out (smoothing only) or solved
//     accurately (600 sweeps), which also harvests training samples.
// The V-cycle is not a symmetric linear map (the network is not linear, and
// pre- and post-smoothing visit the colours in the same order), hence
// flexible PCG.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "fsim_internal.hpp"

namespace fsim {

constexpr float kMgOverCorrect = 1.6f;
constexpr int kMgPreSweeps = 2, kMgPostSweeps = 2;
constexpr int kMgExactSweeps = 600;

// ---- layout helpers -------------------------------------------------------------
struct MgGrid {
    int nx, ny, nz;
    const float *diag, *ox, *oy, *oz;
    float *b, *x;
};
__device__ inline int64_t lin(int nx, int ny, int nz, int x, int y, int z) { return ((int64_t)x * ny + y) * nz + z; }
__device__ inline void unlin(int64_t i, int ny, int nz, int& x, int& y, int& z) {
    z = (int)(i % nz);
    y = (int)((i / nz) % ny);
    x = (int)(i / ((int64_t)nz * ny));
}

struct Mg0 {               // level 0: the solve's operator in the rotated layout
    int nx, ny, nz, off63;
    int64_t nc;
    const uint16_t* nbinfo;
    const float* b;        // r of the PCG iteration
    float* x;              // z
    float binv;            // 1 / s: level 0 solves L0 x = r / s
};
__device__ inline SolveArgs mg0_args(const Mg0& A) {
    SolveArgs S;
    S.nx = A.nx;
    S.ny = A.ny;
    S.nz = A.nz;
    S.off63 = A.off63;
    S.nc = A.nc;
    S.scale_f = 0.f;
    S.scale = 0.0;
    return S;
}
__device__ inline int nb_nonsolid(uint16_t nb) { return (nb >> NB_NONSOLID_SHIFT) & 7; }

// ---- coarsening ---------------------------------------------------------------------
// Level 1 from level 0: an aggregate's diagonal sums its fluid children's
// non-SOLID counts less twice each WATER-WATER coupling inside it; its +a
// coupling counts the couplings of its upper children to the next aggregate.
__global__ void __launch_bounds__(kBlock) k_mg_coarsen0(Mg0 F, int cx, int cy, int cz, float* __restrict__ diag,
                                                         float* __restrict__ ox, float* __restrict__ oy,
                                                         float* __restrict__ oz) {
    const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= (int64_t)cx * cy * cz) return;
    int X, Y, Z;
    unlin(I, cy, cz, X, Y, Z);
    float d = 0.f, c[3] = {0.f, 0.f, 0.f};
    for (int k = 0; k < 8; k++) {
        const int a = k >> 2, b = (k >> 1) & 1, e = k & 1;
        const int x = 2 * X + a, y = 2 * Y + b, z = 2 * Z + e;
        if (x >= F.nx || y >= F.ny || z >= F.nz) continue;
        const uint16_t nb = F.nbinfo[ridx(x, y, z, F.nx, F.ny, F.nz)];
        if (!(nb & NB_FLUID)) continue;
        d += (float)nb_nonsolid(nb);
        const bool up[3] = {(nb & NB_PX) != 0, (nb & NB_PY) != 0, (nb & NB_PZ) != 0};
        const int loc[3] = {a, b, e};
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
            if (!up[ax]) continue;
            if (loc[ax] == 0) d -= 2.f;   // the +ax neighbour is a sibling
            else c[ax] += 1.f;            // it belongs to the next aggregate
        }
    }
    diag[I] = d;
    ox[I] = c[0];
    oy[I] = c[1];
    oz[I] = c[2];
}

// Level l+1 from level l (coefficient arrays)
__global__ void __launch_bounds__(kBlock) k_mg_coarsen(MgGrid F, int cx, int cy, int cz, float* __restrict__ diag,
                                                        float* __restrict__ ox, float* __restrict__ oy,
                                                        float* __restrict__ oz) {
    const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= (int64_t)cx * cy * cz) return;
    int X, Y, Z;
    unlin(I, cy, cz, X, Y, Z);
    float d = 0.f, c[3] = {0.f, 0.f, 0.f};
    for (int k = 0; k < 8; k++) {
        const int a = k >> 2, b = (k >> 1) & 1, e = k & 1;
        const int x = 2 * X + a, y = 2 * Y + b, z = 2 * Z + e;
        if (x >= F.nx || y >= F.ny || z >= F.nz) continue;
        const int64_t i = lin(F.nx, F.ny, F.nz, x, y, z);
        const float di = F.diag[i];
        if (di == 0.f) continue;
        d += di;
        const float o[3] = {F.ox[i], F.oy[i], F.oz[i]};
        const int loc[3] = {a, b, e};
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
            if (loc[ax] == 0) d -= 2.f * o[ax];
            else c[ax] += o[ax];
        }
    }
    diag[I] = d;
    ox[I] = c[0];
    oy[I] = c[1];
    oz[I] = c[2];
}

// ---- smoothing: one red-black half-sweep --------------------------------------------
// Launch shape of the per-cell kernels: a block is kMgBy rows (y) x kMgBz
// columns (z) of one plane (level 0: rotated plane p; coarse levels: x), the
// grid (z chunks, y chunks, planes), so no thread divides its index.  In the
// rotated layout x + z is constant along each 64-column z-tile of a plane
// row (x = p - ((z - 1) mod 64) up to the wrap by nx), so for even nx a row
// has one colour and a half-sweep's other-colour rows leave at once.
constexpr int kMgBz = 64, kMgBy = 4;
struct MgLaunch {
    dim3 grid, block;
};
static MgLaunch mg_launch(int nx, int ny, int nz) {
    int bz = kMgBz;
    while (bz > 16 && bz / 2 >= nz) bz /= 2;   // small coarse levels: narrower z, more rows per block
    const int by = kMgBz * kMgBy / bz;
    return MgLaunch{dim3((unsigned)((nz + bz - 1) / bz), (unsigned)((ny + by - 1) / by), (unsigned)nx),
                    dim3((unsigned)bz, (unsigned)by)};
}

// FIRST: the first half-sweep from x = 0 (no neighbour terms; the other colour
// is set to 0 on the level's unknowns).  ROWS (even nx, not FIRST): the grid
// covers only the rows of colour PAR: x + z = p + 1 (mod 2) on every rotated
// row, since (z - 1) mod 64 has the parity of z - 1 and the wrap adds nx.
template <int PAR, bool FIRST, bool ROWS = false>
__global__ void __launch_bounds__(kMgBz * kMgBy) k_mg_rbgs0(Mg0 A) {
    static_assert(!(FIRST && ROWS), "the first half-sweep also clears the other colour");
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.z;
    const int yr = blockIdx.y * blockDim.y + threadIdx.y;
    const int y = ROWS ? 2 * yr + ((PAR + p + 1) & 1) : yr;
    if (z >= A.nz || y >= A.ny) return;
    const int x = unrot_x(p, z, A.nx);
    const bool mine = ROWS || ((x + y + z) & 1) == PAR;
    if (!FIRST && !mine) return;
    const int64_t i = ((int64_t)p * A.ny + y) * A.nz + z;
    const uint16_t nb = A.nbinfo[i];
    if (!(nb & NB_FLUID)) return;
    if (!mine) {
        A.x[i] = 0.f;   // FIRST only
        return;
    }
    float s = A.b[i] * A.binv;
    if (!FIRST) {
        const RNeighbours n = rneighbours(mg0_args(A), i, p, y, z);
        if (nb & NB_MX) s += A.x[n.mx];
        if (nb & NB_PX) s += A.x[n.px];
        if (nb & NB_MY) s += A.x[n.my];
        if (nb & NB_PY) s += A.x[n.py];
        if (nb & NB_MZ) s += A.x[n.mz];
        if (nb & NB_PZ) s += A.x[n.pz];
    }
    A.x[i] = s / (float)nb_nonsolid(nb);
}

template <int PAR, bool FIRST>
__global__ void __launch_bounds__(kMgBz * kMgBy) k_mg_rbgs(MgGrid G) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int x = blockIdx.z;
    if (z >= G.nz || y >= G.ny) return;
    const bool mine = ((x + y + z) & 1) == PAR;
    if (!FIRST && !mine) return;
    const int64_t i = lin(G.nx, G.ny, G.nz, x, y, z);
    const float d = G.diag[i];
    if (d == 0.f) return;
    if (!mine) {
        G.x[i] = 0.f;   // FIRST only
        return;
    }
    float s = G.b[i];
    if (!FIRST) {
        const int64_t sx = (int64_t)G.ny * G.nz, sy = G.nz;
        float o;
        if ((o = G.ox[i]) != 0.f) s += o * G.x[i + sx];
        if (x > 0 && (o = G.ox[i - sx]) != 0.f) s += o * G.x[i - sx];
        if ((o = G.oy[i]) != 0.f) s += o * G.x[i + sy];
        if (y > 0 && (o = G.oy[i - sy]) != 0.f) s += o * G.x[i - sy];
        if ((o = G.oz[i]) != 0.f) s += o * G.x[i + 1];
        if (z > 0 && (o = G.oz[i - 1]) != 0.f) s += o * G.x[i - 1];
    }
    G.x[i] = s / d;
}

// ---- residual + restriction -----------------------------------------------------------
// b_{l+1}[I] = sum over the fluid children i of (b_i - (L x)_i).
// Level 0 takes two passes.  The children of a coarse cell lie in different
// rotated planes (x + ((z - 1) mod 64) differs between them), so a thread per
// coarse cell gathering them reads every word from another plane.  Instead,
// k_mg_resid0 writes the level-0 residual in the rotated layout (0 off the
// unknowns), a block-coalesced pass like the smoothing; then each block of
// k_mg_restrict_rot reads the rotated rows that hold a 64 (x) x 2 (y) x 64 (z)
// fine box, each row a contiguous run along z, into LDS, and sums every
// coarse cell's 8 children there in a fixed order.
// The residual follows the last pre-smoothing half-sweep, which solved the
// rows of colour 0 exactly: their residual is 0 (up to rounding) and ROWS
// (even nx, whose rotated rows each have one colour) computes only the rows
// of colour 1; the buffer's colour-0 rows keep the zeros of its allocation.
template <bool ROWS>
__global__ void __launch_bounds__(kMgBz * kMgBy) k_mg_resid0(Mg0 A, float* __restrict__ res) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.z;
    const int yr = blockIdx.y * blockDim.y + threadIdx.y;
    const int y = ROWS ? 2 * yr + (p & 1) : yr;   // colour 1: p + 1 + y odd
    if (z >= A.nz || y >= A.ny) return;
    const int64_t i = ((int64_t)p * A.ny + y) * A.nz + z;
    const uint16_t nb = A.nbinfo[i];
    float r = 0.f;
    if (nb & NB_FLUID) {
        const RNeighbours n = rneighbours(mg0_args(A), i, p, y, z);
        float ax = (float)nb_nonsolid(nb) * A.x[i];
        if (nb & NB_MX) ax -= A.x[n.mx];
        if (nb & NB_PX) ax -= A.x[n.px];
        if (nb & NB_MY) ax -= A.x[n.my];
        if (nb & NB_PY) ax -= A.x[n.py];
        if (nb & NB_MZ) ax -= A.x[n.mz];
        if (nb & NB_PZ) ax -= A.x[n.pz];
        r = A.b[i] * A.binv - ax;
    }
    res[i] = r;
}

constexpr int kRtC = 32;   // coarse cells per block along x and along z (a 64 x 2 x 64 fine box)
constexpr int kRtBatch = 8;   // rotated rows a wave loads before storing them
__global__ void __launch_bounds__(256) k_mg_restrict_rot(Mg0 A, const float* __restrict__ res, MgGrid C) {
    // fine box [yy][x - fx0][z - fz0]; a row of 64 words: the lanes of one rotated
    // row store with a stride of -63 words, a different bank each
    __shared__ float F[2][2 * kRtC][2 * kRtC];
    const int X0 = blockIdx.x * kRtC, Y = blockIdx.y, Z0 = blockIdx.z * kRtC;
    const int fx0 = 2 * X0, fz0 = 2 * Z0;
    for (int k = threadIdx.x; k < 2 * 2 * kRtC * 2 * kRtC; k += blockDim.x) (&F[0][0][0])[k] = 0.f;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const int z = fz0 + lane;
    const int rz = rot_of(z);
    // the cells x in [fx0, fx0 + 64) of column z are in the rotated rows fx0 + rz .. fx0 + rz + 63:
    // rows fx0 .. fx0 + 126 cover the box; each wave loads kRtBatch rows, then stores them
    constexpr int kRows = 2 * kRtC + 63;
    for (int k0 = wave * kRtBatch; k0 < 2 * kRows; k0 += nwaves * kRtBatch) {
        float v[kRtBatch];
        int dst[kRtBatch];
#pragma unroll
        for (int j = 0; j < kRtBatch; j++) {
            const int k = k0 + j;
            const int yy = k / kRows, pr = fx0 + k % kRows;
            const int y = 2 * Y + yy;
            const int x = pr - rz;
            const bool ok = k < 2 * kRows && y < A.ny && z < A.nz && x >= fx0 && x < fx0 + 2 * kRtC && x < A.nx;
            const int prow = pr < A.nx ? pr : pr % A.nx;   // rot_plane(x, z)
            v[j] = ok ? res[((int64_t)prow * A.ny + y) * A.nz + z] : 0.f;
            dst[j] = ok ? (yy * 2 * kRtC + (x - fx0)) * 2 * kRtC + lane : -1;
        }
#pragma unroll
        for (int j = 0; j < kRtBatch; j++)
            if (dst[j] >= 0) (&F[0][0][0])[dst[j]] = v[j];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kRtC * kRtC; k += blockDim.x) {
        const int xl = k / kRtC, zl = k % kRtC;
        const int X = X0 + xl, Z = Z0 + zl;
        if (X >= C.nx || Z >= C.nz) continue;
        const int64_t I = lin(C.nx, C.ny, C.nz, X, Y, Z);
        float sum = 0.f;
        if (C.diag[I] != 0.f) {
#pragma unroll
            for (int c = 0; c < 8; c++) sum += F[(c >> 1) & 1][2 * xl + (c >> 2)][2 * zl + (c & 1)];
        }
        C.b[I] = sum;
    }
}

__global__ void __launch_bounds__(kBlock) k_mg_restrict(MgGrid F, MgGrid C) {
    const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= (int64_t)C.nx * C.ny * C.nz) return;
    if (C.diag[I] == 0.f) {
        C.b[I] = 0.f;
        return;
    }
    int X, Y, Z;
    unlin(I, C.ny, C.nz, X, Y, Z);
    const int64_t sx = (int64_t)F.ny * F.nz, sy = F.nz;
    float sum = 0.f;
    for (int k = 0; k < 8; k++) {
        const int x = 2 * X + (k >> 2), y = 2 * Y + ((k >> 1) & 1), z = 2 * Z + (k & 1);
        if (x >= F.nx || y >= F.ny || z >= F.nz) continue;
        const int64_t i = lin(F.nx, F.ny, F.nz, x, y, z);
        const float d = F.diag[i];
        if (d == 0.f) continue;
        float ax = d * F.x[i];
        float o;
        if ((o = F.ox[i]) != 0.f) ax -= o * F.x[i + sx];
        if (x > 0 && (o = F.ox[i - sx]) != 0.f) ax -= o * F.x[i - sx];
        if ((o = F.oy[i]) != 0.f) ax -= o * F.x[i + sy];
        if (y > 0 && (o = F.oy[i - sy]) != 0.f) ax -= o * F.x[i - sy];
        if ((o = F.oz[i]) != 0.f) ax -= o * F.x[i + 1];
        if (z > 0 && (o = F.oz[i - 1]) != 0.f) ax -= o * F.x[i - 1];
        sum += F.b[i] - ax;
    }
    C.b[I] = sum;
}

// ---- prolongation: x_i += 1.6 x_{l+1}[parent(i)] on the level's unknowns ----------------
__global__ void __launch_bounds__(kMgBz * kMgBy) k_mg_prolong0(Mg0 A, MgGrid C) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int p = blockIdx.z;
    if (z >= A.nz || y >= A.ny) return;
    const int64_t i = ((int64_t)p * A.ny + y) * A.nz + z;
    if (!(A.nbinfo[i] & NB_FLUID)) return;
    const int x = unrot_x(p, z, A.nx);
    A.x[i] += kMgOverCorrect * C.x[lin(C.nx, C.ny, C.nz, x >> 1, y >> 1, z >> 1)];
}

__global__ void __launch_bounds__(kMgBz * kMgBy) k_mg_prolong(MgGrid F, MgGrid C) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int x = blockIdx.z;
    if (z >= F.nz || y >= F.ny) return;
    const int64_t i = lin(F.nx, F.ny, F.nz, x, y, z);
    if (F.diag[i] == 0.f) return;
    F.x[i] += kMgOverCorrect * C.x[lin(C.nx, C.ny, C.nz, x >> 1, y >> 1, z >> 1)];
}

// ---- host side ------------------------------------------------------------------------------
struct MgLevelBuf {
    int nx = 0, ny = 0, nz = 0;
    int64_t nc = 0;
    float *diag = nullptr, *ox = nullptr, *oy = nullptr, *oz = nullptr, *b = nullptr, *x = nullptr;
};
struct MgState {
    int nlev = 0;   // levels 0 .. nlev-1; nlev-1 is the coarse (network) level
    MgLevelBuf lv[kMgMaxLevels];
    float* res0 = nullptr;   // level-0 residual, rotated layout
    std::vector<void*> allocs;
    // harvest: up to cap coarse (b, x) pairs of the running solve
    int harvest_cap = 0, harvest_n = 0;
    float *hb = nullptr, *hx = nullptr;
    // the V-cycle captured as a HIP graph, with what its kernel arguments were made from
    hipGraphExec_t exec = nullptr;
    struct Key {
        const float* r;
        float* z;
        double scale;
        int mode, nu, conv, net_gen;
        bool operator==(const Key& o) const {
            return r == o.r && z == o.z && scale == o.scale && mode == o.mode && nu == o.nu && conv == o.conv &&
                   net_gen == o.net_gen;
        }
    } key{};
};

static MgGrid grid_of(const MgLevelBuf& L) { return MgGrid{L.nx, L.ny, L.nz, L.diag, L.ox, L.oy, L.oz, L.b, L.x}; }

static void drop_graph(MgState* M) {
    if (M->exec) (void)hipGraphExecDestroy(M->exec);
    M->exec = nullptr;
}

void mg_free(fsim_ctx* ctx) {
    MgState* M = ctx->mg;
    if (!M) return;
    drop_graph(M);
    for (void* p : M->allocs) (void)hipFree(p);
    delete M;
    ctx->mg = nullptr;
}

int mg_levels_for(int nx, int ny, int nz) {
    int l = 0;
    while (std::max({nx, ny, nz}) > kMgCoarseMaxSide && l < kMgMaxLevels - 1) {
        nx = (nx + 1) / 2;
        ny = (ny + 1) / 2;
        nz = (nz + 1) / 2;
        l++;
    }
    return l + 1;
}

static int mg_alloc(fsim_ctx* ctx) {
    if (ctx->mg) return FSIM_OK;
    if (ctx->g.two_d) return set_err(ctx, FSIM_ERR_INVALID, "the multigrid preconditioner is built for 3D grids only");
    MgState* M = new MgState();
    ctx->mg = M;
    M->nlev = mg_levels_for(ctx->g.nx, ctx->g.ny, ctx->g.nz);
    M->lv[0].nx = ctx->g.nx;
    M->lv[0].ny = ctx->g.ny;
    M->lv[0].nz = ctx->g.nz;
    M->lv[0].nc = ctx->g.nc;
    if (M->nlev > 1) {
        if (hipMalloc(reinterpret_cast<void**>(&M->res0), sizeof(float) * (size_t)ctx->g.nc) != hipSuccess)
            return set_err(ctx, FSIM_ERR_OOM, "multigrid residual allocation failed");
        M->allocs.push_back(M->res0);
        FSIM_TRY(fill_f32(ctx, M->res0, 0.0f, ctx->g.nc));   // the colour-0 rows stay 0 (k_mg_resid0)
    }
    for (int l = 1; l < M->nlev; l++) {
        MgLevelBuf& L = M->lv[l];
        L.nx = (M->lv[l - 1].nx + 1) / 2;
        L.ny = (M->lv[l - 1].ny + 1) / 2;
        L.nz = (M->lv[l - 1].nz + 1) / 2;
        L.nc = (int64_t)L.nx * L.ny * L.nz;
        float** arrs[6] = {&L.diag, &L.ox, &L.oy, &L.oz, &L.b, &L.x};
        for (float** a : arrs) {
            if (hipMalloc(reinterpret_cast<void**>(a), sizeof(float) * (size_t)L.nc) != hipSuccess)
                return set_err(ctx, FSIM_ERR_OOM, "multigrid level allocation failed");
            M->allocs.push_back(*a);
            FSIM_TRY(fill_f32(ctx, *a, 0.0f, L.nc));   // x must be 0 where a level has no unknown
        }
    }
    return FSIM_OK;
}

static Mg0 mg0(fsim_ctx* ctx, const float* r, float* z, double scale) {
    Mg0 A;
    A.nx = ctx->g.nx;
    A.ny = ctx->g.ny;
    A.nz = ctx->g.nz;
    A.off63 = 63 % ctx->g.nx;
    A.nc = ctx->g.nc;
    A.nbinfo = ctx->nbinfo;
    A.b = r;
    A.x = z;
    A.binv = (float)(1.0 / scale);
    return A;
}

// Builds the coarse operators of the current labels (once per solve, after nbinfo).
int mg_setup(fsim_ctx* ctx) {
    FSIM_TRY(mg_alloc(ctx));
    MgState* M = ctx->mg;
    if (ctx->mg_coarse == FSIM_MG_COARSE_NETWORK) {   // the network's buffers exist before any capture
        const MgLevelBuf& L = M->lv[M->nlev - 1];
        if (M->nlev > 1) FSIM_TRY(learned_prepare(ctx, L.nx, L.ny, L.nz));
        else FSIM_TRY(learned_prepare(ctx, ctx->g.nx, ctx->g.ny, ctx->g.nz));
    }
    FSIM_PROF(ctx, KID_MG);
    if (M->nlev > 1) {
        MgLevelBuf& C = M->lv[1];
        hipLaunchKernelGGL(k_mg_coarsen0, dim3(nblocks(C.nc)), dim3(kBlock), 0, ctx->stream, mg0(ctx, nullptr, nullptr, 1.0),
                           C.nx, C.ny, C.nz, C.diag, C.ox, C.oy, C.oz);
        for (int l = 2; l < M->nlev; l++) {
            MgLevelBuf& L = M->lv[l];
            hipLaunchKernelGGL(k_mg_coarsen, dim3(nblocks(L.nc)), dim3(kBlock), 0, ctx->stream, grid_of(M->lv[l - 1]), L.nx,
                               L.ny, L.nz, L.diag, L.ox, L.oy, L.oz);
        }
    }
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

static void smooth0(fsim_ctx* ctx, const Mg0& A, int sweeps, bool first) {
    const MgLaunch L = mg_launch(A.nx, A.ny, A.nz);
    const dim3 g = L.grid, b = L.block;
    // even nx: the non-first half-sweeps launch the rows of their colour only
    const MgLaunch H = mg_launch(A.nx, (A.ny + 1) / 2, A.nz);
    const bool rows = (A.nx & 1) == 0;
    for (int k = 0; k < sweeps; k++) {
        if (first && k == 0) hipLaunchKernelGGL((k_mg_rbgs0<1, true>), g, b, 0, ctx->stream, A);
        else if (rows) hipLaunchKernelGGL((k_mg_rbgs0<1, false, true>), H.grid, H.block, 0, ctx->stream, A);
        else hipLaunchKernelGGL((k_mg_rbgs0<1, false>), g, b, 0, ctx->stream, A);
        if (rows) hipLaunchKernelGGL((k_mg_rbgs0<0, false, true>), H.grid, H.block, 0, ctx->stream, A);
        else hipLaunchKernelGGL((k_mg_rbgs0<0, false>), g, b, 0, ctx->stream, A);
    }
}
static void smooth(fsim_ctx* ctx, const MgLevelBuf& L, int sweeps, bool first) {
    const MgGrid G = grid_of(L);
    const MgLaunch S = mg_launch(L.nx, L.ny, L.nz);
    const dim3 g = S.grid, b = S.block;
    for (int k = 0; k < sweeps; k++) {
        if (first && k == 0) hipLaunchKernelGGL((k_mg_rbgs<1, true>), g, b, 0, ctx->stream, G);
        else hipLaunchKernelGGL((k_mg_rbgs<1, false>), g, b, 0, ctx->stream, G);
        hipLaunchKernelGGL((k_mg_rbgs<0, false>), g, b, 0, ctx->stream, G);
    }
}

// Solves the coarsest level (lv[c], or level 0 itself when the grid is small).
static int coarse_solve(fsim_ctx* ctx, const Mg0& A, int c) {
    MgState* M = ctx->mg;
    const int mode = ctx->mg_coarse;
    const int nu = ctx->learned_smoothing;
    const double norm = std::ldexp(1.0, -2 * c);   // 4^-c: the level's operator on the network's scale
    if (mode == FSIM_MG_COARSE_NETWORK) {
        if (c == 0) {
            FSIM_TRY(learned_apply_level0(ctx, A.b, A.x, A.binv));
            smooth0(ctx, A, nu, false);
        } else {
            const MgLevelBuf& L = M->lv[c];
            FSIM_TRY(learned_apply_level(ctx, L.nx, L.ny, L.nz, L.b, L.diag, L.ox, L.oy, L.oz, norm, L.x));
            smooth(ctx, L, nu, false);
        }
        return FSIM_OK;
    }
    const int sweeps = mode == FSIM_MG_COARSE_ACCURATE ? kMgExactSweeps : std::max(nu, 1);
    if (c == 0) smooth0(ctx, A, sweeps, true);
    else smooth(ctx, M->lv[c], sweeps, true);
    if (mode == FSIM_MG_COARSE_ACCURATE && c > 0 && M->harvest_n < M->harvest_cap) {
        const MgLevelBuf& L = M->lv[c];
        const size_t bytes = sizeof(float) * (size_t)L.nc;
        FSIM_HIP(ctx, hipMemcpyAsync(M->hb + (size_t)M->harvest_n * L.nc, L.b, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        FSIM_HIP(ctx, hipMemcpyAsync(M->hx + (size_t)M->harvest_n * L.nc, L.x, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        M->harvest_n++;
    }
    return FSIM_OK;
}

// z = V-cycle(r) / s: level-0 vectors in the solver's rotated layout.
static int vcycle(fsim_ctx* ctx, const float* r, float* z, double scale) {
    MgState* M = ctx->mg;
    const Mg0 A = mg0(ctx, r, z, scale);
    const int c = M->nlev - 1;
    if (c == 0) {   // the grid itself is the coarse level
        FSIM_TRY(coarse_solve(ctx, A, 0));
        FSIM_HIP(ctx, hipGetLastError());
        return FSIM_OK;
    }
    // down: level 0, then 1 .. c-1
    smooth0(ctx, A, kMgPreSweeps, true);
    {
        if ((A.nx & 1) == 0) {
            const MgLaunch R = mg_launch(A.nx, (A.ny + 1) / 2, A.nz);
            hipLaunchKernelGGL(k_mg_resid0<true>, R.grid, R.block, 0, ctx->stream, A, M->res0);
        } else {
            const MgLaunch R = mg_launch(A.nx, A.ny, A.nz);
            hipLaunchKernelGGL(k_mg_resid0<false>, R.grid, R.block, 0, ctx->stream, A, M->res0);
        }
        const MgLevelBuf& L1 = M->lv[1];
        const dim3 g((unsigned)((L1.nx + kRtC - 1) / kRtC), (unsigned)L1.ny, (unsigned)((L1.nz + kRtC - 1) / kRtC));
        hipLaunchKernelGGL(k_mg_restrict_rot, g, dim3(256), 0, ctx->stream, A, M->res0, grid_of(L1));
    }
    for (int l = 1; l < c; l++) {
        smooth(ctx, M->lv[l], kMgPreSweeps, true);
        hipLaunchKernelGGL(k_mg_restrict, dim3(nblocks(M->lv[l + 1].nc)), dim3(kBlock), 0, ctx->stream,
                           grid_of(M->lv[l]), grid_of(M->lv[l + 1]));
    }
    FSIM_TRY(coarse_solve(ctx, A, c));
    // up
    for (int l = c - 1; l >= 1; l--) {
        const MgLaunch P = mg_launch(M->lv[l].nx, M->lv[l].ny, M->lv[l].nz);
        hipLaunchKernelGGL(k_mg_prolong, P.grid, P.block, 0, ctx->stream, grid_of(M->lv[l]), grid_of(M->lv[l + 1]));
        smooth(ctx, M->lv[l], kMgPostSweeps, false);
    }
    const MgLaunch P0 = mg_launch(A.nx, A.ny, A.nz);
    hipLaunchKernelGGL(k_mg_prolong0, P0.grid, P0.block, 0, ctx->stream, A, grid_of(M->lv[1]));
    smooth0(ctx, A, kMgPostSweeps, false);
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

// The V-cycle is some 50 launches, most of them small: enqueued one by one they
// keep the GPU waiting on the host.  With FSIM_OPT_MG_GRAPH (default) it is
// captured once into a HIP graph and replayed; the kernels' arguments depend
// only on Key (the buffers are the context's, allocated before the capture),
// so a new capture is made only when Key changes.  The V-cycle records no
// memset node (the network clears its max|b| with a fill kernel).  Profiling
// (per-kernel events) and harvesting (host-side sample counts) run it
// directly.
int mg_apply(fsim_ctx* ctx, const float* r, float* z, double scale) {
    MgState* M = ctx->mg;
    FSIM_PROF(ctx, KID_MG);
    if (!ctx->mg_graph || ctx->prof_on || ctx->mg_harvesting) return vcycle(ctx, r, z, scale);
    const MgState::Key key{r, z, scale, ctx->mg_coarse, ctx->learned_smoothing, ctx->learned_conv_lds, ctx->learned_gen};
    if (!M->exec || !(key == M->key)) {
        drop_graph(M);
        FSIM_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const int rc = vcycle(ctx, r, z, scale);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
        if (rc != FSIM_OK || ec != hipSuccess || !g) {
            if (g) (void)hipGraphDestroy(g);
            return rc != FSIM_OK ? rc : set_err(ctx, FSIM_ERR_HIP, std::string("V-cycle capture: ") + hipGetErrorString(ec));
        }
        const hipError_t ei = hipGraphInstantiate(&M->exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            M->exec = nullptr;
            return set_err(ctx, FSIM_ERR_HIP, std::string("V-cycle graph: ") + hipGetErrorString(ei));
        }
        M->key = key;
    }
    FSIM_HIP(ctx, hipGraphLaunch(M->exec, ctx->stream));
    return FSIM_OK;
}

int mg_info(fsim_ctx* ctx, int* levels, int* dims) {
    const int n = mg_levels_for(ctx->g.nx, ctx->g.ny, ctx->g.nz);
    int d[3] = {ctx->g.nx, ctx->g.ny, ctx->g.nz};
    for (int l = 1; l < n; l++)
        for (int a = 0; a < 3; a++) d[a] = (d[a] + 1) / 2;
    if (levels) *levels = n;
    if (dims)
        for (int a = 0; a < 3; a++) dims[a] = d[a];
    return FSIM_OK;
}

// Harvest buffers for the next solve: cap coarse (b, x) pairs.
int mg_harvest_begin(fsim_ctx* ctx, int cap) {
    FSIM_TRY(mg_alloc(ctx));
    MgState* M = ctx->mg;
    if (M->nlev < 2) return set_err(ctx, FSIM_ERR_INVALID, "the grid has no coarse multigrid level to harvest");
    const int64_t nc = M->lv[M->nlev - 1].nc;
    if (cap > M->harvest_cap) {
        float *hb = nullptr, *hx = nullptr;
        if (hipMalloc(reinterpret_cast<void**>(&hb), sizeof(float) * (size_t)cap * nc) != hipSuccess ||
            hipMalloc(reinterpret_cast<void**>(&hx), sizeof(float) * (size_t)cap * nc) != hipSuccess) {
            if (hb) (void)hipFree(hb);
            return set_err(ctx, FSIM_ERR_OOM, "harvest buffer allocation failed");
        }
        FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (M->hb) (void)hipFree(M->hb);
        if (M->hx) (void)hipFree(M->hx);
        M->hb = hb;
        M->hx = hx;
        M->harvest_cap = cap;
    }
    M->harvest_n = 0;
    return FSIM_OK;
}

// Copies the harvested pairs and the coarse operator (normalised by 4^-level) to the host.
int mg_harvest_end(fsim_ctx* ctx, int cap, float* b, float* op, float* x, int* n_out) {
    MgState* M = ctx->mg;
    const int c = M->nlev - 1;
    const MgLevelBuf& L = M->lv[c];
    const int n = std::min(M->harvest_n, cap);
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const size_t nc = (size_t)L.nc;
    if (n > 0) {
        FSIM_HIP(ctx, hipMemcpy(b, M->hb, sizeof(float) * nc * n, hipMemcpyDeviceToHost));
        FSIM_HIP(ctx, hipMemcpy(x, M->hx, sizeof(float) * nc * n, hipMemcpyDeviceToHost));
    }
    std::vector<float> tmp(nc);
    const float* src[4] = {L.diag, L.ox, L.oy, L.oz};
    const float norm = std::ldexp(1.0f, -2 * c);
    for (int k = 0; k < 4; k++) {
        FSIM_HIP(ctx, hipMemcpy(tmp.data(), src[k], sizeof(float) * nc, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < nc; i++) op[k * nc + i] = tmp[i] * norm;
    }
    // the samples' x solve L x = b; on the network's scale (L / 4^c) x / 4^c solves it
    for (size_t i = 0; i < nc * (size_t)n; i++) x[i] /= norm;
    M->harvest_n = 0;
    *n_out = n;
    return FSIM_OK;
}

}  // namespace fsim
