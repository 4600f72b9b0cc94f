// fsim_grid.hip — grid-side stages of the step on gfx950.
//
// All grid kernels run over the dense cell array in the reference's flat
// order (z fastest).  P2G is a deterministic gather over the particles
// binned by cell (no float atomics): each cell's three faces sum the
// contributions of the particles of its 3x3x3 cell neighbourhood in a fixed
// order (neighbour cells in x,y,z order, particles by id inside a cell).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "fsim_internal.hpp"

namespace fsim {

// ---- fills (graph-capture safe replacements for hipMemsetAsync) -------------
template <typename T>
__global__ void __launch_bounds__(kBlock) k_fill(T* __restrict__ p, T v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}
template <typename T>
static int fill(fsim_ctx* ctx, T* p, T v, int64_t n) {
    if (n <= 0) return FSIM_OK;
    unsigned nb = (unsigned)std::min<int64_t>(nblocks(n), 65536);
    hipLaunchKernelGGL(k_fill<T>, dim3(nb), dim3(kBlock), 0, ctx->stream, p, v, n);
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}
int fill_u8(fsim_ctx* ctx, uint8_t* p, uint8_t v, int64_t n) { return fill(ctx, p, v, n); }
int fill_f32(fsim_ctx* ctx, float* p, float v, int64_t n) { return fill(ctx, p, v, n); }
int fill_f64(fsim_ctx* ctx, double* p, double v, int64_t n) { return fill(ctx, p, v, n); }
int fill_i32(fsim_ctx* ctx, int32_t* p, int32_t v, int64_t n) { return fill(ctx, p, v, n); }

struct GridArgs {
    int nx, ny, nz;
    int64_t nc;
    double cd, cdinv;     // cellD.x == cellD.y
    double cdz, cdzinv;   // cellD.z (cd except in 2D)
    float cdf, invf, invfz;
    int two_d;
};
// cellD, cellDInv of axis d (macGrid.cpp:10).  The P2G kernels take the 2D
// switch as a template argument (TWOD): their 3D instances use the x scale on
// every axis and hold no registers for the z scale (fewer scalar registers,
// one more wave per SIMD for the row gather).
template <int TWOD>
__device__ inline double cdinv_t(const GridArgs& G, int d) { return TWOD && d == 2 ? G.cdzinv : G.cdinv; }
template <int TWOD>
__device__ inline float invf_t(const GridArgs& G, int d) { return TWOD && d == 2 ? G.invfz : G.invf; }
__host__ __device__ inline double cd_of(const GridArgs& G, int d) { return d == 2 ? G.cdz : G.cd; }
static GridArgs grid_args(const fsim_ctx* ctx) {
    GridArgs G;
    G.nx = ctx->g.nx;
    G.ny = ctx->g.ny;
    G.nz = ctx->g.nz;
    G.nc = ctx->g.nc;
    G.cd = ctx->g.cd;
    G.cdinv = ctx->g.cdinv;
    G.cdz = ctx->g.cdz;
    G.cdzinv = ctx->g.cdzinv;
    G.cdf = (float)ctx->g.cd;
    G.invf = (float)ctx->g.cdinv;
    G.invfz = (float)ctx->g.cdzinv;
    G.two_d = ctx->g.two_d;
    return G;
}

__device__ inline int64_t cidx(const GridArgs& G, int x, int y, int z) { return ((int64_t)x * G.ny + y) * G.nz + z; }

// Base cell of the 8 faces of axis `a` around a particle, getFacesAround
// (macGrid.cpp:142-148): ivec3(pos*cellDInv - offset_a), then coord[a] -= 1.
// Evaluated in FP64 so stencil membership is decided exactly as in FP64.
__device__ inline void face_base(const GridArgs& G, float px, float py, float pz, int a, int b[3]) {
    const double p[3] = {(double)px * G.cdinv, (double)py * G.cdinv, (double)pz * G.cdzinv};
#pragma unroll
    for (int d = 0; d < 3; d++) b[d] = (int)(p[d] - (d == a ? 0.0 : 0.5));
    b[a] -= 1;
}

// Face centre of axis a of cell (x,y,z) (macGrid.cpp:25-27), rounded to FP32.
template <int TWOD = 1>
__device__ inline void face_pos(const GridArgs& G, int a, int x, int y, int z, float f[3]) {
    f[0] = (float)((x + (a == 0 ? 1.0 : 0.5)) * G.cd);
    f[1] = (float)((y + (a == 1 ? 1.0 : 0.5)) * G.cd);
    f[2] = (float)((z + (a == 2 ? 1.0 : 0.5)) * (TWOD ? G.cdz : G.cd));
}

// trilinearInterpoll, interpolation.h:7-20
__device__ inline float tent3(const float f[3], float px, float py, float pz, float inv) {
    float x = 1.0f - fabsf((px - f[0]) * inv);
    float y = 1.0f - fabsf((py - f[1]) * inv);
    float z = 1.0f - fabsf((pz - f[2]) * inv);
    return x * y * z;
}

// ---- P2G: resetGridValues + p2gTransfer (simulator.cpp:314-352) ------------------
struct P2GArgs {
    GridArgs G;
    const float* pos[3];
    const float* vel[3];
    const float* c[9];
    const int32_t* start;
    const int32_t* count;
    float* uvw[3];
    float* uvw2[3];
    uint8_t* label;
    float* density;
    int apic;
    int fuse_mark;   // 1: also label WATER where the bin count is > 0
};

// Gather form of the transfer: the thread of cell (x,y,z) owns that cell's
// three faces and sums, in a fixed order (neighbour cells dx, dy, dz, then
// particles in sorted order), every particle whose 2x2x2 face stencil of an
// axis holds the face (base from face_base, as in the reference's scatter).
// The same walk sums the particle density of the cell (avgPNum, the
// cell-centre stencil of markFluidCellsAndCalculateParticleDensities), which
// the reference computes on every step whether or not the pressure term of
// the right-hand side reads it.
//
// A block owns 256 consecutive cells.  For one (dx, dy) the three dz cells of
// every thread are consecutive, so the particles the block needs are one
// contiguous range of the sorted arrays.  The block stages that range in LDS
// in chunks (coalesced loads; each particle's (int)(p / dx) and half-cell
// bits computed once), then every thread walks its own sub-range of the
// chunk in order.
constexpr int P2G_CAP_FLIP = 512, P2G_CAP_APIC = 512;   // particles per staged chunk

// One PIC/FLIP particle's terms for a cell's three faces and its density.
// al_d / ce_d: the tent weights 1 - |p_d - f_d| / dx of the axis-aligned and
// the cell-centred face coordinate, each 0 where the stencil (face_base) does
// not hold this cell on axis d.  Face a takes w = (tw_x tw_y) tw_z with tw_d =
// al_d for d == a else ce_d, the density ce_x ce_y ce_z; each sum is one fused
// multiply-add per term.  A term outside the stencil adds an exact zero (all
// sums start at +0 and never become -0), so the sums equal those that skip it.
__device__ inline void p2g_terms(float alx, float cex, float aly, float cey, float alz, float cez, float vx, float vy,
                                 float vz, float swv[3], float sw[3], float& dens) {
    const float c01 = cex * cey;
    dens = __builtin_fmaf(c01, cez, dens);
    const float t0 = alx * cey;
    swv[0] = __builtin_fmaf(vx, t0 * cez, swv[0]);
    sw[0] = __builtin_fmaf(t0, cez, sw[0]);
    const float t1 = cex * aly;
    swv[1] = __builtin_fmaf(vy, t1 * cez, swv[1]);
    sw[1] = __builtin_fmaf(t1, cez, sw[1]);
    swv[2] = __builtin_fmaf(vz, c01 * alz, swv[2]);
    sw[2] = __builtin_fmaf(c01, alz, sw[2]);
}

template <int APIC, int TWOD>
__global__ void __launch_bounds__(kBlock) k_p2g_gather(P2GArgs A) {
    constexpr int CAP = APIC ? P2G_CAP_APIC : P2G_CAP_FLIP;
    constexpr int NC = APIC ? 9 : 1;
    __shared__ float s_pos[3][CAP], s_vel[3][CAP];
    __shared__ int s_t[3][CAP];          // (int)(p_d / dx)
    __shared__ uint8_t s_h[CAP];          // bit d: (int)(p_d / dx - 0.5) == (int)(p_d / dx)
    __shared__ float s_c[NC][APIC ? CAP : 1];
    const GridArgs& G = A.G;
    const int64_t c0 = (int64_t)blockIdx.x * kBlock;
    const int64_t i = c0 + threadIdx.x;
    const bool valid = i < G.nc;
    const int64_t ic = valid ? i : G.nc - 1;
    const int z = (int)(ic % G.nz);
    const int y = (int)((ic / G.nz) % G.ny);
    const int x = (int)(ic / ((int64_t)G.nz * G.ny));
    // face centres: component d of axis a's face is "aligned" (x + 1) for d == a, else the cell centre
    float fal[3], fce[3];
    {
        float f[3][3];
#pragma unroll
        for (int a = 0; a < 3; a++) face_pos<TWOD>(G, a, x, y, z, f[a]);
        fal[0] = f[0][0], fal[1] = f[1][1], fal[2] = f[2][2];
        fce[0] = f[1][0], fce[1] = f[0][1], fce[2] = f[0][2];
    }
    const int cc[3] = {x, y, z};
    float swv[3] = {0.f, 0.f, 0.f}, sw[3] = {0.f, 0.f, 0.f};
    float dens = 0.f;   // avgPNum: tent weights of the cell centre (simulator.cpp:362-366)
    const int64_t plane = (int64_t)G.ny * G.nz;
    for (int dx = -1; dx <= 1; dx++) {
        for (int dy = -1; dy <= 1; dy++) {
            const int64_t o = dx * plane + dy * G.nz;
            // the block's particle range: cells c0 + o - 1 .. c0 + o + kBlock (clamped)
            const int64_t lo = min(max(c0 + o - 1, (int64_t)0), G.nc);
            const int64_t hi = min(max(c0 + o + kBlock + 1, (int64_t)0), G.nc);
            const int u0 = A.start[lo], u1 = A.start[hi];
            if (u0 >= u1) continue;   // block-uniform
            // this thread's range: cells (x+dx, y+dy, z-1 .. z+1) inside the grid
            int j0 = 0, j1 = 0;
            const int cx = x + dx, cy = y + dy;
            if (valid && cx >= 0 && cx < G.nx && cy >= 0 && cy < G.ny) {
                const int64_t n = ic + o;
                j0 = A.start[z > 0 ? n - 1 : n];
                j1 = A.start[z < G.nz - 1 ? n + 2 : n + 1];
            }
            for (int base = u0; base < u1; base += CAP) {
                const int cnt = min(CAP, u1 - base);
                __syncthreads();   // the previous chunk is consumed
                for (int k = threadIdx.x; k < cnt; k += kBlock) {
                    const int j = base + k;
                    uint8_t h = 0;
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const float pv = A.pos[d][j];
                        const double v = (double)pv * cdinv_t<TWOD>(G, d);
                        const int t = (int)v;
                        s_pos[d][k] = pv;
                        s_vel[d][k] = A.vel[d][j];
                        s_t[d][k] = t;
                        if ((int)(v - 0.5) == t) h |= (uint8_t)(1u << d);
                    }
                    s_h[k] = h;
                    if (APIC) {
#pragma unroll
                        for (int m = 0; m < NC; m++) s_c[m][k] = A.c[m][j];
                    }
                }
                __syncthreads();
                const int e = min(j1, base + cnt);
#pragma unroll 4
                for (int j = max(j0, base); j < e; j++) {
                    const int k = j - base;
                    const float p[3] = {s_pos[0][k], s_pos[1][k], s_pos[2][k]};
                    const uint32_t h = s_h[k];
                    // face_base: b_a = t_a - 1, b_d = (int)(p_d/dx - 0.5) = t_d - 1 + h_d (d != a);
                    // the face of this cell is in the stencil when cell - b is 0 or 1 on every axis
                    int ua[3], uc[3];
                    float tal[3], tce[3];
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        ua[d] = cc[d] - s_t[d][k] + 1;
                        uc[d] = ua[d] - (int)((h >> d) & 1u);
                        tal[d] = 1.0f - fabsf((p[d] - fal[d]) * invf_t<TWOD>(G, d));
                        tce[d] = 1.0f - fabsf((p[d] - fce[d]) * invf_t<TWOD>(G, d));
                    }
                    if (!APIC) {
                        // the weights of the axes where the stencil misses this cell are 0
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            tal[d] = (unsigned)ua[d] <= 1u ? tal[d] : 0.f;
                            tce[d] = (unsigned)uc[d] <= 1u ? tce[d] : 0.f;
                        }
                        p2g_terms(tal[0], tce[0], tal[1], tce[1], tal[2], tce[2], s_vel[0][k], s_vel[1][k],
                                  s_vel[2][k], swv, sw, dens);
                        continue;
                    }
                    // getCellsAround (macGrid.cpp:188-202): base (int)(p/dx - 0.5) on every axis
                    if ((unsigned)uc[0] <= 1u && (unsigned)uc[1] <= 1u && (unsigned)uc[2] <= 1u)
                        dens += tce[0] * tce[1] * tce[2];
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        bool in = true;
                        float tw[3];
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            const int u = d == a ? ua[d] : uc[d];
                            in = in && (unsigned)u <= 1u;
                            tw[d] = d == a ? tal[d] : tce[d];
                        }
                        if (!in) continue;
                        const float w = tw[0] * tw[1] * tw[2];
                        float va = s_vel[a][k];
                        if (APIC) {
                            // (v_a + c[a] . (face - p)) * w, simulator.cpp:326-328
                            const float f0 = a == 0 ? fal[0] : fce[0], f1 = a == 1 ? fal[1] : fce[1],
                                        f2 = a == 2 ? fal[2] : fce[2];
                            const float d0 = f0 - p[0], d1 = f1 - p[1], d2 = f2 - p[2];
                            va = va + (s_c[3 * a][k] * d0 + s_c[3 * a + 1][k] * d1 + s_c[3 * a + 2][k] * d2);
                        }
                        swv[a] += va * w;
                        sw[a] += w;
                    }
                }
            }
        }
    }
    if (!valid) return;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        A.uvw[a][i] = sw[a] > 1e-6f ? swv[a] / sw[a] : 0.0f;
        A.uvw2[a][i] = 0.0f;
    }
    A.label[i] = (A.fuse_mark && A.count[i] > 0) ? (uint8_t)FSIM_WATER : (uint8_t)FSIM_AIR;
    A.density[i] = dens;
}

// Tiled gather: a block owns TX x TY cell rows (x, y) of 64 cells along z,
// one wave per cell row and one thread per cell, and stages each of the
// (TX + 2) x (TY + 2) particle rows around them once, 66 cells of each
// (k_p2g_gather stages 9 rows for every cell row: each particle crossed HBM
// about 9 times).  A wave works on a staged row when its cell row's 3 x 3
// (x, y) neighbourhood holds it.  Rows are staged in (x, y) order, so every
// cell adds the rows of its neighbourhood in k_p2g_gather's (dx, dy) order,
// and each row's particles in sorted order, with the same arithmetic: the
// sums are bit-identical.  It serves APIC; PIC and FLIP take k_p2g_tile_w.
#ifndef FSIM_P2G_TX
#define FSIM_P2G_TX 2
#endif
#ifndef FSIM_P2G_TY
#define FSIM_P2G_TY 4
#endif
constexpr int P2G_TX = FSIM_P2G_TX, P2G_TY = FSIM_P2G_TY, P2G_TZ = 64;
constexpr int P2G_TILE_CAP = 1024;   // particles per staged chunk (a row segment is 66 cells)
template <int APIC, int TX, int TY, int TWOD>
__global__ void __launch_bounds__(P2G_TZ * TX * TY) k_p2g_tile(P2GArgs A) {
    constexpr int CAP = P2G_TILE_CAP;
    constexpr int NC = APIC ? 9 : 1;
    constexpr int NT = P2G_TZ * TX * TY;
    __shared__ float s_pos[3][CAP], s_vel[3][CAP];
    __shared__ int s_t[3][CAP];
    __shared__ uint8_t s_h[CAP];
    __shared__ float s_c[NC][APIC ? CAP : 1];
    const GridArgs& G = A.G;
    const int tid = threadIdx.x + P2G_TZ * threadIdx.y;
    const int zc0 = blockIdx.x * P2G_TZ;
    const int ri = threadIdx.y / TY, rj = threadIdx.y % TY;   // this wave's cell row in the tile
    const int x0 = blockIdx.z * TX, y0 = blockIdx.y * TY;
    const int x = x0 + ri, y = y0 + rj, z = zc0 + threadIdx.x;
    const bool valid = x < G.nx && y < G.ny && z < G.nz;
    const int xc = min(x, G.nx - 1), yc = min(y, G.ny - 1), zc = min(z, G.nz - 1);
    float fal[3], fce[3];
    {
        float f[3][3];
#pragma unroll
        for (int a = 0; a < 3; a++) face_pos<TWOD>(G, a, xc, yc, zc, f[a]);
        fal[0] = f[0][0], fal[1] = f[1][1], fal[2] = f[2][2];
        fce[0] = f[1][0], fce[1] = f[0][1], fce[2] = f[0][2];
    }
    const int cc[3] = {xc, yc, zc};
    float swv[3] = {0.f, 0.f, 0.f}, sw[3] = {0.f, 0.f, 0.f};
    float dens = 0.f;
    for (int rx = x0 - 1; rx <= x0 + TX; rx++) {
        if (rx < 0 || rx >= G.nx) continue;
        for (int ry = y0 - 1; ry <= y0 + TY; ry++) {
            if (ry < 0 || ry >= G.ny) continue;
            const int64_t rb = ((int64_t)rx * G.ny + ry) * G.nz;
            // the block's particle range: cells zc0 - 1 .. zc0 + 64 of the row (clamped)
            const int zlo = max(zc0 - 1, 0), zhi = min(zc0 + P2G_TZ, G.nz - 1);
            const int u0 = A.start[rb + zlo], u1 = A.start[rb + zhi + 1];
            if (u0 >= u1) continue;   // block-uniform
            // this wave works on the row if it is in its cell row's neighbourhood (wave-uniform)
            const bool near = valid && rx >= x - 1 && rx <= x + 1 && ry >= y - 1 && ry <= y + 1;
            int j0 = 0, j1 = 0;
            if (near) {
                j0 = A.start[rb + (z > 0 ? z - 1 : z)];
                j1 = A.start[rb + (z < G.nz - 1 ? z + 2 : z + 1)];
            }
            for (int base = u0; base < u1; base += CAP) {
                const int cnt = min(CAP, u1 - base);
                __syncthreads();   // the previous chunk is consumed
                for (int k = tid; k < cnt; k += NT) {
                    const int j = base + k;
                    uint8_t h = 0;
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const float pv = A.pos[d][j];
                        const double v = (double)pv * cdinv_t<TWOD>(G, d);
                        const int t = (int)v;
                        s_pos[d][k] = pv;
                        s_vel[d][k] = A.vel[d][j];
                        s_t[d][k] = t;
                        if ((int)(v - 0.5) == t) h |= (uint8_t)(1u << d);
                    }
                    s_h[k] = h;
                    if (APIC) {
#pragma unroll
                        for (int m = 0; m < NC; m++) s_c[m][k] = A.c[m][j];
                    }
                }
                __syncthreads();
                const int e = min(j1, base + cnt);
#pragma unroll 4
                for (int j = max(j0, base); j < e; j++) {
                    const int k = j - base;
                    const float p[3] = {s_pos[0][k], s_pos[1][k], s_pos[2][k]};
                    const uint32_t h = s_h[k];
                    int ua[3], uc[3];
                    float tal[3], tce[3];
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        ua[d] = cc[d] - s_t[d][k] + 1;
                        uc[d] = ua[d] - (int)((h >> d) & 1u);
                        tal[d] = 1.0f - fabsf((p[d] - fal[d]) * invf_t<TWOD>(G, d));
                        tce[d] = 1.0f - fabsf((p[d] - fce[d]) * invf_t<TWOD>(G, d));
                    }
                    if ((unsigned)uc[0] <= 1u && (unsigned)uc[1] <= 1u && (unsigned)uc[2] <= 1u)
                        dens += tce[0] * tce[1] * tce[2];
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        bool in = true;
                        float tw[3];
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            const int u = d == a ? ua[d] : uc[d];
                            in = in && (unsigned)u <= 1u;
                            tw[d] = d == a ? tal[d] : tce[d];
                        }
                        if (!in) continue;
                        const float w = tw[0] * tw[1] * tw[2];
                        float va = s_vel[a][k];
                        if (APIC) {
                            const float f0 = a == 0 ? fal[0] : fce[0], f1 = a == 1 ? fal[1] : fce[1],
                                        f2 = a == 2 ? fal[2] : fce[2];
                            const float d0 = f0 - p[0], d1 = f1 - p[1], d2 = f2 - p[2];
                            va = va + (s_c[3 * a][k] * d0 + s_c[3 * a + 1][k] * d1 + s_c[3 * a + 2][k] * d2);
                        }
                        swv[a] += va * w;
                        sw[a] += w;
                    }
                }
            }
        }
    }
    if (!valid) return;
    const int64_t i = cidx(G, x, y, z);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        A.uvw[a][i] = sw[a] > 1e-6f ? swv[a] / sw[a] : 0.0f;
        A.uvw2[a][i] = 0.0f;
    }
    A.label[i] = (A.fuse_mark && A.count[i] > 0) ? (uint8_t)FSIM_WATER : (uint8_t)FSIM_AIR;
    A.density[i] = dens;
}

// The tiled gather for PIC/FLIP with the x and y weights computed at staging.
// Every cell of a wave has the same (x, y), and a staged row holds the
// particles of one (x, y) cell row, so a particle's x weights (al_x, ce_x)
// for the three cells x = row - 1 .. row + 1, and its y weights likewise, are
// all the walk needs of its x and y coordinates: the staging computes them
// once (face_pos's face coordinates, 0 outside the stencil) and each wave
// reads the pair for its own offset.  The walk computes only the z weights,
// and p2g_terms sums without branches.  Same rows, same particle order, same
// weights and terms as the row gather: the sums are bit-identical.
// The walk is latency-bound, so occupancy decides: 36 KB of LDS per block
// allows 8 waves per SIMD, and the register budget is held to that (60
// VGPRs with the walk not unrolled).  Measured at 256^3 (FLIP double dam
// break): 3.75 ms, against 4.1 ms at 69 VGPRs, 4.6-4.8 ms unrolled by 2 and
// 5.6 ms unrolled by 4 (94 VGPRs); 4 x 4 tiles took 4.9 ms.
constexpr int P2G_W_CAP = 512;   // particles per staged chunk (72 bytes of LDS each)
template <int TX, int TY, int TWOD>
__global__ void __launch_bounds__(P2G_TZ * TX * TY) __attribute__((amdgpu_waves_per_eu(8)))
k_p2g_tile_w(P2GArgs A) {
    constexpr int CAP = P2G_W_CAP;
    constexpr int NT = P2G_TZ * TX * TY;
    __shared__ float4 s_pv[CAP];      // z position, velocity x, y, z
    __shared__ int2 s_zz[CAP];        // t_z - 1, t_z - 1 + h_z (face_base along z of the two face kinds)
    __shared__ float2 s_wx[3][CAP];   // [x - row + 1]: (al_x, ce_x)
    __shared__ float2 s_wy[3][CAP];   // [y - row + 1]: (al_y, ce_y)
    const GridArgs& G = A.G;
    const int tid = threadIdx.x + P2G_TZ * threadIdx.y;
    const int zc0 = blockIdx.x * P2G_TZ;
    const int ri = threadIdx.y / TY, rj = threadIdx.y % TY;
    const int x0 = blockIdx.z * TX, y0 = blockIdx.y * TY;
    const int x = x0 + ri, y = y0 + rj, z = zc0 + threadIdx.x;
    const bool valid = x < G.nx && y < G.ny && z < G.nz;
    const int zc = min(z, G.nz - 1);
    float fz[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++) face_pos<TWOD>(G, a, 0, 0, zc, fz[a]);
    const float falz = fz[2][2], fcez = fz[0][2];
    const float invz = invf_t<TWOD>(G, 2);
    float swv[3] = {0.f, 0.f, 0.f}, sw[3] = {0.f, 0.f, 0.f};
    float dens = 0.f;
    for (int rx = x0 - 1; rx <= x0 + TX; rx++) {
        if (rx < 0 || rx >= G.nx) continue;
        for (int ry = y0 - 1; ry <= y0 + TY; ry++) {
            if (ry < 0 || ry >= G.ny) continue;
            const int64_t rb = ((int64_t)rx * G.ny + ry) * G.nz;
            const int zlo = max(zc0 - 1, 0), zhi = min(zc0 + P2G_TZ, G.nz - 1);
            const int u0 = A.start[rb + zlo], u1 = A.start[rb + zhi + 1];
            if (u0 >= u1) continue;   // block-uniform
            const bool near = valid && rx >= x - 1 && rx <= x + 1 && ry >= y - 1 && ry <= y + 1;
            int j0 = 0, j1 = 0;
            if (near) {
                j0 = A.start[rb + (z > 0 ? z - 1 : z)];
                j1 = A.start[rb + (z < G.nz - 1 ? z + 2 : z + 1)];
            }
            const int ex = near ? x - rx + 1 : 0, ey = near ? y - ry + 1 : 0;   // wave-uniform
            for (int base = u0; base < u1; base += CAP) {
                const int cnt = min(CAP, u1 - base);
                __syncthreads();   // the previous chunk is consumed
                for (int k = tid; k < cnt; k += NT) {
                    const int j = base + k;
                    const float p[3] = {A.pos[0][j], A.pos[1][j], A.pos[2][j]};
                    int t[3], h[3];
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const double v = (double)p[d] * cdinv_t<TWOD>(G, d);
                        t[d] = (int)v;
                        h[d] = (int)(v - 0.5) == t[d] ? 1 : 0;
                    }
                    const int row[2] = {rx, ry};
#pragma unroll
                    for (int d = 0; d < 2; d++) {
#pragma unroll
                        for (int e = 0; e < 3; e++) {
                            const int cellc = row[d] - 1 + e;   // the wave's cell coordinate on axis d
                            const int ua = cellc - t[d] + 1, uc = ua - h[d];
                            const float fal = (float)((cellc + 1.0) * G.cd), fce = (float)((cellc + 0.5) * G.cd);
                            const float al = (unsigned)ua <= 1u ? 1.0f - fabsf((p[d] - fal) * G.invf) : 0.f;
                            const float ce = (unsigned)uc <= 1u ? 1.0f - fabsf((p[d] - fce) * G.invf) : 0.f;
                            (d == 0 ? s_wx : s_wy)[e][k] = make_float2(al, ce);
                        }
                    }
                    s_zz[k] = make_int2(t[2] - 1, t[2] - 1 + h[2]);
                    s_pv[k] = make_float4(p[2], A.vel[0][j], A.vel[1][j], A.vel[2][j]);
                }
                __syncthreads();
                const int e = min(j1, base + cnt);
#pragma unroll 1
                for (int j = max(j0, base); j < e; j++) {
                    const int k = j - base;
                    const float4 pv = s_pv[k];
                    const int2 zz = s_zz[k];
                    const float2 wx = s_wx[ex][k], wy = s_wy[ey][k];
                    const int ua = zc - zz.x, uc = zc - zz.y;
                    const float alz = (unsigned)ua <= 1u ? 1.0f - fabsf((pv.x - falz) * invz) : 0.f;
                    const float cez = (unsigned)uc <= 1u ? 1.0f - fabsf((pv.x - fcez) * invz) : 0.f;
                    p2g_terms(wx.x, wx.y, wy.x, wy.y, alz, cez, pv.y, pv.z, pv.w, swv, sw, dens);
                }
            }
        }
    }
    if (!valid) return;
    const int64_t i = cidx(G, x, y, z);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        A.uvw[a][i] = sw[a] > 1e-6f ? swv[a] / sw[a] : 0.0f;
        A.uvw2[a][i] = 0.0f;
    }
    A.label[i] = (A.fuse_mark && A.count[i] > 0) ? (uint8_t)FSIM_WATER : (uint8_t)FSIM_AIR;
    A.density[i] = dens;
}

static int p2g(fsim_ctx* ctx, int fuse_mark) {
    if (ctx->prm.transfer == FSIM_APIC && !ctx->c_valid) {
        // APIC after PIC/FLIP steps: the matrices were not carried, start from 0
        FSIM_TRY(fill_f32(ctx, ctx->P.c[0], 0.0f, (int64_t)ctx->P.cap * 9));
        ctx->c_valid = true;
    }
    FSIM_PROF(ctx, KID_P2G);
    P2GArgs A{};
    A.G = grid_args(ctx);
    for (int a = 0; a < 3; a++) {
        A.pos[a] = ctx->P.pos[a];
        A.vel[a] = ctx->P.vel[a];
    }
    for (int k = 0; k < 9; k++) A.c[k] = ctx->P.c[k];
    A.start = ctx->cell_start;
    A.count = ctx->cell_count;
    A.uvw[0] = ctx->u; A.uvw[1] = ctx->v; A.uvw[2] = ctx->w;
    A.uvw2[0] = ctx->u2; A.uvw2[1] = ctx->v2; A.uvw2[2] = ctx->w2;
    A.label = ctx->label;
    A.density = ctx->density;
    A.apic = ctx->prm.transfer == FSIM_APIC;
    A.fuse_mark = fuse_mark;
    if (ctx->p2g_rows) {   // FSIM_OPT_P2G_KERNEL = 1: one cell row per block
        const dim3 g(nblocks(ctx->g.nc)), b(kBlock);
        if (ctx->g.two_d) {
            if (A.apic) hipLaunchKernelGGL((k_p2g_gather<1, 1>), g, b, 0, ctx->stream, A);
            else hipLaunchKernelGGL((k_p2g_gather<0, 1>), g, b, 0, ctx->stream, A);
        } else {
            if (A.apic) hipLaunchKernelGGL((k_p2g_gather<1, 0>), g, b, 0, ctx->stream, A);
            else hipLaunchKernelGGL((k_p2g_gather<0, 0>), g, b, 0, ctx->stream, A);
        }
    } else {
        const dim3 g((unsigned)((ctx->g.nz + P2G_TZ - 1) / P2G_TZ), (unsigned)((ctx->g.ny + P2G_TY - 1) / P2G_TY),
                     (unsigned)((ctx->g.nx + P2G_TX - 1) / P2G_TX));
        const dim3 b(P2G_TZ, P2G_TX * P2G_TY);
        if (ctx->g.two_d) {
            if (A.apic) hipLaunchKernelGGL((k_p2g_tile<1, P2G_TX, P2G_TY, 1>), g, b, 0, ctx->stream, A);
            else hipLaunchKernelGGL((k_p2g_tile_w<P2G_TX, P2G_TY, 1>), g, b, 0, ctx->stream, A);
        } else {
            if (A.apic) hipLaunchKernelGGL((k_p2g_tile<1, P2G_TX, P2G_TY, 0>), g, b, 0, ctx->stream, A);
            else hipLaunchKernelGGL((k_p2g_tile_w<P2G_TX, P2G_TY, 0>), g, b, 0, ctx->stream, A);
        }
    }
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

int stage_p2g(fsim_ctx* ctx) { return p2g(ctx, 0); }
int stage_p2g_fused(fsim_ctx* ctx) { return p2g(ctx, 1); }

// ---- markFluidCellsAndCalculateParticleDensities (simulator.cpp:354-368) --------
// WATER where the cell holds a particle (the bin key is the same
// ivec3(pos*cellDInv)).  The density half of the reference function is summed
// by the P2G gather above, which walks the same particles in the same order.
__global__ void __launch_bounds__(kBlock) k_mark(const int32_t* __restrict__ count, uint8_t* __restrict__ label,
                                                 int64_t nc) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    if (count[i] > 0) label[i] = FSIM_WATER;
}

int stage_mark(fsim_ctx* ctx) {
    FSIM_PROF(ctx, KID_MARK);
    hipLaunchKernelGGL(k_mark, dim3(nblocks(ctx->g.nc)), dim3(kBlock), 0, ctx->stream, ctx->cell_count, ctx->label,
                       ctx->g.nc);
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

// ---- addObstaclesToGrid / MacGrid::addObstacle (macGrid.cpp:231-292) --------------
// The reference visits the obstacle's bounding box serially in x,y,z order and
// sets the face velocity towards each WATER neighbour, reading the labels as
// they are at that moment: a +neighbour is visited later, so it still has its
// label from before this obstacle; a -neighbour inside this obstacle's box
// and shape has already been made SOLID.  The face pass below reads the
// labels from before the obstacle and applies exactly that rule; the label
// pass runs after it.
struct ObsArgs {
    GridArgs G;
    int mn[3], ext[3];       // first cell and extent of the visited box
    int sphere;
    double ctr[3], r2;
    float speed[3];
    uint8_t* label;
    float* uvw[3];
};

__device__ inline bool obs_inside(const ObsArgs& A, int x, int y, int z) {
#pragma clang fp contract(off)
    if (x < A.mn[0] || y < A.mn[1] || z < A.mn[2] || x >= A.mn[0] + A.ext[0] || y >= A.mn[1] + A.ext[1] ||
        z >= A.mn[2] + A.ext[2])
        return false;
    if (!A.sphere) return true;
    double c0 = x + 0.5 - A.ctr[0], c1 = y + 0.5 - A.ctr[1], c2 = z + 0.5 - A.ctr[2];
    return c0 * c0 + c1 * c1 + c2 * c2 < A.r2;
}

__global__ void __launch_bounds__(kBlock) k_obstacle_faces(ObsArgs A) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t vol = (int64_t)A.ext[0] * A.ext[1] * A.ext[2];
    if (t >= vol) return;
    int z = A.mn[2] + (int)(t % A.ext[2]);
    int y = A.mn[1] + (int)((t / A.ext[2]) % A.ext[1]);
    int x = A.mn[0] + (int)(t / ((int64_t)A.ext[2] * A.ext[1]));
    if (!obs_inside(A, x, y, z)) return;
    const GridArgs& G = A.G;
    int64_t i = cidx(G, x, y, z);
    const int64_t st[3] = {(int64_t)G.ny * G.nz, G.nz, 1};
    const int p[3] = {x, y, z};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (A.label[i + st[a]] == FSIM_WATER) A.uvw[a][i] = A.speed[a];
        int m[3] = {p[0], p[1], p[2]};
        m[a] -= 1;
        if (A.label[i - st[a]] == FSIM_WATER && !obs_inside(A, m[0], m[1], m[2])) A.uvw[a][i - st[a]] = A.speed[a];
    }
}

__global__ void __launch_bounds__(kBlock) k_obstacle_labels(ObsArgs A) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t vol = (int64_t)A.ext[0] * A.ext[1] * A.ext[2];
    if (t >= vol) return;
    int z = A.mn[2] + (int)(t % A.ext[2]);
    int y = A.mn[1] + (int)((t / A.ext[2]) % A.ext[1]);
    int x = A.mn[0] + (int)(t / ((int64_t)A.ext[2] * A.ext[1]));
    if (obs_inside(A, x, y, z)) A.label[cidx(A.G, x, y, z)] = FSIM_SOLID;
}

int stage_obstacles(fsim_ctx* ctx) {
    FSIM_PROF(ctx, KID_OBSTACLES);
    const int n[3] = {ctx->g.nx, ctx->g.ny, ctx->g.nz};
    for (const fsim_obstacle& o : ctx->obs) {
        ObsArgs A{};
        A.G = grid_args(ctx);
        A.label = ctx->label;
        A.uvw[0] = ctx->u; A.uvw[1] = ctx->v; A.uvw[2] = ctx->w;
        for (int a = 0; a < 3; a++) {
            A.ctr[a] = o.pos[a] * (a == 2 ? ctx->g.cdzinv : ctx->g.cdinv);
            A.speed[a] = (float)o.speed[a];
        }
        int mx[3];
        if (o.kind == FSIM_OBS_SPHERE || o.kind == FSIM_OBS_SOURCE) {
            double r = ctx->g.cdinv * o.r;
            A.r2 = r * r;
            A.sphere = 1;
            for (int a = 0; a < 3; a++) {
                A.mn[a] = (int)std::fmax(A.ctr[a] - r, 1.0);
                mx[a] = (int)std::fmin(A.ctr[a] + r, n[a] - 2.0);   // inclusive
            }
        } else if (o.kind == FSIM_OBS_BOX) {
            A.sphere = 0;
            for (int a = 0; a < 3; a++) {
                double d = o.size[a] * (a == 2 ? ctx->g.cdzinv : ctx->g.cdinv) * 0.5;
                A.mn[a] = (int)std::fmax(std::round(A.ctr[a] - d), 1.0);
                mx[a] = (int)std::fmin(std::round(A.ctr[a] + d), n[a] - 1.0) - 1;   // exclusive -> inclusive
            }
        } else {
            continue;   // sinks are not voxelised
        }
        int64_t vol = 1;
        for (int a = 0; a < 3; a++) {
            A.ext[a] = mx[a] - A.mn[a] + 1;
            if (A.ext[a] <= 0) vol = 0;
        }
        if (vol == 0) continue;
        vol = (int64_t)A.ext[0] * A.ext[1] * A.ext[2];
        hipLaunchKernelGGL(k_obstacle_faces, dim3(nblocks(vol)), dim3(kBlock), 0, ctx->stream, A);
        hipLaunchKernelGGL(k_obstacle_labels, dim3(nblocks(vol)), dim3(kBlock), 0, ctx->stream, A);
        FSIM_HIP(ctx, hipGetLastError());
    }
    return FSIM_OK;
}

// ---- restoreBorderingSolidCellsAndSpeeds (macGrid.cpp:75-140) ---------------------
// Face pass with the labels from before the pass, reproducing which border
// cells the reference has already made SOLID when it tests a neighbour:
// x-walls first; then, per x, the y-walls (and top) before the z-walls.
struct BorderArgs {
    GridArgs G;
    const uint8_t* label;
    float* uvw[3];
    int top;
};

__global__ void __launch_bounds__(kBlock) k_border_faces(BorderArgs A) {
    const GridArgs& G = A.G;
    const int X = G.nx, Y = G.ny, Z = G.nz;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int wall = blockIdx.y;
    if (wall == 0) {   // x-walls, loop over (y, z): labels as before the pass
        if (t >= (int64_t)Y * Z) return;
        int y = (int)(t / Z), z = (int)(t % Z);
        if (A.label[cidx(G, 1, y, z)] == FSIM_WATER) A.uvw[0][cidx(G, 0, y, z)] = 0.f;
        if (A.label[cidx(G, X - 2, y, z)] == FSIM_WATER) A.uvw[0][cidx(G, X - 2, y, z)] = 0.f;
    } else if (wall == 1) {   // bottom (and top) y-walls, loop over (x, z)
        if (t >= (int64_t)X * Z) return;
        int x = (int)(t / Z), z = (int)(t % Z);
        const bool xb = x == 0 || x == X - 1;
        uint8_t l1 = xb ? (uint8_t)FSIM_SOLID : A.label[cidx(G, x, 1, z)];
        if (l1 == FSIM_WATER) A.uvw[1][cidx(G, x, 0, z)] = 0.f;
        if (A.top) {
            uint8_t l2 = xb ? (uint8_t)FSIM_SOLID : A.label[cidx(G, x, Y - 2, z)];
            if (l2 == FSIM_WATER) A.uvw[1][cidx(G, x, Y - 2, z)] = 0.f;
        }
    } else {   // z-walls, loop over (x, y) after this x's y-walls
        if (t >= (int64_t)X * Y) return;
        int x = (int)(t / Y), y = (int)(t % Y);
        const bool done = x == 0 || x == X - 1 || y == 0 || (A.top && y == Y - 1);
        uint8_t l1 = done ? (uint8_t)FSIM_SOLID : A.label[cidx(G, x, y, 1)];
        uint8_t l2 = done ? (uint8_t)FSIM_SOLID : A.label[cidx(G, x, y, Z - 2)];
        if (l1 == FSIM_WATER) A.uvw[2][cidx(G, x, y, 0)] = 0.f;
        if (l2 == FSIM_WATER) A.uvw[2][cidx(G, x, y, Z - 2)] = 0.f;
    }
}

__global__ void __launch_bounds__(kBlock) k_border_labels(GridArgs G, uint8_t* label, int top) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int X = G.nx, Y = G.ny, Z = G.nz;
    const int wall = blockIdx.y;
    if (wall == 0) {
        if (t >= (int64_t)Y * Z) return;
        int y = (int)(t / Z), z = (int)(t % Z);
        label[cidx(G, 0, y, z)] = FSIM_SOLID;
        label[cidx(G, X - 1, y, z)] = FSIM_SOLID;
    } else if (wall == 1) {
        if (t >= (int64_t)X * Z) return;
        int x = (int)(t / Z), z = (int)(t % Z);
        label[cidx(G, x, 0, z)] = FSIM_SOLID;
        if (top) label[cidx(G, x, Y - 1, z)] = FSIM_SOLID;
    } else {
        if (t >= (int64_t)X * Y) return;
        int x = (int)(t / Y), y = (int)(t % Y);
        label[cidx(G, x, y, 0)] = FSIM_SOLID;
        label[cidx(G, x, y, Z - 1)] = FSIM_SOLID;
    }
}

int stage_borders(fsim_ctx* ctx) {
    FSIM_PROF(ctx, KID_BORDERS);
    BorderArgs A{};
    A.G = grid_args(ctx);
    A.label = ctx->label;
    A.uvw[0] = ctx->u; A.uvw[1] = ctx->v; A.uvw[2] = ctx->w;
    A.top = ctx->prm.top_solid ? 1 : 0;
    int64_t m = std::max({(int64_t)ctx->g.ny * ctx->g.nz, (int64_t)ctx->g.nx * ctx->g.nz, (int64_t)ctx->g.nx * ctx->g.ny});
    hipLaunchKernelGGL(k_border_faces, dim3(nblocks(m), 3), dim3(kBlock), 0, ctx->stream, A);
    hipLaunchKernelGGL(k_border_labels, dim3(nblocks(m), 3), dim3(kBlock), 0, ctx->stream, A.G, ctx->label, A.top);
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

// ---- postP2GUpdate (macGrid.cpp:204-219) -----------------------------------------
// v2 = v, plus gravity on the y-face when the cell and cell<1,1> are not
// SOLID.  cell<1,1> is the flat index +Nz, which for the top row is the
// SOLID border cell (x+1, 0, z): top-row y-faces get no gravity.
__global__ void __launch_bounds__(kBlock) k_post_p2g(GridArgs G, const float* __restrict__ u, const float* __restrict__ v,
                                                     const float* __restrict__ w, float* __restrict__ u2,
                                                     float* __restrict__ v2, float* __restrict__ w2,
                                                     const uint8_t* __restrict__ label, float g) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G.nc) return;
    int64_t nb = i + G.nz;
    bool grav = label[i] != FSIM_SOLID && nb < G.nc && label[nb] != FSIM_SOLID;
    u2[i] = u[i];
    v2[i] = grav ? v[i] + g : v[i];
    w2[i] = w[i];
}

int stage_post_p2g(fsim_ctx* ctx, double grav_inc) {
    FSIM_PROF(ctx, KID_POST_P2G);
    hipLaunchKernelGGL(k_post_p2g, dim3(nblocks(ctx->g.nc)), dim3(kBlock), 0, ctx->stream, grid_args(ctx), ctx->u,
                       ctx->v, ctx->w, ctx->u2, ctx->v2, ctx->w2, ctx->label, (float)grav_inc);
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

// ---- extrapolateVelocities (macGrid.cpp:294-336) -----------------------------------
// Two layers.  A cell that is not yet valid takes, for each of its three
// faces, the mean of the same-component face of its valid neighbours (a face
// is kept when its + neighbour is WATER or outside the grid).  Within one
// layer the cells written are not valid and the cells read are valid, so the
// pass is race-free in place.
__global__ void __launch_bounds__(kBlock) k_valid_init(GridArgs G, const uint8_t* __restrict__ label, uint8_t* __restrict__ valid) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G.nc) return;
    const int z = (int)(i % G.nz);
    const int y = (int)((i / G.nz) % G.ny);
    const int x = (int)(i / ((int64_t)G.nz * G.ny));
    bool interior = x > 0 && y > 0 && z > 0 && x < G.nx - 1 && y < G.ny - 1 && z < G.nz - 1;
    valid[i] = (interior && label[i] == FSIM_WATER) ? 0 : 100;
}

__global__ void __launch_bounds__(kBlock) k_extrapolate(GridArgs G, float* __restrict__ u2, float* __restrict__ v2,
                                                        float* __restrict__ w2, const uint8_t* __restrict__ label,
                                                        uint8_t* __restrict__ valid, int it) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G.nc) return;
    if (valid[i] <= it) return;
    const int z = (int)(i % G.nz);
    const int y = (int)((i / G.nz) % G.ny);
    const int x = (int)(i / ((int64_t)G.nz * G.ny));
    const int64_t st[3] = {(int64_t)G.ny * G.nz, G.nz, 1};
    const int p[3] = {x, y, z};
    const int n[3] = {G.nx, G.ny, G.nz};
    int cnt = 0;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int off = 0; off < 2; off++) {
            int q = p[a] + off * 2 - 1;
            if (q < 0 || q >= n[a]) continue;
            int64_t j = i + (off ? st[a] : -st[a]);
            if (valid[j] > it) continue;
            s0 += u2[j];
            s1 += v2[j];
            s2 += w2[j];
            cnt++;
        }
    if (cnt > 0) {
        float fc = (float)cnt;
        if (x + 1 < G.nx && label[i + st[0]] != FSIM_WATER) u2[i] = s0 / fc;
        if (y + 1 < G.ny && label[i + st[1]] != FSIM_WATER) v2[i] = s1 / fc;
        if (z + 1 < G.nz && label[i + st[2]] != FSIM_WATER) w2[i] = s2 / fc;
        valid[i] = (uint8_t)(it + 1);
    }
}

int stage_extrapolate(fsim_ctx* ctx) {
    FSIM_PROF(ctx, KID_EXTRAPOLATE);
    GridArgs G = grid_args(ctx);
    hipLaunchKernelGGL(k_valid_init, dim3(nblocks(G.nc)), dim3(kBlock), 0, ctx->stream, G, ctx->label, ctx->valid);
    for (int it = 0; it < 2; it++)
        hipLaunchKernelGGL(k_extrapolate, dim3(nblocks(G.nc)), dim3(kBlock), 0, ctx->stream, G, ctx->u2, ctx->v2, ctx->w2,
                           ctx->label, ctx->valid, it);
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

// ---- g2pTransfer (simulator.cpp:375-418) --------------------------------------------
struct G2PArgs {
    GridArgs G;
    float* pos[3];
    float* vel[3];
    float* c[9];
    const float* uvw[3];
    const float* uvw2[3];
    int64_t n;
    int transfer;
    float flip, one_minus_flip;
};

// TT: the transfer type, a template argument so that the 8-corner loops have
// no branches and all their loads are issued before the first is used.
template <int TT>
__global__ void __launch_bounds__(kBlock) k_g2p(G2PArgs A) {
    const GridArgs& G = A.G;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const float px = A.pos[0][i], py = A.pos[1][i], pz = A.pos[2][i];
    const float inv = G.invf, invz = G.invfz;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (a == 2 && G.two_d) {   // 2D: v.z = 0 and no z row (simulator.cpp:382-385)
            A.vel[2][i] = 0.0f;
            break;
        }
        int b[3];
        face_base(G, px, py, pz, a, b);
        // stay inside the grid for positions outside the clamp box (the step
        // never produces them; the reference would read out of bounds)
        b[0] = min(max(b[0], 0), G.nx - 2);
        b[1] = min(max(b[1], 0), G.ny - 2);
        b[2] = min(max(b[2], 0), G.nz - 2);
        // tent factors of the two face rows per dimension
        float tx[2], ty[2], tz[2], dxs[2], dys[2], dzs[2];
#pragma unroll
        for (int o = 0; o < 2; o++) {
            float fx = (float)((b[0] + o + (a == 0 ? 1.0 : 0.5)) * G.cd);
            float fy = (float)((b[1] + o + (a == 1 ? 1.0 : 0.5)) * G.cd);
            float fz = (float)((b[2] + o + (a == 2 ? 1.0 : 0.5)) * G.cdz);
            dxs[o] = (px - fx) * inv;
            dys[o] = (py - fy) * inv;
            dzs[o] = (pz - fz) * invz;
            tx[o] = 1.0f - fabsf(dxs[o]);
            ty[o] = 1.0f - fabsf(dys[o]);
            tz[o] = 1.0f - fabsf(dzs[o]);
        }
        float pic = 0.f, flip = 0.f, g0 = 0.f, g1 = 0.f, g2 = 0.f;
#pragma unroll
        for (int o = 0; o < 8; o++) {
            const int ox = (o >> 2) & 1, oy = (o >> 1) & 1, oz = o & 1;
            int64_t f = cidx(G, b[0] + ox, b[1] + oy, b[2] + oz);
            float w = tx[ox] * ty[oy] * tz[oz];
            float v2 = A.uvw2[a][f];
            pic += v2 * w;
            if constexpr (TT == FSIM_FLIP) flip += (v2 - A.uvw[a][f]) * w;
            if constexpr (TT == FSIM_APIC) {
                // trilinearInterpollGradient, interpolation.h:23-35 (sign(0) = +1)
                float sx = dxs[ox] > 0.f ? -1.f : 1.f, sy = dys[oy] > 0.f ? -1.f : 1.f, sz = dzs[oz] > 0.f ? -1.f : 1.f;
                g0 += sx * ty[oy] * tz[oz] * inv * v2;
                g1 += sy * tx[ox] * tz[oz] * inv * v2;
                g2 += sz * tx[ox] * ty[oy] * invz * v2;
            }
        }
        if constexpr (TT == FSIM_PIC) {
            A.vel[a][i] = pic;
        } else if constexpr (TT == FSIM_FLIP) {
            A.vel[a][i] = pic * A.one_minus_flip + (flip + A.vel[a][i]) * A.flip;
        } else {
            A.vel[a][i] = pic;
            A.c[3 * a][i] = g0;
            A.c[3 * a + 1][i] = g1;
            A.c[3 * a + 2][i] = g2;
        }
    }
}

int stage_g2p(fsim_ctx* ctx) {
    if (ctx->P.n == 0) return FSIM_OK;
    G2PArgs A{};
    A.G = grid_args(ctx);
    for (int a = 0; a < 3; a++) {
        A.pos[a] = ctx->P.pos[a];
        A.vel[a] = ctx->P.vel[a];
    }
    for (int k = 0; k < 9; k++) A.c[k] = ctx->P.c[k];
    A.uvw[0] = ctx->u; A.uvw[1] = ctx->v; A.uvw[2] = ctx->w;
    A.uvw2[0] = ctx->u2; A.uvw2[1] = ctx->v2; A.uvw2[2] = ctx->w2;
    A.n = ctx->P.n;
    A.transfer = ctx->prm.transfer;
    A.flip = ctx->prm.flip_ratio;
    A.one_minus_flip = 1.0f - ctx->prm.flip_ratio;
    if (A.transfer == FSIM_APIC) ctx->c_valid = true;
    FSIM_PROF(ctx, KID_G2P);
    if (A.transfer == FSIM_PIC) hipLaunchKernelGGL(k_g2p<FSIM_PIC>, dim3(nblocks(A.n)), dim3(kBlock), 0, ctx->stream, A);
    else if (A.transfer == FSIM_FLIP) hipLaunchKernelGGL(k_g2p<FSIM_FLIP>, dim3(nblocks(A.n)), dim3(kBlock), 0, ctx->stream, A);
    else hipLaunchKernelGGL(k_g2p<FSIM_APIC>, dim3(nblocks(A.n)), dim3(kBlock), 0, ctx->stream, A);
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

}  // namespace fsim
