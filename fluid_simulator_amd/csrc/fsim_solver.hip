// fsim_solver.hip — Bridson pressure projection on gfx950.
//
// BridsonSolverGrid::solveIncompressibility (bridsonSolverGrid.cpp:244-293)
// over dense vectors with a WATER mask:
//   * b = -(1/dx) div(v2) [+ density term] evaluated in FP64, stored FP32;
//   * A is matrix-free: diag = scale * #non-SOLID neighbours, off-diagonal
//     -scale towards each WATER neighbour (calculateAMatrix, :40-77);
//   * MIC(0) factor (calculatePreconditioner, :91-124) in FP64 with the
//     reference's expression order, level-scheduled on i+j+k planes: every
//     cell depends only on its -x/-y/-z neighbours, so any topological order
//     gives the serial loop's values;
//   * MIC(0) apply (:126-163): forward/backward triangular sweeps in FP32,
//     level-scheduled the same way;
//   * PCG with FP32 vectors, FP64 pressure, FP64 dot products (deterministic
//     block partials summed in a fixed order) and FP64 reliable updates: when
//     the recursive residual meets the tolerance, the true residual b - A x is
//     recomputed in FP64; the solve stops only if that meets the tolerance,
//     otherwise r is replaced by it and the iteration continues.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "fsim_internal.hpp"

namespace fsim {

struct SolveArgs {
    int nx, ny, nz;
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
__device__ inline int pmod(int a, int n) {
    a %= n;
    return a < 0 ? a + n : a;
}
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
    n.pz = (rz == WF_TZ - 1 ? pmod(p - (WF_TZ - 1), S.nx) : pn) * plane + yz + 1;
    n.mz = (rz == 0 ? pmod(p + (WF_TZ - 1), S.nx) : pp) * plane + yz - 1;
    return n;
}
// rotated indices of the +a and -a neighbours of interior cell (x, y, z)
__device__ inline int64_t rnb_plus(const SolveArgs& S, int x, int y, int z, int a) {
    return rsidx(S, x + (a == 0), y + (a == 1), z + (a == 2));
}
__device__ inline int64_t rnb_minus(const SolveArgs& S, int x, int y, int z, int a) {
    return rsidx(S, x - (a == 0), y - (a == 1), z - (a == 2));
}

// ---- block reductions (FP64, fixed order) ------------------------------------
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_down(v, d, 64);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmax(v, __shfl_down(v, d, 64));
    return v;
}
// Writes the block's sum to partials[blockIdx.x] and, for aux_mode 1 (max)
// or 2 (sum), the reduced aux value to partials[gridDim.x + blockIdx.x].
constexpr int AUX_NONE = 0, AUX_MAX = 1, AUX_SUM = 2;
__device__ inline void block_reduce_store(double sum, double mx, double* partials, int aux_mode) {
    __shared__ double lds_s[kBlock / 64], lds_m[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool with_max = aux_mode != AUX_NONE;
    sum = wave_sum(sum);
    if (aux_mode == AUX_MAX) mx = wave_max(mx);
    if (aux_mode == AUX_SUM) mx = wave_sum(mx);
    if (lane == 0) {
        lds_s[wave] = sum;
        lds_m[wave] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0, m = 0;
        for (int w = 0; w < kBlock / 64; w++) {
            s += lds_s[w];
            m = aux_mode == AUX_SUM ? m + lds_m[w] : fmax(m, lds_m[w]);
        }
        partials[blockIdx.x] = s;
        if (with_max) partials[gridDim.x + blockIdx.x] = m;
    }
}

enum FinOp { FIN_BB = 0, FIN_SIGMA0 = 1, FIN_ALPHA = 2, FIN_RES = 3, FIN_BETA = 4, FIN_TRUE = 5 };

__global__ void __launch_bounds__(kBlock) k_finalize(const double* __restrict__ partials, int nparts, int op,
                                                     SolveScalars* __restrict__ sc) {
    __shared__ double lds_s[kBlock], lds_m[kBlock];
    double s = 0, m = 0;
    const bool with_max = op == FIN_RES || op == FIN_TRUE;
    const bool aux_sum = op == FIN_BB;
    for (int i = threadIdx.x; i < nparts; i += kBlock) {
        s += partials[i];
        if (with_max) m = fmax(m, partials[nparts + i]);
        if (aux_sum) m += partials[nparts + i];
    }
    lds_s[threadIdx.x] = s;
    lds_m[threadIdx.x] = m;
    __syncthreads();
    // pairwise tree over the threads' sums, fixed order
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            lds_s[threadIdx.x] += lds_s[threadIdx.x + off];
            lds_m[threadIdx.x] = aux_sum ? lds_m[threadIdx.x] + lds_m[threadIdx.x + off]
                                         : fmax(lds_m[threadIdx.x], lds_m[threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    s = lds_s[0];
    m = lds_m[0];
    switch (op) {
    case FIN_BB: sc->bb = s; sc->nfluid = m; break;
    case FIN_SIGMA0: sc->sigma = s; break;
    case FIN_ALPHA: {
        sc->sq = s;
        double alpha = sc->sigma / s;
        sc->alpha = alpha;
        if (alpha != alpha) sc->nan_flag |= 1;
        break;
    }
    case FIN_RES: sc->rr = s; sc->rmax = m; break;
    case FIN_BETA: {
        sc->zr = s;
        if (sc->sigma != sc->sigma) sc->nan_flag |= 2;   // the reference tests the old sigma (:285-286)
        sc->beta = s / sc->sigma;
        sc->sigma = s;
        break;
    }
    case FIN_TRUE: sc->tr_ss = s; sc->tr_max = m; break;
    }
}

// ---- RHS and neighbour masks (calculateRHS :79-89, calculateAMatrix :40-77) ------
struct RhsArgs {
    SolveArgs S;
    const uint8_t* label;
    const float* u2;
    const float* v2;
    const float* w2;
    const float* density;
    uint16_t* nbinfo;
    float* b;
    double inv_dx;
    double avg_p, pk;
    int pressure_enabled;
    double* partials;
};

__device__ inline bool is_fluid(const SolveArgs& S, const uint8_t* label, int x, int y, int z) {
    return x > 0 && y > 0 && z > 0 && x < S.nx - 1 && y < S.ny - 1 && z < S.nz - 1 && label[sidx(S, x, y, z)] == FSIM_WATER;
}

__global__ void __launch_bounds__(kBlock) k_nbinfo_rhs(RhsArgs A, int with_rhs) {
#pragma clang fp contract(off)
    const SolveArgs& S = A.S;
    double bsq = 0.0, nfl = 0.0;
    const int nrows = S.nx * S.ny;
    // one grid row (x, y) per block iteration, lanes along z
    for (int row = blockIdx.x; row < nrows; row += gridDim.x)
    for (int z = threadIdx.x; z < S.nz; z += blockDim.x) {
        const int x = row / S.ny, y = row - x * S.ny;
        const int64_t i = (int64_t)row * S.nz + z;
        uint16_t info = 0;
        float bv = 0.f;
        if (is_fluid(S, A.label, x, y, z)) {
            const int64_t st[3] = {(int64_t)S.ny * S.nz, S.nz, 1};
            info = NB_FLUID;
            nfl += 1.0;
            int nonsolid = 0;
            const int p[3] = {x, y, z};
            for (int a = 0; a < 3; a++) {
                int qp[3] = {p[0], p[1], p[2]}, qm[3] = {p[0], p[1], p[2]};
                qp[a] += 1;
                qm[a] -= 1;
                uint8_t lp = A.label[i + st[a]], lm = A.label[i - st[a]];
                if (lp != FSIM_SOLID) nonsolid++;
                if (lm != FSIM_SOLID) nonsolid++;
                if (is_fluid(S, A.label, qp[0], qp[1], qp[2])) info |= (uint16_t)(NB_PX << a);
                if (is_fluid(S, A.label, qm[0], qm[1], qm[2])) info |= (uint16_t)(NB_MX << a);
            }
            info |= (uint16_t)(nonsolid << NB_NONSOLID_SHIFT);
            // for each -a neighbour that is WATER: are its +(a+1)%3 and
            // +(a+2)%3 neighbours WATER (the MIC(0) "other" off-diagonals)
            for (int a = 0; a < 3; a++) {
                if (!(info & (NB_MX << a))) continue;
                for (int k = 0; k < 2; k++) {
                    const int b = (a + 1 + k) % 3;
                    int q[3] = {p[0], p[1], p[2]};
                    q[a] -= 1;
                    q[b] += 1;
                    if (is_fluid(S, A.label, q[0], q[1], q[2])) info |= (uint16_t)(1u << (NB_UP_SHIFT + 2 * a + k));
                }
            }
            if (with_rhs) {
                double dv = (double)A.u2[i] + (double)A.v2[i] + (double)A.w2[i] - (double)A.u2[i - st[0]] -
                            (double)A.v2[i - st[1]] - (double)A.w2[i - st[2]];
                double bd = -A.inv_dx * dv +
                            (A.pressure_enabled ? ((double)A.density[i] - A.avg_p) * A.pk : 0.0);
                bsq += bd * bd;
                bv = (float)bd;
            }
        }
        const int64_t ri = rsidx(S, x, y, z);
        A.nbinfo[ri] = info;
        if (with_rhs) A.b[ri] = bv;
    }
    if (with_rhs) block_reduce_store(bsq, nfl, A.partials, AUX_SUM);
}

// ---- MIC(0) factor, one i+j+k plane per launch (FP64) -----------------------------
struct PlaneArgs {
    SolveArgs S;
    const uint16_t* nbinfo;
    int level;
};

__device__ inline bool plane_cell(const PlaneArgs& A, int64_t t, int& x, int& y, int& z) {
    const SolveArgs& S = A.S;
    const int ny_in = S.ny - 2;
    x = 1 + (int)(t / ny_in);
    y = 1 + (int)(t % ny_in);
    z = A.level - x - y;
    return x < S.nx - 1 && z >= 1 && z < S.nz - 1;
}

__global__ void __launch_bounds__(kBlock) k_mic_factor_plane(PlaneArgs A, double* __restrict__ P64, float* __restrict__ P32) {
#pragma clang fp contract(off)
    const SolveArgs& S = A.S;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int x, y, z;
    if (!plane_cell(A, t, x, y, z)) return;
    int64_t i = rsidx(S, x, y, z);
    uint16_t info = A.nbinfo[i];
    if (!(info & NB_FLUID)) return;
    const double tau = 0.97, sigma = 0.25, scale = S.scale;
    double eneg = 0, eneg_tau = 0;
    for (int a = 0; a < 3; a++) {
        if (!(info & (NB_MX << a))) continue;
        int64_t j = rnb_minus(S, x, y, z, a);
        double pj = P64[j];
        double ap = -scale * pj;   // xWater of the -neighbour times its precon
        double o1 = (info & (1u << (NB_UP_SHIFT + 2 * a))) ? -scale : 0.0;
        double o2 = (info & (1u << (NB_UP_SHIFT + 2 * a + 1))) ? -scale : 0.0;
        eneg += ap * ap;
        eneg_tau += ap * (o1 + o2) * pj;
    }
    int cnt = (info >> NB_NONSOLID_SHIFT) & 7;
    double diag = 0;
    for (int k = 0; k < cnt; k++) diag += scale;
    double e = diag - eneg - eneg_tau * tau;
    if (e < sigma * diag) e = diag < 1e-6 ? 1.0 : diag;
    double p = 1.0 / sqrt(e);
    P64[i] = p;
    P32[i] = (float)p;
}

// (P32, nbinfo) packed per cell for the sweeps
__global__ void __launch_bounds__(kBlock) k_pack_w(int64_t n, const uint16_t* __restrict__ nbinfo,
                                                   const float* __restrict__ P32, uint64_t* __restrict__ W) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    W[i] = ((uint64_t)nbinfo[i] << 32) | __float_as_uint(P32[i]);
}

// ---- MIC(0) apply: forward and backward sweeps, one plane per launch (FP32) -------
// Contraction is off in both sweep implementations so that they round alike.
__global__ void __launch_bounds__(kBlock) k_fwd_plane(PlaneArgs A, const float* __restrict__ r, float* __restrict__ q,
                                                      const float* __restrict__ P) {
#pragma clang fp contract(off)
    const SolveArgs& S = A.S;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int x, y, z;
    if (!plane_cell(A, t, x, y, z)) return;
    int64_t i = rsidx(S, x, y, z);
    uint16_t info = A.nbinfo[i];
    if (!(info & NB_FLUID)) return;
    const float ms = -S.scale_f;
    float qneg = 0.f;
#pragma unroll
    for (int a = 0; a < 3; a++)
        if (info & (NB_MX << a)) {
            const int64_t j = rnb_minus(S, x, y, z, a);
            qneg += ms * (q[j] * P[j]);
        }
    q[i] = (r[i] - qneg) * P[i];
}

__global__ void __launch_bounds__(kBlock) k_bwd_plane(PlaneArgs A, const float* __restrict__ q, float* __restrict__ z,
                                                      const float* __restrict__ P) {
#pragma clang fp contract(off)
    const SolveArgs& S = A.S;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int x, y, zc;
    if (!plane_cell(A, t, x, y, zc)) return;
    int64_t i = rsidx(S, x, y, zc);
    uint16_t info = A.nbinfo[i];
    if (!(info & NB_FLUID)) return;
    const float ms = -S.scale_f;
    float tneg = 0.f;
#pragma unroll
    for (int a = 0; a < 3; a++)
        if (info & (NB_PX << a)) tneg += ms * z[rnb_plus(S, x, y, zc, a)];
    z[i] = (q[i] - tneg * P[i]) * P[i];
}

// ---- SpMV (applyAMatrix :165-198) with the s.q partial dot -------------------------
__global__ void __launch_bounds__(kBlock) k_spmv_dot(SolveArgs S, const uint16_t* __restrict__ nbinfo,
                                                     const float* __restrict__ s, float* __restrict__ q,
                                                     double* __restrict__ partials) {
    double d = 0.0;
    const int nrows = S.nx * S.ny;
    const float ms = -S.scale_f;
    // one rotated row (plane p, row y) per block iteration, lanes along z
    for (int row = blockIdx.x; row < nrows; row += gridDim.x) {
        const int p = row / S.ny, y = row - p * S.ny;
        for (int z = threadIdx.x; z < S.nz; z += blockDim.x) {
            const int64_t i = (int64_t)row * S.nz + z;
            const uint16_t info = nbinfo[i];
            float val = 0.f;
            if (info & NB_FLUID) {
                const RNeighbours n = rneighbours(S, i, p, y, z);
                const int cnt = (info >> NB_NONSOLID_SHIFT) & 7;
                val = (S.scale_f * (float)cnt) * s[i];
                if (info & NB_PX) val += ms * s[n.px];
                if (info & NB_PY) val += ms * s[n.py];
                if (info & NB_PZ) val += ms * s[n.pz];
                if (info & NB_MX) val += ms * s[n.mx];
                if (info & NB_MY) val += ms * s[n.my];
                if (info & NB_MZ) val += ms * s[n.mz];
                d += (double)s[i] * (double)val;
            }
            q[i] = val;
        }
    }
    block_reduce_store(d, 0.0, partials, AUX_NONE);
}

// x += alpha s (FP64), r -= alpha q, with ||r||^2 and max|r| partials
__global__ void __launch_bounds__(kBlock) k_update(int64_t n, const uint16_t* __restrict__ nbinfo, const float* __restrict__ s,
                                                   const float* __restrict__ q, double* __restrict__ x, float* __restrict__ r,
                                                   const SolveScalars* __restrict__ sc, double* __restrict__ partials) {
    double ss = 0.0, mx = 0.0;
    const double alpha = sc->alpha;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (!(nbinfo[i] & NB_FLUID)) continue;
        x[i] += (double)s[i] * alpha;
        float rn = r[i] + q[i] * (float)(-alpha);
        r[i] = rn;
        ss += (double)rn * (double)rn;
        mx = fmax(mx, fabs((double)rn));
    }
    block_reduce_store(ss, mx, partials, AUX_MAX);
}

__global__ void __launch_bounds__(kBlock) k_dot(int64_t n, const uint16_t* __restrict__ nbinfo, const float* __restrict__ a,
                                                const float* __restrict__ b, double* __restrict__ partials) {
    double d = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        if (nbinfo[i] & NB_FLUID) d += (double)a[i] * (double)b[i];
    block_reduce_store(d, 0.0, partials, AUX_NONE);
}

// s = z + beta s (multSelfAndAdd :229-242)
__global__ void __launch_bounds__(kBlock) k_xpay(int64_t n, const uint16_t* __restrict__ nbinfo, const float* __restrict__ z,
                                                 float* __restrict__ s, const SolveScalars* __restrict__ sc) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !(nbinfo[i] & NB_FLUID)) return;
    s[i] = s[i] * (float)sc->beta + z[i];
}

// FP64 true residual t = b - A x; optionally r = (float)t.
__global__ void __launch_bounds__(kBlock) k_true_residual(SolveArgs S, const uint16_t* __restrict__ nbinfo,
                                                          const float* __restrict__ b, const double* __restrict__ x,
                                                          float* __restrict__ r, int replace, double* __restrict__ partials) {
#pragma clang fp contract(off)
    double ss = 0.0, mx = 0.0;
    const int nrows = S.nx * S.ny;
    for (int row = blockIdx.x; row < nrows; row += gridDim.x) {
        const int p = row / S.ny, y = row - p * S.ny;
        for (int z = threadIdx.x; z < S.nz; z += blockDim.x) {
            const int64_t i = (int64_t)row * S.nz + z;
            const uint16_t info = nbinfo[i];
            if (!(info & NB_FLUID)) continue;
            const RNeighbours n = rneighbours(S, i, p, y, z);
            const int cnt = (info >> NB_NONSOLID_SHIFT) & 7;
            double diag = 0;
            for (int k = 0; k < cnt; k++) diag += S.scale;
            double val = diag * x[i];
            if (info & NB_PX) val += -S.scale * x[n.px];
            if (info & NB_PY) val += -S.scale * x[n.py];
            if (info & NB_PZ) val += -S.scale * x[n.pz];
            if (info & NB_MX) val += -S.scale * x[n.mx];
            if (info & NB_MY) val += -S.scale * x[n.my];
            if (info & NB_MZ) val += -S.scale * x[n.mz];
            const double t = (double)b[i] - val;
            ss += t * t;
            mx = fmax(mx, fabs(t));
            if (replace) r[i] = (float)t;
        }
    }
    block_reduce_store(ss, mx, partials, AUX_MAX);
}

// applyPressureToVelocities (:295-327), FP64 pressure
__global__ void __launch_bounds__(kBlock) k_apply_pressure(SolveArgs S, const uint16_t* __restrict__ nbinfo,
                                                           const uint8_t* __restrict__ label, const double* __restrict__ x,
                                                           float* __restrict__ u2, float* __restrict__ v2,
                                                           float* __restrict__ w2, double sc) {
#pragma clang fp contract(off)
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.nc) return;
    const int cz = (int)(i % S.nz), cy = (int)((i / S.nz) % S.ny), cx = (int)(i / ((int64_t)S.nz * S.ny));
    const int64_t ri = rsidx(S, cx, cy, cz);
    if (!(nbinfo[ri] & NB_FLUID)) return;
    const int64_t st[3] = {(int64_t)S.ny * S.nz, S.nz, 1};
    float* f[3] = {u2, v2, w2};
    const double pi = x[ri];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        uint8_t lp = label[i + st[a]];
        if (lp == FSIM_AIR) f[a][i] = (float)((double)f[a][i] + sc * pi);
        else if (lp == FSIM_WATER) f[a][i] = (float)((double)f[a][i] + sc * (pi - x[rnb_plus(S, cx, cy, cz, a)]));
        if (label[i - st[a]] == FSIM_AIR) f[a][i - st[a]] = (float)((double)f[a][i - st[a]] - sc * pi);
    }
}

__global__ void k_copy_f32(int64_t n, const float* __restrict__ a, float* __restrict__ b) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// ---- host driver --------------------------------------------------------------------
static SolveArgs solve_args(const fsim_ctx* ctx, double dt) {
    SolveArgs S;
    S.nx = ctx->g.nx;
    S.ny = ctx->g.ny;
    S.nz = ctx->g.nz;
    S.nc = ctx->g.nc;
    S.scale = dt / (ctx->prm.fluid_density * ctx->g.cd * ctx->g.cd);
    S.scale_f = (float)S.scale;
    return S;
}

// A wavefront kernel whose boundary poll ran out of spins sets wave_err; its
// results are then invalid and the call fails.
static int check_wave_err(fsim_ctx* ctx) {
    if (ctx->sweep_mode != 0) return FSIM_OK;
    int err = 0;
    FSIM_HIP(ctx, hipMemcpyAsync(&err, ctx->wave_err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (err) {
        FSIM_HIP(ctx, hipMemsetAsync(ctx->wave_err, 0, sizeof(int), ctx->stream));
        return set_err(ctx, FSIM_ERR_TIMEOUT, "wavefront sweep boundary wait timed out");
    }
    return FSIM_OK;
}

// Reductions use at most this many blocks (grid-stride): enough waves to
// keep HBM busy (8 blocks of 4 waves per CU), few enough partials for the
// single-block finalize, which sums them in a fixed order.
static unsigned red_blocks(int64_t n) { return std::min<unsigned>(nblocks(n), 8192u); }

static int read_scalars(fsim_ctx* ctx) {
    FSIM_HIP(ctx, hipMemcpyAsync(ctx->scal_host, ctx->scal, sizeof(SolveScalars), hipMemcpyDeviceToHost, ctx->stream));
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FSIM_OK;
}

static int finalize(fsim_ctx* ctx, int nparts, int op) {
    FSIM_PROF(ctx, KID_FINALIZE);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->partials, nparts, op, ctx->scal);
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

static int build_nbinfo(fsim_ctx* ctx, const SolveArgs& S, bool with_rhs) {
    FSIM_PROF(ctx, KID_RHS);
    RhsArgs A{};
    A.S = S;
    A.label = ctx->label;
    A.u2 = ctx->u2;
    A.v2 = ctx->v2;
    A.w2 = ctx->w2;
    A.density = ctx->density;
    A.nbinfo = ctx->nbinfo;
    A.b = ctx->b;
    A.inv_dx = 1.0 / ctx->g.cd;
    A.avg_p = ctx->prm.average_pressure;
    A.pk = ctx->prm.pressure_k;
    A.pressure_enabled = ctx->prm.pressure_enabled;
    A.partials = ctx->partials;
    hipLaunchKernelGGL(k_nbinfo_rhs, dim3(red_blocks(S.nc)), dim3(kBlock), 0, ctx->stream, A, with_rhs ? 1 : 0);
    FSIM_HIP(ctx, hipGetLastError());
    if (with_rhs) FSIM_TRY(finalize(ctx, (int)red_blocks(S.nc), FIN_BB));
    return FSIM_OK;
}

static int nplane_threads(const SolveArgs& S) { return (S.nx - 2) * (S.ny - 2); }

static int mic_factor(fsim_ctx* ctx, const SolveArgs& S) {
    FSIM_PROF(ctx, KID_MIC_FACTOR);
    FSIM_TRY(fill_f64(ctx, ctx->P64, 0.0, S.nc));
    FSIM_TRY(fill_f32(ctx, ctx->P32, 0.0f, S.nc));
    if (ctx->sweep_mode == 0) return wave_launch(ctx, 0, nullptr, nullptr, S.scale);
    PlaneArgs A{S, ctx->nbinfo, 0};
    const int nt = nplane_threads(S);
    const int lmax = (S.nx - 2) + (S.ny - 2) + (S.nz - 2);
    for (int L = 3; L <= lmax; L++) {
        A.level = L;
        hipLaunchKernelGGL(k_mic_factor_plane, dim3(nblocks(nt)), dim3(kBlock), 0, ctx->stream, A, ctx->P64, ctx->P32);
    }
    hipLaunchKernelGGL(k_pack_w, dim3(nblocks(S.nc)), dim3(kBlock), 0, ctx->stream, S.nc, ctx->nbinfo, ctx->P32, ctx->W);
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

// z = M^-1 r (q is scratch)
static int mic_apply(fsim_ctx* ctx, const SolveArgs& S, const float* r, float* q, float* z) {
    if (ctx->sweep_mode == 0) {
        {
            FSIM_PROF(ctx, KID_SWEEP_FWD);
            FSIM_TRY(wave_launch(ctx, 1, r, q, S.scale));
        }
        FSIM_PROF(ctx, KID_SWEEP_BWD);
        return wave_launch(ctx, 2, q, z, S.scale);
    }
    PlaneArgs A{S, ctx->nbinfo, 0};
    const int nt = nplane_threads(S);
    const int lmax = (S.nx - 2) + (S.ny - 2) + (S.nz - 2);
    for (int L = 3; L <= lmax; L++) {
        A.level = L;
        hipLaunchKernelGGL(k_fwd_plane, dim3(nblocks(nt)), dim3(kBlock), 0, ctx->stream, A, r, q, ctx->P32);
    }
    for (int L = lmax; L >= 3; L--) {
        A.level = L;
        hipLaunchKernelGGL(k_bwd_plane, dim3(nblocks(nt)), dim3(kBlock), 0, ctx->stream, A, q, z, ctx->P32);
    }
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

static int true_residual(fsim_ctx* ctx, const SolveArgs& S, bool replace) {
    {
    FSIM_PROF(ctx, KID_TRUE_RES);
    hipLaunchKernelGGL(k_true_residual, dim3(red_blocks(S.nc)), dim3(kBlock), 0, ctx->stream, S, ctx->nbinfo, ctx->b, ctx->x,
                       ctx->r, replace ? 1 : 0, ctx->partials);
    }
    FSIM_HIP(ctx, hipGetLastError());
    return finalize(ctx, (int)red_blocks(S.nc), FIN_TRUE);
}

static bool stop_met(const fsim_ctx* ctx, double ss, double mx, double bnorm) {
    if (ctx->prm.stop_rule == FSIM_STOP_REL_L2) return std::sqrt(ss) <= ctx->prm.residual_tolerance * bnorm;
    return mx < ctx->prm.residual_tolerance;
}

int stage_solve(fsim_ctx* ctx, double dt) {
    const SolveArgs S = solve_args(ctx, dt);
    const int64_t nc = S.nc;
    const unsigned nb = nblocks(nc);
    const unsigned nr = red_blocks(nc);
    ctx->last_dt = dt;
    fsim_step_stats& st = ctx->stats;
    st.early_exit = 0;
    st.fallback = 0;
    st.reliable_updates = 0;
    st.pcg_iterations = 0;
    FSIM_HIP(ctx, hipMemsetAsync(ctx->scal, 0, sizeof(SolveScalars), ctx->stream));
    FSIM_TRY(fill_f64(ctx, ctx->x, 0.0, nc));
    FSIM_TRY(build_nbinfo(ctx, S, true));
    FSIM_TRY(read_scalars(ctx));
    const double bb = ctx->scal_host->bb;
    const double bnorm = std::sqrt(bb);
    st.b_norm_l2 = bnorm;
    st.num_fluid_cells = (int64_t)ctx->scal_host->nfluid;
    if (bb < 1e-7) {   // early exit without pressure (bridsonSolverGrid.cpp:254-258)
        st.early_exit = 1;
        FSIM_TRY(true_residual(ctx, S, false));
        FSIM_TRY(read_scalars(ctx));
        st.residual_abs_inf = ctx->scal_host->tr_max;
        st.residual_rel_l2 = 1.0;
        return FSIM_OK;
    }
    FSIM_TRY(mic_factor(ctx, S));
    hipLaunchKernelGGL(k_copy_f32, dim3(nb), dim3(kBlock), 0, ctx->stream, nc, ctx->b, ctx->r);
    FSIM_TRY(fill_f32(ctx, ctx->q, 0.0f, nc));
    FSIM_TRY(fill_f32(ctx, ctx->z, 0.0f, nc));
    FSIM_TRY(mic_apply(ctx, S, ctx->r, ctx->y, ctx->z));
    hipLaunchKernelGGL(k_copy_f32, dim3(nb), dim3(kBlock), 0, ctx->stream, nc, ctx->z, ctx->s);
    {
        FSIM_PROF(ctx, KID_DOT);
        hipLaunchKernelGGL(k_dot, dim3(nr), dim3(kBlock), 0, ctx->stream, nc, ctx->nbinfo, ctx->z, ctx->r, ctx->partials);
    }
    FSIM_TRY(finalize(ctx, nr, FIN_SIGMA0));
    const int max_it = ctx->prm.max_iterations;
    const int max_reliable = 64;
    // Reliable updates (restarted mixed-precision PCG): when the FP32
    // recursive residual meets the tolerance, or has dropped by kDelta since
    // the last FP64 true residual, the true residual b - A x is recomputed in
    // FP64.  The solve stops if it meets the tolerance; otherwise r takes its
    // value and the search direction restarts from z = M^-1 r.
    const double kDelta = 1e-3;
    double last_true = bnorm;
    int it = 0;
    for (; it < max_it; it++) {
        {
            FSIM_PROF(ctx, KID_SPMV);
            hipLaunchKernelGGL(k_spmv_dot, dim3(nr), dim3(kBlock), 0, ctx->stream, S, ctx->nbinfo, ctx->s, ctx->q,
                               ctx->partials);
        }
        FSIM_TRY(finalize(ctx, nr, FIN_ALPHA));
        FSIM_TRY(read_scalars(ctx));
        if (ctx->scal_host->nan_flag & 1) break;
        {
            FSIM_PROF(ctx, KID_UPDATE);
            hipLaunchKernelGGL(k_update, dim3(nr), dim3(kBlock), 0, ctx->stream, nc, ctx->nbinfo, ctx->s, ctx->q, ctx->x,
                               ctx->r, ctx->scal, ctx->partials);
        }
        FSIM_TRY(finalize(ctx, nr, FIN_RES));
        FSIM_TRY(read_scalars(ctx));
        const double rr = ctx->scal_host->rr, rmax = ctx->scal_host->rmax;
        bool restart = false;
        if (stop_met(ctx, rr, rmax, bnorm) || std::sqrt(rr) < kDelta * last_true) {
            const bool replace = st.reliable_updates < max_reliable;
            FSIM_TRY(true_residual(ctx, S, replace));
            FSIM_TRY(read_scalars(ctx));
            if (stop_met(ctx, ctx->scal_host->tr_ss, ctx->scal_host->tr_max, bnorm)) break;
            if (!replace) {
                if (stop_met(ctx, rr, rmax, bnorm)) break;
            } else {
                st.reliable_updates++;
                last_true = std::sqrt(ctx->scal_host->tr_ss);
                restart = true;
            }
        }
        FSIM_TRY(mic_apply(ctx, S, ctx->r, ctx->y, ctx->z));
        {
            FSIM_PROF(ctx, KID_DOT);
            hipLaunchKernelGGL(k_dot, dim3(nr), dim3(kBlock), 0, ctx->stream, nc, ctx->nbinfo, ctx->z, ctx->r, ctx->partials);
        }
        if (restart) {
            FSIM_TRY(finalize(ctx, nr, FIN_SIGMA0));
            hipLaunchKernelGGL(k_copy_f32, dim3(nb), dim3(kBlock), 0, ctx->stream, nc, ctx->z, ctx->s);
            FSIM_HIP(ctx, hipGetLastError());
            continue;
        }
        FSIM_TRY(finalize(ctx, nr, FIN_BETA));
        FSIM_TRY(read_scalars(ctx));
        if (ctx->scal_host->nan_flag & 2) break;
        {
            FSIM_PROF(ctx, KID_XPAY);
            hipLaunchKernelGGL(k_xpay, dim3(nb), dim3(kBlock), 0, ctx->stream, nc, ctx->nbinfo, ctx->z, ctx->s, ctx->scal);
        }
        FSIM_HIP(ctx, hipGetLastError());
    }
    st.pcg_iterations = it;
    const double sc = dt / (ctx->prm.fluid_density * ctx->g.cd);
    {
        FSIM_PROF(ctx, KID_PRESSURE);
        hipLaunchKernelGGL(k_apply_pressure, dim3(nb), dim3(kBlock), 0, ctx->stream, S, ctx->nbinfo, ctx->label, ctx->x,
                           ctx->u2, ctx->v2, ctx->w2, sc);
    }
    FSIM_HIP(ctx, hipGetLastError());
    FSIM_TRY(true_residual(ctx, S, false));
    FSIM_TRY(read_scalars(ctx));
    st.residual_abs_inf = ctx->scal_host->tr_max;
    st.residual_rel_l2 = bnorm > 0 ? std::sqrt(ctx->scal_host->tr_ss) / bnorm : 0.0;
    FSIM_TRY(check_wave_err(ctx));
    if (ctx->scal_host->nan_flag) return set_err(ctx, FSIM_ERR_NUMERIC, "NaN in the pressure solve");
    return FSIM_OK;
}

int wave_init(fsim_ctx* ctx) {
    WaveArgs& A = ctx->wave;
    A.nx = ctx->g.nx;
    A.ny = ctx->g.ny;
    A.nz = ctx->g.nz;
    A.TJ = (A.ny - 2 + WF_TY - 1) / WF_TY;
    A.TK = (A.nz - 2 + WF_TZ - 1) / WF_TZ;
    // packet rows: one per step of a tile; two words per value (the factor's
    // FP64), the sweeps use one
    A.SP = wave_steps(A.nx - 2);
    const int64_t ny_pk = 2 * (int64_t)A.TJ * A.TK * A.SP * WF_TZ;   // per tile (J, K)
    const int64_t nz_pk = 2 * (int64_t)A.TJ * A.TK * A.SP * WF_TY;
    if (ctx->g.nc > kWaveMaxCells) ctx->sweep_mode = 1;
    if (hipMalloc(&ctx->W, sizeof(uint64_t) * ctx->g.nc) != hipSuccess ||
        hipMalloc(&ctx->ybnd, sizeof(uint64_t) * ny_pk) != hipSuccess ||
        hipMalloc(&ctx->zbnd, sizeof(uint64_t) * nz_pk) != hipSuccess ||
        hipMalloc(&ctx->wave_ticket, sizeof(unsigned long long) * 3) != hipSuccess ||
        hipMalloc(&ctx->wave_err, sizeof(int)) != hipSuccess)
        return set_err(ctx, FSIM_ERR_OOM, "wavefront buffer allocation failed");
    FSIM_HIP(ctx, hipMemsetAsync(ctx->W, 0, sizeof(uint64_t) * ctx->g.nc, ctx->stream));
    FSIM_HIP(ctx, hipMemsetAsync(ctx->ybnd, 0, sizeof(uint64_t) * ny_pk, ctx->stream));
    FSIM_HIP(ctx, hipMemsetAsync(ctx->zbnd, 0, sizeof(uint64_t) * nz_pk, ctx->stream));
    FSIM_HIP(ctx, hipMemsetAsync(ctx->wave_ticket, 0, sizeof(unsigned long long) * 3, ctx->stream));
    FSIM_HIP(ctx, hipMemsetAsync(ctx->wave_err, 0, sizeof(int), ctx->stream));
    A.nbinfo = ctx->nbinfo;
    A.P64 = ctx->P64;
    A.P32 = ctx->P32;
    A.W = ctx->W;
    A.ybnd = ctx->ybnd;
    A.zbnd = ctx->zbnd;
    A.err = ctx->wave_err;
    return FSIM_OK;
}

int solver_apply(fsim_ctx* ctx, int op, double dt, const double* in, double* out) {
    const SolveArgs S = solve_args(ctx, dt);
    const int64_t nc = S.nc;
    FSIM_TRY(build_nbinfo(ctx, S, false));
    // host vectors are in flat order, device solver vectors in the rotated one
    std::vector<int64_t> rmap(nc);
    for (int x = 0; x < S.nx; x++)
        for (int y = 0; y < S.ny; y++)
            for (int z = 0; z < S.nz; z++) rmap[((int64_t)x * S.ny + y) * S.nz + z] = ridx(x, y, z, S.nx, S.ny, S.nz);
    std::vector<float> hin(nc), hout(nc);
    if (op == FSIM_OP_FACTOR_MIC0) {
        FSIM_TRY(mic_factor(ctx, S));
        FSIM_TRY(check_wave_err(ctx));
        std::vector<double> hp(nc);
        FSIM_HIP(ctx, hipMemcpyAsync(hp.data(), ctx->P64, sizeof(double) * nc, hipMemcpyDeviceToHost, ctx->stream));
        FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (int64_t i = 0; i < nc; i++) out[i] = hp[rmap[i]];
        return FSIM_OK;
    }
    for (int64_t i = 0; i < nc; i++) hin[rmap[i]] = (float)in[i];
    FSIM_HIP(ctx, hipMemcpyAsync(ctx->r, hin.data(), sizeof(float) * nc, hipMemcpyHostToDevice, ctx->stream));
    if (op == FSIM_OP_APPLY_A) {
        hipLaunchKernelGGL(k_spmv_dot, dim3(red_blocks(nc)), dim3(kBlock), 0, ctx->stream, S, ctx->nbinfo, ctx->r, ctx->q,
                           ctx->partials);
        FSIM_HIP(ctx, hipGetLastError());
        FSIM_HIP(ctx, hipMemcpyAsync(hout.data(), ctx->q, sizeof(float) * nc, hipMemcpyDeviceToHost, ctx->stream));
    } else if (op == FSIM_OP_APPLY_MIC0) {
        FSIM_TRY(mic_factor(ctx, S));
        FSIM_TRY(fill_f32(ctx, ctx->y, 0.0f, nc));
        FSIM_TRY(fill_f32(ctx, ctx->z, 0.0f, nc));
        FSIM_TRY(mic_apply(ctx, S, ctx->r, ctx->y, ctx->z));
        FSIM_TRY(check_wave_err(ctx));
        FSIM_HIP(ctx, hipMemcpyAsync(hout.data(), ctx->z, sizeof(float) * nc, hipMemcpyDeviceToHost, ctx->stream));
    } else {
        return set_err(ctx, FSIM_ERR_INVALID, "unknown solver op");
    }
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < nc; i++) out[i] = hout[rmap[i]];
    return FSIM_OK;
}

}  // namespace fsim
