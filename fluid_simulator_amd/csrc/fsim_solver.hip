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
#include <type_traits>

#include "fsim_internal.hpp"

namespace fsim {

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
// ---- in-kernel finish of a reduction -------------------------------------------------
// Each block reduces its values to one partial (plus an aux value: max or sum)
// and stores it with device-scope stores; once they are acknowledged the block
// takes a ticket from a device-scope counter.  The block with the last ticket
// sums all partials in a fixed order (thread t: partials t, t + kBlock, ...,
// then a pairwise tree), applies the reduction's scalar update and resets the
// counter.  The sum's order does not depend on which block is last, so the
// result is deterministic, and no separate finalize launch is needed.
constexpr int AUX_NONE = 0, AUX_MAX = 1, AUX_SUM = 2;
constexpr unsigned kMaxParts = 2048;   // reduction grids have at most this many blocks, a multiple of 8
enum FinOp {
    FIN_NONE = -1, FIN_BB = 0, FIN_SIGMA0 = 1, FIN_ALPHA = 2, FIN_RES = 3, FIN_BETA = 4, FIN_TRUE = 5, FIN_ZQ = 6,
    FIN_FBETA = 7
};

struct Reduce {
    double* partials;   // 2 x gridDim.x
    unsigned* count;    // zero between launches
    SolveScalars* sc;
    int op;             // FinOp
    SolveScalars* host = nullptr;   // pinned host copy written by the finishing block, or null
};

__device__ inline void apply_fin(int op, double s, double m, SolveScalars* sc) {
    switch (op) {
    case FIN_BB: sc->bb = s; sc->nfluid = m; break;
    case FIN_SIGMA0: sc->zr = s; break;
    case FIN_ALPHA: {   // sigma is the last z.r (initial, restart or beta step)
        sc->sigma = sc->zr;
        sc->sq = s;
        const double alpha = sc->sigma / s;
        sc->alpha = alpha;
        if (alpha != alpha) sc->nan_flag |= 1;
        break;
    }
    case FIN_RES: sc->rr = s; sc->rmax = m; break;
    case FIN_BETA: {
        if (sc->sigma != sc->sigma) sc->nan_flag |= 2;   // the reference tests the old sigma (:285-286)
        sc->zr = s;
        sc->beta = s / sc->sigma;
        break;
    }
    case FIN_TRUE: sc->tr_ss = s; sc->tr_max = m; break;
    case FIN_ZQ: sc->zq = s; break;
    case FIN_FBETA: {   // flexible PCG: beta = z'.(r' - r) / z.r = -alpha z'.q / z.r
        if (sc->sigma != sc->sigma) sc->nan_flag |= 2;
        sc->zr = s;
        sc->beta = -sc->alpha * sc->zq / sc->sigma;
        break;
    }
    default: break;
    }
}

__device__ inline void put_dev(double* p, double v) {
    __hip_atomic_store(reinterpret_cast<uint64_t*>(p), __builtin_bit_cast(uint64_t, v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double get_dev(const double* p) {
    return __builtin_bit_cast(double, __hip_atomic_load(reinterpret_cast<const uint64_t*>(p), __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT));
}

template <int AUX>
__device__ void block_finish(double sum, double aux, const Reduce& R) {
    __shared__ double lds_s[kBlock], lds_m[kBlock];
    __shared__ int last;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    sum = wave_sum(sum);
    if (AUX == AUX_MAX) aux = wave_max(aux);
    if (AUX == AUX_SUM) aux = wave_sum(aux);
    if (lane == 0) {
        lds_s[wave] = sum;
        lds_m[wave] = aux;
    }
    __syncthreads();
    const unsigned nparts = gridDim.x;
    if (threadIdx.x == 0) {
        double s = 0, m = 0;
        for (int w = 0; w < kBlock / 64; w++) {
            s += lds_s[w];
            m = AUX == AUX_SUM ? m + lds_m[w] : fmax(m, lds_m[w]);
        }
        put_dev(&R.partials[blockIdx.x], s);
        if (AUX != AUX_NONE) put_dev(&R.partials[nparts + blockIdx.x], m);
        __builtin_amdgcn_s_waitcnt(0);   // the partials are acknowledged before the ticket
        last = __hip_atomic_fetch_add(R.count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == nparts - 1;
    }
    __syncthreads();
    if (!last) return;
    // all loads first (at most kMaxParts / kBlock per thread), then the sums in order
    constexpr int kPer = kMaxParts / kBlock;
    double vs[kPer], vm[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        const unsigned i = threadIdx.x + k * kBlock;
        vs[k] = i < nparts ? get_dev(&R.partials[i]) : 0.0;
        vm[k] = AUX != AUX_NONE && i < nparts ? get_dev(&R.partials[nparts + i]) : 0.0;
    }
    double s = 0, m = 0;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        s += vs[k];
        m = AUX == AUX_SUM ? m + vm[k] : fmax(m, vm[k]);
    }
    lds_s[threadIdx.x] = s;
    lds_m[threadIdx.x] = m;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            lds_s[threadIdx.x] += lds_s[threadIdx.x + off];
            lds_m[threadIdx.x] = AUX == AUX_SUM ? lds_m[threadIdx.x] + lds_m[threadIdx.x + off]
                                                : fmax(lds_m[threadIdx.x], lds_m[threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        apply_fin(R.op, lds_s[0], lds_m[0], R.sc);
        __hip_atomic_store(R.count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (R.host) {
            // the words the host tests after the r update (rr, rmax, nan_flag) straight into pinned
            // host memory with system-scope stores; the kernel's end and the event after it order
            // them before the host's read
            auto put_host = [](void* dst, const void* src) {
                __hip_atomic_store(reinterpret_cast<uint64_t*>(dst), *reinterpret_cast<const uint64_t*>(src),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            };
            put_host(&R.host->rr, &R.sc->rr);
            put_host(&R.host->rmax, &R.sc->rmax);
            static_assert(offsetof(SolveScalars, pad) == offsetof(SolveScalars, nan_flag) + 4, "one 8-byte word");
            put_host(&R.host->nan_flag, &R.sc->nan_flag);
        }
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
    uint8_t* colw;   // set to 1 for the (y, z) column of each WATER cell (cleared before)
    float* b;
    double inv_dx;
    double avg_p, pk;
    int pressure_enabled;
    Reduce red;   // FIN_BB: ||b||^2 and the WATER cell count
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
            A.colw[y * S.nz + z] = 1;
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
    if (with_rhs) block_finish<AUX_SUM>(bsq, nfl, A.red);
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

// ---- SpMV (applyAMatrix :165-198) with the s.q dot ---------------------------------
constexpr int SPMV_PLAIN = 0, SPMV_COPY = 1, SPMV_XPAY = 2;

// Work split of the stencil kernels: one work item is a 256-cell chunk of an
// interior rotated row (plane p, row y in 1..ny-2, z in 1..nz-2), handled by
// one wave with 4 cells per lane (z = z0 + lane + 64k, coalesced for each k).
// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx mod 8), so
// the items are cut into 8 contiguous ranges, one per XCD: the +-x/+-y
// neighbour rows a wave reads were loaded by waves of the same XCD (its L2).
constexpr int kRowCells = 4 * 64;
struct RowItems {
    int nchunk, ny2;
    int end;           // end of this XCD's item range (items fit in 32 bits: nc < 2^31)
    int first, step;   // this wave's first item and stride
};
__device__ inline RowItems row_items(const SolveArgs& S) {
    RowItems it;
    it.nchunk = (S.nz - 2 + kRowCells - 1) / kRowCells;
    it.ny2 = S.ny - 2;
    const int total = S.nx * it.ny2 * it.nchunk;
    const int per_xcd = (total + 7) / 8;
    const int xcd = blockIdx.x & 7;
    const int begin = xcd * per_xcd;
    it.end = min(total, begin + per_xcd);
    const int waves_per_block = blockDim.x / 64;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    it.first = begin + (int)(blockIdx.x >> 3) * waves_per_block + wave;
    it.step = (int)(gridDim.x >> 3) * waves_per_block;
    return it;
}
// (plane, row, first z) of item k
__device__ inline void item_cells(const RowItems& it, int k, int& p, int& y, int& z0) {
    const int c = k % it.nchunk;
    const int rest = k / it.nchunk;
    y = 1 + rest % it.ny2;
    p = rest / it.ny2;
    z0 = 1 + c * kRowCells;
}

// SPMV_PLAIN: q = A s.  SPMV_COPY: s' = z, q = A s'.  SPMV_XPAY: s' = s * beta + z
// (multSelfAndAdd :229-242), q = A s'.  The new search direction s' goes to a
// second buffer and is recomputed for the neighbours from z and the old s, so
// no block reads a value another block has already replaced.  Only WATER
// cells are written: q is zeroed before a solve and never read elsewhere.
template <int MODE>
__global__ void __launch_bounds__(kBlock) k_spmv(SolveArgs S, const uint16_t* __restrict__ nbinfo,
                                                 const float* __restrict__ z, const float* __restrict__ s,
                                                 float* __restrict__ s_new, float* __restrict__ q, Reduce red) {
    double d = 0.0;
    const float ms = -S.scale_f;
    const float beta = MODE == SPMV_XPAY ? (float)red.sc->beta : 0.f;
    auto dir = [&](int64_t j) {
        if constexpr (MODE == SPMV_PLAIN) return s[j];
        else if constexpr (MODE == SPMV_COPY) return z[j];
        else return s[j] * beta + z[j];
    };
    const RowItems it = row_items(S);
    const int lane = threadIdx.x & 63;
    // the nbinfo words of the next item are loaded while this item's gathers are in flight
    auto load_info = [&](int k, uint16_t (&inf)[4]) {
        if (k >= it.end) return;
        int p, y, z0;
        item_cells(it, k, p, y, z0);
        const int64_t rowbase = ((int64_t)p * S.ny + y) * S.nz;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int zc = z0 + lane + 64 * c;
            inf[c] = zc < S.nz - 1 ? nbinfo[rowbase + zc] : (uint16_t)0;
        }
    };
    uint16_t info_next[4] = {0, 0, 0, 0};
    load_info(it.first, info_next);
    for (int k = it.first; k < it.end; k += it.step) {
        int p, y, z0;
        item_cells(it, k, p, y, z0);
        const int64_t rowbase = ((int64_t)p * S.ny + y) * S.nz;
        uint16_t info[4];
        int zc[4];
        bool any = false;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            info[c] = info_next[c];
            zc[c] = min(z0 + lane + 64 * c, S.nz - 2);   // lanes past the row gather an interior cell, unused
            any |= (info[c] & NB_FLUID) != 0;
        }
        load_info(k + it.step, info_next);
        if (__builtin_amdgcn_ballot_w64(any) == 0) continue;   // no WATER cell in the wave's chunk
        // all gathers of the 4 cells first (interior cells: every neighbour index is in range),
        // then the sums with the reference's term order, by mask
        float v[4][7];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int64_t i = rowbase + zc[c];
            const RNeighbours n = rneighbours(S, i, p, y, zc[c]);
            v[c][0] = dir(i);
            v[c][1] = dir(n.px);
            v[c][2] = dir(n.py);
            v[c][3] = dir(n.pz);
            v[c][4] = dir(n.mx);
            v[c][5] = dir(n.my);
            v[c][6] = dir(n.mz);
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const uint16_t f = info[c];
            if (!(f & NB_FLUID)) continue;
            const int64_t i = rowbase + zc[c];
            const int cnt = (f >> NB_NONSOLID_SHIFT) & 7;
            const float si = v[c][0];
            float val = (S.scale_f * (float)cnt) * si;
            val = (f & NB_PX) ? val + ms * v[c][1] : val;
            val = (f & NB_PY) ? val + ms * v[c][2] : val;
            val = (f & NB_PZ) ? val + ms * v[c][3] : val;
            val = (f & NB_MX) ? val + ms * v[c][4] : val;
            val = (f & NB_MY) ? val + ms * v[c][5] : val;
            val = (f & NB_MZ) ? val + ms * v[c][6] : val;
            d += (double)si * (double)val;
            q[i] = val;
            if (MODE != SPMV_PLAIN) s_new[i] = si;
        }
    }
    block_finish<AUX_NONE>(d, 0.0, red);
}

// Streaming kernels without per-cell masks: the solver vectors r, q and z
// are exactly zero off WATER (zeroed at the start of a solve, written on
// WATER cells only), so the extra terms are exact zeros.  The vectors are cut
// into chunks of CH cells (256 with float4 accesses when nc is a multiple of
// 4, else 64); a chunk without WATER is skipped.  Its flag is one 32-bit word
// (built with nbinfo for every solve), so a wave reads it with a scalar load;
// a wave takes kUnits chunks per trip (grid stride) and issues all their
// loads before using any, so several memory round trips are in flight.
constexpr int kUnits = 4;
__host__ __device__ constexpr int stream_chunk(bool vec4) { return vec4 ? 256 : 64; }
// 1024 measured fastest at 256^3 (the dot: 38 us per launch with 2048 blocks, 29 us with 1024;
// 512 is slower again)
#ifndef FSIM_STREAM_BLOCKS
#define FSIM_STREAM_BLOCKS 1024
#endif
static_assert(FSIM_STREAM_BLOCKS % 8 == 0 && FSIM_STREAM_BLOCKS <= (int)kMaxParts, "stream grid within the reduction");
// blocks of the flagged-chunk kernels (at most kMaxParts, a multiple of 8)
static unsigned stream_blocks(int64_t nc) {
    const unsigned need = (unsigned)((nc / stream_chunk(true) + kUnits * 4 - 1) / (kUnits * 4));
    return std::max(8u, std::min<unsigned>((need + 7) / 8 * 8, FSIM_STREAM_BLOCKS));
}

template <int CH>
__global__ void __launch_bounds__(kBlock) k_chunk_flags(int64_t nc, const uint16_t* __restrict__ nbinfo,
                                                        uint32_t* __restrict__ flags) {
    const int lane = threadIdx.x & 63;
    const int nchunks = (int)((nc + CH - 1) / CH);
    const int nwaves = (int)(gridDim.x * (blockDim.x / 64));
    for (int u = (int)(blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6)); u < nchunks; u += nwaves) {
        bool any = false;
#pragma unroll
        for (int k = 0; k < CH / 64; k++) {
            const int64_t i = (int64_t)u * CH + k * 64 + lane;
            any = any || (i < nc && (nbinfo[i] & NB_FLUID));
        }
        any = __builtin_amdgcn_ballot_w64(any) != 0;
        if (lane == 0) flags[u] = any ? 1u : 0u;
    }
}

template <bool VEC4, int NV, class Load, class Use>
__device__ inline void for_flagged_chunks(int64_t nc, const uint32_t* __restrict__ flags, Load&& load, Use&& use) {
    constexpr int CH = stream_chunk(VEC4);
    using V = typename std::conditional<VEC4, float4, float>::type;
    const int lane = threadIdx.x & 63;
    const int nchunks = (int)((nc + CH - 1) / CH);
    const int nwaves = (int)(gridDim.x * (blockDim.x / 64));
    const int w0 = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6)));
    for (int u0 = w0; u0 < nchunks; u0 += kUnits * nwaves) {
        uint32_t f[kUnits];
#pragma unroll
        for (int k = 0; k < kUnits; k++) {
            const int u = u0 + k * nwaves;
            f[k] = u < nchunks ? flags[u] : 0u;
        }
        // loads without branches (a skipped lane reads cell 0 and uses nothing), so that all
        // kUnits x NV loads are issued before the first wait
        V v[kUnits][NV];
        bool on[kUnits];
        int64_t at[kUnits];
#pragma unroll
        for (int k = 0; k < kUnits; k++) {
            const int64_t i = (int64_t)(u0 + k * nwaves) * CH + (VEC4 ? 4 : 1) * lane;
            on[k] = f[k] != 0u && i < nc;
            at[k] = on[k] ? i : 0;
            load(at[k], v[k]);
        }
#pragma unroll
        for (int k = 0; k < kUnits; k++)
            if (on[k]) use(at[k], v[k]);
    }
}

// r -= alpha q, with ||r||^2 and max|r|
template <bool VEC4>
__global__ void __launch_bounds__(kBlock) k_update_r(int64_t nc, const uint32_t* __restrict__ flags,
                                                     const float* __restrict__ q, float* __restrict__ r, Reduce red) {
    using V = typename std::conditional<VEC4, float4, float>::type;
    double ss = 0.0, mx = 0.0;
    const float malpha = (float)(-red.sc->alpha);
    auto cell = [&](float qv, float& rv) {
        rv = rv + qv * malpha;
        ss += (double)rv * (double)rv;
        mx = fmax(mx, fabs((double)rv));
    };
    for_flagged_chunks<VEC4, 2>(
        nc, flags,
        [&](int64_t i, V(&v)[2]) {
            v[0] = *reinterpret_cast<const V*>(q + i);
            v[1] = *reinterpret_cast<const V*>(r + i);
        },
        [&](int64_t i, V(&v)[2]) {
            if constexpr (VEC4) {
                cell(v[0].x, v[1].x);
                cell(v[0].y, v[1].y);
                cell(v[0].z, v[1].z);
                cell(v[0].w, v[1].w);
            } else {
                cell(v[0], v[1]);
            }
            *reinterpret_cast<V*>(r + i) = v[1];
        });
    block_finish<AUX_MAX>(ss, mx, red);
}

// a . b (FP64 products of the FP32 values)
template <bool VEC4>
__global__ void __launch_bounds__(kBlock) k_dot(int64_t nc, const uint32_t* __restrict__ flags,
                                                const float* __restrict__ a, const float* __restrict__ b, Reduce red) {
    using V = typename std::conditional<VEC4, float4, float>::type;
    double d = 0.0;
    for_flagged_chunks<VEC4, 2>(
        nc, flags,
        [&](int64_t i, V(&v)[2]) {
            v[0] = *reinterpret_cast<const V*>(a + i);
            v[1] = *reinterpret_cast<const V*>(b + i);
        },
        [&](int64_t i, V(&v)[2]) {
            if constexpr (VEC4) {
                d += (double)v[0].x * (double)v[1].x;
                d += (double)v[0].y * (double)v[1].y;
                d += (double)v[0].z * (double)v[1].z;
                d += (double)v[0].w * (double)v[1].w;
            } else {
                d += (double)v[0] * (double)v[1];
            }
        });
    block_finish<AUX_NONE>(d, 0.0, red);
}

// x += alpha s (FP64) for WATER cells.  One cell per thread: it runs on a
// second stream beside the r update, and a lighter HBM load there costs the
// r update less than a vectorised kernel does (measured at 256^3).  A NaN
// alpha leaves x as it is: the reference breaks before multAdd
// (bridsonSolverGrid.cpp:270-272) and applies the last pressure.
__global__ void __launch_bounds__(kBlock) k_update_x(int64_t n, const uint16_t* __restrict__ nbinfo,
                                                     const float* __restrict__ s, double* __restrict__ x,
                                                     const SolveScalars* __restrict__ sc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !(nbinfo[i] & NB_FLUID) || (sc->nan_flag & 1)) return;
    x[i] += (double)s[i] * sc->alpha;
}


// FP64 true residual t = b - A x; optionally r = (float)t.
__global__ void __launch_bounds__(kBlock) k_true_residual(SolveArgs S, const uint16_t* __restrict__ nbinfo,
                                                          const float* __restrict__ b, const double* __restrict__ x,
                                                          float* __restrict__ r, int replace, Reduce red) {
#pragma clang fp contract(off)
    double ss = 0.0, mx = 0.0;
    const RowItems it = row_items(S);
    const int lane = threadIdx.x & 63;
    for (int k = it.first; k < it.end; k += it.step) {
        int p, y, z0;
        item_cells(it, k, p, y, z0);
        const int64_t rowbase = ((int64_t)p * S.ny + y) * S.nz;
        uint16_t info[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int zc = z0 + lane + 64 * c;
            info[c] = zc < S.nz - 1 ? nbinfo[rowbase + zc] : (uint16_t)0;
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            if (!(info[c] & NB_FLUID)) continue;
            const int zc = z0 + lane + 64 * c;
            const int64_t i = rowbase + zc;
            const RNeighbours n = rneighbours(S, i, p, y, zc);
            // all 7 loads first (a WATER cell is interior: every neighbour index is in range), then the
            // reference's terms in order, each kept by its flag
            const double xi = x[i], xpx = x[n.px], xpy = x[n.py], xpz = x[n.pz], xmx = x[n.mx], xmy = x[n.my],
                         xmz = x[n.mz];
            const uint16_t f = info[c];
            const int cnt = (f >> NB_NONSOLID_SHIFT) & 7;
            double diag = 0;
            for (int j = 0; j < cnt; j++) diag += S.scale;
            double val = diag * xi;
            val = (f & NB_PX) ? val + -S.scale * xpx : val;
            val = (f & NB_PY) ? val + -S.scale * xpy : val;
            val = (f & NB_PZ) ? val + -S.scale * xpz : val;
            val = (f & NB_MX) ? val + -S.scale * xmx : val;
            val = (f & NB_MY) ? val + -S.scale * xmy : val;
            val = (f & NB_MZ) ? val + -S.scale * xmz : val;
            const double t = (double)b[i] - val;
            ss += t * t;
            mx = fmax(mx, fabs(t));
            if (replace) r[i] = (float)t;
        }
    }
    block_finish<AUX_MAX>(ss, mx, red);
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
    // only interior cells are WATER (their six neighbours exist); the test keeps
    // the accesses below in range whatever the mask holds
    if (cx < 1 || cy < 1 || cz < 1 || cx > S.nx - 2 || cy > S.ny - 2 || cz > S.nz - 2) return;
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
    S.off63 = (WF_TZ - 1) % S.nx;
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

// Reduction grids: at most kMaxParts blocks (grid-stride; 8 blocks of 4 waves
// per CU keep HBM busy), a multiple of 8 for the XCD-aware row split, and
// the last block sums the partials in one pass.
static unsigned red_blocks(int64_t n) { return std::min<unsigned>((nblocks(n) + 7) / 8 * 8, kMaxParts); }
// 2048 and 1024 measure alike at 256^3, 512 is slower
#ifndef FSIM_SPMV_BLOCKS
#define FSIM_SPMV_BLOCKS 2048
#endif
static_assert(FSIM_SPMV_BLOCKS % 8 == 0 && FSIM_SPMV_BLOCKS <= (int)kMaxParts, "SpMV grid within the reduction");
static unsigned spmv_blocks(int64_t n) { return std::min<unsigned>(red_blocks(n), FSIM_SPMV_BLOCKS); }

static Reduce reduce_to(fsim_ctx* ctx, int op) { return Reduce{ctx->partials, ctx->red_count, ctx->scal, op}; }
// the same, with the finished scalars also written to the pinned host copy (read after an event)
static Reduce reduce_to_host(fsim_ctx* ctx, int op) {
    return Reduce{ctx->partials, ctx->red_count, ctx->scal, op, ctx->scal_host};
}

static int read_scalars(fsim_ctx* ctx) {
    FSIM_HIP(ctx, hipMemcpyAsync(ctx->scal_host, ctx->scal, sizeof(SolveScalars), hipMemcpyDeviceToHost, ctx->stream));
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
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
    A.colw = ctx->colw;
    A.b = ctx->b;
    A.inv_dx = 1.0 / ctx->g.cd;
    A.avg_p = ctx->prm.average_pressure;
    A.pk = ctx->prm.pressure_k;
    A.pressure_enabled = ctx->prm.pressure_enabled;
    A.red = reduce_to(ctx, FIN_BB);
    FSIM_TRY(fill_u8(ctx, ctx->colw, 0, (int64_t)S.ny * S.nz));
    hipLaunchKernelGGL(k_nbinfo_rhs, dim3(red_blocks(S.nc)), dim3(kBlock), 0, ctx->stream, A, with_rhs ? 1 : 0);
    if (S.nc % 4 == 0)
        hipLaunchKernelGGL(k_chunk_flags<stream_chunk(true)>, dim3(nblocks(S.nc / 4)), dim3(kBlock), 0, ctx->stream, S.nc,
                           ctx->nbinfo, ctx->chunk_flags);
    else
        hipLaunchKernelGGL(k_chunk_flags<stream_chunk(false)>, dim3(nblocks(S.nc)), dim3(kBlock), 0, ctx->stream, S.nc,
                           ctx->nbinfo, ctx->chunk_flags);
    FSIM_HIP(ctx, hipGetLastError());
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
    FSIM_PROF(ctx, KID_TRUE_RES);
    hipLaunchKernelGGL(k_true_residual, dim3(red_blocks(S.nc)), dim3(kBlock), 0, ctx->stream, S, ctx->nbinfo, ctx->b, ctx->x,
                       ctx->r, replace ? 1 : 0, reduce_to(ctx, FIN_TRUE));
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

static int dot_with(fsim_ctx* ctx, const float* a, const float* b, int op) {
    FSIM_PROF(ctx, KID_DOT);
    const int64_t nc = ctx->g.nc;
    if (nc % 4 == 0)
        hipLaunchKernelGGL(k_dot<true>, dim3(stream_blocks(nc)), dim3(kBlock), 0, ctx->stream, nc, ctx->chunk_flags, a, b,
                           reduce_to(ctx, op));
    else
        hipLaunchKernelGGL(k_dot<false>, dim3(stream_blocks(nc)), dim3(kBlock), 0, ctx->stream, nc, ctx->chunk_flags, a, b,
                           reduce_to(ctx, op));
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}
static int dot_zr(fsim_ctx* ctx, int op) { return dot_with(ctx, ctx->z, ctx->r, op); }

// New search direction s (z, or z + beta s) into the second buffer, q = A s,
// s.q -> alpha.
static int direction_step(fsim_ctx* ctx, const SolveArgs& S, bool copy) {
    {
        FSIM_PROF(ctx, KID_SPMV);
        const dim3 grid(spmv_blocks(S.nc)), block(kBlock);
        const Reduce red = reduce_to(ctx, FIN_ALPHA);
        if (copy)
            hipLaunchKernelGGL(k_spmv<SPMV_COPY>, grid, block, 0, ctx->stream, S, ctx->nbinfo, ctx->z, ctx->s, ctx->s2,
                               ctx->q, red);
        else
            hipLaunchKernelGGL(k_spmv<SPMV_XPAY>, grid, block, 0, ctx->stream, S, ctx->nbinfo, ctx->z, ctx->s, ctx->s2,
                               ctx->q, red);
    }
    FSIM_HIP(ctx, hipGetLastError());
    std::swap(ctx->s, ctx->s2);
    return FSIM_OK;
}

// x += alpha s of the direction step just queued, on the second stream: it
// runs beside the r update instead of inside the next direction step
// (measured at 256^3: started after the r update instead, it slows the
// latency-bound forward sweep by more than it slows the r update here).  The
// next direction step replaces alpha and, one step later, this s buffer, so
// the main stream waits for it (join_x) before the next direction step and
// before anything else reads x.
static int update_x_beside(fsim_ctx* ctx) {
    FSIM_HIP(ctx, hipEventRecord(ctx->dir_ev, ctx->stream));
    FSIM_HIP(ctx, hipStreamWaitEvent(ctx->xstream, ctx->dir_ev, 0));
    hipLaunchKernelGGL(k_update_x, dim3(nblocks(ctx->g.nc)), dim3(kBlock), 0, ctx->xstream, ctx->g.nc, ctx->nbinfo,
                       ctx->s, ctx->x, ctx->scal);
    FSIM_HIP(ctx, hipGetLastError());
    FSIM_HIP(ctx, hipEventRecord(ctx->x_ev, ctx->xstream));
    return FSIM_OK;
}
static int join_x(fsim_ctx* ctx) {
    FSIM_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->x_ev, 0));
    return FSIM_OK;
}
// joins the x stream on every exit of a solve, error returns included, so that no later work on the
// main stream (the next solve clears x) runs beside a pending x update
struct JoinXOnExit {
    fsim_ctx* ctx;
    ~JoinXOnExit() { (void)hipStreamWaitEvent(ctx->stream, ctx->x_ev, 0); }
};

static bool stop_met(const fsim_ctx* ctx, double ss, double mx, double bnorm) {
    if (ctx->prm.stop_rule == FSIM_STOP_REL_L2) return std::sqrt(ss) <= ctx->prm.residual_tolerance * bnorm;
    return mx < ctx->prm.residual_tolerance;
}

// z = M^-1 r with the solve's preconditioner: the MIC(0) sweeps, or the
// learned multigrid V-cycle (fsim_mg.hip), whose coarsest level the network
// solves.
static int precondition(fsim_ctx* ctx, const SolveArgs& S, bool flexible) {
    if (!flexible) return mic_apply(ctx, S, ctx->r, ctx->y, ctx->z);
    return mg_apply(ctx, ctx->r, ctx->z, S.scale);
}

// After a new z: z.r (and, flexible, z.q first), giving beta.  Plain PCG:
// beta = z'.r' / z.r (Fletcher-Reeves form of the reference).  Flexible PCG
// (the preconditioner is not a fixed linear map): beta = z'.(r' - r) / z.r
// = -alpha z'.q / z.r, since r' - r = -alpha q.
static int next_beta(fsim_ctx* ctx, bool flexible) {
    if (flexible) FSIM_TRY(dot_with(ctx, ctx->z, ctx->q, FIN_ZQ));
    return dot_zr(ctx, flexible ? FIN_FBETA : FIN_BETA);
}

struct PcgResult {
    int iterations = 0;
    bool converged = false;   // stopped on the FP64 true residual
    bool nan = false;         // stopped on a NaN alpha or sigma
};

// PCG from x = 0, r = b.  Reliable updates (restarted mixed-precision PCG):
// when the FP32 recursive residual meets the tolerance, or has dropped by
// kDelta (FSIM_OPT_RELIABLE_DROP) since the last FP64 true residual, the true residual b - A x is
// recomputed in FP64.  The solve stops if it meets the tolerance; otherwise r
// takes its value and the search direction restarts from z = M^-1 r.  (Keeping
// the direction instead, with beta from the replaced residual, was measured:
// 185 iterations against 145 at the 256^3 double dam break, 1e-6.)
//
// One host round trip per iteration: the iteration's scalars are copied after
// the r update, and the next preconditioner application and its dots run
// while the host waits for them.  If the solve stops or restarts, those
// results are discarded.  x += alpha s runs on the second stream
// (update_x_beside) and is joined before the next direction step.
static int pcg_run(fsim_ctx* ctx, const SolveArgs& S, double bnorm, bool flexible, int max_it, PcgResult* res) {
    const JoinXOnExit join_on_exit{ctx};
    const int64_t nc = S.nc;
    fsim_step_stats& st = ctx->stats;
    FSIM_TRY(fill_f64(ctx, ctx->x, 0.0, nc));
    hipLaunchKernelGGL(k_copy_f32, dim3(nblocks(nc)), dim3(kBlock), 0, ctx->stream, nc, ctx->b, ctx->r);
    FSIM_TRY(fill_f32(ctx, ctx->q, 0.0f, nc));
    FSIM_TRY(fill_f32(ctx, ctx->z, 0.0f, nc));
    FSIM_TRY(fill_f32(ctx, ctx->s, 0.0f, nc));
    FSIM_TRY(fill_f32(ctx, ctx->s2, 0.0f, nc));
    FSIM_TRY(precondition(ctx, S, flexible));
    FSIM_TRY(dot_zr(ctx, FIN_SIGMA0));
    const int max_reliable = 64;
    const double kDelta = ctx->reliable_drop > 0 ? std::pow(10.0, -ctx->reliable_drop) : 0.0;
    double last_true = bnorm;
    bool copy = true;
    int it = 0;
    res->converged = false;
    res->nan = false;
    for (; it < max_it; it++) {
        FSIM_TRY(join_x(ctx));
        FSIM_TRY(direction_step(ctx, S, copy));
        FSIM_TRY(update_x_beside(ctx));
        {
            FSIM_PROF(ctx, KID_UPDATE);
            if (nc % 4 == 0)
                hipLaunchKernelGGL(k_update_r<true>, dim3(stream_blocks(nc)), dim3(kBlock), 0, ctx->stream, nc,
                                   ctx->chunk_flags, ctx->q, ctx->r, reduce_to_host(ctx, FIN_RES));
            else
                hipLaunchKernelGGL(k_update_r<false>, dim3(stream_blocks(nc)), dim3(kBlock), 0, ctx->stream, nc,
                                   ctx->chunk_flags, ctx->q, ctx->r, reduce_to_host(ctx, FIN_RES));
        }
        FSIM_HIP(ctx, hipGetLastError());
        // the r update's finishing block wrote rr, rmax and nan_flag into scal_host; the event orders the read
        FSIM_HIP(ctx, hipEventRecord(ctx->res_ev, ctx->stream));
        FSIM_TRY(precondition(ctx, S, flexible));
        FSIM_TRY(next_beta(ctx, flexible));
        FSIM_HIP(ctx, hipEventSynchronize(ctx->res_ev));
        const SolveScalars h = *ctx->scal_host;
        // the previous iteration's old-sigma test, then this iteration's alpha test
        if (h.nan_flag & 2) {
            it--;
            res->nan = true;
            break;
        }
        if (h.nan_flag & 1) {
            res->nan = true;
            break;
        }
        bool restart = false;
        if (stop_met(ctx, h.rr, h.rmax, bnorm) || std::sqrt(h.rr) < kDelta * last_true) {
            FSIM_TRY(join_x(ctx));
            const bool replace = st.reliable_updates < max_reliable;
            FSIM_TRY(true_residual(ctx, S, replace));
            FSIM_TRY(read_scalars(ctx));
            if (stop_met(ctx, ctx->scal_host->tr_ss, ctx->scal_host->tr_max, bnorm)) {
                res->converged = true;
                break;
            }
            if (!replace) {
                if (stop_met(ctx, h.rr, h.rmax, bnorm)) break;
            } else {
                st.reliable_updates++;
                last_true = std::sqrt(ctx->scal_host->tr_ss);
                restart = true;
            }
        }
        if (restart) {
            FSIM_TRY(precondition(ctx, S, flexible));
            FSIM_TRY(dot_zr(ctx, FIN_SIGMA0));
        }
        copy = restart;
    }
    FSIM_TRY(join_x(ctx));
    res->iterations = it;
    return FSIM_OK;
}

// ---- BasicMacGrid::solveIncompressibility, basicMacGrid.cpp:5-102 -------------
// The reference's release (parallel) branch: each iteration one pass over the
// cells with x + y + z odd, then one over the even ones.  A WATER cell with s
// non-SOLID neighbours computes d = (-u - v - w + u(-x) + v(-y) + w(-z)
// [+ (avgPNum - averagePressure) pressureK]) * 1.98 / s and moves its six
// faces by +-d where the cell across is not SOLID.  Cells of one colour share
// no face, so one thread per cell of the colour needs no atomics and the
// result does not depend on the schedule.  The faces are FP64 during the
// sweeps (the reference's Face::v2 type): over many iterations of an
// over-relaxed sweep, FP32 rounding would grow far beyond the oracle's.
struct SorArgs {
    int nx, ny, nz, kz;
    const uint8_t* label;
    const float* density;
    double* f[3];
    double pk, avg_p;
    int pressure_enabled;
};

template <int ODD>
__global__ void __launch_bounds__(kBlock) k_sor_pass(SorArgs A) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t rows = (int64_t)(A.nx - 2) * (A.ny - 2);
    if (t >= rows * A.kz) return;
    const int k = (int)(t % A.kz);
    const int64_t r = t / A.kz;
    const int y = 1 + (int)(r % (A.ny - 2));
    const int x = 1 + (int)(r / (A.ny - 2));
    const int z = (ODD ? 1 + ((x + y) & 1) : 2 - ((x + y) & 1)) + 2 * k;
    if (z > A.nz - 2) return;
    const int64_t sx = (int64_t)A.ny * A.nz, sy = A.nz;
    const int64_t i = ((int64_t)x * A.ny + y) * A.nz + z;
    if (A.label[i] != FSIM_WATER) return;
    const int s1 = A.label[i + 1] != FSIM_SOLID, s2 = A.label[i - 1] != FSIM_SOLID;
    const int s3 = A.label[i + sy] != FSIM_SOLID, s4 = A.label[i - sy] != FSIM_SOLID;
    const int s5 = A.label[i + sx] != FSIM_SOLID, s6 = A.label[i - sx] != FSIM_SOLID;
    const int sum = s1 + s2 + s3 + s4 + s5 + s6;
    if (sum == 0) return;
    double* const u = A.f[0];
    double* const v = A.f[1];
    double* const w = A.f[2];
    double d = -u[i] - v[i] - w[i] + u[i - sx] + v[i - sy] + w[i - 1] +
               (A.pressure_enabled ? ((double)A.density[i] - A.avg_p) * A.pk : 0.0);
    d = d * 1.98 / sum;   // overRelaxation, basicMacGrid.cpp:13
    if (s1) w[i] += d;
    if (s2) w[i - 1] -= d;
    if (s3) v[i] += d;
    if (s4) v[i - sy] -= d;
    if (s5) u[i] += d;
    if (s6) u[i - sx] -= d;
}

__global__ void __launch_bounds__(kBlock) k_f32_to_f64(int64_t n, const float* __restrict__ a, double* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}
__global__ void __launch_bounds__(kBlock) k_f64_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

static int stage_basic_sor(fsim_ctx* ctx, const SolveArgs& S) {
    fsim_step_stats& st = ctx->stats;
    const int64_t nc = S.nc;
    if (!ctx->sor64) {
        if (hipMalloc(reinterpret_cast<void**>(&ctx->sor64), sizeof(double) * 3 * (size_t)nc) != hipSuccess) {
            ctx->sor64 = nullptr;
            return set_err(ctx, FSIM_ERR_OOM, "SOR face buffer allocation failed");
        }
    }
    FSIM_HIP(ctx, hipMemsetAsync(ctx->scal, 0, sizeof(SolveScalars), ctx->stream));
    FSIM_TRY(fill_f64(ctx, ctx->x, 0.0, nc));   // no pressure field
    FSIM_TRY(build_nbinfo(ctx, S, true));
    FSIM_TRY(read_scalars(ctx));
    const double b0 = std::sqrt(ctx->scal_host->bb);
    st.b_norm_l2 = b0;
    st.num_fluid_cells = (int64_t)ctx->scal_host->nfluid;
    SorArgs A{};
    A.nx = S.nx;
    A.ny = S.ny;
    A.nz = S.nz;
    A.kz = (S.nz - 1) / 2;   // z values of one colour in a row: at most this many
    A.label = ctx->label;
    A.density = ctx->density;
    float* f32[3] = {ctx->u2, ctx->v2, ctx->w2};
    for (int a = 0; a < 3; a++) A.f[a] = ctx->sor64 + (size_t)a * nc;
    A.pk = ctx->prm.pressure_k;
    A.avg_p = ctx->prm.average_pressure;
    A.pressure_enabled = ctx->prm.pressure_enabled;
    const int iters = ctx->prm.max_iterations;
    {
        FSIM_PROF(ctx, KID_SOR);
        for (int a = 0; a < 3; a++)
            hipLaunchKernelGGL(k_f32_to_f64, dim3(nblocks(nc)), dim3(kBlock), 0, ctx->stream, nc, f32[a], A.f[a]);
        const int64_t work = (int64_t)(S.nx - 2) * (S.ny - 2) * A.kz;
        for (int n = 0; n < iters; n++) {
            hipLaunchKernelGGL(k_sor_pass<1>, dim3(nblocks(work)), dim3(kBlock), 0, ctx->stream, A);
            hipLaunchKernelGGL(k_sor_pass<0>, dim3(nblocks(work)), dim3(kBlock), 0, ctx->stream, A);
        }
        for (int a = 0; a < 3; a++)
            hipLaunchKernelGGL(k_f64_to_f32, dim3(nblocks(nc)), dim3(kBlock), 0, ctx->stream, nc, A.f[a], f32[a]);
        FSIM_HIP(ctx, hipGetLastError());
    }
    // the right-hand side left on the projected faces (FSIM_FIELD_RHS reads it)
    FSIM_HIP(ctx, hipMemsetAsync(ctx->scal, 0, sizeof(SolveScalars), ctx->stream));
    FSIM_TRY(build_nbinfo(ctx, S, true));
    FSIM_TRY(read_scalars(ctx));
    st.residual_rel_l2 = b0 > 0 ? std::sqrt(ctx->scal_host->bb) / b0 : 0.0;
    st.residual_abs_inf = 0.0;
    st.pcg_iterations = iters;   // BasicMacGrid returns its iteration count (basicMacGrid.cpp:10)
    return FSIM_OK;
}

int stage_solve(fsim_ctx* ctx, double dt) {
    const SolveArgs S = solve_args(ctx, dt);
    const int64_t nc = S.nc;
    const unsigned nb = nblocks(nc);
    ctx->last_dt = dt;
    fsim_step_stats& st = ctx->stats;
    st.early_exit = 0;
    st.fallback = 0;
    st.reliable_updates = 0;
    st.pcg_iterations = 0;
    st.learned_iterations = 0;
    st.learned_stop = 0;
    st.solve_nan = 0;
    st.learned_budget = 0;
    if (ctx->prm.precond == FSIM_SOLVER_BASIC_SOR) return stage_basic_sor(ctx, S);
    const bool learned = ctx->prm.precond == FSIM_PRECOND_LEARNED;
    if (learned && ctx->mg_coarse == FSIM_MG_COARSE_NETWORK && !learned_ready(ctx))
        return set_err(ctx, FSIM_ERR_INVALID, "learned preconditioner selected but it has no parameters "
                                              "(fsim_learned_init or fsim_learned_set_params)");
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
    const int max_it = ctx->prm.max_iterations;
    PcgResult res;
    bool use_mic = !learned;
    if (learned && !ctx->mg_harvesting &&
        (ctx->last_mic_iterations <= 0 || (ctx->learned_refresh > 0 && ctx->learned_since_mic >= ctx->learned_refresh))) {
        // no MIC(0) count to budget against yet (or it is old): this solve is
        // MIC(0)-PCG, which measures it
        use_mic = true;
        st.learned_stop = 3;
    } else if (learned) {
        // flexible PCG with the network; MIC(0)-PCG from scratch when it does not
        // reach the tolerance within twice the last MIC(0) iteration count
        const int budget = ctx->mg_harvesting ? max_it : std::min(max_it, 2 * ctx->last_mic_iterations);
        st.learned_budget = budget;
        if (!ctx->mg_harvesting) ctx->learned_since_mic++;
        FSIM_TRY(mg_setup(ctx));
        FSIM_TRY(pcg_run(ctx, S, bnorm, true, budget, &res));
        st.learned_iterations = res.iterations;
        st.learned_stop = res.converged ? 0 : res.nan ? 2 : 1;
        if (!res.converged) {
            st.fallback = 1;
            use_mic = true;
            FSIM_HIP(ctx, hipMemsetAsync(ctx->scal, 0, sizeof(SolveScalars), ctx->stream));
        }
    }
    if (use_mic) {
        FSIM_TRY(mic_factor(ctx, S));
        FSIM_TRY(pcg_run(ctx, S, bnorm, false, max_it, &res));
        if (res.converged) {
            ctx->last_mic_iterations = res.iterations;
            ctx->learned_since_mic = 0;
        }
    }
    st.pcg_iterations = res.iterations;
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
    // a NaN alpha or sigma ended the solve: like the reference, the last
    // finite pressure has been applied and the step goes on
    st.solve_nan = res.nan ? 1 : 0;
    FSIM_TRY(check_wave_err(ctx));
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
    if (hipMalloc(&ctx->ybnd, sizeof(uint64_t) * ny_pk) != hipSuccess ||
        hipMalloc(&ctx->zbnd, sizeof(uint64_t) * nz_pk) != hipSuccess ||
        hipMalloc(&ctx->wave_ticket, sizeof(unsigned long long) * 3) != hipSuccess ||
        hipMalloc(&ctx->wave_err, sizeof(int)) != hipSuccess ||
        hipMalloc(&ctx->colw, (size_t)A.ny * A.nz) != hipSuccess)
        return set_err(ctx, FSIM_ERR_OOM, "wavefront buffer allocation failed");
    FSIM_TRY(fill_u8(ctx, ctx->colw, 0, (int64_t)A.ny * A.nz));
    FSIM_HIP(ctx, hipMemsetAsync(ctx->ybnd, 0, sizeof(uint64_t) * ny_pk, ctx->stream));
    FSIM_HIP(ctx, hipMemsetAsync(ctx->zbnd, 0, sizeof(uint64_t) * nz_pk, ctx->stream));
    FSIM_HIP(ctx, hipMemsetAsync(ctx->wave_ticket, 0, sizeof(unsigned long long) * 3, ctx->stream));
    FSIM_HIP(ctx, hipMemsetAsync(ctx->wave_err, 0, sizeof(int), ctx->stream));
    A.nbinfo = ctx->nbinfo;
    A.P64 = ctx->P64;
    A.P32 = ctx->P32;
    A.ybnd = ctx->ybnd;
    A.zbnd = ctx->zbnd;
    A.err = ctx->wave_err;
    A.colw = ctx->colw;
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
        FSIM_TRY(fill_f32(ctx, ctx->q, 0.0f, nc));   // the SpMV writes WATER cells only
        hipLaunchKernelGGL(k_spmv<SPMV_PLAIN>, dim3(red_blocks(nc)), dim3(kBlock), 0, ctx->stream, S, ctx->nbinfo, nullptr,
                           ctx->r, nullptr, ctx->q, reduce_to(ctx, FIN_NONE));
        FSIM_HIP(ctx, hipGetLastError());
        FSIM_HIP(ctx, hipMemcpyAsync(hout.data(), ctx->q, sizeof(float) * nc, hipMemcpyDeviceToHost, ctx->stream));
    } else if (op == FSIM_OP_APPLY_MIC0) {
        FSIM_TRY(mic_factor(ctx, S));
        FSIM_TRY(fill_f32(ctx, ctx->y, 0.0f, nc));
        FSIM_TRY(fill_f32(ctx, ctx->z, 0.0f, nc));
        FSIM_TRY(mic_apply(ctx, S, ctx->r, ctx->y, ctx->z));
        FSIM_TRY(check_wave_err(ctx));
        FSIM_HIP(ctx, hipMemcpyAsync(hout.data(), ctx->z, sizeof(float) * nc, hipMemcpyDeviceToHost, ctx->stream));
    } else if (op == FSIM_OP_APPLY_MG) {
        if (ctx->mg_coarse == FSIM_MG_COARSE_NETWORK && !learned_ready(ctx))
            return set_err(ctx, FSIM_ERR_INVALID, "the multigrid's coarse network has no parameters");
        FSIM_TRY(mg_setup(ctx));
        FSIM_TRY(fill_f32(ctx, ctx->z, 0.0f, nc));
        FSIM_TRY(mg_apply(ctx, ctx->r, ctx->z, S.scale));
        FSIM_HIP(ctx, hipMemcpyAsync(hout.data(), ctx->z, sizeof(float) * nc, hipMemcpyDeviceToHost, ctx->stream));
    } else {
        return set_err(ctx, FSIM_ERR_INVALID, "unknown solver op");
    }
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < nc; i++) out[i] = hout[rmap[i]];
    return FSIM_OK;
}

}  // namespace fsim
