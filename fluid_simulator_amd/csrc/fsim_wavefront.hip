// fsim_wavefront.hip — persistent wavefront kernels for the MIC(0) factor
// and the two triangular sweeps (bridsonSolverGrid.cpp:91-163).
//
// Each cell depends only on its three upstream neighbours (-x,-y,-z for the
// factor and the forward sweep, +x,+y,+z for the backward sweep), so the
// serial lexicographic loop of the reference can run on i+j+k planes.
//
// Decomposition.  One workgroup owns a tile of WF_TY x WF_TZ interior (y, z)
// columns and marches along x.  Lane (ty, tz) of the tile processes
// x = t - ty - tz at step t, so its three upstream values were produced at
// step t-1: the -x value by the lane itself, the -z value by lane tz-1 of the
// same wave (DPP wave shift) and the -y value by wave ty-1 (LDS).
//
// Inside a workgroup the WF_TY compute waves form a pipeline along y with
// point-to-point synchronisation: wave ty publishes its row of step t into an
// LDS ring and then its progress counter; wave ty+1 waits only for that
// counter, and wave ty waits for wave ty+1's counter before reusing a ring
// slot.  There is no workgroup barrier in the main loop.
//
// Across workgroups, values at the tile edges travel through boundary
// buffers in HBM as 64-bit packets (launch epoch << 32 | value bits) written
// with device-scope stores; a reader accepts a packet when its epoch tag
// matches, so no flags or fences are needed and a packet is never seen
// half-written.  Two communication waves per workgroup (one for the -y edge,
// one for the -z edge) read the upstream tile's packets in batches and
// republish them into LDS rings tagged with the step that needs them.  This
// keeps every global read whose result is needed "now" out of the compute
// waves: on gfx9 the memory counter completes in order, so such a read would
// drain the compute waves' prefetch stream (per-cell inputs are loaded
// kPrefetch steps ahead into registers).
//
// Every wait is bounded: past kSpinLimit polls, or as soon as any wave in the
// grid has given up, a wave raises the error flag and carries on, so the
// grid always drains.  Workgroups take tiles from a ticket counter in
// row-major (J, K) order, so the tiles a workgroup waits on were taken by
// workgroups already running.
#include <hip/hip_runtime.h>

#include "fsim_internal.hpp"

namespace fsim {

constexpr int kSpinLimit = 1 << 21;
constexpr int kPrefetch = 8;     // compute-wave register prefetch (steps)
constexpr int kRingV = 8;        // LDS ring of exchanged values between compute waves (steps)
constexpr int kRingP = 64;       // LDS ring of boundary packets (steps)

template <int KIND>
struct KindTraits {   // 1 = forward sweep, 2 = backward sweep
    using T = float;
    static constexpr bool FWD = KIND == 1;
    static constexpr int NP = 1;      // packets per boundary value
    static constexpr int BATCH = 32;  // comm-wave batch (steps)
};
template <>
struct KindTraits<0> {   // MIC(0) factor
    using T = double;
    static constexpr bool FWD = true;
    static constexpr int NP = 2;
    static constexpr int BATCH = 16;
};

// f(integral_constant<int, 0>), ..., f(integral_constant<int, N-1>) in order
template <int N, int I = 0, class F>
__device__ __forceinline__ void unroll(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        unroll<N, I + 1>(f);
    }
}

__device__ inline uint64_t pack(uint32_t bits, uint32_t tag) { return ((uint64_t)tag << 32) | bits; }
__device__ inline void gput(uint64_t* p, uint64_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline uint64_t gpeek(const uint64_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void lput(uint64_t* p, uint64_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }
__device__ inline uint64_t lpeek(const uint64_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }
__device__ inline void lput32(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }
__device__ inline int lpeek32(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }
__device__ inline bool gave_up(const int* err) { return __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0; }

// Wait for all of this wave's LDS operations (vector memory counter untouched).
__device__ inline void lds_drain() {
    __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0), vmcnt/expcnt at their maxima
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Spin until the LDS counter *p reaches v (bounded).
__device__ inline void wait_counter(const int* p, int v, int* err) {
    int spins = 0;
    while (lpeek32(p) < v) {
        if (++spins > kSpinLimit) {
            atomicOr(err, 1);
            break;
        }
        if ((spins & 1023) == 0 && gave_up(err)) break;
        __builtin_amdgcn_s_sleep(1);
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Spin until the LDS packet slot carries step tag t; returns its payload.
__device__ inline uint32_t wait_slot(const uint64_t* p, int t, int* err) {
    uint64_t v = lpeek(p);
    int spins = 0;
    while ((int)(uint32_t)(v >> 32) != t) {
        if (++spins > kSpinLimit) {
            atomicOr(err, 1);
            break;
        }
        if ((spins & 1023) == 0 && gave_up(err)) break;
        __builtin_amdgcn_s_sleep(1);
        v = lpeek(p);
    }
    return (uint32_t)v;
}

// lane i receives lane i-1's value, lane 0 receives 0 (DPP wave_shr:1)
__device__ inline float shr1(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
}
__device__ inline double shr1(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, (int)__double2loint(v), 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)__double2hiint(v), 0x138, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

__device__ inline uint32_t word(float v, int) { return __float_as_uint(v); }
__device__ inline uint32_t word(double v, int k) { return k ? (uint32_t)__double2hiint(v) : (uint32_t)__double2loint(v); }
__device__ inline void from_words(float& v, const uint32_t* w) { v = __uint_as_float(w[0]); }
__device__ inline void from_words(double& v, const uint32_t* w) { v = __hiloint2double((int)w[1], (int)w[0]); }

// Ticket in [0, TJ*TK), or -1 (with *err raised) if the counter is out of
// step with the launch (only after a failed launch); the caller then exits.
__device__ inline int take_tile(const WaveArgs& A) {
    __shared__ int s_tile;
    if (threadIdx.x == 0) {
        long long t = (long long)(atomicAdd(A.ticket, 1ull) - A.ticket_base);
        if (t < 0 || t >= (long long)A.TJ * A.TK) {
            atomicOr(A.err, 2);
            t = -1;
        }
        s_tile = (int)t;
    }
    __syncthreads();
    return s_tile;
}

// Geometry of one tile position (local coordinates mirrored for the backward
// direction so that "upstream" is always ty-1, tz-1, xl-1).
struct Pos {
    int y, z;
    bool col;
};
template <bool FWD>
__device__ inline Pos pos_of(const WaveArgs& A, int J, int K, int ty, int tz) {
    Pos P;
    const int ly = J * WF_TY + ty, lz = K * WF_TZ + tz;
    P.col = ly < A.ny - 2 && lz < A.nz - 2;
    P.y = FWD ? 1 + ly : A.ny - 2 - ly;
    P.z = FWD ? 1 + lz : A.nz - 2 - lz;
    return P;
}
template <bool FWD>
__device__ inline int x_of(const WaveArgs& A, int xl) { return FWD ? 1 + xl : A.nx - 2 - xl; }

template <int KIND>
struct WaveLds {
    using T = typename KindTraits<KIND>::T;
    static constexpr int NP = KindTraits<KIND>::NP;
    T v[kRingV][WF_TY][WF_TZ];
    uint64_t yp[kRingP][NP][WF_TZ];
    uint64_t zp[kRingP][NP][WF_TY];
    int prog[WF_TY];
};

// ---- communication wave: upstream tile edge -> LDS packet ring ------------
// Lane l serves one consumer lane: for the -y edge the consumer is (ty 0,
// tz l) of this tile, for the -z edge it is (ty l, tz 0).  The consumer at
// step s needs the upstream value at x-index xl = s - l.
template <int KIND, bool YEDGE>
__device__ void comm_wave(const WaveArgs& A, WaveLds<KIND>& S, int J, int K) {
    using TR = KindTraits<KIND>;
    constexpr bool FWD = TR::FWD;
    constexpr int NP = TR::NP, B = TR::BATCH;
    const int l = threadIdx.x % 64;
    const int nxi = A.nx - 2;
    const bool has_up = YEDGE ? J > 0 : K > 0;
    const int cty = YEDGE ? 0 : l, ctz = YEDGE ? l : 0;
    const bool lane_ok = has_up && (YEDGE || l < WF_TY) && pos_of<FWD>(A, J, K, cty, ctz).col;
    const Pos P = pos_of<FWD>(A, J, K, min(cty, WF_TY - 1), ctz);
    const uint64_t* base = YEDGE ? A.ybnd + (int64_t)NP * ((int64_t)(J - 1) * A.nx) * A.nz
                                 : A.zbnd + (int64_t)NP * ((int64_t)(K - 1) * A.nx) * A.ny;
    const int64_t stride = YEDGE ? A.nz : A.ny;          // per x
    const int64_t off = YEDGE ? P.z : P.y;
    const int* cprog = &S.prog[YEDGE ? 0 : min(l, WF_TY - 1)];
    int s = lane_ok ? l : 0;                              // next step to publish
    const int end = lane_ok ? l + nxi : 0;
    int idle = 0;
    while (__any(s < end)) {
        uint64_t buf[B][NP];
#pragma unroll
        for (int b = 0; b < B; b++) {
            const int xl = min(max(s + b - l, 0), nxi - 1);
            const uint64_t* p = lane_ok ? base + NP * ((int64_t)x_of<FWD>(A, xl) * stride + off) : A.ybnd;
#pragma unroll
            for (int k = 0; k < NP; k++) buf[b][k] = gpeek(p + (lane_ok ? k : 0));
        }
        const int limit = lpeek32(cprog) + kRingP;        // slots free up to this step
        int published = 0;
        bool stop = false;
#pragma unroll
        for (int b = 0; b < B; b++) {
            const int sb = s + b;
            bool ok = !stop && sb < end && sb <= limit;
#pragma unroll
            for (int k = 0; k < NP; k++) ok = ok && (uint32_t)(buf[b][k] >> 32) == A.epoch;
            if (ok) {
                uint64_t* slot = YEDGE ? &S.yp[sb % kRingP][0][l] : &S.zp[sb % kRingP][0][min(l, WF_TY - 1)];
#pragma unroll
                for (int k = 0; k < NP; k++)
                    lput(slot + k * (YEDGE ? WF_TZ : WF_TY), pack((uint32_t)buf[b][k], (uint32_t)sb));
                published++;
            } else {
                stop = true;
            }
        }
        s += published;
        if (__any(published > 0)) {
            idle = 0;
        } else {
            if (++idle > kSpinLimit / 64) {
                atomicOr(A.err, 1);
                break;
            }
            if ((idle & 15) == 0 && gave_up(A.err)) break;
            __builtin_amdgcn_s_sleep(2);
        }
    }
}

// ---- compute waves -----------------------------------------------------------
template <int KIND>
__device__ void compute_wave(const WaveArgs& A, WaveLds<KIND>& S, int J, int K) {
#pragma clang fp contract(off)
    using TR = KindTraits<KIND>;
    using T = typename TR::T;
    constexpr bool FWD = TR::FWD;
    constexpr int NP = TR::NP;
    const int ty = threadIdx.x / 64, tz = threadIdx.x % 64;
    const Pos P = pos_of<FWD>(A, J, K, ty, tz);
    const int nxi = A.nx - 2;
    const int nsteps = nxi + WF_TY + WF_TZ - 2;
    const bool cy = ty == 0 && J > 0, cz = tz == 0 && K > 0;
    uint64_t* ybout = A.ybnd + (int64_t)NP * ((int64_t)J * A.nx) * A.nz;
    uint64_t* zbout = A.zbnd + (int64_t)NP * ((int64_t)K * A.nx) * A.ny;
    const uint32_t mx = FWD ? NB_MX : NB_PX, my = FWD ? NB_MY : NB_PY, mz = FWD ? NB_MZ : NB_PZ;
    const float ms = A.ms;
    const double scale = A.scale;
    // per-cell inputs: packed (nbinfo << 32 | P32) words for the sweeps (plus
    // the input vector), nbinfo for the factor; loads always issued
    // (clamped to a valid cell) so the compiler can count them statically
    uint64_t wring[kPrefetch];
    float iring[kPrefetch];
    auto fetch = [&](int xl, uint64_t& w, float& in) {
        const int xc = min(max(xl, 0), nxi - 1);
        const int64_t i = ((int64_t)x_of<FWD>(A, xc) * A.ny + (P.col ? P.y : 1)) * A.nz + (P.col ? P.z : 1);
        if (KIND == 0) {
            w = A.nbinfo[i];
            in = 0.f;
        } else {
            w = A.W[i];
            in = A.in[i];
        }
    };
#pragma unroll
    for (int d = 0; d < kPrefetch; d++) fetch(d - ty - tz, wring[d], iring[d]);
    T vprev = 0;
    for (int t0 = 0; t0 < nsteps; t0 += kPrefetch) {
        // compile-time step index, so each ring slot is its own register
        unroll<kPrefetch>([&](auto dc) {
            constexpr int d = decltype(dc)::value;
            const int t = t0 + d;
            const int xl = t - ty - tz;
            const bool active = P.col && xl >= 0 && xl < nxi && t < nsteps;
            const uint64_t w = wring[d];
            const float in = iring[d];
            fetch(xl + kPrefetch, wring[d], iring[d]);
            T vz = shr1(vprev);
            T vy = 0;
            if (ty > 0) {
                wait_counter(&S.prog[ty - 1], t - 1, A.err);
                vy = S.v[(t + kRingV - 1) % kRingV][ty - 1][tz];
            }
            T vnew = 0;
            if (active) {
                const int x = x_of<FWD>(A, xl);
                const int64_t i = ((int64_t)x * A.ny + P.y) * A.nz + P.z;
                if (tz == 0) {
                    vz = 0;
                    if (cz) {
                        uint32_t wd[NP];
#pragma unroll
                        for (int k = 0; k < NP; k++) wd[k] = wait_slot(&S.zp[t % kRingP][k][ty], t, A.err);
                        from_words(vz, wd);
                    }
                }
                if (ty == 0) {
                    vy = 0;
                    if (cy) {
                        uint32_t wd[NP];
#pragma unroll
                        for (int k = 0; k < NP; k++) wd[k] = wait_slot(&S.yp[t % kRingP][k][tz], t, A.err);
                        from_words(vy, wd);
                    }
                }
                const uint32_t info = KIND == 0 ? (uint32_t)w : (uint32_t)(w >> 32);
                if (KIND == 0) {
                    if (info & NB_FLUID) {
                        const double tau = 0.97, sigma = 0.25;
                        const double up[3] = {(double)vprev, (double)vy, (double)vz};
                        double eneg = 0, eneg_tau = 0;
#pragma unroll
                        for (int a = 0; a < 3; a++) {
                            if (!(info & (NB_MX << a))) continue;
                            const double pj = up[a];
                            const double ap = -scale * pj;
                            const double o1 = (info & (1u << (NB_UP_SHIFT + 2 * a))) ? -scale : 0.0;
                            const double o2 = (info & (1u << (NB_UP_SHIFT + 2 * a + 1))) ? -scale : 0.0;
                            eneg += ap * ap;
                            eneg_tau += ap * (o1 + o2) * pj;
                        }
                        const int cnt = (info >> NB_NONSOLID_SHIFT) & 7;
                        double diag = 0;
                        for (int k = 0; k < cnt; k++) diag += scale;
                        double e = diag - eneg - eneg_tau * tau;
                        if (e < sigma * diag) e = diag < 1e-6 ? 1.0 : diag;
                        const double p = 1.0 / sqrt(e);
                        vnew = (T)p;
                        A.P64[i] = p;
                        const float p32 = (float)p;
                        A.P32[i] = p32;
                        A.W[i] = ((uint64_t)info << 32) | __float_as_uint(p32);
                    } else {
                        A.W[i] = (uint64_t)info << 32;
                    }
                } else if (info & NB_FLUID) {
                    const float p = __uint_as_float((uint32_t)w);
                    float acc = 0.f;
                    if (info & mx) acc += ms * (float)vprev;
                    if (info & my) acc += ms * (float)vy;
                    if (info & mz) acc += ms * (float)vz;
                    if (FWD) {
                        const float q = (in - acc) * p;
                        A.out[i] = q;
                        vnew = (T)(q * p);
                    } else {
                        const float zz = (in - acc * p) * p;
                        A.out[i] = zz;
                        vnew = (T)zz;
                    }
                }
                if (ty == WF_TY - 1) {
#pragma unroll
                    for (int k = 0; k < NP; k++)
                        gput(ybout + NP * ((int64_t)x * A.nz + P.z) + k, pack(word(vnew, k), A.epoch));
                }
                if (tz == WF_TZ - 1) {
#pragma unroll
                    for (int k = 0; k < NP; k++)
                        gput(zbout + NP * ((int64_t)x * A.ny + P.y) + k, pack(word(vnew, k), A.epoch));
                }
            }
            // reuse of ring slot t: the downstream wave must have consumed step t - kRingV
            if (ty < WF_TY - 1) wait_counter(&S.prog[ty + 1], t - kRingV + 1, A.err);
            S.v[t % kRingV][ty][tz] = vnew;
            lds_drain();
            if (tz == 0) lput32(&S.prog[ty], t);
            vprev = vnew;
        });
    }
}

template <int KIND>
__global__ void __launch_bounds__(WF_THREADS) k_wave(WaveArgs A) {
    __shared__ WaveLds<KIND> S;
    const int tile = take_tile(A);
    if (tile < 0) return;
    const int J = tile / A.TK, K = tile % A.TK;
    constexpr int NP = KindTraits<KIND>::NP;
    for (int i = threadIdx.x; i < kRingP * NP * WF_TZ; i += WF_THREADS) (&S.yp[0][0][0])[i] = ~0ull;
    for (int i = threadIdx.x; i < kRingP * NP * WF_TY; i += WF_THREADS) (&S.zp[0][0][0])[i] = ~0ull;
    if (threadIdx.x < WF_TY) S.prog[threadIdx.x] = -1;
    __syncthreads();
    const int wave = threadIdx.x / 64;
    if (wave < WF_TY) compute_wave<KIND>(A, S, J, K);
    else if (wave == WF_TY) comm_wave<KIND, true>(A, S, J, K);
    else comm_wave<KIND, false>(A, S, J, K);
}

int wave_launch(fsim_ctx* ctx, int kind, const float* in, float* out, double scale) {
    WaveArgs A = ctx->wave;
    A.in = in;
    A.out = out;
    A.scale = scale;
    A.ms = -(float)scale;
    const unsigned ntiles = (unsigned)(A.TJ * A.TK);
    A.ticket = ctx->wave_ticket + kind;
    A.ticket_base = ctx->wave_ticket_base[kind];
    ctx->wave_ticket_base[kind] += ntiles;
    A.epoch = ++ctx->wave_epoch;
    if (kind == 0) hipLaunchKernelGGL(k_wave<0>, dim3(ntiles), dim3(WF_THREADS), 0, ctx->stream, A);
    else if (kind == 1) hipLaunchKernelGGL(k_wave<1>, dim3(ntiles), dim3(WF_THREADS), 0, ctx->stream, A);
    else hipLaunchKernelGGL(k_wave<2>, dim3(ntiles), dim3(WF_THREADS), 0, ctx->stream, A);
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

}  // namespace fsim
