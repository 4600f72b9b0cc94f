// fsim_wavefront.hip — persistent wavefront kernels for the MIC(0) factor
// and the two triangular sweeps (bridsonSolverGrid.cpp:91-163).
//
// Each cell depends only on its three upstream neighbours (-x,-y,-z for the
// factor and the forward sweep, +x,+y,+z for the backward sweep), so the
// serial lexicographic loop of the reference can run on i+j+k planes; every
// cell performs the same operations in the same order as the serial loop, so
// the result is bit-identical to it whatever the schedule.
//
// Decomposition.  One workgroup owns a tile of WF_TY x 64 interior (y, z)
// columns and marches along x.  Lane tz of tile row ty processes
// x = t - ty - tz at step t, so its three upstream values were produced at
// step t-1: the -x value by the lane itself, the -z value by lane tz-1 of the
// same row (DPP wave shift) and the -y value by row ty-1: a register when the
// two rows belong to the same wave (WF_R rows per wave), otherwise the LDS
// ring below.  The backward sweep mirrors y, and z inside each 64-wide
// z-tile, so that its upstream is also ty-1, tz-1, x-1 in local terms.
//
// Memory.  The solver vectors use the rotated layout (ridx): the 64 cells a
// row visits at one step share one rotated plane, so every load and store
// of a step covers 64 contiguous words.  Each row keeps its plane offset in
// a scalar register (advanced and wrapped per step) and each lane a constant
// in-plane offset, so addressing costs no vector instructions; lanes outside
// the grid carry an out-of-range buffer offset (loads read 0, stores drop).
//
// Inside a workgroup the compute waves form a pipeline along y.  The last
// row of each wave is written into an LDS ring as 64-bit words
// (step << 32 | value bits); the next wave reads the word it needs and
// accepts it when the step tag matches, so a step costs one LDS round trip
// and needs no counters, fences or drains (an aligned 64-bit LDS access is
// never seen half-written).  Ring reuse is checked every kBpEvery steps
// against the downstream wave's own tags.
//
// Across workgroups, values at the tile edges travel through boundary
// buffers in HBM as 64-bit packets (launch epoch << 32 | value bits) written
// with device-scope stores, one packet row per step of the consuming tile
// (ybnd[tile][s][64], zbnd[tile][s][WF_TY] of the producing tile), so a
// step's packets are contiguous.
// Four communication waves per workgroup keep all of this traffic out of
// the compute waves: two read the upstream tiles' packets, a few steps at a
// time, and republish them into LDS rings tagged with the step; two take this
// tile's edge values from LDS out-rings (the last compute wave's row, lane 63
// of every row) and store them as packets.  On gfx9 the memory counter
// completes in order, so a global read whose result is needed "now", or a
// slowly acknowledged device-scope store, would stall the compute waves'
// prefetch stream (per-cell inputs are loaded kPrefetch steps ahead, each
// prefetch slot in a fixed register).
//
// Every wait is bounded: past kSpinLimit polls, or as soon as any wave in the
// grid has given up, a wave raises the error flag and carries on, so the
// grid always drains.  Workgroups take tiles from a ticket counter in
// row-major (J, K) order, so the tiles a workgroup waits on were taken by
// workgroups already running.
#include <hip/hip_runtime.h>

#include <climits>
#include <type_traits>

#include "fsim_internal.hpp"

namespace fsim {

constexpr int kSpinLimit = 1 << 21;
constexpr int kBusyPolls = 32;   // polls before a waiting wave starts to sleep
constexpr int kPrefetch = WF_PREFETCH;   // compute-wave register prefetch (steps)
constexpr int kRingV = kPrefetch;   // LDS ring of row values between compute waves (steps)
constexpr int kBpEvery = 4;      // ring-reuse check period (steps); allowed lag kRingV - kBpEvery
constexpr int kRingO = 32;       // LDS rings of this tile's +y / +z edge values (steps), a power of two
constexpr uint32_t kOob = 0x80000000u;        // buffer offset beyond every array: reads 0, drops stores
constexpr int kRawBufferFlags = 0x00020000;   // buffer resource word 3 (raw 32-bit access, gfx9)

template <int KIND>
struct Kind {   // 1 = forward sweep, 2 = backward sweep (FP32)
    using T = float;
    static constexpr bool FWD = KIND == 1;
    static constexpr int NP = 1;       // 32-bit words per value
    static constexpr int BATCH = 4;    // comm-wave poll batch (steps): short polls keep hop latency low
    static constexpr int RINGP = 64;   // LDS ring of boundary packets (steps)
};
template <>
struct Kind<0> {   // MIC(0) factor (FP64)
    using T = double;
    static constexpr bool FWD = true;
    static constexpr int NP = 2;
    static constexpr int BATCH = 4;
    static constexpr int RINGP = 32;
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
__device__ inline int tag_of(uint64_t v) { return (int)(uint32_t)(v >> 32); }
__device__ inline uint64_t gpeek(const uint64_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void lput(uint64_t* p, uint64_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }
__device__ inline uint64_t lpeek(const uint64_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }
__device__ inline void lput32(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }
__device__ inline int lpeek32(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }
__device__ inline bool gave_up(const int* err) { return __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0; }

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
// 16-byte LDS accesses (ds_read_b128 / ds_write_b128) through an explicit LDS pointer
typedef __attribute__((address_space(3))) volatile u64x2 lds_u64x2;
__device__ inline u64x2 lpeek2(const uint64_t* p) { return *(lds_u64x2*)(p); }
__device__ inline void lput2(uint64_t* p, u64x2 v) { *(lds_u64x2*)(p) = v; }
__device__ inline __amdgpu_buffer_rsrc_t rsrc(const void* p, int64_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, kRawBufferFlags);
}
__device__ inline uint64_t bload64(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ inline float bload32f(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ inline uint32_t bload16(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0);
}
template <int AUX = 0>
__device__ inline void bstore64(uint64_t v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, AUX);
}
__device__ inline void bstore32(uint32_t v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(v, r, voff, soff, 0);
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
__device__ inline void from_words(float& v, const uint64_t* w) { v = __uint_as_float((uint32_t)w[0]); }
__device__ inline void from_words(double& v, const uint64_t* w) {
    v = __hiloint2double((int)(uint32_t)w[1], (int)(uint32_t)w[0]);
}

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

// Grid y and z of local row ty, lane tz of tile (J, K).
template <bool FWD>
__device__ inline int y_of(const WaveArgs& A, int J, int ty) { return FWD ? 1 + J * WF_TY + ty : A.ny - 2 - (J * WF_TY + ty); }
template <bool FWD>
__device__ inline int z_of(const WaveArgs& A, int K, int tz) {
    return FWD ? 1 + K * WF_TZ + tz : 1 + (A.TK - 1 - K) * WF_TZ + (WF_TZ - 1 - tz);
}

// 1 if column (ty, tz) of tile (J, K) is interior and holds a WATER cell at any x.
template <bool FWD>
__device__ inline int col_water(const WaveArgs& A, int J, int K, int ty, int tz) {
    const int z = z_of<FWD>(A, K, tz);
    if (J * WF_TY + ty >= A.ny - 2 || z > A.nz - 2) return 0;
    return A.colw[y_of<FWD>(A, J, ty) * A.nz + z];
}

template <int KIND>
struct WaveLds {
    static constexpr int NP = Kind<KIND>::NP, RINGP = Kind<KIND>::RINGP;
    // tagged last-row values of compute waves 0 .. WF_NW-2 (the last wave's go to yo)
    uint64_t v[WF_NW - 1][kRingV][NP][WF_TZ];
    uint64_t yp[RINGP][NP][WF_TZ];          // tagged -y packets for row 0
    alignas(16) uint64_t zp[RINGP][NP][WF_TY];   // tagged -z packets for lane 0 of each row
    int prog[WF_NW];                        // a step whose packets the compute wave has consumed
    uint64_t yo[kRingO][NP][WF_TZ];         // tagged last row of the last compute wave: the +y edge
    alignas(16) uint64_t zo[kRingO][NP][WF_TY];   // tagged lane-63 value of every row: the +z edge
    // sweeps: +z writes of lanes 0..62, never read
    alignas(16) uint64_t zsink[KIND == 0 ? 1 : WF_NW][WF_TZ - 1][NP * WF_TY];
    int sprog[2];                           // compute step up to which the +y / +z edge has been sent
};

// ---- communication wave: upstream tile edge -> LDS packet ring ------------
// The -y edge carries 64 values per step (row 0, lane l), the -z edge WF_TY
// (lane 0 of row l).  The upstream tile writes every packet of a step at
// once, so the wave forwards whole steps: step s once all its packets carry
// this launch's epoch and the ring slot s mod RINGP is free.
template <int KIND, bool YEDGE>
__device__ void comm_wave(const WaveArgs& A, WaveLds<KIND>& S, int J, int K, bool yup, bool zup) {
    using KD = Kind<KIND>;
    constexpr int NP = KD::NP, B = KD::BATCH, RINGP = KD::RINGP;
    constexpr int NL = YEDGE ? WF_TZ : WF_TY;    // values per step
    if (YEDGE ? !yup : !zup) return;   // no upstream tile, or one without WATER (its edge is all zeros)
    const int l = threadIdx.x % 64;
    const bool on = l < NL;
    const int li = on ? l : 0;
    const int nxi = A.nx - 2;
    // steps 0 .. end-1 are consumed: row 0 lane l at l .. l+nxi-1 (-y edge),
    // lane 0 of the rows of wave w at w .. w+nxi-1 (-z edge)
    const int end = (YEDGE ? WF_TZ - 1 : WF_NW - 1) + nxi;
    // packets written by the upstream tile (J-1, K) or (J, K-1)
    const int64_t up = YEDGE ? (int64_t)(J - 1) * A.TK + K : (int64_t)J * A.TK + K - 1;
    const uint64_t* base = YEDGE ? A.ybnd + up * A.SP * WF_TZ * NP : A.zbnd + up * A.SP * WF_TY * NP;
    // the consumer of lane l: wave 0 for the -y edge, the wave holding row l for the -z edge
    const int* cprog = &S.prog[YEDGE ? 0 : li / WF_R];
    int s0 = 0, idle = 0;
    while (s0 < end) {
        uint64_t buf[B][NP];
#pragma unroll
        for (int b = 0; b < B; b++) {
            const int sb = min(s0 + b, end - 1);
            const uint64_t* p = base + ((int64_t)sb * NL + li) * NP;
#pragma unroll
            for (int k = 0; k < NP; k++) buf[b][k] = gpeek(p + k);
        }
        const int limit = lpeek32(cprog) + RINGP;   // slot of step sb is free if sb - RINGP was consumed
        int published = 0;
        bool stop = false;
#pragma unroll
        for (int b = 0; b < B; b++) {
            const int sb = s0 + b;
            bool ok = !on || sb <= limit;
#pragma unroll
            for (int k = 0; k < NP; k++) ok = ok && (!on || (uint32_t)(buf[b][k] >> 32) == A.epoch);
            stop = stop || sb >= end || !__all(ok);
            if (!stop) {
                if (on) {
                    uint64_t* slot = YEDGE ? &S.yp[sb % RINGP][0][li] : &S.zp[sb % RINGP][0][li];
#pragma unroll
                    for (int k = 0; k < NP; k++) lput(slot + k * NL, pack((uint32_t)buf[b][k], (uint32_t)sb));
                }
                published++;
            }
        }
        s0 += published;
        if (published > 0) {
            idle = 0;
        } else {
            if (++idle > kSpinLimit / 64) {
                if (l == 0) atomicOr(A.err, 1);
                break;
            }
            if ((idle & 15) == 0 && gave_up(A.err)) break;
            __builtin_amdgcn_s_sleep(2);
        }
    }
}

// ---- sending wave: this tile's edge (LDS out-ring) -> HBM packets ----------
// The packet of downstream step s is the value the compute waves produced at
// step s + lag (lag NW-1 for the +y edge: the last row of wave NW-1; 63 for
// the +z edge: lane 63 of every row).  The wave waits until every value of
// that step carries its step tag, stores the step's packets with
// device-scope stores and reports its progress every 4 steps, which the
// compute waves check before reusing an out-ring slot.  The device-scope
// stores stay out of the compute waves, whose memory counter also orders
// their prefetch loads.
template <int KIND, bool YEDGE>
__device__ void send_wave(const WaveArgs& A, WaveLds<KIND>& S, int J, int K) {
    constexpr int NP = Kind<KIND>::NP;
    constexpr int NL = YEDGE ? WF_TZ : WF_TY;
    int* prog = &S.sprog[YEDGE ? 0 : 1];
    const int l = threadIdx.x % 64;
    if (YEDGE ? J == A.TJ - 1 : K == A.TK - 1) return;   // no downstream tile: progress preset to "all sent"
    const bool on = l < NL;
    const int li = on ? l : 0;
    const int lag = YEDGE ? WF_NW - 1 : WF_TZ - 1;
    const int end = (YEDGE ? WF_TZ - 1 : WF_NW - 1) + (A.nx - 2);   // downstream steps consumed
    const int64_t tile = (int64_t)J * A.TK + K;
    uint64_t* out = YEDGE ? A.ybnd + tile * A.SP * WF_TZ * NP : A.zbnd + tile * A.SP * WF_TY * NP;
    for (int s = 0; s < end; s++) {
        const int t = s + lag;
        const uint64_t* src = YEDGE ? &S.yo[t & (kRingO - 1)][0][li] : &S.zo[t & (kRingO - 1)][0][li];
        uint64_t v[NP];
        int spins = 0;
        for (;;) {
            bool ok = true;
#pragma unroll
            for (int k = 0; k < NP; k++) {
                v[k] = lpeek(src + k * NL);
                ok = ok && (!on || tag_of(v[k]) == t);
            }
            if (__all(ok)) break;
            if (++spins > kSpinLimit || ((spins & 1023) == 0 && gave_up(A.err))) {
                if (l == 0) {
                    atomicOr(A.err, 1);
                    lput32(prog, INT_MAX);   // never block the compute waves
                }
                return;
            }
            if (spins > kBusyPolls) __builtin_amdgcn_s_sleep(1);
        }
        if (on) {
#pragma unroll
            for (int k = 0; k < NP; k++)
                __hip_atomic_store(out + ((int64_t)s * NL + li) * NP + k, pack((uint32_t)v[k], A.epoch), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        if ((s & 3) == 3 && l == 0) lput32(prog, t);
        for (int i = 0; i < A.send_delay; i++) __builtin_amdgcn_s_sleep(127);
    }
    if (l == 0) lput32(prog, INT_MAX);
}

// ---- compute waves -------------------------------------------------------------
// Factor (KIND 0, FP64):   e = diag - sum_a (A_a p_a)^2 - tau sum_a A_a (o1 + o2) p_a^2,
//                          guarded by sigma, p = 1/sqrt(e)  (calculatePreconditioner)
// forward sweep (KIND 1):  q = (r - sum_{-a fluid} A_a q_a p_a) * p, wave value q*p
// backward sweep (KIND 2): z = (q - p * sum_{+a fluid} A_a z_a) * p, wave value z
// with A_a = -scale, products and sums in the reference's order (x, y, z).
// In the sweeps a neighbour that is not WATER carries exactly +0, whose term
// -scale * (+0) = -0 leaves any partial sum unchanged, so they sum all three.
template <int KIND>
__device__ void compute_wave(const WaveArgs& A, WaveLds<KIND>& S, int J, int K, bool yup, bool zup) {
#pragma clang fp contract(off)
    using KD = Kind<KIND>;
    using T = typename KD::T;
    constexpr bool FWD = KD::FWD;
    constexpr int NP = KD::NP, RINGP = KD::RINGP, R = WF_R, NW = WF_NW, TY = WF_TY;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const int tz = threadIdx.x % 64;
    const int nx = A.nx, nxi = A.nx - 2;
    const int nsteps = wave_steps(nxi);
    const uint32_t plane = (uint32_t)(A.ny * A.nz);
    const uint32_t nc = (uint32_t)nx * plane;   // below 2^28 (kWaveMaxCells)
    // Plane offsets are byte offsets for the widest element a kernel addresses: 8-byte (the
    // factor's P64) or 4-byte (the sweeps), so the sweeps use them without a shift.
    constexpr uint32_t kUnit = KIND == 0 ? 8u : 4u;
    const uint32_t plane8 = plane * kUnit, nc8 = nc * kUnit;
    auto off4 = [](uint32_t o) { return KIND == 0 ? o >> 1 : o; };   // 4-byte elements
    // buffer resources.  factor: rA nbinfo, rC P64, rD P32;  sweeps: rA P32, rB in, rC out.
    // The sweeps read P32 alone: it is 0 for every non-WATER cell (cleared before each factor,
    // which writes WATER cells only), so P != 0 is the WATER test and the neighbour flags are
    // not needed (non-WATER values are 0 on the upstream side).
    __amdgpu_buffer_rsrc_t rA, rB, rC, rD;
    if constexpr (KIND == 0) {
        rA = rsrc(A.nbinfo, (int64_t)nc * 2);
        rB = rA;
        rC = rsrc(A.P64, (int64_t)nc * 8);
        rD = rsrc(A.P32, (int64_t)nc * 4);
    } else {
        // ranges one and two words longer than the arrays (never addressed): distinct descriptor
        // words keep the compiler from rebuilding one descriptor from another at every access
        rA = rsrc(A.P32, (int64_t)nc * 4 + 8);
        rB = rsrc(A.in, (int64_t)nc * 4);
        rC = rsrc(A.out, (int64_t)nc * 4 + 4);
        // opaque to the compiler from here on, so that it keeps each descriptor whole in
        // four SGPRs instead of sharing equal words and reassembling them at every access
        asm volatile("" : "+s"(rA), "+s"(rB), "+s"(rC));
        rD = rC;
    }
    const float ms = A.ms;
    const double scale = A.scale;

    // All R rows of wave w process x-index xl = t - w - tz at step t (row r
    // uses row r-1's value of the same step), so the lane enters the grid
    // along x at step w + tz, and the wave's cells of one step lie in one
    // rotated plane: FWD x + rot(z) = 1 + t - w;  BWD x + rot(z) = nx + 61 + w - t.
    // Per row: the lane's in-plane byte offsets for 8-, 4- and 2-byte
    // elements (kOob, shifted, stays out of range for every array).  Per
    // wave: the byte offsets (8-byte elements) of the rotated plane of the
    // current step and of the prefetch step, in scalar registers.
    uint32_t lo8[R], lo4[R], lo2[R];
    bool colr[R], colz[R];
    const int z = z_of<FWD>(A, K, tz);
    const int base = w + tz;
    uint32_t oc = (uint32_t)(FWD ? ((1 - w) % nx + nx) % nx : (61 + w) % nx) * plane8;
    uint32_t op = oc;
    unroll<R>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        const int ty = w * R + r;
        const int y = y_of<FWD>(A, J, ty);
        const bool rowin = J * TY + ty < A.ny - 2;
        colr[r] = rowin && z <= A.nz - 2;
        lo8[r] = colr[r] ? (uint32_t)(y * A.nz + z) * 8u : kOob;
        lo4[r] = lo8[r] >> 1;
        lo2[r] = lo8[r] >> 2;
        // lane 0 of the row reads -z packets when there is an upstream z tile
        colz[r] = zup && rowin;
    });
    // the rotated plane of the next step; readfirstlane keeps the offset in a
    // scalar register (a per-lane copy turns each buffer access into a loop)
    auto advance = [&](uint32_t& o) {
        if (FWD) {
            o += plane8;
            if (o >= nc8) o -= nc8;
        } else {
            // o - plane8 wraps above every offset when o < plane8: the smaller of the two is the
            // wrapped-around plane (scalar subtract, add and min, as in the forward direction)
            const uint32_t below = o - plane8, wrapped = o + (nc8 - plane8);
            o = below < wrapped ? below : wrapped;
        }
        o = __builtin_amdgcn_readfirstlane(o);
    };
    // prefetch ring: per-cell inputs of the next kPrefetch steps
    uint32_t wring[kPrefetch][R];   // factor: nbinfo; sweeps: P32 bits
    float iring[kPrefetch][R];
    auto fetch = [&](int d, int r) {
        if constexpr (KIND == 0) {
            wring[d][r] = bload16(rA, lo2[r], op >> 2);
            iring[d][r] = 0.f;
        } else {
            wring[d][r] = __float_as_uint(bload32f(rA, lo4[r], off4(op)));
            iring[d][r] = bload32f(rB, lo4[r], off4(op));
        }
    };
    // issue order pinned, and each prologue step issues as many memory
    // operations as the loop steps do (the stores of a step, here dropped as
    // out of range, then the loads): the wait for slot d then allows the same
    // number of younger operations on entry as at every later step
    constexpr int kStores = (KIND == 0 ? 2 : 1) * R;
    unroll<kPrefetch>([&](auto dc) {
        constexpr int d = decltype(dc)::value;
        unroll<kStores>([&](auto) { bstore32(0u, rC, kOob, 0); });
        unroll<R>([&](auto rc) { fetch(d, decltype(rc)::value); });
        advance(op);
        __builtin_amdgcn_sched_barrier(0);
    });
    // Value rings: wave w < NW-1 writes its last row into v[w] (slot t mod kRingV), the last wave
    // into the +y out-ring yo (slot t mod kRingO), which the sending wave forwards and wave NW-2's
    // reuse check reads.  Both have kSlotV words per slot, so a step's slot is a constant offset
    // from the block's base.
    constexpr int kSlotV = NP * WF_TZ;   // words per ring slot
    const uint64_t* up_base = &S.v[w > 0 ? w - 1 : 0][0][0][tz];
    const uint64_t* dn_base = &S.v[w < NW - 2 ? w + 1 : 0][0][0][tz];
    const bool dn_yo = w == NW - 2;   // the downstream wave's ring is yo
    // +z out-ring: lane 63 writes its rows into zo[slot][k][w R + r]; the other lanes write the same
    // pattern into a sink of their own, so the store needs no lane mask (slot stride 0 there)
    const bool zl = tz == WF_TZ - 1;
    uint64_t* const zo_lane = zl ? &S.zo[0][0][w * R] : &S.zsink[w][tz][0];
    const uint32_t zo_stride = zl ? NP * TY : 0;   // words per ring slot
    const uint64_t* yp_base = &S.yp[0][0][tz];
    const uint64_t* zp_base = &S.zp[0][0][w * R];
    const bool cy = w == 0 && yup;
    // Sweeps: the readiness test as 64-bit lane masks.  Each tag test is one
    // compare whose ballot is ANDed with a mask of the lanes that need it:
    // the uniform conditions select a whole mask in scalar code, the lane
    // conditions (row in the grid, x in the grid) are ballots of single
    // compares, so no boolean is turned into a lane value and back.
    // Row 0's value is needed by every lane for w > 0 (at t = 0 the ring's
    // initial tag -1 matches t - 1), and for wave 0 by the lanes of its first
    // row that are in the grid, when there is an upstream y tile.
    // Row 0's value is needed by every lane for w > 0 (at t = 0 the ring's
    // initial tag -1 matches t - 1), and for wave 0 by the lanes of its first
    // row that are in the grid, when there is an upstream y tile.  The -z
    // packets are needed by the rows with an upstream z tile while the row's
    // lane 0 is in the grid.
    const bool cy_u = __builtin_amdgcn_readfirstlane(cy ? 1 : 0) != 0;
    const uint64_t ma_all = w > 0 ? ~0ull : 0ull;
    const uint64_t ma_row = cy_u ? __builtin_amdgcn_ballot_w64(colr[0]) : 0ull;
    uint64_t colz_m[R];
    unroll<R>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        colz_m[r] = __builtin_amdgcn_ballot_w64(colz[r]);
    });
    T prev[R];
    unroll<R>([&](auto rc) { prev[decltype(rc)::value] = 0; });
    // Upstream values of one step: the previous wave's last row (ring slot),
    // the -y and -z packets and, every kBpEvery steps, the next wave's ring
    // slot (reuse check).  The reads of step t+1 are issued right after step
    // t's ring write, so their latency overlaps the rest of step t; the tag
    // test at the top of step t+1 re-reads only if a value was not there yet.
    // Sweeps: row 0's -y value comes from one ring, the previous wave's
    // (w > 0, tagged t - 1) or, for wave 0, the -y packets (tagged t; a first
    // y tile's ring holds zeros), so one read, one tag test and no select.
    uint64_t a[NP], c[NP], zq[R][NP], bp = 0;
    const uint64_t* a0_base = (KIND != 0 && w == 0) ? yp_base : up_base;
    int spy = 0, spz = 0;   // progress of the sending waves (read on reuse-check steps)
    // ring slots of the current block of kPrefetch steps: step t0 + d uses slot pz0 + d of the
    // packet rings and so0 + d of the out-rings (no wrap inside a block), so the slot offsets of
    // the unrolled steps are constants
    static_assert(RINGP % kPrefetch == 0 && kRingO % kPrefetch == 0, "blocks do not wrap the rings");
    int pz0 = 0, so0 = 0;
    // sweeps: the block's base for row 0's -y value, so that step t0 + d2 (d2 = 1 .. kPrefetch-1)
    // reads at a constant offset (d2 - 1 slots): the upstream wave's ring (slot d2 - 1) or, for
    // wave 0, the packet ring (slot pz0 + d2)
    const uint64_t* a_blk = up_base;
    auto rd_at = [&](auto dc2, int tt) {
        constexpr int d2 = decltype(dc2)::value;   // tt mod kPrefetch (= tt mod kRingV)
        constexpr int s_up = (d2 + kRingV - 1) % kRingV, s_bp = (d2 + kBpEvery) % kRingV;
        // tt = t0 + d2 inside the block (d2 > 0), or the first step of the next block
        const int ps = d2 != 0 ? pz0 + d2 : tt % RINGP;
        if constexpr (KIND != 0) {
            static_assert(NP == 1, "one word per sweep value");
            if constexpr (d2 != 0) {
                a[0] = lpeek(a_blk + (d2 - 1) * kSlotV);
            } else {   // tt = t0 + kPrefetch, or tt = t0 (re-read) / 0 (first read)
                const int aoff = w > 0 ? s_up * kSlotV : ps * NP * WF_TZ;
                a[0] = lpeek(a0_base + aoff);
            }
        } else {
#pragma unroll
            for (int k = 0; k < NP; k++) {
                a[k] = lpeek(up_base + s_up * kSlotV + k * WF_TZ);
                c[k] = lpeek(yp_base + ps * NP * WF_TZ + k * WF_TZ);
            }
        }
        if constexpr (KIND != 0 && R == 2) {
            // both rows' words in one 16-byte LDS read (each word carries its own tag)
            const u64x2 q = lpeek2(zp_base + ps * TY);
            zq[0][0] = q.x;
            zq[1][0] = q.y;
        } else {
            unroll<R>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
#pragma unroll
                for (int k = 0; k < NP; k++) zq[r][k] = lpeek(zp_base + ps * NP * TY + k * TY + r);
            });
        }
        if constexpr (d2 % kBpEvery == 0) {
            // the downstream wave's slot of step tt - (kRingV - kBpEvery)
            const int so_bp = ((d2 != 0 ? so0 + d2 : tt) - (kRingV - kBpEvery)) & (kRingO - 1);
            bp = lpeek(dn_yo ? &S.yo[so_bp][0][tz] : dn_base + s_bp * kSlotV);
            spy = lpeek32(&S.sprog[0]);
            spz = lpeek32(&S.sprog[1]);
        }
    };
    static_assert(kRingV == kPrefetch, "ring slots are indexed by the unrolled step");
    rd_at(std::integral_constant<int, 0>{}, 0);
    for (int t0 = 0; t0 < nsteps; t0 += kPrefetch) {
        pz0 = t0 % RINGP;
        so0 = t0 & (kRingO - 1);
        uint64_t* const my_blk = w == NW - 1 ? &S.yo[so0][0][tz] : &S.v[w][0][0][tz];
        a_blk = w > 0 ? up_base : yp_base + (pz0 + 1) * NP * WF_TZ;
        uint64_t* const zblk = zo_lane + (uint32_t)so0 * zo_stride;
        // compile-time step index, so each prefetch slot stays in one register
        // and the value-ring slots are constants
        unroll<kPrefetch>([&](auto dc) {
            constexpr int d = decltype(dc)::value;
            constexpr bool bp_step = d % kBpEvery == 0;
            const int t = t0 + d;
            // the tail of the previous step (prefetch, plane advance) stays ahead of the
            // tag checks, so the latency of the values read ahead is overlapped
            __builtin_amdgcn_sched_barrier(0);
            // ---- receive: tag checks of the values read ahead (re-read while not ready)
            const bool need_up = w > 0 && t > 0;
            const bool need_bp = bp_step && w < NW - 1 && t >= kRingV - kBpEvery;
            const int t_bp = t - (kRingV - kBpEvery);
            // out-ring slots of steps t .. t + kBpEvery - 1 are free once the sending waves passed
            // step t + kBpEvery - 1 - kRingO (the last wave feeds the +y ring, every wave the +z ring)
            const int t_bo = t + kBpEvery - 1 - kRingO;
            // the lane's x-index is in the grid (same for all rows of the wave),
            // and the same test for lane 0 (the -z packet consumer)
            const bool xin = (unsigned)(t - base) < (unsigned)nxi;
            const bool xin0 = (unsigned)(t - w) < (unsigned)nxi;
            bool act[R], need_z[R];
            unroll<R>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                act[r] = colr[r] && xin;
                need_z[r] = colz[r] && xin0;
            });
            const bool need_y = cy && act[0];
            // branch-free: per-lane masks, one ballot (the -z packet slots are
            // read by every lane from one address, so their test is uniform).
            // The out-ring test reads the sending waves' progress as last read,
            // so each re-read of the spin loop below can pass it.
            auto ready = [&]() {
                const bool bo_ok = !bp_step || (spz >= t_bo && (w != NW - 1 || spy >= t_bo));
                if constexpr (KIND != 0) {
                    const int ta = w > 0 ? t - 1 : t;
                    const uint64_t m_a = ma_all | (ma_row & __builtin_amdgcn_ballot_w64(xin));
                    const bool xin0_u = (unsigned)(t - w) < (unsigned)nxi;
                    uint64_t bad = __builtin_amdgcn_ballot_w64(tag_of(a[0]) != ta) & m_a;
                    if constexpr (bp_step) {
                        if (w < NW - 1 && t >= kRingV - kBpEvery) bad |= __builtin_amdgcn_ballot_w64(tag_of(bp) < t_bp);
                    }
                    unroll<R>([&](auto rc) {
                        constexpr int r = decltype(rc)::value;
                        bad |= __builtin_amdgcn_ballot_w64(tag_of(zq[r][0]) != t) & (xin0_u ? colz_m[r] : 0ull);
                    });
                    return bo_ok && bad == 0;
                }
                unsigned lane = (unsigned)(tag_of(bp) >= t_bp) | (unsigned)!need_bp;
#pragma unroll
                for (int k = 0; k < NP; k++) {
                    lane &= ((unsigned)(tag_of(a[k]) == t - 1) | (unsigned)!need_up) &
                            ((unsigned)(tag_of(c[k]) == t) | (unsigned)!need_y);
                    unroll<R>([&](auto rc) {
                        constexpr int r = decltype(rc)::value;
                        lane &= (unsigned)(tag_of(zq[r][k]) == t) | (unsigned)!need_z[r];
                    });
                }
                return bo_ok && __builtin_amdgcn_ballot_w64(lane == 0u) == 0;
            };
            if (__builtin_expect(!ready(), 0)) {   // the spin path out of line: the ready step falls through
                int spins = 0;
                do {
                    if (++spins > kSpinLimit) {
                        if (tz == 0) atomicOr(A.err, 1);
                        break;
                    }
                    if ((spins & 1023) == 0 && gave_up(A.err)) break;
                    if (spins > kBusyPolls) __builtin_amdgcn_s_sleep(1);
                    rd_at(dc, t);
                } while (!ready());
            }
            // ---- compute the R rows (all upstream values are from step t-1)
            T vnew[R];
            unroll<R>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                T vy = 0, vz, vin;
                if constexpr (KIND != 0) {
                    // -z: lane tz - 1's previous value; lane 0 keeps the -z packet (DPP "old" operand;
                    // zeros on a first z tile).  Lane 0 is active exactly when its packet is needed.
                    vz = __builtin_bit_cast(float, __builtin_amdgcn_update_dpp((int)(uint32_t)zq[r][0],
                                                                               __builtin_bit_cast(int, prev[r]),
                                                                               0x138, 0xf, 0xf, false));
                    if constexpr (r > 0) vy = vnew[r - 1];   // the row above, same step (same x)
                    else from_words(vin, a), vy = vin;
                } else {
                    vz = shr1(prev[r]);
                    if constexpr (r > 0) {
                        vy = vnew[r - 1];
                    } else {
                        if (w > 0) from_words(vin, a), vy = vin;
                        else if (need_y) from_words(vin, c), vy = vin;
                    }
                    if (tz == 0 && need_z[r]) from_words(vin, zq[r]), vz = vin;
                }
                if constexpr (KIND == 0) {
                    const uint32_t info = act[r] ? (uint32_t)wring[d][r] : 0u;
                    double v = 0;
                    if (info & NB_FLUID) {
                        const double tau = 0.97, sigma = 0.25;
                        const double upv[3] = {prev[r], vy, vz};
                        double eneg = 0, eneg_tau = 0;
#pragma unroll
                        for (int ax = 0; ax < 3; ax++) {
                            if (!(info & (NB_MX << ax))) continue;
                            const double pj = upv[ax];
                            const double ap = -scale * pj;
                            const double o1 = (info & (1u << (NB_UP_SHIFT + 2 * ax))) ? -scale : 0.0;
                            const double o2 = (info & (1u << (NB_UP_SHIFT + 2 * ax + 1))) ? -scale : 0.0;
                            eneg += ap * ap;
                            eneg_tau += ap * (o1 + o2) * pj;
                        }
                        const int cnt = (info >> NB_NONSOLID_SHIFT) & 7;
                        double diag = 0;
                        for (int k = 0; k < cnt; k++) diag += scale;
                        double e = diag - eneg - eneg_tau * tau;
                        if (e < sigma * diag) e = diag < 1e-6 ? 1.0 : diag;
                        v = 1.0 / sqrt(e);
                    }
                    const bool fl = (info & NB_FLUID) != 0;
                    const float p32 = (float)v;
                    vnew[r] = v;
                    // P64/P32 for WATER cells (both were cleared before the factor)
                    bstore64(__builtin_bit_cast(uint64_t, v), rC, fl ? lo8[r] : kOob, oc);
                    bstore32(__float_as_uint(p32), rD, fl ? lo4[r] : kOob, off4(oc));
                } else {
                    const float p = __uint_as_float(wring[d][r]);
                    // stores of WATER cells only: storing every interior cell doubles the stores, and
                    // the in-order memory counter then delays the prefetch waits (measured: slower)
                    const bool fl = act[r] && wring[d][r] != 0u;
                    // 0 + ms * prev in one rounding: fma(ms, prev, +0) equals the product, and +0 for -0
                    float acc = __builtin_fmaf(ms, prev[r], 0.f);
                    acc += ms * vy;
                    acc += ms * vz;
                    float outv, v;
                    if (FWD) {
                        outv = (iring[d][r] - acc) * p;
                        v = outv * p;
                    } else {
                        outv = (iring[d][r] - acc * p) * p;
                        v = outv;
                    }
                    vnew[r] = fl ? v : 0.f;
                    bstore32(__float_as_uint(outv), rC, fl ? lo4[r] : kOob, off4(oc));
                }
            });
            // ---- publish: the last row for wave w+1 (the last wave's into the +y out-ring), lane 63 of
            // every row into the +z out-ring (both sent by the sending waves), then read ahead,
            // advance and prefetch step t + kPrefetch
#pragma unroll
            for (int k = 0; k < NP; k++) lput(my_blk + d * kSlotV + k * WF_TZ, pack(word(vnew[R - 1], k), (uint32_t)t));
            // (the factor's FP64 values would make the sink writes of all lanes cost more LDS
            // bandwidth than the lane mask costs instructions: it keeps the mask)
            if (KIND != 0 || zl) {
                uint64_t* zdst = zblk + (uint32_t)d * zo_stride;
                if constexpr (KIND != 0 && R == 2) {
                    // both rows in one 16-byte LDS write (the sending wave reads them word by word)
                    u64x2 q;
                    q.x = pack(word(vnew[0], 0), (uint32_t)t);
                    q.y = pack(word(vnew[1], 0), (uint32_t)t);
                    lput2(zdst, q);
                } else {
                    unroll<R>([&](auto rc) {
                        constexpr int r = decltype(rc)::value;
#pragma unroll
                        for (int k = 0; k < NP; k++) lput(zdst + k * TY + r, pack(word(vnew[r], k), (uint32_t)t));
                    });
                }
            }
            rd_at(std::integral_constant<int, (d + 1) % kPrefetch>{}, t + 1);
            // the packet reads of steps <= t have completed (their values were used above)
            if (d == kPrefetch - 1 && tz == 0) lput32(&S.prog[w], t);
            unroll<R>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                prev[r] = vnew[r];
                fetch(d, r);
            });
            advance(oc);
            advance(op);
        });
    }
    if (tz == 0) lput32(&S.prog[w], nsteps);
}

template <int KIND>
__global__ void __launch_bounds__(WF_THREADS) k_wave(WaveArgs A) {
    using KD = Kind<KIND>;
    constexpr int NP = KD::NP, RINGP = KD::RINGP;
    __shared__ WaveLds<KIND> S;
    const int tile = take_tile(A);
    if (tile < 0) return;
    const int J = tile / A.TK, K = tile % A.TK;
    // A tile none of whose (y, z) columns holds WATER has nothing to compute
    // (its cells' values are 0, and only WATER cells are stored): it leaves at
    // once.  Its downstream tiles take its edge values as the zeros they are,
    // exactly as on a first tile, instead of waiting for packets.
    int mine = 0, upy = 0, upz = 0;
    for (int c = threadIdx.x; c < WF_TY * WF_TZ; c += WF_THREADS) {
        const int ty = c / WF_TZ, tz = c % WF_TZ;
        mine |= col_water<KD::FWD>(A, J, K, ty, tz);
        if (J > 0) upy |= col_water<KD::FWD>(A, J - 1, K, ty, tz);
        if (K > 0) upz |= col_water<KD::FWD>(A, J, K - 1, ty, tz);
    }
    if (!__syncthreads_or(mine)) return;
    const bool yup = __syncthreads_or(upy) != 0, zup = __syncthreads_or(upz) != 0;
    for (int i = threadIdx.x; i < (WF_NW - 1) * kRingV * NP * WF_TZ; i += WF_THREADS) (&S.v[0][0][0][0])[i] = ~0ull;
    // sweeps on a first y / z tile, or below / beside a tile without WATER, read their -y / -z
    // values from a ring of zeros (never refilled)
    const uint64_t yp_init = KIND != 0 && !yup ? 0ull : ~0ull, zp_init = KIND != 0 && !zup ? 0ull : ~0ull;
    for (int i = threadIdx.x; i < RINGP * NP * WF_TZ; i += WF_THREADS) (&S.yp[0][0][0])[i] = yp_init;
    for (int i = threadIdx.x; i < RINGP * NP * WF_TY; i += WF_THREADS) (&S.zp[0][0][0])[i] = zp_init;
    for (int i = threadIdx.x; i < kRingO * NP * WF_TZ; i += WF_THREADS) (&S.yo[0][0][0])[i] = ~0ull;
    for (int i = threadIdx.x; i < kRingO * NP * WF_TY; i += WF_THREADS) (&S.zo[0][0][0])[i] = ~0ull;
    if (threadIdx.x < WF_NW) S.prog[threadIdx.x] = -1;
    // steps below a sending wave's lag are never sent; without a downstream tile nothing is
    if (threadIdx.x == 0) {
        S.sprog[0] = J == A.TJ - 1 ? INT_MAX : WF_NW - 2;
        S.sprog[1] = K == A.TK - 1 ? INT_MAX : WF_TZ - 2;
    }
    __syncthreads();
    const int wave = threadIdx.x / 64;
    if (wave < WF_NW) compute_wave<KIND>(A, S, J, K, yup, zup);
    else if (wave == WF_NW) comm_wave<KIND, true>(A, S, J, K, yup, zup);
    else if (wave == WF_NW + 1) comm_wave<KIND, false>(A, S, J, K, yup, zup);
    else if (wave == WF_NW + 2) send_wave<KIND, true>(A, S, J, K);
    else send_wave<KIND, false>(A, S, J, K);
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
    A.send_delay = ctx->wave_send_delay;
    if (kind == 0) hipLaunchKernelGGL(k_wave<0>, dim3(ntiles), dim3(WF_THREADS), 0, ctx->stream, A);
    else if (kind == 1) hipLaunchKernelGGL(k_wave<1>, dim3(ntiles), dim3(WF_THREADS), 0, ctx->stream, A);
    else hipLaunchKernelGGL(k_wave<2>, dim3(ntiles), dim3(WF_THREADS), 0, ctx->stream, A);
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

}  // namespace fsim
