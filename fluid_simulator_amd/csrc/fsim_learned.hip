// fsim_learned.hip — learned preconditioner: a 4-level 3D conv encoder-decoder
// (no counterpart in the reference) used as M^-1 in flexible PCG.
//
// Network (fully convolutional, so weights trained on small grids apply to
// any grid whose padded size is a multiple of 8):
//   in  = (b / s, unknown mask, diag / 6, cx, cy, cz, 0, 0)   level 0, 8 ch
//         (the operator L of the grid it runs on: diagonal diag, couplings
//         -cx, -cy, -cz to the +x, +y, +z neighbour; out ~ L^-1 b / s)
//   e0  = relu(conv3(in, 8 -> 16))                           level 0
//   e1  = relu(conv3(pool(e0), 16 -> 32))                    level 1
//   e2  = relu(conv3(pool(e1), 32 -> 64))                    level 2
//   e3  = relu(conv3(pool(e2), 64 -> 128))                   level 3
//   d2  = relu(conv3(cat(up(e3), e2), 192 -> 64))            level 2
//   d1  = relu(conv3(cat(up(d2), e1), 96 -> 32))             level 1
//   d0  = relu(conv3(cat(up(d1), e0), 48 -> 16))             level 0
//   out = conv1(d0, 16 -> 1)
// with s = max|r|, pool = 2x2x2 average, up = nearest, conv3 = 3x3x3, stride 1,
// zero padding.  Activations and weights are BF16, every sum is accumulated in
// FP32 (MFMA), biases and the output are FP32.
//
// Layout: channels last, voxel v = (x * H + y) * W + z at each level.  Conv
// weights are stored per output channel, Wt[co][t * Cin + ci] with tap
// t = ((dx+1) * 3 + (dy+1)) * 3 + (dz+1), so a K run of 8 is 8 channels of one tap.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fsim_internal.hpp"

namespace fsim {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ---- architecture --------------------------------------------------------------------
struct LayerDef {
    int cin, cout, level;
};
constexpr int kNetLayers = 7;   // 3x3x3 convolutions; the 1x1 output conv follows
constexpr LayerDef kLayers[kNetLayers] = {
    {8, 16, 0}, {16, 32, 1}, {32, 64, 2}, {64, 128, 3}, {192, 64, 2}, {96, 32, 1}, {48, 16, 0}};
constexpr int kInCh = 8, kOutIn = 16;

__host__ __device__ inline int64_t conv_weights(int l) { return (int64_t)kLayers[l].cout * 27 * kLayers[l].cin; }
// offsets of each layer's weights and biases in the flat parameter vector
static void param_offsets(int64_t* woff, int64_t* boff, int64_t* out_w, int64_t* out_b, int64_t* total) {
    int64_t o = 0;
    for (int l = 0; l < kNetLayers; l++) {
        woff[l] = o;
        o += conv_weights(l);
        boff[l] = o;
        o += kLayers[l].cout;
    }
    *out_w = o;
    o += kOutIn;
    *out_b = o;
    o += 1;
    *total = o;
}
int64_t learned_num_params() {
    int64_t w[kNetLayers], b[kNetLayers], ow, ob, t;
    param_offsets(w, b, &ow, &ob, &t);
    return t;
}

// ---- implicit-GEMM 3x3x3 convolution on MFMA ------------------------------------------
// Y[v][co] = bias[co] + sum_{t, ci} X[v + off(t)][ci] * Wt[co][t * Cin + ci]
// One wave: MT tiles of 16 consecutive voxels x NT tiles of 16 output
// channels (output channel group blockIdx.y), K = 27 * CIN in chunks of 32
// with v_mfma_f32_16x16x32_bf16.  Operand lanes: A (16 x 32): lane l holds
// row l % 16, k = 8 (l / 16) .. + 7; B (32 x 16): lane l holds column l % 16,
// the same k run; D (16 x 16): lane l holds rows 4 (l / 16) .. + 3 of column
// l % 16.  A K run of 8 is 8 channels of one tap (channel counts are
// multiples of 8), so each operand is one 16-byte load; a chunk's tap
// geometry is computed once for all MT tiles, each B fragment feeds MT MFMAs
// and each A fragment NT.  Out-of-grid taps read zeros.
// With CU > 0 the input is cat(up(coarse), skip) without materialising it:
// channels [0, CU) are read from the coarse parent voxel (x/2, y/2, z/2) of
// xc, channels [CU, CIN) from x (skip, CIN - CU channels) -- the same BF16
// values a concatenated copy would hold.
struct ConvArgs {
    const __bf16* x;      // [V][cin] (or the skip part [V][cin - cu])
    const __bf16* w;      // [cout][27 * cin]
    const float* bias;    // [cout] or null
    void* y;              // [V][cout], bf16 or float
    int D, H, W;          // grid of this level
    int cin, cout;
    int relu;
    const __bf16* xc;     // coarse parent [V/8][cu] when cu > 0
    int cu;
    int lds;              // 1: the LDS-staged kernel where the layer and the level fit it
    float* part;          // split-K partial sums [slice][V][cout] (null: no split)
    int64_t part_cap;     // floats in part
};

template <int CIN, int CU, int NT, int MT, bool OUT_F32>
__global__ void __launch_bounds__(256) k_conv3(ConvArgs a) {
    constexpr int KTOT = 27 * CIN;
    constexpr int NCH = (KTOT + 31) / 32;
    constexpr int CS = CIN - CU;   // channels read from x
    const int lane = threadIdx.x & 63;
    const int64_t V = (int64_t)a.D * a.H * a.W;
    const int64_t m0 = ((int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6)) * (16 * MT);
    if (m0 >= V) return;   // wave-uniform
    const int co0 = 16 * NT * (int)blockIdx.y;
    const int kq = lane >> 4;   // which 8-run of the 32-wide K chunk
    const __bf16* xl[MT];       // this lane's A row (voxel) of each tile in x, at channel 0
    int xs[MT], ys[MT], zs[MT];
    bool rin[MT];
#pragma unroll
    for (int m = 0; m < MT; m++) {
        const int64_t vr = m0 + 16 * m + (lane & 15);
        rin[m] = vr < V;
        const int64_t v = rin[m] ? vr : 0;
        zs[m] = (int)(v % a.W);
        ys[m] = (int)((v / a.W) % a.H);
        xs[m] = (int)(v / ((int64_t)a.W * a.H));
        xl[m] = a.x + v * CS;
    }
    const int Hc = a.H >> 1, Wc = a.W >> 1;
    const __bf16* wl = a.w + (int64_t)(co0 + (lane & 15)) * KTOT + 8 * kq;
    const bf16x8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int j = 0; j < NT; j++) acc[m][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    // split K: block z sums the K chunks of its slice (one slice: all of them)
    const int cps = (NCH + (int)gridDim.z - 1) / (int)gridDim.z;
    const int c_lo = (int)blockIdx.z * cps, c_hi = min(NCH, c_lo + cps);
#pragma unroll 2
    for (int c = c_lo; c < c_hi; c++) {
        const int k0 = 32 * c + 8 * kq;
        const bool kin = (KTOT % 32 == 0) || k0 < KTOT;
        const int t = k0 / CIN, ci = k0 - t * CIN;
        const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
        const bool up = CU > 0 && ci < CU;
        const int64_t toff = (((int64_t)dx * a.H + dy) * a.W + dz) * CS + (ci - CU);
        bf16x8 av[MT], bv[NT];
#pragma unroll
        for (int m = 0; m < MT; m++) {
            const int xx = xs[m] + dx, yy = ys[m] + dy, zz = zs[m] + dz;
            const bool ok = kin && rin[m] && (unsigned)xx < (unsigned)a.D && (unsigned)yy < (unsigned)a.H &&
                            (unsigned)zz < (unsigned)a.W;
            const __bf16* src = xl[m] + toff;
            if (up) src = a.xc + (((int64_t)(xx >> 1) * Hc + (yy >> 1)) * Wc + (zz >> 1)) * (CU > 0 ? CU : 1) + ci;
            av[m] = ok ? *reinterpret_cast<const bf16x8*>(src) : zero;
        }
#pragma unroll
        for (int j = 0; j < NT; j++)
            bv[j] = kin ? *reinterpret_cast<const bf16x8*>(wl + (int64_t)16 * j * KTOT + 32 * c) : zero;
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int j = 0; j < NT; j++) acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], bv[j], acc[m][j], 0, 0, 0);
    }
    if (gridDim.z > 1) {   // raw partial sums; k_conv3_reduce adds the slices, bias and ReLU
        float* P = a.part + (int64_t)blockIdx.z * V * a.cout;
#pragma unroll
        for (int j = 0; j < NT; j++)
#pragma unroll
            for (int m = 0; m < MT; m++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int64_t v = m0 + 16 * m + 4 * kq + i;
                    if (v < V) P[v * a.cout + co0 + 16 * j + (lane & 15)] = acc[m][j][i];
                }
        return;
    }
#pragma unroll
    for (int j = 0; j < NT; j++) {
        const int co = co0 + 16 * j + (lane & 15);
        const float b = a.bias ? a.bias[co] : 0.f;
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int64_t v = m0 + 16 * m + 4 * kq + i;
                if (v >= V) continue;
                float o = acc[m][j][i] + b;
                if (a.relu) o = fmaxf(o, 0.f);
                if (OUT_F32) static_cast<float*>(a.y)[v * a.cout + co] = o;
                else static_cast<__bf16*>(a.y)[v * a.cout + co] = (__bf16)o;
            }
    }
}

// Sum of the split-K slices in slice order, then bias and ReLU (deterministic).
template <bool OUT_F32>
__global__ void __launch_bounds__(256) k_conv3_reduce(ConvArgs a, int slices) {
    const int64_t n = (int64_t)a.D * a.H * a.W * a.cout;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int k = 0; k < slices; k++) s += a.part[k * n + i];
    float o = s + (a.bias ? a.bias[i % a.cout] : 0.f);
    if (a.relu) o = fmaxf(o, 0.f);
    if (OUT_F32) static_cast<float*>(a.y)[i] = o;
    else static_cast<__bf16*>(a.y)[i] = (__bf16)o;
}

// Small levels (the coarsest multigrid level's network runs at 16^3 and below)
// give a layer a handful of waves, each walking up to 162 K chunks with its
// loads in series: such layers split K over blockIdx.z, about kSplitWaves
// waves in all with at least 3 chunks per slice, as far as the partial-sum
// buffer holds them.  Layers with fewer than kSplitMinChunks chunks do not
// split: the reduction's launch costs more than their short K loop (measured
// at 16^3: 8 -> 16, 7 chunks, 10.0 us unsplit, 7.6 + 7.3 us split).
constexpr int64_t kSplitWaves = 1024;
constexpr int kSplitMinChunks = 24;
constexpr int64_t kConvPartFloats = int64_t(4) << 20;   // 16 MB of split-K partial sums per context
template <int CIN, int CU, int COUT, int NT, int MT, bool OUT_F32>
static void conv3_launch(hipStream_t st, const ConvArgs& a) {
    static_assert(COUT % (16 * NT) == 0 && CIN % 8 == 0 && CU % 8 == 0, "conv3 tile shape");
    constexpr int NCH = (27 * CIN + 31) / 32;
    const int64_t V = (int64_t)a.D * a.H * a.W;
    const int64_t waves = (V + 16 * MT - 1) / (16 * MT);
    const int64_t all = waves * (COUT / (16 * NT));
    int64_t slices = 1;
    if (a.part && all < kSplitWaves && NCH >= kSplitMinChunks) {
        slices = std::min<int64_t>({(kSplitWaves + all - 1) / all, NCH / 3, a.part_cap / (V * COUT)});
        slices = std::max<int64_t>(slices, 1);
        const int64_t cps = (NCH + slices - 1) / slices;   // every slice gets at least one chunk
        slices = (NCH + cps - 1) / cps;
    }
    hipLaunchKernelGGL((k_conv3<CIN, CU, NT, MT, OUT_F32>),
                       dim3((unsigned)((waves + 3) / 4), COUT / (16 * NT), (unsigned)slices), dim3(256), 0, st, a);
    if (slices > 1)
        hipLaunchKernelGGL((k_conv3_reduce<OUT_F32>), dim3((unsigned)((V * COUT + 255) / 256)), dim3(256), 0, st, a,
                           (int)slices);
}

// LDS-staged variant for the large levels (rows a multiple of 32 voxels).  A block owns TY rows x 32
// voxels of XS consecutive x planes and slides along x with a ring of four halo planes in LDS, each
// (TY + 2) x 34 voxels x all CIN channels (zeros outside the grid; a decoder layer's channels below
// CU from the coarse parent): per output plane it loads the one new input plane x + 1, waits at one
// barrier, and its waves read the MFMA A operands from the planes x - 1 .. x + 1, one 16-byte LDS read
// per lane and K chunk.  The slot written at plane x (ring index x + 1) was last read at plane x - 2,
// which every wave finished before the previous barrier.  The voxel stride is padded by 8 channels when
// CIN / 2 dwords is a multiple of 16, so that the eight lanes of a 16-byte read hit distinct banks.
// K order and MFMA sequence are k_conv3's, so the output is bit-identical.
constexpr int kLdsTZ = 32, kLdsSlots = 4;
template <int CIN>
constexpr int conv3_lds_stride() { return (CIN / 2) % 16 == 0 ? CIN + 8 : CIN; }
template <int CIN, int TY>
constexpr size_t conv3_lds_bytes() {
    return (size_t)kLdsSlots * (TY + 2) * (kLdsTZ + 2) * conv3_lds_stride<CIN>() * sizeof(__bf16);
}

template <int CIN, int CU, int NT, int TY, int XS>
__global__ void __launch_bounds__(256) k_conv3_lds(ConvArgs a) {
    constexpr int TZ = kLdsTZ, HY = TY + 2, HZ = TZ + 2, PL = HY * HZ, VS = conv3_lds_stride<CIN>();
    constexpr int MT = TY * TZ / (16 * 4);   // voxel tiles per wave
    constexpr int KTOT = 27 * CIN, NCH = (KTOT + 31) / 32, CS = CIN - CU, RUNS = CIN / 8;
    static_assert(CIN % 8 == 0 && MT >= 1 && TY * TZ == 64 * MT, "conv3_lds tile shape");
    extern __shared__ uint4 lds_raw[];
    __bf16* tile = reinterpret_cast<__bf16*>(lds_raw);   // [kLdsSlots][HY][HZ][VS]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nzb = a.W / TZ, nyb = a.H / TY;
    const int zb = (int)(blockIdx.x % nzb), yb = (int)((blockIdx.x / nzb) % nyb);
    const int x0 = (int)(blockIdx.x / (nzb * nyb)) * XS;
    const int y0 = yb * TY, z0 = zb * TZ;
    const int Hc = a.H >> 1, Wc = a.W >> 1;
    const bf16x8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
    auto load_plane = [&](int X) {
        __bf16* slot = tile + (size_t)(X & (kLdsSlots - 1)) * PL * VS;
        for (int e = threadIdx.x; e < PL * RUNS; e += blockDim.x) {
            const int run = e % RUNS, hv = e / RUNS;
            const int Y = y0 + hv / HZ - 1, Z = z0 + hv % HZ - 1;
            const int ci = 8 * run;
            bf16x8 val = zero;
            if ((unsigned)X < (unsigned)a.D && (unsigned)Y < (unsigned)a.H && (unsigned)Z < (unsigned)a.W) {
                if (CU > 0 && ci < CU)
                    val = *reinterpret_cast<const bf16x8*>(
                        a.xc + (((int64_t)(X >> 1) * Hc + (Y >> 1)) * Wc + (Z >> 1)) * (CU > 0 ? CU : 1) + ci);
                else
                    val = *reinterpret_cast<const bf16x8*>(a.x + (((int64_t)X * a.H + Y) * a.W + Z) * CS + (ci - CU));
            }
            *reinterpret_cast<bf16x8*>(slot + (size_t)hv * VS + ci) = val;
        }
    };
    const int kq = lane >> 4;
    int hbase[MT];   // in-plane halo voxel (ty + 1, tz + 1) of this lane's A row in each tile
#pragma unroll
    for (int m = 0; m < MT; m++) {
        const int f = 16 * (wave * MT + m) + (lane & 15);
        hbase[m] = (f / TZ + 1) * HZ + f % TZ + 1;
    }
    // the block's weights stay in registers for all XS planes: B fragment of every K chunk
    const __bf16* wl = a.w + (int64_t)(lane & 15) * KTOT + 8 * kq;
    bf16x8 wreg[NCH][NT];
#pragma unroll
    for (int c = 0; c < NCH; c++)
#pragma unroll
        for (int j = 0; j < NT; j++)
            wreg[c][j] = ((KTOT % 32 == 0) || 32 * c + 8 * kq < KTOT)
                             ? *reinterpret_cast<const bf16x8*>(wl + (int64_t)16 * j * KTOT + 32 * c) : zero;
    load_plane(x0 - 1);
    load_plane(x0);
    for (int X = x0; X < x0 + XS; X++) {
        load_plane(X + 1);
        __syncthreads();
        f32x4 acc[MT][NT];
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int j = 0; j < NT; j++) acc[m][j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            const int k0 = 32 * c + 8 * kq;
            const bool kin = (KTOT % 32 == 0) || k0 < KTOT;
            const int t = k0 / CIN, ci = k0 - t * CIN;
            const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
            // a K run past the last tap reads nothing
            const int hoff = kin ? ((X + dx) & (kLdsSlots - 1)) * PL + dy * HZ + dz : 0;
            bf16x8 av[MT];
#pragma unroll
            for (int m = 0; m < MT; m++)
                av[m] = kin ? *reinterpret_cast<const bf16x8*>(tile + (size_t)(hbase[m] + hoff) * VS + ci) : zero;
#pragma unroll
            for (int m = 0; m < MT; m++)
#pragma unroll
                for (int j = 0; j < NT; j++)
                    acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], wreg[c][j], acc[m][j], 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < NT; j++) {
            const int co = 16 * j + (lane & 15);
            const float b = a.bias ? a.bias[co] : 0.f;
#pragma unroll
            for (int m = 0; m < MT; m++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int f = 16 * (wave * MT + m) + 4 * kq + i;
                    const int64_t v = ((int64_t)X * a.H + y0 + f / TZ) * a.W + z0 + f % TZ;
                    float o = acc[m][j][i] + b;
                    if (a.relu) o = fmaxf(o, 0.f);
                    static_cast<__bf16*>(a.y)[v * a.cout + co] = (__bf16)o;
                }
        }
    }
}

template <int CIN, int CU, int COUT, int TY, int XS>
static bool conv3_lds_launch(hipStream_t st, const ConvArgs& a) {
    const size_t lds = conv3_lds_bytes<CIN, TY>();
    // once per process: the ring must fit the device's per-block LDS (raised above the 64 KiB default
    // for the kernel when needed); otherwise the layer stays on k_conv3
    static const bool fits = [&] {
        int dev = 0, max_lds = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
            (size_t)max_lds < lds) {
            (void)hipGetLastError();
            return false;
        }
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_conv3_lds<CIN, CU, COUT / 16, TY, XS>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        return true;
    }();
    if (!fits || !a.lds || a.W % kLdsTZ != 0 || a.H % TY != 0 || a.D % XS != 0) return false;
    const unsigned blocks = (unsigned)((int64_t)(a.D / XS) * (a.H / TY) * (a.W / kLdsTZ));
    hipLaunchKernelGGL((k_conv3_lds<CIN, CU, COUT / 16, TY, XS>), dim3(blocks), dim3(256), lds, st, a);
    return true;
}

// The convolutions of the network, by (Cin, Cu, Cout) with their tile shapes
// (NT output tiles x MT voxel tiles per wave): the 7 layers' forward
// convolutions (BF16 out; the decoder reads cat(up, skip) directly), and the
// data gradients of layers 1-6 (transposed: Cin = the layer's Cout; FP32
// out).  Returns false for any other shape.
template <bool OUT_F32>
static bool launch_conv3(hipStream_t st, const ConvArgs& a) {
#define FSIM_CONV3(CI_, CU_, CO_, NT_, MT_)                     \
    if (a.cin == CI_ && a.cu == CU_ && a.cout == CO_) {         \
        conv3_launch<CI_, CU_, CO_, NT_, MT_, OUT_F32>(st, a);  \
        return true;                                            \
    }
    if constexpr (!OUT_F32) {
        // the full-resolution layers through LDS when their rows are a multiple of 32 voxels
        if (a.cin == 8 && a.cu == 0 && a.cout == 16 && conv3_lds_launch<8, 0, 16, 4, 16>(st, a)) return true;
        if (a.cin == 48 && a.cu == 32 && a.cout == 16 && conv3_lds_launch<48, 32, 16, 4, 32>(st, a)) return true;
        FSIM_CONV3(8, 0, 16, 1, 4) FSIM_CONV3(16, 0, 32, 2, 4) FSIM_CONV3(32, 0, 64, 2, 2)
        FSIM_CONV3(64, 0, 128, 2, 1) FSIM_CONV3(192, 128, 64, 4, 2) FSIM_CONV3(96, 64, 32, 2, 4)
        FSIM_CONV3(48, 32, 16, 1, 4)
    } else {
        FSIM_CONV3(32, 0, 16, 1, 4) FSIM_CONV3(64, 0, 32, 2, 4) FSIM_CONV3(128, 0, 64, 4, 2)
        FSIM_CONV3(64, 0, 192, 4, 2) FSIM_CONV3(32, 0, 96, 2, 4) FSIM_CONV3(16, 0, 48, 1, 4)
    }
#undef FSIM_CONV3
    return false;
}

// ---- elementwise pieces -------------------------------------------------------------
struct Dims {
    int D, H, W;
    __host__ __device__ int64_t V() const { return (int64_t)D * H * W; }
};

// 2x2x2 average pool (FP32 sum of the 8 BF16 values, / 8, rounded to BF16)
__global__ void k_pool2(const __bf16* __restrict__ in, __bf16* __restrict__ out, Dims fine, int C) {
    const Dims cs{fine.D / 2, fine.H / 2, fine.W / 2};
    const int64_t n = cs.V() * C;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % C);
    const int64_t v = i / C;
    const int z = (int)(v % cs.W), y = (int)((v / cs.W) % cs.H), x = (int)(v / ((int64_t)cs.W * cs.H));
    float s = 0.f;
    for (int d = 0; d < 8; d++) {
        const int64_t fv = ((int64_t)(2 * x + (d >> 2)) * fine.H + 2 * y + ((d >> 1) & 1)) * fine.W + 2 * z + (d & 1);
        s += (float)in[fv * C + c];
    }
    out[i] = (__bf16)(s * 0.125f);
}

// cat(up(coarse), skip) at the fine level: channels [0, Cu) from the coarse parent, [Cu, Cu+Cs) from skip
__global__ void k_upcat(const __bf16* __restrict__ coarse, const __bf16* __restrict__ skip, __bf16* __restrict__ out,
                        Dims fine, int Cu, int Cs) {
    const int C = Cu + Cs;
    const int64_t n = fine.V() * C;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % C);
    const int64_t v = i / C;
    if (c >= Cu) {
        out[i] = skip[v * Cs + (c - Cu)];
        return;
    }
    const int z = (int)(v % fine.W), y = (int)((v / fine.W) % fine.H), x = (int)(v / ((int64_t)fine.W * fine.H));
    const int64_t pv = ((int64_t)(x / 2) * (fine.H / 2) + y / 2) * (fine.W / 2) + z / 2;
    out[i] = coarse[pv * Cu + c];
}

// Network input: (b / s, unknown, diag / 6, cx, cy, cz, 0, 0) per voxel, from
// the solver's r and nbinfo (rotated layout: the 7-point operator of the
// WATER cells, couplings 1 towards WATER neighbours) ...
__global__ void k_net_input(const float* __restrict__ r, const uint16_t* __restrict__ nbinfo,
                            const float* __restrict__ scale, __bf16* __restrict__ x0, Dims nd, int nx, int ny, int nz) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nd.V()) return;
    const int z = (int)(v % nd.W), y = (int)((v / nd.W) % nd.H), x = (int)(v / ((int64_t)nd.W * nd.H));
    float val = 0.f, fl = 0.f, dg = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
    if (x < nx && y < ny && z < nz) {
        const int64_t i = ridx(x, y, z, nx, ny, nz);
        const uint16_t f = nbinfo[i];
        if (f & NB_FLUID) {
            fl = 1.f;
            val = r[i];
            dg = (float)((f >> NB_NONSOLID_SHIFT) & 7);
            cx = (f & NB_PX) ? 1.f : 0.f;
            cy = (f & NB_PY) ? 1.f : 0.f;
            cz = (f & NB_PZ) ? 1.f : 0.f;
        }
    }
    const float s = *scale;
    const float inv = s > 0.f ? 1.f / s : 0.f;
    bf16x8 o = {(__bf16)(val * inv), (__bf16)fl, (__bf16)(dg * (1.f / 6.f)), (__bf16)cx, (__bf16)cy, (__bf16)cz, 0, 0};
    *reinterpret_cast<bf16x8*>(x0 + v * kInCh) = o;
}
// ... or from flat arrays of b and an operator (diag, cx, cy, cz) times norm
__global__ void k_net_input_op(const float* __restrict__ b, const float* __restrict__ diag, const float* __restrict__ ox,
                               const float* __restrict__ oy, const float* __restrict__ oz, float norm,
                               const float* __restrict__ scale, __bf16* __restrict__ x0, Dims nd, int nx, int ny, int nz) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nd.V()) return;
    const int z = (int)(v % nd.W), y = (int)((v / nd.W) % nd.H), x = (int)(v / ((int64_t)nd.W * nd.H));
    float val = 0.f, fl = 0.f, dg = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
    if (x < nx && y < ny && z < nz) {
        const int64_t i = ((int64_t)x * ny + y) * nz + z;
        if (diag[i] > 0.f) {
            fl = 1.f;
            val = b[i];
            dg = diag[i] * norm;
            cx = ox[i] * norm;
            cy = oy[i] * norm;
            cz = oz[i] * norm;
        }
    }
    const float s = *scale;
    const float inv = s > 0.f ? 1.f / s : 0.f;
    bf16x8 o = {(__bf16)(val * inv), (__bf16)fl, (__bf16)(dg * (1.f / 6.f)), (__bf16)cx, (__bf16)cy, (__bf16)cz, 0, 0};
    *reinterpret_cast<bf16x8*>(x0 + v * kInCh) = o;
}

// s = max |r| over WATER cells (non-negative floats order like their bit patterns)
__global__ void k_absmax(const float* __restrict__ r, const uint16_t* __restrict__ nbinfo, int64_t n,
                         unsigned* __restrict__ out) {
    unsigned m = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (nbinfo[i] & NB_FLUID) m = max(m, __float_as_uint(fabsf(r[i])));
    for (int d = 32; d >= 1; d >>= 1) m = max(m, (unsigned)__shfl_down((int)m, d, 64));
    if ((threadIdx.x & 63) == 0 && m) atomicMax(out, m);
}
__global__ void k_absmax_op(const float* __restrict__ b, const float* __restrict__ diag, int64_t n,
                            unsigned* __restrict__ out) {
    unsigned m = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (diag[i] > 0.f) m = max(m, __float_as_uint(fabsf(b[i])));
    for (int d = 32; d >= 1; d >>= 1) m = max(m, (unsigned)__shfl_down((int)m, d, 64));
    if ((threadIdx.x & 63) == 0 && m) atomicMax(out, m);
}

// out[v] = b + sum_c d0[v][c] * w[c]  (FP32)
__global__ void k_conv1_fwd(const __bf16* __restrict__ d0, const float* __restrict__ w, const float* __restrict__ b,
                            float* __restrict__ out, int64_t V) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    float s = *b;
    for (int c = 0; c < kOutIn; c++) s += (float)d0[v * kOutIn + c] * (float)(__bf16)w[c];
    out[v] = s;
}

// z (solver layout) = WATER ? out * s * mult : 0
__global__ void k_net_output(const float* __restrict__ out, const uint16_t* __restrict__ nbinfo,
                             const float* __restrict__ scale, float mult, float* __restrict__ z, Dims nd, int nx, int ny,
                             int nz) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nc = (int64_t)nx * ny * nz;
    if (i >= nc) return;
    // i is a flat cell index; write its rotated slot
    const int zc = (int)(i % nz), yc = (int)((i / nz) % ny), xc = (int)(i / ((int64_t)nz * ny));
    const int64_t ri = ridx(xc, yc, zc, nx, ny, nz);
    const int64_t v = ((int64_t)xc * nd.H + yc) * nd.W + zc;
    if (nbinfo[ri] & NB_FLUID) z[ri] = out[v] * *scale * mult;
}
// x (flat level) = unknown ? out * s * mult : 0
__global__ void k_net_output_op(const float* __restrict__ out, const float* __restrict__ diag,
                                const float* __restrict__ scale, float mult, float* __restrict__ x, Dims nd, int nx,
                                int ny, int nz) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nx * ny * nz) return;
    const int zc = (int)(i % nz), yc = (int)((i / nz) % ny), xc = (int)(i / ((int64_t)nz * ny));
    const int64_t v = ((int64_t)xc * nd.H + yc) * nd.W + zc;
    x[i] = diag[i] > 0.f ? out[v] * *scale * mult : 0.f;
}

// ---- backward pieces (training) ----------------------------------------------------------
// g = dy * (y > 0), as FP32 and BF16 (the BF16 copy feeds the MFMA data-gradient convolution)
__global__ void k_relu_bwd(const float* __restrict__ dy, const __bf16* __restrict__ y, float* __restrict__ g,
                           __bf16* __restrict__ g16, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = (float)y[i] > 0.f ? dy[i] : 0.f;
    g[i] = v;
    g16[i] = (__bf16)v;
}

// encoder gradient: g = (dskip + dpool[parent] / 8) * (y > 0)
__global__ void k_pool_bwd_relu(const float* __restrict__ dskip, const float* __restrict__ dpool,
                                const __bf16* __restrict__ y, float* __restrict__ g, __bf16* __restrict__ g16, Dims fine,
                                int C) {
    const int64_t n = fine.V() * C;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % C);
    const int64_t v = i / C;
    const int z = (int)(v % fine.W), yy = (int)((v / fine.W) % fine.H), x = (int)(v / ((int64_t)fine.W * fine.H));
    const int64_t pv = ((int64_t)(x / 2) * (fine.H / 2) + yy / 2) * (fine.W / 2) + z / 2;
    float d = dpool ? dpool[pv * C + c] * 0.125f : 0.f;
    if (dskip) d += dskip[i];
    const float o = (float)y[i] > 0.f ? d : 0.f;
    g[i] = o;
    g16[i] = (__bf16)o;
}

// gradient of cat(up(coarse), skip): dcoarse = sum over the 8 children, dskip = slice
__global__ void k_upcat_bwd_up(const float* __restrict__ dcat, float* __restrict__ dcoarse, Dims fine, int Cu, int Cs) {
    const Dims cs{fine.D / 2, fine.H / 2, fine.W / 2};
    const int C = Cu + Cs;
    const int64_t n = cs.V() * Cu;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % Cu);
    const int64_t v = i / Cu;
    const int z = (int)(v % cs.W), y = (int)((v / cs.W) % cs.H), x = (int)(v / ((int64_t)cs.W * cs.H));
    float s = 0.f;
    for (int d = 0; d < 8; d++) {
        const int64_t fv = ((int64_t)(2 * x + (d >> 2)) * fine.H + 2 * y + ((d >> 1) & 1)) * fine.W + 2 * z + (d & 1);
        s += dcat[fv * C + c];
    }
    dcoarse[i] = s;
}
__global__ void k_upcat_bwd_skip(const float* __restrict__ dcat, float* __restrict__ dskip, Dims fine, int Cu, int Cs) {
    const int64_t n = fine.V() * Cs;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % Cs);
    const int64_t v = i / Cs;
    dskip[i] = dcat[v * (Cu + Cs) + Cu + c];
}

// data-gradient weights: Wb[ci][t' * Cout + co] = W[co][(26 - t') * Cin + ci]
__global__ void k_flip_weights(const __bf16* __restrict__ w, __bf16* __restrict__ wb, int cin, int cout) {
    const int64_t n = (int64_t)cout * 27 * cin;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int co = (int)(i % cout);
    const int64_t r = i / cout;
    const int tp = (int)(r % 27), ci = (int)(r / 27);
    wb[i] = w[(int64_t)co * 27 * cin + (int64_t)(26 - tp) * cin + ci];
}

// Weight gradient of one 3x3x3 convolution on MFMA: dW[co][t][ci] = sum_v g[v][co] x[v + off(t)][ci],
// one GEMM per tap with M = Cout, N = Cin, K = voxels.  Block (voxel slice, tap t): its 4 waves take
// the slice's 32-voxel chunks in turn, each accumulating every 16 x 16 (co, ci) tile with
// v_mfma_f32_16x16x32_bf16; a wave's partial goes to partials[slice * 4 + wave][co][t][ci] and the
// partials are summed in that order (k_sum_parts).  A (16 co x 32 v): lane l holds co = 16 mi + l % 16
// at voxels v0 + 8 (l / 16) .. + 7; B (32 v x 16 ci): ci = 16 ni + l % 16 at the same voxels,
// shifted by the tap (zero outside the grid; Cin = 8 is padded to one 16-wide tile).  Operands are
// the layer's BF16 output gradient and its BF16 input activations, sums in FP32.
constexpr int kWgWaves = 4;
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_conv3_wgrad(const __bf16* __restrict__ g, const __bf16* __restrict__ x,
                                                     float* __restrict__ partials, Dims d, int64_t vox_per_block) {
    constexpr int MT = COUT / 16, NT = (CIN + 15) / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = blockIdx.y;
    const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
    const int64_t V = d.V();
    const int64_t vb = (int64_t)blockIdx.x * vox_per_block, ve = min(V, vb + vox_per_block);
    const int cl = lane & 15, kq = lane >> 4;
    const __bf16 zero = (__bf16)0.f;
    f32x4 acc[MT][NT];
#pragma unroll
    for (int mi = 0; mi < MT; mi++)
#pragma unroll
        for (int ni = 0; ni < NT; ni++) acc[mi][ni] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int64_t v0 = vb + 32 * wave; v0 < ve; v0 += 32 * kWgWaves) {
        bf16x8 a[MT], b[NT];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int64_t v = v0 + 8 * kq + j;
            const bool vin = v < ve;
            int64_t sv = -1;   // the tap's neighbour of voxel v
            if (vin) {
                const int z = (int)(v % d.W), y = (int)((v / d.W) % d.H), xx = (int)(v / ((int64_t)d.W * d.H));
                const int X = xx + dx, Y = y + dy, Z = z + dz;
                if ((unsigned)X < (unsigned)d.D && (unsigned)Y < (unsigned)d.H && (unsigned)Z < (unsigned)d.W)
                    sv = ((int64_t)X * d.H + Y) * d.W + Z;
            }
#pragma unroll
            for (int mi = 0; mi < MT; mi++) a[mi][j] = vin ? g[v * COUT + 16 * mi + cl] : zero;
#pragma unroll
            for (int ni = 0; ni < NT; ni++) {
                const int ci = 16 * ni + cl;
                b[ni][j] = sv >= 0 && ci < CIN ? x[sv * CIN + ci] : zero;
            }
        }
#pragma unroll
        for (int mi = 0; mi < MT; mi++)
#pragma unroll
            for (int ni = 0; ni < NT; ni++) acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mi], b[ni], acc[mi][ni], 0, 0, 0);
    }
    float* out = partials + ((int64_t)blockIdx.x * kWgWaves + wave) * COUT * 27 * CIN;
#pragma unroll
    for (int mi = 0; mi < MT; mi++)
#pragma unroll
        for (int ni = 0; ni < NT; ni++) {
            const int ci = 16 * ni + cl;
            if (ci >= CIN) continue;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int co = 16 * mi + 4 * kq + i;
                out[(int64_t)co * 27 * CIN + (int64_t)t * CIN + ci] = acc[mi][ni][i];
            }
        }
}

// sum of nparts partial vectors of length n, in partial order
__global__ void k_sum_parts(const float* __restrict__ parts, int nparts, int64_t n, float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int p = 0; p < nparts; p++) s += parts[(int64_t)p * n + i];
    out[i] = s;
}

// bias gradient: db[c] = sum_v g[v][c] (one block per channel, fixed order)
__global__ void __launch_bounds__(256) k_bias_grad(const float* __restrict__ g, int64_t V, int C, float* __restrict__ db) {
    __shared__ float red[256];
    const int c = blockIdx.x;
    float s = 0.f;
    for (int64_t v = threadIdx.x; v < V; v += blockDim.x) s += g[v * C + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = red[0];
}

// output layer backward: dd0[v][c] = dout[v] * w[c];  dw[c] = sum_v dout[v] d0[v][c];  db = sum dout
__global__ void k_conv1_bwd_data(const float* __restrict__ dout, const float* __restrict__ w, float* __restrict__ dd0,
                                 int64_t V) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V * kOutIn) return;
    dd0[i] = dout[i / kOutIn] * (float)(__bf16)w[i % kOutIn];
}
__global__ void __launch_bounds__(256) k_conv1_bwd_weight(const float* __restrict__ dout, const __bf16* __restrict__ d0,
                                                          int64_t V, float* __restrict__ dw) {
    __shared__ float red[256];
    const int c = blockIdx.x;   // 0..15 weights, 16 = bias
    float s = 0.f;
    for (int64_t v = threadIdx.x; v < V; v += blockDim.x) s += dout[v] * (c < kOutIn ? (float)d0[v * kOutIn + c] : 1.f);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) dw[c] = red[0];
}

// masked MSE over WATER cells: loss = sum (out - t)^2 / n_water, dout = 2 (out - t) / n_water
__global__ void k_mse(const float* __restrict__ out, const float* __restrict__ target, const uint8_t* __restrict__ fluid,
                      int64_t V, float inv_n, float* __restrict__ dout, float* __restrict__ loss_parts) {
    __shared__ float red[256];
    float s = 0.f;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        const float e = fluid[v] ? out[v] - target[v] : 0.f;
        dout[v] = 2.f * e * inv_n;
        s += e * e;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_parts[blockIdx.x] = red[0] * inv_n;
}

// Energy-norm loss: loss = e^T A e / t^T A t with e = (out - t) on the
// sample's unknowns, dout = 2 A e / t^T A t.  A is the sample's operator in
// network layout: diagonal aop[0], couplings -aop[1..3] to the +x, +y, +z
// neighbour (symmetric).  Step 1: e into a buffer.
__global__ void k_energy_err(const float* __restrict__ out, const float* __restrict__ target,
                             const uint8_t* __restrict__ fluid, int64_t V, float* __restrict__ e) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) e[v] = fluid[v] ? out[v] - target[v] : 0.f;
}
// Step 2: A e (e is zero off the unknowns, whose couplings are zero).
__global__ void k_energy(const float* __restrict__ e, const float* __restrict__ aop, Dims d, float inv_et,
                         float* __restrict__ dout, float* __restrict__ loss_parts) {
    __shared__ float red[256];
    const int64_t V = d.V();
    const float *dg = aop, *cx = aop + V, *cy = aop + 2 * V, *cz = aop + 3 * V;
    float s = 0.f;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        float ae = 0.f;
        if (dg[v] > 0.f) {
            const int z = (int)(v % d.W), y = (int)((v / d.W) % d.H), x = (int)(v / ((int64_t)d.W * d.H));
            const int64_t sx = (int64_t)d.H * d.W;
            ae = dg[v] * e[v];
            if (x > 0) ae -= cx[v - sx] * e[v - sx];
            if (x < d.D - 1) ae -= cx[v] * e[v + sx];
            if (y > 0) ae -= cy[v - d.W] * e[v - d.W];
            if (y < d.H - 1) ae -= cy[v] * e[v + d.W];
            if (z > 0) ae -= cz[v - 1] * e[v - 1];
            if (z < d.W - 1) ae -= cz[v] * e[v + 1];
        }
        dout[v] = 2.f * ae * inv_et;
        s += e[v] * ae;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_parts[blockIdx.x] = red[0] * inv_et;
}

// Adam on FP32 master weights; BF16 copy of the conv weights for the MFMA kernels
__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v, const float* __restrict__ g,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2, float gscale) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] * gscale;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
__global__ void k_to_bf16(const float* __restrict__ a, __bf16* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (__bf16)a[i];
}

// ---- host side ----------------------------------------------------------------------------
// Parameter-sized buffers are allocated once per context; the activation and
// training buffers are sized for the grid the network runs on (the coarsest
// multigrid level of a solve, or the context's own grid for training
// samples) and are reallocated when that grid changes.
struct LearnedNet {
    int64_t woff[kNetLayers], boff[kNetLayers], out_w = 0, out_b = 0, nparams = 0;
    bool has_weights = false;
    int adam_t = 0;
    int gx = -1, gy = -1, gz = -1;   // the grid the per-size buffers are for
    Dims lv[4];
    float *w32 = nullptr, *adam_m = nullptr, *adam_v = nullptr, *grad = nullptr, *grad_acc = nullptr;
    __bf16* w16 = nullptr;
    __bf16* wflip = nullptr;
    float* wparts = nullptr;
    int64_t wparts_cap = 0;
    float* loss_parts = nullptr;
    bool train_params_ready = false;
    // forward (per size)
    __bf16 *x0 = nullptr, *e[4] = {}, *pl[3] = {}, *cat[3] = {}, *dd[3] = {};   // cat/dd index = level
    float *out = nullptr, *scale = nullptr;
    unsigned* scale_bits = nullptr;
    float* part = nullptr;            // split-K partial sums of the small levels' convolutions
    int64_t part_cap = 0;
    // training (per size)
    bool train_ready = false;
    float *in_b = nullptr, *in_op = nullptr, *target = nullptr, *dout = nullptr, *dd0 = nullptr;
    uint8_t* fluid = nullptr;
    float* aop = nullptr;             // the sample's operator in network layout (energy loss)
    float* eres = nullptr;            // energy loss: out - target on the unknowns
    double energy_t = 0.0;            // energy loss: t^T A t of the uploaded sample
    float* g32[kNetLayers] = {};
    __bf16* g16[kNetLayers] = {};
    float *dcat[3] = {}, *dcoarse[3] = {}, *dskip[3] = {}, *dpool[3] = {};
    std::vector<void*> pallocs, allocs;   // parameter-sized, per-size
};

template <typename T>
static int nalloc(fsim_ctx* ctx, std::vector<void*>& list, T** p, int64_t n) {
    if (hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (size_t)std::max<int64_t>(n, 1)) != hipSuccess)
        return set_err(ctx, FSIM_ERR_OOM, "learned preconditioner: device allocation failed");
    list.push_back(*p);
    return FSIM_OK;
}

static void free_sized(LearnedNet* N) {
    for (void* p : N->allocs) (void)hipFree(p);
    N->allocs.clear();
    N->gx = N->gy = N->gz = -1;
    N->train_ready = false;
}

void learned_free(fsim_ctx* ctx) {
    LearnedNet* N = ctx->learned;
    if (!N) return;
    free_sized(N);
    for (void* p : N->pallocs) (void)hipFree(p);
    delete N;
    ctx->learned = nullptr;
}

static unsigned gb(int64_t n) { return (unsigned)((n + 255) / 256); }

static int learned_alloc_params(fsim_ctx* ctx) {
    if (ctx->learned) return FSIM_OK;
    LearnedNet* N = new LearnedNet();
    ctx->learned = N;
    param_offsets(N->woff, N->boff, &N->out_w, &N->out_b, &N->nparams);
    int rc;
    N->part_cap = kConvPartFloats;
    if ((rc = nalloc(ctx, N->pallocs, &N->w32, N->nparams)) || (rc = nalloc(ctx, N->pallocs, &N->w16, N->nparams)) ||
        (rc = nalloc(ctx, N->pallocs, &N->part, N->part_cap)))
        return rc;
    return FSIM_OK;
}

// forward buffers for a network running on an nx x ny x nz grid (padded to multiples of 8)
static int learned_alloc(fsim_ctx* ctx, int nx, int ny, int nz) {
    FSIM_TRY(learned_alloc_params(ctx));
    LearnedNet* N = ctx->learned;
    if (N->gx == nx && N->gy == ny && N->gz == nz) return FSIM_OK;
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    free_sized(N);
    ctx->learned_gen++;   // a captured V-cycle that used the old buffers is stale
    auto r8 = [](int n) { return (n + 7) / 8 * 8; };
    N->lv[0] = Dims{r8(nx), r8(ny), r8(nz)};
    for (int l = 1; l < 4; l++) N->lv[l] = Dims{N->lv[l - 1].D / 2, N->lv[l - 1].H / 2, N->lv[l - 1].W / 2};
    const int64_t V0 = N->lv[0].V(), V1 = N->lv[1].V(), V2 = N->lv[2].V(), V3 = N->lv[3].V();
    std::vector<void*>& A = N->allocs;
    int rc;
    if ((rc = nalloc(ctx, A, &N->x0, V0 * kInCh)) || (rc = nalloc(ctx, A, &N->e[0], V0 * 16)) ||
        (rc = nalloc(ctx, A, &N->e[1], V1 * 32)) || (rc = nalloc(ctx, A, &N->e[2], V2 * 64)) ||
        (rc = nalloc(ctx, A, &N->e[3], V3 * 128)) || (rc = nalloc(ctx, A, &N->pl[0], V1 * 16)) ||
        (rc = nalloc(ctx, A, &N->pl[1], V2 * 32)) || (rc = nalloc(ctx, A, &N->pl[2], V3 * 64)) ||
        (rc = nalloc(ctx, A, &N->dd[2], V2 * 64)) || (rc = nalloc(ctx, A, &N->dd[1], V1 * 32)) ||
        (rc = nalloc(ctx, A, &N->dd[0], V0 * 16)) || (rc = nalloc(ctx, A, &N->out, V0)) ||
        (rc = nalloc(ctx, A, &N->scale, 1)) || (rc = nalloc(ctx, A, &N->scale_bits, 1)))
        return rc;
    N->gx = nx;
    N->gy = ny;
    N->gz = nz;
    return FSIM_OK;
}

// training buffers for samples on the context's own grid
static int learned_alloc_train(fsim_ctx* ctx) {
    FSIM_TRY(learned_alloc(ctx, ctx->g.nx, ctx->g.ny, ctx->g.nz));
    LearnedNet* N = ctx->learned;
    int rc;
    if (!N->train_params_ready) {
        std::vector<void*>& P = N->pallocs;
        if ((rc = nalloc(ctx, P, &N->adam_m, N->nparams)) || (rc = nalloc(ctx, P, &N->adam_v, N->nparams)) ||
            (rc = nalloc(ctx, P, &N->grad, N->nparams)) || (rc = nalloc(ctx, P, &N->grad_acc, N->nparams)) ||
            (rc = nalloc(ctx, P, &N->loss_parts, 1024)) || (rc = nalloc(ctx, P, &N->wflip, 64 * 27 * 192)))
            return rc;
        N->wparts_cap = 64LL * 64 * 27 * 192;
        if ((rc = nalloc(ctx, P, &N->wparts, N->wparts_cap))) return rc;
        FSIM_HIP(ctx, hipMemsetAsync(N->adam_m, 0, sizeof(float) * N->nparams, ctx->stream));
        FSIM_HIP(ctx, hipMemsetAsync(N->adam_v, 0, sizeof(float) * N->nparams, ctx->stream));
        N->adam_t = 0;
        N->train_params_ready = true;
    }
    if (N->train_ready) return FSIM_OK;
    const int64_t V0 = N->lv[0].V(), V1 = N->lv[1].V(), V2 = N->lv[2].V(), V3 = N->lv[3].V();
    const int64_t Vl[4] = {V0, V1, V2, V3};
    std::vector<void*>& A = N->allocs;
    if ((rc = nalloc(ctx, A, &N->in_b, ctx->g.nc)) || (rc = nalloc(ctx, A, &N->in_op, 4 * ctx->g.nc)) ||
        (rc = nalloc(ctx, A, &N->target, V0)) || (rc = nalloc(ctx, A, &N->fluid, V0)) ||
        (rc = nalloc(ctx, A, &N->dout, V0)) || (rc = nalloc(ctx, A, &N->dd0, V0 * 16)) ||
        (rc = nalloc(ctx, A, &N->cat[2], V2 * 192)) || (rc = nalloc(ctx, A, &N->cat[1], V1 * 96)) ||
        (rc = nalloc(ctx, A, &N->cat[0], V0 * 48)) || (rc = nalloc(ctx, A, &N->aop, 4 * V0)) ||
        (rc = nalloc(ctx, A, &N->eres, V0)))
        return rc;
    for (int l = 0; l < kNetLayers; l++) {
        const int64_t n = Vl[kLayers[l].level] * kLayers[l].cout;
        if ((rc = nalloc(ctx, A, &N->g32[l], n)) || (rc = nalloc(ctx, A, &N->g16[l], n))) return rc;
    }
    // concat data gradients at levels 2, 1, 0 and their splits
    const int cu[3] = {32, 64, 128}, cs[3] = {16, 32, 64};   // index = level
    for (int l = 0; l < 3; l++) {
        if ((rc = nalloc(ctx, A, &N->dcat[l], Vl[l] * (cu[l] + cs[l]))) ||
            (rc = nalloc(ctx, A, &N->dcoarse[l], Vl[l + 1] * cu[l])) || (rc = nalloc(ctx, A, &N->dskip[l], Vl[l] * cs[l])))
            return rc;
    }
    // data gradients of the pooled inputs: dpool[l] at level l + 1 with the channels of e[l]
    const int pc[3] = {16, 32, 64};
    for (int l = 0; l < 3; l++)
        if ((rc = nalloc(ctx, A, &N->dpool[l], Vl[l + 1] * pc[l]))) return rc;
    N->train_ready = true;
    return FSIM_OK;
}

// layer l reading x (or, for a decoder layer, the skip x and the coarse parent xc)
static ConvArgs conv_args(const LearnedNet* N, int l, const __bf16* x, void* y, int lds, const __bf16* xc = nullptr) {
    const Dims d = N->lv[kLayers[l].level];
    const int cu = xc ? kLayers[l].cin - kLayers[l].cin / 3 : 0;
    return ConvArgs{x, N->w16 + N->woff[l], N->w32 + N->boff[l], y, d.D, d.H, d.W, kLayers[l].cin, kLayers[l].cout, 1,
                    xc, cu, lds, N->part, N->part_cap};
}

// forward from x0 (already packed) to out; for training the concatenated
// decoder inputs are also stored (the weight gradients read them)
static int net_forward(fsim_ctx* ctx, bool keep_cat) {
    LearnedNet* N = ctx->learned;
    hipStream_t st = ctx->stream;
    const Dims* L = N->lv;
    const int lds = ctx->learned_conv_lds;
    bool ok = launch_conv3<false>(st, conv_args(N, 0, N->x0, N->e[0], lds));
    for (int l = 1; l < 4; l++) {
        hipLaunchKernelGGL(k_pool2, dim3(gb(L[l].V() * kLayers[l].cin)), dim3(256), 0, st, N->e[l - 1], N->pl[l - 1],
                           L[l - 1], kLayers[l].cin);
        ok = ok && launch_conv3<false>(st, conv_args(N, l, N->pl[l - 1], N->e[l], lds));
    }
    // decoder: layer 4 at level 2 (up e3 + e2), 5 at level 1 (up d2 + e1), 6 at level 0 (up d1 + e0)
    const __bf16* coarse = N->e[3];
    for (int l = 4; l < kNetLayers; l++) {
        const int lev = kLayers[l].level;
        const int Cs = kLayers[l].cin / 3, Cu = kLayers[l].cin - Cs;
        if (keep_cat)
            hipLaunchKernelGGL(k_upcat, dim3(gb(L[lev].V() * kLayers[l].cin)), dim3(256), 0, st, coarse, N->e[lev],
                               N->cat[lev], L[lev], Cu, Cs);
        ok = ok && launch_conv3<false>(st, conv_args(N, l, N->e[lev], N->dd[lev], lds, coarse));
        coarse = N->dd[lev];
    }
    hipLaunchKernelGGL(k_conv1_fwd, dim3(gb(L[0].V())), dim3(256), 0, st, N->dd[0], N->w32 + N->out_w, N->w32 + N->out_b,
                       N->out, L[0].V());
    FSIM_HIP(ctx, hipGetLastError());
    if (!ok) return set_err(ctx, FSIM_ERR_INVALID, "learned network: convolution shape without a kernel");
    return FSIM_OK;
}

static int sync_bf16(fsim_ctx* ctx) {
    LearnedNet* N = ctx->learned;
    hipLaunchKernelGGL(k_to_bf16, dim3(gb(N->nparams)), dim3(256), 0, ctx->stream, N->w32, N->w16, N->nparams);
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

bool learned_ready(const fsim_ctx* ctx) { return ctx->learned && ctx->learned->has_weights; }

int learned_set_params(fsim_ctx* ctx, const float* params) {
    FSIM_TRY(learned_alloc_params(ctx));
    LearnedNet* N = ctx->learned;
    FSIM_HIP(ctx, hipMemcpyAsync(N->w32, params, sizeof(float) * N->nparams, hipMemcpyHostToDevice, ctx->stream));
    FSIM_TRY(sync_bf16(ctx));
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    N->has_weights = true;
    N->adam_t = 0;
    if (N->train_params_ready) {
        FSIM_HIP(ctx, hipMemsetAsync(N->adam_m, 0, sizeof(float) * N->nparams, ctx->stream));
        FSIM_HIP(ctx, hipMemsetAsync(N->adam_v, 0, sizeof(float) * N->nparams, ctx->stream));
    }
    return FSIM_OK;
}

int learned_get_params(fsim_ctx* ctx, float* params) {
    if (!learned_ready(ctx)) return set_err(ctx, FSIM_ERR_INVALID, "learned preconditioner has no parameters");
    LearnedNet* N = ctx->learned;
    FSIM_HIP(ctx, hipMemcpyAsync(params, N->w32, sizeof(float) * N->nparams, hipMemcpyDeviceToHost, ctx->stream));
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FSIM_OK;
}

// He-normal initial parameters from a counter RNG (splitmix64 + Box-Muller), biases 0
int learned_init(fsim_ctx* ctx, uint64_t seed) {
    int64_t woff[kNetLayers], boff[kNetLayers], ow, ob, total;
    param_offsets(woff, boff, &ow, &ob, &total);
    std::vector<float> p((size_t)total, 0.f);
    uint64_t ctr = 0;
    auto next = [&]() {
        uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (++ctr);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);   // (0, 1)
    };
    auto normal = [&]() { return std::sqrt(-2.0 * std::log(next())) * std::cos(6.283185307179586 * next()); };
    for (int l = 0; l < kNetLayers; l++) {
        const double sd = std::sqrt(2.0 / (27.0 * kLayers[l].cin));
        for (int64_t i = 0; i < conv_weights(l); i++) p[(size_t)(woff[l] + i)] = (float)(sd * normal());
    }
    for (int c = 0; c < kOutIn; c++) p[(size_t)(ow + c)] = (float)(std::sqrt(1.0 / kOutIn) * normal());
    return learned_set_params(ctx, p.data());
}

// The network on a coarse multigrid level (flat arrays): x = unknown ? out * max|b| * norm : 0
int learned_apply_level(fsim_ctx* ctx, int nx, int ny, int nz, const float* b, const float* diag, const float* ox,
                        const float* oy, const float* oz, double norm, float* x) {
    FSIM_PROF(ctx, KID_LEARNED);
    FSIM_TRY(learned_alloc(ctx, nx, ny, nz));
    LearnedNet* N = ctx->learned;
    hipStream_t st = ctx->stream;
    const int64_t nc = (int64_t)nx * ny * nz;
    FSIM_TRY(fill_i32(ctx, reinterpret_cast<int32_t*>(N->scale_bits), 0, 1));   // a kernel: graph-capturable
    hipLaunchKernelGGL(k_absmax_op, dim3(std::min<unsigned>(gb(nc), 2048u)), dim3(256), 0, st, b, diag, nc, N->scale_bits);
    const float* sc = reinterpret_cast<const float*>(N->scale_bits);   // the max as a float
    hipLaunchKernelGGL(k_net_input_op, dim3(gb(N->lv[0].V())), dim3(256), 0, st, b, diag, ox, oy, oz, (float)norm, sc,
                       N->x0, N->lv[0], nx, ny, nz);
    FSIM_TRY(net_forward(ctx, false));
    hipLaunchKernelGGL(k_net_output_op, dim3(gb(nc)), dim3(256), 0, st, N->out, diag, sc, (float)norm, x, N->lv[0], nx, ny,
                       nz);
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

// The network's buffers for a level of nx x ny x nz cells, made before a
// multigrid V-cycle is captured into a graph (nothing may allocate then).
int learned_prepare(fsim_ctx* ctx, int nx, int ny, int nz) { return learned_alloc(ctx, nx, ny, nz); }

// The network on the solve's own grid (rotated r and z): z = WATER ? out * max|r| * binv : 0
int learned_apply_level0(fsim_ctx* ctx, const float* r, float* z, float binv) {
    FSIM_PROF(ctx, KID_LEARNED);
    FSIM_TRY(learned_alloc(ctx, ctx->g.nx, ctx->g.ny, ctx->g.nz));
    LearnedNet* N = ctx->learned;
    hipStream_t st = ctx->stream;
    const int64_t nc = ctx->g.nc;
    FSIM_TRY(fill_i32(ctx, reinterpret_cast<int32_t*>(N->scale_bits), 0, 1));   // a kernel: graph-capturable
    hipLaunchKernelGGL(k_absmax, dim3(std::min<unsigned>(gb(nc), 2048u)), dim3(256), 0, st, r, ctx->nbinfo, nc,
                       N->scale_bits);
    static_assert(sizeof(unsigned) == sizeof(float), "scale bits");
    const float* sc = reinterpret_cast<const float*>(N->scale_bits);
    hipLaunchKernelGGL(k_net_input, dim3(gb(N->lv[0].V())), dim3(256), 0, st, r, ctx->nbinfo, sc, N->x0, N->lv[0],
                       ctx->g.nx, ctx->g.ny, ctx->g.nz);
    FSIM_TRY(net_forward(ctx, false));
    hipLaunchKernelGGL(k_net_output, dim3(gb(nc)), dim3(256), 0, st, N->out, ctx->nbinfo, sc, binv, z, N->lv[0], ctx->g.nx,
                       ctx->g.ny, ctx->g.nz);
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

// The 7-point operator of a label field (flat order): diag = non-SOLID
// neighbours of a WATER cell (cells outside the grid count as SOLID),
// couplings 1 towards a WATER +x / +y / +z neighbour.
void label_operator(int nx, int ny, int nz, const uint8_t* label, float* op) {
    const int64_t nc = (int64_t)nx * ny * nz;
    auto lab = [&](int x, int y, int z) -> int {
        if (x < 0 || x >= nx || y < 0 || y >= ny || z < 0 || z >= nz) return FSIM_SOLID;
        return label[((int64_t)x * ny + y) * nz + z];
    };
    static const int nb6[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++)
            for (int z = 0; z < nz; z++) {
                const int64_t i = ((int64_t)x * ny + y) * nz + z;
                float d = 0.f, c[3] = {0.f, 0.f, 0.f};
                if (label[i] == FSIM_WATER) {
                    for (const auto& o : nb6) d += lab(x + o[0], y + o[1], z + o[2]) != FSIM_SOLID ? 1.f : 0.f;
                    c[0] = lab(x + 1, y, z) == FSIM_WATER ? 1.f : 0.f;
                    c[1] = lab(x, y + 1, z) == FSIM_WATER ? 1.f : 0.f;
                    c[2] = lab(x, y, z + 1) == FSIM_WATER ? 1.f : 0.f;
                }
                op[i] = d;
                op[nc + i] = c[0];
                op[2 * nc + i] = c[1];
                op[3 * nc + i] = c[2];
            }
}

// Uploads one sample in flat grid order: input (b / s, the operator's
// channels), and for training the target p / s, the unknown mask and the
// operator in network order.  s = max |b| over the unknowns (diag > 0).
static int upload_sample(fsim_ctx* ctx, const float* b, const float* op, const double* p, float* s_out) {
    LearnedNet* N = ctx->learned;
    const int nx = ctx->g.nx, ny = ctx->g.ny, nz = ctx->g.nz;
    const int64_t nc = ctx->g.nc;
    float s = 0.f;
    for (int64_t i = 0; i < nc; i++)
        if (op[i] > 0.f) s = std::max(s, std::fabs(b[i]));
    *s_out = s;
    FSIM_HIP(ctx, hipMemcpyAsync(N->in_b, b, sizeof(float) * nc, hipMemcpyHostToDevice, ctx->stream));
    FSIM_HIP(ctx, hipMemcpyAsync(N->in_op, op, sizeof(float) * 4 * nc, hipMemcpyHostToDevice, ctx->stream));
    FSIM_HIP(ctx, hipMemcpyAsync(N->scale, &s, sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_net_input_op, dim3(gb(N->lv[0].V())), dim3(256), 0, ctx->stream, N->in_b, N->in_op,
                       N->in_op + nc, N->in_op + 2 * nc, N->in_op + 3 * nc, 1.0f, N->scale, N->x0, N->lv[0], nx, ny, nz);
    FSIM_HIP(ctx, hipGetLastError());
    if (p) {
        const Dims d = N->lv[0];
        const int64_t V = d.V();
        std::vector<float> t((size_t)V, 0.f), a((size_t)(4 * V), 0.f);
        std::vector<uint8_t> f((size_t)V, 0);
        const float inv = s > 0.f ? 1.f / s : 0.f;
        for (int x = 0; x < nx; x++)
            for (int y = 0; y < ny; y++)
                for (int z = 0; z < nz; z++) {
                    const int64_t i = ((int64_t)x * ny + y) * nz + z, v = ((int64_t)x * d.H + y) * d.W + z;
                    if (!(op[i] > 0.f)) continue;
                    f[(size_t)v] = 1;
                    t[(size_t)v] = (float)p[i] * inv;
                    for (int k = 0; k < 4; k++) a[(size_t)(k * V + v)] = op[k * nc + i];
                }
        // t^T A t in FP64 over the unknowns
        double et = 0.0;
        const int64_t sx = (int64_t)d.H * d.W, sy = d.W;
        for (int x = 0; x < nx; x++)
            for (int y = 0; y < ny; y++)
                for (int z = 0; z < nz; z++) {
                    const int64_t v = ((int64_t)x * d.H + y) * d.W + z;
                    if (!f[(size_t)v]) continue;
                    double at = (double)a[(size_t)v] * t[(size_t)v];
                    if (x > 0) at -= (double)a[(size_t)(V + v - sx)] * t[(size_t)(v - sx)];
                    if (x < d.D - 1) at -= (double)a[(size_t)(V + v)] * t[(size_t)(v + sx)];
                    if (y > 0) at -= (double)a[(size_t)(2 * V + v - sy)] * t[(size_t)(v - sy)];
                    if (y < d.H - 1) at -= (double)a[(size_t)(2 * V + v)] * t[(size_t)(v + sy)];
                    if (z > 0) at -= (double)a[(size_t)(3 * V + v - 1)] * t[(size_t)(v - 1)];
                    if (z < d.W - 1) at -= (double)a[(size_t)(3 * V + v)] * t[(size_t)(v + 1)];
                    et += (double)t[(size_t)v] * at;
                }
        N->energy_t = et;
        FSIM_HIP(ctx, hipMemcpyAsync(N->target, t.data(), sizeof(float) * t.size(), hipMemcpyHostToDevice, ctx->stream));
        FSIM_HIP(ctx, hipMemcpyAsync(N->fluid, f.data(), f.size(), hipMemcpyHostToDevice, ctx->stream));
        FSIM_HIP(ctx, hipMemcpyAsync(N->aop, a.data(), sizeof(float) * a.size(), hipMemcpyHostToDevice, ctx->stream));
        // the host vectors go out of scope: wait for the copies
        FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return FSIM_OK;
}

// network output for one flat-order sample (raw network output, flat grid order)
int learned_forward_host(fsim_ctx* ctx, const float* b, const float* op, float* out) {
    if (!learned_ready(ctx)) return set_err(ctx, FSIM_ERR_INVALID, "learned preconditioner has no parameters");
    FSIM_TRY(learned_alloc_train(ctx));
    LearnedNet* N = ctx->learned;
    float s;
    FSIM_TRY(upload_sample(ctx, b, op, nullptr, &s));
    FSIM_TRY(net_forward(ctx, false));
    const Dims d = N->lv[0];
    std::vector<float> o((size_t)d.V());
    FSIM_HIP(ctx, hipMemcpyAsync(o.data(), N->out, sizeof(float) * o.size(), hipMemcpyDeviceToHost, ctx->stream));
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const int nx = ctx->g.nx, ny = ctx->g.ny, nz = ctx->g.nz;
    for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++)
            for (int z = 0; z < nz; z++)
                out[((int64_t)x * ny + y) * nz + z] = o[(size_t)(((int64_t)x * d.H + y) * d.W + z)];
    return FSIM_OK;
}

// weight and bias gradients of layer l from its output gradient (FP32 g32, BF16 g16) and input x
static int wgrad(fsim_ctx* ctx, int l, const float* g32, const __bf16* g16, const __bf16* x) {
    LearnedNet* N = ctx->learned;
    const Dims d = N->lv[kLayers[l].level];
    const int cin = kLayers[l].cin, cout = kLayers[l].cout;
    const int64_t V = d.V();
    const int64_t n = (int64_t)cout * 27 * cin;
    // voxel slices: at least one 32-voxel chunk per wave, at most 16, within the partial buffer
    int nblk = (int)std::max<int64_t>(1, std::min<int64_t>(16, V / (32 * kWgWaves)));
    while (nblk > 1 && (int64_t)nblk * kWgWaves * n > N->wparts_cap) nblk--;
    if ((int64_t)nblk * kWgWaves * n > N->wparts_cap)
        return set_err(ctx, FSIM_ERR_INVALID, "learned network: weight-gradient partials exceed their buffer");
    const int64_t per = (V + nblk - 1) / nblk;
    const dim3 grid(nblk, 27), block(256);
    bool ok = true;
    switch (l) {
    case 0: hipLaunchKernelGGL((k_conv3_wgrad<8, 16>), grid, block, 0, ctx->stream, g16, x, N->wparts, d, per); break;
    case 1: hipLaunchKernelGGL((k_conv3_wgrad<16, 32>), grid, block, 0, ctx->stream, g16, x, N->wparts, d, per); break;
    case 2: hipLaunchKernelGGL((k_conv3_wgrad<32, 64>), grid, block, 0, ctx->stream, g16, x, N->wparts, d, per); break;
    case 3: hipLaunchKernelGGL((k_conv3_wgrad<64, 128>), grid, block, 0, ctx->stream, g16, x, N->wparts, d, per); break;
    case 4: hipLaunchKernelGGL((k_conv3_wgrad<192, 64>), grid, block, 0, ctx->stream, g16, x, N->wparts, d, per); break;
    case 5: hipLaunchKernelGGL((k_conv3_wgrad<96, 32>), grid, block, 0, ctx->stream, g16, x, N->wparts, d, per); break;
    case 6: hipLaunchKernelGGL((k_conv3_wgrad<48, 16>), grid, block, 0, ctx->stream, g16, x, N->wparts, d, per); break;
    default: ok = false;
    }
    if (!ok) return set_err(ctx, FSIM_ERR_INVALID, "learned network: no weight-gradient kernel for this layer");
    hipLaunchKernelGGL(k_sum_parts, dim3(gb(n)), dim3(256), 0, ctx->stream, N->wparts, nblk * kWgWaves, n,
                       N->grad + N->woff[l]);
    hipLaunchKernelGGL(k_bias_grad, dim3(cout), dim3(256), 0, ctx->stream, g32, V, cout, N->grad + N->boff[l]);
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

// data gradient of layer l's input: dx = conv3(g16, flipped weights)
static int dgrad(fsim_ctx* ctx, int l, const __bf16* g16, float* dx) {
    LearnedNet* N = ctx->learned;
    const Dims d = N->lv[kLayers[l].level];
    const int cin = kLayers[l].cin, cout = kLayers[l].cout;
    hipLaunchKernelGGL(k_flip_weights, dim3(gb((int64_t)cout * 27 * cin)), dim3(256), 0, ctx->stream, N->w16 + N->woff[l],
                       N->wflip, cin, cout);
    ConvArgs a{g16, N->wflip, nullptr, dx, d.D, d.H, d.W, cout, cin, 0, nullptr, 0, 0, N->part, N->part_cap};
    if (!launch_conv3<true>(ctx->stream, a))
        return set_err(ctx, FSIM_ERR_INVALID, "learned network: convolution shape without a kernel");
    FSIM_HIP(ctx, hipGetLastError());
    return FSIM_OK;
}

// loss and gradient of one uploaded sample into N->grad; returns the loss
static int net_backward(fsim_ctx* ctx, double* loss) {
    LearnedNet* N = ctx->learned;
    hipStream_t st = ctx->stream;
    const Dims* L = N->lv;
    const int64_t V0 = L[0].V();
    int64_t nfl = 0;
    {
        std::vector<uint8_t> f((size_t)V0);
        FSIM_HIP(ctx, hipMemcpyAsync(f.data(), N->fluid, f.size(), hipMemcpyDeviceToHost, st));
        FSIM_HIP(ctx, hipStreamSynchronize(st));
        for (uint8_t c : f) nfl += c;
    }
    const float inv_n = nfl > 0 ? 1.f / (float)nfl : 0.f;
    const int nlb = 256;
    if (ctx->learned_loss == 1) {
        const float inv_et = N->energy_t > 0.0 ? (float)(1.0 / N->energy_t) : 0.f;
        hipLaunchKernelGGL(k_energy_err, dim3(gb(V0)), dim3(256), 0, st, N->out, N->target, N->fluid, V0, N->eres);
        hipLaunchKernelGGL(k_energy, dim3(nlb), dim3(256), 0, st, N->eres, N->aop, L[0], inv_et, N->dout, N->loss_parts);
    } else {
        hipLaunchKernelGGL(k_mse, dim3(nlb), dim3(256), 0, st, N->out, N->target, N->fluid, V0, inv_n, N->dout,
                           N->loss_parts);
    }
    // output layer
    hipLaunchKernelGGL(k_conv1_bwd_data, dim3(gb(V0 * kOutIn)), dim3(256), 0, st, N->dout, N->w32 + N->out_w, N->dd0, V0);
    hipLaunchKernelGGL(k_conv1_bwd_weight, dim3(kOutIn + 1), dim3(256), 0, st, N->dout, N->dd[0], V0, N->grad + N->out_w);
    // decoder: layer 6 (level 0), 5 (level 1), 4 (level 2)
    const float* dy = N->dd0;
    for (int l = 6; l >= 4; l--) {
        const int lev = kLayers[l].level;
        const int64_t n = L[lev].V() * kLayers[l].cout;
        hipLaunchKernelGGL(k_relu_bwd, dim3(gb(n)), dim3(256), 0, st, dy, N->dd[lev], N->g32[l], N->g16[l], n);
        FSIM_TRY(wgrad(ctx, l, N->g32[l], N->g16[l], N->cat[lev]));
        FSIM_TRY(dgrad(ctx, l, N->g16[l], N->dcat[lev]));
        const int Cs = kLayers[l].cin / 3, Cu = kLayers[l].cin - Cs;
        hipLaunchKernelGGL(k_upcat_bwd_up, dim3(gb(L[lev + 1].V() * Cu)), dim3(256), 0, st, N->dcat[lev], N->dcoarse[lev],
                           L[lev], Cu, Cs);
        hipLaunchKernelGGL(k_upcat_bwd_skip, dim3(gb(L[lev].V() * Cs)), dim3(256), 0, st, N->dcat[lev], N->dskip[lev],
                           L[lev], Cu, Cs);
        dy = N->dcoarse[lev];   // gradient of the coarse decoder output (d1, d2) or of e3
    }
    // bottleneck, layer 3 at level 3: its output e3's gradient is dcoarse[2]
    {
        const int64_t n = L[3].V() * 128;
        hipLaunchKernelGGL(k_relu_bwd, dim3(gb(n)), dim3(256), 0, st, N->dcoarse[2], N->e[3], N->g32[3], N->g16[3], n);
        FSIM_TRY(wgrad(ctx, 3, N->g32[3], N->g16[3], N->pl[2]));
        FSIM_TRY(dgrad(ctx, 3, N->g16[3], N->dpool[2]));
    }
    // encoder layers 2, 1, 0: gradient of e[l] = skip gradient + pooled-input gradient of layer l + 1
    for (int l = 2; l >= 0; l--) {
        const int C = kLayers[l].cout;
        hipLaunchKernelGGL(k_pool_bwd_relu, dim3(gb(L[l].V() * C)), dim3(256), 0, st, N->dskip[l], N->dpool[l], N->e[l],
                           N->g32[l], N->g16[l], L[l], C);
        FSIM_TRY(wgrad(ctx, l, N->g32[l], N->g16[l], l == 0 ? N->x0 : N->pl[l - 1]));
        if (l > 0) FSIM_TRY(dgrad(ctx, l, N->g16[l], N->dpool[l - 1]));
    }
    FSIM_HIP(ctx, hipGetLastError());
    std::vector<float> lp(nlb);
    FSIM_HIP(ctx, hipMemcpyAsync(lp.data(), N->loss_parts, sizeof(float) * nlb, hipMemcpyDeviceToHost, st));
    FSIM_HIP(ctx, hipStreamSynchronize(st));
    double s = 0;
    for (float v : lp) s += v;
    *loss = s;
    return FSIM_OK;
}

int learned_grad_host(fsim_ctx* ctx, const float* b, const float* op, const double* p, float* grad, double* loss) {
    if (!learned_ready(ctx)) return set_err(ctx, FSIM_ERR_INVALID, "learned preconditioner has no parameters");
    FSIM_TRY(learned_alloc_train(ctx));
    LearnedNet* N = ctx->learned;
    float s;
    FSIM_TRY(upload_sample(ctx, b, op, p, &s));
    FSIM_TRY(net_forward(ctx, true));
    FSIM_TRY(net_backward(ctx, loss));
    FSIM_HIP(ctx, hipMemcpyAsync(grad, N->grad, sizeof(float) * N->nparams, hipMemcpyDeviceToHost, ctx->stream));
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FSIM_OK;
}

__global__ void k_axpy_f(float* __restrict__ y, const float* __restrict__ x, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += x[i];
}

// One Adam step on the mean gradient of n samples (flat grid order, stride nc each, operators 4 nc each).
int learned_train_batch(fsim_ctx* ctx, int n, const float* b, const float* op, const double* p, float lr,
                        double* mean_loss) {
    if (!learned_ready(ctx)) return set_err(ctx, FSIM_ERR_INVALID, "learned preconditioner has no parameters");
    if (n <= 0) return set_err(ctx, FSIM_ERR_INVALID, "empty training batch");
    FSIM_TRY(learned_alloc_train(ctx));
    LearnedNet* N = ctx->learned;
    const int64_t nc = ctx->g.nc;
    FSIM_HIP(ctx, hipMemsetAsync(N->grad_acc, 0, sizeof(float) * N->nparams, ctx->stream));
    double lsum = 0;
    for (int k = 0; k < n; k++) {
        float s;
        double l;
        FSIM_TRY(upload_sample(ctx, b + k * nc, op + 4 * k * nc, p + k * nc, &s));
        FSIM_TRY(net_forward(ctx, true));
        FSIM_TRY(net_backward(ctx, &l));
        lsum += l;
        hipLaunchKernelGGL(k_axpy_f, dim3(gb(N->nparams)), dim3(256), 0, ctx->stream, N->grad_acc, N->grad, N->nparams);
    }
    N->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)N->adam_t), c2 = 1.f - std::pow(b2, (float)N->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(gb(N->nparams)), dim3(256), 0, ctx->stream, N->w32, N->adam_m, N->adam_v, N->grad_acc,
                       N->nparams, lr, b1, b2, 1e-8f, c1, c2, 1.f / (float)n);
    FSIM_TRY(sync_bf16(ctx));
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *mean_loss = lsum / n;
    return FSIM_OK;
}

}  // namespace fsim
