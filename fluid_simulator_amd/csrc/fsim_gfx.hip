// fsim_gfx.hip — the particle snapshot the reference's SimulationManager
// builds for the renderer before every step (simulationManager.cpp:218-231,
// returned by getParticleGfxData, :40-43).
//
// One kernel over the device particles writes, at each particle's index in
// the reference's particle vector (its id), the packed record
// {float3 pos, |v|, density}; one device-to-host copy hands it over.  The
// density is the trilinear interpolation of the cells' particle density
// (avgPNum) from the 8 cell centres of getCellsAround (macGrid.cpp:188-202)
// with the weights of trilinearInterpoll (interpolation.h:7-20), summed in
// FP64 per term into an FP32 total in the reference's cell order, as its
// `float density += double` does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "fsim_internal.hpp"

namespace fsim {

// Staging of the allocation-free snapshot (fsim_gfx_snapshot_noalloc): a
// fixed array in every GPU's copy of this code object, held until the
// process ends.  Host code never passes `scratch` itself to a kernel or a
// copy (on the host it is only a stand-in for the symbol): it uses the
// device address hipGetSymbolAddress returns for the current device, kept
// per device in g_scratch.  Calls that overlap on one GPU share the array
// and race; each call overwrites the records it reads back.
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int64_t kScratchParticles = (int64_t)(kScratchBytes / sizeof(fsim_gfx_particle));
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxScratchDevices = 64;
static void* g_scratch[kMaxScratchDevices];
static std::mutex g_scratch_mu;

// Resolves this device's scratch address once (with the device current).
int scratch_init(fsim_ctx* ctx) {
    const int dev = ctx->device;
    if (dev < 0 || dev >= kMaxScratchDevices) return set_err(ctx, FSIM_ERR_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[dev]) return FSIM_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_err(ctx, FSIM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = p;
    return FSIM_OK;
}

int64_t scratch_particles() { return kScratchParticles; }

struct GfxArgs {
    const float* pos[3];
    const float* vel[3];
    const int32_t* pid;
    const float* density;
    fsim_gfx_particle* out;
    int64_t n;
    int nx, ny, nz;
    double cd[3], cdinv[3];   // per axis (z differs in 2D)
};

// getCellsAround's order of the 8 cells (macGrid.cpp:192-201)
__constant__ int kGfxCorner[8][3] = {{1, 1, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0},
                                     {1, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 0, 0}};

__global__ void __launch_bounds__(kBlock) k_gfx_snapshot(GfxArgs A) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const double p[3] = {(double)A.pos[0][i], (double)A.pos[1][i], (double)A.pos[2][i]};
    const double v[3] = {(double)A.vel[0][i], (double)A.vel[1][i], (double)A.vel[2][i]};
    int base[3];
#pragma unroll
    for (int d = 0; d < 3; d++) base[d] = (int)(p[d] * A.cdinv[d] - 0.5);
    float dens = 0.f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int c[3] = {base[0] + kGfxCorner[k][0], base[1] + kGfxCorner[k][1], base[2] + kGfxCorner[k][2]};
        // particles stay inside the clamp box, so the 8 cells exist; any other position skips the missing ones
        if ((unsigned)c[0] >= (unsigned)A.nx || (unsigned)c[1] >= (unsigned)A.ny || (unsigned)c[2] >= (unsigned)A.nz)
            continue;
        double w = 1.0;
#pragma unroll
        for (int d = 0; d < 3; d++) w *= 1.0 - fabs(((c[d] + 0.5) * A.cd[d] - p[d]) * A.cdinv[d]);
        const float rho = A.density[((int64_t)c[0] * A.ny + c[1]) * A.nz + c[2]];
        dens = (float)((double)dens + w * (double)rho);
    }
    fsim_gfx_particle r;
    r.pos[0] = A.pos[0][i];
    r.pos[1] = A.pos[1][i];
    r.pos[2] = A.pos[2][i];
    r.v = (float)sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);   // glm::length
    r.density = dens;
    A.out[A.pid[i]] = r;
}

static GfxArgs gfx_args(fsim_ctx* ctx, fsim_gfx_particle* dev_out) {
    GfxArgs A{};
    for (int a = 0; a < 3; a++) {
        A.pos[a] = ctx->P.pos[a];
        A.vel[a] = ctx->P.vel[a];
    }
    A.pid = ctx->P.pid;
    A.density = ctx->density;
    A.out = dev_out;
    A.n = ctx->P.n;
    A.nx = ctx->g.nx;
    A.ny = ctx->g.ny;
    A.nz = ctx->g.nz;
    for (int d = 0; d < 3; d++) {
        A.cd[d] = d == 2 ? ctx->g.cdz : ctx->g.cd;
        A.cdinv[d] = d == 2 ? ctx->g.cdzinv : ctx->g.cdinv;
    }
    return A;
}

static int gfx_run(fsim_ctx* ctx, fsim_gfx_particle* staging, fsim_gfx_particle* host_out) {
    const int64_t n = ctx->P.n;
    if (n == 0) return FSIM_OK;
    hipLaunchKernelGGL(k_gfx_snapshot, dim3(nblocks(n)), dim3(kBlock), 0, ctx->stream, gfx_args(ctx, staging));
    FSIM_HIP(ctx, hipGetLastError());
    FSIM_HIP(ctx, hipMemcpyAsync(host_out, staging, sizeof(fsim_gfx_particle) * (size_t)n, hipMemcpyDeviceToHost,
                                 ctx->stream));
    FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FSIM_OK;
}

// The context's staging buffer, grown to the particle count on demand.
int gfx_snapshot(fsim_ctx* ctx, fsim_gfx_particle* host_out) {
    const int64_t n = ctx->P.n;
    if (n > ctx->gfx_cap) {
        fsim_gfx_particle* buf = nullptr;
        const int64_t cap = n + n / 8;
        if (hipMalloc(reinterpret_cast<void**>(&buf), sizeof(fsim_gfx_particle) * (size_t)cap) != hipSuccess)
            return set_err(ctx, FSIM_ERR_OOM, "gfx staging allocation failed");
        FSIM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        (void)hipFree(ctx->gfx_staging);
        ctx->gfx_staging = buf;
        ctx->gfx_cap = cap;
    }
    return gfx_run(ctx, ctx->gfx_staging, host_out);
}

// The same kernel and copy with the staging in this GPU's scratch array:
// no allocation, at most kScratchParticles particles.
int gfx_snapshot_noalloc(fsim_ctx* ctx, fsim_gfx_particle* host_out) {
    if (ctx->P.n > kScratchParticles)
        return set_err(ctx, FSIM_ERR_INVALID, "fsim_gfx_snapshot_noalloc: more than " +
                                                  std::to_string(kScratchParticles) +
                                                  " particles (64 MiB scratch); use fsim_gfx_snapshot");
    void* p = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        p = g_scratch[ctx->device];
    }
    if (!p) return set_err(ctx, FSIM_ERR_INVALID, "scratch address of this device was not resolved");
    return gfx_run(ctx, static_cast<fsim_gfx_particle*>(p), host_out);
}

void gfx_free(fsim_ctx* ctx) {
    if (ctx->gfx_staging) (void)hipFree(ctx->gfx_staging);
    ctx->gfx_staging = nullptr;
    ctx->gfx_cap = 0;
}

}  // namespace fsim
