// Leapfrog (kick-drift-kick) updates and the device-side bounding box of a resident particle system.
#include "common.hpp"

namespace ryo
{

void cubicBox(const float* mn, const float* mx, float* limits);

// particle i: v += c_kick * a; then, if c_drift != 0, x += c_drift * v; periodic axes wrap into [lo, lo + L)
__device__ inline void kickDriftOne(size_t i, float* __restrict__ x, float* __restrict__ y, float* __restrict__ z,
                                    float* __restrict__ vx, float* __restrict__ vy, float* __restrict__ vz,
                                    const float* __restrict__ ax, const float* __restrict__ ay,
                                    const float* __restrict__ az, float cKick, float cDrift, int periodic, float3 lo,
                                    float3 L)
{
    {
        float u = vx[i] + cKick * ax[i], v = vy[i] + cKick * ay[i], w = vz[i] + cKick * az[i];
        vx[i] = u, vy[i] = v, vz[i] = w;
        if (cDrift != 0.0f)
        {
            float px = x[i] + cDrift * u, py = y[i] + cDrift * v, pz = z[i] + cDrift * w;
            if (periodic)
            {
                px -= L.x * floorf((px - lo.x) / L.x);
                py -= L.y * floorf((py - lo.y) / L.y);
                pz -= L.z * floorf((pz - lo.z) / L.z);
                // rounding can land exactly on the upper bound
                if (px >= lo.x + L.x) px = lo.x;
                if (py >= lo.y + L.y) py = lo.y;
                if (pz >= lo.z + L.z) pz = lo.z;
            }
            x[i] = px, y[i] = py, z[i] = pz;
        }
    }
}

__global__ __launch_bounds__(256) void kickDriftKernel(float* __restrict__ x, float* __restrict__ y,
                                                      float* __restrict__ z, float* __restrict__ vx,
                                                      float* __restrict__ vy, float* __restrict__ vz,
                                                      const float* __restrict__ ax, const float* __restrict__ ay,
                                                      const float* __restrict__ az, size_t n, float cKick,
                                                      float cDrift, int periodic, float3 lo, float3 L)
{
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        kickDriftOne(i, x, y, z, vx, vy, vz, ax, ay, az, cKick, cDrift, periodic, lo, L);
}

// the same update for systems packed one after the other: system s owns particles [P[s], P[s+1]) and updates with
// its own coefficients coef[s] = {c_kick, c_drift} and, when wrap is given, its own periodic box wrap[s]
__global__ __launch_bounds__(256) void packedKickDriftKernel(float* __restrict__ x, float* __restrict__ y,
                                                            float* __restrict__ z, float* __restrict__ vx,
                                                            float* __restrict__ vy, float* __restrict__ vz,
                                                            const float* __restrict__ ax, const float* __restrict__ ay,
                                                            const float* __restrict__ az, const uint32_t* __restrict__ P,
                                                            int S, const float2* __restrict__ coef,
                                                            const PackedWrap* __restrict__ wrap)
{
    const size_t n = P[S];
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    {
        const int s    = segmentOf(P, S, uint32_t(i));
        const float2 c = coef[s];
        int periodic   = 0;
        float3 lo = make_float3(0, 0, 0), L = make_float3(1, 1, 1);
        if (wrap)
        {
            const PackedWrap w = wrap[s];
            periodic = w.periodic;
            lo = make_float3(w.lo[0], w.lo[1], w.lo[2]), L = make_float3(w.L[0], w.L[1], w.L[2]);
        }
        kickDriftOne(i, x, y, z, vx, vy, vz, ax, ay, az, c.x, c.y, periodic, lo, L);
    }
}

// per-system min/max of x, y, z of packed systems: block s reduces system s into out[s][6]
__global__ __launch_bounds__(256) void packedBoxKernel(const float* __restrict__ x, const float* __restrict__ y,
                                                      const float* __restrict__ z, const uint32_t* __restrict__ P,
                                                      float* __restrict__ out)
{
    const int sys = blockIdx.x;
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = P[sys] + threadIdx.x; i < P[sys + 1]; i += blockDim.x)
    {
        float p[3] = {x[i], y[i], z[i]};
        for (int d = 0; d < 3; ++d)
            mn[d] = fminf(mn[d], p[d]), mx[d] = fmaxf(mx[d], p[d]);
    }
    __shared__ float s[6][4];
    for (int d = 0; d < 3; ++d)
    {
        mn[d] = waveMin(mn[d]);
        mx[d] = waveMax(mx[d]);
    }
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int d = 0; d < 3; ++d)
            s[d][w] = mn[d], s[3 + d][w] = mx[d];
    __syncthreads();
    if (threadIdx.x < 6)
    {
        float v = s[threadIdx.x][0];
        for (int k = 1; k < 4; ++k)
            v = threadIdx.x < 3 ? fminf(v, s[threadIdx.x][k]) : fmaxf(v, s[threadIdx.x][k]);
        out[sys * 6 + threadIdx.x] = v;
    }
}

// per-block min/max of x,y,z -> out[block][6]
__global__ __launch_bounds__(256) void bboxKernel(const float* __restrict__ x, const float* __restrict__ y,
                                                 const float* __restrict__ z, size_t n, float* __restrict__ out)
{
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    {
        float p[3] = {x[i], y[i], z[i]};
        for (int d = 0; d < 3; ++d)
            mn[d] = fminf(mn[d], p[d]), mx[d] = fmaxf(mx[d], p[d]);
    }
    __shared__ float s[6][4];
    for (int d = 0; d < 3; ++d)
    {
        mn[d] = waveMin(mn[d]);
        mx[d] = waveMax(mx[d]);
    }
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int d = 0; d < 3; ++d)
            s[d][w] = mn[d], s[3 + d][w] = mx[d];
    __syncthreads();
    if (threadIdx.x < 6)
    {
        float v = s[threadIdx.x][0];
        for (int k = 1; k < 4; ++k)
            v = threadIdx.x < 3 ? fminf(v, s[threadIdx.x][k]) : fmaxf(v, s[threadIdx.x][k]);
        out[blockIdx.x * 6 + threadIdx.x] = v;
    }
}

int kickDrift(Context* ctx, float* x, float* y, float* z, float* vx, float* vy, float* vz, const float* ax,
              const float* ay, const float* az, size_t n, float cKick, float cDrift, int periodic, const float* limits)
{
    if (n == 0) return RYO_OK;
    float3 lo = make_float3(limits[0], limits[2], limits[4]);
    float3 L  = make_float3(limits[1] - limits[0], limits[3] - limits[2], limits[5] - limits[4]);
    unsigned blocks = std::min(iceil(n, 256), unsigned(ctx->numCUs * 8));
    kickDriftKernel<<<blocks, 256, 0, ctx->stream>>>(x, y, z, vx, vy, vz, ax, ay, az, n, cKick, cDrift, periodic, lo,
                                                     L);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

// cubic box around all particles, padded by a relative 1e-4 (host float[6])
int fitBox(Context* ctx, const float* x, const float* y, const float* z, size_t n, float* limits)
{
    if (n == 0) return setError(ctx, RYO_ERR_ARG, "bbox: no particles");
    unsigned blocks = std::min(iceil(n, 256), unsigned(ctx->numCUs * 2));
    auto* part      = static_cast<float*>(ctx->getScratch(kSlotMisc, blocks * 6 * sizeof(float)));
    if (!part) return RYO_ERR_HIP;
    bboxKernel<<<blocks, 256, 0, ctx->stream>>>(x, y, z, n, part);
    RYO_LAUNCH_CHECK(ctx);
    std::vector<float> h(blocks * 6);
    RYO_HIP(ctx, hipMemcpyAsync(h.data(), part, h.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    RYO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (unsigned b = 0; b < blocks; ++b)
        for (int d = 0; d < 3; ++d)
            mn[d] = std::min(mn[d], h[b * 6 + d]), mx[d] = std::max(mx[d], h[b * 6 + 3 + d]);
    cubicBox(mn, mx, limits);
    return RYO_OK;
}

// the cubic box of an extent, padded by a relative 1e-4 (host float[6]); all three lengths are the same float
void cubicBox(const float* mn, const float* mx, float* limits)
{
    float ext = 0, c[3];
    for (int d = 0; d < 3; ++d)
    {
        ext  = std::max(ext, mx[d] - mn[d]);
        c[d] = 0.5f * (mx[d] + mn[d]);
    }
    float half = 0.5f * ext * (1.0f + 1e-4f) + 1e-30f;
    for (int d = 0; d < 3; ++d)
        limits[2 * d] = c[d] - half, limits[2 * d + 1] = c[d] + half;
    // a cubic box in float: make all three lengths identical
    float Lx = limits[1] - limits[0];
    for (int d = 1; d < 3; ++d)
        limits[2 * d + 1] = limits[2 * d] + Lx;
}

// v += coef[s].x * a, then x += coef[s].y * v for the packed systems of the offset table P (device, S + 1 entries);
// wrap (device, S entries, may be null): the periodic systems' boxes, whose positions wrap as a resident system's
int packedKickDrift(Context* ctx, float* x, float* y, float* z, float* vx, float* vy, float* vz, const float* ax,
                    const float* ay, const float* az, size_t n, const uint32_t* P, int S, const float2* coef,
                    const PackedWrap* wrap)
{
    if (n == 0 || S < 1) return RYO_OK;
    unsigned blocks = std::min(iceil(n, 256), unsigned(ctx->numCUs * 8));
    packedKickDriftKernel<<<blocks, 256, 0, ctx->stream>>>(x, y, z, vx, vy, vz, ax, ay, az, P, S, coef, wrap);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

// per-system extents of packed systems into out (device, 6 S floats)
int packedBoxes(Context* ctx, const float* x, const float* y, const float* z, const uint32_t* P, int S, float* out)
{
    if (S < 1) return RYO_OK;
    packedBoxKernel<<<S, 256, 0, ctx->stream>>>(x, y, z, P, out);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

} // namespace ryo
