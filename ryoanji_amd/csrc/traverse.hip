// Barnes-Hut traversal for wave64 (gfx950): one wavefront per target group, one target per lane.
//
// Semantics follow ryoanji's traversal (nbody/traversal_gpu.cu:170-570; CPU traversal_cpu.hpp:67-230):
//   * the target group is described by the bounding box of its (valid) targets;
//   * a source cell is "close" when the squared distance from its expansion center to the box is below its
//     squared MAC radius (applyMAC, traversal_gpu.cu:77-84 / macs.hpp:93-101);
//   * far cells apply their quadrupole (M2P, cartesian_qpole.hpp:152-176) to every target, close leaves
//     apply their particles (P2P with the softening clamp, kernel.hpp:99-120), close internal cells are
//     opened; periodic images repeat the walk with shifted targets.
// The walk itself is designed for CDNA: the 64 lanes evaluate 8 sibling groups (64 cells) per step, opened
// cells are pushed to a per-wave LIFO in LDS (the level-ordered LIFO holds at most 64 entries per tree
// level, so 21 x 64 entries can never overflow), and the interactions are applied right away with the
// source index uniform across the wave, so the source data of a cell or body is fetched once per wave
// with scalar loads and broadcast to all 64 targets.
#include "common.hpp"

#include <cstring>
#include <mutex>

namespace ryo
{

constexpr int kTravWaves     = 4;       // default waves per block (2, 4 or 8: ryo_ctx_set_traversal_waves)
constexpr int kMaxStack      = 22 * 64; // LIFO entries per wave that can never overflow (>= 21 levels x 64)
constexpr int kDefaultStack  = kMaxStack; // default LIFO entries per wave (a smaller one may take the retry path)
constexpr int kWavesPerCU    = 24;        // default resident waves per CU: 6 per SIMD (the VGPR budget below)
constexpr int kXcds          = 8;       // MI355X: 8 XCDs, workgroups dispatched round-robin over them

// device control words (ctx->dStats, 32 x u64): u32 [0..7] per-XCD group counters, u64 [8..14] statistics
// {sumP2P, maxP2P, sumM2P, maxM2P, maxStack, failed, corrupt}, u64 [16] FP64 potential sum, u32 [40..41]
// softening range {min, max} as order-preserving integers
constexpr int kCtlHRange = 40;

struct TravArgs
{
    int numGroups;
    const uint32_t* gStart;
    const uint32_t* gEnd;
    int initNode;
    const float *xt, *yt, *zt, *mt, *ht;
    const float *xs, *ys, *zs, *ms, *hs;
    const int32_t* child;
    const int32_t* i2l;
    const uint32_t* layout;
    const float4* centers;
    const float4* mp; // two float4 per node: {M, qxx, qxy, qxz}, {qyy, qyz, qzz, trace}
    float G;
    int numShells;
    float boxL[3];
    float *p, *ax, *ay, *az;
    double* groupPot;
    unsigned* work;            // kXcds group counters
    const unsigned* hRange;    // softening range of targets and sources (order-preserving integers)
    unsigned long long* stats;
    uint32_t* failed;
    const uint32_t* groupList; // when set, work item i is group groupList[i] (a retry of overflowed groups)
    int stackCap;
    size_t hCount; // targets and sources are h[0..hCount) (resident systems), 0 when unknown
    int ranges;    // contiguous group ranges, one per XCD (1: a single global counter)
    bool mdq;      // multipoles are CartesianMDQpole (12 floats per node: the quadrupole's 8, then px, py, pz, 0)
    const uint32_t* abort; // device, may be null: nonzero when the tree was not built (traverse nothing)
};

struct Acc
{
    float p, x, y, z;
};

// the source tree and source particles a walk reads (one system's slices in a batch)
struct TreeRef
{
    const int32_t* child;
    const int32_t* i2l;
    const uint32_t* layout;
    const float4* centers;
    const float4* mp;
    const float *xs, *ys, *zs, *ms, *hs;
};

// v_rsq_f32 (about 1 ulp) without the denormal rescaling that rsqrtf adds; the arguments are squared distances
// clamped by the softening, never denormal for the systems served here
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }

// quadrupole M2P (cartesian_qpole.hpp:152-176), with the products folded into fused multiply-adds
__device__ __forceinline__ void m2p(Acc& a, float tx, float ty, float tz, float4 c, float4 q0, float4 q1)
{
    float rx = tx - c.x, ry = ty - c.y, rz = tz - c.z;
    float r2  = rx * rx + ry * ry + rz * rz;
    float r1  = rsq(r2);
    float ri2 = r1 * r1;
    float r5  = ri2 * ri2 * r1;
    float M = q0.x, qxx = q0.y, qxy = q0.z, qxz = q0.w, qyy = q1.x, qyz = q1.y, qzz = q1.z;
    float Qrx  = rx * qxx + ry * qxy + rz * qxz;
    float Qry  = rx * qxy + ry * qyy + rz * qyz;
    float Qrz  = rx * qxz + ry * qyz + rz * qzz;
    float rQr  = rx * Qrx + ry * Qry + rz * Qrz;
    float u    = M * r1;
    float t    = rQr * r5;
    float mono = __builtin_fmaf(-2.5f, t, -u) * ri2;
    a.p = __builtin_fmaf(-0.5f, t, a.p - u);
    a.x = __builtin_fmaf(r5, Qrx, __builtin_fmaf(mono, rx, a.x));
    a.y = __builtin_fmaf(r5, Qry, __builtin_fmaf(mono, ry, a.y));
    a.z = __builtin_fmaf(r5, Qrz, __builtin_fmaf(mono, rz, a.z));
}

// M2P of a monopole-dipole-quadrupole (CartesianMDQpole, cartesian_qpole.hpp:318-344): the quadrupole M2P plus
// the dipole p about the expansion center, potential -(p.r) / r^3 and field (p r^2 - 3 (p.r) r) / r^5
__device__ __forceinline__ void m2pMDQ(Acc& a, float tx, float ty, float tz, float4 c, float4 q0, float4 q1,
                                       float4 d)
{
    float rx = tx - c.x, ry = ty - c.y, rz = tz - c.z;
    float r2  = rx * rx + ry * ry + rz * rz;
    float r1  = rsq(r2);
    float ri2 = r1 * r1;
    float r5  = ri2 * ri2 * r1;
    float M = q0.x, qxx = q0.y, qxy = q0.z, qxz = q0.w, qyy = q1.x, qyz = q1.y, qzz = q1.z;
    float Qrx  = rx * qxx + ry * qxy + rz * qxz;
    float Qry  = rx * qxy + ry * qyy + rz * qyz;
    float Qrz  = rx * qxz + ry * qyz + rz * qzz;
    float rQr  = rx * Qrx + ry * Qry + rz * Qrz;
    float PR   = rx * d.x + ry * d.y + rz * d.z;
    float mono = (-2.5f * rQr * r5 - M * r1) * ri2;
    float p3   = 3.0f * PR;
    a.p += (-M - PR * ri2) * r1 - 0.5f * r5 * rQr;
    a.x += r5 * Qrx + mono * rx + r5 * (d.x * r2 - p3 * rx);
    a.y += r5 * Qry + mono * ry + r5 * (d.y * r2 - p3 * ry);
    a.z += r5 * Qrz + mono * rz + r5 * (d.z * r2 - p3 * rz);
}

// softened P2P (kernel.hpp:99-120). UniformH: every target and source has the same softening h, so
// hij2 = (h + h)^2 is the value the general path computes for each pair.
template<bool UniformH>
__device__ __forceinline__ void p2p(Acc& a, float tx, float ty, float tz, float th, float hij2u, float sx, float sy,
                                    float sz, float sm, float sh)
{
    float dx = sx - tx, dy = sy - ty, dz = sz - tz;
    float r2 = dx * dx + dy * dy + dz * dz;
    float hij2;
    if constexpr (UniformH)
        hij2 = hij2u;
    else
    {
        float hij = th + sh;
        hij2      = hij * hij;
    }
    float r2eff  = fmaxf(r2, hij2); // the softening clamp (R2 < h_ij^2 ? h_ij^2 : R2) for non-NaN inputs
    float invR   = rsq(r2eff);
    float invR3m = sm * invR * (invR * invR);
    a.p -= invR3m * r2;
    a.x += dx * invR3m;
    a.y += dy * invR3m;
    a.z += dz * invR3m;
}

// MAC as the reference evaluates it: positive part of |tc - com| - ts per axis, squared norm (right fold),
// compared with mac^2. Unfused operations so that equal inputs give the CPU's decision.
__device__ __forceinline__ bool macClose(float4 c, float tcx, float tcy, float tcz, float tsx, float tsy, float tsz)
{
#pragma clang fp contract(off)
    float dx = fabsf(tcx - c.x) - tsx;
    float dy = fabsf(tcy - c.y) - tsy;
    float dz = fabsf(tcz - c.z) - tsz;
    dx = (dx + fabsf(dx)) * 0.5f;
    dy = (dy + fabsf(dy)) * 0.5f;
    dz = (dz + fabsf(dz)) * 0.5f;
    float R2 = dx * dx + (dy * dy + dz * dz);
    return R2 < fabsf(c.w);
}

struct WalkOut
{
    unsigned m2p, p2p, maxStack;
    bool overflow, corrupt;
};

// Wave-uniform loads through the constant address space: the source data of one cell or body is the same
// for all 64 lanes, so it is fetched with scalar memory instructions into SGPRs (read-only during the kernel).
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
typedef float f8u __attribute__((ext_vector_type(8), aligned(16)));
__device__ __forceinline__ f4u uload4(const float* p) { return *((const __attribute__((address_space(4))) f4u*)p); }
__device__ __forceinline__ f8u uload8(const float* p) { return *((const __attribute__((address_space(4))) f8u*)p); }
__device__ __forceinline__ float uload1(const float* p)
{
    return *((const __attribute__((address_space(4))) float*)p);
}

struct Target
{
    float x, y, z, h, hij2u;
};

// apply all bodies of leaf ranges flagged in `dmask` (lane -> [rb, re)) to the wave's targets
template<bool UH>
__device__ __forceinline__ unsigned applyLeaves(const TreeRef& a, uint64_t dmask, uint32_t rb, uint32_t re, Acc& acc,
                                                const Target& t)
{
    unsigned cnt = 0;
    while (dmask)
    {
        int l = __ffsll((unsigned long long)dmask) - 1;
        dmask &= dmask - 1;
        uint32_t b = readlane(rb, l), e = readlane(re, l);
        cnt += e - b;
        uint32_t j = b;
        // four bodies per scalar load of each array (dword-aligned 16-byte scalar loads)
        for (; j + 4 <= e; j += 4)
        {
            f4u X = uload4(a.xs + j), Y = uload4(a.ys + j), Z = uload4(a.zs + j), M = uload4(a.ms + j);
            f4u H = UH ? f4u{0.f, 0.f, 0.f, 0.f} : uload4(a.hs + j);
            p2p<UH>(acc, t.x, t.y, t.z, t.h, t.hij2u, X.x, Y.x, Z.x, M.x, H.x);
            p2p<UH>(acc, t.x, t.y, t.z, t.h, t.hij2u, X.y, Y.y, Z.y, M.y, H.y);
            p2p<UH>(acc, t.x, t.y, t.z, t.h, t.hij2u, X.z, Y.z, Z.z, M.z, H.z);
            p2p<UH>(acc, t.x, t.y, t.z, t.h, t.hij2u, X.w, Y.w, Z.w, M.w, H.w);
        }
        for (; j < e; ++j)
            p2p<UH>(acc, t.x, t.y, t.z, t.h, t.hij2u, uload1(a.xs + j), uload1(a.ys + j), uload1(a.zs + j),
                    uload1(a.ms + j), UH ? 0.f : uload1(a.hs + j));
    }
    return cnt;
}

// apply the multipoles of the cells flagged in `amask` (lane -> cell) in ascending lane order
template<bool MDQ>
__device__ __forceinline__ unsigned applyCells(const TreeRef& a, uint64_t amask, int cell, Acc& acc, const Target& t)
{
    unsigned cnt = __popcll(amask);
    const float* cbase = reinterpret_cast<const float*>(a.centers);
    const float* qbase = reinterpret_cast<const float*>(a.mp);
    constexpr int stride = MDQ ? 12 : 8;
    while (amask)
    {
        int l = __ffsll((unsigned long long)amask) - 1;
        amask &= amask - 1;
        int c = readlane(cell, l);
        f4u C = uload4(cbase + 4 * c);
        f8u Q = uload8(qbase + stride * c);
        if constexpr (MDQ)
        {
            f4u D = uload4(qbase + stride * c + 8);
            m2pMDQ(acc, t.x, t.y, t.z, make_float4(C.x, C.y, C.z, C.w), make_float4(Q.s0, Q.s1, Q.s2, Q.s3),
                   make_float4(Q.s4, Q.s5, Q.s6, Q.s7), make_float4(D.x, D.y, D.z, D.w));
        }
        else
            m2p(acc, t.x, t.y, t.z, make_float4(C.x, C.y, C.z, C.w), make_float4(Q.s0, Q.s1, Q.s2, Q.s3),
                make_float4(Q.s4, Q.s5, Q.s6, Q.s7));
    }
    return cnt;
}

template<bool UH, bool MDQ>
__device__ WalkOut walk(const TravArgs& ta, const TreeRef& a, int* stack, const Target& t, float tcx, float tcy,
                        float tcz, float tsx, float tsy, float tsz, Acc& acc)
{
    const int lane    = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    WalkOut o{0, 0, 0, false, false};
    int sp = 0;

    const int rootChild = a.child[0];
    if (ta.initNode == 0 || rootChild == 0)
    {
        float4 c   = a.centers[0];
        bool close = macClose(c, tcx, tcy, tcz, tsx, tsy, tsz);
        if (!close)
        {
            if (c.w != 0.0f) o.m2p += applyCells<MDQ>(a, 1ull, 0, acc, t);
            return o;
        }
        if (rootChild == 0)
        {
            o.p2p += applyLeaves<UH>(a, 1ull, a.layout[0], a.layout[1], acc, t);
            return o;
        }
    }
    if (lane == 0) stack[0] = rootChild;
    sp = 1;

    // a valid tree is walked in at most numNodes/8 + 1 steps; the cap only guards against a corrupt tree
    int steps = 0;
    while (sp > 0)
    {
        if (++steps > (1 << 24))
        {
            o.overflow = true;
            o.corrupt  = true;
            return o;
        }
        const int k   = sp < 8 ? sp : 8;
        const int e   = lane >> 3;
        const bool ok = e < k;
        __builtin_amdgcn_wave_barrier();
        int cell = ok ? stack[sp - k + e] + (lane & 7) : 0;
        sp -= k;

        float4 c    = a.centers[cell];
        int ch      = a.child[cell];
        bool close  = ok && macClose(c, tcx, tcy, tcz, tsx, tsy, tsz);
        bool isNode = ch != 0;
        uint32_t rb = 0, re = 0;
        if (close && !isNode)
        {
            int li = a.i2l[cell];
            rb = a.layout[li], re = a.layout[li + 1];
        }
        // a leaf without local particles but with sources elsewhere (remote) is always approximated
        bool remote = close && !isNode && rb == re && c.w != 0.0f;
        bool split  = close && isNode;
        bool approx = ok && ((!close && c.w != 0.0f) || remote);
        bool direct = close && !isNode && rb != re;

        uint64_t sm = __ballot(split);
        int cnt     = __popcll(sm);
        if (sp + cnt > ta.stackCap)
        {
            o.overflow = true;
            o.maxStack = max(o.maxStack, unsigned(sp + cnt));
            return o;
        }
        if (split) stack[sp + __popcll(sm & lt)] = ch;
        sp += cnt;
        o.maxStack = max(o.maxStack, unsigned(sp));

        o.m2p += applyCells<MDQ>(a, __ballot(approx), cell, acc, t);
        o.p2p += applyLeaves<UH>(a, __ballot(direct), rb, re, acc, t);
    }
    return o;
}

__device__ __forceinline__ unsigned floatToOrdered(float f)
{
    unsigned b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float orderedToFloat(unsigned o)
{
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

// min and max softening over targets and sources (order-preserving integers; out = {~0u, 0} on entry); the
// arrays are read as float4 where 16-byte aligned, the rest element by element
__device__ inline void hRangeOf(const float* __restrict__ h, size_t n, unsigned& lo, unsigned& hi)
{
    const size_t stride = size_t(gridDim.x) * blockDim.x, t0 = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t head = (reinterpret_cast<uintptr_t>(h) & 15u) ? (16u - (reinterpret_cast<uintptr_t>(h) & 15u)) / 4u : 0u;
    head = head < n ? head : n;
    for (size_t i = t0; i < head; i += stride)
    {
        unsigned o = floatToOrdered(h[i]);
        lo = min(lo, o), hi = max(hi, o);
    }
    const float4* h4 = reinterpret_cast<const float4*>(h + head);
    const size_t n4  = (n - head) / 4;
    for (size_t i = t0; i < n4; i += stride)
    {
        float4 v = h4[i];
        unsigned a = floatToOrdered(v.x), b = floatToOrdered(v.y), c = floatToOrdered(v.z), d = floatToOrdered(v.w);
        lo = min(lo, min(min(a, b), min(c, d)));
        hi = max(hi, max(max(a, b), max(c, d)));
    }
    for (size_t i = head + 4 * n4 + t0; i < n; i += stride)
    {
        unsigned o = floatToOrdered(h[i]);
        lo = min(lo, o), hi = max(hi, o);
    }
}

__global__ __launch_bounds__(256) void hRangeKernel(const float* __restrict__ h1, size_t n1,
                                                    const float* __restrict__ h2, size_t n2, unsigned* out)
{
    unsigned lo = ~0u, hi = 0;
    hRangeOf(h1, n1, lo, hi);
    if (h2) hRangeOf(h2, n2, lo, hi);
    for (int off = 32; off > 0; off >>= 1)
    {
        lo = min(lo, unsigned(__shfl_xor(int(lo), off)));
        hi = max(hi, unsigned(__shfl_xor(int(hi), off)));
    }
    // one atomic pair per block: thousands of atomics on the same two words serialize in L2
    __shared__ unsigned wlo[4], whi[4];
    if ((threadIdx.x & 63) == 0) wlo[threadIdx.x >> 6] = lo, whi[threadIdx.x >> 6] = hi;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        for (int w = 1; w < int(blockDim.x >> 6); ++w)
            lo = min(lo, wlo[w]), hi = max(hi, whi[w]);
        atomicMin(out, lo);
        atomicMax(out + 1, hi);
    }
}

// reset the control words: group counters, statistics and potential sum to 0, softening range to {~0u, 0}
__global__ void travInitKernel(unsigned* ctl)
{
    int i = threadIdx.x;
    if (i < 64) ctl[i] = (i == kCtlHRange) ? ~0u : 0u;
}

// smallest and largest softening of h[0..n) (one read-back; synchronizes the stream)
int softeningRange(Context* ctx, const float* h, size_t n, float* hmin, float* hmax)
{
    if (n == 0) return setError(ctx, RYO_ERR_ARG, "softening range: no particles");
    auto* w = static_cast<unsigned*>(ctx->getScratch(kSlotSoft, 64 * sizeof(unsigned)));
    if (!w) return RYO_ERR_HIP;
    hipStream_t s = ctx->stream;
    travInitKernel<<<1, 64, 0, s>>>(w); // w[kCtlHRange..+1] = {~0u, 0}
    hRangeKernel<<<std::max(1u, std::min(iceil(n, 4096), unsigned(ctx->numCUs * 2))), 256, 0, s>>>(
        h, n, nullptr, 0, w + kCtlHRange);
    RYO_LAUNCH_CHECK(ctx);
    unsigned o[2];
    RYO_HIP(ctx, hipMemcpyAsync(o, w + kCtlHRange, sizeof(o), hipMemcpyDeviceToHost, s));
    RYO_HIP(ctx, hipStreamSynchronize(s));
    // the order-preserving map back to float (the inverse of floatToOrdered)
    for (int k = 0; k < 2; ++k)
    {
        unsigned u = (o[k] & 0x80000000u) ? (o[k] ^ 0x80000000u) : ~o[k];
        float f;
        std::memcpy(&f, &u, sizeof(f));
        (k == 0 ? *hmin : *hmax) = f;
    }
    return RYO_OK;
}

// per-group source tree of a batch of systems: group g belongs to system gSys[g]; the tree arrays and the source
// particles are the system's slices, the targets are global particle indices
struct TravBatch
{
    BatchView bv;
    const int32_t* gSys;
    const float* G;           // per system
    const uint32_t* abort;    // nonzero: the trees were not built (arrays too small), traverse nothing
    const SysImages* images;  // per system (null: every system open)
};

// 64 * Waves threads; the second bound is the minimum number of waves per SIMD: 6 (<= 80 VGPRs; more would spill)
template<bool Batched, int Waves, bool MDQ = false>
__global__ __launch_bounds__(64 * Waves, kWavesPerCU / 4) void traverseKernelT(TravArgs a, TravBatch tb)
{
    extern __shared__ int stackMem[];
    const int w    = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    int* stack     = stackMem + w * a.stackCap;
    if (a.abort && *a.abort) return;
    if constexpr (Batched)
        if (tb.abort && *tb.abort) return;

    // uniform softening: one hij^2 for every pair (the same value the general path computes)
    const unsigned hlo = a.hRange[0], hhi = a.hRange[1];
    const bool uniformH = hlo == hhi;
    float hij = orderedToFloat(hlo);
    hij       = hij + hij;
    const float hij2u = hij * hij;

    // XCD-aware work distribution: groups are split into kXcds contiguous ranges and the blocks of an XCD
    // (blockIdx % kXcds under round-robin dispatch) take groups from their own range first, so the groups in
    // flight on one XCD are SFC neighbours sharing source cells in that XCD's L2; exhausted ranges are then
    // helped from in order
    const int R    = a.ranges;
    const int home = blockIdx.x % R;
    int src        = 0;
    for (;;)
    {
        unsigned g = ~0u;
        if (lane == 0)
        {
            while (src < R)
            {
                const int x       = (home + src) % R;
                const unsigned lo = unsigned((unsigned long long)a.numGroups * x / R);
                const unsigned hi = unsigned((unsigned long long)a.numGroups * (x + 1) / R);
                if (lo < hi)
                {
                    const unsigned i = atomicAdd(a.work + x, 1u);
                    if (i < hi - lo)
                    {
                        g = lo + i;
                        break;
                    }
                }
                ++src;
            }
        }
        // batched: lane 0 is the first active lane here, so the group index becomes a scalar value and the
        // group's system and that system's tree pointers are scalar loads and scalar registers (as vector values
        // the ten pointers spill); a single system keeps the broadcast, measured 0.25 % faster there
        if constexpr (Batched) g = unsigned(__builtin_amdgcn_readfirstlane(int(g)));
        else g = __shfl(g, 0);
        if (g == ~0u) return;
        if (a.groupList) g = a.groupList[g];

        TreeRef tr{a.child, a.i2l, a.layout, a.centers, a.mp, a.xs, a.ys, a.zs, a.ms, a.hs};
        float Gs = a.G;
        int S    = a.numShells;
        float bx = a.boxL[0], by = a.boxL[1], bz = a.boxL[2];
        if constexpr (Batched)
        {
            const int s    = tb.gSys[g];
            const size_t p = tb.bv.P[s], no = tb.bv.NO[s];
            tr = TreeRef{a.child + no, a.i2l + no, a.layout + tb.bv.LO[s], a.centers + no, a.mp + 2 * no,
                         a.xs + p, a.ys + p, a.zs + p, a.ms + p, a.hs + p};
            Gs = tb.G[s];
            if (tb.images)
            {
                const SysImages im = tb.images[s];
                S = im.shells, bx = im.L[0], by = im.L[1], bz = im.L[2];
            }
        }
        const uint32_t b = a.gStart[g], e = a.gEnd[g];
        const bool valid = b + lane < e;
        const uint32_t ti = valid ? b + lane : e - 1;
        Target t{a.xt[ti], a.yt[ti], a.zt[ti], a.ht[ti], hij2u};

        float mnx = waveMin(t.x), mny = waveMin(t.y), mnz = waveMin(t.z);
        float mxx = waveMax(t.x), mxy = waveMax(t.y), mxz = waveMax(t.z);
        float tcx = (mxx + mnx) * 0.5f, tcy = (mxy + mny) * 0.5f, tcz = (mxz + mnz) * 0.5f;
        float tsx = (mxx - mnx) * 0.5f, tsy = (mxy - mny) * 0.5f, tsz = (mxz - mnz) * 0.5f;

        Acc acc{0.f, 0.f, 0.f, 0.f};
        unsigned nM2P = 0, nP2P = 0, maxStack = 0;
        bool overflow = false, corrupt = false;
        for (int iz = -S; iz <= S && !overflow; ++iz)
            for (int iy = -S; iy <= S && !overflow; ++iy)
                for (int ix = -S; ix <= S && !overflow; ++ix)
                {
                    float sx = ix * bx, sy = iy * by, sz = iz * bz;
                    Target ts{t.x + sx, t.y + sy, t.z + sz, t.h, t.hij2u};
                    WalkOut o = uniformH
                                    ? walk<true, MDQ>(a, tr, stack, ts, tcx + sx, tcy + sy, tcz + sz, tsx, tsy, tsz,
                                                      acc)
                                    : walk<false, MDQ>(a, tr, stack, ts, tcx + sx, tcy + sy, tcz + sz, tsx, tsy, tsz,
                                                       acc);
                    nM2P += o.m2p;
                    nP2P += o.p2p;
                    maxStack = max(maxStack, o.maxStack);
                    overflow = o.overflow;
                    corrupt  = o.corrupt;
                }

        if (overflow)
        {
            if (lane == 0)
            {
                unsigned slot = atomicAdd(reinterpret_cast<unsigned*>(&a.stats[5]), 1u);
                a.failed[slot] = g;
                atomicMax(&a.stats[4], (unsigned long long)maxStack);
                if (corrupt) atomicAdd(&a.stats[6], 1ull);
            }
            continue;
        }

        double myPot = 0;
        if (valid)
        {
            float m = a.mt[ti];
            if (a.p) a.p[ti] += Gs * m * acc.p;
            a.ax[ti] += Gs * acc.x;
            a.ay[ti] += Gs * acc.y;
            a.az[ti] += Gs * acc.z;
            myPot = double(m) * double(acc.p);
        }
        myPot = waveSum(myPot);
        if (lane == 0)
        {
            a.groupPot[g] = myPot;
            unsigned long long nt = e - b;
            atomicAdd(&a.stats[0], nt * nP2P);
            atomicMax(&a.stats[1], (unsigned long long)nP2P);
            atomicAdd(&a.stats[2], nt * nM2P);
            atomicMax(&a.stats[3], (unsigned long long)nM2P);
            atomicMax(&a.stats[4], (unsigned long long)maxStack);
        }
    }
}

// waves per block of this context's traversal launches, resident blocks per CU, and the launch itself
int travWaves(const Context* ctx) { return ctx->travWaves == 2 || ctx->travWaves == 8 ? ctx->travWaves : kTravWaves; }

int travBlocks(const Context* ctx, size_t groups)
{
    const int W     = travWaves(ctx);
    const int perCU = ctx->travBlocksPerCU > 0 ? ctx->travBlocksPerCU : kWavesPerCU / W;
    return std::max(1, std::min(ctx->numCUs * perCU, int(iceil(groups, W))));
}

template<bool Batched>
void launchTraversal(const Context* ctx, size_t groups, int stackCap, hipStream_t s, const TravArgs& a,
                     const TravBatch& tb)
{
    const int W       = travWaves(ctx);
    const int blocks  = travBlocks(ctx, groups);
    const size_t lds  = size_t(W) * stackCap * sizeof(int);
    if constexpr (!Batched)
        if (a.mdq)
        {
            // monopole-dipole-quadrupole multipoles: the same block shapes and register budget (80 VGPRs, no
            // spills at 6 waves per SIMD)
            if (W == 2) traverseKernelT<false, 2, true><<<blocks, 128, lds, s>>>(a, tb);
            else if (W == 8) traverseKernelT<false, 8, true><<<blocks, 512, lds, s>>>(a, tb);
            else traverseKernelT<false, 4, true><<<blocks, 256, lds, s>>>(a, tb);
            return;
        }
    if (W == 2) traverseKernelT<Batched, 2><<<blocks, 128, lds, s>>>(a, tb);
    else if (W == 8) traverseKernelT<Batched, 8><<<blocks, 512, lds, s>>>(a, tb);
    else traverseKernelT<Batched, 4><<<blocks, 256, lds, s>>>(a, tb);
}

// fixed-order FP64 sum (Neumaier) of v[0..n): block b reduces the contiguous chunk b of `chunk` values into out[b]
// (one block with chunk = n gives the total); the chunking depends only on n, so the result is deterministic
__global__ __launch_bounds__(256) void sumKernel(const double* __restrict__ v, int n, double* __restrict__ out,
                                                 int chunk)
{
    __shared__ double ss[256], sc[256];
    double s = 0, c = 0;
    const int lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
    for (int i = lo + threadIdx.x; i < hi; i += 256)
    {
        double x = v[i], t = s + x;
        c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    ss[threadIdx.x] = s;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if (int(threadIdx.x) < o)
        {
            double x = ss[threadIdx.x + o], y = ss[threadIdx.x];
            double t = x + y;
            sc[threadIdx.x] += sc[threadIdx.x + o] + (fabs(y) >= fabs(x) ? (y - t) + x : (x - t) + y);
            ss[threadIdx.x] = t;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = ss[0] + sc[0];
}

// deterministic two-stage FP64 sum of v[0..n) into *out (partials: scratch of 256 doubles)
void sumDeterministic(hipStream_t s, const double* v, int n, double* partials, double* out)
{
    const int blocks = 256;
    const int chunk  = (n + blocks - 1) / blocks;
    sumKernel<<<blocks, 256, 0, s>>>(v, n, partials, std::max(chunk, 1));
    sumKernel<<<1, 256, 0, s>>>(partials, blocks, out, blocks);
}

__global__ void zeroKernelU(unsigned* a, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        a[i] = 0;
}

int traverse(Context* ctx, TravArgs a, double* totalPot)
{
    hipStream_t s = ctx->stream;
    for (int k = 0; k < 6; ++k)
        ctx->hStats[k] = 0;
    ctx->retriedGroups = 0;
    if (a.numGroups <= 0)
    {
        if (totalPot) *totalPot = 0;
        return RYO_OK;
    }
    // LIFO capacity per wave (LDS) and resident blocks per CU: the context's launch shape, optionally reduced
    // by the caller's stack capacity (the overflow path splits groups and retries them)
    const int ldsStack = ctx->travStack > 8 ? std::min(ctx->travStack, kMaxStack) : kDefaultStack;
    a.stackCap         = (a.stackCap > 8 && a.stackCap < ldsStack) ? a.stackCap : ldsStack;
    a.ranges           = ctx->travRanges > 0 ? std::min(ctx->travRanges, kXcds) : kXcds;

    unsigned* ctl             = reinterpret_cast<unsigned*>(ctx->dStats);
    unsigned long long* stats = ctx->dStats + 8;
    double* potSum            = reinterpret_cast<double*>(ctx->dStats + 16);
    a.work   = ctl;
    a.stats  = stats;
    a.hRange = ctl + kCtlHRange;
    travInitKernel<<<1, 64, 0, s>>>(ctl);
    // the softening range over all targets and sources selects the uniform-softening P2P; without the particle
    // count (hCount = 0) the range stays empty and the general P2P runs
    if (a.hCount > 0)
        hRangeKernel<<<std::max(1u, std::min(iceil(a.hCount, 4096), unsigned(ctx->numCUs * 2))), 256, 0, s>>>(a.ht, a.hCount,
                                                                            a.hs == a.ht ? nullptr : a.hs,
                                                                            a.hs == a.ht ? 0 : a.hCount,
                                                                            ctl + kCtlHRange);

    double pot              = 0;
    unsigned long long h[6] = {0, 0, 0, 0, 0, 0};
    std::vector<uint32_t> ns, ne;
    for (int pass = 0;; ++pass)
    {
        size_t G             = size_t(a.numGroups);
        const size_t partOff = (12 * G + 255) & ~size_t(255);
        auto* buf            = static_cast<char*>(ctx->getScratch(kSlotTraverse, partOff + 256 * sizeof(double)));
        if (!buf) return RYO_ERR_HIP;
        a.groupPot       = reinterpret_cast<double*>(buf);
        a.failed         = reinterpret_cast<uint32_t*>(buf + 8 * G);
        double* partials = reinterpret_cast<double*>(buf + partOff);
        if (pass > 0)
        {
            zeroKernelU<<<1, 64, 0, s>>>(ctl, kXcds);
            zeroKernelU<<<1, 64, 0, s>>>(reinterpret_cast<unsigned*>(stats + 5), 4);
        }
        zeroKernelU<<<std::min(iceil(2 * G, 256), 4096u), 256, 0, s>>>(reinterpret_cast<unsigned*>(a.groupPot),
                                                                       int(2 * G));
        if (ctx->timing && pass == 0) RYO_HIP(ctx, hipEventRecord(ctx->ev[0], s));
        launchTraversal<false>(ctx, G, ldsStack, s, a, TravBatch{});
        if (ctx->timing && pass == 0) RYO_HIP(ctx, hipEventRecord(ctx->ev[1], s));
        RYO_LAUNCH_CHECK(ctx);
        sumDeterministic(s, a.groupPot, a.numGroups, partials, potSum);
        double hp = 0;
        RYO_HIP(ctx, hipMemcpyAsync(h, stats, sizeof(h), hipMemcpyDeviceToHost, s));
        RYO_HIP(ctx, hipMemcpyAsync(&hp, potSum, sizeof(double), hipMemcpyDeviceToHost, s));
        RYO_HIP(ctx, hipStreamSynchronize(s));
        if (ctx->timing && pass == 0)
        {
            float ms = 0;
            RYO_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
            ctx->lastTraverseMs = ms;
        }
        pot += hp; // failed groups left their (zeroed) slot untouched
        unsigned numFailed = unsigned(h[5] & 0xffffffffu);
        if (numFailed == 0) break;
        unsigned long long corrupt = 0;
        RYO_HIP(ctx, hipMemcpy(&corrupt, stats + 6, sizeof(corrupt), hipMemcpyDeviceToHost));
        if (corrupt) return setError(ctx, RYO_ERR_TRAVERSAL, "traversal: step limit exceeded (corrupt tree?)");
        if (pass > 40) return setError(ctx, RYO_ERR_TRAVERSAL, "traversal: overflow not resolved by splitting");
        ctx->retriedGroups += numFailed;
        // splitFailedGroups: halve each failed group along the SFC and traverse the halves again
        std::vector<uint32_t> hf(numFailed), hs(G), he(G);
        RYO_HIP(ctx, hipMemcpyAsync(hf.data(), a.failed, numFailed * 4, hipMemcpyDeviceToHost, s));
        RYO_HIP(ctx, hipMemcpyAsync(hs.data(), a.gStart, G * 4, hipMemcpyDeviceToHost, s));
        RYO_HIP(ctx, hipMemcpyAsync(he.data(), a.gEnd, G * 4, hipMemcpyDeviceToHost, s));
        RYO_HIP(ctx, hipStreamSynchronize(s));
        ns.clear(), ne.clear();
        for (uint32_t g : hf)
        {
            if (he[g] - hs[g] < 2)
                return setError(ctx, RYO_ERR_TRAVERSAL, "traversal: a single target exhausted the stack");
            uint32_t mid = hs[g] + (he[g] - hs[g]) / 2;
            ns.push_back(hs[g]), ns.push_back(mid);
            ne.push_back(mid), ne.push_back(he[g]);
        }
        auto* sb = static_cast<uint32_t*>(ctx->getScratch(kSlotGroups, 8 * ns.size() + 64));
        if (!sb) return RYO_ERR_HIP;
        RYO_HIP(ctx, hipMemcpyAsync(sb, ns.data(), 4 * ns.size(), hipMemcpyHostToDevice, s));
        RYO_HIP(ctx, hipMemcpyAsync(sb + ns.size(), ne.data(), 4 * ne.size(), hipMemcpyHostToDevice, s));
        RYO_HIP(ctx, hipStreamSynchronize(s));
        a.numGroups = int(ns.size());
        a.gStart    = sb;
        a.gEnd      = sb + ns.size();
    }
    for (int k = 0; k < 5; ++k)
        ctx->hStats[k] = h[k];
    if (totalPot) *totalPot = 0.5 * double(a.G) * pot;
    return RYO_OK;
}

/* ------------------------------------------------------------------ allocation-free traversal */

// Work memory of ryo_traverse_f32_noalloc, one copy per GPU for the life of the process: 256 bytes of control
// words (the layout of ctx->dStats: group counters, statistics, potential sum, softening range), 256 FP64
// reduction partials, then per-group FP64 potentials and the overflow list for up to kNoAllocMaxGroups groups
// (12 MiB in all). In host code `scratch` only names the symbol; its device address comes from
// hipGetSymbolAddress, once per device (initNoAllocScratch), and only that address is passed to kernels.
constexpr size_t kNoAllocMaxGroups = size_t(RYO_NOALLOC_MAX_GROUPS);
constexpr size_t kNoAllocPartials  = 256;
constexpr size_t kNoAllocPotOff    = 256 + kNoAllocPartials * sizeof(double);
constexpr size_t kNoAllocFailedOff = kNoAllocPotOff + kNoAllocMaxGroups * sizeof(double);
constexpr size_t kNoAllocBytes     = kNoAllocFailedOff + kNoAllocMaxGroups * sizeof(uint32_t);
constexpr int kNoAllocMaxDevices   = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace
{
std::mutex g_noAllocMutex;
void* g_noAllocPtr[kNoAllocMaxDevices] = {};
} // namespace

// resolve this device's copy of `scratch` (the context's device is current: ryo_ctx_create calls this)
int initNoAllocScratch(Context* ctx)
{
    if (ctx->device < 0 || ctx->device >= kNoAllocMaxDevices)
        return setError(ctx, RYO_ERR_UNSUPPORTED, "device %d: at most %d devices", ctx->device, kNoAllocMaxDevices);
    std::lock_guard<std::mutex> lock(g_noAllocMutex);
    if (g_noAllocPtr[ctx->device]) return RYO_OK;
    void* ptr    = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return setError(ctx, RYO_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", ctx->device,
                        hipGetErrorString(e));
    g_noAllocPtr[ctx->device] = ptr;
    return RYO_OK;
}

// the traversal of traverse() with every work array in this device's `scratch`: no allocation, no split/retry
// (the full-depth LIFO cannot overflow), one read-back of statistics and potential
int traverseNoAlloc(Context* ctx, TravArgs a, double* totalPot)
{
    hipStream_t s = ctx->stream;
    for (int k = 0; k < 6; ++k)
        ctx->hStats[k] = 0;
    ctx->retriedGroups = 0;
    if (a.numGroups <= 0)
    {
        if (totalPot) *totalPot = 0;
        return RYO_OK;
    }
    if (size_t(a.numGroups) > kNoAllocMaxGroups)
        return setError(ctx, RYO_ERR_CAPACITY, "allocation-free traversal: %d groups, at most %zu", a.numGroups,
                        kNoAllocMaxGroups);
    char* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_noAllocMutex);
        if (ctx->device >= 0 && ctx->device < kNoAllocMaxDevices)
            base = static_cast<char*>(g_noAllocPtr[ctx->device]);
    }
    if (!base) return setError(ctx, RYO_ERR_HIP, "allocation-free traversal: scratch of device %d not resolved",
                               ctx->device);

    a.stackCap                = kMaxStack;
    a.ranges                  = ctx->travRanges > 0 ? std::min(ctx->travRanges, kXcds) : kXcds;
    unsigned* ctl             = reinterpret_cast<unsigned*>(base);
    unsigned long long* stats = reinterpret_cast<unsigned long long*>(base) + 8;
    double* potSum            = reinterpret_cast<double*>(base) + 16;
    double* partials          = reinterpret_cast<double*>(base + 256);
    a.work                    = ctl;
    a.stats                   = stats;
    a.hRange                  = ctl + kCtlHRange;
    a.groupPot                = reinterpret_cast<double*>(base + kNoAllocPotOff);
    a.failed                  = reinterpret_cast<uint32_t*>(base + kNoAllocFailedOff);
    a.groupList               = nullptr;

    const size_t G = size_t(a.numGroups);
    travInitKernel<<<1, 64, 0, s>>>(ctl); // counters, statistics and potential 0; softening range empty
    if (a.hCount > 0)
        hRangeKernel<<<std::max(1u, std::min(iceil(a.hCount, 4096), unsigned(ctx->numCUs * 2))), 256, 0, s>>>(
            a.ht, a.hCount, a.hs == a.ht ? nullptr : a.hs, a.hs == a.ht ? 0 : a.hCount, ctl + kCtlHRange);
    zeroKernelU<<<std::min(iceil(2 * G, 256), 4096u), 256, 0, s>>>(reinterpret_cast<unsigned*>(a.groupPot),
                                                                   int(2 * G));
    if (ctx->timing) RYO_HIP(ctx, hipEventRecord(ctx->ev[0], s));
    launchTraversal<false>(ctx, G, kMaxStack, s, a, TravBatch{});
    if (ctx->timing) RYO_HIP(ctx, hipEventRecord(ctx->ev[1], s));
    RYO_LAUNCH_CHECK(ctx);
    sumDeterministic(s, a.groupPot, a.numGroups, partials, potSum);
    RYO_LAUNCH_CHECK(ctx);

    unsigned long long h[7] = {0, 0, 0, 0, 0, 0, 0};
    double pot              = 0;
    RYO_HIP(ctx, hipMemcpyAsync(h, stats, sizeof(h), hipMemcpyDeviceToHost, s));
    RYO_HIP(ctx, hipMemcpyAsync(&pot, potSum, sizeof(double), hipMemcpyDeviceToHost, s));
    RYO_HIP(ctx, hipStreamSynchronize(s));
    if (ctx->timing)
    {
        float ms = 0;
        RYO_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        ctx->lastTraverseMs = ms;
    }
    if (h[6] != 0) return setError(ctx, RYO_ERR_TRAVERSAL, "traversal: step limit exceeded (corrupt tree?)");
    if ((h[5] & 0xffffffffu) != 0)
        return setError(ctx, RYO_ERR_TRAVERSAL, "allocation-free traversal: %u groups overflowed the full LIFO",
                        unsigned(h[5] & 0xffffffffu));
    for (int k = 0; k < 5; ++k)
        ctx->hStats[k] = h[k];
    if (totalPot) *totalPot = 0.5 * double(a.G) * pot;
    return RYO_OK;
}

// per-system FP64 potential: block s sums the partials of groups [GO[s], GO[s+1]) in a fixed order (Neumaier)
__global__ __launch_bounds__(256) void systemPotKernel(const double* __restrict__ groupPot, BatchView bv,
                                                       const float* __restrict__ G, double* __restrict__ out)
{
    __shared__ double ss[256], sc[256];
    const int sys = blockIdx.x;
    double sum = 0, c = 0;
    for (uint32_t i = bv.GO[sys] + threadIdx.x; i < bv.GO[sys + 1]; i += 256)
    {
        double x = groupPot[i], t = sum + x;
        c += fabs(sum) >= fabs(x) ? (sum - t) + x : (x - t) + sum;
        sum = t;
    }
    ss[threadIdx.x] = sum;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if (int(threadIdx.x) < o)
        {
            double x = ss[threadIdx.x + o], y = ss[threadIdx.x];
            double t = x + y;
            sc[threadIdx.x] += sc[threadIdx.x + o] + (fabs(y) >= fabs(x) ? (y - t) + x : (x - t) + y);
            ss[threadIdx.x] = t;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[sys] = 0.5 * double(G[sys]) * (ss[0] + sc[0]);
}

// device memory a batch traversal of numGroups groups works in: per-group potentials (8 B), the overflow list
// (4 B) and the retry list (4 B)
size_t batchTraverseWorkBytes(size_t numGroups) { return 16 * numGroups + 256; }

namespace
{
// launch shape and work arrays of a batch traversal (the same for its launch and its finish)
void batchShape(Context* ctx, TravArgs& a, void* work)
{
    a.stackCap    = ctx->travStack > 8 ? std::min(ctx->travStack, kMaxStack) : kMaxStack;
    a.ranges      = ctx->travRanges > 0 ? std::min(ctx->travRanges, kXcds) : kXcds;
    unsigned* ctl = reinterpret_cast<unsigned*>(ctx->dStats);
    a.work        = ctl;
    a.stats       = ctx->dStats + 8;
    a.hRange      = ctl + kCtlHRange;
    auto* buf     = static_cast<char*>(work);
    a.groupPot    = reinterpret_cast<double*>(buf);
    a.failed      = reinterpret_cast<uint32_t*>(buf + 8 * size_t(a.numGroups));
    a.groupList   = nullptr;
}

} // namespace

// the capturable part: init, softening range, zeroed partials, traversal, per-system potentials
int traverseBatchLaunch(Context* ctx, const TravArgs& a, const TravBatch& tb, double* sysPot, bool events)
{
    hipStream_t s = ctx->stream;
    if (a.numGroups <= 0) return RYO_OK;
    unsigned* ctl = reinterpret_cast<unsigned*>(ctx->dStats);
    travInitKernel<<<1, 64, 0, s>>>(ctl);
    if (a.hCount > 0)
        hRangeKernel<<<std::max(1u, std::min(iceil(a.hCount, 4096), unsigned(ctx->numCUs * 2))), 256, 0, s>>>(
            a.ht, a.hCount, nullptr, 0, ctl + kCtlHRange);
    const size_t G = size_t(a.numGroups);
    zeroKernelU<<<std::min(iceil(2 * G, 256), 4096u), 256, 0, s>>>(reinterpret_cast<unsigned*>(a.groupPot),
                                                                   int(2 * G));
    if (events && ctx->timing) RYO_HIP(ctx, hipEventRecord(ctx->ev[0], s));
    launchTraversal<true>(ctx, G, a.stackCap, s, a, tb);
    if (events && ctx->timing) RYO_HIP(ctx, hipEventRecord(ctx->ev[1], s));
    RYO_LAUNCH_CHECK(ctx);
    systemPotKernel<<<tb.bv.S, 256, 0, s>>>(a.groupPot, tb.bv, tb.G, sysPot);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

// statistics of a finished batch traversal (one host synchronization); groups that overflowed the LIFO wrote
// nothing and are traversed again at full depth, after which the per-system potentials are summed again
int traverseBatchFinish(Context* ctx, TravArgs a, const TravBatch& tb, double* sysPot, bool events)
{
    hipStream_t s = ctx->stream;
    for (int k = 0; k < 6; ++k)
        ctx->hStats[k] = 0;
    ctx->retriedGroups = 0;
    if (a.numGroups <= 0) return RYO_OK;
    unsigned long long h[7] = {0, 0, 0, 0, 0, 0, 0};
    RYO_HIP(ctx, hipMemcpyAsync(h, a.stats, sizeof(h), hipMemcpyDeviceToHost, s));
    RYO_HIP(ctx, hipStreamSynchronize(s));
    if (events && ctx->timing)
    {
        float ms = 0;
        RYO_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        ctx->lastTraverseMs = ms;
    }
    if (h[6] != 0) return setError(ctx, RYO_ERR_TRAVERSAL, "batch traversal: step limit exceeded (corrupt tree?)");
    const unsigned numFailed = unsigned(h[5] & 0xffffffffu);
    if (numFailed > 0)
    {
        if (a.stackCap >= kMaxStack)
            return setError(ctx, RYO_ERR_TRAVERSAL, "batch traversal: %u groups overflowed a full-depth LIFO",
                            numFailed);
        ctx->retriedGroups = numFailed;
        // the retry reads its groups from a copy of the overflow list (it records overflows into the list)
        auto* list = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(a.groupPot) + 12 * size_t(a.numGroups));
        RYO_HIP(ctx, hipMemcpyAsync(list, a.failed, numFailed * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        zeroKernelU<<<1, 64, 0, s>>>(a.work, kXcds);
        zeroKernelU<<<1, 64, 0, s>>>(reinterpret_cast<unsigned*>(a.stats + 5), 4);
        a.groupList = list;
        a.numGroups = int(numFailed);
        a.stackCap  = kMaxStack;
        launchTraversal<true>(ctx, numFailed, kMaxStack, s, a, tb);
        RYO_LAUNCH_CHECK(ctx);
        systemPotKernel<<<tb.bv.S, 256, 0, s>>>(a.groupPot, tb.bv, tb.G, sysPot);
        RYO_LAUNCH_CHECK(ctx);
        unsigned long long r[7] = {0, 0, 0, 0, 0, 0, 0};
        RYO_HIP(ctx, hipMemcpyAsync(r, a.stats, sizeof(r), hipMemcpyDeviceToHost, s));
        RYO_HIP(ctx, hipStreamSynchronize(s));
        if ((r[5] & 0xffffffffu) != 0 || r[6] != 0)
            return setError(ctx, RYO_ERR_TRAVERSAL, "batch traversal: retry at full depth failed");
        h[0] = r[0], h[2] = r[2]; // sums now include the retried groups
        for (int k : {1, 3, 4})
            h[k] = std::max(h[k], r[k]);
    }
    for (int k = 0; k < 5; ++k)
        ctx->hStats[k] = h[k];
    return RYO_OK;
}

/* traversal of every system of a batch: the batch-wide arrays (targets = sources = the packed particles, tree
 * arrays, groups of all systems) and the offset tables; sysPot (device, S entries) receives each system's
 * 0.5 G sum(m phi); work: batchTraverseWorkBytes(numGroups) bytes of device memory; abort (device, optional):
 * nonzero when the trees were not built. Open boundaries. mode 0: launch and finish; 1: launch only (no host
 * synchronization, e.g. under graph capture); 2: finish only, after a mode-1 sequence with the same arguments. */
int batchTraverse(Context* ctx, int numGroups, const uint32_t* gStart, const uint32_t* gEnd, const int32_t* gSys,
                  const BatchView& bv, const float* G, const float* x, const float* y, const float* z,
                  const float* m, const float* h, size_t n, const int32_t* child, const int32_t* i2l,
                  const uint32_t* layout, const float* centers, const float* mp, float* p, float* ax, float* ay,
                  float* az, double* sysPot, int initNode, void* work, const uint32_t* abort, int mode,
                  const SysImages* images)
{
    TravArgs a{};
    a.numGroups = numGroups;
    a.gStart = gStart, a.gEnd = gEnd, a.initNode = initNode;
    a.xt = x, a.yt = y, a.zt = z, a.mt = m, a.ht = h;
    a.xs = x, a.ys = y, a.zs = z, a.ms = m, a.hs = h;
    a.child = child, a.i2l = i2l, a.layout = layout;
    a.centers = reinterpret_cast<const float4*>(centers);
    a.mp      = reinterpret_cast<const float4*>(mp);
    a.G = 1.0f, a.numShells = 0;
    a.p = p, a.ax = ax, a.ay = ay, a.az = az;
    a.hCount = n;
    batchShape(ctx, a, work);
    TravBatch tb{bv, gSys, G, abort, images};
    int rc = RYO_OK;
    if (mode != 2) rc = traverseBatchLaunch(ctx, a, tb, sysPot, mode == 0);
    if (!rc && mode != 1) rc = traverseBatchFinish(ctx, a, tb, sysPot, mode == 0);
    return rc;
}

int traverseEntry(Context* ctx, int numGroups, const uint32_t* gs, const uint32_t* ge, int initNode, const float* xt,
                  const float* yt, const float* zt, const float* mt, const float* ht, const float* xs,
                  const float* ys, const float* zs, const float* ms, const float* hs, const int32_t* child,
                  const int32_t* i2l, const uint32_t* layout, const float* centers, const float* mp, float G,
                  int numShells, const float* boxL, float* p, float* ax, float* ay, float* az, double* totalPot,
                  int stackCap, size_t hCount, bool mdq, bool noAlloc, const uint32_t* abort)
{
    if (numGroups > 0 && (!gs || !ge || !xt || !yt || !zt || !mt || !ht || !xs || !ys || !zs || !ms || !hs ||
                          !child || !i2l || !layout || !centers || !mp || !ax || !ay || !az))
        return setError(ctx, RYO_ERR_ARG, "traverse: null pointer");
    TravArgs a{};
    a.numGroups = numGroups;
    a.gStart = gs, a.gEnd = ge, a.initNode = initNode;
    a.xt = xt, a.yt = yt, a.zt = zt, a.mt = mt, a.ht = ht;
    a.xs = xs, a.ys = ys, a.zs = zs, a.ms = ms, a.hs = hs;
    a.child = child, a.i2l = i2l, a.layout = layout;
    a.centers = reinterpret_cast<const float4*>(centers);
    a.mp      = reinterpret_cast<const float4*>(mp);
    a.G = G, a.numShells = numShells;
    for (int d = 0; d < 3; ++d)
        a.boxL[d] = boxL ? boxL[d] : 0.0f;
    a.p = p, a.ax = ax, a.ay = ay, a.az = az;
    a.stackCap = stackCap;
    a.hCount   = hCount;
    a.mdq      = mdq;
    a.abort    = abort;
    return noAlloc ? traverseNoAlloc(ctx, a, totalPot) : traverse(ctx, a, totalPot);
}

} // namespace ryo
