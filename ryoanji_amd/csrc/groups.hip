// Spatial target groups (gfx950): fixed windows of 64 consecutive SFC-ordered targets split where the targets
// are not compact (cstone computeGroupSplits, traversal/groups_gpu.cu:60-151 and groups_gpu.cuh:29-270).
//
// One wavefront per window, one target per lane (the window is exactly one wave64):
//   * each lane finds its leaf (binary search of the layout) and the leaf's cell box in a unit box with the
//     system's per-axis key bits; the window's distance criterion is tolFactor * cbrt(smallest leaf volume);
//   * a split follows target l when its distance to target l + 1 exceeds min(criterion, 2 h_l) (positions
//     and h in units of the longest box edge);
//   * a window whose extent along some axis exceeds extentRatio x tolFactor x the smallest leaf edge along that
//     axis is also cut into equal power-of-two chunks within that limit.
// The split masks become group boundaries through a two-level scan of the per-window group counts.
#include "common.hpp"
#include "sfc.hpp"

namespace ryo
{

constexpr int kWindow      = 64;
constexpr int kScanItems   = 16;                 // windows per thread in the scan/expand kernel
constexpr int kScanChunk   = 256 * kScanItems;   // windows per block

struct SplitArgs
{
    uint32_t first, last;
    int numWindows;
    const float *x, *y, *z, *h;
    const uint64_t* leaves;
    int numLeaves;
    const uint32_t* layout;
    Bits3 bits;
    float invMaxExtent;
    float tolFactor;
    float extentRatio;
    uint64_t* masks;
    uint32_t* counts;
};

// leaf index of particle i: the last leaf whose layout entry is <= i (upper_bound - 1)
__device__ inline int leafOf(const uint32_t* layout, int numLeaves, uint32_t i)
{
    int lo = 0, hi = numLeaves; // answer in [0, numLeaves)
    while (hi - lo > 1)
    {
        int mid = (lo + hi) >> 1;
        if (layout[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void groupSplitsKernel(SplitArgs a)
{
#pragma clang fp contract(off) // the split criteria round every operation as the reference CPU does
    const int w    = int((size_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (w >= a.numWindows) return; // whole waves exit together

    const uint32_t ws   = a.first + uint32_t(w) * kWindow;
    const uint32_t body = min(ws + uint32_t(lane), a.last - 1);
    const int leaf      = leafOf(a.layout, a.numLeaves, body);

    // the leaf's cell in the unit box whose edges are 2^-(21 - bits[d]): half-unit length 2^-22 on every axis
    const uint64_t k0 = a.leaves[leaf], k1 = a.leaves[leaf + 1];
    IBox3 ib          = cellIBox(k0, treeLevel(k1 - k0), a.bits);
    const float hul   = 1.0f / float(1u << (kMaxLevel + 1));
    float sz[3];
    for (int d = 0; d < 3; ++d)
        sz[d] = float(ib.len[d]) * hul;
    float vol    = ((8.0f * sz[0]) * sz[1]) * sz[2];
    float volMin = waveMin(vol > 0.0f ? fminf(vol, 1.0f) : 1.0f);
    float edge[3];
    for (int d = 0; d < 3; ++d)
        edge[d] = waveMin(fminf(1.0f, 2.0f * sz[d]));
    const float distCrit   = (cbrtf(volMin) * a.tolFactor);
    const float distCritSq = (distCrit * distCrit);

    // positions and interaction radii in units of the longest box edge
    const float px = (a.x[body] * a.invMaxExtent), py = (a.y[body] * a.invMaxExtent),
                pz = (a.z[body] * a.invMaxExtent);
    const float pr = a.h ? ((2.0f * a.h[body]) * a.invMaxExtent) : 1.0f;

    // consecutive-distance criterion (the last lane compares with itself: never a split)
    const float nx = __shfl_down(px, 1), ny = __shfl_down(py, 1), nz = __shfl_down(pz, 1);
    const float dx = (nx - px), dy = (ny - py), dz = (nz - pz);
    const float d2 = dx * dx + (dy * dy + dz * dz);
    uint64_t mask  = __ballot(d2 > fminf(distCritSq, (pr * pr)));

    // window-extent criterion
    float ratio       = 0.0f;
    const float pos[3] = {px, py, pz};
    for (int d = 0; d < 3; ++d)
    {
        const float allowed = (a.tolFactor * edge[d]);
        const float extent  = (waveMax(pos[d]) - waveMin(pos[d]));
        if (allowed > 0.0f) ratio = fmaxf(ratio, (extent / allowed));
    }
    if (a.extentRatio > 0.0f && ratio > a.extentRatio)
    {
        unsigned chunk = kWindow;
        while (chunk > 1 && (float(chunk) * ratio) > (float(kWindow) * a.extentRatio))
            chunk >>= 1;
        const uint32_t inWindow = min(uint32_t(kWindow), a.last - ws);
        const uint32_t next     = uint32_t(lane) + 1;
        mask |= __ballot(next < inWindow && next % chunk == 0);
    }
    if (lane == 0)
    {
        a.masks[w]  = mask;
        a.counts[w] = 1u + unsigned(__popcll(mask));
    }
}

// sum of the group counts of each chunk of kScanChunk windows
__global__ __launch_bounds__(256) void chunkSumKernel(const uint32_t* __restrict__ counts, int numWindows,
                                                      uint32_t* __restrict__ sums)
{
    __shared__ uint32_t sw[4];
    const int base = blockIdx.x * kScanChunk + threadIdx.x * kScanItems;
    uint32_t v     = 0;
    for (int k = 0; k < kScanItems; ++k)
        v += base + k < numWindows ? counts[base + k] : 0u;
    uint32_t tot;
    blockExclusiveScan256(v, sw, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// group starts: window w writes its first particle and the particle after each split at its scanned offset;
// the entry after the last group is `last`
__global__ __launch_bounds__(256) void expandGroupsKernel(SplitArgs a, const uint32_t* __restrict__ chunkBase,
                                                          uint32_t* __restrict__ groups)
{
    __shared__ uint32_t sw[4];
    const int base = blockIdx.x * kScanChunk + threadIdx.x * kScanItems;
    uint32_t v     = 0;
    for (int k = 0; k < kScanItems; ++k)
        v += base + k < a.numWindows ? a.counts[base + k] : 0u;
    uint32_t tot;
    uint32_t off = chunkBase[blockIdx.x] + blockExclusiveScan256(v, sw, &tot);
    for (int k = 0; k < kScanItems && base + k < a.numWindows; ++k)
    {
        const int w       = base + k;
        const uint32_t ws = a.first + uint32_t(w) * kWindow;
        uint64_t m        = a.masks[w];
        groups[off++]     = ws;
        while (m)
        {
            groups[off++] = ws + uint32_t(__ffsll((unsigned long long)m)); // split after bit l: start l + 1
            m &= m - 1;
        }
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) groups[chunkBase[gridDim.x]] = a.last;
}

__global__ __launch_bounds__(256) void scanChunksKernel(uint32_t* __restrict__ sums, int numChunks)
{
    __shared__ uint32_t sw[4];
    uint32_t carry = 0;
    for (int b0 = 0; b0 < numChunks; b0 += 256)
    {
        int b       = b0 + threadIdx.x;
        uint32_t v  = b < numChunks ? sums[b] : 0, tot;
        uint32_t ex = blockExclusiveScan256(v, sw, &tot);
        if (b < numChunks) sums[b] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) sums[numChunks] = carry;
}

int groupSplits(Context* ctx, uint32_t first, uint32_t last, const float* x, const float* y, const float* z,
                const float* h, const uint64_t* leaves, int numLeaves, const uint32_t* layout, const float* limits,
                unsigned groupSize, float tolFactor, float extentRatio, uint32_t* groups, size_t capacity,
                int* numGroups)
{
    if (groupSize != kWindow)
        return setError(ctx, RYO_ERR_UNSUPPORTED, "group splits: group size %u (64 targets per wavefront)", groupSize);
    if (!x || !y || !z || !leaves || !layout || !limits || !groups || !numGroups || numLeaves < 1 || last < first)
        return setError(ctx, RYO_ERR_ARG, "group splits: invalid arguments");
    if (last == first)
    {
        *numGroups = 0;
        if (capacity < 1) return setError(ctx, RYO_ERR_CAPACITY, "group splits: capacity 0");
        RYO_HIP(ctx, hipMemcpyAsync(groups, &last, sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
        RYO_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return RYO_OK;
    }
    SplitArgs a{};
    a.first = first, a.last = last;
    a.numWindows = int((last - first + kWindow - 1) / kWindow);
    a.x = x, a.y = y, a.z = z, a.h = h;
    a.leaves = leaves, a.numLeaves = numLeaves, a.layout = layout;
    a.bits = boxBits(limits);
    float len[3] = {limits[1] - limits[0], limits[3] - limits[2], limits[5] - limits[4]};
    a.invMaxExtent = 1.0f / std::max(std::max(len[0], len[1]), len[2]);
    a.tolFactor = tolFactor, a.extentRatio = extentRatio;

    const int numChunks = int(iceil(size_t(a.numWindows), kScanChunk));
    const size_t maskOff = 0, cntOff = (8 * size_t(a.numWindows) + 255) & ~size_t(255);
    const size_t sumOff  = cntOff + ((4 * size_t(a.numWindows) + 255) & ~size_t(255));
    auto* buf = static_cast<char*>(ctx->getScratch(kSlotSplits, sumOff + 4 * (size_t(numChunks) + 1)));
    if (!buf) return RYO_ERR_HIP;
    a.masks       = reinterpret_cast<uint64_t*>(buf + maskOff);
    a.counts      = reinterpret_cast<uint32_t*>(buf + cntOff);
    uint32_t* sum = reinterpret_cast<uint32_t*>(buf + sumOff);

    hipStream_t s = ctx->stream;
    groupSplitsKernel<<<iceil(size_t(a.numWindows) * kWindow, 256), 256, 0, s>>>(a);
    chunkSumKernel<<<numChunks, 256, 0, s>>>(a.counts, a.numWindows, sum);
    scanChunksKernel<<<1, 256, 0, s>>>(sum, numChunks);
    RYO_LAUNCH_CHECK(ctx);
    uint32_t total = 0;
    RYO_HIP(ctx, hipMemcpyAsync(&total, sum + numChunks, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    RYO_HIP(ctx, hipStreamSynchronize(s));
    *numGroups = int(total);
    if (size_t(total) + 1 > capacity)
        return setError(ctx, RYO_ERR_CAPACITY, "group splits: %u groups need %u entries > capacity %zu", total,
                        total + 1, capacity);
    expandGroupsKernel<<<numChunks, 256, 0, s>>>(a, sum, groups);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

} // namespace ryo
