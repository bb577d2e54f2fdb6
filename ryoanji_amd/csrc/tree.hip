// Cornerstone leaf array, per-leaf layout, the linked octree and fixed target groups (gfx950).
//
// Leaves. The reference reaches its leaf array by iterating split/merge passes with host syncs
// (csarray.hpp:267-440, csarray_gpu.cu:86-231). The fixed point of those passes is unique: a node is internal
// iff it holds more than `bucket` particles and its level is below 21, and the leaves are the children of
// internal nodes that are not internal themselves. This file builds that tree directly from the sorted keys
// in one parallel pass:
//   c_j = common octree level of keys j and j+bucket (the window of bucket+1 particles starting at j),
//   M_i = max c_j over the windows containing particle i (its deepest ancestor with > bucket particles),
//   the leaf of particle i is at level min(M_i + 1, 21) and starts at key_i truncated to that level.
// Distinct particle leaves are the non-empty leaves; the gaps between them are covered by the coarsest
// aligned octree nodes (the empty leaves, as cstone's spanSfcRange, common.hpp:413-461).
//
// Linked octree. The same algorithm as buildOctreeGpu (octree_gpu.cu:138-200): Warren-Salmon prefixes of
// leaves and internal nodes in the binary-tree layout (octree.hpp:77-100, binaryKeyWeight :55-65), a sort by
// prefix, the inverse permutation, level ranges and child links -- the arrays it defines are part of the
// interface (internalToLeaf / leafToInternal expose the unsorted layout), so they are produced identically.
#include "common.hpp"

namespace ryo
{

int sortPairs(Context* ctx, uint64_t* keys, int32_t* vals, size_t n, uint64_t* keyBuf, int32_t* valBuf, void* tmp,
              size_t tmpBytes);
int sortPairsDigits(Context* ctx, uint64_t* keys, int32_t* vals, size_t n, uint64_t* keyBuf, int32_t* valBuf,
                    void* tmp, size_t tmpBytes, int digits, const uint32_t* dN);
size_t sortTempBytes(size_t n);

/* ============================================================================ leaves */

constexpr int kLeafThreads = 256;
constexpr int kLeafTile    = 1024; // particles per block
constexpr int kMaxBucket   = 1024;

// finest-first decomposition of [a, b) into aligned octree nodes; the level of the node starting at a
__device__ __forceinline__ unsigned spanLevel(uint64_t a, uint64_t b)
{
    unsigned l = a == 0 ? 0u : kMaxLevel - unsigned(ctz64(a)) / 3u;
    while (a + nodeRange(l) > b)
        ++l;
    return l;
}
__device__ inline int spanCount(uint64_t a, uint64_t b)
{
    int c = 0;
    while (a < b)
    {
        a += nodeRange(spanLevel(a, b));
        ++c;
    }
    return c;
}

// LDS of one leaf tile for buckets up to MaxB (21 KB for MaxB = 64, so 7 blocks fit a CU; 45 KB for 1024)
template<int MaxB>
struct LeafTileSmem
{
    uint64_t keys[kLeafTile + 2 * MaxB + 2];
    int8_t c[2][kLeafTile + MaxB + 2]; // common levels -1 .. 21
    uint64_t start[kLeafTile + 1];     // leaf start key of particles t0-1 .. t1-1
    unsigned char level[kLeafTile + 1];
    uint32_t sw[4];
};
constexpr int kSmallBucket = 64;

/* Per tile [t0, t1): leaf start/level of particles t0-1..t1-1 into s.start/s.level, and the number of
 * leaves each particle of the tile emits (returned for the first kLeafTile/256 items of each thread). */
template<int MaxB>
__device__ __forceinline__ void leafTile(const uint64_t* __restrict__ k, long n, int B, long t0, LeafTileSmem<MaxB>& s, int em[4])
{
    const long t1    = min(t0 + long(kLeafTile), n);
    const long kbase = t0 - 1 - B; // s.keys[j - kbase] = k[j]
    const int nk     = int(t1 + B - kbase);
    for (int j = threadIdx.x; j < nk; j += kLeafThreads)
    {
        long g      = kbase + j;
        s.keys[j] = (g >= 0 && g < n) ? k[g] : 0;
    }
    __syncthreads();
    // c over windows j in [kbase, t1)
    const int nc = int(t1 - kbase);
    const long lastWindow = n - 1 - B;
    for (int j = threadIdx.x; j < nc; j += kLeafThreads)
    {
        long g = kbase + j;
        int v  = -1;
        if (g >= 0 && g <= lastWindow)
        {
            int cp = commonPrefix(s.keys[j], s.keys[j + B]);
            v      = min(cp / 3, int(kMaxLevel));
        }
        s.c[0][j] = int8_t(v);
    }
    __syncthreads();
    // doubling: c[P][j] = max(c[j .. j+P-1]) for the largest power of two P <= B+1
    const int W = B + 1;
    int P       = 1;
    int cur     = 0;
    while (2 * P <= W)
    {
        for (int j = threadIdx.x; j < nc; j += kLeafThreads)
        {
            int a = s.c[cur][j];
            int b = (j + P < nc) ? s.c[cur][j + P] : -1;
            s.c[cur ^ 1][j] = int8_t(max(a, b));
        }
        __syncthreads();
        cur ^= 1;
        P *= 2;
    }
    // particles i in [t0-1, t1): window max over j in [i-B, i] = max(c[i-B .. i-B+P-1], c[i+1-P .. i])
    const int ni = int(t1 - (t0 - 1));
    for (int q = threadIdx.x; q < ni; q += kLeafThreads)
    {
        long i = t0 - 1 + q;
        if (i < 0) continue;
        int ja = int(i - B - kbase), jb = int(i + 1 - P - kbase);
        int M  = max(s.c[cur][ja], s.c[cur][jb]);
        unsigned L = unsigned(min(M + 1, int(kMaxLevel)));
        uint64_t key = s.keys[int(i - kbase)];
        s.start[q] = key & ~(nodeRange(L) - 1);
        s.level[q] = (unsigned char)L;
    }
    __syncthreads();
    for (int r = 0; r < kLeafTile / kLeafThreads; ++r)
    {
        int q  = r * kLeafThreads + threadIdx.x; // particle t0 + q, s index q + 1
        long i = t0 + q;
        int e  = 0;
        if (i < t1)
        {
            uint64_t si = s.start[q + 1];
            bool boundary = (i == 0) || si != s.start[q];
            if (boundary)
            {
                uint64_t prevEnd = (i == 0) ? 0 : s.start[q] + nodeRange(s.level[q]);
                e                = spanCount(prevEnd, si) + 1;
            }
            if (i == n - 1) e += spanCount(si + nodeRange(s.level[q + 1]), kKeyEnd);
        }
        em[r] = e;
    }
}

// exclusive scan of the block sums in place (one block); total into blockSums[numBlocks]
__global__ __launch_bounds__(256) void scanBlocksKernel(uint32_t* __restrict__ sums, int numBlocks)
{
    // thread t scans the contiguous run [t k, (t + 1) k): independent loads, one block-wide scan of the run sums
    __shared__ uint32_t sw[4];
    const int k  = (numBlocks + 255) / 256;
    const int b0 = int(threadIdx.x) * k, b1 = min(numBlocks, b0 + k);
    uint32_t run = 0;
    for (int b = b0; b < b1; ++b)
        run += sums[b];
    uint32_t tot;
    uint32_t carry = blockExclusiveScan256(run, sw, &tot);
    for (int b = b0; b < b1; ++b)
    {
        const uint32_t v = sums[b];
        sums[b] = carry;
        carry += v;
    }
    if (threadIdx.x == 0) sums[numBlocks] = tot;
}

// exclusive scan of many block sums: totals of chunks of kScanChunk sums, a scan of the totals, then each chunk
// scanned with its offset (coalesced; a single block reading runs of 256 strided sums took 115 us for 65536)
constexpr int kScanChunk = 1024;

__global__ __launch_bounds__(256) void chunkTotalsKernel(const uint32_t* __restrict__ v, int n,
                                                         uint32_t* __restrict__ totals)
{
    __shared__ uint32_t sw[4];
    const int base = blockIdx.x * kScanChunk, end = min(n, base + kScanChunk);
    uint32_t s = 0;
    for (int i = base + int(threadIdx.x); i < end; i += 256)
        s += v[i];
    uint32_t tot;
    blockExclusiveScan256(s, sw, &tot);
    if (threadIdx.x == 0) totals[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void chunkScanKernel(uint32_t* __restrict__ v, int n,
                                                       const uint32_t* __restrict__ offsets)
{
    __shared__ uint32_t sw[4];
    const int base = blockIdx.x * kScanChunk + int(threadIdx.x) * 4;
    uint32_t a[4], run = 0;
    for (int k = 0; k < 4; ++k)
        a[k] = base + k < n ? v[base + k] : 0u, run += a[k];
    uint32_t tot;
    uint32_t ex = blockExclusiveScan256(run, sw, &tot) + offsets[blockIdx.x];
    for (int k = 0; k < 4; ++k)
    {
        if (base + k < n) v[base + k] = ex;
        ex += a[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) v[n] = offsets[gridDim.x];
}

// exclusive scan of sums[0, n) in place, the total into sums[n]; `work` holds n / kScanChunk + 2 words
int scanBlockSums(Context* ctx, uint32_t* sums, int n, uint32_t* work)
{
    hipStream_t s = ctx->stream;
    if (n <= 4 * kScanChunk)
    {
        scanBlocksKernel<<<1, 256, 0, s>>>(sums, n);
        RYO_LAUNCH_CHECK(ctx);
        return RYO_OK;
    }
    const int nc = (n + kScanChunk - 1) / kScanChunk;
    chunkTotalsKernel<<<nc, 256, 0, s>>>(sums, n, work);
    scanBlocksKernel<<<1, 256, 0, s>>>(work, nc);
    chunkScanKernel<<<nc, 256, 0, s>>>(sums, n, work);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

// leaves of the tile starting at particle t0, written from position `base` on
template<int MaxB>
__device__ __forceinline__ uint32_t leafEmitTile(const uint64_t* __restrict__ k, long n, int B, long t0,
                                                 uint32_t base, uint64_t* __restrict__ leaves,
                                                 uint32_t* __restrict__ layout, long capacity, LeafTileSmem<MaxB>& s);


// returns the number of leaves the tile emits (the terminators of the last tile not counted)
template<int MaxB>
__device__ __forceinline__ uint32_t leafEmitTile(const uint64_t* __restrict__ k, long n, int B, long t0,
                                                 uint32_t base, uint64_t* __restrict__ leaves,
                                                 uint32_t* __restrict__ layout, long capacity, LeafTileSmem<MaxB>& s)
{
    const uint32_t base0 = base;
    int em[4];
    leafTile(k, n, B, t0, s, em);
    // thread-contiguous order: item r of thread t is particle t0 + r*256 + t -> scan in that order
    // (the emission order must follow particle order, so scan over q = r*256 + t by rows)
    for (int r = 0; r < kLeafTile / kLeafThreads; ++r)
    {
        uint32_t tot;
        uint32_t ex = blockExclusiveScan256(uint32_t(em[r]), s.sw, &tot);
        uint32_t pos = base + ex;
        base += tot;
        int q  = r * kLeafThreads + threadIdx.x;
        long i = t0 + q;
        // the last particle always writes the terminators leaves[L], layout[L], even when it emits no leaf
        if (em[r] == 0 && i != n - 1) continue;
        uint64_t si = s.start[q + 1];
        bool boundary = (i == 0) || si != s.start[q];
        if (boundary)
        {
            uint64_t a = (i == 0) ? 0 : s.start[q] + nodeRange(s.level[q]);
            while (a < si)
            {
                if (pos < capacity) leaves[pos] = a, layout[pos] = uint32_t(i);
                ++pos;
                a += nodeRange(spanLevel(a, si));
            }
            if (pos < capacity) leaves[pos] = si, layout[pos] = uint32_t(i);
            ++pos;
        }
        if (i == n - 1)
        {
            uint64_t a = si + nodeRange(s.level[q + 1]);
            while (a < kKeyEnd)
            {
                if (pos < capacity) leaves[pos] = a, layout[pos] = uint32_t(n);
                ++pos;
                a += nodeRange(spanLevel(a, kKeyEnd));
            }
            if (pos <= capacity) leaves[pos] = kKeyEnd, layout[pos] = uint32_t(n);
        }
    }
    return base - base0;
}

/* Leaves through a staging area: the count pass emits each tile's leaves at their positions within the tile into a
 * per-tile area of `stage` entries (as many as fit) and records the tile's count; once the counts are scanned, the
 * place pass copies every tile that fit to its offset and emits the others (and the last tile, which also writes
 * the terminators) again from the keys. The emission is computed once for almost every tile instead of twice. */
__host__ __device__ inline int leafStageSize(int B) { return 2 * kLeafTile / B + 32 < 256 ? 2 * kLeafTile / B + 32 : 256; }

template<int MaxB>
__global__ __launch_bounds__(kLeafThreads) void leafCountStageKernel(const uint64_t* __restrict__ k, long n, int B,
                                                                    uint32_t* __restrict__ blockSums,
                                                                    uint64_t* __restrict__ stageLeaves,
                                                                    uint32_t* __restrict__ stageLayout, int stage)
{
    __shared__ LeafTileSmem<MaxB> s;
    // capacity stage - 1: the last tile's terminator (written at the position past its leaves) stays in the area
    const uint32_t tot = leafEmitTile(k, n, B, long(blockIdx.x) * kLeafTile, 0u,
                                      stageLeaves + size_t(blockIdx.x) * stage,
                                      stageLayout + size_t(blockIdx.x) * stage, long(stage - 1), s);
    if (threadIdx.x == 0) blockSums[blockIdx.x] = tot;
}

template<int MaxB>
__global__ __launch_bounds__(kLeafThreads) void leafPlaceKernel(const uint64_t* __restrict__ k, long n, int B,
                                                               const uint32_t* __restrict__ blockOffsets,
                                                               const uint64_t* __restrict__ stageLeaves,
                                                               const uint32_t* __restrict__ stageLayout, int stage,
                                                               uint64_t* __restrict__ leaves,
                                                               uint32_t* __restrict__ layout, long capacity)
{
    __shared__ LeafTileSmem<MaxB> s;
    const uint32_t t = blockIdx.x, base = blockOffsets[t], tot = blockOffsets[t + 1] - base;
    if (t + 1 == gridDim.x || tot >= uint32_t(stage - 1))
    {
        leafEmitTile(k, n, B, long(t) * kLeafTile, base, leaves, layout, capacity, s);
        return;
    }
    const uint64_t* sl = stageLeaves + size_t(t) * stage;
    const uint32_t* sy = stageLayout + size_t(t) * stage;
    for (uint32_t j = threadIdx.x; j < tot; j += kLeafThreads)
    {
        const long pos = long(base) + long(j);
        if (pos < capacity) leaves[pos] = sl[j], layout[pos] = sy[j];
    }
}

__global__ void leafCountsKernel(const uint32_t* __restrict__ layout, uint32_t* __restrict__ counts, int L,
                                 const uint32_t* __restrict__ stat = nullptr)
{
    if (stat) L = int(stat[kStLc]);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < L; i += gridDim.x * blockDim.x)
        counts[i] = layout[i + 1] - layout[i];
}

__global__ void singleLeafKernel(uint64_t* leaves, uint32_t* counts, uint32_t* layout, uint32_t n)
{
    leaves[0] = 0, leaves[1] = kKeyEnd;
    counts[0] = n;
    layout[0] = 0, layout[1] = n;
}

int buildLeaves(Context* ctx, const uint64_t* keys, size_t n, unsigned bucket, uint64_t* leaves, uint32_t* counts,
                uint32_t* layout, size_t capacity, int* numLeaves)
{
    hipStream_t s = ctx->stream;
    if (bucket == 0) return setError(ctx, RYO_ERR_ARG, "leaves: bucket size must be positive");
    if (n <= bucket)
    {
        *numLeaves = 1;
        if (capacity < 1) return setError(ctx, RYO_ERR_CAPACITY, "leaves: capacity 0 < 1");
        singleLeafKernel<<<1, 1, 0, s>>>(leaves, counts, layout, uint32_t(n));
        RYO_LAUNCH_CHECK(ctx);
        return RYO_OK;
    }
    if (bucket > unsigned(kMaxBucket))
        return setError(ctx, RYO_ERR_UNSUPPORTED, "leaves: bucket size %u > %d", bucket, kMaxBucket);
    int numBlocks = int((n + kLeafTile - 1) / kLeafTile);
    const size_t scanWork = size_t(numBlocks) / kScanChunk + 2;
    const int stage = leafStageSize(int(bucket));
    const size_t sumBytes   = ((numBlocks + 1 + scanWork) * sizeof(uint32_t) + 255) & ~size_t(255);
    const size_t stageCount = size_t(numBlocks) * stage;
    auto* buf = static_cast<char*>(ctx->getScratch(kSlotLeafBlocks, sumBytes + stageCount * 12));
    if (!buf) return RYO_ERR_HIP;
    auto* sums        = reinterpret_cast<uint32_t*>(buf);
    auto* stageLeaves = reinterpret_cast<uint64_t*>(buf + sumBytes);
    auto* stageLayout = reinterpret_cast<uint32_t*>(buf + sumBytes + stageCount * 8);
    if (bucket <= unsigned(kSmallBucket))
        leafCountStageKernel<kSmallBucket><<<numBlocks, kLeafThreads, 0, s>>>(keys, long(n), int(bucket), sums,
                                                                               stageLeaves, stageLayout, stage);
    else
        leafCountStageKernel<kMaxBucket><<<numBlocks, kLeafThreads, 0, s>>>(keys, long(n), int(bucket), sums,
                                                                             stageLeaves, stageLayout, stage);
    if (int rc = scanBlockSums(ctx, sums, numBlocks, sums + numBlocks + 1)) return rc;
    uint32_t L = 0;
    RYO_HIP(ctx, hipMemcpyAsync(&L, sums + numBlocks, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    RYO_HIP(ctx, hipStreamSynchronize(s));
    *numLeaves = int(L);
    if (size_t(L) > capacity) return setError(ctx, RYO_ERR_CAPACITY, "leaves: %u leaves > capacity %zu", L, capacity);
    if (bucket <= unsigned(kSmallBucket))
        leafPlaceKernel<kSmallBucket><<<numBlocks, kLeafThreads, 0, s>>>(keys, long(n), int(bucket), sums, stageLeaves,
                                                                          stageLayout, stage, leaves, layout, long(L));
    else
        leafPlaceKernel<kMaxBucket><<<numBlocks, kLeafThreads, 0, s>>>(keys, long(n), int(bucket), sums, stageLeaves,
                                                                        stageLayout, stage, leaves, layout, long(L));
    leafCountsKernel<<<iceil(L, 256), 256, 0, s>>>(layout, counts, int(L));
    RYO_LAUNCH_CHECK(ctx);
    // the layout drives every later particle access: check its terminator before anything reads it
    uint32_t last = 0;
    RYO_HIP(ctx, hipMemcpyAsync(&last, layout + L, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    RYO_HIP(ctx, hipStreamSynchronize(s));
    if (last != uint32_t(n)) return setError(ctx, RYO_ERR_HIP, "leaves: layout[L] = %u != n = %zu", last, n);
    return RYO_OK;
}

// the build's counts from the scanned leaf total (one thread)
__global__ void leafStatusKernel(const uint32_t* __restrict__ total, const uint32_t* __restrict__ layout, uint32_t n,
                                 uint32_t capacity, uint32_t* __restrict__ stat)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t L = *total;
    // more leaves than the arrays hold, or a layout that does not end at n: every later stage sees an empty tree
    // and the traversal an abort flag, so none reads through the clipped or inconsistent leaf array; the caller
    // builds again with grown arrays, or reports the layout
    const bool over   = L > capacity;
    const uint32_t last = over ? n : layout[L];
    const bool abort  = over || last != n;
    const uint32_t Lc = abort ? 0u : L;
    const uint32_t nI = Lc > 0 ? (Lc - 1) / 7 : 0;
    stat[kStL] = L, stat[kStNodes] = Lc + nI, stat[kStInternal] = nI, stat[kStOverflow] = over;
    stat[kStLast] = last, stat[kStLc] = Lc, stat[kStAbort] = abort;
}

/* The leaves of a system of more than `bucket` particles sized on the device: count, scan, emission into arrays of
 * `capacity` leaves and the counts, with no host round trip; stat (device, kStWords words) receives the leaf count,
 * node counts, an overflow flag (more than `capacity` leaves: the caller builds again with larger arrays) and the
 * layout terminator for the caller's check (layout[L] must be n). Every later stage sized from stat stays inside
 * arrays sized for `capacity` leaves. */
int buildLeavesDevice(Context* ctx, const uint64_t* keys, size_t n, unsigned bucket, uint64_t* leaves,
                      uint32_t* counts, uint32_t* layout, size_t capacity, uint32_t* stat)
{
    hipStream_t s = ctx->stream;
    if (bucket == 0 || n <= bucket || bucket > unsigned(kMaxBucket))
        return setError(ctx, RYO_ERR_ARG, "leaves: device-sized build needs n > bucket and bucket <= %d", kMaxBucket);
    int numBlocks = int((n + kLeafTile - 1) / kLeafTile);
    const size_t scanWork = size_t(numBlocks) / kScanChunk + 2;
    const int stage = leafStageSize(int(bucket));
    const size_t sumBytes   = ((numBlocks + 1 + scanWork) * sizeof(uint32_t) + 255) & ~size_t(255);
    const size_t stageCount = size_t(numBlocks) * stage;
    auto* buf = static_cast<char*>(ctx->getScratch(kSlotLeafBlocks, sumBytes + stageCount * 12));
    if (!buf) return RYO_ERR_HIP;
    auto* sums        = reinterpret_cast<uint32_t*>(buf);
    auto* stageLeaves = reinterpret_cast<uint64_t*>(buf + sumBytes);
    auto* stageLayout = reinterpret_cast<uint32_t*>(buf + sumBytes + stageCount * 8);
    if (bucket <= unsigned(kSmallBucket))
        leafCountStageKernel<kSmallBucket><<<numBlocks, kLeafThreads, 0, s>>>(keys, long(n), int(bucket), sums,
                                                                               stageLeaves, stageLayout, stage);
    else
        leafCountStageKernel<kMaxBucket><<<numBlocks, kLeafThreads, 0, s>>>(keys, long(n), int(bucket), sums,
                                                                             stageLeaves, stageLayout, stage);
    if (int rc = scanBlockSums(ctx, sums, numBlocks, sums + numBlocks + 1)) return rc;
    if (bucket <= unsigned(kSmallBucket))
        leafPlaceKernel<kSmallBucket><<<numBlocks, kLeafThreads, 0, s>>>(keys, long(n), int(bucket), sums, stageLeaves,
                                                                          stageLayout, stage, leaves, layout,
                                                                          long(capacity));
    else
        leafPlaceKernel<kMaxBucket><<<numBlocks, kLeafThreads, 0, s>>>(keys, long(n), int(bucket), sums, stageLeaves,
                                                                        stageLayout, stage, leaves, layout,
                                                                        long(capacity));
    leafStatusKernel<<<1, 64, 0, s>>>(sums + numBlocks, layout, uint32_t(n), uint32_t(capacity), stat);
    leafCountsKernel<<<std::max(1u, std::min(iceil(capacity, 256), unsigned(ctx->numCUs * 8))), 256, 0, s>>>(
        layout, counts, int(capacity), stat);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

/* ============================================================================ linked octree */

__device__ __forceinline__ int digitWeight(int d)
{
    int ge4 = -int(d >= 4);
    return ((7 - d) & ge4) - (d & ~ge4);
}

// prefixes and unsorted-layout indices of leaves (at numInternal + leaf) and internal nodes (at the
// binary-tree index derived from the leaf boundary between the node's 4th and 5th child)
__global__ void unsortedLayoutKernel(const uint64_t* __restrict__ leaves, int L, int nI, uint64_t* __restrict__ pre,
                                     int32_t* __restrict__ idx, const uint32_t* __restrict__ stat = nullptr)
{
    if (stat) L = int(stat[kStLc]), nI = int(stat[kStInternal]);
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < L; t += gridDim.x * blockDim.x)
    {
        uint64_t key  = leaves[t], next = leaves[t + 1];
        unsigned lvl  = treeLevel(next - key);
        pre[t + nI]   = encodePlaceholder(key, 3 * lvl);
        idx[t + nI]   = t + nI;
        int pl        = commonPrefix(key, next);
        if (pl % 3 == 0 && t < L - 1)
        {
            unsigned level = unsigned(pl) / 3u;
            int w          = 0;
            for (unsigned l = 1; l <= level + 1; ++l)
                w += digitWeight(int(octalDigit(key, l)));
            int o = (t + w) / 7;
            if (o >= 0 && o < nI)
            {
                pre[o] = encodePlaceholder(key, pl);
                idx[o] = o;
            }
        }
    }
}

__global__ void invertKernel(int32_t* __restrict__ i2l, int32_t* __restrict__ l2i, int n, int nI,
                             const uint32_t* __restrict__ stat = nullptr)
{
    if (stat) n = int(stat[kStNodes]), nI = int(stat[kStInternal]);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        int u = i2l[i];
        if (u >= 0 && u < n) l2i[u] = i;
        i2l[i] = u - nI;
    }
}

__device__ inline int lowerBound(const uint64_t* a, int lo, int hi, uint64_t v)
{
    while (lo < hi)
    {
        int mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void levelRangeKernel(const uint64_t* __restrict__ pre, int n, int32_t* __restrict__ levelRange,
                                 const uint32_t* __restrict__ stat = nullptr)
{
    if (stat) n = int(stat[kStNodes]);
    int l = threadIdx.x;
    if (l <= int(kMaxLevel)) levelRange[l] = lowerBound(pre, 0, n, encodePlaceholder(0, 3 * l));
    if (l == int(kMaxLevel) + 1) levelRange[l] = n;
}

__global__ void zeroI32Kernel(int32_t* a, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        a[i] = 0;
}

__global__ void linkKernel(const uint64_t* __restrict__ pre, int n, int nI, const int32_t* __restrict__ l2i,
                           const int32_t* __restrict__ levelRange, int32_t* __restrict__ child,
                           int32_t* __restrict__ parents, const uint32_t* __restrict__ stat = nullptr)
{
    if (stat) n = int(stat[kStNodes]), nI = int(stat[kStInternal]);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nI; i += gridDim.x * blockDim.x)
    {
        int a = l2i[i];
        if (a < 0 || a >= n) continue;
        uint64_t prefix = pre[a];
        unsigned pl     = prefixLength(prefix);
        unsigned level  = pl / 3;
        if (level >= kMaxLevel) continue;
        uint64_t cp     = encodePlaceholder(decodePlaceholder(prefix), int(pl) + 3);
        int lo = levelRange[level + 1], hi = levelRange[level + 2];
        int c = lowerBound(pre, lo, hi, cp);
        if (c != hi && pre[c] == cp)
        {
            child[a]             = c;
            parents[(c - 1) / 8] = a;
        }
    }
}

int linkOctree(Context* ctx, const uint64_t* leaves, int L, uint64_t* prefixes, int32_t* child, int32_t* parents,
               int32_t* levelRange, int32_t* i2l, int32_t* l2i, int32_t* levelRangeHost)
{
    if (L < 1) return setError(ctx, RYO_ERR_ARG, "octree: need at least one leaf");
    hipStream_t s = ctx->stream;
    int nI = (L - 1) / 7, n = L + nI;
    unsigned g = std::min(iceil(n, 256), unsigned(ctx->numCUs * 8));
    unsortedLayoutKernel<<<g, 256, 0, s>>>(leaves, L, nI, prefixes, i2l);
    size_t tb   = sortTempBytes(n);
    size_t need = n * (sizeof(uint64_t) + sizeof(int32_t)) + tb + 512;
    auto* buf   = static_cast<char*>(ctx->getScratch(kSlotLinkPairs, need));
    if (!buf) return RYO_ERR_HIP;
    auto* kb = reinterpret_cast<uint64_t*>(buf);
    auto* vb = reinterpret_cast<int32_t*>(buf + n * sizeof(uint64_t));
    void* tmp = buf + ((n * (sizeof(uint64_t) + sizeof(int32_t)) + 255) & ~size_t(255));
    int rc = sortPairs(ctx, prefixes, i2l, n, kb, vb, tmp, tb);
    if (rc) return rc;
    invertKernel<<<g, 256, 0, s>>>(i2l, l2i, n, nI);
    levelRangeKernel<<<1, 64, 0, s>>>(prefixes, n, levelRange);
    zeroI32Kernel<<<g, 256, 0, s>>>(child, n + 1);
    int np = std::max(1, (n - 1) / 8);
    zeroI32Kernel<<<iceil(np, 256), 256, 0, s>>>(parents, np);
    if (nI > 0) linkKernel<<<std::max(1u, std::min(iceil(nI, 256), unsigned(ctx->numCUs * 8))), 256, 0, s>>>(
        prefixes, n, nI, l2i, levelRange, child, parents);
    RYO_LAUNCH_CHECK(ctx);
    if (levelRangeHost)
    {
        RYO_HIP(ctx, hipMemcpyAsync(levelRangeHost, levelRange, (kMaxLevel + 2) * sizeof(int32_t),
                                    hipMemcpyDeviceToHost, s));
        RYO_HIP(ctx, hipStreamSynchronize(s));
    }
    return RYO_OK;
}

/* linkOctree for a device-sized build: launches and arrays for `capacity` leaves, the actual counts from stat
 * (buildLeavesDevice); no host round trip. */
int linkOctreeDevice(Context* ctx, const uint64_t* leaves, size_t capacity, const uint32_t* stat, uint64_t* prefixes,
                     int32_t* child, int32_t* parents, int32_t* levelRange, int32_t* i2l, int32_t* l2i)
{
    if (capacity < 1) return setError(ctx, RYO_ERR_ARG, "octree: need at least one leaf");
    hipStream_t s = ctx->stream;
    const int Lc = int(capacity), nI = (Lc - 1) / 7, n = Lc + nI;
    unsigned g = std::max(1u, std::min(iceil(n, 256), unsigned(ctx->numCUs * 8)));
    unsortedLayoutKernel<<<g, 256, 0, s>>>(leaves, Lc, nI, prefixes, i2l, stat);
    size_t tb   = sortTempBytes(n);
    size_t need = n * (sizeof(uint64_t) + sizeof(int32_t)) + tb + 512;
    auto* buf   = static_cast<char*>(ctx->getScratch(kSlotLinkPairs, need));
    if (!buf) return RYO_ERR_HIP;
    auto* kb  = reinterpret_cast<uint64_t*>(buf);
    auto* vb  = reinterpret_cast<int32_t*>(buf + n * sizeof(uint64_t));
    void* tmp = buf + ((n * (sizeof(uint64_t) + sizeof(int32_t)) + 255) & ~size_t(255));
    int rc    = sortPairsDigits(ctx, prefixes, i2l, n, kb, vb, tmp, tb, 8, stat + kStNodes);
    if (rc) return rc;
    invertKernel<<<g, 256, 0, s>>>(i2l, l2i, n, nI, stat);
    levelRangeKernel<<<1, 64, 0, s>>>(prefixes, n, levelRange, stat);
    zeroI32Kernel<<<g, 256, 0, s>>>(child, n + 1);
    int np = std::max(1, (n - 1) / 8);
    zeroI32Kernel<<<std::max(1u, std::min(iceil(np, 256), unsigned(ctx->numCUs * 8))), 256, 0, s>>>(parents, np);
    if (nI > 0)
        linkKernel<<<std::max(1u, std::min(iceil(nI, 256), unsigned(ctx->numCUs * 8))), 256, 0, s>>>(
            prefixes, n, nI, l2i, levelRange, child, parents, stat);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

/* ============================================================================ batches of systems */

// Each kernel below runs the single-system computation of one stage for every system of a batch in one launch:
// a block (or thread, or node) finds its system in the offset tables and works on that system's slice with
// local indices, so every system gets exactly the arrays it would get on its own.

// leaf tiles: block b is tile b - TS[s] of system s; the last tile of a system also counts its terminator slot
template<int MaxB>
__global__ __launch_bounds__(kLeafThreads) void batchLeafCountKernel(const uint64_t* __restrict__ sk, BatchView bv,
                                                                    int B, uint32_t* __restrict__ blockSums)
{
    __shared__ LeafTileSmem<MaxB> s;
    const int sys = segmentOf(bv.TS, bv.S, blockIdx.x);
    const uint32_t t = blockIdx.x - bv.TS[sys];
    const uint32_t p0 = bv.P[sys];
    int em[4];
    leafTile(sk + p0, long(bv.P[sys + 1] - p0), B, long(t) * kLeafTile, s, em);
    uint32_t v = em[0] + em[1] + em[2] + em[3], tot;
    blockExclusiveScan256(v, s.sw, &tot);
    if (threadIdx.x == 0) blockSums[blockIdx.x] = tot + (t + 1 == bv.TS[sys + 1] - bv.TS[sys] ? 1u : 0u);
}

template<int MaxB>
__global__ __launch_bounds__(kLeafThreads) void batchLeafEmitKernel(const uint64_t* __restrict__ sk, BatchView bv,
                                                                   int B, const uint32_t* __restrict__ blockOffsets,
                                                                   uint64_t* __restrict__ leaves,
                                                                   uint32_t* __restrict__ layout, long capacity)
{
    __shared__ LeafTileSmem<MaxB> s;
    const int sys = segmentOf(bv.TS, bv.S, blockIdx.x);
    const uint32_t t = blockIdx.x - bv.TS[sys];
    const uint32_t p0 = bv.P[sys];
    leafEmitTile(sk + p0, long(bv.P[sys + 1] - p0), B, long(t) * kLeafTile, blockOffsets[blockIdx.x], leaves, layout,
                 capacity, s);
}

// per-system leaf-slot offsets LO[s] = offset of the system's first tile (LO[S] = the total)
__global__ void batchLeafOffsetsKernel(BatchView bv, const uint32_t* __restrict__ blockOffsets, int numBlocks,
                                       uint32_t* __restrict__ LO)
{
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s <= bv.S; s += gridDim.x * blockDim.x)
        LO[s] = s < bv.S ? blockOffsets[bv.TS[s]] : blockOffsets[numBlocks];
}

// node, internal-node and parent-slot offsets of every system from its leaf-slot offsets (a system with L leaves
// has L + 1 slots, (L - 1) / 7 internal nodes, L + (L - 1) / 7 nodes and max(1, (nodes - 1) / 8) parent slots).
// *overflow = whether a total exceeds the capacity the arrays were sized for; on overflow every table is
// cleared, so each later stage sees empty trees and stays inside its arrays (the traversal then does nothing).
__global__ __launch_bounds__(256) void batchNodeTablesKernel(int S, uint32_t* __restrict__ LO, uint32_t* __restrict__ NO,
                                                             uint32_t* __restrict__ NIO, uint32_t* __restrict__ PO,
                                                             uint32_t capSlots, uint32_t capNodes, uint32_t capParents,
                                                             uint32_t* __restrict__ overflow)
{
    __shared__ uint32_t sw[4];
    __shared__ uint32_t over;
    uint32_t cN = 0, cI = 0, cP = 0;
    for (int s0 = 0; s0 < S; s0 += 256)
    {
        const int s      = s0 + int(threadIdx.x);
        const uint32_t L = s < S ? LO[s + 1] - LO[s] - 1 : 1u;
        uint32_t nI = s < S ? (L - 1) / 7 : 0u, nn = s < S ? L + nI : 0u;
        uint32_t np = s < S ? max(1u, (nn - 1) / 8) : 0u;
        uint32_t tN, tI, tP;
        const uint32_t eN = blockExclusiveScan256(nn, sw, &tN);
        const uint32_t eI = blockExclusiveScan256(nI, sw, &tI);
        const uint32_t eP = blockExclusiveScan256(np, sw, &tP);
        if (s < S) NO[s] = cN + eN, NIO[s] = cI + eI, PO[s] = cP + eP;
        cN += tN, cI += tI, cP += tP;
    }
    if (threadIdx.x == 0)
    {
        NO[S] = cN, NIO[S] = cI, PO[S] = cP;
        over      = (LO[S] > capSlots || cN > capNodes || cP > capParents) ? 1u : 0u;
        *overflow = over;
    }
    __syncthreads();
    if (over)
        for (int s = threadIdx.x; s <= S; s += blockDim.x)
            LO[s] = 0, NO[s] = 0, NIO[s] = 0, PO[s] = 0;
}

// particle counts per leaf slot (0 at the terminator slots)
__global__ void batchLeafCountsKernel(const uint32_t* __restrict__ layout, BatchView bv, uint32_t* __restrict__ counts)
{
    const uint32_t total = bv.LO[bv.S];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < total; j += gridDim.x * blockDim.x)
    {
        int s     = segmentOf(bv.LO, bv.S, j);
        counts[j] = j + 1 == bv.LO[s + 1] ? 0u : layout[j + 1] - layout[j];
    }
}

// unsorted node layout of every system; the values are global node slots (NO[s] + local unsorted index)
__global__ void batchUnsortedLayoutKernel(const uint64_t* __restrict__ leaves, BatchView bv,
                                          uint64_t* __restrict__ pre, int32_t* __restrict__ idx)
{
    const uint32_t total = bv.LO[bv.S];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < total; j += gridDim.x * blockDim.x)
    {
        const int s    = segmentOf(bv.LO, bv.S, j);
        const int L    = int(bv.LO[s + 1] - bv.LO[s]) - 1;
        const int t    = int(j - bv.LO[s]);
        if (t >= L) continue; // terminator slot
        const int nI   = int(bv.NIO[s + 1] - bv.NIO[s]);
        const int base = int(bv.NO[s]);
        const uint64_t* lv = leaves + bv.LO[s];
        uint64_t key = lv[t], next = lv[t + 1];
        unsigned lvl = treeLevel(next - key);
        pre[base + t + nI] = encodePlaceholder(key, 3 * lvl);
        idx[base + t + nI] = base + t + nI;
        int pl = commonPrefix(key, next);
        if (pl % 3 == 0 && t < L - 1)
        {
            unsigned level = unsigned(pl) / 3u;
            int w          = 0;
            for (unsigned l = 1; l <= level + 1; ++l)
                w += digitWeight(int(octalDigit(key, l)));
            int o = (t + w) / 7;
            if (o >= 0 && o < nI)
            {
                pre[base + o] = encodePlaceholder(key, pl);
                idx[base + o] = base + o;
            }
        }
    }
}

// sorted slot j of system s: l2i (local) and i2l (local, minus the internal count) from the global unsorted slot
__global__ void batchInvertKernel(const int32_t* __restrict__ slot, BatchView bv, int32_t* __restrict__ i2l,
                                  int32_t* __restrict__ l2i)
{
    const uint32_t total = bv.NO[bv.S];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < total; j += gridDim.x * blockDim.x)
    {
        const int s  = segmentOf(bv.NO, bv.S, j);
        const int nn = int(bv.NO[s + 1] - bv.NO[s]);
        const int nI = int(bv.NIO[s + 1] - bv.NIO[s]);
        const int u  = slot[j] - int(bv.NO[s]);
        if (u >= 0 && u < nn) l2i[bv.NO[s] + u] = int(j - bv.NO[s]);
        i2l[j] = u - nI;
    }
}

// level ranges of system blockIdx.x (local node indices)
__global__ void batchLevelRangeKernel(const uint64_t* __restrict__ pre, BatchView bv, int32_t* __restrict__ LR)
{
    const int s = blockIdx.x, l = threadIdx.x;
    const int nn = int(bv.NO[s + 1] - bv.NO[s]);
    const uint64_t* p = pre + bv.NO[s];
    if (l <= int(kMaxLevel)) LR[23 * s + l] = lowerBound(p, 0, nn, encodePlaceholder(0, 3 * l));
    if (l == int(kMaxLevel) + 1) LR[23 * s + l] = nn;
}

// per level l (block l): LP[(S+1) l + s] = number of level-l nodes in systems before s
__global__ __launch_bounds__(256) void batchLevelPrefixKernel(const int32_t* __restrict__ LR, int S,
                                                              uint32_t* __restrict__ LP)
{
    __shared__ uint32_t sw[4];
    const int l    = blockIdx.x;
    uint32_t carry = 0;
    for (int s0 = 0; s0 < S; s0 += 256)
    {
        int s       = s0 + threadIdx.x;
        uint32_t v  = s < S ? uint32_t(LR[23 * s + l + 1] - LR[23 * s + l]) : 0u, tot;
        uint32_t ex = blockExclusiveScan256(v, sw, &tot);
        if (s < S) LP[(S + 1) * l + s] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) LP[(S + 1) * l + S] = carry;
}

__global__ void batchLinkKernel(const uint64_t* __restrict__ pre, BatchView bv, const int32_t* __restrict__ l2i,
                                int32_t* __restrict__ child, int32_t* __restrict__ parents)
{
    const uint32_t total = bv.NIO[bv.S];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x)
    {
        const int s       = segmentOf(bv.NIO, bv.S, i);
        const int base    = int(bv.NO[s]);
        const int nn      = int(bv.NO[s + 1] - bv.NO[s]);
        const int32_t* lr = bv.LR + 23 * s;
        int a = l2i[base + int(i - bv.NIO[s])];
        if (a < 0 || a >= nn) continue;
        const uint64_t* p = pre + base;
        uint64_t prefix = p[a];
        unsigned pl     = prefixLength(prefix);
        unsigned level  = pl / 3;
        if (level >= kMaxLevel) continue;
        uint64_t cp = encodePlaceholder(decodePlaceholder(prefix), int(pl) + 3);
        int lo = lr[level + 1], hi = lr[level + 2];
        int c  = lowerBound(p, lo, hi, cp);
        if (c != hi && p[c] == cp)
        {
            child[base + a]                  = c;
            parents[bv.PO[s] + (c - 1) / 8] = a;
        }
    }
}

// fixed groups of every system: group g of system s covers its particles [64 (g - GO[s]), ...), as global
// particle indices, and remembers its system
__global__ void batchGroupsKernel(BatchView bv, unsigned gs, uint32_t* __restrict__ gStart,
                                  uint32_t* __restrict__ gEnd, int32_t* __restrict__ gSys)
{
    const uint32_t total = bv.GO[bv.S];
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x)
    {
        const int s = segmentOf(bv.GO, bv.S, g);
        uint32_t b  = bv.P[s] + (g - bv.GO[s]) * gs;
        gStart[g]   = b;
        gEnd[g]     = min(b + gs, bv.P[s + 1]);
        gSys[g]     = s;
    }
}

// host side of the batch stages (bv's device tables must hold what each stage reads)
int batchLeavesCount(Context* ctx, const uint64_t* sk, const BatchView& bv, int numTiles, int B,
                     uint32_t* blockSums, uint32_t* LO)
{
    if (B < 1 || B > kMaxBucket) return setError(ctx, RYO_ERR_UNSUPPORTED, "batch leaves: bucket size %d", B);
    hipStream_t s = ctx->stream;
    if (B <= kSmallBucket)
        batchLeafCountKernel<kSmallBucket><<<numTiles, kLeafThreads, 0, s>>>(sk, bv, B, blockSums);
    else
        batchLeafCountKernel<kMaxBucket><<<numTiles, kLeafThreads, 0, s>>>(sk, bv, B, blockSums);
    scanBlocksKernel<<<1, 256, 0, s>>>(blockSums, numTiles);
    batchLeafOffsetsKernel<<<iceil(bv.S + 1, 256), 256, 0, s>>>(bv, blockSums, numTiles, LO);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

int batchNodeTables(Context* ctx, const BatchView& bv, uint32_t capSlots, uint32_t capNodes, uint32_t capParents,
                    uint32_t* overflow)
{
    auto* t = const_cast<uint32_t*>(bv.LO);
    batchNodeTablesKernel<<<1, 256, 0, ctx->stream>>>(bv.S, t, const_cast<uint32_t*>(bv.NO),
                                                      const_cast<uint32_t*>(bv.NIO), const_cast<uint32_t*>(bv.PO),
                                                      capSlots, capNodes, capParents, overflow);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

int batchLeavesEmit(Context* ctx, const uint64_t* sk, const BatchView& bv, int numTiles, int B,
                    const uint32_t* blockSums, uint64_t* leaves, uint32_t* layout, uint32_t* counts,
                    uint32_t totalSlots)
{
    hipStream_t s = ctx->stream;
    if (B <= kSmallBucket)
        batchLeafEmitKernel<kSmallBucket><<<numTiles, kLeafThreads, 0, s>>>(sk, bv, B, blockSums, leaves, layout,
                                                                             long(totalSlots));
    else
        batchLeafEmitKernel<kMaxBucket><<<numTiles, kLeafThreads, 0, s>>>(sk, bv, B, blockSums, leaves, layout,
                                                                           long(totalSlots));
    batchLeafCountsKernel<<<std::max(1u, std::min(iceil(totalSlots, 256), unsigned(ctx->numCUs * 8))), 256, 0, s>>>(
        layout, bv, counts);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

int batchUnsortedLayout(Context* ctx, const uint64_t* leaves, const BatchView& bv, uint32_t totalSlots,
                        uint64_t* pre, int32_t* idx)
{
    batchUnsortedLayoutKernel<<<std::max(1u, std::min(iceil(totalSlots, 256), unsigned(ctx->numCUs * 8))), 256, 0,
                                ctx->stream>>>(leaves, bv, pre, idx);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

/* after the segmented sort of the node prefixes (slot = global unsorted slot of each sorted node): i2l, l2i,
 * level ranges LR (S x 23), per-level prefixes LP (22 x (S+1)), child links and parents */
int batchLinkFinish(Context* ctx, const uint64_t* pre, const int32_t* slot, const BatchView& bv, uint32_t totalNodes,
                    uint32_t totalInternal, uint32_t totalParents, int32_t* i2l, int32_t* l2i, int32_t* LR,
                    uint32_t* LP, int32_t* child, int32_t* parents)
{
    hipStream_t s = ctx->stream;
    const unsigned g = std::max(1u, std::min(iceil(totalNodes, 256), unsigned(ctx->numCUs * 8)));
    batchInvertKernel<<<g, 256, 0, s>>>(slot, bv, i2l, l2i);
    batchLevelRangeKernel<<<bv.S, 64, 0, s>>>(pre, bv, LR);
    batchLevelPrefixKernel<<<int(kMaxLevel) + 1, 256, 0, s>>>(LR, bv.S, LP);
    zeroI32Kernel<<<g, 256, 0, s>>>(child, int(totalNodes) + 1);
    zeroI32Kernel<<<std::max(1u, iceil(totalParents, 256)), 256, 0, s>>>(parents, int(std::max(1u, totalParents)));
    if (totalInternal > 0)
        batchLinkKernel<<<std::max(1u, std::min(iceil(totalInternal, 256), unsigned(ctx->numCUs * 8))), 256, 0, s>>>(
            pre, bv, l2i, child, parents);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

int batchGroups(Context* ctx, const BatchView& bv, uint32_t totalGroups, unsigned gs, uint32_t* gStart,
                uint32_t* gEnd, int32_t* gSys)
{
    if (totalGroups == 0) return RYO_OK;
    batchGroupsKernel<<<std::min(iceil(totalGroups, 256), unsigned(ctx->numCUs * 8)), 256, 0, ctx->stream>>>(
        bv, gs, gStart, gEnd, gSys);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

/* ============================================================================ fixed groups */

__global__ void fixedGroupsKernel(uint32_t first, uint32_t last, unsigned gs, uint32_t* groups, int numGroups)
{
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g <= numGroups; g += gridDim.x * blockDim.x)
        groups[g] = g == numGroups ? last : first + uint32_t(g) * gs;
}

int fixedGroups(Context* ctx, uint32_t first, uint32_t last, unsigned gs, uint32_t* groups, int* numGroups)
{
    if (gs == 0 || last < first) return setError(ctx, RYO_ERR_ARG, "groups: invalid range or group size");
    int G      = int((last - first + gs - 1) / gs);
    *numGroups = G;
    fixedGroupsKernel<<<iceil(G + 1, 256), 256, 0, ctx->stream>>>(first, last, gs, groups, G);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

} // namespace ryo
