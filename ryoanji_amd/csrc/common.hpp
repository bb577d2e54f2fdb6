// Shared device/host helpers of the gfx950 Barnes-Hut library: key arithmetic, wave64 primitives,
// error handling and the per-context state every entry point runs on.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/ryoanji_amd.h"

namespace ryo
{

constexpr unsigned kMaxLevel = 21;           // 64-bit keys: 21 octree levels, 63 key bits
constexpr uint64_t kKeyEnd   = 1ull << 63;   // end of the key range; also the "remove" key
constexpr int kWave          = 64;
// segmented sorts of at most this many segments sort each segment on its own (MSD rounds and LDS range sorts);
// more segments take a full sort and a second stable pass on the segment of each value
constexpr int kMaxSortSegments = 1 << 16;

/* ------------------------------------------------------------------ key arithmetic (host+device) */

__host__ __device__ inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }
__host__ __device__ inline int ctz64(uint64_t x) { return x ? __builtin_ctzll(x) : 64; }
__host__ __device__ inline uint64_t nodeRange(unsigned level) { return 1ull << (3u * (kMaxLevel - level)); }
// level of a node whose key range is `range` (a power of 8)
__host__ __device__ inline unsigned treeLevel(uint64_t range) { return unsigned(clz64(range - 1) - 1) / 3u; }
// number of equal leading key bits (the unused top bit excluded)
__host__ __device__ inline int commonPrefix(uint64_t a, uint64_t b) { return clz64(a ^ b) - 1; }
// Warren-Salmon placeholder-bit prefix of the node starting at `key` with `prefixLength` key bits
__host__ __device__ inline uint64_t encodePlaceholder(uint64_t key, int prefixLength)
{
    return (1ull << prefixLength) | (key >> (63 - prefixLength));
}
__host__ __device__ inline unsigned prefixLength(uint64_t code) { return 63u - unsigned(clz64(code)); }
__host__ __device__ inline uint64_t decodePlaceholder(uint64_t code)
{
    unsigned pl = prefixLength(code);
    return (code ^ (1ull << pl)) << (63 - pl);
}
__host__ __device__ inline unsigned octalDigit(uint64_t key, unsigned level)
{
    return unsigned(key >> (3u * (kMaxLevel - level))) & 7u;
}

/* ------------------------------------------------------------------ wave64 primitives */

__device__ inline unsigned laneId() { return __lane_id(); }
__device__ inline uint64_t lanemaskLt() { return (1ull << laneId()) - 1ull; }
__device__ inline uint64_t ballot(bool p) { return __ballot(p); }
__device__ inline int popc64(uint64_t m) { return __popcll(m); }

template<class T>
__device__ inline T readlane(T v, int lane);
template<>
__device__ inline int readlane(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
template<>
__device__ inline unsigned readlane(unsigned v, int lane) { return unsigned(__builtin_amdgcn_readlane(int(v), lane)); }
template<>
__device__ inline float readlane(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ inline float waveMin(float v)
{
    for (int o = 32; o >= 1; o >>= 1)
        v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ inline float waveMax(float v)
{
    for (int o = 32; o >= 1; o >>= 1)
        v = fmaxf(v, __shfl_xor(v, o));
    return v;
}
__device__ inline double waveSum(double v)
{
    for (int o = 32; o >= 1; o >>= 1)
        v += __shfl_xor(v, o);
    return v;
}
__device__ inline int waveInclusiveScan(int v)
{
    for (int o = 1; o < 64; o <<= 1)
    {
        int t = __shfl_up(v, o);
        if (int(laneId()) >= o) v += t;
    }
    return v;
}

// block-wide exclusive scan of one value per thread (256 threads = 4 waves); *total gets the block sum
__device__ inline uint32_t blockExclusiveScan256(uint32_t v, uint32_t* sWave /* [4] */, uint32_t* total)
{
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = v;
    for (int o = 1; o < 64; o <<= 1)
    {
        uint32_t t = __shfl_up(inc, o);
        if (lane >= o) inc += t;
    }
    if (lane == 63) sWave[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (int k = 0; k < 4; ++k)
    {
        if (k < w) pre += sWave[k];
        tot += sWave[k];
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

/* ------------------------------------------------------------------ batches of systems */

// Offset tables of a batch of S particle systems packed one after the other (all on the device). System s owns
// particles [P[s], P[s+1]), leaf slots [LO[s], LO[s+1]) (its L_s leaves plus the terminator slot), tree nodes
// [NO[s], NO[s+1]) of which NIO[s+1] - NIO[s] are internal, parent slots from PO[s], level ranges LR[23 s ..]
// (local node indices), per-level node prefixes LP[(S+1) l + s], target groups [GO[s], GO[s+1]) and leaf
// tiles [TS[s], TS[s+1]). Every index stored inside a system's arrays is local to that system.
struct BatchView
{
    int S;
    const uint32_t *P, *LO, *NO, *NIO, *PO, *LP, *GO, *TS;
    const int32_t* LR;
};

// the counts of a device-sized build (device, kStWords words): {L, nodes, internal nodes, L > capacity, layout[L]
// (n after an overflow), the leaf count the later stages use (L, or 0 when the build aborts), abort (an overflow or
// a layout not ending at n)}; every later stage of that build takes its sizes from here, within the capacity its
// launches and arrays are sized for
enum BuildStat : int
{
    kStL = 0, kStNodes, kStInternal, kStOverflow, kStLast, kStLc, kStAbort, kStWords
};

// the periodic images of one system of a batch: (2 shells + 1)^3 replicas of its box lengths L (shells 0: open)
struct SysImages
{
    int shells;
    float L[3];
};

// the box of one packed system that wraps periodically (periodic 0: open, positions not wrapped)
struct PackedWrap
{
    float lo[3], L[3];
    int periodic, pad;
};

// the segment holding element i: the largest s with off[s] <= i (empty segments are skipped)
__device__ inline int segmentOf(const uint32_t* off, int S, uint32_t i)
{
    int lo = 0, hi = S; // answer in [0, S)
    while (hi - lo > 1)
    {
        int mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

/* ------------------------------------------------------------------ host side */

struct Context
{
    hipStream_t stream = nullptr;
    int device         = 0;
    int numCUs         = 256;
    std::string lastError;

    // grow-only device scratch, by slot
    std::vector<void*> scratch;
    std::vector<size_t> scratchBytes;

    // traversal statistics: {sumP2P, maxP2P, sumM2P, maxM2P, maxStack, failedGroups}
    unsigned long long* dStats = nullptr;
    unsigned long long hStats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint64_t retriedGroups = 0;

    // traversal launch shape: LIFO entries per wave (LDS) and resident blocks per CU (0: defaults)
    int travStack       = 0;
    int travBlocksPerCU = 0;
    int travRanges      = 0; // group ranges with their own counters (0: one per XCD)
    int travWaves       = 0; // waves per traversal block (2, 4 or 8; 0: 4)

    // event timing of the last traversal kernel launch (ms), when enabled
    bool timing          = false;
    hipEvent_t ev[2]     = {nullptr, nullptr};
    double lastTraverseMs = 0;
    double lastTraverseFlops[2] = {0, 0};

    void* getScratch(int slot, size_t bytes);
};

enum ScratchSlot : int
{
    kSlotSortTiles = 0,
    kSlotSortHist,
    kSlotLeafBlocks,
    kSlotLinkPairs,
    kSlotLinkTmp,
    kSlotUpsweep,
    kSlotTraverse,
    kSlotPot,
    kSlotGroups,
    kSlotEwald,
    kSlotMisc,
    kSlotSplits,
    kSlotAccuracy,
    kSlotPack,
    kSlotSoft,
    kSlotStages,
    kNumSlots
};

int setError(Context* ctx, int code, const char* fmt, ...);

#define RYO_HIP(ctx, call)                                                                                             \
    do                                                                                                                 \
    {                                                                                                                  \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess)                                                                                          \
            return ::ryo::setError(ctx, RYO_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,                       \
                                   hipGetErrorString(e_));                                                             \
    } while (0)

#define RYO_LAUNCH_CHECK(ctx) RYO_HIP(ctx, hipGetLastError())

inline unsigned iceil(size_t a, size_t b) { return unsigned((a + b - 1) / b); }

} // namespace ryo

struct ryo_ctx
{
    ryo::Context c;
};
