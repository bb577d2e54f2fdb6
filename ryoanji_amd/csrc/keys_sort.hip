// SFC keys, the stable LSD radix sort of (key, index) pairs and the fused permutation (gfx950).
//
// Keys: cstone/include/cstone/sfc/sfc.hpp:153-207 (cell index with the fork's integer/fraction split) and
// hilbert.hpp:43-94 (3D Hilbert), morton.hpp:94-110. Sort: the semantics of cub::DeviceRadixSort::SortPairs
// as used by cstone::sortByKey (primitives_gpu.cu:310-358), i.e. std::stable_sort order on all 64 bits.
#include "common.hpp"
#include "sfc.hpp"

#include <cstdarg>

namespace ryo
{

/* ============================================================================ keys */

// Cell index floor((x - xmin) * m) with x*m and xmin*m rounded once each and split into integer and
// fraction (sfc.hpp:153-159). Contraction is off: fused into a*b - c the products would not be rounded, and a
// point whose exact product lies just below an integer would get the cell below the reference's.
__device__ __forceinline__ int cellIndex(float x, float xmin, float m)
{
#pragma clang fp contract(off)
    const float a = x * m, b = xmin * m;
    const float fa = floorf(a), fb = floorf(b);
    return int(fa - fb) - int(a - fa < b - fb);
}
__device__ __forceinline__ int cellIndex(double x, double xmin, double m)
{
#pragma clang fp contract(off)
    const double a = x * m, b = xmin * m;
    const double fa = floor(a), fb = floor(b);
    return int(fa - fb) - int(a - fa < b - fb);
}

__device__ __forceinline__ uint64_t spread3(unsigned v)
{
    uint64_t x = v & 0x1fffffu;
    x = (x | x << 32) & 0x001f00000000ffffull;
    x = (x | x << 16) & 0x001f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}


// sfc.hpp:161-207: cell indices clamped to 2^bits - 1, then the Hilbert (mixed-dimension when bits differ) or
// Morton key
__constant__ HilbertLut kHilbertLutDev = detail::makeHilbertLut();

// a freshly loaded system also writes, in the same pass, the identity permutation the sort starts from and one
// 32-byte record {x, y, z, m}, {h, o, 0, 0} per particle for the permutation gather
struct PackOut
{
    const float *m, *h, *o;
    int32_t* perm;
    float4* rec;
};

// KeyT: 64-bit keys (21 levels) or 32-bit keys (10 levels, HilbertKey<uint32_t> / MortonKey<uint32_t>); the
// removal sentinel is 2^(3 levels) (removeKey, tree/definitions.h:71-75)
template<class T, bool Cubic, bool Pack, class KeyT = uint64_t>
__global__ __launch_bounds__(256) void sfcKeysKernel(const T* __restrict__ x, const T* __restrict__ y,
                                                     const T* __restrict__ z, KeyT* __restrict__ keys,
                                                     size_t n, KeyBoxT<T> box, int kind, int skipRemoved, PackOut pk)
{
    static_assert(!Pack || sizeof(T) == 4, "the gather records hold float coordinates");
    static_assert(!Cubic || sizeof(KeyT) == 8, "the Hilbert tables encode 21 levels");
    constexpr KeyT kRemove = KeyT(1) << (sizeof(KeyT) == 8 ? 63 : 30);
    // cubic boxes: the Hilbert state tables in LDS (6.9 KB), read with per-lane indices
    __shared__ HilbertLut lut;
    if constexpr (Cubic)
    {
        static_assert(sizeof(HilbertLut) % 4 == 0, "tables are copied in words");
        auto* dst       = reinterpret_cast<uint32_t*>(&lut);
        const auto* src = reinterpret_cast<const uint32_t*>(&kHilbertLutDev);
        for (int w = threadIdx.x; w < int(sizeof(HilbertLut) / 4); w += blockDim.x)
            dst[w] = src[w];
        __syncthreads();
    }
    const int mx = int((1u << box.bits.b[0]) - 1), my = int((1u << box.bits.b[1]) - 1),
              mz = int((1u << box.bits.b[2]) - 1);
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    {
        if (!Pack && skipRemoved && keys[i] == kRemove) continue;
        const T px = x[i], py = y[i], pz = z[i];
        int ix = min(cellIndex(px, box.lo[0], box.m[0]), mx);
        int iy = min(cellIndex(py, box.lo[1], box.m[1]), my);
        int iz = min(cellIndex(pz, box.lo[2], box.m[2]), mz);
        uint64_t k;
        if (kind != 0)
            k = (spread3(ix) << 2) | (spread3(iy) << 1) | spread3(iz);
        else if constexpr (Cubic)
            k = hilbert3DLut(ix, iy, iz, lut.two, lut.one);
        else
            k = hilbertMixD(ix, iy, iz, box.bits);
        keys[i] = KeyT(k);
        if constexpr (Pack)
        {
            if (pk.perm) pk.perm[i] = int32_t(i);
            pk.rec[2 * i]     = make_float4(float(px), float(py), float(pz), pk.m[i]);
            pk.rec[2 * i + 1] = make_float4(pk.h[i], pk.o[i], 0.0f, 0.0f);
        }
    }
}

/* ============================================================================ radix sort */

constexpr int kSortThreads = 256;
constexpr int kSortItems   = 8;  // 2048-key tiles: 30 KB of LDS and 74 VGPRs in the scatter, 5 waves per SIMD
constexpr int kSortTile    = kSortThreads * kSortItems; // keys per tile
constexpr int kDigits      = 256;
constexpr int kPasses      = 8;
constexpr int kScanChunk   = 32; // tiles per chunk of the tile-offset scan
constexpr int kSmallScanTiles = 1024; // up to this many tiles the offsets take one launch

// element count of a sort: n, or the count held in device memory (at most n) when dN is given, so that a
// sequence whose sizes are produced on the device can run without a host round trip
__device__ inline size_t deviceCount(const uint32_t* dN, size_t n)
{
    return dN ? min(size_t(*dN), n) : n;
}

__global__ void zeroWordsKernel(uint32_t* a, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        a[i] = 0;
}

// all 8 digit histograms in one read of the keys
__global__ __launch_bounds__(256) void histAllKernel(const uint64_t* __restrict__ keys, size_t n,
                                                    const uint32_t* __restrict__ dN,
                                                    uint32_t* __restrict__ hist /* [8][256] */)
{
    __shared__ uint32_t h[kPasses * kDigits];
    n = deviceCount(dN, n);
    for (int i = threadIdx.x; i < kPasses * kDigits; i += blockDim.x)
        h[i] = 0;
    __syncthreads();
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    {
        uint64_t k = keys[i];
#pragma unroll
        for (int p = 0; p < kPasses; ++p)
            atomicAdd(&h[p * kDigits + ((k >> (8 * p)) & 255u)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kPasses * kDigits; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

// digit bases per pass: exclusive scan over digits of each pass's histogram
__global__ __launch_bounds__(256) void digitBaseKernel(const uint32_t* __restrict__ hist, uint32_t* __restrict__ base)
{
    __shared__ uint32_t sw[4];
    int p = blockIdx.x, d = threadIdx.x;
    uint32_t tot;
    base[p * kDigits + d] = blockExclusiveScan256(hist[p * kDigits + d], sw, &tot);
}

// The LSD passes to run: a pass whose digit is the same for every key (one digit bucket holds them all) leaves the
// order as it is and is skipped. plan[p] = 1 for the passes that run, plan[kPasses + p] = the buffer pass p reads
// (0: keys/vals, 1: keyBuf/valBuf), plan[2 kPasses] = the buffer holding the result. Node prefixes, whose high
// digits are zero, skip most passes; full-range keys skip none.
constexpr int kPlanWords = 2 * kPasses + 1;
__global__ __launch_bounds__(256) void lsdPlanKernel(const uint32_t* __restrict__ hist, size_t n,
                                                    const uint32_t* __restrict__ dN, int digits,
                                                    uint32_t* __restrict__ plan)
{
    __shared__ uint32_t constant[kPasses];
    const size_t count = deviceCount(dN, n);
    if (threadIdx.x < kPasses) constant[threadIdx.x] = 0;
    __syncthreads();
    for (int p = 0; p < kPasses; ++p)
        if (hist[p * kDigits + threadIdx.x] == count) constant[p] = 1; // also every pass of an empty sort
    __syncthreads();
    if (threadIdx.x == 0)
    {
        uint32_t buf = 0;
        for (int p = 0; p < kPasses; ++p)
        {
            const uint32_t run = p < digits && !constant[p];
            plan[p]           = run;
            plan[kPasses + p] = buf;
            buf ^= run;
        }
        plan[2 * kPasses] = buf;
    }
}

// per-tile digit counts of one pass, one contiguous row per tile: tileHist[tile * 256 + d] (a coalesced 1 KB write
// per tile, read back the same way by the offset scan and the scatter)
// With a plan (LSD passes), pass `shift / 8` reads keys or keys1 as the plan says and does nothing when skipped.
__global__ __launch_bounds__(kSortThreads) void tileHistKernel(const uint64_t* __restrict__ keys, size_t n,
                                                              const uint32_t* __restrict__ dN, int shift,
                                                              uint32_t* __restrict__ tileHist, int numTiles,
                                                              const uint64_t* __restrict__ keys1 = nullptr,
                                                              const uint32_t* __restrict__ plan = nullptr)
{
    if (plan)
    {
        if (!plan[shift / 8]) return;
        if (plan[kPasses + shift / 8]) keys = keys1;
    }
    __shared__ uint32_t h[kDigits];
    n = deviceCount(dN, n); // tiles past the count record empty histograms
    h[threadIdx.x] = 0;
    __syncthreads();
    size_t base = size_t(blockIdx.x) * kSortTile;
#pragma unroll 4
    for (int r = 0; r < kSortItems; ++r)
    {
        size_t i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    tileHist[size_t(blockIdx.x) * kDigits + threadIdx.x] = h[threadIdx.x];
}

// tile offsets in three coalesced steps: (1) the digit sums of each chunk of kScanChunk tiles, (2) per digit an
// exclusive scan over the chunks from the digit's base, (3) per chunk the running offsets of its tiles, written in
// place of the counts: tileHist[tile * 256 + d] = base[d] + (keys of digit d in earlier tiles)
__global__ __launch_bounds__(256) void tileChunkSumKernel(const uint32_t* __restrict__ tileHist, int numTiles,
                                                     uint32_t* __restrict__ chunkSum,
                                                     const uint32_t* __restrict__ run = nullptr)
{
    if (run && !*run) return; // a skipped pass
    const int c = blockIdx.x, d = threadIdx.x;
    const int t1 = min(numTiles, (c + 1) * kScanChunk);
    uint32_t sum = 0;
    for (int t = c * kScanChunk; t < t1; ++t)
        sum += tileHist[size_t(t) * kDigits + d];
    chunkSum[size_t(c) * kDigits + d] = sum;
}

// digitBase null: offsets from the start of the digit's run, and totals[d] = keys of digit d (the scatter then adds
// the runs' starts)
__global__ __launch_bounds__(256) void tileChunkScanKernel(uint32_t* __restrict__ chunkSum, int numChunks,
                                                      const uint32_t* __restrict__ digitBase,
                                                      uint32_t* __restrict__ totals,
                                                     const uint32_t* __restrict__ run = nullptr)
{
    if (run && !*run) return; // a skipped pass
    __shared__ uint32_t sw[4];
    const int d = blockIdx.x;
    uint32_t carry = digitBase ? digitBase[d] : 0u;
    for (int c0 = 0; c0 < numChunks; c0 += 256)
    {
        const int c = c0 + threadIdx.x;
        const uint32_t v = c < numChunks ? chunkSum[size_t(c) * kDigits + d] : 0u;
        uint32_t tot;
        const uint32_t ex = blockExclusiveScan256(v, sw, &tot);
        if (c < numChunks) chunkSum[size_t(c) * kDigits + d] = carry + ex;
        carry += tot;
    }
    if (totals && threadIdx.x == 0) totals[d] = carry;
}

// the same offsets in one launch for few tiles (one block per digit; its reads are strided, which costs little
// when there are few tiles and saves two launches per pass)
__global__ __launch_bounds__(256) void tileScanSmallKernel(uint32_t* __restrict__ tileHist, int numTiles,
                                                          const uint32_t* __restrict__ digitBase,
                                                          uint32_t* __restrict__ totals,
                                                     const uint32_t* __restrict__ run = nullptr)
{
    if (run && !*run) return; // a skipped pass
    __shared__ uint32_t sw[4];
    const int d = blockIdx.x;
    uint32_t carry = digitBase ? digitBase[d] : 0u;
    for (int t0 = 0; t0 < numTiles; t0 += 256)
    {
        const int t = t0 + threadIdx.x;
        const uint32_t v = t < numTiles ? tileHist[size_t(t) * kDigits + d] : 0u;
        uint32_t tot;
        const uint32_t ex = blockExclusiveScan256(v, sw, &tot);
        if (t < numTiles) tileHist[size_t(t) * kDigits + d] = carry + ex;
        carry += tot;
    }
    if (totals && threadIdx.x == 0) totals[d] = carry;
}

__global__ __launch_bounds__(256) void tileOffsetKernel(uint32_t* __restrict__ tileHist, int numTiles,
                                                       const uint32_t* __restrict__ chunkSum,
                                                     const uint32_t* __restrict__ run = nullptr)
{
    if (run && !*run) return; // a skipped pass
    const int c = blockIdx.x, d = threadIdx.x;
    const int t1 = min(numTiles, (c + 1) * kScanChunk);
    uint32_t carry = chunkSum[size_t(c) * kDigits + d];
    for (int t = c * kScanChunk; t < t1; ++t)
    {
        const uint32_t v = tileHist[size_t(t) * kDigits + d];
        tileHist[size_t(t) * kDigits + d] = carry;
        carry += v;
    }
}

// the tile a scatter block works on: blocks are dispatched round-robin over the 8 XCDs (block b on XCD b % 8), and
// each XCD takes a contiguous range of tiles in order, so the digit runs that consecutive tiles write next to each
// other (tile t's run of digit d ends where tile t+1's begins) meet in the same L2 and leave it as whole lines
__device__ inline int xcdTile(int b, int numTiles)
{
    const int x = b & 7, k = b >> 3;
    const int per = numTiles >> 3, extra = numTiles & 7;
    return x * per + min(x, extra) + k;
}

// The lanes of the wave that hold this lane's 8-bit digit d, among the lanes of the wave-uniform mask `valid`: one
// ballot per bit, after which each lane rules out the lanes whose bit differs from its own (the ballot xor the bit
// spread over a whole word, per 32-lane half). 4 VALU per bit on gfx950: the sign-extended bit field, the compare
// that makes the ballot from it, and one three-input bit operation per half (acc | (ballot ^ bit)). `below`: how
// many of the lanes sit below this one (its rank among them); `count`: how many there are. The highest of them is
// the lane with below + 1 == count.
struct WaveMatch
{
    uint32_t below, count;
};
__device__ __forceinline__ WaveMatch waveMatch(unsigned d, uint64_t valid)
{
    uint32_t lo = ~uint32_t(valid), hi = ~uint32_t(valid >> 32); // lanes ruled out
#pragma unroll
    for (int b = 0; b < 8; ++b)
    {
        const uint32_t bit = uint32_t(__builtin_amdgcn_sbfe(int(d), b, 1)); // all ones where bit b of d is set
        const uint64_t bb  = __builtin_amdgcn_uicmp(bit, 0u, 33);            // ballot of bit != 0 (33: ne)
        lo = __builtin_amdgcn_bitop3_b32(lo, uint32_t(bb), bit, 0xF6);       // lo | (ballot ^ bit)
        hi = __builtin_amdgcn_bitop3_b32(hi, uint32_t(bb >> 32), bit, 0xF6);
    }
    lo = ~lo, hi = ~hi;
    return {__builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u)), uint32_t(__popc(lo) + __popc(hi))};
}

// LDS of one tile's scatter
struct ScatterLds
{
    uint64_t sk[kSortTile];
    int32_t sv[kSortTile];
    uint32_t wcnt[4][kDigits];
    uint32_t tstart[kDigits];
    uint32_t gbase[kDigits];
    uint32_t sw[4];
};

// stable scatter of the nValid (<= kSortTile) pairs from `base` (valsIn null: each value is the pair's index, the
// identity permutation a sort of fresh keys starts from): wave-level match ranking (8 ballots per key),
// block-level offsets, staging in LDS so that the writes of one digit run are contiguous. offs[d] is where this
// tile's first pair of digit d goes; a destination at or past nOut (offsets that do not describe these keys) is
// not written and raises *err instead. Starts and ends with the block in step, so that a block can run it for
// one tile after another.
__device__ __forceinline__ void scatterTile(const uint64_t* __restrict__ keysIn, const int32_t* __restrict__ valsIn,
                                   uint64_t* __restrict__ keysOut, int32_t* __restrict__ valsOut, size_t base,
                                   int nValid, int shift, const uint32_t* __restrict__ offs,
                                   const uint32_t* __restrict__ digitBase, size_t nOut, uint32_t* err, ScatterLds& L)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int i = threadIdx.x; i < 4 * kDigits; i += kSortThreads)
        (&L.wcnt[0][0])[i] = 0;
    __syncthreads();

    uint64_t k[kSortItems];
    int32_t v[kSortItems];
    uint32_t rank[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r)
    {
        const int idx    = w * (kSortItems * 64) + r * 64 + lane;
        const bool valid = idx < nValid;
        k[r] = valid ? keysIn[base + idx] : 0;
        v[r] = valid ? (valsIn ? valsIn[base + idx] : int32_t(base + idx)) : 0; // no values: the index
        unsigned d = unsigned(k[r] >> shift) & 255u;
        const WaveMatch m = waveMatch(d, __ballot(valid));
        uint32_t pre = valid ? L.wcnt[w][d] : 0;
        rank[r] = pre + m.below;
        if (valid && m.below + 1 == m.count) L.wcnt[w][d] = pre + m.count;
    }
    __syncthreads();
    {
        int d = threadIdx.x;
        uint32_t c0 = L.wcnt[0][d], c1 = L.wcnt[1][d], c2 = L.wcnt[2][d], c3 = L.wcnt[3][d];
        uint32_t tot;
        uint32_t start = blockExclusiveScan256(c0 + c1 + c2 + c3, L.sw, &tot);
        L.wcnt[0][d] = 0;
        L.wcnt[1][d] = c0;
        L.wcnt[2][d] = c0 + c1;
        L.wcnt[3][d] = c0 + c1 + c2;
        L.tstart[d] = start;
        L.gbase[d]  = offs[d] + (digitBase ? digitBase[d] : 0u);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r)
    {
        const int idx = w * (kSortItems * 64) + r * 64 + lane;
        if (idx < nValid)
        {
            unsigned d = unsigned(k[r] >> shift) & 255u;
            uint32_t pos = L.tstart[d] + L.wcnt[w][d] + rank[r];
            L.sk[pos] = k[r];
            L.sv[pos] = v[r];
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < nValid; j += kSortThreads)
    {
        uint64_t key = L.sk[j];
        unsigned d = unsigned(key >> shift) & 255u;
        size_t out = size_t(L.gbase[d]) + (j - L.tstart[d]);
        if (out < nOut)
        {
            keysOut[out] = key;
            valsOut[out] = L.sv[j];
        }
        else
            atomicOr(err, 1u);
    }
    __syncthreads();
}

__global__ __launch_bounds__(kSortThreads) void scatterKernel(uint64_t* keysIn, int32_t* valsIn, uint64_t* keysOut,
                                                             int32_t* valsOut, size_t n,
                                                             const uint32_t* __restrict__ dN, int shift,
                                                             const uint32_t* __restrict__ tileOffsets, int numTiles,
                                                             const uint32_t* __restrict__ digitBase, uint32_t* err,
                                                             const uint32_t* __restrict__ plan = nullptr)
{
    // with a plan (LSD passes) the two buffers are (keysIn, valsIn) and (keysOut, valsOut), and the plan says which
    // one pass `shift / 8` reads, or that it is skipped
    if (plan)
    {
        if (!plan[shift / 8]) return;
        if (plan[kPasses + shift / 8])
        {
            uint64_t* k = keysIn;
            int32_t* v  = valsIn;
            keysIn = keysOut, valsIn = valsOut;
            keysOut = k, valsOut = v;
        }
    }
    n = deviceCount(dN, n);
    const int tile = xcdTile(blockIdx.x, numTiles);
    if (size_t(tile) * kSortTile >= n) return; // a tile past the count (the whole block leaves)
    __shared__ ScatterLds L;
    const size_t base = size_t(tile) * kSortTile;
    const int nValid  = int(n - base < size_t(kSortTile) ? n - base : size_t(kSortTile));
    scatterTile(keysIn, valsIn, keysOut, valsOut, base, nValid, shift, tileOffsets + size_t(tile) * kDigits, digitBase,
                n, err, L);
}

// Sort of at most ITEMS * THREADS pairs inside one block on their lowest `digits` 8-bit digits: the pairs stay in
// registers (ITEMS per thread) and LDS for every pass; the wave-level match ranking of scatterTile gives each
// wave's digit counts, whose sums are the pass's histogram, and the stable scatter goes to LDS, from where each
// thread reloads its items. One global read and one write in all. Item r of thread (w, lane) is element
// w * 64 ITEMS + 64 r + lane (wave order, then item order, then lanes), so waves past the count skip their work.
//
// With more than kFastDigits digits only the top kFastDigits are sorted first. Pairs whose keys differ in those
// bits are then in their final order; the runs of keys equal in them (rare: the keys of one range already share
// every digit above) are sorted on the whole key by insertion, stable, by the thread owning the run's first pair.
// A run longer than kTieRunMax sends the whole range through every digit instead, from the pairs as they were read
// (nothing is written before the end). Reads kin/vin, writes kout/vout (which may be the same arrays). Starts and
// ends with the block in step.
constexpr int kFastDigits  = 3;
constexpr int kTieRunMax   = 32;
constexpr int kLongRunsMax = 8; // more runs longer than kTieRunMax: the whole range goes through every digit

template<int THREADS, int ITEMS>
struct RangeSortLds
{
    static constexpr int kWaves = THREADS / 64;
    uint64_t sk[THREADS * ITEMS];
    int32_t sv[THREADS * ITEMS];
    uint32_t wcnt[kWaves][kDigits]; // per-wave digit counts, then their prefixes over waves
    uint32_t offs[kDigits];         // digit totals, then their exclusive scan
    uint32_t numLong;               // runs longer than kTieRunMax after the top digits, and where they start
    uint32_t longStart[kLongRunsMax];
    int runEnd;
};

// stable LSD passes on digits [d0, d1) (d1 > d0) of the items in registers (read from global memory); the sorted
// range is in L.sk/L.sv afterwards
template<int THREADS, int ITEMS>
__device__ __forceinline__ void rangePasses(uint64_t (&k)[ITEMS], int32_t (&v)[ITEMS], int n, int d0, int d1,
                                            RangeSortLds<THREADS, ITEMS>& L, int base = 0)
{
    uint64_t* sk = L.sk + base; // the n pairs sorted here sit at [base, base + n) of the LDS arrays
    int32_t* sv  = L.sv + base;
    constexpr int kWaves = THREADS / 64;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int p = d0; p < d1; ++p)
    {
        const int shift = 8 * p;
        if (p > d0)
        {
#pragma unroll
            for (int r = 0; r < ITEMS; ++r)
            {
                const int idx = w * (ITEMS * 64) + r * 64 + lane;
                if (idx < n) k[r] = sk[idx], v[r] = sv[idx];
            }
        }
        for (int i = threadIdx.x; i < kWaves * kDigits; i += THREADS)
            (&L.wcnt[0][0])[i] = 0;
        __syncthreads();
        uint32_t rank[ITEMS];
        unsigned dig[ITEMS];
#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
        {
            rank[r] = 0;
            dig[r]  = unsigned(k[r] >> shift) & 255u;
            if (w * (ITEMS * 64) + r * 64 >= n) continue; // the whole wave: nothing in this slot
            const bool valid  = w * (ITEMS * 64) + r * 64 + lane < n;
            const unsigned d  = dig[r];
            const WaveMatch m = waveMatch(d, __ballot(valid));
            const uint32_t pre = valid ? L.wcnt[w][d] : 0;
            rank[r] = pre + m.below;
            if (valid && m.below + 1 == m.count) L.wcnt[w][d] = pre + m.count;
        }
        __syncthreads();
        for (int d = threadIdx.x; d < kDigits; d += THREADS)
        {
            uint32_t c = 0;
            for (int ww = 0; ww < kWaves; ++ww)
            {
                const uint32_t x = L.wcnt[ww][d];
                L.wcnt[ww][d] = c;
                c += x;
            }
            L.offs[d] = c;
        }
        __syncthreads();
        if (w == 0)
        {
            // exclusive scan of the 256 digit totals, 4 per lane
            const uint32_t c0 = L.offs[4 * lane], c1 = L.offs[4 * lane + 1], c2 = L.offs[4 * lane + 2],
                           c3 = L.offs[4 * lane + 3];
            const uint32_t sum = c0 + c1 + c2 + c3;
            uint32_t inc = sum;
            for (int o = 1; o < 64; o <<= 1)
            {
                const uint32_t t = __shfl_up(inc, o);
                if (lane >= o) inc += t;
            }
            const uint32_t ex = inc - sum;
            L.offs[4 * lane] = ex, L.offs[4 * lane + 1] = ex + c0, L.offs[4 * lane + 2] = ex + c0 + c1;
            L.offs[4 * lane + 3] = ex + c0 + c1 + c2;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
            if (w * (ITEMS * 64) + r * 64 + lane < n)
            {
                const unsigned d  = dig[r];
                const uint32_t at = L.offs[d] + L.wcnt[w][d] + rank[r];
                sk[at] = k[r];
                sv[at] = v[r];
            }
        __syncthreads();
    }
}

// the runs of pairs in L.sk/L.sv whose keys are equal in bits [sh, 64), sorted on the whole key by insertion
// (stable) by the thread owning each run's first pair; a run longer than kTieRunMax is left as it is and its start
// recorded in longStart (the first kLongRunsMax of them; *numLong counts them all). Keys in a run share those bits,
// so a run's neighbours stay recognisable while other runs are being sorted.
template<int THREADS>
__device__ __forceinline__ void sortTieRuns(uint64_t* sk, int32_t* sv, int n, int sh, uint32_t* numLong,
                                            uint32_t* longStart)
{
#pragma unroll 1
    for (int idx = threadIdx.x; idx + 1 < n; idx += THREADS)
    {
        const uint64_t pfx = sk[idx] >> sh;
        if ((sk[idx + 1] >> sh) != pfx || (idx > 0 && (sk[idx - 1] >> sh) == pfx)) continue;
        int end = idx + 2;
        while (end < n && end - idx <= kTieRunMax && (sk[end] >> sh) == pfx)
            ++end;
        if (end - idx > kTieRunMax)
        {
            const uint32_t i = atomicAdd(numLong, 1u);
            if (i < uint32_t(kLongRunsMax)) longStart[i] = uint32_t(idx);
            continue;
        }
        for (int i = idx + 1; i < end; ++i)
        {
            const uint64_t key = sk[i];
            const int32_t val  = sv[i];
            int j              = i;
            while (j > idx && sk[j - 1] > key)
            {
                sk[j] = sk[j - 1];
                sv[j] = sv[j - 1];
                --j;
            }
            sk[j] = key;
            sv[j] = val;
        }
    }
}

// the end of the run of keys equal to sk[s] in bits [sh, 64) that starts at s (the whole block, which returns it
// to every thread): chunks of THREADS pairs, each thread testing one
template<int THREADS, int ITEMS>
__device__ __forceinline__ int longRunEnd(int s, int n, int sh, RangeSortLds<THREADS, ITEMS>& L)
{
    const uint64_t pfx = L.sk[s] >> sh;
    for (int c = s + 1; c < n; c += THREADS)
    {
        if (threadIdx.x == 0) L.runEnd = n;
        __syncthreads();
        const int i = c + int(threadIdx.x);
        if (i < n && (L.sk[i] >> sh) != pfx) atomicMin(&L.runEnd, i);
        __syncthreads();
        const int e = L.runEnd;
        __syncthreads();
        if (e < n) return e;
    }
    return n;
}

template<int THREADS, int ITEMS>
__device__ __forceinline__ void sortRangeLds(const uint64_t* kin, const int32_t* vin, uint64_t* kout, int32_t* vout,
                                             int n, int digits, RangeSortLds<THREADS, ITEMS>& L)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (digits == 0)
    {
        for (int i = threadIdx.x; i < n; i += THREADS)
            kout[i] = kin[i], vout[i] = vin[i];
        return;
    }
    int d0 = digits > kFastDigits ? digits - kFastDigits : 0;
    for (;;)
    {
        uint64_t k[ITEMS];
        int32_t v[ITEMS];
#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
        {
            const int idx = w * (ITEMS * 64) + r * 64 + lane;
            k[r] = idx < n ? kin[idx] : 0;
            v[r] = idx < n ? vin[idx] : 0;
        }
        if (threadIdx.x == 0) L.numLong = 0;
        rangePasses<THREADS, ITEMS>(k, v, n, d0, digits, L);
        if (d0 == 0) break;
        // the top digits only: the runs of keys equal in them are sorted on the lower digits, short ones by
        // insertion, the few long ones each by the block in LDS; many long runs: every digit, from the start
        sortTieRuns<THREADS>(L.sk, L.sv, n, 8 * d0, &L.numLong, L.longStart);
        __syncthreads();
        const uint32_t numLong = L.numLong;
        if (numLong > uint32_t(kLongRunsMax))
        {
            __syncthreads();
            d0 = 0;
            continue;
        }
        for (uint32_t r = 0; r < numLong; ++r)
        {
            const int s0  = int(L.longStart[r]);
            const int len = longRunEnd<THREADS, ITEMS>(s0, n, 8 * d0, L) - s0;
#pragma unroll
            for (int q = 0; q < ITEMS; ++q)
            {
                const int idx = w * (ITEMS * 64) + q * 64 + lane;
                k[q] = idx < len ? L.sk[s0 + idx] : 0;
                v[q] = idx < len ? L.sv[s0 + idx] : 0;
            }
            rangePasses<THREADS, ITEMS>(k, v, len, 0, d0, L, s0);
            __syncthreads();
        }
        break;
    }
    __syncthreads(); // the sorted range is in LDS; kout (possibly kin) is written after every read of kin
    for (int i = threadIdx.x; i < n; i += THREADS)
        kout[i] = L.sk[i], vout[i] = L.sv[i];
    __syncthreads();
}

// sorts of at most kBlockSortMax pairs in one launch: one block of 1024 threads runs every pass (one launch instead
// of three per pass plus three, which is most of a small sort's time)
constexpr int kBlockSortThreads = 1024;
constexpr int kBlockSortItems   = 8;
constexpr size_t kBlockSortMax  = size_t(kBlockSortThreads) * kBlockSortItems;

__global__ __launch_bounds__(kBlockSortThreads) void blockSortKernel(uint64_t* __restrict__ keys,
                                                                     int32_t* __restrict__ vals, size_t n,
                                                                     const uint32_t* __restrict__ dN, int digits)
{
    __shared__ RangeSortLds<kBlockSortThreads, kBlockSortItems> L;
    n = deviceCount(dN, n);
    sortRangeLds<kBlockSortThreads, kBlockSortItems>(keys, vals, keys, vals, int(n), digits, L);
}

/* ---------------------------------------------------------------------------- hybrid MSD sort of large inputs
 *
 * A full 64-bit sort of more than kBlockSortMax pairs, stable (the order of std::stable_sort on the keys):
 *  round 0: one stable pass on the top digit (bits 56-63) with the tile-histogram kernels above, into the
 *           second buffer. Every bucket of equal top digits (a "child") is then classified:
 *  - a child of more than kLocalMax pairs becomes a segment of the next round, which makes a stable pass on the
 *    next lower digit inside each segment (tiles never straddle segments, offsets scanned per segment), from one
 *    buffer into the other; rounds 1..7 repeat this on digits 6..0 for whatever is still large;
 *  - smaller children are finished in LDS by one block (sortRangeLds) on their remaining lower digits, read from
 *    the buffer the round wrote and written to the first buffer. Runs of consecutive small children are packed
 *    into ranges of at most kPackMax pairs, which are then sorted on one more digit (the one that separates
 *    the packed children: they are in order of it already, and a stable sort on it keeps that).
 * Each pass is stable and the passes go from the most significant digit down, so equal keys keep their input order
 * throughout. Uniform 2^26 keys take round 0, round 1 and one LDS sort of ~2048-pair buckets: 3 reads and writes
 * of the pairs plus 2 reads of the keys, against 8 of each for the LSD passes. Rounds with no large segment left
 * cost three launches that exit at once; every count lives in device memory, so there is no host round trip.
 */
constexpr int kLocalMax  = 8192; // children up to this size are finished in LDS
constexpr int kPackMax   = 4096; // runs of small children are packed into ranges up to this size
constexpr int kMsdRounds = 8;    // round r splits on digit 7 - r
constexpr unsigned kMsdScatterPerCU = 5; // MSD round scatter blocks per CU (the scatter's LDS fits 5)
// ranges by size class: up to 1024 pairs (128 threads x 8), 4096 (512 x 8), 8192 (1024 x 8)
constexpr int kNumRangeClasses = 3;
__device__ inline int rangeClass(uint32_t len) { return len <= 1024 ? 0 : len <= 4096 ? 1 : 2; }

struct MsdSeg
{
    uint32_t start, len, tileStart, pad;
};

// device counters of an MSD sort (words)
enum MsdCtl : int
{
    kMsdBig   = 0,                      // [8]: large segments of round r
    kMsdTiles = kMsdBig + kMsdRounds,   // [8]: tiles of round r
    kMsdRange = kMsdTiles + kMsdRounds, // [4]: ranges of class c
    kMsdErr   = kMsdRange + kNumRangeClasses,
    kMsdCtlWords = 32
};

struct MsdLists
{
    uint32_t* ctl;
    MsdSeg* segs[2];     // large segments of odd / even rounds
    uint32_t* owner[2];  // segment of each tile, odd / even rounds
    uint2* ranges[kNumRangeClasses]; // {start, len | digits << 16 | source buffer << 24}
    uint32_t maxSegs, maxTiles, maxRanges;
};

__global__ void msdInitKernel(uint32_t* ctl, uint32_t* hist)
{
    for (int i = threadIdx.x; i < kMsdCtlWords; i += blockDim.x)
        ctl[i] = 0;
    for (int i = threadIdx.x; i < kDigits; i += blockDim.x)
        hist[i] = 0;
}

// Children of one segment after its pass on the digit of round r (cnt[d], start[d] in LDS, at least 64 threads):
// large ones become segments of round r + 1 (with their tiles), runs of small ones ranges for the LDS sorts. Wave 0
// walks the 256 children in order; one reservation per list, then the block writes the entries.
__device__ __forceinline__ void emitChildren(const uint32_t* cnt, const uint32_t* start, int r, MsdLists L)
{
    const int rem = kMsdRounds - 1 - r; // digits below the one this round split on
    const uint32_t src = (r & 1) ? 0u : 1u; // round r wrote the second buffer when r is even
    __shared__ uint2 packs[kDigits];
    __shared__ MsdSeg bigs[kDigits];
    __shared__ int nPack, nBig;
    __shared__ uint32_t bigTiles, segBase, tileBase, rangeBase[kNumRangeClasses], classOf[kDigits],
        classIdx[kDigits];
    if (threadIdx.x < 64)
    {
        // wave 0 walks the children in order: lane l holds the counts and starts of children l, l + 64, l + 128 and
        // l + 192, which the walk reads with readlane, so its state and branches are wave-uniform (scalar) and
        // only lane 0 writes the lists (one LDS round trip per child in a one-thread walk cost ~40 us per call)
        const int lane = int(threadIdx.x);
        uint32_t cv[4], sv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
            cv[j] = cnt[64 * j + lane], sv[j] = start[64 * j + lane];
        int np = 0, nb = 0;
        uint32_t tiles = 0, pStart = 0, pLen = 0, pKids = 0;
        uint32_t nc[kNumRangeClasses] = {};
        auto put = [&](uint32_t s, uint32_t len, uint32_t digits) {
            // a range already in the first buffer with nothing left to sort needs no work
            if (len == 0 || (digits == 0 && src == 0)) return;
            const int c = rangeClass(len);
            if (lane == 0)
            {
                packs[np]    = make_uint2(s, len | (digits << 16) | (src << 24));
                classOf[np]  = c;
                classIdx[np] = nc[c];
            }
            ++nc[c];
            ++np;
        };
        auto flush = [&]() {
            if (pLen) put(pStart, pLen, rem == 0 ? 0u : uint32_t(rem) + (pKids > 1 ? 1u : 0u));
            pLen = 0, pKids = 0;
        };
#pragma unroll
        for (int j = 0; j < 4; ++j)
            for (int l = 0; l < 64; ++l)
            {
                const uint32_t c = uint32_t(__builtin_amdgcn_readlane(int(cv[j]), l));
                if (c == 0) continue;
                const uint32_t st = uint32_t(__builtin_amdgcn_readlane(int(sv[j]), l));
                if (rem > 0 && c > uint32_t(kLocalMax))
                {
                    flush();
                    if (lane == 0) bigs[nb] = MsdSeg{st, c, tiles, 0};
                    ++nb;
                    tiles += (c + kSortTile - 1) / kSortTile;
                }
                else if (c > uint32_t(kPackMax))
                {
                    flush();
                    // only a child with no digits left can be larger than kLocalMax: it is sorted already
                    put(st, c > uint32_t(kLocalMax) ? 0u : c, uint32_t(rem));
                }
                else
                {
                    if (pLen + c > uint32_t(kPackMax)) flush();
                    if (pLen == 0) pStart = st;
                    pLen += c;
                    ++pKids;
                }
            }
        flush();
        if (lane == 0)
        {
            // the reservations in flight together (an add of zero only reads the counter)
            nPack = np, nBig = nb, bigTiles = tiles;
            uint32_t base[kNumRangeClasses];
#pragma unroll
            for (int c = 0; c < kNumRangeClasses; ++c)
                base[c] = atomicAdd(&L.ctl[kMsdRange + c], nc[c]);
            const uint32_t sb = atomicAdd(&L.ctl[kMsdBig + r + 1], uint32_t(nb));
            const uint32_t tb = atomicAdd(&L.ctl[kMsdTiles + r + 1], tiles);
#pragma unroll
            for (int c = 0; c < kNumRangeClasses; ++c)
                rangeBase[c] = base[c];
            segBase = sb, tileBase = tb;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nPack; i += blockDim.x)
    {
        const uint32_t at = rangeBase[classOf[i]] + classIdx[i];
        if (at < L.maxRanges) L.ranges[classOf[i]][at] = packs[i];
        else atomicOr(&L.ctl[kMsdErr], 2u);
    }
    if (nBig)
    {
        const int nxt = (r + 1) & 1;
        if (segBase + nBig > L.maxSegs || tileBase + bigTiles > L.maxTiles)
        {
            if (threadIdx.x == 0) atomicOr(&L.ctl[kMsdErr], 4u);
        }
        else
        {
            for (int i = threadIdx.x; i < nBig; i += blockDim.x)
            {
                MsdSeg g = bigs[i];
                g.tileStart += tileBase;
                L.segs[nxt][segBase + i] = g;
            }
            // one wave per segment at a time
            const int waves = int(blockDim.x >> 6), w = int(threadIdx.x >> 6), lane = int(threadIdx.x & 63);
            for (int i = w; i < nBig; i += waves)
            {
                const uint32_t t0 = tileBase + bigs[i].tileStart;
                const uint32_t nt = (bigs[i].len + kSortTile - 1) / kSortTile;
                for (uint32_t t = uint32_t(lane); t < nt; t += 64u)
                    L.owner[nxt][t0 + t] = segBase + i;
            }
        }
    }
    __syncthreads();
}

// round 0's children: the buckets of the top digit, from its histogram and their exclusive scan
__global__ __launch_bounds__(256) void msdRootKernel(const uint32_t* __restrict__ hist,
                                                    const uint32_t* __restrict__ base, MsdLists L)
{
    __shared__ uint32_t cnt[kDigits], st[kDigits];
    cnt[threadIdx.x] = hist[threadIdx.x];
    st[threadIdx.x]  = base[threadIdx.x];
    __syncthreads();
    emitChildren(cnt, st, 0, L);
}

// the tiles of round r's segments, spread over the grid so that each XCD takes a contiguous range of tiles in
// order (as xcdTile does for the one-pass kernels); the grid is a multiple of 8 blocks
#define MSD_FOR_EACH_TILE(t, T)                                                                                       \
    for (uint32_t t_per_ = (T) >> 3, t_x_ = blockIdx.x & 7u, t_k_ = blockIdx.x >> 3, t_g_ = gridDim.x >> 3,             \
                  t_end_ = t_x_ * t_per_ + min(t_x_, (T) & 7u) + t_per_ + (t_x_ < ((T) & 7u) ? 1u : 0u),               \
                  t = t_x_ * t_per_ + min(t_x_, (T) & 7u) + t_k_;                                                       \
         t < t_end_; t += t_g_)

// Start of a segmented sort: segment s holds pairs [off[s], off[s+1]) (within the count) and is sorted on its own.
// A segment of more than kLocalMax pairs is one of round 0's segments, a smaller one a range sorted in LDS on all 8
// digits in place; different segments never meet in a tile or a range, so each one's order is that of a sort of
// the segment alone. One block; every counter of the sort starts here.
__global__ __launch_bounds__(256) void msdSegInitKernel(const uint32_t* __restrict__ off, int S, size_t n,
                                                       const uint32_t* __restrict__ dN, MsdLists L)
{
    __shared__ uint32_t nBig, nTiles, nRange[kNumRangeClasses];
    if (threadIdx.x == 0)
    {
        nBig = 0, nTiles = 0;
        for (int c = 0; c < kNumRangeClasses; ++c)
            nRange[c] = 0;
    }
    for (int i = threadIdx.x; i < kMsdCtlWords; i += blockDim.x)
        L.ctl[i] = 0;
    __syncthreads();
    const size_t count = deviceCount(dN, n);
    for (int s = threadIdx.x; s < S; s += blockDim.x)
    {
        const uint32_t b = uint32_t(min(size_t(off[s + 1]), count)), a = min(off[s], b);
        const uint32_t len = b - a;
        if (len > uint32_t(kLocalMax))
        {
            const uint32_t nt = (len + kSortTile - 1) / kSortTile;
            const uint32_t i = atomicAdd(&nBig, 1u), t0 = atomicAdd(&nTiles, nt);
            if (i < L.maxSegs && t0 + nt <= L.maxTiles) L.segs[0][i] = MsdSeg{a, len, t0, 0};
            else
            {
                if (i < L.maxSegs) L.segs[0][i] = MsdSeg{a, 0, 0, 0}; // no tiles of its own (and an error)
                atomicOr(&L.ctl[kMsdErr], 64u);
            }
        }
        else if (len > 1)
        {
            const int c      = rangeClass(len);
            const uint32_t i = atomicAdd(&nRange[c], 1u);
            if (i < L.maxRanges) L.ranges[c][i] = make_uint2(a, len | (uint32_t(kMsdRounds) << 16));
            else atomicOr(&L.ctl[kMsdErr], 64u);
        }
    }
    __syncthreads();
    // the owners of the large segments' tiles, written by the whole block (a segment may have thousands of tiles)
    const uint32_t big = min(nBig, L.maxSegs);
    for (uint32_t i = 0; i < big; ++i)
    {
        const MsdSeg g    = L.segs[0][i];
        const uint32_t nt = (g.len + kSortTile - 1) / kSortTile;
        for (uint32_t t = threadIdx.x; t < nt; t += blockDim.x)
            if (g.tileStart + t < L.maxTiles) L.owner[0][g.tileStart + t] = i;
    }
    if (threadIdx.x == 0)
    {
        L.ctl[kMsdBig]   = big;
        L.ctl[kMsdTiles] = min(nTiles, L.maxTiles);
        for (int c = 0; c < kNumRangeClasses; ++c)
            L.ctl[kMsdRange + c] = min(nRange[c], L.maxRanges);
    }
}

// the segment of tile t of round r, checked against the lists and the element count (an entry that does not
// describe tile t raises the error word and is skipped, so no launch reads or writes past the arrays)
__device__ __forceinline__ bool msdTileSeg(const MsdLists& L, int r, uint32_t t, size_t n, MsdSeg& g, uint32_t& off,
                                           int& cnt)
{
    const uint32_t S = min(L.ctl[kMsdBig + r], L.maxSegs);
    const uint32_t o = L.owner[r & 1][t];
    if (o >= S) return false;
    g = L.segs[r & 1][o];
    if (size_t(g.start) + g.len > n || t < g.tileStart || size_t(t - g.tileStart) * kSortTile >= g.len) return false;
    off = (t - g.tileStart) * kSortTile;
    cnt = int(min(uint32_t(kSortTile), g.len - off));
    return true;
}

__global__ __launch_bounds__(256) void msdTileHistKernel(const uint64_t* __restrict__ keys, size_t n, MsdLists L,
                                                        int r, uint32_t* __restrict__ tileHist)
{
    __shared__ uint32_t h[kDigits];
    const uint32_t T = min(L.ctl[kMsdTiles + r], L.maxTiles);
    const int shift  = 8 * (kMsdRounds - 1 - r);
    MSD_FOR_EACH_TILE(t, T)
    {
        __syncthreads();
        h[threadIdx.x] = 0;
        MsdSeg g{};
        uint32_t off = 0;
        int cnt      = 0;
        if (!msdTileSeg(L, r, t, n, g, off, cnt))
        {
            if (threadIdx.x == 0) atomicOr(&L.ctl[kMsdErr], 16u);
            cnt = 0;
        }
        const uint64_t* k = keys + g.start + off;
        __syncthreads();
#pragma unroll 4
        for (int i = threadIdx.x; i < cnt; i += kSortThreads)
            atomicAdd(&h[(k[i] >> shift) & 255u], 1u);
        __syncthreads();
        tileHist[size_t(t) * kDigits + threadIdx.x] = h[threadIdx.x];
    }
}

// per segment of round r: the digit totals over its tiles, their exclusive scan from the segment's start, the
// running offsets of its tiles (in place of their counts), then its children. 1024 threads: thread (q, d) sums and
// then offsets digit d over the q-th quarter of the segment's tiles.
constexpr int kMsdScanThreads = 1024;
__global__ __launch_bounds__(kMsdScanThreads) void msdScanKernel(size_t n, MsdLists L, int r,
                                                                uint32_t* __restrict__ tileHist)
{
    __shared__ uint32_t cnt[kDigits], st[kDigits], part[4][kDigits];
    const uint32_t S   = min(L.ctl[kMsdBig + r], L.maxSegs);
    const MsdSeg* segs = L.segs[r & 1];
    const int d = threadIdx.x & 255, q = threadIdx.x >> 8, lane = threadIdx.x & 63;
    for (uint32_t s = blockIdx.x; s < S; s += gridDim.x)
    {
        const MsdSeg g    = segs[s];
        const uint32_t nt = (g.len + kSortTile - 1) / kSortTile;
        if (size_t(g.start) + g.len > n || size_t(g.tileStart) + nt > L.maxTiles)
        {
            if (threadIdx.x == 0) atomicOr(&L.ctl[kMsdErr], 32u);
            continue;
        }
        const uint32_t t0 = nt * q / 4, t1 = nt * (q + 1) / 4;
        uint32_t* th      = tileHist + size_t(g.tileStart) * kDigits + d;
        uint32_t sum      = 0;
#pragma unroll 8
        for (uint32_t t = t0; t < t1; ++t)
            sum += th[size_t(t) * kDigits];
        part[q][d] = sum;
        __syncthreads();
        if (q == 0) cnt[d] = part[0][d] + part[1][d] + part[2][d] + part[3][d];
        __syncthreads();
        if (threadIdx.x < 64)
        {
            // exclusive scan of the 256 digit totals from the segment's start, 4 per lane
            const uint32_t c0 = cnt[4 * lane], c1 = cnt[4 * lane + 1], c2 = cnt[4 * lane + 2], c3 = cnt[4 * lane + 3];
            const uint32_t tot = c0 + c1 + c2 + c3;
            uint32_t inc       = tot;
            for (int o = 1; o < 64; o <<= 1)
            {
                const uint32_t x = __shfl_up(inc, o);
                if (lane >= o) inc += x;
            }
            const uint32_t ex = g.start + inc - tot;
            st[4 * lane] = ex, st[4 * lane + 1] = ex + c0, st[4 * lane + 2] = ex + c0 + c1;
            st[4 * lane + 3] = ex + c0 + c1 + c2;
        }
        __syncthreads();
        uint32_t carry = st[d];
        for (int qq = 0; qq < q; ++qq)
            carry += part[qq][d];
        for (uint32_t t = t0; t < t1; ++t)
        {
            const uint32_t v = th[size_t(t) * kDigits];
            th[size_t(t) * kDigits] = carry;
            carry += v;
        }
        __syncthreads();
        emitChildren(cnt, st, r, L);
    }
}

__global__ __launch_bounds__(kSortThreads, 5) void msdScatterKernel(const uint64_t* __restrict__ keysIn,
                                                                const int32_t* __restrict__ valsIn,
                                                                uint64_t* __restrict__ keysOut,
                                                                int32_t* __restrict__ valsOut, size_t n, MsdLists L,
                                                                int r, const uint32_t* __restrict__ tileHist)
{
    __shared__ ScatterLds S;
    const uint32_t T = min(L.ctl[kMsdTiles + r], L.maxTiles);
    const int shift  = 8 * (kMsdRounds - 1 - r);
    MSD_FOR_EACH_TILE(t, T)
    {
        MsdSeg g{};
        uint32_t off = 0;
        int cnt      = 0;
        if (!msdTileSeg(L, r, t, n, g, off, cnt)) continue; // the same for the whole block
        scatterTile(keysIn, valsIn, keysOut, valsOut, size_t(g.start) + off, cnt, shift,
                    tileHist + size_t(t) * kDigits, nullptr, n, &L.ctl[kMsdErr], S);
    }
}

// the ranges of one class: each sorted in LDS on its digits from the buffer its round wrote into the first buffer
template<int THREADS, int ITEMS>
__global__ __launch_bounds__(THREADS) void msdLocalSortKernel(MsdLists L, int cls, uint64_t* __restrict__ k0,
                                                              int32_t* __restrict__ v0, const uint64_t* __restrict__ k1,
                                                              const int32_t* __restrict__ v1, size_t n)
{
    __shared__ RangeSortLds<THREADS, ITEMS> S;
    const uint32_t num = min(L.ctl[kMsdRange + cls], L.maxRanges);
    const uint2* ranges = L.ranges[cls];
    for (uint32_t i = blockIdx.x; i < num; i += gridDim.x)
    {
        const uint2 rg        = ranges[i];
        const uint32_t len    = rg.y & 0xffffu, digits = (rg.y >> 16) & 0xffu, src = rg.y >> 24;
        if (len > uint32_t(THREADS * ITEMS) || size_t(rg.x) + len > n || digits > 8u)
        {
            if (threadIdx.x == 0) atomicOr(&L.ctl[kMsdErr], 8u);
            continue;
        }
        const uint64_t* kin = src ? k1 : k0;
        const int32_t* vin  = src ? v1 : v0;
        sortRangeLds<THREADS, ITEMS>(kin + rg.x, vin + rg.x, k0 + rg.x, v0 + rg.x, int(len), int(digits), S);
    }
}

__global__ void copyPairsKernel(const uint64_t* __restrict__ kin, const int32_t* __restrict__ vin, size_t n,
                                const uint32_t* __restrict__ dN, uint64_t* __restrict__ kout,
                                int32_t* __restrict__ vout, const uint32_t* __restrict__ plan = nullptr)
{
    if (plan && !plan[2 * kPasses]) return; // the LSD passes left the result in place
    n = deviceCount(dN, n);
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        kout[i] = kin[i], vout[i] = vin[i];
}

__global__ void iotaKernel(int32_t* out, size_t n, int32_t start)
{
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        out[i] = start + int32_t(i);
}

template<class T>
struct GatherArgsT
{
    const T* in[10];
    T* out[10];
    int num;
};
using GatherArgs = GatherArgsT<float>;

template<class T>
__global__ __launch_bounds__(256) void gatherKernel(const int32_t* __restrict__ perm, size_t n, GatherArgsT<T> a)
{
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    {
        int32_t j = perm[i];
#pragma unroll
        for (int k = 0; k < 10; ++k)
            if (k < a.num) a.out[k][i] = a.in[k][j];
    }
}

// the six arrays of a freshly loaded system from the 32-byte records the keys pass wrote (sfcKeysPacked): the
// gather of a random permutation fetches one line per particle instead of one per array and particle
__global__ __launch_bounds__(256) void gatherRecordsKernel(const int32_t* __restrict__ perm, size_t n,
                                                           const float4* __restrict__ rec, float* __restrict__ x,
                                                           float* __restrict__ y, float* __restrict__ z,
                                                           float* __restrict__ m, float* __restrict__ h,
                                                           float* __restrict__ o)
{
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    {
        const size_t j = size_t(perm[i]);
        const float4 a = rec[2 * j], b = rec[2 * j + 1];
        x[i] = a.x, y[i] = a.y, z[i] = a.z, m[i] = a.w;
        h[i] = b.x, o[i] = b.y;
    }
}

static unsigned gridFor(size_t n, int threads, int maxBlocks)
{
    size_t b = (n + threads - 1) / threads;
    if (b > size_t(maxBlocks)) b = maxBlocks;
    return unsigned(b ? b : 1);
}

static size_t lsdTempBytes(size_t n)
{
    size_t numTiles  = (n + kSortTile - 1) / kSortTile;
    size_t numChunks = (numTiles + kScanChunk - 1) / kScanChunk;
    return 256 * sizeof(uint32_t) * (2 * kPasses) + size_t(kDigits) * (numTiles + numChunks) * sizeof(uint32_t) + 256;
}

// temp storage of the MSD sort: counters, top-digit histogram and bases, the tile histograms (round 0's tiles, or
// the most tiles a later round can have: segments longer than kLocalMax are disjoint, each adds at most one partial
// tile), round 0's chunk sums, two segment lists, two tile-owner arrays and the four range lists (every range but
// the last of a run of small children shares more than kPackMax pairs with the next one)
struct MsdLayout
{
    size_t ctl, hist, base, tileHist, chunkSum, segs[2], owner[2], ranges[kNumRangeClasses], bytes;
    uint32_t maxSegs, maxTiles, maxRanges, numTiles0, numChunks0;
};

static MsdLayout msdLayout(size_t n)
{
    MsdLayout l{};
    l.numTiles0  = uint32_t((n + kSortTile - 1) / kSortTile);
    l.numChunks0 = (l.numTiles0 + kScanChunk - 1) / kScanChunk;
    l.maxSegs    = uint32_t(n / (kLocalMax + 1) + 1);
    l.maxTiles   = l.numTiles0 + l.maxSegs;
    // ranges made by the rounds, plus one per segment of a segmented sort (at most kMaxSortSegments)
    l.maxRanges  = uint32_t(n / 256 + 64 + std::min(n / 2, size_t(kMaxSortSegments)));
    size_t at    = 0;
    auto take    = [&](size_t bytes) {
        const size_t r = at;
        at += (bytes + 255) & ~size_t(255);
        return r;
    };
    l.ctl      = take(kMsdCtlWords * 4);
    l.hist     = take(kDigits * 4);
    l.base     = take(kDigits * 4);
    l.tileHist = take(size_t(l.maxTiles) * kDigits * 4);
    l.chunkSum = take(size_t(l.numChunks0) * kDigits * 4);
    for (int i = 0; i < 2; ++i)
        l.segs[i] = take(size_t(l.maxSegs) * sizeof(MsdSeg));
    for (int i = 0; i < 2; ++i)
        l.owner[i] = take(size_t(l.maxTiles) * 4);
    for (int c = 0; c < kNumRangeClasses; ++c)
        l.ranges[c] = take(size_t(l.maxRanges) * sizeof(uint2));
    l.bytes = at;
    return l;
}

size_t sortTempBytes(size_t n) { return std::max(lsdTempBytes(n), msdLayout(n).bytes); }

// stable LSD sort on the lowest `digits` 8-bit digits of the keys; the result is in keys/vals. With dN, n is the
// capacity the launches are sized for and *dN (device memory) the number of elements sorted; the sequence then
// has no host dependence on the count and can be captured in a graph.
int sortPairsDigits(Context* ctx, uint64_t* keys, int32_t* vals, size_t n, uint64_t* keyBuf, int32_t* valBuf,
                    void* tmp, size_t tmpBytes, int digits, const uint32_t* dN)
{
    if (n == 0) return RYO_OK;
    if (digits < 1 || digits > kPasses) return setError(ctx, RYO_ERR_ARG, "sort: %d digits", digits);
    if (tmpBytes < sortTempBytes(n))
        return setError(ctx, RYO_ERR_ARG, "sort: temp storage too small (%zu < %zu)", tmpBytes, sortTempBytes(n));
    hipStream_t s = ctx->stream;
    if (n <= kBlockSortMax)
    {
        blockSortKernel<<<1, kBlockSortThreads, 0, s>>>(keys, vals, n, dN, digits);
        RYO_LAUNCH_CHECK(ctx);
        return RYO_OK;
    }
    const int numTiles = int((n + kSortTile - 1) / kSortTile);
    uint32_t* hist     = reinterpret_cast<uint32_t*>(tmp);
    uint32_t* dbase    = hist + kPasses * kDigits;
    uint32_t* tileHist = dbase + kPasses * kDigits;
    const int numChunks = (numTiles + kScanChunk - 1) / kScanChunk;
    uint32_t* chunkSum  = tileHist + size_t(kDigits) * numTiles;
    uint32_t* plan      = chunkSum + size_t(kDigits) * numChunks; // in the last 256 bytes, with the error word
    uint32_t* err       = plan + kPlanWords;
    zeroWordsKernel<<<8, 256, 0, s>>>(hist, kPasses * kDigits); // a kernel, not a memset node, under capture
    histAllKernel<<<gridFor(n, 256, ctx->numCUs * 4), 256, 0, s>>>(keys, n, dN, hist);
    digitBaseKernel<<<kPasses, 256, 0, s>>>(hist, dbase);
    lsdPlanKernel<<<1, 256, 0, s>>>(hist, n, dN, digits, plan);
    // every pass is launched; the plan, made on the device, skips the passes whose digit is constant and says which
    // buffer each of the others reads
    for (int p = 0; p < digits; ++p)
    {
        int shift = 8 * p;
        tileHistKernel<<<numTiles, kSortThreads, 0, s>>>(keys, n, dN, shift, tileHist, numTiles, keyBuf, plan);
        if (numTiles <= kSmallScanTiles)
            tileScanSmallKernel<<<kDigits, 256, 0, s>>>(tileHist, numTiles, dbase + p * kDigits, nullptr, plan + p);
        else
        {
            tileChunkSumKernel<<<numChunks, 256, 0, s>>>(tileHist, numTiles, chunkSum, plan + p);
            tileChunkScanKernel<<<kDigits, 256, 0, s>>>(chunkSum, numChunks, dbase + p * kDigits, nullptr, plan + p);
            tileOffsetKernel<<<numChunks, 256, 0, s>>>(tileHist, numTiles, chunkSum, plan + p);
        }
        scatterKernel<<<numTiles, kSortThreads, 0, s>>>(keys, vals, keyBuf, valBuf, n, dN, shift, tileHist, numTiles,
                                                        nullptr, err, plan);
    }
    copyPairsKernel<<<gridFor(n, 256, ctx->numCUs * 16), 256, 0, s>>>(keyBuf, valBuf, n, dN, keys, vals, plan);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

int sequence(Context* ctx, int32_t* out, size_t n, int32_t start);

static MsdLists msdLists(const MsdLayout& l, void* tmp)
{
    char* t = static_cast<char*>(tmp);
    MsdLists L;
    L.ctl = reinterpret_cast<uint32_t*>(t + l.ctl);
    for (int i = 0; i < 2; ++i)
    {
        L.segs[i]  = reinterpret_cast<MsdSeg*>(t + l.segs[i]);
        L.owner[i] = reinterpret_cast<uint32_t*>(t + l.owner[i]);
    }
    for (int c = 0; c < kNumRangeClasses; ++c)
        L.ranges[c] = reinterpret_cast<uint2*>(t + l.ranges[c]);
    L.maxSegs = l.maxSegs, L.maxTiles = l.maxTiles, L.maxRanges = l.maxRanges;
    return L;
}

// rounds r0..7 on the segments still larger than kLocalMax (round r reads keyBuf when r is odd), then the LDS sorts
// of every range into keys/vals
static void msdRoundsAndRanges(Context* ctx, const MsdLayout& l, const MsdLists& L, int r0, uint64_t* keys,
                               int32_t* vals, uint64_t* keyBuf, int32_t* valBuf, size_t n, uint32_t* tileHist)
{
    hipStream_t s = ctx->stream;
    // grids of what is resident at once (5 scatter blocks per CU with 30 KB of LDS each, 8 histogram blocks), so
    // that no second wave of blocks starts late; multiples of 8 for the XCD mapping
    const unsigned histGrid    = std::max(8u, std::min(l.maxTiles, unsigned(ctx->numCUs) * 8u) & ~7u);
    const unsigned scatterGrid = std::max(8u, std::min(l.maxTiles, unsigned(ctx->numCUs) * kMsdScatterPerCU) & ~7u);
    const unsigned segGrid     = std::max(1u, std::min(l.maxSegs, unsigned(ctx->numCUs) * 2u));
    for (int r = r0; r < kMsdRounds; ++r)
    {
        uint64_t* kin  = (r & 1) ? keyBuf : keys;
        int32_t* vin   = (r & 1) ? valBuf : vals;
        uint64_t* kout = (r & 1) ? keys : keyBuf;
        int32_t* vout  = (r & 1) ? vals : valBuf;
        msdTileHistKernel<<<histGrid, 256, 0, s>>>(kin, n, L, r, tileHist);
        msdScanKernel<<<segGrid, kMsdScanThreads, 0, s>>>(n, L, r, tileHist);
        msdScatterKernel<<<scatterGrid, kSortThreads, 0, s>>>(kin, vin, kout, vout, n, L, r, tileHist);
    }
    // the ranges, by size class, into the first buffer
    const unsigned cus = unsigned(ctx->numCUs);
    msdLocalSortKernel<128, 8><<<cus * 10, 128, 0, s>>>(L, 0, keys, vals, keyBuf, valBuf, n);
    msdLocalSortKernel<512, 8><<<cus * 2, 512, 0, s>>>(L, 1, keys, vals, keyBuf, valBuf, n);
    msdLocalSortKernel<1024, 8><<<cus, 1024, 0, s>>>(L, 2, keys, vals, keyBuf, valBuf, n);
}

// the full sort of more than kBlockSortMax pairs (fewer than 2^31) by MSD rounds and LDS sorts
static int sortPairsMsd(Context* ctx, uint64_t* keys, int32_t* vals, size_t n, uint64_t* keyBuf, int32_t* valBuf,
                        void* tmp, bool valsAreIndex)
{
    const MsdLayout l  = msdLayout(n);
    const MsdLists L   = msdLists(l, tmp);
    char* t            = static_cast<char*>(tmp);
    uint32_t* hist     = reinterpret_cast<uint32_t*>(t + l.hist);
    uint32_t* base     = reinterpret_cast<uint32_t*>(t + l.base);
    uint32_t* tileHist = reinterpret_cast<uint32_t*>(t + l.tileHist);
    uint32_t* chunkSum = reinterpret_cast<uint32_t*>(t + l.chunkSum);
    const int numTiles = int(l.numTiles0), numChunks = int(l.numChunks0);
    constexpr int kTop = 8 * (kMsdRounds - 1);
    hipStream_t s = ctx->stream;

    // round 0: the stable pass on the top digit into the second buffer, then its children
    // (the tile scan gives each tile's offsets within the digit's run and the digit totals, whose exclusive scan
    // the scatter adds)
    msdInitKernel<<<1, 256, 0, s>>>(L.ctl, hist);
    tileHistKernel<<<numTiles, kSortThreads, 0, s>>>(keys, n, nullptr, kTop, tileHist, numTiles);
    if (numTiles <= kSmallScanTiles)
        tileScanSmallKernel<<<kDigits, 256, 0, s>>>(tileHist, numTiles, nullptr, hist);
    else
    {
        tileChunkSumKernel<<<numChunks, 256, 0, s>>>(tileHist, numTiles, chunkSum);
        tileChunkScanKernel<<<kDigits, 256, 0, s>>>(chunkSum, numChunks, nullptr, hist);
        tileOffsetKernel<<<numChunks, 256, 0, s>>>(tileHist, numTiles, chunkSum);
    }
    digitBaseKernel<<<1, 256, 0, s>>>(hist, base);
    scatterKernel<<<numTiles, kSortThreads, 0, s>>>(keys, valsAreIndex ? nullptr : vals, keyBuf, valBuf, n, nullptr,
                                                    kTop, tileHist, numTiles, base, L.ctl + kMsdErr);
    msdRootKernel<<<1, 256, 0, s>>>(hist, base, L);
    // rounds 1..7: the segments still larger than kLocalMax, on the next lower digit, between the buffers
    msdRoundsAndRanges(ctx, l, L, 1, keys, vals, keyBuf, valBuf, n, tileHist);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

// Each segment [off[s], off[s+1]) sorted on its own (S <= kMaxSortSegments, fewer than 2^31 pairs): MSD rounds from
// round 0 for the segments larger than kLocalMax, an LDS sort for each smaller one. With dN, n is the capacity and
// *dN the count (offsets past it are clipped).
static int segmentedSortMsd(Context* ctx, uint64_t* keys, int32_t* vals, size_t n, const uint32_t* off, int S,
                            uint64_t* keyBuf, int32_t* valBuf, void* tmp, const uint32_t* dN, size_t maxSegment)
{
    const MsdLayout l  = msdLayout(n);
    const MsdLists L   = msdLists(l, tmp);
    uint32_t* tileHist = reinterpret_cast<uint32_t*>(static_cast<char*>(tmp) + l.tileHist);
    msdSegInitKernel<<<1, 256, 0, ctx->stream>>>(off, S, n, dN, L);
    // no segment larger than an LDS sort: the rounds would find nothing to split, so they are not launched
    msdRoundsAndRanges(ctx, l, L, maxSegment <= size_t(kLocalMax) ? kMsdRounds : 0, keys, vals, keyBuf, valBuf, n,
                       tileHist);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

// the guard word of the last sortPairs / sortPairsIndex of n pairs in tmp (the caller has synchronized the stream):
// nonzero when a kernel found offsets or list entries that do not describe the keys (it then skipped the access)
int sortCheck(Context* ctx, const void* tmp, size_t n)
{
    if (n <= kBlockSortMax || n >= (size_t(1) << 31)) return RYO_OK; // the one-block sort has no lists to check
    uint32_t err = 0;
    const char* w = static_cast<const char*>(tmp) + msdLayout(n).ctl + sizeof(uint32_t) * kMsdErr;
    RYO_HIP(ctx, hipMemcpy(&err, w, sizeof(err), hipMemcpyDeviceToHost));
    if (err) return setError(ctx, RYO_ERR_HIP, "sort: inconsistent offsets or lists (guard bits 0x%x)", err);
    return RYO_OK;
}

// the full 64-bit stable sort; tmp: sortTempBytes(n) bytes
int sortPairs(Context* ctx, uint64_t* keys, int32_t* vals, size_t n, uint64_t* keyBuf, int32_t* valBuf, void* tmp,
              size_t tmpBytes)
{
    if (n > kBlockSortMax && n < (size_t(1) << 31))
    {
        if (tmpBytes < sortTempBytes(n))
            return setError(ctx, RYO_ERR_ARG, "sort: temp storage too small (%zu < %zu)", tmpBytes, sortTempBytes(n));
        return sortPairsMsd(ctx, keys, vals, n, keyBuf, valBuf, tmp, false);
    }
    return sortPairsDigits(ctx, keys, vals, n, keyBuf, valBuf, tmp, tmpBytes, kPasses, nullptr);
}

// the same sort of keys paired with their indices 0..n-1 (vals is written, not read: the sort's first pass makes
// the identity permutation instead of reading one)
int sortPairsIndex(Context* ctx, uint64_t* keys, int32_t* vals, size_t n, uint64_t* keyBuf, int32_t* valBuf,
                   void* tmp, size_t tmpBytes)
{
    if (n > kBlockSortMax && n < (size_t(1) << 31))
    {
        if (tmpBytes < sortTempBytes(n))
            return setError(ctx, RYO_ERR_ARG, "sort: temp storage too small (%zu < %zu)", tmpBytes, sortTempBytes(n));
        return sortPairsMsd(ctx, keys, vals, n, keyBuf, valBuf, tmp, true);
    }
    if (int rc = sequence(ctx, vals, n, 0)) return rc;
    return sortPairsDigits(ctx, keys, vals, n, keyBuf, valBuf, tmp, tmpBytes, kPasses, nullptr);
}

/* ============================================================================ batches of systems */

// keys of packed systems: particle i of system s uses that system's box; vals = particle index
__global__ __launch_bounds__(256) void batchKeysKernel(const float* __restrict__ x, const float* __restrict__ y,
                                                       const float* __restrict__ z, size_t n,
                                                       const uint32_t* __restrict__ P, int S,
                                                       const KeyBox* __restrict__ boxes, uint64_t* __restrict__ keys,
                                                       int32_t* __restrict__ vals)
{
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    {
        const KeyBox& box = boxes[segmentOf(P, S, uint32_t(i))];
        int ix = min(cellIndex(x[i], box.lo[0], box.m[0]), int((1u << box.bits.b[0]) - 1));
        int iy = min(cellIndex(y[i], box.lo[1], box.m[1]), int((1u << box.bits.b[1]) - 1));
        int iz = min(cellIndex(z[i], box.lo[2], box.m[2]), int((1u << box.bits.b[2]) - 1));
        keys[i] = hilbertMixD(ix, iy, iz, box.bits);
        vals[i] = int32_t(i);
    }
}



KeyBox makeKeyBox(const float* limits)
{
    KeyBox b;
    b.bits = boxBits(limits);
    for (int d = 0; d < 3; ++d)
    {
        float lo = limits[2 * d], hi = limits[2 * d + 1];
        b.lo[d]  = lo;
        b.m[d]   = float(1u << b.bits.b[d]) * (1.0f / (hi - lo));
    }
    return b;
}

int batchKeys(Context* ctx, const float* x, const float* y, const float* z, size_t n, const uint32_t* P, int S,
              const KeyBox* boxes, uint64_t* keys, int32_t* vals)
{
    if (n == 0) return RYO_OK;
    batchKeysKernel<<<gridFor(n, 256, ctx->numCUs * 16), 256, 0, ctx->stream>>>(x, y, z, n, P, S, boxes, keys,
                                                                                 vals);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

/* stable sort of (keys, vals) grouped by segment: each segment [off[s], off[s+1]) of the pairs (1 <= S <=
 * kMaxSortSegments, fewer than 2^31 pairs) sorted on its own, keys ascending and equal keys in their order.
 * keyBuf/valBuf: n entries each; tmp: sortTempBytes(n). With dN, n is the capacity and *dN (device) the element
 * count (offsets past it are clipped). maxSegment: a bound the caller knows on the longest segment (SIZE_MAX if
 * none). */
int segmentedSortPairs(Context* ctx, uint64_t* keys, int32_t* vals, size_t n, const uint32_t* off, int S,
                       uint64_t* keyBuf, int32_t* valBuf, void* tmp, size_t tmpBytes, const uint32_t* dN,
                       size_t maxSegment)
{
    if (n == 0) return RYO_OK;
    if (S < 1 || S > kMaxSortSegments || n >= (size_t(1) << 31))
        return setError(ctx, RYO_ERR_ARG, "segmented sort: %d segments, %zu pairs", S, n);
    if (tmpBytes < sortTempBytes(n))
        return setError(ctx, RYO_ERR_ARG, "sort: temp storage too small (%zu < %zu)", tmpBytes, sortTempBytes(n));
    return segmentedSortMsd(ctx, keys, vals, n, off, S, keyBuf, valBuf, tmp, dN, maxSegment);
}

template<class T>
static int makeKeyBox(Context* ctx, const T* limits, KeyBoxT<T>& b, unsigned maxLevel = kMaxLevel)
{
    b.bits = boxBits(limits, maxLevel); // box.hpp:160-198: an axis shorter than 0.75 x the longest gets fewer bits
    for (int d = 0; d < 3; ++d)
    {
        T lo = limits[2 * d], hi = limits[2 * d + 1];
        if (!(hi > lo)) return setError(ctx, RYO_ERR_ARG, "sfc keys: empty box along axis %d", d);
        T il    = T(1) / (hi - lo); // Box<T>::inverseLengths_ (box.hpp:107), in T on the host
        b.lo[d] = lo;
        b.m[d]  = T(1u << b.bits.b[d]) * il;
    }
    return RYO_OK;
}

template<class T>
int sfcKeysT(Context* ctx, const T* x, const T* y, const T* z, uint64_t* keys, size_t n, const T* limits, int kind,
             int skipRemoved)
{
    if (n == 0) return RYO_OK;
    KeyBoxT<T> b;
    if (int rc = makeKeyBox(ctx, limits, b)) return rc;
    const unsigned grid = gridFor(n, 256, ctx->numCUs * 16);
    if (isCubic(b.bits))
        sfcKeysKernel<T, true, false><<<grid, 256, 0, ctx->stream>>>(x, y, z, keys, n, b, kind, skipRemoved,
                                                                      PackOut{});
    else
        sfcKeysKernel<T, false, false><<<grid, 256, 0, ctx->stream>>>(x, y, z, keys, n, b, kind, skipRemoved,
                                                                       PackOut{});
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

int sfcKeys(Context* ctx, const float* x, const float* y, const float* z, uint64_t* keys, size_t n,
            const float* limits, int kind, int skipRemoved)
{
    return sfcKeysT(ctx, x, y, z, keys, n, limits, kind, skipRemoved);
}

int sfcKeys64(Context* ctx, const double* x, const double* y, const double* z, uint64_t* keys, size_t n,
              const double* limits, int kind, int skipRemoved)
{
    return sfcKeysT(ctx, x, y, z, keys, n, limits, kind, skipRemoved);
}

// 32-bit keys: 10 levels, the box bits reduced from 10 (maxTreeLevel<uint32_t>, sfc.hpp:89-96)
template<class T>
int sfcKeys32(Context* ctx, const T* x, const T* y, const T* z, uint32_t* keys, size_t n, const T* limits, int kind,
              int skipRemoved)
{
    if (n == 0) return RYO_OK;
    KeyBoxT<T> b;
    if (int rc = makeKeyBox(ctx, limits, b, 10u)) return rc;
    sfcKeysKernel<T, false, false, uint32_t><<<gridFor(n, 256, ctx->numCUs * 16), 256, 0, ctx->stream>>>(
        x, y, z, keys, n, b, kind, skipRemoved, PackOut{});
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}
template int sfcKeys32<float>(Context*, const float*, const float*, const float*, uint32_t*, size_t, const float*, int,
                              int);
template int sfcKeys32<double>(Context*, const double*, const double*, const double*, uint32_t*, size_t,
                               const double*, int, int);

// Hilbert keys of a freshly loaded system together with the identity permutation and the gather records of
// in = {x, y, z, m, h, o} (the records stay in the context's pack scratch until gatherRecords)
int sfcKeysPacked(Context* ctx, const float* const* in, uint64_t* keys, int32_t* perm, size_t n,
                  const float* limits, float4** rec)
{
    *rec = nullptr;
    if (n == 0) return RYO_OK;
    KeyBox b;
    if (int rc = makeKeyBox(ctx, limits, b)) return rc;
    auto* r = static_cast<float4*>(ctx->getScratch(kSlotPack, 32 * n));
    if (!r) return RYO_ERR_HIP;
    const PackOut pk{in[3], in[4], in[5], perm, r}; // perm may be null (the sort makes the identity permutation)
    const unsigned grid = gridFor(n, 256, ctx->numCUs * 16);
    if (isCubic(b.bits))
        sfcKeysKernel<float, true, true><<<grid, 256, 0, ctx->stream>>>(in[0], in[1], in[2], keys, n, b, 0, 0, pk);
    else
        sfcKeysKernel<float, false, true><<<grid, 256, 0, ctx->stream>>>(in[0], in[1], in[2], keys, n, b, 0, 0, pk);
    RYO_LAUNCH_CHECK(ctx);
    *rec = r;
    return RYO_OK;
}

__global__ void copyWordsKernel(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src, size_t n)
{
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

__global__ void zeroWordsKernelN(uint32_t* __restrict__ dst, size_t n)
{
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        dst[i] = 0;
}

// 4-byte word copies and fills as kernels (capturable in a graph without memcpy or memset nodes)
int copyWords(Context* ctx, void* dst, const void* src, size_t words)
{
    if (words == 0) return RYO_OK;
    copyWordsKernel<<<gridFor(words, 256, ctx->numCUs * 16), 256, 0, ctx->stream>>>(
        static_cast<uint32_t*>(dst), static_cast<const uint32_t*>(src), words);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

int zeroWords(Context* ctx, void* dst, size_t words)
{
    if (words == 0) return RYO_OK;
    zeroWordsKernelN<<<gridFor(words, 256, ctx->numCUs * 16), 256, 0, ctx->stream>>>(static_cast<uint32_t*>(dst),
                                                                                    words);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

int sequence(Context* ctx, int32_t* out, size_t n, int32_t start)
{
    if (n == 0) return RYO_OK;
    iotaKernel<<<gridFor(n, 256, ctx->numCUs * 16), 256, 0, ctx->stream>>>(out, n, start);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

/* out[k][i] = in[k][perm[i]] for the six arrays x, y, z, m, h, order of a system in random order, through
 * 32-byte records (scratch: 32 n bytes) */
// out[k][i] = field k of record perm[i] (records from sfcKeysPacked)
int gatherRecords(Context* ctx, const int32_t* perm, size_t n, const float4* rec, float* const* out)
{
    if (n == 0) return RYO_OK;
    gatherRecordsKernel<<<gridFor(n, 256, ctx->numCUs * 16), 256, 0, ctx->stream>>>(perm, n, rec, out[0], out[1],
                                                                                     out[2], out[3], out[4], out[5]);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

template<class T>
static int gatherT(Context* ctx, const int32_t* perm, size_t n, int num, const T* const* in, T* const* out)
{
    if (num < 0 || num > 10) return setError(ctx, RYO_ERR_ARG, "gather: at most 10 arrays");
    if (n == 0 || num == 0) return RYO_OK;
    GatherArgsT<T> a{};
    a.num = num;
    for (int k = 0; k < num; ++k)
        a.in[k] = in[k], a.out[k] = out[k];
    gatherKernel<<<gridFor(n, 256, ctx->numCUs * 16), 256, 0, ctx->stream>>>(perm, n, a);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

int gather(Context* ctx, const int32_t* perm, size_t n, int num, const float* const* in, float* const* out)
{
    return gatherT(ctx, perm, n, num, in, out);
}

int gather64(Context* ctx, const int32_t* perm, size_t n, int num, const double* const* in, double* const* out)
{
    return gatherT(ctx, perm, n, num, in, out);
}

} // namespace ryo
