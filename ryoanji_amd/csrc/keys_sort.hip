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
            pk.perm[i]       = int32_t(i);
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

// per-tile digit counts of one pass, one contiguous row per tile: tileHist[tile * 256 + d] (a coalesced 1 KB write
// per tile, read back the same way by the offset scan and the scatter)
__global__ __launch_bounds__(kSortThreads) void tileHistKernel(const uint64_t* __restrict__ keys, size_t n,
                                                              const uint32_t* __restrict__ dN, int shift,
                                                              uint32_t* __restrict__ tileHist, int numTiles)
{
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
                                                     uint32_t* __restrict__ chunkSum)
{
    const int c = blockIdx.x, d = threadIdx.x;
    const int t1 = min(numTiles, (c + 1) * kScanChunk);
    uint32_t sum = 0;
    for (int t = c * kScanChunk; t < t1; ++t)
        sum += tileHist[size_t(t) * kDigits + d];
    chunkSum[size_t(c) * kDigits + d] = sum;
}

__global__ __launch_bounds__(256) void tileChunkScanKernel(uint32_t* __restrict__ chunkSum, int numChunks,
                                                      const uint32_t* __restrict__ digitBase)
{
    __shared__ uint32_t sw[4];
    const int d = blockIdx.x;
    uint32_t carry = digitBase[d];
    for (int c0 = 0; c0 < numChunks; c0 += 256)
    {
        const int c = c0 + threadIdx.x;
        const uint32_t v = c < numChunks ? chunkSum[size_t(c) * kDigits + d] : 0u;
        uint32_t tot;
        const uint32_t ex = blockExclusiveScan256(v, sw, &tot);
        if (c < numChunks) chunkSum[size_t(c) * kDigits + d] = carry + ex;
        carry += tot;
    }
}

// the same offsets in one launch for few tiles (one block per digit; its reads are strided, which costs little
// when there are few tiles and saves two launches per pass)
__global__ __launch_bounds__(256) void tileScanSmallKernel(uint32_t* __restrict__ tileHist, int numTiles,
                                                          const uint32_t* __restrict__ digitBase)
{
    __shared__ uint32_t sw[4];
    const int d = blockIdx.x;
    uint32_t carry = digitBase[d];
    for (int t0 = 0; t0 < numTiles; t0 += 256)
    {
        const int t = t0 + threadIdx.x;
        const uint32_t v = t < numTiles ? tileHist[size_t(t) * kDigits + d] : 0u;
        uint32_t tot;
        const uint32_t ex = blockExclusiveScan256(v, sw, &tot);
        if (t < numTiles) tileHist[size_t(t) * kDigits + d] = carry + ex;
        carry += tot;
    }
}

__global__ __launch_bounds__(256) void tileOffsetKernel(uint32_t* __restrict__ tileHist, int numTiles,
                                                       const uint32_t* __restrict__ chunkSum)
{
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

// stable scatter of one tile: wave-level match ranking (8 ballots per key), block-level offsets, staging in
// LDS so that the writes of one digit run are contiguous
__global__ __launch_bounds__(kSortThreads) void scatterKernel(const uint64_t* __restrict__ keysIn,
                                                             const int32_t* __restrict__ valsIn,
                                                             uint64_t* __restrict__ keysOut,
                                                             int32_t* __restrict__ valsOut, size_t n,
                                                             const uint32_t* __restrict__ dN, int shift,
                                                             const uint32_t* __restrict__ tileOffsets, int numTiles)
{
    n = deviceCount(dN, n);
    const int tile = xcdTile(blockIdx.x, numTiles);
    if (size_t(tile) * kSortTile >= n) return; // a tile past the count (the whole block leaves)
    __shared__ uint64_t sk[kSortTile];
    __shared__ int32_t sv[kSortTile];
    __shared__ uint32_t wcnt[4][kDigits];
    __shared__ uint32_t tstart[kDigits];
    __shared__ uint32_t gbase[kDigits];
    __shared__ uint32_t sw[4];

    const size_t base = size_t(tile) * kSortTile;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int i = threadIdx.x; i < 4 * kDigits; i += kSortThreads)
        (&wcnt[0][0])[i] = 0;
    __syncthreads();

    uint64_t k[kSortItems];
    int32_t v[kSortItems];
    uint32_t rank[kSortItems];
    const uint64_t ltMask = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < kSortItems; ++r)
    {
        size_t idx = base + size_t(w) * (kSortItems * 64) + r * 64 + lane;
        bool valid = idx < n;
        k[r] = valid ? keysIn[idx] : 0;
        v[r] = valid ? valsIn[idx] : 0;
        unsigned d = unsigned(k[r] >> shift) & 255u;
        uint64_t match = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b)
        {
            bool bit = (d >> b) & 1u;
            uint64_t bb = __ballot(bit);
            match &= bit ? bb : ~bb;
        }
        uint32_t pre = valid ? wcnt[w][d] : 0;
        rank[r] = pre + __popcll(match & ltMask);
        bool leader = valid && (match >> lane) == 1ull;
        if (leader) wcnt[w][d] = pre + __popcll(match);
    }
    __syncthreads();
    {
        int d = threadIdx.x;
        uint32_t c0 = wcnt[0][d], c1 = wcnt[1][d], c2 = wcnt[2][d], c3 = wcnt[3][d];
        uint32_t tot;
        uint32_t start = blockExclusiveScan256(c0 + c1 + c2 + c3, sw, &tot);
        wcnt[0][d] = 0;
        wcnt[1][d] = c0;
        wcnt[2][d] = c0 + c1;
        wcnt[3][d] = c0 + c1 + c2;
        tstart[d] = start;
        gbase[d] = tileOffsets[size_t(tile) * kDigits + d];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r)
    {
        size_t idx = base + size_t(w) * (kSortItems * 64) + r * 64 + lane;
        if (idx < n)
        {
            unsigned d = unsigned(k[r] >> shift) & 255u;
            uint32_t pos = tstart[d] + wcnt[w][d] + rank[r];
            sk[pos] = k[r];
            sv[pos] = v[r];
        }
    }
    __syncthreads();
    size_t nValid = n - base < size_t(kSortTile) ? n - base : size_t(kSortTile);
    for (int j = threadIdx.x; j < int(nValid); j += kSortThreads)
    {
        uint64_t key = sk[j];
        unsigned d = unsigned(key >> shift) & 255u;
        size_t out = size_t(gbase[d]) + (j - tstart[d]);
        keysOut[out] = key;
        valsOut[out] = sv[j];
    }
}

// sorts of at most kBlockSortMax pairs in one launch: a single block of 1024 threads holds the pairs in registers
// (8 per thread) and LDS for every pass: the wave-level match ranking of scatterKernel gives each wave's digit
// counts, whose sums are the pass's histogram, and the stable scatter goes to LDS, from where each thread reloads
// its 8 items. One global read and one write in all, and one launch instead of three per pass plus three, which is
// most of a small sort's time.
constexpr int kBlockSortThreads = 1024;
constexpr int kBlockSortWaves   = kBlockSortThreads / 64;
constexpr int kBlockSortItems   = 8;
constexpr size_t kBlockSortMax  = size_t(kBlockSortThreads) * kBlockSortItems;

__global__ __launch_bounds__(kBlockSortThreads) void blockSortKernel(uint64_t* __restrict__ keys,
                                                                     int32_t* __restrict__ vals, size_t n,
                                                                     const uint32_t* __restrict__ dN, int digits)
{
    __shared__ uint64_t sk[kBlockSortMax];
    __shared__ int32_t sv[kBlockSortMax];
    __shared__ uint32_t wcnt[kBlockSortWaves][kDigits]; // per-wave digit counts, then their prefixes over waves
    __shared__ uint32_t offs[kDigits];                   // digit totals, then their exclusive scan
    n = deviceCount(dN, n);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t ltMask = (1ull << lane) - 1ull;
    // item r of thread (w, lane) is element w * 512 + r * 64 + lane: in wave order, then item order, then lanes
    uint64_t k[kBlockSortItems];
    int32_t v[kBlockSortItems];
#pragma unroll
    for (int r = 0; r < kBlockSortItems; ++r)
    {
        const size_t idx = size_t(w) * (kBlockSortItems * 64) + r * 64 + lane;
        k[r] = idx < n ? keys[idx] : 0;
        v[r] = idx < n ? vals[idx] : 0;
    }
    for (int p = 0; p < digits; ++p)
    {
        const int shift = 8 * p;
        for (int i = threadIdx.x; i < kBlockSortWaves * kDigits; i += kBlockSortThreads)
            (&wcnt[0][0])[i] = 0;
        __syncthreads();
        uint32_t rank[kBlockSortItems];
#pragma unroll
        for (int r = 0; r < kBlockSortItems; ++r)
        {
            const bool valid = size_t(w) * (kBlockSortItems * 64) + r * 64 + lane < n;
            const unsigned d = unsigned(k[r] >> shift) & 255u;
            uint64_t match = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; ++b)
            {
                const bool bit = (d >> b) & 1u;
                const uint64_t bb = __ballot(bit);
                match &= bit ? bb : ~bb;
            }
            const uint32_t pre = valid ? wcnt[w][d] : 0;
            rank[r] = pre + __popcll(match & ltMask);
            if (valid && (match >> lane) == 1ull) wcnt[w][d] = pre + __popcll(match);
        }
        __syncthreads();
        if (threadIdx.x < kDigits)
        {
            const int d = threadIdx.x;
            uint32_t c = 0;
            for (int ww = 0; ww < kBlockSortWaves; ++ww)
            {
                const uint32_t x = wcnt[ww][d];
                wcnt[ww][d] = c;
                c += x;
            }
            offs[d] = c;
        }
        __syncthreads();
        if (w == 0)
        {
            // exclusive scan of the 256 digit totals, 4 per lane
            const uint32_t c0 = offs[4 * lane], c1 = offs[4 * lane + 1], c2 = offs[4 * lane + 2],
                           c3 = offs[4 * lane + 3];
            const uint32_t sum = c0 + c1 + c2 + c3;
            uint32_t inc = sum;
            for (int o = 1; o < 64; o <<= 1)
            {
                const uint32_t t = __shfl_up(inc, o);
                if (lane >= o) inc += t;
            }
            const uint32_t ex = inc - sum;
            offs[4 * lane] = ex, offs[4 * lane + 1] = ex + c0, offs[4 * lane + 2] = ex + c0 + c1;
            offs[4 * lane + 3] = ex + c0 + c1 + c2;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kBlockSortItems; ++r)
            if (size_t(w) * (kBlockSortItems * 64) + r * 64 + lane < n)
            {
                const unsigned d  = unsigned(k[r] >> shift) & 255u;
                const uint32_t at = offs[d] + wcnt[w][d] + rank[r];
                sk[at] = k[r];
                sv[at] = v[r];
            }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kBlockSortItems; ++r)
        {
            const size_t idx = size_t(w) * (kBlockSortItems * 64) + r * 64 + lane;
            if (idx < n) k[r] = sk[idx], v[r] = sv[idx];
        }
    }
#pragma unroll
    for (int r = 0; r < kBlockSortItems; ++r)
    {
        const size_t idx = size_t(w) * (kBlockSortItems * 64) + r * 64 + lane;
        if (idx < n) keys[idx] = k[r], vals[idx] = v[r];
    }
}

__global__ void copyPairsKernel(const uint64_t* __restrict__ kin, const int32_t* __restrict__ vin, size_t n,
                                const uint32_t* __restrict__ dN, uint64_t* __restrict__ kout,
                                int32_t* __restrict__ vout)
{
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

size_t sortTempBytes(size_t n)
{
    size_t numTiles  = (n + kSortTile - 1) / kSortTile;
    size_t numChunks = (numTiles + kScanChunk - 1) / kScanChunk;
    return 256 * sizeof(uint32_t) * (2 * kPasses) + size_t(kDigits) * (numTiles + numChunks) * sizeof(uint32_t) + 256;
}

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
    zeroWordsKernel<<<8, 256, 0, s>>>(hist, kPasses * kDigits); // a kernel, not a memset node, under capture
    histAllKernel<<<gridFor(n, 256, ctx->numCUs * 4), 256, 0, s>>>(keys, n, dN, hist);
    digitBaseKernel<<<kPasses, 256, 0, s>>>(hist, dbase);
    uint64_t* kin = keys;
    uint64_t* kout = keyBuf;
    int32_t* vin = vals;
    int32_t* vout = valBuf;
    for (int p = 0; p < digits; ++p)
    {
        int shift = 8 * p;
        tileHistKernel<<<numTiles, kSortThreads, 0, s>>>(kin, n, dN, shift, tileHist, numTiles);
        if (numTiles <= kSmallScanTiles)
            tileScanSmallKernel<<<kDigits, 256, 0, s>>>(tileHist, numTiles, dbase + p * kDigits);
        else
        {
            tileChunkSumKernel<<<numChunks, 256, 0, s>>>(tileHist, numTiles, chunkSum);
            tileChunkScanKernel<<<kDigits, 256, 0, s>>>(chunkSum, numChunks, dbase + p * kDigits);
            tileOffsetKernel<<<numChunks, 256, 0, s>>>(tileHist, numTiles, chunkSum);
        }
        scatterKernel<<<numTiles, kSortThreads, 0, s>>>(kin, vin, kout, vout, n, dN, shift, tileHist, numTiles);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (kin != keys)
        copyPairsKernel<<<gridFor(n, 256, ctx->numCUs * 16), 256, 0, s>>>(kin, vin, n, dN, keys, vals);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

int sortPairs(Context* ctx, uint64_t* keys, int32_t* vals, size_t n, uint64_t* keyBuf, int32_t* valBuf, void* tmp,
              size_t tmpBytes)
{
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

// the segment of each (sorted) value as the key of a second, stable pass; pos = current position
__global__ void segmentKeysKernel(const int32_t* __restrict__ vals, size_t n, const uint32_t* __restrict__ dN,
                                  const uint32_t* __restrict__ off, int S, uint64_t* __restrict__ segKeys,
                                  int32_t* __restrict__ pos)
{
    n = deviceCount(dN, n);
    for (size_t j = size_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n; j += size_t(gridDim.x) * blockDim.x)
    {
        segKeys[j] = uint64_t(segmentOf(off, S, uint32_t(vals[j])));
        pos[j]     = int32_t(j);
    }
}

__global__ void applyPosKernel(const int32_t* __restrict__ pos, size_t n, const uint32_t* __restrict__ dN,
                               const uint64_t* __restrict__ kin, const int32_t* __restrict__ vin,
                               uint64_t* __restrict__ kout, int32_t* __restrict__ vout)
{
    n = deviceCount(dN, n);
    for (size_t j = size_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n; j += size_t(gridDim.x) * blockDim.x)
    {
        kout[j] = kin[pos[j]];
        vout[j] = vin[pos[j]];
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

/* stable sort of (keys, vals) grouped by segment: a full sort on the keys, then a stable pass on the segment of
 * each value (segment = the s with off[s] <= val < off[s+1]); keys ascend inside each segment and equal keys
 * keep their order. keyBuf/valBuf/segKeys/segBuf/pos/posBuf: n entries each; tmp: sortTempBytes(n). With dN, n is
 * the capacity and *dN (device) the element count, as in sortPairsDigits. */
int segmentedSortPairs(Context* ctx, uint64_t* keys, int32_t* vals, size_t n, const uint32_t* off, int S,
                       uint64_t* keyBuf, int32_t* valBuf, uint64_t* segKeys, uint64_t* segBuf, int32_t* pos,
                       int32_t* posBuf, void* tmp, size_t tmpBytes, const uint32_t* dN)
{
    if (n == 0) return RYO_OK;
    int rc = sortPairsDigits(ctx, keys, vals, n, keyBuf, valBuf, tmp, tmpBytes, kPasses, dN);
    if (rc) return rc;
    if (S <= 1) return RYO_OK;
    hipStream_t s = ctx->stream;
    const unsigned grid = gridFor(n, 256, ctx->numCUs * 16);
    segmentKeysKernel<<<grid, 256, 0, s>>>(vals, n, dN, off, S, segKeys, pos);
    const int digits = S <= 256 ? 1 : (S <= 65536 ? 2 : 3);
    rc = sortPairsDigits(ctx, segKeys, pos, n, segBuf, posBuf, tmp, tmpBytes, digits, dN);
    if (rc) return rc;
    applyPosKernel<<<grid, 256, 0, s>>>(pos, n, dN, keys, vals, keyBuf, valBuf);
    copyPairsKernel<<<grid, 256, 0, s>>>(keyBuf, valBuf, n, dN, keys, vals);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
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
    const PackOut pk{in[3], in[4], in[5], perm, r};
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
