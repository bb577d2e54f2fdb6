// Reference-signature tree stages for callers that drive the pipeline step by step, with the call sequence of
// ryoanji/test/demo.cu and ryoanji/interface/treebuilder.cuh:
//   * cornerstone leaf arrays by iterated count / rebalance steps: computeNodeCountsGpu, computeNodeOpsGpu,
//     rebalanceTreeGpu (cstone/tree/csarray_gpu.cu:86-231; logic of csarray.hpp:148-366);
//   * source centers and MAC radii: computeLeafSourceCenterGpu, upsweepCentersGpu (one level per launch from the
//     caller's host level ranges), setMacGpu (cstone/focus/source_center_gpu.cu:103-310);
//   * multipoles: computeLeafMultipoles and upsweepMultipoles for one level (ryoanji/nbody/upsweep_gpu.cu:33-128).
// The fused entry points (ryo_build_leaves_u64, ryo_upsweep_*) produce the same trees and expansions in fewer
// launches; these keep the reference's intermediate states between the steps (centers rounded to the caller's
// type before the multipoles are formed about them, counts and node operations as arrays), so a caller that
// inspects or replaces one step sees what it expects. Sums are accumulated in FP64 and stored in the caller's
// type; the 8-lane groups (one leaf or one cell each) reduce by butterfly shuffles.
#include "common.hpp"
#include "mac.hpp"
#include "sfc.hpp"

#include <type_traits>

namespace ryo
{

namespace
{

/* ------------------------------------------------------------------ cornerstone leaf arrays */

__device__ inline size_t lowerBound64(const uint64_t* a, size_t n, uint64_t v)
{
    size_t lo = 0, hi = n;
    while (lo < hi)
    {
        size_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// particles of each node [tree[i], tree[i+1]) among the sorted keys, capped at maxCount (computeNodeCounts,
// csarray.hpp:148-230: the search guesses of the reference only shorten its searches)
__global__ __launch_bounds__(256) void nodeCountsKernel(const uint64_t* __restrict__ tree, int numNodes,
                                                        const uint64_t* __restrict__ keys, size_t n,
                                                        unsigned maxCount, uint32_t* __restrict__ counts)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < numNodes; i += gridDim.x * blockDim.x)
    {
        size_t lo   = lowerBound64(keys, n, tree[i]);
        size_t hi   = lowerBound64(keys, n, tree[i + 1]);
        unsigned c  = unsigned(hi - lo);
        counts[i]   = c < maxCount ? c : maxCount;
    }
}

// calculateNodeOp (csarray.hpp:236-290): 0 merge with the 7 siblings, 1 keep, 8 / 64 / 512 / 4096 split
__device__ inline int nodeOp(const uint64_t* tree, int i, int numNodes, const uint32_t* counts, unsigned bucket)
{
    const uint64_t thisNode = tree[i];
    const unsigned level    = treeLevel(tree[i + 1] - thisNode);
    int siblingIdx          = -1;
    if (level > 0)
    {
        siblingIdx      = int(octalDigit(thisNode, level));
        const int first = i - siblingIdx;
        const bool sib  = first >= 0 && first + 8 <= numNodes && tree[first + 8] == tree[first] + nodeRange(level - 1);
        if (!sib) siblingIdx = -1;
    }
    if (siblingIdx > 0)
    {
        const uint32_t* g = counts + (i - siblingIdx);
        size_t parentCount = 0;
        for (int k = 0; k < 8; ++k)
            parentCount += size_t(g[k]);
        if (parentCount <= size_t(bucket)) return 0;
    }
    const unsigned c = counts[i];
    if (c > bucket * 512u && level + 3 < kMaxLevel) return 4096;
    if (c > bucket * 64u && level + 2 < kMaxLevel) return 512;
    if (c > bucket * 8u && level + 1 < kMaxLevel) return 64;
    if (c > bucket && level < kMaxLevel) return 8;
    return 1;
}

__global__ __launch_bounds__(256) void nodeOpsKernel(const uint64_t* __restrict__ tree, int numNodes,
                                                     const uint32_t* __restrict__ counts, unsigned bucket,
                                                     int32_t* __restrict__ ops, int* changed)
{
    bool any = false;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < numNodes; i += gridDim.x * blockDim.x)
    {
        const int op = nodeOp(tree, i, numNodes, counts, bucket);
        ops[i]       = op;
        any |= op != 1;
    }
    if (__ballot(any) && (threadIdx.x & 63) == 0) atomicOr(changed, 1);
}

constexpr int kScanItems = 4;
constexpr int kScanBlock = 256 * kScanItems;

// exclusive scan of int32 in place (values and totals fit int32, as the reference's TreeNodeIndex): per-block
// sums, one block scanning them, then every block scans its slice with its offset
__global__ __launch_bounds__(256) void scanSumsKernel(const int32_t* __restrict__ a, size_t n, int32_t* sums)
{
    __shared__ uint32_t sw[4];
    const size_t base = size_t(blockIdx.x) * kScanBlock + size_t(threadIdx.x) * kScanItems;
    uint32_t v = 0;
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) v += uint32_t(a[base + k]);
    uint32_t total;
    blockExclusiveScan256(v, sw, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = int32_t(total);
}

__global__ __launch_bounds__(256) void scanOffsetsKernel(int32_t* sums, int numBlocks)
{
    __shared__ uint32_t sw[4];
    uint32_t running = 0;
    for (int b0 = 0; b0 < numBlocks; b0 += 256)
    {
        const int b      = b0 + int(threadIdx.x);
        const uint32_t v = b < numBlocks ? uint32_t(sums[b]) : 0u;
        uint32_t total;
        const uint32_t ex = blockExclusiveScan256(v, sw, &total);
        if (b < numBlocks) sums[b] = int32_t(running + ex);
        running += total;
    }
}

__global__ __launch_bounds__(256) void scanApplyKernel(int32_t* __restrict__ a, size_t n,
                                                       const int32_t* __restrict__ offsets)
{
    __shared__ uint32_t sw[4];
    const size_t base = size_t(blockIdx.x) * kScanBlock + size_t(threadIdx.x) * kScanItems;
    uint32_t v[kScanItems], sum = 0;
    for (int k = 0; k < kScanItems; ++k)
    {
        v[k] = base + k < n ? uint32_t(a[base + k]) : 0u;
        sum += v[k];
    }
    uint32_t total;
    uint32_t run = blockExclusiveScan256(sum, sw, &total) + uint32_t(offsets[blockIdx.x]);
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n)
        {
            a[base + k] = int32_t(run);
            run += v[k];
        }
}

// processNode (csarray.hpp:339-366): the new nodes of old node i, at nodeOps[i] after the exclusive scan
__global__ __launch_bounds__(256) void processNodesKernel(const uint64_t* __restrict__ tree, int numNodes,
                                                          const int32_t* __restrict__ ops, int newNumNodes,
                                                          uint64_t* __restrict__ newTree)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < numNodes; i += gridDim.x * blockDim.x)
    {
        const uint64_t thisNode = tree[i];
        const unsigned level    = treeLevel(tree[i + 1] - thisNode);
        const int op            = ops[i + 1] - ops[i];
        const int base          = ops[i];
        if (op == 1) newTree[base] = thisNode;
        else if (op >= 8)
        {
            const unsigned levelDiff = unsigned(31 - __builtin_clz(unsigned(op))) / 3u; // log8 of 8, 64, 512, 4096
            for (int k = 0; k < op; ++k)
                newTree[base + k] = thisNode + uint64_t(k) * nodeRange(level + levelDiff);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) newTree[newNumNodes] = kKeyEnd;
}

__global__ void clearFlagKernel(int* f) { *f = 0; }

/* ------------------------------------------------------------------ source centers and MAC */

__device__ inline double sum8(double v)
{
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    return v;
}

// |m|-weighted center of each leaf, stored at its internal-layout node as {com, sum|m|} (addBody / normalizeMass,
// source_center.hpp:46-68)
template<class Tc, class Tm, class Tf>
__global__ __launch_bounds__(256) void leafCentersKernel(const Tc* __restrict__ x, const Tc* __restrict__ y,
                                                         const Tc* __restrict__ z, const Tm* __restrict__ m,
                                                         const int32_t* __restrict__ l2i, int numLeaves,
                                                         const uint32_t* __restrict__ layout, Tf* __restrict__ centers)
{
    const int leaf = int((size_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 3), sub = threadIdx.x & 7;
    const bool ok  = leaf < numLeaves;
    uint32_t b = 0, e = 0;
    if (ok) b = layout[leaf], e = layout[leaf + 1];
    double c[4] = {0, 0, 0, 0};
    for (uint32_t p = b + sub; p < e; p += 8)
    {
        const double w = fabs(double(m[p]));
        c[0] += w * double(x[p]);
        c[1] += w * double(y[p]);
        c[2] += w * double(z[p]);
        c[3] += w;
    }
    for (int k = 0; k < 4; ++k)
        c[k] = sum8(c[k]);
    const double inv = c[3] != 0.0 ? 1.0 / c[3] : 1.0;
    if (ok && sub == 0)
    {
        Tf* o = centers + 4 * size_t(l2i[leaf]);
        o[0] = Tf(c[0] * inv), o[1] = Tf(c[1] * inv), o[2] = Tf(c[2] * inv), o[3] = Tf(c[3]);
    }
}

// the cells of one level: the |m|-weighted combination of their 8 children (CombineSourceCenter,
// source_center.hpp:84-98)
template<class T>
__global__ __launch_bounds__(256) void levelCentersKernel(int first, int last, const int32_t* __restrict__ child,
                                                          T* __restrict__ centers)
{
    const int cell = first + int(blockIdx.x * blockDim.x + threadIdx.x);
    if (cell >= last) return;
    const int ch = child[cell];
    if (ch == 0) return;
    double c[4] = {0, 0, 0, 0};
    for (int k = 0; k < 8; ++k)
    {
        const T* s     = centers + 4 * size_t(ch + k);
        const double w = fabs(double(s[3]));
        c[0] += w * double(s[0]);
        c[1] += w * double(s[1]);
        c[2] += w * double(s[2]);
        c[3] += w;
    }
    const double inv = c[3] != 0.0 ? 1.0 / c[3] : 1.0;
    T* o = centers + 4 * size_t(cell);
    o[0] = T(c[0] * inv), o[1] = T(c[1] * inv), o[2] = T(c[2] * inv), o[3] = T(c[3]);
}

// centers[i][3] = squared vector-MAC radius, 0 for empty cells (setMacKernel, source_center_gpu.cu:284-297)
template<class T>
__global__ __launch_bounds__(256) void setMacKernel(const uint64_t* __restrict__ prefixes, int numNodes,
                                                    T* __restrict__ centers, MacBox box)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < numNodes; i += gridDim.x * blockDim.x)
    {
        T* c        = centers + 4 * size_t(i);
        double mac2 = double(c[3]) != 0.0 ? macSquared(prefixes[i], double(c[0]), double(c[1]), double(c[2]), box)
                                          : 0.0;
        c[3] = T(mac2);
    }
}

/* ------------------------------------------------------------------ multipoles */

template<bool MDQ>
constexpr int kN = MDQ ? 12 : 8;

// P2M of each leaf about its stored center (P2M_add / P2M_finalize, cartesian_qpole.hpp:67-117; CartesianMDQpole
// adds the dipole sum m (r - com), :237-295)
template<class Tc, class Tm, class Tf, class Tq, bool MDQ>
__global__ __launch_bounds__(256) void leafMultipolesKernel(const Tc* __restrict__ x, const Tc* __restrict__ y,
                                                            const Tc* __restrict__ z, const Tm* __restrict__ m,
                                                            const int32_t* __restrict__ l2i, int numLeaves,
                                                            const uint32_t* __restrict__ layout,
                                                            const Tf* __restrict__ centers, Tq* __restrict__ mp)
{
    const int leaf = int((size_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 3), sub = threadIdx.x & 7;
    const bool ok  = leaf < numLeaves;
    uint32_t b = 0, e = 0;
    double cx = 0, cy = 0, cz = 0;
    int node = 0;
    if (ok)
    {
        b = layout[leaf], e = layout[leaf + 1];
        node = l2i[leaf];
        const Tf* c = centers + 4 * size_t(node);
        cx = double(c[0]), cy = double(c[1]), cz = double(c[2]);
    }
    double q[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t p = b + sub; p < e; p += 8)
    {
        const double rx = double(x[p]) - cx, ry = double(y[p]) - cy, rz = double(z[p]) - cz;
        const double mi = double(m[p]);
        q[0] += mi;
        q[1] += rx * rx * mi;
        q[2] += rx * ry * mi;
        q[3] += rx * rz * mi;
        q[4] += ry * ry * mi;
        q[5] += ry * rz * mi;
        q[6] += rz * rz * mi;
        if constexpr (MDQ) q[8] += mi * rx, q[9] += mi * ry, q[10] += mi * rz;
    }
    for (int k = 0; k < kN<MDQ>; ++k)
        q[k] = sum8(q[k]);
    const double tr = q[1] + q[4] + q[6];
    q[7] = tr;
    q[1] = 3 * q[1] - tr;
    q[4] = 3 * q[4] - tr;
    q[6] = 3 * q[6] - tr;
    q[2] *= 3, q[3] *= 3, q[5] *= 3;
    if (ok && sub == 0)
    {
        Tq* o = mp + kN<MDQ> * size_t(node);
        for (int k = 0; k < kN<MDQ>; ++k)
            o[k] = Tq(q[k]);
    }
}

// M2M of the cells [first, last) from their 8 children about the stored centers (addQuadrupole / M2M,
// cartesian_qpole.hpp:187-235; CartesianMDQpole :359-417): lane k of a cell's 8-lane group shifts child k
template<class Tf, class Tq, bool MDQ>
__global__ __launch_bounds__(256) void levelMultipolesKernel(int first, int last, const int32_t* __restrict__ child,
                                                             const Tf* __restrict__ centers, Tq* __restrict__ mp)
{
    const int cell = first + int((size_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 3), sub = threadIdx.x & 7;
    const int ch   = cell < last ? child[cell] : 0; // uniform over the 8 lanes of the cell
    if (ch == 0) return;
    const Tf* pc  = centers + 4 * size_t(cell);
    const Tf* sc  = centers + 4 * size_t(ch + sub);
    const Tq* cqp = mp + kN<MDQ> * size_t(ch + sub);
    double cq[12];
    for (int k = 0; k < kN<MDQ>; ++k)
        cq[k] = double(cqp[k]);
    const double rx = double(pc[0]) - double(sc[0]), ry = double(pc[1]) - double(sc[1]),
                 rz = double(pc[2]) - double(sc[2]);
    const double rx2 = rx * rx, ry2 = ry * ry, rz2 = rz * rz;
    const double r2  = (rx2 + ry2 + rz2) * (1.0 / 3.0);
    const double ml  = cq[0] * 3;
    double pq[12];
    if constexpr (!MDQ)
    {
        pq[0] = cq[0];
        pq[1] = cq[1] + ml * (rx2 - r2);
        pq[2] = cq[2] + ml * rx * ry;
        pq[3] = cq[3] + ml * rx * rz;
        pq[4] = cq[4] + ml * (ry2 - r2);
        pq[5] = cq[5] + ml * ry * rz;
        pq[6] = cq[6] + ml * (rz2 - r2);
        pq[7] = cq[7] + ml * r2;
    }
    else
    {
        const double px = cq[8], py = cq[9], pz = cq[10];
        const double dip = rx * px + ry * py + rz * pz;
        pq[0]  = cq[0];
        pq[1]  = cq[1] + ml * (rx2 - r2) - 6.0 * rx * px + 2.0 * dip;
        pq[2]  = cq[2] + ml * rx * ry - 3.0 * (rx * py + ry * px);
        pq[3]  = cq[3] + ml * rx * rz - 3.0 * (rx * pz + rz * px);
        pq[4]  = cq[4] + ml * (ry2 - r2) - 6.0 * ry * py + 2.0 * dip;
        pq[5]  = cq[5] + ml * ry * rz - 3.0 * (ry * pz + rz * py);
        pq[6]  = cq[6] + ml * (rz2 - r2) - 6.0 * rz * pz + 2.0 * dip;
        pq[7]  = cq[7] + ml * r2 - 2.0 * dip;
        pq[8]  = px - cq[0] * rx;
        pq[9]  = py - cq[0] * ry;
        pq[10] = pz - cq[0] * rz;
        pq[11] = 0.0;
    }
    for (int k = 0; k < kN<MDQ>; ++k)
        pq[k] = sum8(pq[k]);
    if (sub == 0)
    {
        Tq* o = mp + kN<MDQ> * size_t(cell);
        for (int k = 0; k < kN<MDQ>; ++k)
            o[k] = Tq(pq[k]);
    }
}

template<class T>
MacBox macBoxOf(const T* limits, float invTheta)
{
    MacBox b;
    b.bits = boxBits(limits);
    for (int d = 0; d < 3; ++d)
    {
        const double lo = double(limits[2 * d]), hi = double(limits[2 * d + 1]);
        b.lo[d]   = lo;
        b.half[d] = 0.5 * (hi - lo) / double(1u << b.bits.b[d]);
    }
    b.invTheta = invTheta;
    return b;
}

unsigned gridFor(const Context* ctx, size_t items, unsigned perBlock = 256)
{
    return std::max(1u, std::min(iceil(items, perBlock), unsigned(ctx->numCUs) * 16u));
}

} // namespace

// the change flag of the last node-operation pass lives in the context (the reference's rebalanceChangeCounter)
static int* stageFlag(Context* ctx) { return static_cast<int*>(ctx->getScratch(kSlotStages, 4096)); }

int nodeCounts(Context* ctx, const uint64_t* tree, int numNodes, const uint64_t* keys, size_t n, unsigned maxCount,
               uint32_t* counts)
{
    if (numNodes < 1 || !tree || !counts || (n && !keys)) return setError(ctx, RYO_ERR_ARG, "node counts: bad args");
    nodeCountsKernel<<<gridFor(ctx, size_t(numNodes)), 256, 0, ctx->stream>>>(tree, numNodes, keys, n, maxCount,
                                                                              counts);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

int nodeOps(Context* ctx, const uint64_t* tree, int numNodes, const uint32_t* counts, unsigned bucket, int32_t* ops,
            int* newNumNodes)
{
    if (numNodes < 1 || !tree || !counts || !ops || !newNumNodes || bucket == 0)
        return setError(ctx, RYO_ERR_ARG, "node ops: bad args");
    int* flag = stageFlag(ctx);
    if (!flag) return RYO_ERR_HIP;
    hipStream_t s = ctx->stream;
    clearFlagKernel<<<1, 1, 0, s>>>(flag);
    nodeOpsKernel<<<gridFor(ctx, size_t(numNodes)), 256, 0, s>>>(tree, numNodes, counts, bucket, ops, flag);
    // exclusive scan over numNodes + 1 entries: ops[numNodes] becomes the new node count
    const size_t n     = size_t(numNodes) + 1;
    const int blocks   = int(iceil(n, kScanBlock));
    auto* sums         = static_cast<int32_t*>(ctx->getScratch(kSlotMisc, size_t(blocks) * sizeof(int32_t) + 64));
    if (!sums) return RYO_ERR_HIP;
    scanSumsKernel<<<blocks, 256, 0, s>>>(ops, n, sums);
    scanOffsetsKernel<<<1, 256, 0, s>>>(sums, blocks);
    scanApplyKernel<<<blocks, 256, 0, s>>>(ops, n, sums);
    RYO_LAUNCH_CHECK(ctx);
    int32_t total = 0;
    RYO_HIP(ctx, hipMemcpyAsync(&total, ops + numNodes, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RYO_HIP(ctx, hipStreamSynchronize(s));
    *newNumNodes = total;
    return RYO_OK;
}

int rebalanceTree(Context* ctx, const uint64_t* tree, int numNodes, int newNumNodes, const int32_t* ops,
                  uint64_t* newTree, int* converged)
{
    if (numNodes < 1 || newNumNodes < 1 || !tree || !ops || !newTree || !converged)
        return setError(ctx, RYO_ERR_ARG, "rebalance: bad args");
    int* flag = stageFlag(ctx);
    if (!flag) return RYO_ERR_HIP;
    hipStream_t s = ctx->stream;
    processNodesKernel<<<gridFor(ctx, size_t(numNodes)), 256, 0, s>>>(tree, numNodes, ops, newNumNodes, newTree);
    RYO_LAUNCH_CHECK(ctx);
    int changed = 0;
    RYO_HIP(ctx, hipMemcpyAsync(&changed, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    RYO_HIP(ctx, hipStreamSynchronize(s));
    *converged = changed == 0;
    return RYO_OK;
}

template<class Tc, class Tm, class Tf>
static int leafCentersT(Context* ctx, const void* x, const void* y, const void* z, const void* m, const int32_t* l2i,
                        int numLeaves, const uint32_t* layout, void* centers)
{
    if (numLeaves <= 0) return RYO_OK;
    leafCentersKernel<Tc, Tm, Tf><<<iceil(size_t(numLeaves) * 8, 256), 256, 0, ctx->stream>>>(
        static_cast<const Tc*>(x), static_cast<const Tc*>(y), static_cast<const Tc*>(z), static_cast<const Tm*>(m),
        l2i, numLeaves, layout, static_cast<Tf*>(centers));
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

int leafCenters(Context* ctx, int types, const void* x, const void* y, const void* z, const void* m,
                const int32_t* l2i, int numLeaves, const uint32_t* layout, void* centers)
{
    if (numLeaves > 0 && (!x || !y || !z || !m || !l2i || !layout || !centers))
        return setError(ctx, RYO_ERR_ARG, "leaf centers: null pointer");
    switch (types)
    {
    case 0: return leafCentersT<float, float, float>(ctx, x, y, z, m, l2i, numLeaves, layout, centers);
    case 1: return leafCentersT<double, double, double>(ctx, x, y, z, m, l2i, numLeaves, layout, centers);
    case 2: return leafCentersT<double, float, double>(ctx, x, y, z, m, l2i, numLeaves, layout, centers);
    default: return setError(ctx, RYO_ERR_ARG, "leaf centers: types %d (0, 1 or 2)", types);
    }
}

int upsweepCenters(Context* ctx, int types, int numLevels, const int32_t* levelRangeHost, const int32_t* child,
                   void* centers)
{
    if (!levelRangeHost || !child || !centers || numLevels < 0 || numLevels > int(kMaxLevel) + 1)
        return setError(ctx, RYO_ERR_ARG, "upsweep centers: bad args");
    if (types < 0 || types > 2) return setError(ctx, RYO_ERR_ARG, "upsweep centers: types %d", types);
    for (int level = numLevels - 1; level >= 0; --level)
    {
        const int first = levelRangeHost[level], last = levelRangeHost[level + 1];
        if (last <= first) continue;
        if (types == 0)
            levelCentersKernel<float><<<iceil(last - first, 256), 256, 0, ctx->stream>>>(
                first, last, child, static_cast<float*>(centers));
        else
            levelCentersKernel<double><<<iceil(last - first, 256), 256, 0, ctx->stream>>>(
                first, last, child, static_cast<double*>(centers));
    }
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

int setMac(Context* ctx, int types, const uint64_t* prefixes, int numNodes, void* centers, float invTheta,
           const void* limits)
{
    if (numNodes <= 0) return RYO_OK;
    if (!prefixes || !centers || !limits) return setError(ctx, RYO_ERR_ARG, "set mac: null pointer");
    if (types == 0)
        setMacKernel<float><<<gridFor(ctx, size_t(numNodes)), 256, 0, ctx->stream>>>(
            prefixes, numNodes, static_cast<float*>(centers), macBoxOf(static_cast<const float*>(limits), invTheta));
    else if (types == 1 || types == 2)
        setMacKernel<double><<<gridFor(ctx, size_t(numNodes)), 256, 0, ctx->stream>>>(
            prefixes, numNodes, static_cast<double*>(centers), macBoxOf(static_cast<const double*>(limits), invTheta));
    else return setError(ctx, RYO_ERR_ARG, "set mac: types %d", types);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

template<class Tc, class Tm, class Tf, class Tq, bool MDQ>
static void launchLeafMultipoles(Context* ctx, const void* x, const void* y, const void* z, const void* m,
                                 const int32_t* l2i, int numLeaves, const uint32_t* layout, const void* centers,
                                 void* mp)
{
    leafMultipolesKernel<Tc, Tm, Tf, Tq, MDQ><<<iceil(size_t(numLeaves) * 8, 256), 256, 0, ctx->stream>>>(
        static_cast<const Tc*>(x), static_cast<const Tc*>(y), static_cast<const Tc*>(z), static_cast<const Tm*>(m),
        l2i, numLeaves, layout, static_cast<const Tf*>(centers), static_cast<Tq*>(mp));
}

int leafMultipoles(Context* ctx, int types, int mdq, const void* x, const void* y, const void* z, const void* m,
                   const int32_t* l2i, int numLeaves, const uint32_t* layout, const void* centers, void* mp)
{
    if (numLeaves <= 0) return RYO_OK;
    if (!x || !y || !z || !m || !l2i || !layout || !centers || !mp)
        return setError(ctx, RYO_ERR_ARG, "leaf multipoles: null pointer");
    const bool q = mdq != 0;
    switch (types)
    {
    case 0:
        q ? launchLeafMultipoles<float, float, float, float, true>(ctx, x, y, z, m, l2i, numLeaves, layout, centers, mp)
          : launchLeafMultipoles<float, float, float, float, false>(ctx, x, y, z, m, l2i, numLeaves, layout, centers,
                                                                    mp);
        break;
    case 1:
        q ? launchLeafMultipoles<double, double, double, double, true>(ctx, x, y, z, m, l2i, numLeaves, layout,
                                                                       centers, mp)
          : launchLeafMultipoles<double, double, double, double, false>(ctx, x, y, z, m, l2i, numLeaves, layout,
                                                                        centers, mp);
        break;
    case 2:
        q ? launchLeafMultipoles<double, float, double, float, true>(ctx, x, y, z, m, l2i, numLeaves, layout, centers,
                                                                     mp)
          : launchLeafMultipoles<double, float, double, float, false>(ctx, x, y, z, m, l2i, numLeaves, layout,
                                                                      centers, mp);
        break;
    default: return setError(ctx, RYO_ERR_ARG, "leaf multipoles: types %d (0, 1 or 2)", types);
    }
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

template<class Tf, class Tq, bool MDQ>
static void launchLevelMultipoles(Context* ctx, int first, int last, const int32_t* child, const void* centers,
                                  void* mp)
{
    levelMultipolesKernel<Tf, Tq, MDQ><<<iceil(size_t(last - first) * 8, 256), 256, 0, ctx->stream>>>(
        first, last, child, static_cast<const Tf*>(centers), static_cast<Tq*>(mp));
}

int upsweepMultipoles(Context* ctx, int types, int mdq, int first, int last, const int32_t* child, const void* centers,
                      void* mp)
{
    if (last <= first) return RYO_OK;
    if (first < 0 || !child || !centers || !mp) return setError(ctx, RYO_ERR_ARG, "upsweep multipoles: bad args");
    const bool q = mdq != 0;
    switch (types)
    {
    case 0:
        q ? launchLevelMultipoles<float, float, true>(ctx, first, last, child, centers, mp)
          : launchLevelMultipoles<float, float, false>(ctx, first, last, child, centers, mp);
        break;
    case 1:
        q ? launchLevelMultipoles<double, double, true>(ctx, first, last, child, centers, mp)
          : launchLevelMultipoles<double, double, false>(ctx, first, last, child, centers, mp);
        break;
    case 2:
        q ? launchLevelMultipoles<double, float, true>(ctx, first, last, child, centers, mp)
          : launchLevelMultipoles<double, float, false>(ctx, first, last, child, centers, mp);
        break;
    default: return setError(ctx, RYO_ERR_ARG, "upsweep multipoles: types %d (0, 1 or 2)", types);
    }
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

} // namespace ryo
