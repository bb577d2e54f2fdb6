// FP64 upsweep (gfx950): leaf mass centers and P2M, |m|-weighted centers and M2M of internal nodes up to the
// root, and the vector-MAC radius.
//
// Each leaf is reduced by 8 lanes (particles strided by 8, shuffle reduction). Internal nodes follow level by
// level from the deepest to the root, 8 lanes per node (lane k reads child k, coalesced); each level launch
// reads its node range from the device copy of the level ranges, so the sequence has no host round trip.
// (Kernel boundaries order the levels; a single-launch bottom-up climb with device-scope acquire/release
// atomics was 20x slower here, as every release writes back the L2.)
//
// Formulas: addBody/normalizeMass/CombineSourceCenter (cstone/focus/source_center.hpp:46-98),
// P2M_add/P2M_finalize (ryoanji/nbody/cartesian_qpole.hpp:67-117), addQuadrupole (:187-210),
// computeVecMacR2 (cstone/traversal/macs.hpp:64-82) with hilbertIBox/centerAndSize (hilbert.hpp:549-585,
// box.hpp:357-381). All arithmetic is FP64; the traversal copies are rounded to FP32 once.
#include "common.hpp"
#include "mac.hpp"
#include "sfc.hpp"

namespace ryo
{

// particles of type Tin (float, or double for the double-coordinate entry point); outputs c32, q32 and cmac are
// optional (NULL: not written)
template<class Tin>
struct UpsweepArgsT
{
    const Tin *x, *y, *z, *m;
    const uint32_t* layout;
    int numLeaves;
    const int32_t* leafNode; // leafToInternal + numInternal
    const int32_t* parents;
    const int32_t* child;
    const uint64_t* prefixes;
    float4* c32;
    float* q32;
    double* c64; // [n][4] {com, sum|m|}
    double* q64; // [n][8]
    MacBox box;
    double4* cmac = nullptr; // [n] {com, mac^2} in double (SourceCenterType<double>)
    // a device-sized build's counts (BuildStat): the leaves are the first stat[kStLc] (at most numLeaves) and
    // leafNode is leafToInternal, offset by stat[kStInternal] on the device
    const uint32_t* dStat = nullptr;
};
// multipole values per node: 8 (CartesianQuadrupole), 12 (CartesianMDQpole: the quadrupole's 8, px, py, pz, 0)
template<bool MDQ>
constexpr int kMpN = MDQ ? 12 : 8;
using UpsweepArgs = UpsweepArgsT<float>;

template<bool MDQ = false, class Tin>
__device__ inline void storeNode(const UpsweepArgsT<Tin>& a, int node, const double c[4], const double* q)
{
    constexpr int N = kMpN<MDQ>;
    double* c64 = a.c64 + 4 * size_t(node);
    double* q64 = a.q64 + N * size_t(node);
    for (int k = 0; k < 4; ++k)
        c64[k] = c[k];
    for (int k = 0; k < N; ++k)
        q64[k] = q[k];
    double mac2 = c[3] != 0.0 ? macSquared(a.prefixes[node], c[0], c[1], c[2], a.box) : 0.0;
    if (a.c32) a.c32[node] = make_float4(float(c[0]), float(c[1]), float(c[2]), float(mac2));
    if (a.q32)
    {
        float* q32 = a.q32 + N * size_t(node);
        for (int k = 0; k < N; ++k)
            q32[k] = float(q[k]);
    }
    if (a.cmac) a.cmac[node] = make_double4(c[0], c[1], c[2], mac2);
}

// value of another lane of the same 8-lane group through DPP (both halves of the double); the 8 lanes of a
// group are always active or inactive together, so every lane read is active
template<int Ctrl>
__device__ __forceinline__ double dppMove(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), Ctrl, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), Ctrl, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}

// sum over the 8-lane group, in every lane: pairs (quad_perm [1,0,3,2]), quads ([2,3,0,1]), then the other quad
// of the group (row_half_mirror: lane l reads 7 - l); the same additions as a butterfly over lane ^ 1, ^ 2, ^ 4
__device__ inline double red8(double v)
{
    v += dppMove<0xB1>(v);
    v += dppMove<0x4E>(v);
    v += dppMove<0x141>(v);
    return v;
}

// leaves: 8 lanes per leaf reduce its particles (|m|-weighted center, then the quadrupole about it; MDQ adds the
// dipole sum m r about the center, P2M for CartesianMDQpole, cartesian_qpole.hpp:237-295); all 8 lanes of a group
// call with the same leaf and `ok`
template<bool MDQ = false, class Tin>
__device__ inline void leafUpsweepOne(const UpsweepArgsT<Tin>& a, int leaf, bool ok, int sub)
{
    uint32_t b = 0, e = 0;
    if (ok) b = a.layout[leaf], e = a.layout[leaf + 1];

    double c[4]  = {0, 0, 0, 0};
    double q[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    auto center = [&](Tin px, Tin py, Tin pz, Tin pm) {
        double w = fabs(double(pm));
        c[0] += w * double(px);
        c[1] += w * double(py);
        c[2] += w * double(pz);
        c[3] += w;
    };
    auto moments = [&](Tin px, Tin py, Tin pz, Tin pm) {
        double rx = double(px) - c[0], ry = double(py) - c[1], rz = double(pz) - c[2];
        double mi = double(pm);
        q[0] += mi;
        q[1] += rx * rx * mi;
        q[2] += rx * ry * mi;
        q[3] += rx * rz * mi;
        q[4] += ry * ry * mi;
        q[5] += ry * rz * mi;
        q[6] += rz * rz * mi;
        if constexpr (MDQ) q[8] += mi * rx, q[9] += mi * ry, q[10] += mi * rz;
    };
    auto centerDone = [&]() {
        for (int k = 0; k < 4; ++k)
            c[k] = red8(c[k]);
        double inv = c[3] != 0.0 ? 1.0 / c[3] : 1.0;
        c[0] *= inv, c[1] *= inv, c[2] *= inv;
    };
    // leaves of at most 64 particles (every leaf of a bucket of at most 64 but those at the deepest level): the
    // lane's 8 particles are loaded at once and kept in registers for both sums, which add the same terms in the
    // same order as the loops below (the slots past the leaf are skipped)
    constexpr int kRegSlots = 8;
    if (e - b <= uint32_t(8 * kRegSlots))
    {
        Tin px[kRegSlots], py[kRegSlots], pz[kRegSlots], pm[kRegSlots];
#pragma unroll
        for (int i = 0; i < kRegSlots; ++i)
        {
            const uint32_t p = b + sub + 8 * i;
            const bool in    = p < e;
            px[i] = in ? a.x[p] : Tin(0);
            py[i] = in ? a.y[p] : Tin(0);
            pz[i] = in ? a.z[p] : Tin(0);
            pm[i] = in ? a.m[p] : Tin(0);
        }
#pragma unroll
        for (int i = 0; i < kRegSlots; ++i)
            if (b + sub + 8 * i < e) center(px[i], py[i], pz[i], pm[i]);
        centerDone();
#pragma unroll
        for (int i = 0; i < kRegSlots; ++i)
            if (b + sub + 8 * i < e) moments(px[i], py[i], pz[i], pm[i]);
    }
    else
    {
        for (uint32_t p = b + sub; p < e; p += 8)
            center(a.x[p], a.y[p], a.z[p], a.m[p]);
        centerDone();
        for (uint32_t p = b + sub; p < e; p += 8)
            moments(a.x[p], a.y[p], a.z[p], a.m[p]);
    }
    for (int k = 0; k < 7; ++k)
        q[k] = red8(q[k]);
    if constexpr (MDQ)
        for (int k = 8; k < 11; ++k)
            q[k] = red8(q[k]);
    double tr = q[1] + q[4] + q[6];
    q[7] = tr;
    q[1] = 3 * q[1] - tr;
    q[4] = 3 * q[4] - tr;
    q[6] = 3 * q[6] - tr;
    q[2] *= 3, q[3] *= 3, q[5] *= 3;

    if (ok && sub == 0) storeNode<MDQ>(a, a.leafNode[leaf], c, q);
}

// one 8-lane group per leaf, grid-stride (the 8 lanes of a group step together)
template<class Tin, bool MDQ = false>
__global__ __launch_bounds__(256) void leafUpsweepKernel(UpsweepArgsT<Tin> a)
{
    if (a.dStat)
    {
        a.numLeaves = min(int(a.dStat[kStLc]), a.numLeaves);
        a.leafNode += a.dStat[kStInternal];
    }
    const int groups = int((size_t(gridDim.x) * blockDim.x) >> 3);
    const int group  = int((size_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 3);
    // the loop bound is uniform over the grid, so the 8-lane reductions always see all 8 lanes
    for (int base = 0; base < a.numLeaves; base += groups)
        leafUpsweepOne<MDQ>(a, base + group, base + group < a.numLeaves, threadIdx.x & 7);
}

// one internal node: lane k of the 8-lane group holds child k; center and M2M by 8-lane reductions. MDQ follows
// addQuadrupole for CartesianMDQpole (cartesian_qpole.hpp:359-403): the child's dipole p about its own center
// shifts the trace by -2 p.dX, the diagonal by -6 dX_i p_i + 2 p.dX, the off-diagonal by -3 (dX_i p_j + dX_j p_i),
// and the parent dipole collects p - M dX (dX = parent center - child center)
template<bool MDQ = false, class Tin>
__device__ inline void levelNodeOne(const UpsweepArgsT<Tin>& a, int node, int sub)
{
    {
        const int ch = a.child[node]; // uniform over the 8 lanes of the node
        if (ch == 0) return;
        const double* s  = a.c64 + 4 * size_t(ch + sub);
        const double* cq = a.q64 + kMpN<MDQ> * size_t(ch + sub);
        double sx = s[0], sy = s[1], sz = s[2], w = fabs(s[3]);
        double pc[4] = {red8(w * sx), red8(w * sy), red8(w * sz), red8(w)};
        double pinv  = pc[3] != 0.0 ? 1.0 / pc[3] : 1.0;
        pc[0] *= pinv, pc[1] *= pinv, pc[2] *= pinv;
        double rx = pc[0] - sx, ry = pc[1] - sy, rz = pc[2] - sz;
        double rx2 = rx * rx, ry2 = ry * ry, rz2 = rz * rz;
        double r2  = (rx2 + ry2 + rz2) * (1.0 / 3.0);
        double ml  = cq[0] * 3;
        double pq[12];
        if constexpr (!MDQ)
        {
            pq[0] = red8(cq[0]);
            pq[1] = red8(cq[1] + ml * (rx2 - r2));
            pq[2] = red8(cq[2] + ml * rx * ry);
            pq[3] = red8(cq[3] + ml * rx * rz);
            pq[4] = red8(cq[4] + ml * (ry2 - r2));
            pq[5] = red8(cq[5] + ml * ry * rz);
            pq[6] = red8(cq[6] + ml * (rz2 - r2));
            pq[7] = red8(cq[7] + ml * r2);
        }
        else
        {
            const double px = cq[8], py = cq[9], pz = cq[10];
            const double dip = rx * px + ry * py + rz * pz;
            pq[0]  = red8(cq[0]);
            pq[1]  = red8(cq[1] + ml * (rx2 - r2) - 6.0 * rx * px + 2.0 * dip);
            pq[2]  = red8(cq[2] + ml * rx * ry - 3.0 * (rx * py + ry * px));
            pq[3]  = red8(cq[3] + ml * rx * rz - 3.0 * (rx * pz + rz * px));
            pq[4]  = red8(cq[4] + ml * (ry2 - r2) - 6.0 * ry * py + 2.0 * dip);
            pq[5]  = red8(cq[5] + ml * ry * rz - 3.0 * (ry * pz + rz * py));
            pq[6]  = red8(cq[6] + ml * (rz2 - r2) - 6.0 * rz * pz + 2.0 * dip);
            pq[7]  = red8(cq[7] + ml * r2 - 2.0 * dip);
            pq[8]  = red8(px - cq[0] * rx);
            pq[9]  = red8(py - cq[0] * ry);
            pq[10] = red8(pz - cq[0] * rz);
            pq[11] = 0.0;
        }
        if (sub == 0) storeNode<MDQ>(a, node, pc, pq);
    }
}

// internal nodes of one level
template<class Tin, bool MDQ = false>
__global__ __launch_bounds__(256) void levelUpsweepKernel(UpsweepArgsT<Tin> a, const int32_t* __restrict__ levelRange,
                                                          int level)
{
    const int first = levelRange[level], last = levelRange[level + 1];
    const int groupsPerGrid = int(gridDim.x * blockDim.x) >> 3;
    for (int node = first + int((size_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 3); node < last;
         node += groupsPerGrid)
        levelNodeOne<MDQ>(a, node, threadIdx.x & 7);
}

// a small tree in one launch: one block of 1024 threads (128 groups of 8 lanes) forms the leaves, then every level
// from the deepest possible parent level to the root, with a barrier between levels (the waves of one work-group
// share the CU's L1, so a level's stores are visible to the next level's loads after it). The same per-node
// arithmetic as leafUpsweepKernel and levelUpsweepKernel, so the results are identical; it saves the 21 level
// launches that dominate a small tree's upsweep.
constexpr int kBlockUpsweepThreads = 1024;
constexpr int kBlockUpsweepLeaves  = 2048;

template<class Tin, bool MDQ = false>
__global__ __launch_bounds__(kBlockUpsweepThreads) void blockUpsweepKernel(UpsweepArgsT<Tin> a,
                                                                          const int32_t* __restrict__ levelRange)
{
    if (a.dStat)
    {
        a.numLeaves = min(int(a.dStat[kStLc]), a.numLeaves);
        a.leafNode += a.dStat[kStInternal];
    }
    constexpr int groups = kBlockUpsweepThreads / 8;
    __shared__ int32_t lr[kMaxLevel + 2];
    if (threadIdx.x < kMaxLevel + 2) lr[threadIdx.x] = levelRange[threadIdx.x];
    const int group = threadIdx.x >> 3, sub = threadIdx.x & 7;
    for (int base = 0; base < a.numLeaves; base += groups)
        leafUpsweepOne<MDQ>(a, base + group, base + group < a.numLeaves, sub);
    __syncthreads();
    for (int l = int(kMaxLevel) - 1; l >= 0; --l)
    {
        const int first = lr[l], last = lr[l + 1];
        for (int node = first + group; node < last; node += groups)
            levelNodeOne<MDQ>(a, node, sub);
        __syncthreads();
    }
}

/* ---------------------------------------------------------------- batches of systems */

// the arguments of system s: the batch-wide arrays offset to the system's slices (local indices inside)
__device__ inline UpsweepArgs systemArgs(const UpsweepArgs& g, const BatchView& bv, const MacBox* boxes, int s)
{
    UpsweepArgs a  = g;
    const size_t p = bv.P[s], no = bv.NO[s];
    a.x = g.x + p, a.y = g.y + p, a.z = g.z + p, a.m = g.m + p;
    a.layout    = g.layout + bv.LO[s];
    a.numLeaves = int(bv.LO[s + 1] - bv.LO[s]) - 1;
    a.leafNode  = g.leafNode + no + (bv.NIO[s + 1] - bv.NIO[s]);
    a.parents   = g.parents + bv.PO[s];
    a.child     = g.child + no;
    a.prefixes  = g.prefixes + no;
    a.c32 = g.c32 + no, a.q32 = g.q32 + 8 * no, a.c64 = g.c64 + 4 * no, a.q64 = g.q64 + 8 * no;
    a.box = boxes[s];
    return a;
}

// 8 lanes per leaf slot of every system (terminator slots idle)
__global__ __launch_bounds__(256) void batchLeafUpsweepKernel(UpsweepArgs g, BatchView bv,
                                                              const MacBox* __restrict__ boxes)
{
    const uint32_t j = uint32_t((size_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 3);
    const uint32_t total = bv.LO[bv.S];
    const int s = segmentOf(bv.LO, bv.S, j < total ? j : total - 1);
    UpsweepArgs a = systemArgs(g, bv, boxes, s);
    const int t   = int(j - bv.LO[s]);
    leafUpsweepOne(a, t, j < total && t < a.numLeaves, threadIdx.x & 7);
}

// level `level` of every system: node group q is node LR[s][level] + q - LP[level][s] of system s
__global__ __launch_bounds__(256) void batchLevelUpsweepKernel(UpsweepArgs g, BatchView bv,
                                                               const MacBox* __restrict__ boxes, int level)
{
    const uint32_t* lp   = bv.LP + size_t(bv.S + 1) * level;
    const uint32_t total = lp[bv.S];
    const uint32_t per   = uint32_t(gridDim.x * blockDim.x) >> 3;
    for (uint32_t q = uint32_t((size_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 3); q < total; q += per)
    {
        const int s   = segmentOf(lp, bv.S, q);
        UpsweepArgs a = systemArgs(g, bv, boxes, s);
        levelNodeOne(a, bv.LR[23 * s + level] + int(q - lp[s]), threadIdx.x & 7);
    }
}

/* upsweep of every system of a batch: g holds the batch-wide arrays (particles, leaf slots, node arrays),
 * boxes the per-system MAC boxes (device) */
int batchUpsweep(Context* ctx, const UpsweepArgs& g, const BatchView& bv, const MacBox* boxes, uint32_t totalSlots,
                 uint32_t totalNodes)
{
    hipStream_t s = ctx->stream;
    if (totalSlots == 0) return RYO_OK;
    batchLeafUpsweepKernel<<<iceil(size_t(totalSlots) * 8, 256), 256, 0, s>>>(g, bv, boxes);
    const unsigned blocks = std::max(1u, std::min(iceil(size_t(totalNodes) * 8, 256), unsigned(ctx->numCUs * 4)));
    for (int l = int(kMaxLevel) - 1; l >= 0; --l)
        batchLevelUpsweepKernel<<<blocks, 256, 0, s>>>(g, bv, boxes, l);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

int batchUpsweepEntry(Context* ctx, const float* x, const float* y, const float* z, const float* m,
                      const uint32_t* layout, const uint64_t* prefixes, const int32_t* child, const int32_t* parents,
                      const int32_t* l2i, float* centers, float* mp, double* c64, double* q64, const BatchView& bv,
                      const MacBox* boxes, uint32_t totalSlots, uint32_t totalNodes)
{
    UpsweepArgs g{};
    g.x = x, g.y = y, g.z = z, g.m = m;
    g.layout = layout, g.numLeaves = 0;
    g.leafNode = l2i, g.parents = parents, g.child = child, g.prefixes = prefixes;
    g.c32 = reinterpret_cast<float4*>(centers), g.q32 = mp, g.c64 = c64, g.q64 = q64;
    return batchUpsweep(ctx, g, bv, boxes, totalSlots, totalNodes);
}

MacBox makeMacBox(const float* limits, float invTheta)
{
    MacBox b;
    b.bits = boxBits(limits);
    for (int d = 0; d < 3; ++d)
    {
        double lo = limits[2 * d], hi = limits[2 * d + 1];
        b.lo[d]   = lo;
        b.half[d] = 0.5 * (hi - lo) / double(1u << b.bits.b[d]);
    }
    b.invTheta = invTheta;
    return b;
}

// leaf kernel, then one kernel per level from the deepest possible parent level to the root; each launch reads
// its node range on the device, so the sequence needs no host copy of the level ranges (and can be captured)
template<class Tin, class Tl, bool MDQ = false>
static int upsweepT(Context* ctx, UpsweepArgsT<Tin> a, int L, const int32_t* levelRange, const Tl* limits,
                    float invTheta, double softReach)
{
    if (L < 1) return setError(ctx, RYO_ERR_ARG, "upsweep: need at least one leaf");
    if (!levelRange) return setError(ctx, RYO_ERR_ARG, "upsweep: level ranges are required");
    int nI = (L - 1) / 7, n = L + nI;
    size_t need = 256;
    if (!a.c64) need += size_t(n) * 4 * sizeof(double);
    if (!a.q64) need += size_t(n) * kMpN<MDQ> * sizeof(double);
    auto* buf = static_cast<char*>(ctx->getScratch(kSlotUpsweep, need));
    if (!buf) return RYO_ERR_HIP;
    size_t off = 0;
    if (!a.c64) a.c64 = reinterpret_cast<double*>(buf + off), off += size_t(n) * 4 * sizeof(double);
    if (!a.q64) a.q64 = reinterpret_cast<double*>(buf + off), off += size_t(n) * kMpN<MDQ> * sizeof(double);
    a.numLeaves = L;
    if (!a.dStat) a.leafNode += nI; // leafToInternal + numInternal (device-sized builds: offset on the device)
    const Bits3 bits = boxBits(limits);
    a.box.bits       = bits;
    for (int d = 0; d < 3; ++d)
    {
        double lo = limits[2 * d], hi = limits[2 * d + 1];
        a.box.lo[d]   = lo;
        a.box.half[d] = 0.5 * (hi - lo) / double(1u << bits.b[d]);
    }
    a.box.invTheta  = invTheta;
    a.box.softReach = softReach;
    hipStream_t s = ctx->stream;
    if (L <= kBlockUpsweepLeaves)
    {
        blockUpsweepKernel<Tin, MDQ><<<1, kBlockUpsweepThreads, 0, s>>>(a, levelRange);
        RYO_LAUNCH_CHECK(ctx);
        return RYO_OK;
    }
    leafUpsweepKernel<Tin, MDQ><<<std::max(1u, std::min(iceil(size_t(L) * 8, 256), unsigned(ctx->numCUs * 64))), 256,
                                   0, s>>>(a);
    const unsigned levelBlocks = std::max(1u, std::min(iceil(size_t(nI) * 8, 256), unsigned(ctx->numCUs * 4)));
    if (nI > 0)
        for (int l = int(kMaxLevel) - 1; l >= 0; --l)
            levelUpsweepKernel<Tin, MDQ><<<levelBlocks, 256, 0, s>>>(a, levelRange, l);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

int upsweep(Context* ctx, const float* x, const float* y, const float* z, const float* m, const uint32_t* layout,
            int L, const uint64_t* prefixes, const int32_t* child, const int32_t* parents, const int32_t* l2i,
            const int32_t* levelRange, const float* limits, float invTheta, float* centers, float* mpoles, double* c64,
            double* q64, double softReach, const uint32_t* dStat)
{
    UpsweepArgs a{};
    a.dStat = dStat;
    a.x = x, a.y = y, a.z = z, a.m = m;
    a.layout = layout;
    a.leafNode = l2i, a.parents = parents, a.child = child, a.prefixes = prefixes;
    a.c32 = reinterpret_cast<float4*>(centers), a.q32 = mpoles, a.c64 = c64, a.q64 = q64;
    return upsweepT(ctx, a, L, levelRange, limits, invTheta, softReach);
}

// CartesianMDQpole multipoles ([n][12] float, optional [n][12] double copy) from float particles
int upsweepMDQ(Context* ctx, const float* x, const float* y, const float* z, const float* m, const uint32_t* layout,
               int L, const uint64_t* prefixes, const int32_t* child, const int32_t* parents, const int32_t* l2i,
               const int32_t* levelRange, const float* limits, float invTheta, float* centers, float* mpoles,
               double* c64, double* q64, const uint32_t* dStat)
{
    UpsweepArgs a{};
    a.dStat = dStat;
    a.x = x, a.y = y, a.z = z, a.m = m;
    a.layout = layout;
    a.leafNode = l2i, a.parents = parents, a.child = child, a.prefixes = prefixes;
    a.c32 = reinterpret_cast<float4*>(centers), a.q32 = mpoles, a.c64 = c64, a.q64 = q64;
    return upsweepT<float, float, true>(ctx, a, L, levelRange, limits, invTheta, 0.0);
}

// double particles and box: centers {com, mac^2} and multipoles in double (mp32: optional float copy)
int upsweep64(Context* ctx, const double* x, const double* y, const double* z, const double* m,
              const uint32_t* layout, int L, const uint64_t* prefixes, const int32_t* child, const int32_t* parents,
              const int32_t* l2i, const int32_t* levelRange, const double* limits, float invTheta, double* centers,
              double* mpoles, float* mp32)
{
    if (!centers || !mpoles) return setError(ctx, RYO_ERR_ARG, "upsweep: null output");
    UpsweepArgsT<double> a{};
    a.x = x, a.y = y, a.z = z, a.m = m;
    a.layout = layout;
    a.leafNode = l2i, a.parents = parents, a.child = child, a.prefixes = prefixes;
    a.c32 = nullptr, a.q32 = mp32, a.c64 = nullptr, a.q64 = mpoles;
    a.cmac = reinterpret_cast<double4*>(centers);
    return upsweepT(ctx, a, L, levelRange, limits, invTheta, 0.0);
}

// the same with CartesianMDQpole multipoles: mpoles double[n][12], mp32 (optional) float[n][12]
int upsweepMDQ64(Context* ctx, const double* x, const double* y, const double* z, const double* m,
                 const uint32_t* layout, int L, const uint64_t* prefixes, const int32_t* child, const int32_t* parents,
                 const int32_t* l2i, const int32_t* levelRange, const double* limits, float invTheta,
                 double* centers, double* mpoles, float* mp32)
{
    if (!centers || !mpoles) return setError(ctx, RYO_ERR_ARG, "upsweep: null output");
    UpsweepArgsT<double> a{};
    a.x = x, a.y = y, a.z = z, a.m = m;
    a.layout = layout;
    a.leafNode = l2i, a.parents = parents, a.child = child, a.prefixes = prefixes;
    a.c32 = nullptr, a.q32 = mp32, a.c64 = nullptr, a.q64 = mpoles;
    a.cmac = reinterpret_cast<double4*>(centers);
    return upsweepT<double, double, true>(ctx, a, L, levelRange, limits, invTheta, 0.0);
}

} // namespace ryo
