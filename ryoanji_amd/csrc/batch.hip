// Batches of particle systems (gfx950): many systems packed one after the other and solved together, every stage
// one launch sequence for all of them, driven by per-system offset tables (BatchView). A small system alone
// occupies a few of the 256 CUs and pays ~90 dependent launches and several host round trips per solve; packed
// into a batch, the whole batch pays them once.
//
// Per step: keys with each system's box -> one stable radix sort of all keys followed by a stable pass on the
// system of each particle (grouping by system, keys ascending inside) -> gather -> leaves of every system (leaf
// tiles never straddle systems) -> per-system leaf-slot, node and parent offsets computed on the device ->
// linked octrees (the node prefixes sorted the same segmented way) -> FP64 upsweep, level by level across all
// systems -> fixed 64-target groups per system -> one traversal launch over all groups -> per-system FP64
// potentials. Every index stored in a system's slice is local to that system, so each system gets exactly the
// keys, tree, multipoles and forces it gets when solved on its own.
//
// Two ways to run it. ryo_batch_solve sizes the node arrays exactly on its first call (the leaf-slot offsets are
// read back once) and after that for capacities, launching every stage without a host round trip and reading
// the overflow flag, tables, statistics and potentials once at the end (a solve that outgrew them is redone
// exactly). ryo_batch_solve_graph runs the whole sequence as one captured HIP graph with no host round trip inside: node
// arrays are sized for a capacity (the totals of the solve that captured the graph plus a margin), the node
// stages read their counts from the device tables, and a replay whose trees outgrow the capacity is detected
// on the device (every later stage then sees empty trees) and solved again the first way; the graph is
// captured again for the grown arrays on the next call.
#include "common.hpp"
#include "sfc.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace ryo
{
size_t sortTempBytes(size_t n);
int gather(Context* ctx, const int32_t* perm, size_t n, int num, const float* const* in, float* const* out);
KeyBox makeKeyBox(const float* limits);
MacBox makeMacBox(const float* limits, float invTheta);
int batchKeys(Context* ctx, const float* x, const float* y, const float* z, size_t n, const uint32_t* P, int S,
              const KeyBox* boxes, uint64_t* keys, int32_t* vals);
int segmentedSortPairs(Context* ctx, uint64_t* keys, int32_t* vals, size_t n, const uint32_t* off, int S,
                       uint64_t* keyBuf, int32_t* valBuf, void* tmp, size_t tmpBytes, const uint32_t* dN,
                       size_t maxSegment);
int copyWords(Context* ctx, void* dst, const void* src, size_t words);
int zeroWords(Context* ctx, void* dst, size_t words);
int batchNodeTables(Context* ctx, const BatchView& bv, uint32_t capSlots, uint32_t capNodes, uint32_t capParents,
                    uint32_t* overflow);
int batchLeavesCount(Context* ctx, const uint64_t* sk, const BatchView& bv, int numTiles, int B,
                     uint32_t* blockSums, uint32_t* LO);
int batchLeavesEmit(Context* ctx, const uint64_t* sk, const BatchView& bv, int numTiles, int B,
                    const uint32_t* blockSums, uint64_t* leaves, uint32_t* layout, uint32_t* counts,
                    uint32_t totalSlots);
int batchUnsortedLayout(Context* ctx, const uint64_t* leaves, const BatchView& bv, uint32_t totalSlots,
                        uint64_t* pre, int32_t* idx);
int batchLinkFinish(Context* ctx, const uint64_t* pre, const int32_t* slot, const BatchView& bv, uint32_t totalNodes,
                    uint32_t totalInternal, uint32_t totalParents, int32_t* i2l, int32_t* l2i, int32_t* LR,
                    uint32_t* LP, int32_t* child, int32_t* parents);
int batchGroups(Context* ctx, const BatchView& bv, uint32_t totalGroups, unsigned gs, uint32_t* gStart,
                uint32_t* gEnd, int32_t* gSys);
int batchTraverse(Context* ctx, int numGroups, const uint32_t* gStart, const uint32_t* gEnd, const int32_t* gSys,
                  const BatchView& bv, const float* G, const float* x, const float* y, const float* z,
                  const float* m, const float* h, size_t n, const int32_t* child, const int32_t* i2l,
                  const uint32_t* layout, const float* centers, const float* mp, float* p, float* ax, float* ay,
                  float* az, double* sysPot, int initNode, void* work, const uint32_t* abort, int mode,
                  const SysImages* images);
size_t batchTraverseWorkBytes(size_t numGroups);
int ewaldBatch(Context* ctx, int K, const uint32_t* first, const uint32_t* last, const float* lim, const double* roots,
               const float* G, const double* settings, const int* mode, const float* x, const float* y,
               const float* z, const float* m, float* ugrav, float* ax, float* ay, float* az, double* pot);

int batchUpsweepEntry(Context* ctx, const float* x, const float* y, const float* z, const float* m,
                      const uint32_t* layout, const uint64_t* prefixes, const int32_t* child, const int32_t* parents,
                      const int32_t* l2i, float* centers, float* mp, double* c64, double* q64, const BatchView& bv,
                      const MacBox* boxes, uint32_t totalSlots, uint32_t totalNodes);

constexpr int kBatchLeafTile = 1024; // particles per leaf tile (the leaf kernels' tile)

// a grow-only device buffer
struct DevBuf
{
    void* p     = nullptr;
    size_t size = 0;
    int ensure(Context* ctx, size_t bytes)
    {
        if (bytes <= size) return RYO_OK;
        if (p) hipFree(p);
        p = nullptr, size = 0;
        size_t sz = std::max<size_t>(bytes + bytes / 8, 256);
        if (hipMalloc(&p, sz) != hipSuccess)
            return setError(ctx, RYO_ERR_HIP, "batch: out of device memory allocating %zu bytes", sz);
        size = sz;
        return RYO_OK;
    }
    template<class T>
    T* as() const
    {
        return static_cast<T*>(p);
    }
    void release()
    {
        if (p) hipFree(p);
        p = nullptr, size = 0;
    }
};

} // namespace ryo

using namespace ryo;

struct ryo_batch
{
    Context* ctx = nullptr;
    size_t capN  = 0;
    int capS     = 0;

    int S    = 0;
    size_t N = 0;
    std::vector<uint32_t> hP, hTS, hLO, hNO, hNIO, hPO, hGO;
    std::vector<float> limits;
    std::vector<int> periodic; // per system (ryo_batch_set_periodic; all open after a load)
    std::vector<int> sysShells, sysEwald; // per-system image shells and Ewald modes (empty: prm's values)
    std::vector<float> hG;     // per system G of the current solve

    // particles: loaded (packed, original order) and sorted (x, y, z, m, h, original index)
    DevBuf in[5], srt[6];
    DevBuf keys, keyBuf, vals, valBuf, sortTmp;
    DevBuf p, ax, ay, az;
    // tables (u32 offsets: P, TS, LO, NO, NIO, PO, GO; LP; LR) and per-system boxes
    DevBuf tab, LP, LR, keyBoxes, macBoxes, G, sysPot, blockSums;
    // trees
    DevBuf leaves, layout, counts, prefixes, child, parents, i2l, l2i, centers, mp, c64, q64;
    DevBuf nodeIdx, nodeKeyBuf, nodeValBuf, nodeTmp;
    DevBuf gStart, gEnd, gSys;
    DevBuf travWork, flags; // traversal work arrays; flags[0]: the trees outgrew the node arrays
    DevBuf images, roots;   // per-system periodic images; the root multipoles of the periodic systems
    std::vector<double> hPot;
    uint32_t totalSlots = 0, totalNodes = 0, totalInternal = 0, totalParents = 0, totalGroups = 0;
    bool solved = false;

    // the captured pipeline: the graph, what it was captured for, the capacities its node arrays have
    hipStream_t capStream = nullptr;
    hipGraph_t graph      = nullptr;
    hipGraphExec_t exec   = nullptr;
    std::vector<uint64_t> graphKey;
    uint64_t graphBufs = 0;
    uint32_t capSlots = 0, capNodes = 0, capParents = 0;
    uint64_t captures = 0, replays = 0, fallbacks = 0;
    // the eager solve on capacity-sized arrays: pinned memory its one read lands in, and the solves it had to
    // redo the exact way because the trees outgrew the capacities
    void* hRead            = nullptr;
    uint64_t sizedFallbacks = 0;
    int capBasisS          = 0; // systems and particles of the exact solve the capacities were taken from
    size_t capBasisN       = 0;

    uint32_t* T(int k) const { return tab.as<uint32_t>() + size_t(k) * (capS + 1); }
    BatchView view() const
    {
        BatchView v;
        v.S = S;
        v.P = T(0), v.TS = T(1), v.LO = T(2), v.NO = T(3), v.NIO = T(4), v.PO = T(5), v.GO = T(6);
        v.LP = LP.as<uint32_t>();
        v.LR = LR.as<int32_t>();
        return v;
    }
};

namespace
{
void allBufs(ryo_batch* b, std::vector<DevBuf*>& out)
{
    for (auto& d : b->in) out.push_back(&d);
    for (auto& d : b->srt) out.push_back(&d);
    DevBuf* rest[] = {&b->keys,     &b->keyBuf,   &b->vals,     &b->valBuf,     &b->sortTmp,    &b->p,
                      &b->ax,       &b->ay,       &b->az,       &b->tab,        &b->LP,         &b->LR,
                      &b->keyBoxes, &b->macBoxes, &b->G,        &b->sysPot,     &b->blockSums,  &b->leaves,
                      &b->layout,   &b->counts,   &b->prefixes, &b->child,      &b->parents,    &b->i2l,
                      &b->l2i,      &b->centers,  &b->mp,       &b->c64,        &b->q64,        &b->nodeIdx,
                      &b->nodeKeyBuf, &b->nodeValBuf, &b->nodeTmp, &b->gStart,  &b->gEnd,       &b->gSys,
                      &b->travWork, &b->flags,    &b->images,   &b->roots};
    for (DevBuf* d : rest) out.push_back(d);
}

std::vector<uint32_t> prefix(const std::vector<uint32_t>& v)
{
    std::vector<uint32_t> r(v.size() + 1, 0);
    for (size_t i = 0; i < v.size(); ++i)
        r[i + 1] = r[i] + v[i];
    return r;
}

// what a solve needs from the host: launch sizes and the per-system parameters that were uploaded
struct Shape
{
    int B          = 64;
    unsigned gs    = 64;
    int initNode   = 1;
    int numTiles   = 0;
    bool periodic  = false; // some system is periodic: per-system images (and the Ewald correction when asked)
};

// host tables known before the build (particle offsets P, leaf tiles TS, groups GO) and the per-system key and
// MAC boxes and G, uploaded to the device; validates the parameters
int uploadTables(ryo_batch* b, const ryo_params* prm, const float* theta, const float* Gs, Shape& sh)
{
    Context* ctx   = b->ctx;
    hipStream_t st = ctx->stream;
    const int S    = b->S;
    sh.B  = int(prm->bucket_size ? prm->bucket_size : 64);
    sh.gs = prm->group_size ? prm->group_size : 64;
    sh.initNode = prm->init_node;
    if (sh.gs > 64) return setError(ctx, RYO_ERR_UNSUPPORTED, "batch: group size %u > 64", sh.gs);
    if (prm->init_node != 0 && prm->init_node != 1) return setError(ctx, RYO_ERR_ARG, "batch: init_node 0 or 1");
    if (prm->num_shells < 0 || prm->ewald < 0 || prm->ewald > 2)
        return setError(ctx, RYO_ERR_ARG, "batch: num_shells >= 0 and ewald 0, 1 or 2");
    if (prm->soft_mac != 0) return setError(ctx, RYO_ERR_UNSUPPORTED, "batch: the reference MAC only");
    if (prm->group_split_tol != 0.0f) return setError(ctx, RYO_ERR_UNSUPPORTED, "batch: fixed target groups only");
    std::vector<uint32_t> tiles(S), ngroups(S);
    std::vector<KeyBox> kb(S);
    std::vector<MacBox> mb(S);
    std::vector<float> gv(S);
    for (int s = 0; s < S; ++s)
    {
        uint32_t n = b->hP[s + 1] - b->hP[s];
        tiles[s]   = (n + kBatchLeafTile - 1) / kBatchLeafTile;
        ngroups[s] = (n + sh.gs - 1) / sh.gs;
        float th   = theta ? theta[s] : prm->theta;
        if (!(th > 0)) return setError(ctx, RYO_ERR_ARG, "batch: theta of system %d", s);
        kb[s] = makeKeyBox(&b->limits[6 * s]);
        mb[s] = makeMacBox(&b->limits[6 * s], 1.0f / th);
        gv[s] = Gs ? Gs[s] : prm->G;
    }
    b->hG = gv;
    // periodic systems: (2 num_shells + 1)^3 images of their box (as a resident periodic system); open ones none
    sh.periodic = std::any_of(b->periodic.begin(), b->periodic.end(), [](int v) { return v != 0; });
    if (sh.periodic)
    {
        std::vector<SysImages> im(S);
        for (int s = 0; s < S; ++s)
        {
            const float* l = &b->limits[6 * s];
            const int sh  = b->sysShells.empty() ? prm->num_shells : b->sysShells[s];
            const int ew  = b->sysEwald.empty() ? prm->ewald : b->sysEwald[s];
            if (sh < 0) return setError(ctx, RYO_ERR_ARG, "batch: num_shells of system %d", s);
            // the Ewald correction's requirements, checked before the solve (ewald.cuh:25-27, ewald.hpp:53)
            if (b->periodic[s] && ew != 0 && !(l[1] - l[0] == l[3] - l[2] && l[3] - l[2] == l[5] - l[4]))
                return setError(ctx, RYO_ERR_UNSUPPORTED, "batch: Ewald gravity requires cubic boxes (system %d)", s);
            if (b->periodic[s] && ew != 0 && std::ceil(prm->ewald_settings[2]) > 3.0)
                return setError(ctx, RYO_ERR_UNSUPPORTED, "batch: Ewald hCut %g > 3", prm->ewald_settings[2]);
            im[s] = SysImages{b->periodic[s] ? sh : 0, {l[1] - l[0], l[3] - l[2], l[5] - l[4]}};
        }
        int rc = b->images.ensure(ctx, S * sizeof(SysImages));
        if (rc) return rc;
        RYO_HIP(ctx, hipMemcpyAsync(b->images.p, im.data(), S * sizeof(SysImages), hipMemcpyHostToDevice, st));
    }
    b->hTS = prefix(tiles);
    b->hGO = prefix(ngroups);
    sh.numTiles = int(b->hTS[S]);
    b->totalGroups = b->hGO[S];
    int rc = b->blockSums.ensure(ctx, (size_t(sh.numTiles) + 1) * 4);
    if (!rc) rc = b->flags.ensure(ctx, 256);
    for (DevBuf* d : {&b->gStart, &b->gEnd, &b->gSys})
        if (!rc) rc = d->ensure(ctx, (size_t(b->totalGroups) + 1) * 4);
    if (!rc) rc = b->travWork.ensure(ctx, batchTraverseWorkBytes(b->totalGroups));
    if (rc) return rc;
    RYO_HIP(ctx, hipMemcpyAsync(b->T(0), b->hP.data(), (S + 1) * 4, hipMemcpyHostToDevice, st));
    RYO_HIP(ctx, hipMemcpyAsync(b->T(1), b->hTS.data(), (S + 1) * 4, hipMemcpyHostToDevice, st));
    RYO_HIP(ctx, hipMemcpyAsync(b->T(6), b->hGO.data(), (S + 1) * 4, hipMemcpyHostToDevice, st));
    RYO_HIP(ctx, hipMemcpyAsync(b->keyBoxes.p, kb.data(), S * sizeof(KeyBox), hipMemcpyHostToDevice, st));
    RYO_HIP(ctx, hipMemcpyAsync(b->macBoxes.p, mb.data(), S * sizeof(MacBox), hipMemcpyHostToDevice, st));
    RYO_HIP(ctx, hipMemcpyAsync(b->G.p, gv.data(), S * 4, hipMemcpyHostToDevice, st));
    return RYO_OK;
}

// the largest system of the batch (particles)
size_t maxSystem(const ryo_batch* b)
{
    uint32_t m = 0;
    for (size_t s = 0; s + 1 < b->hP.size(); ++s)
        m = std::max(m, b->hP[s + 1] - b->hP[s]);
    return m;
}

// a bound on the octree nodes of a system of n particles with bucket B: a node is internal only if it holds more
// than B particles, so a level has at most min(8^l, n / (B + 1)) internal nodes, and nodes = 8 internal + 1
size_t maxNodes(size_t n, unsigned B)
{
    size_t internal = 0, perLevel = n / (size_t(B) + 1), cells = 1;
    for (unsigned l = 0; l < kMaxLevel; ++l, cells = std::min(cells * 8, size_t(1) << 40))
        internal += std::min(cells, perLevel);
    return 8 * internal + 1;
}

// keys, segmented sort, gather, leaf counts and offsets, node tables (device); no host synchronization
int runFront(ryo_batch* b, const Shape& sh, uint32_t capSlots, uint32_t capNodes, uint32_t capParents)
{
    Context* ctx = b->ctx;
    const int S = b->S;
    const size_t N = b->N;
    BatchView bv = b->view();
    auto* keys = b->keys.as<uint64_t>();
    auto* vals = b->vals.as<int32_t>();
    const float* in[5] = {b->in[0].as<float>(), b->in[1].as<float>(), b->in[2].as<float>(), b->in[3].as<float>(),
                          b->in[4].as<float>()};
    float* out[5] = {b->srt[0].as<float>(), b->srt[1].as<float>(), b->srt[2].as<float>(), b->srt[3].as<float>(),
                     b->srt[4].as<float>()};
    int rc;
    if ((rc = batchKeys(ctx, in[0], in[1], in[2], N, bv.P, S, b->keyBoxes.as<KeyBox>(), keys, vals))) return rc;
    if ((rc = segmentedSortPairs(ctx, keys, vals, N, bv.P, S, b->keyBuf.as<uint64_t>(), b->valBuf.as<int32_t>(),
                                 b->sortTmp.p, b->sortTmp.size, nullptr, maxSystem(b))))
        return rc;
    if ((rc = gather(ctx, vals, N, 5, in, out))) return rc;
    if ((rc = copyWords(ctx, b->srt[5].p, vals, N))) return rc;
    if ((rc = batchLeavesCount(ctx, keys, bv, sh.numTiles, sh.B, b->blockSums.as<uint32_t>(), b->T(2)))) return rc;
    return batchNodeTables(ctx, bv, capSlots, capNodes, capParents, b->flags.as<uint32_t>());
}

// node arrays for the given totals (leaf slots, nodes, parents)
int ensureNodes(ryo_batch* b, uint32_t slots1, uint32_t nodes1, uint32_t parents1)
{
    Context* ctx = b->ctx;
    const size_t slots = size_t(slots1) + 1, nn = size_t(nodes1) + 1, np = size_t(parents1) + 1;
    int rc = RYO_OK;
    if (!rc) rc = b->leaves.ensure(ctx, slots * 8);
    if (!rc) rc = b->layout.ensure(ctx, slots * 4);
    if (!rc) rc = b->counts.ensure(ctx, slots * 4);
    for (DevBuf* d : {&b->prefixes, &b->nodeKeyBuf})
        if (!rc) rc = d->ensure(ctx, nn * 8);
    for (DevBuf* d : {&b->child, &b->i2l, &b->l2i, &b->nodeIdx, &b->nodeValBuf})
        if (!rc) rc = d->ensure(ctx, nn * 4);
    if (!rc) rc = b->parents.ensure(ctx, np * 4);
    if (!rc) rc = b->centers.ensure(ctx, nn * 16);
    if (!rc) rc = b->mp.ensure(ctx, nn * 32);
    if (!rc) rc = b->c64.ensure(ctx, nn * 32);
    if (!rc) rc = b->q64.ensure(ctx, nn * 64);
    if (!rc) rc = b->nodeTmp.ensure(ctx, sortTempBytes(nn));
    return rc;
}

// the batch traversal with the batch's arrays (mode 0: launch and finish, 1: launch only, 2: finish only)
int traversal(ryo_batch* b, const Shape& sh, int mode, const uint32_t* abort)
{
    const float* srt[5] = {b->srt[0].as<float>(), b->srt[1].as<float>(), b->srt[2].as<float>(),
                           b->srt[3].as<float>(), b->srt[4].as<float>()};
    return batchTraverse(b->ctx, int(b->totalGroups), b->gStart.as<uint32_t>(), b->gEnd.as<uint32_t>(),
                         b->gSys.as<int32_t>(), b->view(), b->G.as<float>(), srt[0], srt[1], srt[2], srt[3], srt[4],
                         b->N, b->child.as<int32_t>(), b->i2l.as<int32_t>(), b->layout.as<uint32_t>(),
                         b->centers.as<float>(), b->mp.as<float>(), b->p.as<float>(), b->ax.as<float>(),
                         b->ay.as<float>(), b->az.as<float>(), b->sysPot.as<double>(), sh.initNode, b->travWork.p,
                         abort, mode, sh.periodic ? b->images.as<SysImages>() : nullptr);
}

// leaves, linked octrees, upsweep, groups, zeroed outputs and the traversal. slots/nodes/internal/parents:
// exact totals (eager) or capacities (captured; the node sort then reads its count on the device)
int runBack(ryo_batch* b, const Shape& sh, uint32_t slots, uint32_t nodes, uint32_t internal, uint32_t parents,
            bool captured)
{
    Context* ctx = b->ctx;
    const int S = b->S;
    BatchView bv = b->view();
    auto* keys = b->keys.as<uint64_t>();
    int rc;
    if ((rc = batchLeavesEmit(ctx, keys, bv, sh.numTiles, sh.B, b->blockSums.as<uint32_t>(), b->leaves.as<uint64_t>(),
                              b->layout.as<uint32_t>(), b->counts.as<uint32_t>(), slots)))
        return rc;
    auto* pre = b->prefixes.as<uint64_t>();
    auto* idx = b->nodeIdx.as<int32_t>();
    if ((rc = batchUnsortedLayout(ctx, b->leaves.as<uint64_t>(), bv, slots, pre, idx))) return rc;
    if ((rc = segmentedSortPairs(ctx, pre, idx, nodes, bv.NO, S, b->nodeKeyBuf.as<uint64_t>(),
                                 b->nodeValBuf.as<int32_t>(), b->nodeTmp.p, b->nodeTmp.size,
                                 captured ? bv.NO + S : nullptr, maxNodes(maxSystem(b), sh.B))))
        return rc;
    if ((rc = batchLinkFinish(ctx, pre, idx, bv, nodes, internal, parents, b->i2l.as<int32_t>(),
                              b->l2i.as<int32_t>(), b->LR.as<int32_t>(), b->LP.as<uint32_t>(), b->child.as<int32_t>(),
                              b->parents.as<int32_t>())))
        return rc;
    const float* srt[4] = {b->srt[0].as<float>(), b->srt[1].as<float>(), b->srt[2].as<float>(),
                           b->srt[3].as<float>()};
    if ((rc = batchUpsweepEntry(ctx, srt[0], srt[1], srt[2], srt[3], b->layout.as<uint32_t>(), pre,
                                b->child.as<int32_t>(), b->parents.as<int32_t>(), b->l2i.as<int32_t>(),
                                b->centers.as<float>(), b->mp.as<float>(), b->c64.as<double>(), b->q64.as<double>(),
                                bv, b->macBoxes.as<MacBox>(), slots, nodes)))
        return rc;
    if ((rc = batchGroups(ctx, bv, b->totalGroups, sh.gs, b->gStart.as<uint32_t>(), b->gEnd.as<uint32_t>(),
                          b->gSys.as<int32_t>())))
        return rc;
    for (DevBuf* d : {&b->p, &b->ax, &b->ay, &b->az})
        if ((rc = zeroWords(ctx, d->p, b->N))) return rc;
    return traversal(b, sh, captured ? 1 : 0, captured ? b->flags.as<uint32_t>() : nullptr);
}

// node, internal-node and parent totals of every system from its leaf-slot offsets (the formula of the
// device's node-table kernel)
void hostNodeTables(ryo_batch* b)
{
    const int S = b->S;
    std::vector<uint32_t> nodes(S), internal(S), parentsN(S);
    for (int s = 0; s < S; ++s)
    {
        uint32_t L  = b->hLO[s + 1] - b->hLO[s] - 1;
        internal[s] = (L - 1) / 7;
        nodes[s]    = L + internal[s];
        parentsN[s] = std::max(1u, (nodes[s] - 1) / 8);
    }
    b->hNO  = prefix(nodes);
    b->hNIO = prefix(internal);
    b->hPO  = prefix(parentsN);
    b->totalSlots    = b->hLO[S];
    b->totalNodes    = b->hNO[S];
    b->totalInternal = b->hNIO[S];
    b->totalParents  = b->hPO[S];
}

int capacities(ryo_batch* b, uint32_t* capSlots, uint32_t* capNodes, uint32_t* capParents);

// the exact solve: the leaf-slot offsets are read back once to size the node arrays. The arrays are grown to the
// capacities these totals give (not only to the totals), so the capacity-sized solves and graphs that follow
// find them large enough: growing an array frees the old one, and hipFree waits for the whole device, the other
// streams' work included
int solveEager(ryo_batch* b, const Shape& sh)
{
    Context* ctx = b->ctx;
    const int S  = b->S;
    int rc = runFront(b, sh, ~0u, ~0u, ~0u);
    if (rc) return rc;
    b->hLO.resize(S + 1);
    RYO_HIP(ctx, hipMemcpyAsync(b->hLO.data(), b->T(2), (S + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
    RYO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    hostNodeTables(b);
    uint32_t cs, cn, cp;
    if ((rc = capacities(b, &cs, &cn, &cp))) return rc;
    if ((rc = ensureNodes(b, cs, cn, cp))) return rc;
    return runBack(b, sh, b->totalSlots, b->totalNodes, b->totalInternal, b->totalParents, false);
}

// the root multipole {Mroot[8], center[3], 0} of each listed system (node 0 of its slice)
__global__ void batchRootsKernel(const double* __restrict__ c64, const double* __restrict__ q64,
                                 const uint32_t* __restrict__ NO, const int32_t* __restrict__ sys, int K,
                                 double* __restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const size_t node = NO[sys[k]];
    for (int j = 0; j < 8; ++j)
        out[12 * k + j] = q64[8 * node + j];
    for (int j = 0; j < 3; ++j)
        out[12 * k + 8 + j] = c64[4 * node + j];
    out[12 * k + 11] = 0.0;
}

// the Ewald correction of the periodic systems (after the traversal), as a resident periodic system takes it;
// ewPot[s] receives each system's correction potential (0 for the others)
int batchEwald(ryo_batch* b, const ryo_params* prm, std::vector<double>& ewPot)
{
    Context* ctx = b->ctx;
    const int S  = b->S;
    ewPot.assign(S, 0.0);
    std::vector<int32_t> sys;
    std::vector<int> mode;
    for (int s = 0; s < S; ++s)
    {
        const int e = b->sysEwald.empty() ? prm->ewald : b->sysEwald[s];
        if (b->periodic[s] && e != 0) sys.push_back(s), mode.push_back(e);
    }
    const int K = int(sys.size());
    if (K == 0) return RYO_OK;
    int rc = b->roots.ensure(ctx, size_t(K) * (12 * 8 + 4));
    if (rc) return rc;
    auto* dRoots = b->roots.as<double>();
    auto* dSys   = reinterpret_cast<int32_t*>(dRoots + 12 * size_t(K));
    hipStream_t st = ctx->stream;
    RYO_HIP(ctx, hipMemcpyAsync(dSys, sys.data(), K * 4, hipMemcpyHostToDevice, st));
    batchRootsKernel<<<iceil(K, 64), 64, 0, st>>>(b->c64.as<double>(), b->q64.as<double>(), b->T(3), dSys, K, dRoots);
    RYO_LAUNCH_CHECK(ctx);
    std::vector<double> roots(12 * size_t(K));
    RYO_HIP(ctx, hipMemcpyAsync(roots.data(), dRoots, roots.size() * 8, hipMemcpyDeviceToHost, st));
    RYO_HIP(ctx, hipStreamSynchronize(st));
    std::vector<uint32_t> first(K), last(K);
    std::vector<float> lim(6 * size_t(K)), G(K);
    for (int k = 0; k < K; ++k)
    {
        const int s = sys[k];
        first[k] = b->hP[s], last[k] = b->hP[s + 1];
        std::memcpy(&lim[6 * k], &b->limits[6 * s], 6 * sizeof(float));
        G[k] = b->hG[s];
    }
    std::vector<double> pot(K);
    rc = ewaldBatch(ctx, K, first.data(), last.data(), lim.data(), roots.data(), G.data(), prm->ewald_settings,
                    mode.data(), b->srt[0].as<float>(), b->srt[1].as<float>(), b->srt[2].as<float>(),
                    b->srt[3].as<float>(), b->p.as<float>(), b->ax.as<float>(), b->ay.as<float>(), b->az.as<float>(),
                    pot.data());
    if (rc) return rc;
    for (int k = 0; k < K; ++k)
        ewPot[sys[k]] = pot[k];
    return RYO_OK;
}

// per-system potentials to the host (plus the Ewald potentials of periodic systems, when given)
int finishSolve(ryo_batch* b, const Shape&, double* pot, const std::vector<double>* ewPot = nullptr,
                bool potRead = false)
{
    Context* ctx = b->ctx;
    if (!potRead)
    {
        b->hPot.resize(b->S);
        RYO_HIP(ctx, hipMemcpyAsync(b->hPot.data(), b->sysPot.p, b->S * 8, hipMemcpyDeviceToHost, ctx->stream));
        RYO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (ewPot)
        for (int s = 0; s < b->S; ++s)
            b->hPot[s] += (*ewPot)[s];
    if (pot) std::memcpy(pot, b->hPot.data(), b->S * sizeof(double));
    b->solved = true;
    return RYO_OK;
}

// what a captured graph depends on: sizes and launch parameters (theta, G, boxes and positions are read from
// device memory at replay)
std::vector<uint64_t> graphKey(const ryo_batch* b, const Shape& sh)
{
    const Context* ctx = b->ctx;
    std::vector<uint64_t> k = {uint64_t(b->S), uint64_t(b->N), uint64_t(sh.B), uint64_t(sh.gs),
                               uint64_t(sh.initNode), uint64_t(ctx->travStack), uint64_t(ctx->travBlocksPerCU),
                               uint64_t(ctx->travRanges), uint64_t(ctx->travWaves)};
    k.insert(k.end(), b->hP.begin(), b->hP.end());
    return k;
}

// the device addresses a graph holds (any reallocation makes it stale)
uint64_t bufferSignature(ryo_batch* b)
{
    std::vector<DevBuf*> bufs;
    allBufs(b, bufs);
    uint64_t h = 1469598103934665603ull;
    for (DevBuf* d : bufs)
        h = (h ^ uint64_t(reinterpret_cast<uintptr_t>(d->p))) * 1099511628211ull;
    h = (h ^ uint64_t(reinterpret_cast<uintptr_t>(b->ctx->dStats))) * 1099511628211ull;
    return h;
}

void dropGraph(ryo_batch* b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr, b->graph = nullptr;
    b->graphKey.clear();
}

// node-array capacities from the last exact solve's totals: a quarter more and 8 leaf slots per system (the exact
// solve grows the arrays to them before it writes them)
int capacities(ryo_batch* b, uint32_t* capSlots, uint32_t* capNodes, uint32_t* capParents)
{
    const uint64_t slots64 = uint64_t(b->totalSlots) + b->totalSlots / 4 + 8ull * uint64_t(b->S) + 64;
    if (slots64 > 0x7fffffffull) return setError(b->ctx, RYO_ERR_CAPACITY, "batch: too many leaf slots");
    *capSlots   = uint32_t(slots64);
    *capNodes   = *capSlots + *capSlots / 7 + uint32_t(b->S);
    *capParents = *capNodes / 8 + uint32_t(b->S);
    return RYO_OK;
}

// the eager solve with no host round trip before its end (the stages a graph captures, launched for the
// capacities), then one read of the overflow flag, the leaf-slot/node/internal/parent tables, the traversal's
// statistics and the per-system potentials. The first solve, and one whose trees outgrew the capacities (every
// stage after the tables then saw empty trees), run the exact way and set the capacities for the next ones; so
// does a batch with more systems or particles than the one the capacities were taken from (a member joined the
// serving lane), whose trees would most likely not fit. *potRead: the potentials are in hPot. Groups that
// overflowed a shallow LIFO are retried by the traversal's own finish.
int solveSized(ryo_batch* b, const Shape& sh, bool* potRead)
{
    Context* ctx   = b->ctx;
    const int S    = b->S;
    hipStream_t st = ctx->stream;
    *potRead       = false;
    int rc;
    auto exact = [&]() {
        int r = solveEager(b, sh);
        if (!r) r = capacities(b, &b->capSlots, &b->capNodes, &b->capParents);
        b->capBasisS = S, b->capBasisN = b->N;
        return r;
    };
    if (b->capSlots == 0 || b->capNodes == 0 || S > b->capBasisS || b->N > b->capBasisN) return exact();
    if ((rc = ensureNodes(b, b->capSlots, b->capNodes, b->capParents))) return rc;
    if ((rc = runFront(b, sh, b->capSlots, b->capNodes, b->capParents))) return rc;
    if ((rc = runBack(b, sh, b->capSlots, b->capNodes, b->capSlots / 7 + 1, b->capParents, true))) return rc;
    const size_t S1 = size_t(S) + 1;
    // layout of the read: stats[7] (u64), potentials[S] (f64), overflow flag, tables LO, NO, NIO, PO (u32)
    const size_t bytes = 7 * 8 + size_t(b->capS) * 8 + 4 + 4 * (size_t(b->capS) + 1) * 4;
    if (!b->hRead && hipHostMalloc(&b->hRead, bytes) != hipSuccess)
    {
        b->hRead = nullptr;
        return setError(ctx, RYO_ERR_HIP, "batch: pinned host memory");
    }
    auto* hs   = static_cast<unsigned long long*>(b->hRead);
    auto* hp   = reinterpret_cast<double*>(hs + 7);
    auto* hf   = reinterpret_cast<uint32_t*>(hp + b->capS);
    auto* htab = hf + 1;
    RYO_HIP(ctx, hipMemcpyAsync(hs, ctx->dStats + 8, 7 * 8, hipMemcpyDeviceToHost, st));
    RYO_HIP(ctx, hipMemcpyAsync(hp, b->sysPot.p, size_t(S) * 8, hipMemcpyDeviceToHost, st));
    RYO_HIP(ctx, hipMemcpyAsync(hf, b->flags.p, 4, hipMemcpyDeviceToHost, st));
    for (int k = 0; k < 4; ++k)
        RYO_HIP(ctx, hipMemcpyAsync(htab + k * S1, b->T(2 + k), S1 * 4, hipMemcpyDeviceToHost, st));
    RYO_HIP(ctx, hipStreamSynchronize(st));
    if (*hf)
    {
        ++b->sizedFallbacks;
        return exact();
    }
    b->hLO.assign(htab, htab + S1);
    b->hNO.assign(htab + S1, htab + 2 * S1);
    b->hNIO.assign(htab + 2 * S1, htab + 3 * S1);
    b->hPO.assign(htab + 3 * S1, htab + 4 * S1);
    b->totalSlots = b->hLO[S], b->totalNodes = b->hNO[S];
    b->totalInternal = b->hNIO[S], b->totalParents = b->hPO[S];
    if ((hs[5] & 0xffffffffu) != 0 || hs[6] != 0) return traversal(b, sh, 2, nullptr);
    for (int k = 0; k < 6; ++k)
        ctx->hStats[k] = k < 5 ? hs[k] : 0;
    ctx->retriedGroups = 0;
    b->hPot.assign(hp, hp + S);
    *potRead = true;
    return RYO_OK;
}

// capture the whole solve (after the host tables) as one graph, with node arrays sized for capacities
int capture(ryo_batch* b, const Shape& sh)
{
    Context* ctx = b->ctx;
    dropGraph(b);
    uint32_t capSlots, capNodes, capParents;
    int rc = capacities(b, &capSlots, &capNodes, &capParents);
    if (!rc) rc = ensureNodes(b, capSlots, capNodes, capParents);
    if (rc) return rc;
    if (!b->capStream)
        RYO_HIP(ctx, hipStreamCreateWithFlags(&b->capStream, hipStreamNonBlocking));
    RYO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    hipStream_t user = ctx->stream;
    ctx->stream      = b->capStream;
    hipError_t e     = hipStreamBeginCapture(b->capStream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess)
    {
        ctx->stream = user;
        return setError(ctx, RYO_ERR_HIP, "batch graph: begin capture: %s", hipGetErrorString(e));
    }
    rc = runFront(b, sh, capSlots, capNodes, capParents);
    if (!rc)
        rc = runBack(b, sh, capSlots, capNodes, capSlots / 7 + 1, capParents, true);
    hipGraph_t g = nullptr;
    e            = hipStreamEndCapture(b->capStream, &g);
    ctx->stream  = user;
    if (rc || e != hipSuccess || !g)
    {
        if (g) hipGraphDestroy(g);
        return rc ? rc : setError(ctx, RYO_ERR_HIP, "batch graph: end capture: %s", hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess)
    {
        hipGraphDestroy(g);
        b->exec = nullptr;
        return setError(ctx, RYO_ERR_HIP, "batch graph: instantiate: %s", hipGetErrorString(e));
    }
    b->graph     = g;
    b->graphKey  = graphKey(b, sh);
    b->graphBufs = bufferSignature(b);
    b->capSlots = capSlots, b->capNodes = capNodes, b->capParents = capParents;
    b->capBasisS = b->S, b->capBasisN = b->N;
    ++b->captures;
    return RYO_OK;
}
} // namespace

namespace ryo
{
/* per-system boundaries for a caller that mixes them in one batch (the serving lane): periodic flags, image
 * shells and Ewald modes (0 none, 1 the reference's potential, 2 the consistent one) of the loaded systems;
 * they replace ryo_batch_set_periodic's flags and prm's num_shells and ewald until the next load */
int batchSetBoundaries(ryo_batch* b, const int* periodic, const int* shells, const int* ewald)
{
    if (!b || !periodic || !shells || !ewald || b->S < 1) return RYO_ERR_ARG;
    for (int s = 0; s < b->S; ++s)
        if (shells[s] < 0 || ewald[s] < 0 || ewald[s] > 2)
            return setError(b->ctx, RYO_ERR_ARG, "batch: boundaries of system %d", s);
    b->periodic.assign(periodic, periodic + b->S);
    b->sysShells.assign(shells, shells + b->S);
    b->sysEwald.assign(ewald, ewald + b->S);
    b->solved = false;
    return RYO_OK;
}
} // namespace ryo

extern "C" {

int ryo_batch_create(ryo_ctx* h, size_t max_particles, int max_systems, ryo_batch** out)
{
    if (!h || !out || max_particles == 0 || max_systems < 1) return RYO_ERR_ARG;
    Context* ctx = &h->c;
    if (max_systems > kMaxSortSegments)
        return setError(ctx, RYO_ERR_UNSUPPORTED, "batch: at most %d systems (%d asked)", kMaxSortSegments,
                        max_systems);
    auto* b      = new ryo_batch();
    b->ctx  = ctx;
    b->capN = max_particles;
    b->capS = max_systems;
    const size_t n = max_particles, S1 = size_t(max_systems) + 1;
    int rc = RYO_OK;
    for (auto& d : b->in)
        if (!rc) rc = d.ensure(ctx, n * 4);
    for (auto& d : b->srt)
        if (!rc) rc = d.ensure(ctx, n * 4);
    for (DevBuf* d : {&b->keys, &b->keyBuf})
        if (!rc) rc = d->ensure(ctx, n * 8);
    for (DevBuf* d : {&b->vals, &b->valBuf, &b->p, &b->ax, &b->ay, &b->az})
        if (!rc) rc = d->ensure(ctx, n * 4);
    if (!rc) rc = b->sortTmp.ensure(ctx, sortTempBytes(n));
    if (!rc) rc = b->tab.ensure(ctx, 7 * S1 * 4);
    if (!rc) rc = b->LP.ensure(ctx, (kMaxLevel + 1) * S1 * 4);
    if (!rc) rc = b->LR.ensure(ctx, 23 * S1 * 4);
    if (!rc) rc = b->keyBoxes.ensure(ctx, S1 * sizeof(KeyBox));
    if (!rc) rc = b->macBoxes.ensure(ctx, S1 * sizeof(MacBox));
    if (!rc) rc = b->G.ensure(ctx, S1 * 4);
    if (!rc) rc = b->sysPot.ensure(ctx, S1 * 8);
    if (!rc) rc = b->blockSums.ensure(ctx, (n / kBatchLeafTile + S1 + 1) * 4);
    // node arrays for a typical tree of the full capacity (n / 8 leaf slots and 8 per system), so that solves
    // rarely grow them (a growth frees the old array, which waits for the whole device)
    if (!rc)
    {
        const uint64_t slots = n / 8 + 8 * S1 + 64;
        const uint64_t nodes = slots + slots / 7 + S1;
        if (nodes < 0x7fffffffull) rc = ensureNodes(b, uint32_t(slots), uint32_t(nodes), uint32_t(nodes / 8 + S1));
    }
    if (rc)
    {
        ryo_batch_destroy(b);
        return rc;
    }
    *out = b;
    return RYO_OK;
}

int ryo_batch_destroy(ryo_batch* b)
{
    if (!b) return RYO_OK;
    hipStreamSynchronize(b->ctx->stream);
    dropGraph(b);
    if (b->capStream) hipStreamDestroy(b->capStream);
    if (b->hRead) hipHostFree(b->hRead);
    std::vector<DevBuf*> bufs;
    allBufs(b, bufs);
    for (DevBuf* d : bufs)
        d->release();
    delete b;
    return RYO_OK;
}

int ryo_batch_load(ryo_batch* b, int num_systems, const uint32_t* counts, const float* x, const float* y,
                   const float* z, const float* m, const float* hh, const float* limits)
{
    if (!b || !counts || !x || !y || !z || !m || !hh || !limits) return RYO_ERR_ARG;
    Context* ctx = b->ctx;
    if (num_systems < 1 || num_systems > b->capS)
        return setError(ctx, RYO_ERR_CAPACITY, "batch: %d systems (capacity %d)", num_systems, b->capS);
    size_t N = 0;
    for (int s = 0; s < num_systems; ++s)
    {
        if (counts[s] == 0) return setError(ctx, RYO_ERR_ARG, "batch: system %d is empty", s);
        for (int d = 0; d < 3; ++d)
            if (!(limits[6 * s + 2 * d + 1] > limits[6 * s + 2 * d]))
                return setError(ctx, RYO_ERR_ARG, "batch: system %d has an empty box", s);
        N += counts[s];
    }
    if (N > b->capN) return setError(ctx, RYO_ERR_CAPACITY, "batch: %zu particles (capacity %zu)", N, b->capN);
    b->S = num_systems;
    b->N = N;
    b->hP = prefix(std::vector<uint32_t>(counts, counts + num_systems));
    b->limits.assign(limits, limits + 6 * num_systems);
    b->periodic.assign(num_systems, 0);
    b->sysShells.clear(), b->sysEwald.clear();
    const float* src[5] = {x, y, z, m, hh};
    for (int k = 0; k < 5; ++k)
        RYO_HIP(ctx, hipMemcpyAsync(b->in[k].p, src[k], N * 4, hipMemcpyDefault, ctx->stream));
    b->solved = false;
    return RYO_OK;
}

int ryo_batch_set_periodic(ryo_batch* b, const int* periodic)
{
    if (!b || !periodic || b->S < 1) return RYO_ERR_ARG;
    for (int s = 0; s < b->S; ++s)
        b->periodic[s] = periodic[s] != 0;
    b->sysShells.clear(), b->sysEwald.clear();
    b->solved = false;
    return RYO_OK;
}

int ryo_batch_solve(ryo_batch* b, const ryo_params* prm, const float* theta, const float* Gs, double* pot)
{
    if (!b || !prm || b->S < 1) return RYO_ERR_ARG;
    if (prm->multipole != 0)
        return setError(b->ctx, RYO_ERR_UNSUPPORTED, "batch: CartesianQuadrupole multipoles only");
    Shape sh;
    std::vector<double> ewPot;
    bool potRead = false;
    int rc = uploadTables(b, prm, theta, Gs, sh);
    if (!rc) rc = solveSized(b, sh, &potRead);
    if (!rc && sh.periodic) rc = batchEwald(b, prm, ewPot);
    if (!rc) rc = finishSolve(b, sh, pot, sh.periodic ? &ewPot : nullptr, potRead);
    return rc;
}

int ryo_batch_solve_graph(ryo_batch* b, const ryo_params* prm, const float* theta, const float* Gs, double* pot)
{
    if (!b || !prm || b->S < 1) return RYO_ERR_ARG;
    Context* ctx = b->ctx;
    if (prm->multipole != 0) return setError(ctx, RYO_ERR_UNSUPPORTED, "batch: CartesianQuadrupole multipoles only");
    Shape sh;
    int rc = uploadTables(b, prm, theta, Gs, sh);
    if (rc) return rc;
    if (sh.periodic)
        return setError(ctx, RYO_ERR_UNSUPPORTED,
                        "batch graph: open boundaries only (periodic systems take ryo_batch_solve)");
    const bool current = b->exec && b->graphKey == graphKey(b, sh) && b->graphBufs == bufferSignature(b);
    if (!current)
    {
        // an exact solve sizes the capacities, then the graph is captured for them and replayed below
        if ((rc = solveEager(b, sh))) return rc;
        if ((rc = capture(b, sh))) return rc;
    }
    RYO_HIP(ctx, hipGraphLaunch(b->exec, ctx->stream));
    ++b->replays;
    uint32_t overflow = 0;
    RYO_HIP(ctx, hipMemcpyAsync(&overflow, b->flags.p, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    RYO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (overflow)
    {
        // the trees outgrew the captured capacities: solve exactly (the arrays grow, so the next call captures
        // again)
        ++b->fallbacks;
        if ((rc = solveEager(b, sh))) return rc;
        return finishSolve(b, sh, pot);
    }
    // host copies of the device-computed tables (the exact solve reads them before building)
    const size_t S1 = size_t(b->S) + 1;
    b->hLO.resize(S1), b->hNO.resize(S1), b->hNIO.resize(S1), b->hPO.resize(S1);
    RYO_HIP(ctx, hipMemcpyAsync(b->hLO.data(), b->T(2), S1 * 4, hipMemcpyDeviceToHost, ctx->stream));
    RYO_HIP(ctx, hipMemcpyAsync(b->hNO.data(), b->T(3), S1 * 4, hipMemcpyDeviceToHost, ctx->stream));
    RYO_HIP(ctx, hipMemcpyAsync(b->hNIO.data(), b->T(4), S1 * 4, hipMemcpyDeviceToHost, ctx->stream));
    RYO_HIP(ctx, hipMemcpyAsync(b->hPO.data(), b->T(5), S1 * 4, hipMemcpyDeviceToHost, ctx->stream));
    if ((rc = traversal(b, sh, 2, nullptr))) return rc;
    b->totalSlots = b->hLO[b->S], b->totalNodes = b->hNO[b->S];
    b->totalInternal = b->hNIO[b->S], b->totalParents = b->hPO[b->S];
    return finishSolve(b, sh, pot);
}

/* {graphs captured, replays, replays solved again exactly because the trees outgrew the captured arrays,
 * leaf-slot capacity of the current graph} */
int ryo_batch_graph_stats(ryo_batch* b, uint64_t* out)
{
    if (!b || !out) return RYO_ERR_ARG;
    out[0] = b->captures, out[1] = b->replays, out[2] = b->fallbacks, out[3] = b->capSlots;
    return RYO_OK;
}

/* sorted particles and results (device, packed: system s at [P[s], P[s+1])): out[0..3] = x, y, z, m, out[4] = h,
 * out[5..8] = p, ax, ay, az; order = global index into the loaded arrays */
int ryo_batch_results(ryo_batch* b, float** out, int32_t** order)
{
    if (!b || !out || !order || !b->solved) return RYO_ERR_ARG;
    for (int k = 0; k < 5; ++k)
        out[k] = b->srt[k].as<float>();
    out[5] = b->p.as<float>(), out[6] = b->ax.as<float>(), out[7] = b->ay.as<float>(), out[8] = b->az.as<float>();
    *order = b->srt[5].as<int32_t>();
    return RYO_OK;
}

/* host copies of the offset tables, each num_systems + 1 entries: P (particles), LO (leaf slots: L_s leaves +
 * the terminator), NO (nodes), PO (parents), GO (groups) */
int ryo_batch_offsets(ryo_batch* b, uint32_t* P, uint32_t* LO, uint32_t* NO, uint32_t* PO, uint32_t* GO)
{
    if (!b || !b->solved) return RYO_ERR_ARG;
    const size_t sz = (b->S + 1) * sizeof(uint32_t);
    if (P) std::memcpy(P, b->hP.data(), sz);
    if (LO) std::memcpy(LO, b->hLO.data(), sz);
    if (NO) std::memcpy(NO, b->hNO.data(), sz);
    if (PO) std::memcpy(PO, b->hPO.data(), sz);
    if (GO) std::memcpy(GO, b->hGO.data(), sz);
    return RYO_OK;
}

/* device arrays of the trees (indices local to each system's slice): {keys (sorted), leaves, layout, counts,
 * prefixes, childOffsets, parents, internalToLeaf, leafToInternal, levelRange (23 per system), centers (float4),
 * multipoles (8 floats), centers64, multipoles64} */
int ryo_batch_tree(ryo_batch* b, void** a)
{
    if (!b || !a || !b->solved) return RYO_ERR_ARG;
    void* v[14] = {b->keys.p, b->leaves.p, b->layout.p, b->counts.p, b->prefixes.p, b->child.p, b->parents.p,
                   b->i2l.p,  b->l2i.p,    b->LR.p,     b->centers.p, b->mp.p,     b->c64.p,    b->q64.p};
    std::memcpy(a, v, sizeof(v));
    return RYO_OK;
}

} // extern "C"
