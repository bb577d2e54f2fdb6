// Batches of particle systems (gfx950): many systems packed one after the other and solved together, every stage
// one launch sequence for all of them, driven by per-system offset tables (BatchView). A small system alone
// occupies a few of the 256 CUs and pays ~90 dependent launches and several host round trips per solve; packed
// into a batch, the whole batch pays them once.
//
// Per step: keys with each system's box -> one stable radix sort of all keys followed by a stable pass on the
// system of each particle (grouping by system, keys ascending inside) -> gather -> leaves of every system (leaf
// tiles never straddle systems) -> the leaf-slot offsets are read back (the one host round trip, which sizes the
// node arrays) -> linked octrees (the node prefixes sorted the same segmented way) -> FP64 upsweep, level by
// level across all systems -> fixed 64-target groups per system -> one traversal launch over all groups ->
// per-system FP64 potentials. Every index stored in a system's slice is local to that system, so each system
// gets exactly the keys, tree, multipoles and forces it gets when solved on its own.
#include "common.hpp"
#include "sfc.hpp"

#include <algorithm>
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
                       uint64_t* keyBuf, int32_t* valBuf, uint64_t* segKeys, uint64_t* segBuf, int32_t* pos,
                       int32_t* posBuf, void* tmp, size_t tmpBytes);
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
                  float* az, double* sysPot, int initNode);

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

    // particles: loaded (packed, original order) and sorted (x, y, z, m, h, original index)
    DevBuf in[5], srt[6];
    DevBuf keys, keyBuf, segKeys, segBuf, vals, valBuf, pos, posBuf, sortTmp;
    DevBuf p, ax, ay, az;
    // tables (u32 offsets: P, TS, LO, NO, NIO, PO, GO; LP; LR) and per-system boxes
    DevBuf tab, LP, LR, keyBoxes, macBoxes, G, sysPot, blockSums;
    // trees
    DevBuf leaves, layout, counts, prefixes, child, parents, i2l, l2i, centers, mp, c64, q64;
    DevBuf nodeIdx, nodeKeyBuf, nodeValBuf, nodeSeg, nodeSegBuf, nodePos, nodePosBuf, nodeTmp;
    DevBuf gStart, gEnd, gSys;
    std::vector<double> hPot;
    uint32_t totalSlots = 0, totalNodes = 0, totalInternal = 0, totalParents = 0, totalGroups = 0;
    bool solved = false;

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
    DevBuf* rest[] = {&b->keys,     &b->keyBuf,   &b->segKeys,  &b->segBuf,     &b->vals,       &b->valBuf,
                      &b->pos,      &b->posBuf,   &b->sortTmp,  &b->p,          &b->ax,         &b->ay,
                      &b->az,       &b->tab,      &b->LP,       &b->LR,         &b->keyBoxes,   &b->macBoxes,
                      &b->G,        &b->sysPot,   &b->blockSums, &b->leaves,    &b->layout,     &b->counts,
                      &b->prefixes, &b->child,    &b->parents,  &b->i2l,        &b->l2i,        &b->centers,
                      &b->mp,       &b->c64,      &b->q64,      &b->nodeIdx,    &b->nodeKeyBuf, &b->nodeValBuf,
                      &b->nodeSeg,  &b->nodeSegBuf, &b->nodePos, &b->nodePosBuf, &b->nodeTmp,   &b->gStart,
                      &b->gEnd,     &b->gSys};
    for (DevBuf* d : rest) out.push_back(d);
}

std::vector<uint32_t> prefix(const std::vector<uint32_t>& v)
{
    std::vector<uint32_t> r(v.size() + 1, 0);
    for (size_t i = 0; i < v.size(); ++i)
        r[i + 1] = r[i] + v[i];
    return r;
}
} // namespace

extern "C" {

int ryo_batch_create(ryo_ctx* h, size_t max_particles, int max_systems, ryo_batch** out)
{
    if (!h || !out || max_particles == 0 || max_systems < 1) return RYO_ERR_ARG;
    Context* ctx = &h->c;
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
    for (DevBuf* d : {&b->keys, &b->keyBuf, &b->segKeys, &b->segBuf})
        if (!rc) rc = d->ensure(ctx, n * 8);
    for (DevBuf* d : {&b->vals, &b->valBuf, &b->pos, &b->posBuf, &b->p, &b->ax, &b->ay, &b->az})
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
    const float* src[5] = {x, y, z, m, hh};
    for (int k = 0; k < 5; ++k)
        RYO_HIP(ctx, hipMemcpyAsync(b->in[k].p, src[k], N * 4, hipMemcpyDefault, ctx->stream));
    b->solved = false;
    return RYO_OK;
}

/* theta and G per system (host arrays, NULL: prm's value for every system) */
int ryo_batch_solve(ryo_batch* b, const ryo_params* prm, const float* theta, const float* Gs, double* pot)
{
    if (!b || !prm || b->S < 1) return RYO_ERR_ARG;
    Context* ctx    = b->ctx;
    hipStream_t st  = ctx->stream;
    const int S     = b->S;
    const size_t N  = b->N;
    const int B     = int(prm->bucket_size ? prm->bucket_size : 64);
    const unsigned gs = prm->group_size ? prm->group_size : 64;
    if (gs > 64) return setError(ctx, RYO_ERR_UNSUPPORTED, "batch: group size %u > 64", gs);
    if (prm->init_node != 0 && prm->init_node != 1) return setError(ctx, RYO_ERR_ARG, "batch: init_node 0 or 1");
    if (prm->num_shells != 0 || prm->ewald != 0)
        return setError(ctx, RYO_ERR_UNSUPPORTED, "batch: open boundaries only");
    int rc;

    // per-system tables known before the build: particle offsets, leaf tiles, boxes, G
    std::vector<uint32_t> tiles(S), ngroups(S);
    std::vector<KeyBox> kb(S);
    std::vector<MacBox> mb(S);
    std::vector<float> gv(S);
    for (int s = 0; s < S; ++s)
    {
        uint32_t n = b->hP[s + 1] - b->hP[s];
        tiles[s]   = (n + kBatchLeafTile - 1) / kBatchLeafTile;
        ngroups[s] = (n + gs - 1) / gs;
        float th   = theta ? theta[s] : prm->theta;
        if (!(th > 0)) return setError(ctx, RYO_ERR_ARG, "batch: theta of system %d", s);
        kb[s] = makeKeyBox(&b->limits[6 * s]);
        mb[s] = makeMacBox(&b->limits[6 * s], 1.0f / th);
        gv[s] = Gs ? Gs[s] : prm->G;
    }
    b->hTS = prefix(tiles);
    b->hGO = prefix(ngroups);
    const int numTiles = int(b->hTS[S]);
    if ((rc = b->blockSums.ensure(ctx, (size_t(numTiles) + 1) * 4))) return rc;
    RYO_HIP(ctx, hipMemcpyAsync(b->T(0), b->hP.data(), (S + 1) * 4, hipMemcpyHostToDevice, st));
    RYO_HIP(ctx, hipMemcpyAsync(b->T(1), b->hTS.data(), (S + 1) * 4, hipMemcpyHostToDevice, st));
    RYO_HIP(ctx, hipMemcpyAsync(b->T(6), b->hGO.data(), (S + 1) * 4, hipMemcpyHostToDevice, st));
    RYO_HIP(ctx, hipMemcpyAsync(b->keyBoxes.p, kb.data(), S * sizeof(KeyBox), hipMemcpyHostToDevice, st));
    RYO_HIP(ctx, hipMemcpyAsync(b->macBoxes.p, mb.data(), S * sizeof(MacBox), hipMemcpyHostToDevice, st));
    RYO_HIP(ctx, hipMemcpyAsync(b->G.p, gv.data(), S * 4, hipMemcpyHostToDevice, st));
    BatchView bv = b->view();

    // keys, segmented sort, gather
    auto* keys = b->keys.as<uint64_t>();
    auto* vals = b->vals.as<int32_t>();
    const float* in[5] = {b->in[0].as<float>(), b->in[1].as<float>(), b->in[2].as<float>(), b->in[3].as<float>(),
                          b->in[4].as<float>()};
    if ((rc = batchKeys(ctx, in[0], in[1], in[2], N, bv.P, S, b->keyBoxes.as<KeyBox>(), keys, vals))) return rc;
    if ((rc = segmentedSortPairs(ctx, keys, vals, N, bv.P, S, b->keyBuf.as<uint64_t>(), b->valBuf.as<int32_t>(),
                                 b->segKeys.as<uint64_t>(), b->segBuf.as<uint64_t>(), b->pos.as<int32_t>(),
                                 b->posBuf.as<int32_t>(), b->sortTmp.p, b->sortTmp.size)))
        return rc;
    float* out[5] = {b->srt[0].as<float>(), b->srt[1].as<float>(), b->srt[2].as<float>(), b->srt[3].as<float>(),
                     b->srt[4].as<float>()};
    if ((rc = gather(ctx, vals, N, 5, in, out))) return rc;
    RYO_HIP(ctx, hipMemcpyAsync(b->srt[5].p, vals, N * 4, hipMemcpyDeviceToDevice, st));

    // leaves; the slot offsets come back to size the node arrays
    if ((rc = batchLeavesCount(ctx, keys, bv, numTiles, B, b->blockSums.as<uint32_t>(), b->T(2)))) return rc;
    b->hLO.resize(S + 1);
    RYO_HIP(ctx, hipMemcpyAsync(b->hLO.data(), b->T(2), (S + 1) * 4, hipMemcpyDeviceToHost, st));
    RYO_HIP(ctx, hipStreamSynchronize(st));
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
    b->totalGroups   = b->hGO[S];
    RYO_HIP(ctx, hipMemcpyAsync(b->T(3), b->hNO.data(), (S + 1) * 4, hipMemcpyHostToDevice, st));
    RYO_HIP(ctx, hipMemcpyAsync(b->T(4), b->hNIO.data(), (S + 1) * 4, hipMemcpyHostToDevice, st));
    RYO_HIP(ctx, hipMemcpyAsync(b->T(5), b->hPO.data(), (S + 1) * 4, hipMemcpyHostToDevice, st));

    const size_t slots = b->totalSlots + 1, nn = b->totalNodes + 1, np = b->totalParents + 1, G = b->totalGroups + 1;
    rc = RYO_OK;
    if (!rc) rc = b->leaves.ensure(ctx, slots * 8);
    if (!rc) rc = b->layout.ensure(ctx, slots * 4);
    if (!rc) rc = b->counts.ensure(ctx, slots * 4);
    for (DevBuf* d : {&b->prefixes, &b->nodeKeyBuf, &b->nodeSeg, &b->nodeSegBuf})
        if (!rc) rc = d->ensure(ctx, nn * 8);
    for (DevBuf* d : {&b->child, &b->i2l, &b->l2i, &b->nodeIdx, &b->nodeValBuf, &b->nodePos, &b->nodePosBuf})
        if (!rc) rc = d->ensure(ctx, nn * 4);
    if (!rc) rc = b->parents.ensure(ctx, np * 4);
    if (!rc) rc = b->centers.ensure(ctx, nn * 16);
    if (!rc) rc = b->mp.ensure(ctx, nn * 32);
    if (!rc) rc = b->c64.ensure(ctx, nn * 32);
    if (!rc) rc = b->q64.ensure(ctx, nn * 64);
    if (!rc) rc = b->nodeTmp.ensure(ctx, sortTempBytes(nn));
    for (DevBuf* d : {&b->gStart, &b->gEnd, &b->gSys})
        if (!rc) rc = d->ensure(ctx, G * 4);
    if (rc) return rc;

    if ((rc = batchLeavesEmit(ctx, keys, bv, numTiles, B, b->blockSums.as<uint32_t>(), b->leaves.as<uint64_t>(),
                              b->layout.as<uint32_t>(), b->counts.as<uint32_t>(), b->totalSlots)))
        return rc;
    // linked octrees
    auto* pre = b->prefixes.as<uint64_t>();
    auto* idx = b->nodeIdx.as<int32_t>();
    if ((rc = batchUnsortedLayout(ctx, b->leaves.as<uint64_t>(), bv, b->totalSlots, pre, idx))) return rc;
    if ((rc = segmentedSortPairs(ctx, pre, idx, b->totalNodes, bv.NO, S, b->nodeKeyBuf.as<uint64_t>(),
                                 b->nodeValBuf.as<int32_t>(), b->nodeSeg.as<uint64_t>(),
                                 b->nodeSegBuf.as<uint64_t>(), b->nodePos.as<int32_t>(),
                                 b->nodePosBuf.as<int32_t>(), b->nodeTmp.p, b->nodeTmp.size)))
        return rc;
    if ((rc = batchLinkFinish(ctx, pre, idx, bv, b->totalNodes, b->totalInternal, b->totalParents,
                              b->i2l.as<int32_t>(), b->l2i.as<int32_t>(), b->LR.as<int32_t>(), b->LP.as<uint32_t>(),
                              b->child.as<int32_t>(), b->parents.as<int32_t>())))
        return rc;
    // upsweep, groups, traversal
    if ((rc = batchUpsweepEntry(ctx, out[0], out[1], out[2], out[3], b->layout.as<uint32_t>(), pre,
                                b->child.as<int32_t>(), b->parents.as<int32_t>(), b->l2i.as<int32_t>(),
                                b->centers.as<float>(), b->mp.as<float>(), b->c64.as<double>(), b->q64.as<double>(),
                                bv, b->macBoxes.as<MacBox>(), b->totalSlots, b->totalNodes)))
        return rc;
    if ((rc = batchGroups(ctx, bv, b->totalGroups, gs, b->gStart.as<uint32_t>(), b->gEnd.as<uint32_t>(),
                          b->gSys.as<int32_t>())))
        return rc;
    for (DevBuf* d : {&b->p, &b->ax, &b->ay, &b->az})
        RYO_HIP(ctx, hipMemsetAsync(d->p, 0, N * 4, st));
    if ((rc = batchTraverse(ctx, int(b->totalGroups), b->gStart.as<uint32_t>(), b->gEnd.as<uint32_t>(),
                            b->gSys.as<int32_t>(), bv, b->G.as<float>(), out[0], out[1], out[2], out[3], out[4], N,
                            b->child.as<int32_t>(), b->i2l.as<int32_t>(), b->layout.as<uint32_t>(),
                            b->centers.as<float>(), b->mp.as<float>(), b->p.as<float>(), b->ax.as<float>(),
                            b->ay.as<float>(), b->az.as<float>(), b->sysPot.as<double>(), prm->init_node)))
        return rc;
    b->hPot.resize(S);
    RYO_HIP(ctx, hipMemcpyAsync(b->hPot.data(), b->sysPot.p, S * 8, hipMemcpyDeviceToHost, st));
    RYO_HIP(ctx, hipStreamSynchronize(st));
    if (pot) std::memcpy(pot, b->hPot.data(), S * sizeof(double));
    b->solved = true;
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
