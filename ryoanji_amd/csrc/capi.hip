// C ABI (include/ryoanji_amd.h): contexts, the per-stage entry points and the resident-system pipeline.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>

namespace ryo
{

// stage implementations (keys_sort.hip, tree.hip, upsweep.hip, traverse.hip, ewald_direct.hip)
int sfcKeys(Context*, const float*, const float*, const float*, uint64_t*, size_t, const float*, int, int);
int sfcKeys64(Context*, const double*, const double*, const double*, uint64_t*, size_t, const double*, int, int);
template<class T>
int sfcKeys32(Context*, const T*, const T*, const T*, uint32_t*, size_t, const T*, int, int);
size_t sortTempBytes(size_t n);
int sortPairs(Context*, uint64_t*, int32_t*, size_t, uint64_t*, int32_t*, void*, size_t);
int sortPairsIndex(Context*, uint64_t*, int32_t*, size_t, uint64_t*, int32_t*, void*, size_t);
int sortCheck(Context*, const void*, size_t);
int sequence(Context*, int32_t*, size_t, int32_t);
int gather(Context*, const int32_t*, size_t, int, const float* const*, float* const*);
int gather64(Context*, const int32_t*, size_t, int, const double* const*, double* const*);
int sfcKeysPacked(Context*, const float* const*, uint64_t*, int32_t*, size_t, const float*, float4**);
int gatherRecords(Context*, const int32_t*, size_t, const float4*, float* const*);
int buildLeaves(Context*, const uint64_t*, size_t, unsigned, uint64_t*, uint32_t*, uint32_t*, size_t, int*);
int linkOctree(Context*, const uint64_t*, int, uint64_t*, int32_t*, int32_t*, int32_t*, int32_t*, int32_t*, int32_t*);
int upsweep(Context*, const float*, const float*, const float*, const float*, const uint32_t*, int, const uint64_t*,
            const int32_t*, const int32_t*, const int32_t*, const int32_t*, const float*, float, float*, float*,
            double*, double*, double, const uint32_t* dStat = nullptr);
int upsweepMDQ(Context*, const float*, const float*, const float*, const float*, const uint32_t*, int,
               const uint64_t*, const int32_t*, const int32_t*, const int32_t*, const int32_t*, const float*, float,
               float*, float*, double*, double*, const uint32_t* dStat = nullptr);
int buildLeavesDevice(Context* ctx, const uint64_t* keys, size_t n, unsigned bucket, uint64_t* leaves,
                      uint32_t* counts, uint32_t* layout, size_t capacity, uint32_t* stat);
int linkOctreeDevice(Context* ctx, const uint64_t* leaves, size_t capacity, const uint32_t* stat, uint64_t* prefixes,
                     int32_t* child, int32_t* parents, int32_t* levelRange, int32_t* i2l, int32_t* l2i);
int upsweepMDQ64(Context*, const double*, const double*, const double*, const double*, const uint32_t*, int,
                 const uint64_t*, const int32_t*, const int32_t*, const int32_t*, const int32_t*, const double*,
                 float, double*, double*, float*);
int upsweep64(Context*, const double*, const double*, const double*, const double*, const uint32_t*, int,
              const uint64_t*, const int32_t*, const int32_t*, const int32_t*, const int32_t*, const double*, float,
              double*, double*, float*);
int fixedGroups(Context*, uint32_t, uint32_t, unsigned, uint32_t*, int*);
int groupSplits(Context* ctx, uint32_t first, uint32_t last, const float* x, const float* y, const float* z,
                const float* h, const uint64_t* leaves, int numLeaves, const uint32_t* layout, const float* limits,
                unsigned groupSize, float tolFactor, float extentRatio, uint32_t* groups, size_t capacity,
                int* numGroups);
int ewald(Context*, const double*, const double*, uint32_t, uint32_t, const float*, const float*, const float*,
          const float*, const float*, float, const double*, float*, float*, float*, float*, double*, bool);
int softeningRange(Context* ctx, const float* h, size_t n, float* hmin, float* hmax);
int directSum(Context*, const float*, const float*, const float*, const float*, const float*, size_t,
              const int64_t*, size_t, double, const double*, int, double*);
int ewaldPairSum(Context* ctx, const float* x, const float* y, const float* z, const float* h, const float* m,
                 size_t n, const int64_t* targets, size_t nt, double G, double L, double alphaScale, double* out);
int kickDrift(Context*, float*, float*, float*, float*, float*, float*, const float*, const float*, const float*,
              size_t, float, float, int, const float*);
int fitBox(Context*, const float*, const float*, const float*, size_t, float*);

struct TravArgs;
int traverseEntry(Context* ctx, int numGroups, const uint32_t* gs, const uint32_t* ge, int initNode, const float* xt,
                  const float* yt, const float* zt, const float* mt, const float* ht, const float* xs,
                  const float* ys, const float* zs, const float* ms, const float* hs, const int32_t* child,
                  const int32_t* i2l, const uint32_t* layout, const float* centers, const float* mp, float G,
                  int numShells, const float* boxL, float* p, float* ax, float* ay, float* az, double* totalPot,
                  int stackCap, size_t hCount, bool mdq = false, bool noAlloc = false,
                  const uint32_t* abort = nullptr);
int initNoAllocScratch(Context* ctx);
// reference-signature stages (stages.hip, ewald_direct.hip)
int nodeCounts(Context*, const uint64_t*, int, const uint64_t*, size_t, unsigned, uint32_t*);
int nodeOps(Context*, const uint64_t*, int, const uint32_t*, unsigned, int32_t*, int*);
int rebalanceTree(Context*, const uint64_t*, int, int, const int32_t*, uint64_t*, int*);
int leafCenters(Context*, int, const void*, const void*, const void*, const void*, const int32_t*, int,
                const uint32_t*, void*);
int upsweepCenters(Context*, int, int, const int32_t*, const int32_t*, void*);
int setMac(Context*, int, const uint64_t*, int, void*, float, const void*);
int leafMultipoles(Context*, int, int, const void*, const void*, const void*, const void*, const int32_t*, int,
                   const uint32_t*, const void*, void*);
int upsweepMultipoles(Context*, int, int, int, int, const int32_t*, const void*, void*);
int directSumRange(Context*, int, size_t, size_t, size_t, const void*, int, const void*, const void*, const void*,
                   const void*, const void*, void*, void*, void*, void*);

int setError(Context* ctx, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->lastError = buf;
    return code;
}

void* Context::getScratch(int slot, size_t bytes)
{
    if (scratch.size() < size_t(kNumSlots)) scratch.resize(kNumSlots, nullptr), scratchBytes.resize(kNumSlots, 0);
    if (bytes == 0) bytes = 1;
    if (scratchBytes[slot] >= bytes) return scratch[slot];
    if (scratch[slot])
    {
        // the old buffer may still be read by queued work
        hipStreamSynchronize(stream);
        hipFree(scratch[slot]);
        scratch[slot] = nullptr, scratchBytes[slot] = 0;
    }
    size_t sz = std::max(bytes, size_t(4096));
    sz        = sz + sz / 8;
    if (hipMalloc(&scratch[slot], sz) != hipSuccess)
    {
        setError(this, RYO_ERR_HIP, "out of device memory allocating %zu bytes of scratch", sz);
        scratch[slot] = nullptr;
        return nullptr;
    }
    scratchBytes[slot] = sz;
    return scratch[slot];
}

} // namespace ryo

using namespace ryo;

#define CTX_OR_FAIL(h)                                                                                                 \
    if (!(h)) return RYO_ERR_ARG;                                                                                      \
    Context* ctx = &(h)->c

extern "C" {

int ryo_ctx_create(ryo_ctx** out, void* stream)
{
    if (!out) return RYO_ERR_ARG;
    auto* h      = new ryo_ctx();
    Context* c   = &h->c;
    c->stream    = static_cast<hipStream_t>(stream);
    if (hipGetDevice(&c->device) != hipSuccess)
    {
        delete h;
        return RYO_ERR_HIP;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess && cus > 0)
        c->numCUs = cus;
    if (hipMalloc(&c->dStats, 256) != hipSuccess)
    {
        delete h;
        return RYO_ERR_HIP;
    }
    hipMemset(c->dStats, 0, 256);
    // this device's work array of the allocation-free traversal (once per device)
    if (initNoAllocScratch(c) != RYO_OK)
    {
        hipFree(c->dStats);
        delete h;
        return RYO_ERR_HIP;
    }
    hipEventCreate(&c->ev[0]);
    hipEventCreate(&c->ev[1]);
    *out = h;
    return RYO_OK;
}

int ryo_ctx_destroy(ryo_ctx* h)
{
    if (!h) return RYO_OK;
    Context* c = &h->c;
    hipStreamSynchronize(c->stream);
    for (void* p : c->scratch)
        if (p) hipFree(p);
    if (c->dStats) hipFree(c->dStats);
    if (c->ev[0]) hipEventDestroy(c->ev[0]);
    if (c->ev[1]) hipEventDestroy(c->ev[1]);
    delete h;
    return RYO_OK;
}

const char* ryo_last_error(const ryo_ctx* h) { return h ? h->c.lastError.c_str() : "null context"; }

int ryo_ctx_sync(ryo_ctx* h)
{
    CTX_OR_FAIL(h);
    RYO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RYO_OK;
}

int ryo_ctx_set_timing(ryo_ctx* h, int enable)
{
    CTX_OR_FAIL(h);
    ctx->timing = enable != 0;
    return RYO_OK;
}

int ryo_ctx_set_traversal_shape(ryo_ctx* h, int stack_entries_per_wave, int blocks_per_cu, int group_ranges)
{
    CTX_OR_FAIL(h);
    if (stack_entries_per_wave < 0 || (stack_entries_per_wave > 0 && stack_entries_per_wave < 9) ||
        stack_entries_per_wave > 22 * 64 || blocks_per_cu < 0 || blocks_per_cu > 64 || group_ranges < 0 ||
        group_ranges > 8)
        return setError(ctx, RYO_ERR_ARG,
                        "traversal shape: stack entries in [9, 1408], blocks per CU in [1, 64], ranges in [1, 8]");
    ctx->travStack       = stack_entries_per_wave;
    ctx->travBlocksPerCU = blocks_per_cu;
    ctx->travRanges      = group_ranges;
    return RYO_OK;
}

int ryo_ctx_set_traversal_waves(ryo_ctx* h, int waves_per_block)
{
    CTX_OR_FAIL(h);
    if (waves_per_block != 0 && waves_per_block != 2 && waves_per_block != 4 && waves_per_block != 8)
        return setError(ctx, RYO_ERR_ARG, "traversal waves per block: 2, 4 or 8 (0: 4)");
    ctx->travWaves = waves_per_block;
    return RYO_OK;
}

double ryo_last_traverse_ms(const ryo_ctx* h) { return h ? h->c.lastTraverseMs : 0.0; }

int ryo_sfc_keys_f32(ryo_ctx* h, const float* x, const float* y, const float* z, uint64_t* keys, size_t n,
                     const float* limits, int kind, int skip_removed)
{
    CTX_OR_FAIL(h);
    if (!limits || (n && (!x || !y || !z || !keys))) return setError(ctx, RYO_ERR_ARG, "sfc keys: null pointer");
    return sfcKeys(ctx, x, y, z, keys, n, limits, kind, skip_removed);
}

int ryo_sfc_keys_u32_f32(ryo_ctx* h, const float* x, const float* y, const float* z, uint32_t* keys, size_t n,
                         const float* limits, int kind, int skip_removed)
{
    CTX_OR_FAIL(h);
    if (!limits || (n && (!x || !y || !z || !keys))) return setError(ctx, RYO_ERR_ARG, "sfc keys: null pointer");
    return sfcKeys32(ctx, x, y, z, keys, n, limits, kind, skip_removed);
}

int ryo_sfc_keys_u32_f64(ryo_ctx* h, const double* x, const double* y, const double* z, uint32_t* keys, size_t n,
                         const double* limits, int kind, int skip_removed)
{
    CTX_OR_FAIL(h);
    if (!limits || (n && (!x || !y || !z || !keys))) return setError(ctx, RYO_ERR_ARG, "sfc keys: null pointer");
    return sfcKeys32(ctx, x, y, z, keys, n, limits, kind, skip_removed);
}

int ryo_sfc_keys_f64(ryo_ctx* h, const double* x, const double* y, const double* z, uint64_t* keys, size_t n,
                     const double* limits, int kind, int skip_removed)
{
    CTX_OR_FAIL(h);
    if (!limits || (n && (!x || !y || !z || !keys))) return setError(ctx, RYO_ERR_ARG, "sfc keys: null pointer");
    return sfcKeys64(ctx, x, y, z, keys, n, limits, kind, skip_removed);
}

size_t ryo_sort_temp_bytes(size_t n) { return sortTempBytes(n); }

int ryo_sort_pairs_u64_i32(ryo_ctx* h, uint64_t* keys, int32_t* values, size_t n, uint64_t* keyBuf,
                           int32_t* valueBuf, void* tmp, size_t tmpBytes)
{
    CTX_OR_FAIL(h);
    if (n && (!keys || !values || !keyBuf || !valueBuf || !tmp)) return setError(ctx, RYO_ERR_ARG, "sort: null pointer");
    if (n >= (size_t(1) << 31)) return setError(ctx, RYO_ERR_ARG, "sort: n must be < 2^31");
    return sortPairs(ctx, keys, values, n, keyBuf, valueBuf, tmp, tmpBytes);
}

int ryo_sort_check(ryo_ctx* h, const void* tmp, size_t n)
{
    CTX_OR_FAIL(h);
    if (n && !tmp) return setError(ctx, RYO_ERR_ARG, "sort check: null pointer");
    RYO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return sortCheck(ctx, tmp, n);
}

int ryo_sequence_i32(ryo_ctx* h, int32_t* out, size_t n, int32_t start)
{
    CTX_OR_FAIL(h);
    return sequence(ctx, out, n, start);
}

int ryo_gather_f32(ryo_ctx* h, const int32_t* perm, size_t n, int num_arrays, const float* const* in,
                   float* const* out)
{
    CTX_OR_FAIL(h);
    return gather(ctx, perm, n, num_arrays, in, out);
}

int ryo_gather_f64(ryo_ctx* h, const int32_t* perm, size_t n, int num_arrays, const double* const* in,
                   double* const* out)
{
    CTX_OR_FAIL(h);
    return gather64(ctx, perm, n, num_arrays, in, out);
}

int ryo_build_leaves_u64(ryo_ctx* h, const uint64_t* sorted_keys, size_t n, unsigned bucket_size, uint64_t* leaves,
                         uint32_t* counts, uint32_t* layout, size_t capacity, int* num_leaves)
{
    CTX_OR_FAIL(h);
    if (!num_leaves || !leaves || !counts || !layout) return setError(ctx, RYO_ERR_ARG, "leaves: null pointer");
    return buildLeaves(ctx, sorted_keys, n, bucket_size, leaves, counts, layout, capacity, num_leaves);
}

int ryo_link_octree_u64(ryo_ctx* h, const uint64_t* leaves, int num_leaves, uint64_t* prefixes, int32_t* child,
                        int32_t* parents, int32_t* level_range, int32_t* i2l, int32_t* l2i, int32_t* lr_host)
{
    CTX_OR_FAIL(h);
    return linkOctree(ctx, leaves, num_leaves, prefixes, child, parents, level_range, i2l, l2i, lr_host);
}

int ryo_upsweep_f32(ryo_ctx* h, const float* x, const float* y, const float* z, const float* m,
                    const uint32_t* layout, int num_leaves, const uint64_t* prefixes, const int32_t* child,
                    const int32_t* parents, const int32_t* l2i, const int32_t* level_range, const float* limits,
                    float inv_theta, float* centers, float* mp, double* c64, double* q64)
{
    CTX_OR_FAIL(h);
    return upsweep(ctx, x, y, z, m, layout, num_leaves, prefixes, child, parents, l2i, level_range, limits,
                   inv_theta, centers, mp, c64, q64, 0.0);
}

int ryo_upsweep_mdq_f32(ryo_ctx* h, const float* x, const float* y, const float* z, const float* m,
                        const uint32_t* layout, int num_leaves, const uint64_t* prefixes, const int32_t* child,
                        const int32_t* parents, const int32_t* l2i, const int32_t* level_range, const float* limits,
                        float inv_theta, float* centers, float* mp, double* c64, double* q64)
{
    CTX_OR_FAIL(h);
    if (!limits || !centers || !mp) return setError(ctx, RYO_ERR_ARG, "upsweep: null pointer");
    return upsweepMDQ(ctx, x, y, z, m, layout, num_leaves, prefixes, child, parents, l2i, level_range, limits,
                      inv_theta, centers, mp, c64, q64);
}

int ryo_upsweep_f64(ryo_ctx* h, const double* x, const double* y, const double* z, const double* m,
                    const uint32_t* layout, int num_leaves, const uint64_t* prefixes, const int32_t* child,
                    const int32_t* parents, const int32_t* l2i, const int32_t* level_range, const double* limits,
                    float inv_theta, double* centers, double* mp, float* mp32)
{
    CTX_OR_FAIL(h);
    if (!limits) return setError(ctx, RYO_ERR_ARG, "upsweep: null limits");
    return upsweep64(ctx, x, y, z, m, layout, num_leaves, prefixes, child, parents, l2i, level_range, limits,
                     inv_theta, centers, mp, mp32);
}

int ryo_upsweep_mdq_f64(ryo_ctx* h, const double* x, const double* y, const double* z, const double* m,
                        const uint32_t* layout, int num_leaves, const uint64_t* prefixes, const int32_t* child,
                        const int32_t* parents, const int32_t* l2i, const int32_t* level_range, const double* limits,
                        float inv_theta, double* centers, double* mp, float* mp32)
{
    CTX_OR_FAIL(h);
    if (!limits) return setError(ctx, RYO_ERR_ARG, "upsweep: null limits");
    return upsweepMDQ64(ctx, x, y, z, m, layout, num_leaves, prefixes, child, parents, l2i, level_range, limits,
                        inv_theta, centers, mp, mp32);
}

int ryo_fixed_groups(ryo_ctx* h, uint32_t first, uint32_t last, unsigned group_size, uint32_t* groups,
                     int* num_groups)
{
    CTX_OR_FAIL(h);
    if (!groups || !num_groups) return setError(ctx, RYO_ERR_ARG, "groups: null pointer");
    return fixedGroups(ctx, first, last, group_size, groups, num_groups);
}

int ryo_group_splits_f32(ryo_ctx* hc, uint32_t first, uint32_t last, const float* x, const float* y, const float* z,
                         const float* h, const uint64_t* leaves, int num_leaves, const uint32_t* layout,
                         const float* limits, unsigned group_size, float tol_factor, float extent_ratio,
                         uint32_t* groups, size_t capacity, int* num_groups)
{
    CTX_OR_FAIL(hc);
    return groupSplits(ctx, first, last, x, y, z, h, leaves, num_leaves, layout, limits, group_size, tol_factor,
                       extent_ratio, groups, capacity, num_groups);
}

int ryo_traverse_f32(ryo_ctx* h, int num_groups, const uint32_t* gs, const uint32_t* ge, int init_node,
                     const float* xt, const float* yt, const float* zt, const float* mt, const float* ht,
                     const float* xs, const float* ys, const float* zs, const float* ms, const float* hs,
                     const int32_t* child, const int32_t* i2l, const uint32_t* layout, const float* centers,
                     const float* mp, float G, int num_shells, const float* box_l, float* p, float* ax, float* ay,
                     float* az, double* total_pot)
{
    CTX_OR_FAIL(h);
    if (init_node != 0 && init_node != 1) return setError(ctx, RYO_ERR_ARG, "traverse: init_node must be 0 or 1");
    if (num_shells < 0 || (num_shells > 0 && !box_l)) return setError(ctx, RYO_ERR_ARG, "traverse: bad shells");
    return traverseEntry(ctx, num_groups, gs, ge, init_node, xt, yt, zt, mt, ht, xs, ys, zs, ms, hs, child, i2l,
                         layout, centers, mp, G, num_shells, box_l, p, ax, ay, az, total_pot, 0, 0);
}

int ryo_traverse_mdq_f32(ryo_ctx* h, int num_groups, const uint32_t* gs, const uint32_t* ge, int init_node,
                         const float* xt, const float* yt, const float* zt, const float* mt, const float* ht,
                         const float* xs, const float* ys, const float* zs, const float* ms, const float* hs,
                         const int32_t* child, const int32_t* i2l, const uint32_t* layout, const float* centers,
                         const float* mp, float G, int num_shells, const float* box_l, float* p, float* ax, float* ay,
                         float* az, double* total_pot)
{
    CTX_OR_FAIL(h);
    if (init_node != 0 && init_node != 1) return setError(ctx, RYO_ERR_ARG, "traverse: init_node must be 0 or 1");
    if (num_shells < 0 || (num_shells > 0 && !box_l)) return setError(ctx, RYO_ERR_ARG, "traverse: bad shells");
    return traverseEntry(ctx, num_groups, gs, ge, init_node, xt, yt, zt, mt, ht, xs, ys, zs, ms, hs, child, i2l,
                         layout, centers, mp, G, num_shells, box_l, p, ax, ay, az, total_pot, 0, 0, true);
}

int ryo_traverse_f32_noalloc(ryo_ctx* h, int num_groups, const uint32_t* gs, const uint32_t* ge, int init_node,
                             const float* xt, const float* yt, const float* zt, const float* mt, const float* ht,
                             const float* xs, const float* ys, const float* zs, const float* ms, const float* hs,
                             const int32_t* child, const int32_t* i2l, const uint32_t* layout, const float* centers,
                             const float* mp, float G, int num_shells, const float* box_l, float* p, float* ax,
                             float* ay, float* az, double* total_pot)
{
    CTX_OR_FAIL(h);
    if (init_node != 0 && init_node != 1) return setError(ctx, RYO_ERR_ARG, "traverse: init_node must be 0 or 1");
    if (num_shells < 0 || (num_shells > 0 && !box_l)) return setError(ctx, RYO_ERR_ARG, "traverse: bad shells");
    return traverseEntry(ctx, num_groups, gs, ge, init_node, xt, yt, zt, mt, ht, xs, ys, zs, ms, hs, child, i2l,
                         layout, centers, mp, G, num_shells, box_l, p, ax, ay, az, total_pot, 0, 0, false, true);
}

/* test hook: the traversal with a reduced per-wave stack capacity (forces the split/retry path) */
int ryo_traverse_f32_stackcap(ryo_ctx* h, int stack_cap, int num_groups, const uint32_t* gs, const uint32_t* ge,
                              int init_node, const float* xt, const float* yt, const float* zt, const float* mt,
                              const float* ht, const float* xs, const float* ys, const float* zs, const float* ms,
                              const float* hs, const int32_t* child, const int32_t* i2l, const uint32_t* layout,
                              const float* centers, const float* mp, float G, int num_shells, const float* box_l,
                              float* p, float* ax, float* ay, float* az, double* total_pot)
{
    CTX_OR_FAIL(h);
    return traverseEntry(ctx, num_groups, gs, ge, init_node, xt, yt, zt, mt, ht, xs, ys, zs, ms, hs, child, i2l,
                         layout, centers, mp, G, num_shells, box_l, p, ax, ay, az, total_pot, stack_cap, 0);
}

/* ---------------------------------------------------------------- reference-signature stages */

int ryo_node_counts_u64(ryo_ctx* h, const uint64_t* tree, int num_nodes, const uint64_t* keys, size_t n,
                        unsigned max_count, uint32_t* counts)
{
    CTX_OR_FAIL(h);
    return nodeCounts(ctx, tree, num_nodes, keys, n, max_count, counts);
}

int ryo_node_ops_u64(ryo_ctx* h, const uint64_t* tree, int num_nodes, const uint32_t* counts, unsigned bucket_size,
                     int32_t* node_ops, int* new_num_nodes)
{
    CTX_OR_FAIL(h);
    return nodeOps(ctx, tree, num_nodes, counts, bucket_size, node_ops, new_num_nodes);
}

int ryo_rebalance_tree_u64(ryo_ctx* h, const uint64_t* tree, int num_nodes, int new_num_nodes,
                           const int32_t* node_ops, uint64_t* new_tree, int* converged)
{
    CTX_OR_FAIL(h);
    return rebalanceTree(ctx, tree, num_nodes, new_num_nodes, node_ops, new_tree, converged);
}

int ryo_leaf_centers(ryo_ctx* h, int types, const void* x, const void* y, const void* z, const void* m,
                     const int32_t* leaf_to_internal, int num_leaves, const uint32_t* layout, void* centers)
{
    CTX_OR_FAIL(h);
    return leafCenters(ctx, types, x, y, z, m, leaf_to_internal, num_leaves, layout, centers);
}

int ryo_upsweep_centers(ryo_ctx* h, int types, int num_levels, const int32_t* level_range_host,
                        const int32_t* child_offsets, void* centers)
{
    CTX_OR_FAIL(h);
    return upsweepCenters(ctx, types, num_levels, level_range_host, child_offsets, centers);
}

int ryo_set_mac(ryo_ctx* h, int types, const uint64_t* prefixes, int num_nodes, void* centers, float inv_theta,
                const void* limits)
{
    CTX_OR_FAIL(h);
    return setMac(ctx, types, prefixes, num_nodes, centers, inv_theta, limits);
}

int ryo_leaf_multipoles(ryo_ctx* h, int types, int mdq, const void* x, const void* y, const void* z, const void* m,
                        const int32_t* leaf_to_internal, int num_leaves, const uint32_t* layout, const void* centers,
                        void* multipoles)
{
    CTX_OR_FAIL(h);
    return leafMultipoles(ctx, types, mdq, x, y, z, m, leaf_to_internal, num_leaves, layout, centers, multipoles);
}

int ryo_upsweep_multipoles(ryo_ctx* h, int types, int mdq, int first_cell, int last_cell,
                           const int32_t* child_offsets, const void* centers, void* multipoles)
{
    CTX_OR_FAIL(h);
    return upsweepMultipoles(ctx, types, mdq, first_cell, last_cell, child_offsets, centers, multipoles);
}

int ryo_direct_sum_range(ryo_ctx* h, int types, size_t first, size_t last, size_t num_bodies, const void* box_l,
                         int num_shells, const void* x, const void* y, const void* z, const void* m, const void* hh,
                         void* p, void* ax, void* ay, void* az)
{
    CTX_OR_FAIL(h);
    return directSumRange(ctx, types, first, last, num_bodies, box_l, num_shells, x, y, z, m, hh, p, ax, ay, az);
}

int ryo_malloc(ryo_ctx* h, void** ptr, size_t bytes)
{
    CTX_OR_FAIL(h);
    if (!ptr) return setError(ctx, RYO_ERR_ARG, "malloc: null pointer");
    *ptr = nullptr;
    if (bytes == 0) return RYO_OK;
    RYO_HIP(ctx, hipMalloc(ptr, bytes));
    return RYO_OK;
}

int ryo_free(ryo_ctx* h, void* ptr)
{
    CTX_OR_FAIL(h);
    if (!ptr) return RYO_OK;
    RYO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    RYO_HIP(ctx, hipFree(ptr));
    return RYO_OK;
}

int ryo_memcpy(ryo_ctx* h, void* dst, const void* src, size_t bytes)
{
    CTX_OR_FAIL(h);
    if (bytes == 0) return RYO_OK;
    if (!dst || !src) return setError(ctx, RYO_ERR_ARG, "memcpy: null pointer");
    RYO_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream));
    RYO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RYO_OK;
}

/* test hook: the MDQpole traversal with a reduced per-wave stack capacity (forces the split/retry path) */
int ryo_traverse_mdq_f32_stackcap(ryo_ctx* h, int stack_cap, int num_groups, const uint32_t* gs, const uint32_t* ge,
                                  int init_node, const float* xt, const float* yt, const float* zt, const float* mt,
                                  const float* ht, const float* xs, const float* ys, const float* zs,
                                  const float* ms, const float* hs, const int32_t* child, const int32_t* i2l,
                                  const uint32_t* layout, const float* centers, const float* mp, float G,
                                  int num_shells, const float* box_l, float* p, float* ax, float* ay, float* az,
                                  double* total_pot)
{
    CTX_OR_FAIL(h);
    return traverseEntry(ctx, num_groups, gs, ge, init_node, xt, yt, zt, mt, ht, xs, ys, zs, ms, hs, child, i2l,
                         layout, centers, mp, G, num_shells, box_l, p, ax, ay, az, total_pot, stack_cap, 0, true);
}

int ryo_read_stats(ryo_ctx* h, uint64_t* stats)
{
    CTX_OR_FAIL(h);
    for (int k = 0; k < 5; ++k)
        stats[k] = ctx->hStats[k];
    return RYO_OK;
}

uint64_t ryo_traversal_retries(const ryo_ctx* h) { return h ? h->c.retriedGroups : 0; }

int ryo_ewald_f32(ryo_ctx* h, const double* Mroot, const double* center, uint32_t first, uint32_t last,
                  const float* x, const float* y, const float* z, const float* m, const float* limits, float G,
                  const double* settings, float* ugrav, float* ax, float* ay, float* az, double* total_pot)
{
    CTX_OR_FAIL(h);
    return ewald(ctx, Mroot, center, first, last, x, y, z, m, limits, G, settings, ugrav, ax, ay, az, total_pot,
                 false);
}

int ryo_direct_sum_f64(ryo_ctx* h, const float* x, const float* y, const float* z, const float* hh, const float* m,
                       size_t n, const int64_t* targets, size_t nt, double G, const double* box_l, int num_shells,
                       double* out)
{
    CTX_OR_FAIL(h);
    return directSum(ctx, x, y, z, hh, m, n, targets, nt, G, box_l, num_shells, out);
}

int ryo_ewald_pair_sum_f64(ryo_ctx* h, const float* x, const float* y, const float* z, const float* hh,
                           const float* m, size_t n, const int64_t* targets, size_t nt, double G, double box_length,
                           double alpha_scale, double* out)
{
    CTX_OR_FAIL(h);
    if (nt && (!x || !y || !z || !hh || !m || !out)) return setError(ctx, RYO_ERR_ARG, "Ewald pair sum: null pointer");
    return ewaldPairSum(ctx, x, y, z, hh, m, n, targets, nt, G, box_length, alpha_scale, out);
}

} // extern "C"

/* ============================================================================ resident system */

struct ryo_system
{
    Context* ctx;
    ryo_ctx* handle;
    size_t cap = 0, n = 0;
    float limits[6] = {0, 0, 0, 0, 0, 0};
    int periodic     = 0;
    // particle arrays, current (SFC) order, and a second set for permutation
    float* a[2][9] = {}; // x y z m h vx vy vz order(i32 bits)
    int cur = 0;
    uint64_t *keys = nullptr, *keyBuf = nullptr;
    int32_t *perm = nullptr, *permBuf = nullptr;
    void* sortTmp = nullptr;
    size_t sortTmpBytes = 0;
    float *p = nullptr, *ax = nullptr, *ay = nullptr, *az = nullptr;
    // tree
    size_t leafCap = 0;
    int L = 0, nNodes = 0, G = 0;
    uint64_t* leaves = nullptr;
    uint32_t *counts = nullptr, *layout = nullptr;
    size_t nodeCap = 0;
    uint64_t* prefixes = nullptr;
    int32_t *child = nullptr, *parents = nullptr, *levelRange = nullptr, *i2l = nullptr, *l2i = nullptr;
    float *centers = nullptr, *mp = nullptr;
    double *c64 = nullptr, *q64 = nullptr;
    size_t groupCap = 0;
    uint32_t* groups = nullptr;
    int levelRangeHost[24] = {};
    bool haveAcc = false;     // accelerations of the current positions are in ax, ay, az
    bool freshOrder = false;  // the particle arrays are in the order they were loaded in
    float lastG = 1.0f;       // G, replica shells and Ewald use of the last gravity call
    int lastShells = 0, lastEwald = 0;
    bool integrating = false; // velocities are in use (leapfrog): the build permutes them too
    bool mdq         = false; // the last build formed CartesianMDQpole multipoles (12 per node)
    // a device-sized build (BuildStat counts on the device, copied to pinned host memory with the next
    // synchronization): its leaf count, overflow and layout check are read after the gravity pass
    uint32_t* dStat    = nullptr;
    uint32_t* hStat    = nullptr;
    bool statPending   = false;
    bool timingPending = false; // the build's phase times are read at the next synchronization
    ryo_params buildPrm{};      // the parameters of the last build (to build again when its leaves outgrew)
    bool regrowing     = false; // inside the second pass after an outgrown build
    // leaves a device-sized build launches for: the last leaf count with a margin (the particles of a loaded system
    // change little between builds), 0 for n / 8; a build with more leaves overflows and runs again
    size_t leafHint    = 0;
    hipEvent_t ev[9] = {};
    double phaseMs[8] = {};
};

namespace
{
template<class T>
int devAlloc(ryo_system* s, T** p, size_t count)
{
    if (*p) hipFree(*p);
    *p = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess)
        return setError(s->ctx, RYO_ERR_HIP, "system: out of device memory (%zu bytes)", count * sizeof(T));
    return RYO_OK;
}

int growTree(ryo_system* s, int L)
{
    int rc = RYO_OK;
    if (size_t(L) > s->leafCap)
    {
        size_t c = size_t(L) + size_t(L) / 4 + 64;
        if ((rc = devAlloc(s, &s->leaves, c + 1)) || (rc = devAlloc(s, &s->counts, c)) ||
            (rc = devAlloc(s, &s->layout, c + 1)))
            return rc;
        s->leafCap = c;
    }
    // the node arrays hold the nodes of a full leaf capacity (what a device-sized build may produce)
    const size_t lc = std::max(s->leafCap, size_t(L));
    size_t nodes    = lc + (lc - 1) / 7;
    if (nodes > s->nodeCap)
    {
        size_t c = nodes + 64;
        if ((rc = devAlloc(s, &s->prefixes, c)) || (rc = devAlloc(s, &s->child, c + 1)) ||
            (rc = devAlloc(s, &s->parents, c / 8 + 2)) || (rc = devAlloc(s, &s->i2l, c)) ||
            (rc = devAlloc(s, &s->l2i, c)) || (rc = devAlloc(s, &s->centers, 4 * c)) ||
            (rc = devAlloc(s, &s->mp, 12 * c)) || (rc = devAlloc(s, &s->c64, 4 * c)) ||
            (rc = devAlloc(s, &s->q64, 12 * c)))
            return rc;
        s->nodeCap = c;
    }
    return RYO_OK;
}
} // namespace

extern "C" {

void ryo_default_params(ryo_params* p)
{
    p->bucket_size = 64;
    p->theta       = 0.5f;
    p->G           = 1.0f;
    p->group_size  = 64;
    p->num_shells  = 0;
    p->ewald       = 0;
    const double def[5] = {1, 2.6, 2.8, 2.0, 3.0e-3};
    for (int k = 0; k < 5; ++k)
        p->ewald_settings[k] = def[k];
    p->init_node       = 0;
    p->soft_mac        = 0;
    p->group_split_tol = 0.0f;
    p->multipole       = 0;
}

int ryo_system_create(ryo_ctx* h, size_t capacity, ryo_system** out)
{
    CTX_OR_FAIL(h);
    if (!out || capacity == 0) return setError(ctx, RYO_ERR_ARG, "system: capacity must be positive");
    auto* s   = new ryo_system();
    s->ctx    = ctx;
    s->handle = h;
    s->cap    = capacity;
    int rc    = RYO_OK;
    for (int b = 0; b < 2 && !rc; ++b)
        for (int k = 0; k < 9 && !rc; ++k)
            rc = devAlloc(s, &s->a[b][k], capacity);
    if (!rc) rc = devAlloc(s, &s->keys, capacity);
    if (!rc) rc = devAlloc(s, &s->keyBuf, capacity);
    if (!rc) rc = devAlloc(s, &s->perm, capacity);
    if (!rc) rc = devAlloc(s, &s->permBuf, capacity);
    s->sortTmpBytes = sortTempBytes(capacity);
    if (!rc) rc = devAlloc(s, reinterpret_cast<char**>(&s->sortTmp), s->sortTmpBytes);
    if (!rc) rc = devAlloc(s, &s->p, capacity);
    if (!rc) rc = devAlloc(s, &s->ax, capacity);
    if (!rc) rc = devAlloc(s, &s->ay, capacity);
    if (!rc) rc = devAlloc(s, &s->az, capacity);
    for (auto& e : s->ev)
        if (!rc && hipEventCreate(&e) != hipSuccess) rc = setError(ctx, RYO_ERR_HIP, "system: event creation");
    if (!rc) rc = devAlloc(s, &s->dStat, kStWords);
    // tree arrays for n / 8 leaves of the full capacity: a build rarely grows them later (a growth frees the old
    // arrays, and hipFree waits for the whole device, other streams' work included)
    if (!rc) rc = growTree(s, int(std::min<size_t>(std::max<size_t>(64, capacity / 8), 0x7fffffff)));
    if (!rc && hipHostMalloc(reinterpret_cast<void**>(&s->hStat), kStWords * sizeof(uint32_t)) != hipSuccess)
        rc = setError(ctx, RYO_ERR_HIP, "system: pinned host memory");
    if (rc)
    {
        ryo_system_destroy(s);
        return rc;
    }
    *out = s;
    return RYO_OK;
}

int ryo_system_destroy(ryo_system* s)
{
    if (!s) return RYO_OK;
    hipStreamSynchronize(s->ctx->stream);
    for (auto& b : s->a)
        for (auto* p : b)
            if (p) hipFree(p);
    void* ptrs[] = {s->keys,    s->keyBuf, s->perm,    s->permBuf, s->sortTmp, s->p,   s->ax,   s->ay,
                    s->az,      s->leaves, s->counts,  s->layout,  s->prefixes, s->child, s->parents, s->i2l,
                    s->l2i,     s->centers, s->mp,     s->c64,     s->q64,     s->groups, s->levelRange, s->dStat};
    for (void* p : ptrs)
        if (p) hipFree(p);
    if (s->hStat) hipHostFree(s->hStat);
    for (auto& e : s->ev)
        if (e) hipEventDestroy(e);
    delete s;
    return RYO_OK;
}

int ryo_system_arrays(ryo_system* s, float** out)
{
    if (!s || !out) return RYO_ERR_ARG;
    for (int k = 0; k < 8; ++k)
        out[k] = s->a[s->cur][k];
    return RYO_OK;
}

int ryo_system_load(ryo_system* s, size_t n, const float* x, const float* y, const float* z, const float* m,
                    const float* h, const float* limits, int periodic)
{
    if (!s) return RYO_ERR_ARG;
    Context* ctx = s->ctx;
    if (n > s->cap) return setError(ctx, RYO_ERR_CAPACITY, "system: %zu particles > capacity %zu", n, s->cap);
    s->statPending = false; // the last build's tree no longer describes the particles
    if (n != s->n) s->leafHint = 0; // a new system of the same size keeps the launch size (a miss builds again)
    s->n = n;
    std::memcpy(s->limits, limits, sizeof(s->limits));
    s->periodic = periodic;
    s->cur      = 0;
    s->haveAcc  = false;
    s->integrating = false;
    s->freshOrder  = true;
    const float* src[5] = {x, y, z, m, h};
    for (int k = 0; k < 5; ++k)
        RYO_HIP(ctx, hipMemcpyAsync(s->a[0][k], src[k], n * sizeof(float), hipMemcpyDefault, ctx->stream));
    for (int k = 5; k < 8; ++k)
        RYO_HIP(ctx, hipMemsetAsync(s->a[0][k], 0, n * sizeof(float), ctx->stream));
    int rc = sequence(ctx, reinterpret_cast<int32_t*>(s->a[0][8]), n, 0);
    if (rc) return rc;
    RYO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RYO_OK;
}

int ryo_system_store(ryo_system* s, float* x, float* y, float* z, float* m, float* h)
{
    if (!s) return RYO_ERR_ARG;
    Context* ctx = s->ctx;
    float* dst[5] = {x, y, z, m, h};
    for (int k = 0; k < 5; ++k)
        if (dst[k])
            RYO_HIP(ctx, hipMemcpyAsync(dst[k], s->a[s->cur][k], s->n * sizeof(float), hipMemcpyDeviceToDevice,
                                        ctx->stream));
    return RYO_OK;
}

static void mark(ryo_system* s, int i)
{
    if (s->ctx->timing) hipEventRecord(s->ev[i], s->ctx->stream);
}

// the build's phase times, once its events have completed (the caller has synchronized the stream)
static void resolveTiming(ryo_system* s)
{
    if (!s->timingPending) return;
    s->timingPending = false;
    if (hipEventSynchronize(s->ev[6]) != hipSuccess) return;
    for (int k = 0; k < 6; ++k)
    {
        float ms = 0;
        hipEventElapsedTime(&ms, s->ev[k], s->ev[k + 1]);
        s->phaseMs[k] = ms;
    }
}

// the counts of a device-sized build: 1 when the leaves outgrew the arrays (build again after growTree), an error
// when the layout does not end at n, else the leaf and node counts are taken over
constexpr int kBuildRegrow = 1;
static int resolveStatus(ryo_system* s)
{
    if (!s->statPending) return RYO_OK;
    Context* ctx = s->ctx;
    RYO_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s->statPending = false;
    const uint32_t* h = s->hStat;
    s->leafHint       = size_t(h[kStL]) + h[kStL] / 4 + 64;
    if (h[kStOverflow]) return kBuildRegrow;
    if (h[kStLast] != uint32_t(s->n))
        return setError(ctx, RYO_ERR_HIP, "leaves: layout[L] = %u != n = %zu", h[kStLast], s->n);
    s->L      = int(h[kStL]);
    s->nNodes = int(h[kStNodes]);
    return RYO_OK;
}

int ryo_system_build(ryo_system* s, const ryo_params* prm);

// the counts of a device-sized build taken over; leaves that outgrew the arrays are built again with grown arrays
// (from the last build's parameters), so the tree matches the particles before anyone reads it
static int settleBuild(ryo_system* s)
{
    int rc = resolveStatus(s);
    if (rc != kBuildRegrow) return rc;
    s->haveAcc = false;
    if ((rc = growTree(s, int(s->hStat[kStL]))) || (rc = ryo_system_build(s, &s->buildPrm))) return rc;
    rc = resolveStatus(s);
    return rc == kBuildRegrow ? setError(s->ctx, RYO_ERR_CAPACITY, "system: leaves outgrew the arrays after growth")
                              : rc;
}

// target groups of a built system (fixed groups of group_size, or the spatial split of 64-target windows); the
// build's phase times are read at the next synchronization
static int buildGroups(ryo_system* s, const ryo_params* prm, float** A)
{
    Context* ctx   = s->ctx;
    const size_t n = s->n;
    int rc         = RYO_OK;
    const int L    = s->L;
    int gsz = prm->group_size > 0 ? prm->group_size : 64;
    if (gsz > 64) return setError(ctx, RYO_ERR_ARG, "system: group size must be <= 64");
    if (!(prm->group_split_tol >= 0.0f))
        return setError(ctx, RYO_ERR_ARG, "system: group_split_tol must be >= 0");
    const bool spatial = prm->group_split_tol > 0.0f;
    if (spatial && gsz != 64)
        return setError(ctx, RYO_ERR_UNSUPPORTED, "system: spatial groups split windows of 64 targets");
    size_t G = (n + gsz - 1) / gsz;
    for (int attempt = 0;; ++attempt)
    {
        if (G + 1 > s->groupCap)
        {
            s->groupCap = 0;
            if ((rc = devAlloc(s, &s->groups, G + 1 + G / 4))) return rc;
            s->groupCap = G + 1 + G / 4;
        }
        if (!spatial)
        {
            rc = fixedGroups(ctx, 0, uint32_t(n), unsigned(gsz), s->groups, &s->G);
            break;
        }
        // the split count is known only after the split kernel: grow and split again when it exceeds capacity
        rc = groupSplits(ctx, 0, uint32_t(n), A[0], A[1], A[2], A[4], s->leaves, L, s->layout, s->limits, 64u,
                         prm->group_split_tol, 8.0f, s->groups, s->groupCap, &s->G);
        if (rc != RYO_ERR_CAPACITY || attempt > 0) break;
        G = size_t(s->G);
    }
    if (rc) return rc;
    s->timingPending = ctx->timing != 0;
    return RYO_OK;
}

int ryo_system_build(ryo_system* s, const ryo_params* prm)
{
    if (!s || !prm) return RYO_ERR_ARG;
    Context* ctx = s->ctx;
    const size_t n = s->n;
    if (n == 0) return setError(ctx, RYO_ERR_ARG, "system: no particles loaded");
    float** A = s->a[s->cur];
    float** B = s->a[s->cur ^ 1];
    int rc;
    s->haveAcc = false;
    // x, y, z, m, h and the original index always; velocities only once the system integrates
    const int idx[9] = {0, 1, 2, 3, 4, 8, 5, 6, 7};
    const int num    = s->integrating ? 9 : 6;
    const float* in[9];
    float* out[9];
    for (int k = 0; k < num; ++k)
        in[k] = A[idx[k]], out[k] = B[idx[k]];
    // arrays in the caller's order (a fresh load): the keys pass also writes the identity permutation and one
    // 32-byte record per particle, which the gather after the sort reads with one cache line per particle; a
    // time-stepping system is already close to SFC order, where the plain gather of the arrays is cheaper
    const bool packed = s->freshOrder && num == 6;
    float4* rec = nullptr;
    mark(s, 0);
    if ((rc = packed ? sfcKeysPacked(ctx, in, s->keys, nullptr, n, s->limits, &rec)
                     : sfcKeys(ctx, A[0], A[1], A[2], s->keys, n, s->limits, 0, 0)))
        return rc;
    mark(s, 1);
    if ((rc = sortPairsIndex(ctx, s->keys, s->perm, n, s->keyBuf, s->permBuf, s->sortTmp, s->sortTmpBytes)))
        return rc;
    mark(s, 2);
    if ((rc = packed ? gatherRecords(ctx, s->perm, n, rec, out) : gather(ctx, s->perm, n, num, in, out))) return rc;
    s->freshOrder = false;
    s->cur ^= 1;
    A = s->a[s->cur];
    mark(s, 3);
    if (s->leafCap == 0 && (rc = growTree(s, int(std::max<size_t>(64, n / 8))))) return rc;
    if (!s->levelRange && (rc = devAlloc(s, &s->levelRange, 24))) return rc;
    if (prm->multipole != 0 && prm->multipole != 1)
        return setError(ctx, RYO_ERR_ARG, "system: multipole must be 0 (quadrupole) or 1 (MDQpole)");
    if (prm->multipole == 1 && prm->soft_mac)
        return setError(ctx, RYO_ERR_UNSUPPORTED, "system: the softening-safe MAC is for quadrupoles");
    // sized on the device (no host round trip before the traversal): leaves into the current capacity, the linked
    // tree and the upsweep sized from the device counts; the counts are checked after the gravity pass, which
    // builds again with larger arrays in the rare case the leaves outgrew them. The single-leaf system, the
    // softening-safe MAC (which reads the softening range) and spatial groups (which read the leaf count) take the
    // host-sized sequence below.
    const bool deviceSized = n > prm->bucket_size && prm->bucket_size > 0 && prm->bucket_size <= 1024 &&
                             !prm->soft_mac && !(prm->group_split_tol > 0.0f);
    s->statPending = false;
    s->buildPrm    = *prm;
    if (deviceSized)
    {
        // launches for the last leaf count plus a quarter, or n / 8 leaves when none is known yet (the arrays
        // may be far larger: they are sized for the system's capacity)
        const size_t want = s->leafHint ? s->leafHint : std::max<size_t>(64, n / 8);
        const size_t cap  = std::min(want, s->leafCap);
        if ((rc = buildLeavesDevice(ctx, s->keys, n, prm->bucket_size, s->leaves, s->counts, s->layout, cap,
                                    s->dStat)))
            return rc;
        mark(s, 4);
        if ((rc = linkOctreeDevice(ctx, s->leaves, cap, s->dStat, s->prefixes, s->child, s->parents, s->levelRange,
                                   s->i2l, s->l2i)))
            return rc;
        mark(s, 5);
        s->mdq      = prm->multipole == 1;
        const int C = int(cap);
        if ((rc = s->mdq ? upsweepMDQ(ctx, A[0], A[1], A[2], A[3], s->layout, C, s->prefixes, s->child, s->parents,
                                      s->l2i, s->levelRange, s->limits, 1.0f / prm->theta, s->centers, s->mp, s->c64,
                                      s->q64, s->dStat)
                         : upsweep(ctx, A[0], A[1], A[2], A[3], s->layout, C, s->prefixes, s->child, s->parents,
                                   s->l2i, s->levelRange, s->limits, 1.0f / prm->theta, s->centers, s->mp, s->c64,
                                   s->q64, 0.0, s->dStat)))
            return rc;
        RYO_HIP(ctx, hipMemcpyAsync(s->hStat, s->dStat, kStWords * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                    ctx->stream));
        s->statPending = true;
        mark(s, 6);
        return buildGroups(s, prm, A);
    }
    // leaves: count pass + host read of L, then emission into buffers of the right size
    int L = 0;
    rc = buildLeaves(ctx, s->keys, n, prm->bucket_size, s->leaves, s->counts, s->layout, s->leafCap, &L);
    if (rc == RYO_ERR_CAPACITY)
    {
        if ((rc = growTree(s, L))) return rc;
        rc = buildLeaves(ctx, s->keys, n, prm->bucket_size, s->leaves, s->counts, s->layout, s->leafCap, &L);
    }
    if (rc) return rc;
    if ((rc = growTree(s, L))) return rc;
    s->L      = L;
    s->nNodes = L + (L - 1) / 7;
    mark(s, 4);
    if ((rc = linkOctree(ctx, s->leaves, L, s->prefixes, s->child, s->parents, s->levelRange, s->i2l, s->l2i,
                         nullptr)))
        return rc;
    mark(s, 5);
    s->mdq           = prm->multipole == 1;
    double softReach = 0;
    if (prm->soft_mac)
    {
        // the softening-safe MAC needs the largest softening of the system (one read-back)
        float hmin = 0, hmax = 0;
        if ((rc = softeningRange(ctx, A[4], n, &hmin, &hmax))) return rc;
        softReach = 2.0 * double(hmax);
    }
    if ((rc = s->mdq ? upsweepMDQ(ctx, A[0], A[1], A[2], A[3], s->layout, L, s->prefixes, s->child, s->parents,
                                  s->l2i, s->levelRange, s->limits, 1.0f / prm->theta, s->centers, s->mp, s->c64,
                                  s->q64)
                     : upsweep(ctx, A[0], A[1], A[2], A[3], s->layout, L, s->prefixes, s->child, s->parents, s->l2i,
                               s->levelRange, s->limits, 1.0f / prm->theta, s->centers, s->mp, s->c64, s->q64,
                               softReach)))
        return rc;
    mark(s, 6);
    return buildGroups(s, prm, A);
}

int ryo_system_gravity(ryo_system* s, const ryo_params* prm, double* totalPot)
{
    if (!s || !prm) return RYO_ERR_ARG;
    Context* ctx = s->ctx;
    if (prm->ewald < 0 || prm->ewald > 2) return setError(ctx, RYO_ERR_ARG, "system: ewald must be 0, 1 or 2");
    const size_t n = s->n;
    float** A = s->a[s->cur];
    int rc;
    for (float* o : {s->p, s->ax, s->ay, s->az})
        RYO_HIP(ctx, hipMemsetAsync(o, 0, n * sizeof(float), ctx->stream));
    float boxL[3] = {s->limits[1] - s->limits[0], s->limits[3] - s->limits[2], s->limits[5] - s->limits[4]};
    int shells    = s->periodic ? prm->num_shells : 0;
    double pot    = 0;
    if (s->mdq && s->periodic && prm->ewald)
        return setError(ctx, RYO_ERR_UNSUPPORTED, "system: the Ewald correction is for quadrupole systems");
    rc = traverseEntry(ctx, s->G, s->groups, s->groups + 1, prm->init_node, A[0], A[1], A[2], A[3], A[4], A[0], A[1],
                       A[2], A[3], A[4], s->child, s->i2l, s->layout, s->centers, s->mp, prm->G, shells, boxL, s->p,
                       s->ax, s->ay, s->az, &pot, 0, s->n, s->mdq, false,
                       s->statPending ? s->dStat + kStAbort : nullptr);
    // a device-sized build is checked here, with the traversal's synchronization: leaves that outgrew the arrays
    // (whatever the traversal made of the clipped tree) mean one more build with grown arrays and a new pass
    if (s->statPending)
    {
        const int st = resolveStatus(s);
        if (st == kBuildRegrow)
        {
            // the arrays grow past the leaf count just read, so the second build fits
            if (s->regrowing) return setError(ctx, RYO_ERR_CAPACITY, "system: leaves outgrew the arrays after growth");
            s->regrowing = true;
            int brc = growTree(s, int(s->hStat[kStL]));
            if (!brc) brc = ryo_system_build(s, &s->buildPrm);
            if (!brc) brc = ryo_system_gravity(s, prm, totalPot);
            s->regrowing = false;
            return brc;
        }
        if (st) return st;
    }
    resolveTiming(s);
    if (rc) return rc;
    s->phaseMs[6] = ctx->lastTraverseMs;
    s->phaseMs[7] = 0;
    if (s->periodic && prm->ewald)
    {
        double root[8], rc64[4];
        RYO_HIP(ctx, hipMemcpyAsync(root, s->q64, sizeof(root), hipMemcpyDeviceToHost, ctx->stream));
        RYO_HIP(ctx, hipMemcpyAsync(rc64, s->c64, sizeof(rc64), hipMemcpyDeviceToHost, ctx->stream));
        RYO_HIP(ctx, hipStreamSynchronize(ctx->stream));
        double ep = 0;
        mark(s, 7);
        if ((rc = ewald(ctx, root, rc64, 0, uint32_t(n), A[0], A[1], A[2], A[3], s->limits, prm->G,
                        prm->ewald_settings, s->p, s->ax, s->ay, s->az, &ep, prm->ewald == 2)))
            return rc;
        mark(s, 8);
        if (ctx->timing)
        {
            float ms = 0;
            RYO_HIP(ctx, hipEventSynchronize(s->ev[8]));
            hipEventElapsedTime(&ms, s->ev[7], s->ev[8]);
            s->phaseMs[7] = ms;
        }
        pot += ep;
    }
    s->haveAcc    = true;
    s->lastG      = prm->G;
    s->lastShells = shells;
    s->lastEwald  = (s->periodic && prm->ewald) ? 1 : 0;
    if (totalPot) *totalPot = pot;
    return RYO_OK;
}

} // extern "C"

namespace ryo
{
/* Accuracy statistics of solved particles (device arrays x, y, z, m, h and their results p, ax, ay, az) against the
 * FP64 direct sum (shells > 0: over the same replica images of boxL) or, with ewaldL > 0, the FP64 Ewald pair sum of
 * a cubic box of that side, on num_targets splitmix64 draws of the seed: out = {median, 99th percentile, maximum
 * of |a - a_ref| / |a_ref|, relative error of the sampled potential sum}. Synchronizes the stream. */
int accuracyStats(Context* ctx, const float* x, const float* y, const float* z, const float* m, const float* h,
                  size_t n, const float* p, const float* ax, const float* ay, const float* az, double G, int shells,
                  const double* boxL, double ewaldL, int num_targets, uint64_t seed, double* out)
{
    const size_t nt = std::min(size_t(num_targets), n);
    // targets: splitmix64 draws of the seed, with replacement
    std::vector<int64_t> tg(nt);
    for (size_t k = 0; k < nt; ++k)
    {
        uint64_t zz = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
        zz          = (zz ^ (zz >> 30)) * 0xBF58476D1CE4E5B9ull;
        zz          = (zz ^ (zz >> 27)) * 0x94D049BB133111EBull;
        tg[k]       = int64_t((zz ^ (zz >> 31)) % n);
    }
    const size_t refBytes = nt * 4 * sizeof(double), tgBytes = nt * sizeof(int64_t);
    auto* buf = static_cast<char*>(ctx->getScratch(kSlotAccuracy, refBytes + tgBytes + nt * 4 * (sizeof(float) + 4) + 1024));
    if (!buf) return RYO_ERR_HIP;
    auto* ref  = reinterpret_cast<double*>(buf);
    auto* dtg  = reinterpret_cast<int64_t*>(buf + refBytes);
    auto* idx  = reinterpret_cast<int32_t*>(buf + refBytes + tgBytes);
    auto* got  = reinterpret_cast<float*>(buf + refBytes + tgBytes + ((nt * 4 + 255) & ~size_t(255)));
    std::vector<int32_t> ti(tg.begin(), tg.end());
    hipStream_t st = ctx->stream;
    RYO_HIP(ctx, hipMemcpyAsync(dtg, tg.data(), tgBytes, hipMemcpyHostToDevice, st));
    RYO_HIP(ctx, hipMemcpyAsync(idx, ti.data(), nt * 4, hipMemcpyHostToDevice, st));
    int rc = ewaldL > 0 ? ewaldPairSum(ctx, x, y, z, h, m, n, dtg, nt, G, ewaldL, 0.0, ref)
                        : directSum(ctx, x, y, z, h, m, n, dtg, nt, G, shells ? boxL : nullptr, shells, ref);
    if (rc) return rc;
    const float* src[4] = {p, ax, ay, az};
    float* dst[4]       = {got, got + nt, got + 2 * nt, got + 3 * nt};
    if ((rc = gather(ctx, idx, nt, 4, src, dst))) return rc;
    std::vector<double> hr(nt * 4);
    std::vector<float> hg(nt * 4);
    RYO_HIP(ctx, hipMemcpyAsync(hr.data(), ref, refBytes, hipMemcpyDeviceToHost, st));
    RYO_HIP(ctx, hipMemcpyAsync(hg.data(), got, nt * 4 * sizeof(float), hipMemcpyDeviceToHost, st));
    RYO_HIP(ctx, hipStreamSynchronize(st));
    std::vector<double> e(nt);
    double potRef = 0, potGot = 0;
    for (size_t k = 0; k < nt; ++k)
    {
        double dx = hg[nt + k] - hr[4 * k + 1], dy = hg[2 * nt + k] - hr[4 * k + 2], dz = hg[3 * nt + k] - hr[4 * k + 3];
        double r  = std::sqrt(hr[4 * k + 1] * hr[4 * k + 1] + hr[4 * k + 2] * hr[4 * k + 2] + hr[4 * k + 3] * hr[4 * k + 3]);
        e[k]      = std::sqrt(dx * dx + dy * dy + dz * dz) / std::max(r, 1e-300);
        potRef += hr[4 * k];
        potGot += hg[k];
    }
    std::sort(e.begin(), e.end());
    out[0] = e[nt / 2];
    out[1] = e[std::min(nt - 1, size_t(0.99 * double(nt)))];
    out[2] = e[nt - 1];
    out[3] = std::abs(potGot - potRef) / std::max(std::abs(potRef), 1e-300);
    return RYO_OK;
}
} // namespace ryo

extern "C" {

int ryo_system_accuracy(ryo_system* s, int num_targets, uint64_t seed, double* out)
{
    if (!s || !out || num_targets < 1) return RYO_ERR_ARG;
    Context* ctx = s->ctx;
    if (!s->haveAcc) return setError(ctx, RYO_ERR_ARG, "accuracy: no gravity results to check");
    float** A = s->a[s->cur];
    double boxL[3] = {double(s->limits[1] - s->limits[0]), double(s->limits[3] - s->limits[2]),
                      double(s->limits[5] - s->limits[4])};
    double ewaldL = 0;
    if (s->lastEwald)
    {
        // the full periodic result (images + Ewald correction) against the FP64 Ewald pair sum
        if (!(boxL[0] == boxL[1] && boxL[1] == boxL[2]))
            return setError(ctx, RYO_ERR_UNSUPPORTED, "accuracy: the Ewald reference needs a cubic box");
        ewaldL = boxL[0];
    }
    return accuracyStats(ctx, A[0], A[1], A[2], A[3], A[4], s->n, s->p, s->ax, s->ay, s->az, double(s->lastG),
                         s->lastShells, boxL, ewaldL, num_targets, seed, out);
}

int ryo_system_results(ryo_system* s, float** out, int32_t** order)
{
    if (!s) return RYO_ERR_ARG;
    if (out)
    {
        out[0] = s->p, out[1] = s->ax, out[2] = s->ay, out[3] = s->az;
    }
    if (order) *order = reinterpret_cast<int32_t*>(s->a[s->cur][8]);
    return RYO_OK;
}

int ryo_system_info(ryo_system* s, int64_t* info)
{
    if (!s || !info) return RYO_ERR_ARG;
    if (int rc = settleBuild(s)) return rc;
    info[0] = int64_t(s->n), info[1] = s->L, info[2] = s->nNodes, info[3] = s->G;
    return RYO_OK;
}

int ryo_system_tree(ryo_system* s, void** p)
{
    if (!s || !p) return RYO_ERR_ARG;
    if (int rc = settleBuild(s)) return rc;
    void* v[14] = {s->keys, s->leaves, s->counts, s->layout, s->prefixes, s->child, s->parents, s->levelRange,
                   s->i2l,  s->l2i,    s->centers, s->mp,    s->c64,      s->q64};
    for (int k = 0; k < 14; ++k)
        p[k] = v[k];
    return RYO_OK;
}

int ryo_system_leapfrog(ryo_system* s, const ryo_params* prm, float dt, int steps, double* totalPot)
{
    if (!s || !prm || steps < 0) return RYO_ERR_ARG;
    Context* ctx = s->ctx;
    int rc;
    double pot = 0;
    if (!s->integrating)
    {
        // velocities are zero after load and were not carried by earlier builds: zero the current copies, then
        // let every later permutation carry them
        for (int k = 5; k < 8; ++k)
            RYO_HIP(ctx, hipMemsetAsync(s->a[s->cur][k], 0, s->n * sizeof(float), ctx->stream));
        s->integrating = true;
    }
    if (!s->haveAcc)
    {
        if ((rc = ryo_system_build(s, prm)) || (rc = ryo_system_gravity(s, prm, &pot))) return rc;
    }
    for (int k = 0; k < steps; ++k)
    {
        float** A = s->a[s->cur];
        if ((rc = kickDrift(ctx, A[0], A[1], A[2], A[5], A[6], A[7], s->ax, s->ay, s->az, s->n, 0.5f * dt, dt,
                            s->periodic, s->limits)))
            return rc;
        if (!s->periodic && (rc = fitBox(ctx, A[0], A[1], A[2], s->n, s->limits))) return rc;
        if ((rc = ryo_system_build(s, prm)) || (rc = ryo_system_gravity(s, prm, &pot))) return rc;
        A = s->a[s->cur];
        if ((rc = kickDrift(ctx, A[0], A[1], A[2], A[5], A[6], A[7], s->ax, s->ay, s->az, s->n, 0.5f * dt, 0.0f,
                            s->periodic, s->limits)))
            return rc;
    }
    if (totalPot) *totalPot = pot;
    return RYO_OK;
}

int ryo_system_limits(ryo_system* s, float* limits)
{
    if (!s || !limits) return RYO_ERR_ARG;
    std::memcpy(limits, s->limits, sizeof(s->limits));
    return RYO_OK;
}

int ryo_system_phase_ms(ryo_system* s, double* ms)
{
    if (!s || !ms) return RYO_ERR_ARG;
    if (s->timingPending) hipStreamSynchronize(s->ctx->stream);
    resolveTiming(s);
    for (int k = 0; k < 8; ++k)
        ms[k] = s->phaseMs[k];
    return RYO_OK;
}

} // extern "C"
