/*
 * ryoanji_amd.h -- C ABI of the MI355X (gfx950) Barnes-Hut gravity path.
 *
 * Plain pointers and sizes only. Every array argument is DEVICE memory owned by the caller unless
 * it is documented as a host pointer. Every function is stream-ordered on the context's HIP stream,
 * returns an int status (RYO_OK or an error code, message via ryo_last_error) and never exits the
 * process. All per-call state (counters, statistics, scratch) lives in the context, so contexts on
 * different streams are independent; calls that share one context must not overlap.
 *
 * Each entry point names the reference interface it replaces (paths relative to the reference
 * checkout: cstone/include/cstone/..., ryoanji/src/ryoanji/...).
 *
 * Array conventions (reference types): keys are 64-bit SFC keys (HilbertKey<uint64_t>), node
 * indices are int32 (TreeNodeIndex), particle indices uint32 (LocalIndex). Source centers are
 * float4 {x, y, z, mac^2}; quadrupoles are 8 floats {M, qxx, qxy, qxz, qyy, qyz, qzz, trace}
 * (CartesianQuadrupole<float>, ryoanji/nbody/cartesian_qpole.hpp:27-51).
 */
#ifndef RYOANJI_AMD_H
#define RYOANJI_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum
{
    RYO_OK              = 0,
    RYO_ERR_ARG         = 1, /* invalid argument (reference: std::runtime_error)          */
    RYO_ERR_HIP         = 2, /* HIP runtime error (reference: checkGpuErrors -> exit())   */
    RYO_ERR_UNSUPPORTED = 3, /* e.g. non-cubic box for keys or Ewald                      */
    RYO_ERR_CAPACITY    = 4, /* output capacity too small; the needed size is reported    */
    RYO_ERR_TRAVERSAL   = 5  /* a single target exhausted the traversal stack             */
};

typedef struct ryo_ctx ryo_ctx;

/* ---------------------------------------------------------------- context */

/* Replaces the reference's process-global device state (traversal_gpu.cu:351-356, csarray_gpu.cu:68,130,
 * ewald.cu:23) with per-context state. `stream` is a hipStream_t (NULL: the null stream). */
int ryo_ctx_create(ryo_ctx** ctx, void* stream);
/* Destroy every system created on a context before the context itself (systems use its stream and scratch). */
int ryo_ctx_destroy(ryo_ctx* ctx);
const char* ryo_last_error(const ryo_ctx* ctx);
int ryo_ctx_sync(ryo_ctx* ctx);
/* enable HIP-event timing of the traversal kernel (ryo_last_traverse_ms) */
int ryo_ctx_set_timing(ryo_ctx* ctx, int enable);
double ryo_last_traverse_ms(const ryo_ctx* ctx);
/* Launch shape of the traversal on this context: LIFO entries per wave held in LDS (9..1408; a walk that needs
 * more takes the split/retry path), blocks of 4 waves launched per CU, and the number of contiguous group
 * ranges with their own work counters (1..8; blocks start on range blockIdx % ranges, which under round-robin
 * dispatch keeps each XCD's L2 on SFC-neighbouring groups). 0 keeps the default (1408, 24 waves per CU, 8).
 * Replaces the compile-time TravConfig (traversal_gpu.cu:45-64), for the autotuner. */
int ryo_ctx_set_traversal_shape(ryo_ctx* ctx, int stack_entries_per_wave, int blocks_per_cu, int group_ranges);
/* Waves per traversal block: 2, 4 or 8 (0: 4). Every block size keeps 6 waves per SIMD resident; the default
 * blocks per CU follow (12, 6, 3). The block size of the reference's TravConfig (traversal_gpu.cu:45-64). */
int ryo_ctx_set_traversal_waves(ryo_ctx* ctx, int waves_per_block);

/* ---------------------------------------------------------------- a2: SFC keys */

/* cstone::computeSfcKeys(execution::Gpu, x, y, z, keys, n, Box<float>) -- cstone/sfc/sfc_gpu.h:37,
 * sfc_gpu.cu:23-44. limits = host float[6] {xmin,xmax,ymin,ymax,zmin,zmax}; kind 0 = Hilbert, 1 = Morton.
 * With skip_removed != 0 a key already equal to 2^63 (removeKey) is left untouched, as in the reference;
 * with 0 the keys array is write-only. Non-cubic boxes get the mixed-dimension Hilbert keys of
 * hilbert.hpp:117-244 (box bits reduced per axis as Box<float>::getBoxDimBits, box.hpp:160-198). */
int ryo_sfc_keys_f32(ryo_ctx* ctx, const float* x, const float* y, const float* z, uint64_t* keys, size_t n,
                     const float* limits, int kind, int skip_removed);
/* cstone::computeSfcKeys for double coordinates (the Box<double> instantiations, sfc_gpu.cu:46-62): cell indices
 * and box bits in double as Box<double> and sfcCellIndex<double> compute them (box.hpp:160-198, sfc.hpp:153-159);
 * limits = host double[6]; otherwise as ryo_sfc_keys_f32 (mixed-dimension keys for non-cubic boxes included). */
int ryo_sfc_keys_f64(ryo_ctx* ctx, const double* x, const double* y, const double* z, uint64_t* keys, size_t n,
                     const double* limits, int kind, int skip_removed);
/* computeSfcKeys with 32-bit keys (HilbertKey<uint32_t> / MortonKey<uint32_t>, sfc_gpu.cu:46-62): 10 levels,
 * box bits reduced from maxTreeLevel<uint32_t> = 10 (sfc.hpp:89-96), removeKey 2^30 for skip_removed; float or
 * double coordinates with a box of the same type. */
int ryo_sfc_keys_u32_f32(ryo_ctx* ctx, const float* x, const float* y, const float* z, uint32_t* keys, size_t n,
                         const float* limits, int kind, int skip_removed);
int ryo_sfc_keys_u32_f64(ryo_ctx* ctx, const double* x, const double* y, const double* z, uint32_t* keys, size_t n,
                         const double* limits, int kind, int skip_removed);

/* ---------------------------------------------------------------- a3: stable radix sort */

/* cstone::sortByKeyTempStorage<uint64_t, int>(n) -- cstone/primitives/primitives_gpu.h:120-124 */
size_t ryo_sort_temp_bytes(size_t n);
/* cstone::sortByKey(exec, keys, keys + n, values, keyBuf, valueBuf, tmp, tmpBytes) --
 * primitives_gpu.h:126-130, primitives_gpu.cu:310-358: stable sort of (u64 key, i32 value) pairs over all 64
 * key bits (equal keys keep their order); the result is in keys/values, keyBuf/valueBuf are scratch of n elements
 * each, tmp holds ryo_sort_temp_bytes(n) bytes. */
int ryo_sort_pairs_u64_i32(ryo_ctx* ctx, uint64_t* keys, int32_t* values, size_t n, uint64_t* keyBuf,
                           int32_t* valueBuf, void* tmp, size_t tmpBytes);
/* No counterpart in the reference: after ryo_sort_pairs_u64_i32 on the same n and tmp, waits for the context's
 * stream and reports RYO_ERR_HIP if the sort's kernels found offsets or work lists that do not describe the keys
 * (they skip such accesses instead of reading or writing outside the arrays, so the order would be wrong). */
int ryo_sort_check(ryo_ctx* ctx, const void* tmp, size_t n);
/* cstone::sequence (primitives_gpu.cu:265-273) */
int ryo_sequence_i32(ryo_ctx* ctx, int32_t* out, size_t n, int32_t start);

/* ---------------------------------------------------------------- a4: permutation */

/* out_k[i] = in_k[perm[i]] for k < num_arrays (<= 10) -- the thrust::gather of TreeBuilder::update
 * (ryoanji/interface/treebuilder.cuh:74-79), applied to all arrays at once. `in`/`out` are HOST arrays of
 * device pointers. */
int ryo_gather_f32(ryo_ctx* ctx, const int32_t* perm, size_t n, int num_arrays, const float* const* in,
                   float* const* out);
/* the same gather for double arrays (TreeBuilder over double coordinates) */
int ryo_gather_f64(ryo_ctx* ctx, const int32_t* perm, size_t n, int num_arrays, const double* const* in,
                   double* const* out);

/* ---------------------------------------------------------------- a5+a7: cornerstone leaves */

/* cstone::computeOctree / updateOctreeGpu iterated to convergence (csarray.hpp:408-440, update_gpu.cuh:43-68)
 * plus TreeBuilder's layout (treebuilder.cuh:95-98), built directly in one pass:
 *   leaves  [capacity + 1] u64 : cornerstone leaf keys, leaves[L] = 2^63
 *   counts  [capacity]     u32 : particles per leaf
 *   layout  [capacity + 1] u32 : first particle of each leaf, layout[L] = n
 * *num_leaves (host) receives L. If L > capacity, returns RYO_ERR_CAPACITY and *num_leaves = L. */
int ryo_build_leaves_u64(ryo_ctx* ctx, const uint64_t* sorted_keys, size_t n, unsigned bucket_size,
                         uint64_t* leaves, uint32_t* counts, uint32_t* layout, size_t capacity, int* num_leaves);

/* The cornerstone leaf array one count / rebalance step at a time, for callers that iterate
 * updateOctreeGpu themselves (cstone/tree/update_gpu.cuh:43-68; ryo_build_leaves_u64 gives the converged array in
 * one call). `tree` holds num_nodes + 1 keys ending with 2^63.
 * cstone::computeNodeCountsGpu (csarray_gpu.h:41-48, csarray_gpu.cu:86-122): counts[i] = particles of the sorted
 * keys[0..n) in [tree[i], tree[i+1]), capped at max_count (the reference's count guess only speeds its search). */
int ryo_node_counts_u64(ryo_ctx* ctx, const uint64_t* tree, int num_nodes, const uint64_t* keys, size_t n,
                        unsigned max_count, uint32_t* counts);
/* cstone::computeNodeOpsGpu (csarray_gpu.h:59-65, csarray_gpu.cu:182-205): node_ops [num_nodes + 1] receives the
 * exclusive scan of the per-node operation (0 merge, 1 keep, 8/64/512/4096 split; calculateNodeOp,
 * csarray.hpp:267-290); *new_num_nodes (host) = node_ops[num_nodes]. Whether any node changed is kept in the context
 * for the next ryo_rebalance_tree_u64 (the reference's rebalanceChangeCounter). Synchronizes the stream. */
int ryo_node_ops_u64(ryo_ctx* ctx, const uint64_t* tree, int num_nodes, const uint32_t* counts, unsigned bucket_size,
                     int32_t* node_ops, int* new_num_nodes);
/* cstone::rebalanceTreeGpu (csarray_gpu.h:67-73, csarray_gpu.cu:212-231): new_tree [new_num_nodes + 1] from the
 * scanned operations (processNode, csarray.hpp:339-366); *converged (host) = no node changed in the last
 * ryo_node_ops_u64 on this context. Synchronizes the stream. */
int ryo_rebalance_tree_u64(ryo_ctx* ctx, const uint64_t* tree, int num_nodes, int new_num_nodes,
                           const int32_t* node_ops, uint64_t* new_tree, int* converged);

/* ---------------------------------------------------------------- a6: linked octree */

/* cstone::buildOctreeGpu(exec, leaves, OctreeView) -- cstone/tree/octree_gpu.h:35-45, octree_gpu.cu:138-200.
 * numInternal = (L-1)/7, numNodes = L + numInternal. Sizes: prefixes, internal_to_leaf, leaf_to_internal
 * [numNodes]; child_offsets [numNodes + 1]; parents [max(1, (numNodes-1)/8)]; level_range [23] on the
 * device; level_range_host (optional, host int[23]) receives a copy, as OctreeData::levelRange. */
int ryo_link_octree_u64(ryo_ctx* ctx, const uint64_t* leaves, int num_leaves, uint64_t* prefixes,
                        int32_t* child_offsets, int32_t* parents, int32_t* level_range, int32_t* internal_to_leaf,
                        int32_t* leaf_to_internal, int32_t* level_range_host);

/* ---------------------------------------------------------------- a8-a12: centers, MAC, multipoles */

/* Upsweep in FP64 (no host round trip): leaf mass centers and P2M
 * (cstone::computeLeafSourceCenterGpu, ryoanji::computeLeafMultipoles), the |m|-weighted center and M2M
 * combination of internal nodes up to the root (cstone::upsweepCentersGpu, ryoanji::upsweepMultipoles,
 * source_center_gpu.h:66-116, upsweep_gpu.h:25,40) and the vector-MAC radius (cstone::setMacGpu,
 * source_center_gpu.cu:284-310). Particles are SFC-ordered float x,y,z,m; tree arrays from
 * ryo_link_octree_u64 (level_range is its device copy) and the leaf layout. Outputs: centers float4[numNodes] {com, mac^2},
 * multipoles float[numNodes][8]; centers64 double[numNodes][4] {com, sum|m|} and multipoles64
 * double[numNodes][8] are optional (NULL: kept internally). limits = host float[6]. */
int ryo_upsweep_f32(ryo_ctx* ctx, const float* x, const float* y, const float* z, const float* m,
                    const uint32_t* layout, int num_leaves, const uint64_t* prefixes, const int32_t* child_offsets,
                    const int32_t* parents, const int32_t* leaf_to_internal, const int32_t* level_range,
                    const float* limits, float inv_theta, float* centers, float* multipoles, double* centers64,
                    double* multipoles64);
/* The upsweep with CartesianMDQpole<float> multipoles (P2M / addQuadrupole / M2M for CartesianMDQpole,
 * cartesian_qpole.hpp:237-417): multipoles float[numNodes][12] {M, qxx, qxy, qxz, qyy, qyz, qzz, trace, px, py,
 * pz, 0}, the dipole p = sum m (r - com) taken about the |m|-weighted center (nonzero with signed masses);
 * multipoles64 (optional) double[numNodes][12]. Otherwise as ryo_upsweep_f32. */
int ryo_upsweep_mdq_f32(ryo_ctx* ctx, const float* x, const float* y, const float* z, const float* m,
                        const uint32_t* layout, int num_leaves, const uint64_t* prefixes, const int32_t* child_offsets,
                        const int32_t* parents, const int32_t* leaf_to_internal, const int32_t* level_range,
                        const float* limits, float inv_theta, float* centers, float* multipoles, double* centers64,
                        double* multipoles64);
/* The same upsweep for double particles and a Box<double> (computeLeafMultipoles / upsweepMultipoles / setMac for
 * T = double, upsweep_gpu.cu:134-143): centers double[numNodes][4] {com, mac^2} (SourceCenterType<double>),
 * multipoles double[numNodes][8] (CartesianQuadrupole<double>), multipoles32 (optional, NULL: not written) the
 * float copy for the (d,f,f,f,d) traversal. limits = host double[6]. */
int ryo_upsweep_f64(ryo_ctx* ctx, const double* x, const double* y, const double* z, const double* m,
                    const uint32_t* layout, int num_leaves, const uint64_t* prefixes, const int32_t* child_offsets,
                    const int32_t* parents, const int32_t* leaf_to_internal, const int32_t* level_range,
                    const double* limits, float inv_theta, double* centers, double* multipoles, float* multipoles32);
/* The double-particle upsweep with CartesianMDQpole<double> multipoles double[numNodes][12] (the layout of
 * ryo_upsweep_mdq_f32), multipoles32 (optional) the CartesianMDQpole<float> copy; otherwise as ryo_upsweep_f64. */
int ryo_upsweep_mdq_f64(ryo_ctx* ctx, const double* x, const double* y, const double* z, const double* m,
                        const uint32_t* layout, int num_leaves, const uint64_t* prefixes, const int32_t* child_offsets,
                        const int32_t* parents, const int32_t* leaf_to_internal, const int32_t* level_range,
                        const double* limits, float inv_theta, double* centers, double* multipoles,
                        float* multipoles32);

/* The upsweep one reference step at a time (the call sequence of ryoanji/test/demo.cu:192-219), for callers that
 * drive it themselves; ryo_upsweep_f32 / _f64 do the same in one FP64 call. `types` names the reference
 * instantiation: 0 (Tc, Tm, Tf) = (float, float, float) with MType<float>; 1 (double, double, double) with
 * MType<double>; 2 (double, float, double) with MType<float>. mdq != 0: CartesianMDQpole (12 values per node),
 * else CartesianQuadrupole (8). Centers are Tf[n][4]. Sums are formed in FP64 and stored in the caller's types.
 * cstone::computeLeafSourceCenterGpu (source_center_gpu.h:66-75, source_center_gpu.cu:103-162): the |m|-weighted
 * center {com, sum|m|} of each leaf, stored at leaf_to_internal[leaf] (the leaf part of the map, num_leaves
 * entries, as TreeBuilder::leafToInternal() returns it). */
int ryo_leaf_centers(ryo_ctx* ctx, int types, const void* x, const void* y, const void* z, const void* m,
                     const int32_t* leaf_to_internal, int num_leaves, const uint32_t* layout, void* centers);
/* cstone::upsweepCentersGpu (source_center_gpu.h:77-88, source_center_gpu.cu:172-205): levels num_levels - 1
 * down to 0, one launch per level from the HOST level ranges level_range_host[num_levels + 1] (as
 * OctreeData::levelRange); types 0 float centers, 1 or 2 double centers. */
int ryo_upsweep_centers(ryo_ctx* ctx, int types, int num_levels, const int32_t* level_range_host,
                        const int32_t* child_offsets, void* centers);
/* cstone::setMacGpu (source_center_gpu.h:107-116, source_center_gpu.cu:284-310): centers[i][3] = squared vector-MAC
 * radius (computeVecMacR2, macs.hpp:64-82, evaluated in FP64), 0 for empty cells. limits = host T[6] (float for
 * types 0, double otherwise). */
int ryo_set_mac(ryo_ctx* ctx, int types, const uint64_t* prefixes, int num_nodes, void* centers, float inv_theta,
                const void* limits);
/* ryoanji::computeLeafMultipoles (nbody/upsweep_gpu.h:25, upsweep_gpu.cu:33-77): P2M of each leaf about its
 * center (P2M_add / P2M_finalize, cartesian_qpole.hpp:67-117; the MDQpole form :237-295). */
int ryo_leaf_multipoles(ryo_ctx* ctx, int types, int mdq, const void* x, const void* y, const void* z, const void* m,
                        const int32_t* leaf_to_internal, int num_leaves, const uint32_t* layout, const void* centers,
                        void* multipoles);
/* ryoanji::upsweepMultipoles (nbody/upsweep_gpu.h:40, upsweep_gpu.cu:84-128): M2M of the cells
 * [first_cell, last_cell) from their 8 children about the stored centers (addQuadrupole, cartesian_qpole.hpp:187-235;
 * :359-417 for MDQpole); the caller loops over the levels. types: 0 (T, MType) = (float, float), 1 (double,
 * double), 2 (double, float). */
int ryo_upsweep_multipoles(ryo_ctx* ctx, int types, int mdq, int first_cell, int last_cell,
                           const int32_t* child_offsets, const void* centers, void* multipoles);

/* ---------------------------------------------------------------- a13: target groups */

/* cstone::computeFixedGroups (traversal/groups_gpu.cu:45-58): groups[g] = first + g*group_size,
 * groups[G] = last, G = ceil((last-first)/group_size); groups has G+1 entries; *num_groups (host) = G. */
int ryo_fixed_groups(ryo_ctx* ctx, uint32_t first, uint32_t last, unsigned group_size, uint32_t* groups,
                     int* num_groups);

/* cstone::computeGroupSplits (traversal/groups_gpu.cu:121-151, groups_gpu.cuh:149-270) for (Tc, T) = (f, f):
 * windows of group_size (= 64, bhMaxTargetSize) SFC-ordered targets [first, last) are split after target l
 * when |x_{l+1} - x_l| > min(tol_factor * cbrt(smallest leaf volume in the window), 2 h_l) in units of the
 * longest box edge (h may be NULL: radius 1), and windows longer than extent_ratio x tol_factor x the smallest
 * leaf edge along an axis are cut into equal chunks (extent_ratio 0 disables that; the reference default is 8,
 * CSTONE_GROUP_EXTENT_SPLIT). leaves/layout: the cornerstone leaf array ([num_leaves+1]) and its layout.
 * groups (device, capacity entries) receives the G+1 group boundaries; *num_groups (host) = G. Synchronizes
 * the stream (the group count is read back, as in the reference). RYO_ERR_CAPACITY when G+1 > capacity;
 * RYO_ERR_UNSUPPORTED for group sizes other than 64. */
int ryo_group_splits_f32(ryo_ctx* ctx, uint32_t first, uint32_t last, const float* x, const float* y, const float* z,
                         const float* h, const uint64_t* leaves, int num_leaves, const uint32_t* layout,
                         const float* limits, unsigned group_size, float tol_factor, float extent_ratio,
                         uint32_t* groups, size_t capacity, int* num_groups);

/* ---------------------------------------------------------------- a14-a17: traversal */

/* ryoanji::traverse (nbody/traversal_gpu.h:51-56, traversal_gpu.cu:572-623) for the float instantiation
 * (Tc,Th,Tm,Ta,Tf) = (f,f,f,f,f) with CartesianQuadrupole<float>.
 * group_start/group_end: num_groups target ranges (group_end may alias group_start + 1).
 * init_node 0: the traversal starts at the root, as the reference CPU traversal (traversal.hpp:26-69);
 * init_node 1: it starts with the root's children, as the reference GPU kernel.
 * Outputs are ADDED to p (may be NULL; adds G m_i phi_i), ax, ay, az. *total_pot (host, may be NULL)
 * receives 0.5 G sum_i m_i phi_i, reduced in FP64 in a fixed order. A group whose traversal overflows the
 * per-wave stack is split and traversed again (splitFailedGroups, traversal_gpu.cu:647-677); a single
 * target that still overflows gives RYO_ERR_TRAVERSAL. */
int ryo_traverse_f32(ryo_ctx* ctx, int num_groups, const uint32_t* group_start, const uint32_t* group_end,
                     int init_node, const float* xt, const float* yt, const float* zt, const float* mt,
                     const float* ht, const float* xs, const float* ys, const float* zs, const float* ms,
                     const float* hs, const int32_t* child_offsets, const int32_t* internal_to_leaf,
                     const uint32_t* layout, const float* centers, const float* multipoles, float G, int num_shells,
                     const float* box_l /* host float[3] */, float* p, float* ax, float* ay, float* az,
                     double* total_pot);

/* ryoanji::traverse with CartesianMDQpole<float> multipoles (float[n][12] from ryo_upsweep_mdq_f32): M2P adds the
 * dipole terms (M2P for CartesianMDQpole, cartesian_qpole.hpp:320-357); otherwise as ryo_traverse_f32
 * (no batched form). */
int ryo_traverse_mdq_f32(ryo_ctx* ctx, int num_groups, const uint32_t* group_start, const uint32_t* group_end,
                         int init_node, const float* xt, const float* yt, const float* zt, const float* mt,
                         const float* ht, const float* xs, const float* ys, const float* zs, const float* ms,
                         const float* hs, const int32_t* child_offsets, const int32_t* internal_to_leaf,
                         const uint32_t* layout, const float* centers, const float* multipoles, float G,
                         int num_shells, const float* box_l, float* p, float* ax, float* ay, float* az,
                         double* total_pot);

/* ryoanji::traverse for the (Tc,Th,Tm,Ta,Tf) = (d,d,d,d,d) instantiation with CartesianQuadrupole<double>
 * (traversal_gpu.cu:635-641): double coordinates, softenings, masses and outputs; centers double[n][4]
 * {center of mass, mac^2} (SourceCenterType<double>), multipoles double[n][8]. Every interaction is evaluated and
 * accumulated in double except 1/R, which the reference's inverseSquareRoot(double) takes in single precision
 * (kernel.hpp:39-53); same walk, MAC, grouping and output convention as ryo_traverse_f32 (init_node,
 * outputs ADDED, box_l host double[3] when num_shells > 0, *total_pot = 0.5 G sum m phi in FP64). */
int ryo_traverse_f64(ryo_ctx* ctx, int num_groups, const uint32_t* group_start, const uint32_t* group_end,
                     int init_node, const double* xt, const double* yt, const double* zt, const double* mt,
                     const double* ht, const double* xs, const double* ys, const double* zs, const double* ms,
                     const double* hs, const int32_t* child_offsets, const int32_t* internal_to_leaf,
                     const uint32_t* layout, const double* centers, const double* multipoles, double G,
                     int num_shells, const double* box_l, double* p, double* ax, double* ay, double* az,
                     double* total_pot);

/* ryoanji::traverse for (d,d,d,d,d) with CartesianMDQpole<double> and (d,f,f,f,d) with CartesianMDQpole<float>
 * (traversal_gpu.cu:635-642): the arguments of ryo_traverse_f64 / ryo_traverse_f64_f32 with 12 values per
 * node; M2P adds the dipole terms (cartesian_qpole.hpp:320-357). */
int ryo_traverse_mdq_f64(ryo_ctx* ctx, int num_groups, const uint32_t* group_start, const uint32_t* group_end,
                         int init_node, const double* xt, const double* yt, const double* zt, const double* mt,
                         const double* ht, const double* xs, const double* ys, const double* zs, const double* ms,
                         const double* hs, const int32_t* child_offsets, const int32_t* internal_to_leaf,
                         const uint32_t* layout, const double* centers, const double* multipoles, double G,
                         int num_shells, const double* box_l, double* p, double* ax, double* ay, double* az,
                         double* total_pot);
int ryo_traverse_mdq_f64_f32(ryo_ctx* ctx, int num_groups, const uint32_t* group_start, const uint32_t* group_end,
                             int init_node, const double* xt, const double* yt, const double* zt, const float* mt,
                             const float* ht, const double* xs, const double* ys, const double* zs, const float* ms,
                             const float* hs, const int32_t* child_offsets, const int32_t* internal_to_leaf,
                             const uint32_t* layout, const double* centers, const float* multipoles, double G,
                             int num_shells, const double* box_l, float* p, float* ax, float* ay, float* az,
                             double* total_pot);

/* ryoanji::traverse for the (d,f,f,f,d) instantiation with CartesianQuadrupole<float>: double coordinates and
 * centers, float softenings, masses, multipoles and outputs; interactions in double (the reference accumulates in
 * Tf = double), outputs rounded to float. Otherwise as ryo_traverse_f64. */
int ryo_traverse_f64_f32(ryo_ctx* ctx, int num_groups, const uint32_t* group_start, const uint32_t* group_end,
                         int init_node, const double* xt, const double* yt, const double* zt, const float* mt,
                         const float* ht, const double* xs, const double* ys, const double* zs, const float* ms,
                         const float* hs, const int32_t* child_offsets, const int32_t* internal_to_leaf,
                         const uint32_t* layout, const double* centers, const float* multipoles, double G,
                         int num_shells, const double* box_l, float* p, float* ax, float* ay, float* az,
                         double* total_pot);

/* No counterpart in the reference. ryo_traverse_f32 for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): its work memory (per-group FP64
 * potentials, overflow list, control words, reduction partials) is a fixed array that the library holds on each
 * GPU for the life of the process (12.0 MiB), so it serves at most RYO_NOALLOC_MAX_GROUPS target groups
 * (RYO_ERR_CAPACITY beyond). It allocates nothing. It always walks with the full-depth LIFO, which no group can
 * overflow, so there is no split/retry; where ryo_traverse_f32 needs none (its default LIFO depth) the results,
 * statistics and total potential are the same bit for bit. Every call on one GPU uses that GPU's one array:
 * calls that overlap on one GPU (other contexts or streams) race, so the caller must serialize them.
 * Synchronizes the context's stream (statistics and potential read-back). */
#define RYO_NOALLOC_MAX_GROUPS 1048576
int ryo_traverse_f32_noalloc(ryo_ctx* ctx, int num_groups, const uint32_t* group_start, const uint32_t* group_end,
                             int init_node, const float* xt, const float* yt, const float* zt, const float* mt,
                             const float* ht, const float* xs, const float* ys, const float* zs, const float* ms,
                             const float* hs, const int32_t* child_offsets, const int32_t* internal_to_leaf,
                             const uint32_t* layout, const float* centers, const float* multipoles, float G,
                             int num_shells, const float* box_l, float* p, float* ax, float* ay, float* az,
                             double* total_pot);

/* ryoanji::readBhStats (traversal_gpu.cu:688-700): host uint64[5] {sumP2P, maxP2P, sumM2P, maxM2P, maxStack}
 * of the last ryo_traverse_f32 call (sums are interactions summed over targets). */
int ryo_read_stats(ryo_ctx* ctx, uint64_t* stats);
/* ryoanji::readTraversalRetries (traversal_gpu.cu:645) */
uint64_t ryo_traversal_retries(const ryo_ctx* ctx);

/* ---------------------------------------------------------------- a18: Ewald */

/* ryoanji::computeGravityEwaldGpu (interface/ewald.cuh:25-27, ewald.cu:70-99) with an FP64 k-space and
 * real-space sum, one particle per lane. root_multipole = host double[8] (FP64 root quadrupole),
 * root_center = host double[3], limits = host float[6] (cubic), settings = host double[5]
 * {numReplicaShells, lCut, hCut, alpha_scale, small_R_scale_factor} (EwaldSettings, nbody/ewald.h:24-31).
 * Adds G m phi to ugrav (may be NULL) and G a to ax, ay, az for particles [first, last);
 * *total_pot (host, may be NULL) receives 0.5 G sum m phi. */
int ryo_ewald_f32(ryo_ctx* ctx, const double* root_multipole, const double* root_center, uint32_t first,
                  uint32_t last, const float* x, const float* y, const float* z, const float* m, const float* limits,
                  float G, const double* settings, float* ugrav, float* ax, float* ay, float* az, double* total_pot);

/* ---------------------------------------------------------------- a19: direct sum (accuracy reference) */

/* ryoanji::directSum (nbody/direct.cuh:88-107, traversal_cpu.hpp:232-268) with P2P<Exact> evaluated and
 * accumulated in FP64 for a list of targets (targets == NULL: targets 0..num_targets-1) over all n sources and
 * (2*num_shells+1)^3 images of box_l (host double[3]). out = double[num_targets][4] {G m_t phi_t, G a}. */
int ryo_direct_sum_f64(ryo_ctx* ctx, const float* x, const float* y, const float* z, const float* h,
                       const float* m, size_t n, const int64_t* targets, size_t num_targets, double G,
                       const double* box_l, int num_shells, double* out);

/* ryoanji::directSum<T> (interface/multipole_holder.cuh:92, nbody/direct.cuh:88-107) in the reference's form:
 * targets [first, last) against sources [0, num_bodies) and their (2 num_shells + 1)^3 images of box_l (host
 * T[3]); P2P<Exact> accumulated in FP64 and ADDED to p (m_i phi_i), ax, ay, az (G = 1, as the reference).
 * types 0: float arrays, 1: double arrays. */
int ryo_direct_sum_range(ryo_ctx* ctx, int types, size_t first, size_t last, size_t num_bodies, const void* box_l,
                         int num_shells, const void* x, const void* y, const void* z, const void* m, const void* h,
                         void* p, void* ax, void* ay, void* az);

/* The exact periodic potential and acceleration of the softened particles in a cubic box of side box_length,
 * in FP64, for a list of targets: the Ewald sum over all n sources and all their images (the convention of
 * ewald.hpp:201-327: neutralizing background), with pair images closer than h_i + h_j interacting by the
 * softened P2P<Exact> law as in the traversal. No reference counterpart: the reference has no periodic FP64
 * reference (global_forces_gpu.cpp:45-47); this is the accuracy oracle of traversal + Ewald correction.
 * alpha_scale = alpha L in (5, 16] (0: 6.5); the result does not depend on it beyond ~1e-12.
 * out = double[num_targets][4] {G m_t phi_t, G a}. */
int ryo_ewald_pair_sum_f64(ryo_ctx* ctx, const float* x, const float* y, const float* z, const float* h,
                           const float* m, size_t n, const int64_t* targets, size_t num_targets, double G,
                           double box_length, double alpha_scale, double* out);

/* Device memory for host callers that do not link the HIP runtime themselves (cstone::DeviceVector in
 * include/ryoanji_amd.hpp, cstone/cuda/device_vector.h:33-79): allocation, release (after the context's stream has
 * drained) and a copy in any direction (hipMemcpyDefault) that has finished on return. */
int ryo_malloc(ryo_ctx* ctx, void** ptr, size_t bytes);
int ryo_free(ryo_ctx* ctx, void* ptr);
int ryo_memcpy(ryo_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---------------------------------------------------------------- system pipeline */

/* One resident particle system: the ryoanji::TreeBuilder (interface/treebuilder.cuh:35-127) and
 * MultipoleHolder (interface/multipole_holder.cuh:28-89) facades in one object. It owns device buffers for
 * up to `capacity` particles and keeps keys, tree and multipoles on the device across calls. */
typedef struct ryo_system ryo_system;

typedef struct
{
    unsigned bucket_size; /* leaf capacity, 64 in the reference demos                   */
    float theta;          /* opening angle                                              */
    float G;              /* gravitational constant                                     */
    int group_size;       /* targets per traversal group: 64 (one wave) or 16 (parity)  */
    int num_shells;       /* periodic image shells in the traversal (0 = open)          */
    int ewald;            /* periodic boxes: 0 no Ewald correction; 1 the reference's correction
                           * (computeGravityEwald, ewald.hpp:356-389); 2 the same correction with its k-space
                           * potential consistent with its k-space acceleration (ewald.hpp:320 subtracts that
                           * potential term with the opposite sign; accelerations are identical to 1) */
    double ewald_settings[5];
    int init_node;        /* 0: traversal starts at the root, 1: at its children       */
    int soft_mac;         /* 0: the reference's MAC (macs.hpp:64-82). 1: also open every cell nearer than
                           * its radius plus twice the largest softening, so that no softened pair is
                           * approximated by a multipole (resident systems; costs a softening-range read-back) */
    float group_split_tol; /* resident systems: 0 fixed groups of group_size (computeFixedGroups, as the
                            * single-GPU demo, ryoanji/test/demo.cu:180); > 0 spatial groups, computeGroupSplits
                            * with this tolerance factor, the particles' h and extent ratio 8, as
                            * MultipoleHolder::computeSpatialGroups (multipole_holder.cu:36-53, factor 2);
                            * needs group_size 64 (one group-count read-back) */
    int multipole;         /* resident systems: 0 CartesianQuadrupole, 1 CartesianMDQpole (cartesian_qpole.hpp:237-417:
                            * 12 values per node, ryo_system_tree's multipoles hold 12 per node); MDQpole systems run
                            * without the Ewald correction (the reference instantiates computeGravityEwaldGpu for
                            * CartesianQuadrupole only, ewald.cu:107-110) and without soft_mac */
} ryo_params;

void ryo_default_params(ryo_params* p);
int ryo_system_create(ryo_ctx* ctx, size_t capacity, ryo_system** sys);
int ryo_system_destroy(ryo_system* sys);
/* device pointers of the system's particle arrays (x, y, z, m, h, vx, vy, vz), in its current (SFC) order */
int ryo_system_arrays(ryo_system* sys, float** xyzmh_v /* host float*[8] */);
/* copy n particles in (device or host memory, hipMemcpyDefault) */
int ryo_system_load(ryo_system* sys, size_t n, const float* x, const float* y, const float* z, const float* m,
                    const float* h, const float* limits /* host float[6] */, int periodic);
/* copy the system's particles (current SFC order) into caller device arrays; any pointer may be NULL */
int ryo_system_store(ryo_system* sys, float* x, float* y, float* z, float* m, float* h);
/* keys -> sort -> permute (x,y,z,m,h and velocities) -> leaves -> linked octree -> FP64 upsweep */
int ryo_system_build(ryo_system* sys, const ryo_params* params);
/* traversal (+ Ewald); outputs in the system's SFC order, written (not added) to the system's p, ax, ay, az */
int ryo_system_gravity(ryo_system* sys, const ryo_params* params, double* total_pot);
/* device pointers of the results {p, ax, ay, az} and the original index of each particle (int32) */
int ryo_system_results(ryo_system* sys, float** p_ax_ay_az /* host float*[4] */, int32_t** order);
/* sizes: {n, numLeaves, numNodes, numGroups} */
int ryo_system_info(ryo_system* sys, int64_t* info);
/* tree arrays (device pointers) of the last build, for parity checks:
 * {keys(u64), leaves(u64), counts(u32), layout(u32), prefixes(u64), childOffsets, parents, levelRange,
 *  internalToLeaf, leafToInternal, centers(f32x4), multipoles(f32x8), centers64, multipoles64} */
int ryo_system_tree(ryo_system* sys, void** ptrs /* host void*[14] */);
/* `steps` kick-drift-kick leapfrog steps of size dt (velocities live in the system, zero after load); each step
 * rebuilds the tree (open systems first refit a cubic box around the particles) and solves gravity; the first
 * call also solves gravity for the initial positions. *total_pot (host, may be NULL): the last potential. */
int ryo_system_leapfrog(ryo_system* sys, const ryo_params* params, float dt, int steps, double* total_pot);
/* the system's current box (host float[6]) */
int ryo_system_limits(ryo_system* sys, float* limits);
/* phase times (ms) of the last build+gravity: {keys, sort, permute, leaves, link, upsweep, traverse, ewald} */
int ryo_system_phase_ms(ryo_system* sys, double* ms);

/* Accuracy of the last ryo_system_gravity of `sys` against the FP64 direct sum (P2P<Exact>,
 * traversal_cpu.hpp:232-268, the reference's accuracy oracle) on num_targets targets drawn from the system
 * (splitmix64 of seed, with replacement). For a periodic system the reference is the direct sum over the same
 * replica images when the last gravity had no Ewald term, and the FP64 Ewald pair sum
 * (ryo_ewald_pair_sum_f64, cubic boxes) when it had. out[4] = {median, 99th percentile, maximum} of
 * |a - a_ref| / |a_ref| and the relative error of the sampled potential sum. Synchronizes the stream. */
int ryo_system_accuracy(ryo_system* sys, int num_targets, uint64_t seed, double* out);

/* ---------------------------------------------------------------- batches of systems */

/* Many particle systems packed one after the other and solved together: every stage (keys, sort, leaves,
 * linked octree, FP64 upsweep, groups, traversal) runs once for all of them, driven by per-system offset
 * tables. Each system gets exactly the keys, tree, multipoles and forces it gets when solved on its own
 * (GroupView offsets, cstone/traversal/groups.hpp:28-34, are the reference's per-batch table; the reference
 * solves one system per call). Systems are open unless ryo_batch_set_periodic marks them periodic. */
typedef struct ryo_batch ryo_batch;
/* max_systems at most 65536 (RYO_ERR_UNSUPPORTED beyond: each system is a segment of the batch's sorts) */
int ryo_batch_create(ryo_ctx* ctx, size_t max_particles, int max_systems, ryo_batch** batch);
int ryo_batch_destroy(ryo_batch* batch);
/* counts[S] particles per system (each >= 1); x, y, z, m, h: the systems packed one after the other (device or
 * host memory, copied); limits[6 S] the box of each system (cubic or not) */
int ryo_batch_load(ryo_batch* batch, int num_systems, const uint32_t* counts, const float* x, const float* y,
                   const float* z, const float* m, const float* h, const float* limits);
/* periodic[S] (host) for the loaded systems: nonzero marks a periodic system, solved as a resident periodic
 * system is (traversal over the (2 prm.num_shells + 1)^3 images of its box, then, with prm.ewald 1 or 2, the FP64
 * Ewald correction of ryo_ewald_f32 from its root multipole; a cubic box for the Ewald correction); every
 * system is open after ryo_batch_load. (traversal_gpu.cu:477-518 traverses the images of any call.) */
int ryo_batch_set_periodic(ryo_batch* batch, const int* periodic);
/* one full pass for every system; theta[S] and G[S] per system (host; NULL: prm's value), pot[S] (host, may be
 * NULL) receives each system's 0.5 G sum(m phi) (with its Ewald correction's for a periodic system). prm:
 * bucket_size and group_size (<= 64); num_shells, ewald and ewald_settings for the periodic systems. The first
 * solve of a batch object reads the leaf counts back to size the node arrays exactly and sets capacities (its
 * totals plus a margin); later solves launch every stage for the capacities and read the overflow flag, tables,
 * statistics and potentials back once at the end, solving again exactly (and growing the capacities) when the
 * trees outgrew them. The results are the same either way, bit for bit. */
int ryo_batch_solve(ryo_batch* batch, const ryo_params* prm, const float* theta, const float* G, double* pot);
/* The same solve as ryo_batch_solve, bit for bit, run as one captured HIP graph (keys -> sort -> leaves ->
 * octrees -> upsweep -> groups -> traversal, no host round trip inside). The graph is captured on the first call and again when the system
 * counts, bucket or group size, init node, traversal shape or a buffer changed; positions, boxes, theta and G
 * may change between replays. Its node arrays hold the capturing solve's totals plus a margin; a replay whose
 * trees outgrow them is detected on the device and solved again exactly (counted in the statistics). The
 * batch's stream must not be capturing. Synchronizes. Open systems only (RYO_ERR_UNSUPPORTED with a periodic
 * one: the Ewald tables are built on the host from the solved root multipoles). */
int ryo_batch_solve_graph(ryo_batch* batch, const ryo_params* prm, const float* theta, const float* G,
                          double* pot);
/* out[4] = {graphs captured, replays, replays solved again because the trees outgrew the graph's arrays,
 * leaf-slot capacity of the node arrays (the current graph's; ryo_batch_solve's after its first call)} */
int ryo_batch_graph_stats(ryo_batch* batch, uint64_t* out);
/* device arrays, packed like the input but each system in its SFC order: out[0..4] = x, y, z, m, h,
 * out[5..8] = p, ax, ay, az; *order = index into the loaded (packed) arrays */
int ryo_batch_results(ryo_batch* batch, float** out, int32_t** order);
/* host copies of the offset tables (num_systems + 1 entries each; NULL skips): particles P, leaf slots LO (each
 * system's L leaves and its terminator), nodes NO, parents PO, groups GO */
int ryo_batch_offsets(ryo_batch* batch, uint32_t* P, uint32_t* LO, uint32_t* NO, uint32_t* PO, uint32_t* GO);
/* device pointers of the trees, indices local to each system's slice: {sorted keys, leaves, layout, counts,
 * prefixes, childOffsets, parents, internalToLeaf, leafToInternal, levelRange (23 per system), centers,
 * multipoles, centers64, multipoles64} */
int ryo_batch_tree(ryo_batch* batch, void** arrays);

/* ---------------------------------------------------------------- serving (no reference counterpart) */

/* A served request: a particle system (host or device arrays) that runs `steps` leapfrog steps. */
typedef struct
{
    size_t n;
    const float *x, *y, *z, *m, *h;
    float limits[6];
    int periodic;
    float theta, G, dt;
    int steps;
    unsigned bucket_size; /* 0: 64 */
    int group_size;       /* 0: 64 */
    double arrival_ms;    /* relative to the start of ryo_server_run; requests sorted by arrival */
    int accuracy_targets; /* > 0: check the final accelerations against the FP64 direct sum on this many targets */
    int num_shells;       /* periodic requests: replica shells the traversal walks (ryo_params.num_shells) */
    int ewald;            /* periodic requests: ryo_params.ewald (0 none, 1 the reference's correction, 2 with
                             the consistent potential; a cubic box) */
} ryo_request;

typedef struct
{
    double arrival_ms, admit_ms, finish_ms, total_pot;
    int status, slot;
    double err_median, err_p99, err_max; /* ryo_system_accuracy of the final state, when requested (else 0) */
    float limits[6];                      /* the box of the final gravity solve (refitted at every step)    */
} ryo_result;

typedef struct ryo_server ryo_server;

/* `slots` concurrent workers (each: a HIP stream, a context, a resident system of capacity max_n particles) and
 * a batch lane for small requests (below); requests are admitted while their summed HBM footprint stays within
 * hbm_budget bytes (0: unlimited). */
int ryo_server_create(int slots, size_t hbm_budget, size_t max_n, ryo_server** server);
/* The batch lane: requests of at most max_n particles (default 2^16; 0 turns the lane off) with the default bucket
 * and group size, open or periodic (a periodic one with its own image shells and Ewald mode), are packed one after
 * the other into one ryo_batch (capacity particles and max_systems systems in all, defaults 2^22 and 1024) and
 * advanced in lockstep, one batched solve per leapfrog step of every member; requests join at the next step and
 * leave when their steps are done. A member's trajectory is the one it has on a slot (same kick-drift arithmetic,
 * periodic wrap or box refit, and solve). Results of lane members report slot = `slots`. Set before
 * ryo_server_run. */
int ryo_server_set_batching(ryo_server* server, size_t max_n, size_t capacity, int max_systems);
/* Admission order of waiting requests (each still first fit: a request that does not fit the free places or the
 * HBM budget lets later ones in). order 0 (default): arrival. order 1: least work first, work = n x (steps + 1)
 * solves; with max_wait_ms > 0 requests that have waited that long go ahead of the others, in arrival order, so
 * none waits forever behind smaller ones. The order changes when requests run, not what they compute. Set before
 * ryo_server_run. */
int ryo_server_set_order(ryo_server* server, int order, double max_wait_ms);
int ryo_server_destroy(ryo_server* server);
/* estimated HBM bytes of a resident system of n particles (the admission weight) */
size_t ryo_server_footprint(size_t n);
/* run the request stream to completion (continuous batching); out_state may be NULL, else out_state[i] NULL or a
 * host buffer of 7*n floats for request i: x, y, z, ax, ay, az and the original index (int32), in SFC order */
int ryo_server_run(ryo_server* server, int num, const ryo_request* requests, ryo_result* results, float** out_state);

#ifdef __cplusplus
}
#endif

#endif /* RYOANJI_AMD_H */
