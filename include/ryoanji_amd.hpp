/*
 * ryoanji_amd.hpp -- C++ entry points with the reference's names and argument meaning, over the C ABI.
 *
 * Host-only, header-only. A caller written against cstone/ryoanji's GPU functions finds the same names,
 * namespaces and argument order here (the reference's own types are mirrored minimally: Box<float>,
 * GroupView, OctreeView, execution::Gpu, Vec3/Vec4, CartesianQuadrupole<float>). Errors are reported as
 * std::runtime_error (the reference throws for configuration errors and exits on GPU errors; this layer
 * never exits). The context (stream, scratch, statistics) is the thread's current Session.
 *
 * Paths cited are relative to the reference checkout.
 */
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "ryoanji_amd.h"

namespace ryoanji_amd
{
//! the calling thread's C-ABI context (stream, scratch, statistics); created on first use on the null stream
inline ryo_ctx*& threadContext()
{
    thread_local ryo_ctx* c = nullptr;
    return c;
}

inline void check(ryo_ctx* c, int rc)
{
    if (rc != RYO_OK) throw std::runtime_error(std::string("ryoanji_amd: ") + (c ? ryo_last_error(c) : "error"));
}

inline ryo_ctx* ctx()
{
    ryo_ctx*& c = threadContext();
    if (!c) check(nullptr, ryo_ctx_create(&c, nullptr));
    return c;
}

//! run this thread's calls on `hipStream` (a hipStream_t)
inline void bindStream(void* hipStream)
{
    ryo_ctx*& c = threadContext();
    if (c) ryo_ctx_destroy(c), c = nullptr;
    check(nullptr, ryo_ctx_create(&c, hipStream));
}

inline void check(int rc) { check(ctx(), rc); }
} // namespace ryoanji_amd

namespace cstone
{
using TreeNodeIndex = int;      // cstone/tree/definitions.h:25
using LocalIndex    = unsigned; // cstone/tree/definitions.h:27

template<class T>
using Vec3 = std::array<T, 3>;
template<class T>
using Vec4 = std::array<T, 4>;

namespace execution
{
//! a HIP stream (cstone/execution.hpp:37-56); the C-ABI context of the calling thread runs the work
struct Gpu
{
    void* stream = nullptr;
};
inline constexpr Gpu gpuDefaultStream{};
} // namespace execution

enum class BoundaryType : uint8_t
{
    open     = 0,
    periodic = 1
};

//! cstone/sfc/box.hpp:99-214 (cubic boxes)
template<class T>
class Box
{
public:
    Box(T lo, T hi, BoundaryType b = BoundaryType::open)
        : lim_{lo, hi, lo, hi, lo, hi}
        , b_(b)
    {
    }
    Box(T xmin, T xmax, T ymin, T ymax, T zmin, T zmax, BoundaryType b = BoundaryType::open)
        : lim_{xmin, xmax, ymin, ymax, zmin, zmax}
        , b_(b)
    {
    }
    T xmin() const { return lim_[0]; }
    T xmax() const { return lim_[1]; }
    T ymin() const { return lim_[2]; }
    T ymax() const { return lim_[3]; }
    T zmin() const { return lim_[4]; }
    T zmax() const { return lim_[5]; }
    T lx() const { return lim_[1] - lim_[0]; }
    T ly() const { return lim_[3] - lim_[2]; }
    T lz() const { return lim_[5] - lim_[4]; }
    BoundaryType boundaryX() const { return b_; }
    const float* limits() const
    {
        for (int k = 0; k < 6; ++k)
            f_[k] = float(lim_[k]);
        return f_;
    }
    //! the limits in the box's own type (host T[6]), for the double entry points
    const T* limitsT() const { return lim_; }

private:
    T lim_[6];
    BoundaryType b_;
    mutable float f_[6];
};

//! cstone/traversal/groups.hpp:28-34
struct GroupView
{
    LocalIndex firstBody, lastBody;
    LocalIndex numGroups;
    const LocalIndex* groupStart;
    const LocalIndex* groupEnd;
};

//! cstone/tree/octree.hpp:234-260 (device pointers; levelRange on the host, d_levelRange on the device)
template<class KeyType>
struct OctreeView
{
    TreeNodeIndex numLeafNodes, numInternalNodes, numNodes, numParents;
    KeyType* prefixes;
    TreeNodeIndex* childOffsets;
    TreeNodeIndex* parents;
    TreeNodeIndex* levelRange;
    TreeNodeIndex* d_levelRange;
    TreeNodeIndex* internalToLeaf;
    TreeNodeIndex* leafToInternal;
};

//! cstone::computeSfcKeys -- cstone/sfc/sfc_gpu.h:37 (Hilbert keys, 64-bit)
inline void computeSfcKeys(execution::Gpu, const float* x, const float* y, const float* z, uint64_t* keys, size_t n,
                           const Box<float>& box)
{
    ryoanji_amd::check(ryo_sfc_keys_f32(ryoanji_amd::ctx(), x, y, z, keys, n, box.limits(), 0, 1));
}

//! cstone::computeSfcKeys with 32-bit Hilbert keys (10 levels) for float and double boxes (sfc_gpu.cu:46-62)
inline void computeSfcKeys(execution::Gpu, const float* x, const float* y, const float* z, uint32_t* keys, size_t n,
                           const Box<float>& box)
{
    ryoanji_amd::check(ryo_sfc_keys_u32_f32(ryoanji_amd::ctx(), x, y, z, keys, n, box.limitsT(), 0, 1));
}
inline void computeSfcKeys(execution::Gpu, const double* x, const double* y, const double* z, uint32_t* keys,
                           size_t n, const Box<double>& box)
{
    ryoanji_amd::check(ryo_sfc_keys_u32_f64(ryoanji_amd::ctx(), x, y, z, keys, n, box.limitsT(), 0, 1));
}

//! cstone::computeSfcKeys for double coordinates and Box<double> (sfc_gpu.cu:46-62)
inline void computeSfcKeys(execution::Gpu, const double* x, const double* y, const double* z, uint64_t* keys,
                           size_t n, const Box<double>& box)
{
    ryoanji_amd::check(ryo_sfc_keys_f64(ryoanji_amd::ctx(), x, y, z, keys, n, box.limitsT(), 0, 1));
}

//! cstone::sortByKeyTempStorage -- cstone/primitives/primitives_gpu.h:120-124
template<class K, class V>
inline uint64_t sortByKeyTempStorage(uint64_t n)
{
    static_assert(sizeof(K) == 8 && sizeof(V) == 4, "(uint64 key, int32 value) pairs");
    return ryo_sort_temp_bytes(n);
}

//! cstone::sortByKey -- cstone/primitives/primitives_gpu.h:126-130
inline void sortByKey(execution::Gpu, uint64_t* first, uint64_t* last, int* values, uint64_t* keyBuf, int* valueBuf,
                      void* tmp, uint64_t tmpBytes)
{
    ryoanji_amd::check(ryo_sort_pairs_u64_i32(ryoanji_amd::ctx(), first, values, size_t(last - first), keyBuf,
                                              valueBuf, tmp, tmpBytes));
}

//! cstone::sequence -- primitives_gpu.cu:265-273
inline void sequence(execution::Gpu, int* out, size_t n, int start)
{
    ryoanji_amd::check(ryo_sequence_i32(ryoanji_amd::ctx(), out, n, start));
}

/*! converged cornerstone leaves + counts (cstone::computeOctree / updateOctreeGpu to convergence,
 *  csarray.hpp:429-440, update_gpu.cuh:43-68) and TreeBuilder's layout; returns the number of leaves */
inline int computeOctreeGpu(execution::Gpu, const uint64_t* sortedKeys, size_t n, unsigned bucketSize,
                            uint64_t* leaves, unsigned* counts, LocalIndex* layout, size_t capacity)
{
    int L = 0;
    ryoanji_amd::check(
        ryo_build_leaves_u64(ryoanji_amd::ctx(), sortedKeys, n, bucketSize, leaves, counts, layout, capacity, &L));
    return L;
}

//! cstone::buildOctreeGpu(exec, leaves, OctreeView) -- cstone/tree/octree_gpu.h:35-45
inline void buildOctreeGpu(execution::Gpu, const uint64_t* leaves, OctreeView<uint64_t> o)
{
    ryoanji_amd::check(ryo_link_octree_u64(ryoanji_amd::ctx(), leaves, o.numLeafNodes, o.prefixes, o.childOffsets,
                                           o.parents, o.d_levelRange, o.internalToLeaf, o.leafToInternal,
                                           o.levelRange));
}

//! cstone::computeFixedGroups -- cstone/traversal/groups_gpu.cu:45-58 (groups has numGroups + 1 entries)
inline GroupView computeFixedGroups(execution::Gpu, LocalIndex first, LocalIndex last, unsigned groupSize,
                                    LocalIndex* groups)
{
    int G = 0;
    ryoanji_amd::check(ryo_fixed_groups(ryoanji_amd::ctx(), first, last, groupSize, groups, &G));
    return GroupView{first, last, LocalIndex(G), groups, groups + 1};
}

/*! cstone::computeGroupSplits -- cstone/traversal/groups_gpu.cu:121-151 for (Tc, T) = (float, float): target
 *  groups of up to groupSize (64) SFC-ordered particles, split where they are not compact. groups must hold
 *  up to last - first + 1 entries (`capacity`); extentRatio as CSTONE_GROUP_EXTENT_SPLIT (default 8). */
inline GroupView computeGroupSplits(execution::Gpu, LocalIndex first, LocalIndex last, const float* x, const float* y,
                                    const float* z, const float* h, const uint64_t* leaves, TreeNodeIndex numLeaves,
                                    const LocalIndex* layout, const Box<float>& box, unsigned groupSize,
                                    float tolFactor, LocalIndex* groups, size_t capacity, float extentRatio = 8.0f)
{
    int G = 0;
    ryoanji_amd::check(ryo_group_splits_f32(ryoanji_amd::ctx(), first, last, x, y, z, h, leaves, int(numLeaves),
                                            layout, box.limits(), groupSize, tolFactor, extentRatio, groups,
                                            capacity, &G));
    return GroupView{first, last, LocalIndex(G), groups, groups + 1};
}
} // namespace cstone

namespace ryoanji
{
using cstone::GroupView;
using cstone::LocalIndex;
using cstone::TreeNodeIndex;
using cstone::Vec3;
using cstone::Vec4;

template<class T>
using CartesianQuadrupole = std::array<T, 8>; // nbody/cartesian_qpole.hpp:27-51
template<class T>
using CartesianMDQpole = std::array<T, 12>; // nbody/cartesian_qpole.hpp:31-32

/*! fused FP64 replacement of computeLeafSourceCenterGpu + upsweepCentersGpu + computeLeafMultipoles +
 *  upsweepMultipoles + setMacGpu (upsweep_gpu.h:25,40; source_center_gpu.h:66-116). leafToInternal is the leaf
 *  part of the map (numLeaves entries), as TreeBuilder::leafToInternal() and computeLeafMultipoles take it. */
inline void upsweep(const float* x, const float* y, const float* z, const float* m, const LocalIndex* layout,
                    int numLeaves, const uint64_t* prefixes, const TreeNodeIndex* childOffsets,
                    const TreeNodeIndex* parents, const TreeNodeIndex* leafToInternal, const TreeNodeIndex* d_levelRange,
                    const cstone::Box<float>& box, float theta, Vec4<float>* centers,
                    CartesianQuadrupole<float>* multipoles)
{
    const TreeNodeIndex* full = leafToInternal - (numLeaves - 1) / 7;
    ryoanji_amd::check(ryo_upsweep_f32(ryoanji_amd::ctx(), x, y, z, m, layout, numLeaves, prefixes, childOffsets,
                                       parents, full, d_levelRange, box.limits(), 1.0f / theta,
                                       reinterpret_cast<float*>(centers), reinterpret_cast<float*>(multipoles),
                                       nullptr, nullptr));
}

//! the same fused FP64 upsweep for double particles and Box<double>: SourceCenterType<double> centers and
//! CartesianQuadrupole<double> multipoles (upsweep_gpu.cu:134-143)
inline void upsweep(const double* x, const double* y, const double* z, const double* m, const LocalIndex* layout,
                    int numLeaves, const uint64_t* prefixes, const TreeNodeIndex* childOffsets,
                    const TreeNodeIndex* parents, const TreeNodeIndex* leafToInternal, const TreeNodeIndex* d_levelRange,
                    const cstone::Box<double>& box, float theta, Vec4<double>* centers,
                    CartesianQuadrupole<double>* multipoles)
{
    const TreeNodeIndex* full = leafToInternal - (numLeaves - 1) / 7;
    ryoanji_amd::check(ryo_upsweep_f64(ryoanji_amd::ctx(), x, y, z, m, layout, numLeaves, prefixes, childOffsets,
                                       parents, full, d_levelRange, box.limitsT(), 1.0f / theta,
                                       reinterpret_cast<double*>(centers), reinterpret_cast<double*>(multipoles),
                                       nullptr));
}

//! the fused upsweep with CartesianMDQpole<float> multipoles (P2M / M2M for CartesianMDQpole)
inline void upsweep(const float* x, const float* y, const float* z, const float* m, const LocalIndex* layout,
                    int numLeaves, const uint64_t* prefixes, const TreeNodeIndex* childOffsets,
                    const TreeNodeIndex* parents, const TreeNodeIndex* leafToInternal, const TreeNodeIndex* d_levelRange,
                    const cstone::Box<float>& box, float theta, Vec4<float>* centers,
                    CartesianMDQpole<float>* multipoles)
{
    const TreeNodeIndex* full = leafToInternal - (numLeaves - 1) / 7;
    ryoanji_amd::check(ryo_upsweep_mdq_f32(ryoanji_amd::ctx(), x, y, z, m, layout, numLeaves, prefixes, childOffsets,
                                       parents, full, d_levelRange, box.limits(), 1.0f / theta,
                                       reinterpret_cast<float*>(centers), reinterpret_cast<float*>(multipoles),
                                       nullptr, nullptr));
}

//! the FP64 upsweep for double particles with CartesianMDQpole<double> multipoles
inline void upsweep(const double* x, const double* y, const double* z, const double* m, const LocalIndex* layout,
                    int numLeaves, const uint64_t* prefixes, const TreeNodeIndex* childOffsets,
                    const TreeNodeIndex* parents, const TreeNodeIndex* leafToInternal, const TreeNodeIndex* d_levelRange,
                    const cstone::Box<double>& box, float theta, Vec4<double>* centers,
                    CartesianMDQpole<double>* multipoles)
{
    const TreeNodeIndex* full = leafToInternal - (numLeaves - 1) / 7;
    ryoanji_amd::check(ryo_upsweep_mdq_f64(ryoanji_amd::ctx(), x, y, z, m, layout, numLeaves, prefixes, childOffsets,
                                       parents, full, d_levelRange, box.limitsT(), 1.0f / theta,
                                       reinterpret_cast<double*>(centers), reinterpret_cast<double*>(multipoles),
                                       nullptr));
}

//! ryoanji::traverse -- nbody/traversal_gpu.h:51-56, float instantiation; gmPool is not needed (LDS stacks)
inline double traverse(GroupView grp, int initNodeIdx, const float* xt, const float* yt, const float* zt,
                       const float* mt, const float* ht, const float* xs, const float* ys, const float* zs,
                       const float* ms, const float* hs, const TreeNodeIndex* childOffsets,
                       const TreeNodeIndex* internalToLeaf, const LocalIndex* layout, const Vec4<float>* sourceCenter,
                       const CartesianQuadrupole<float>* multipoles, float G, int numShells, Vec3<float> boxL,
                       float* p, float* ax, float* ay, float* az, int* /*gmPool*/ = nullptr)
{
    double pot = 0;
    ryoanji_amd::check(ryo_traverse_f32(
        ryoanji_amd::ctx(), int(grp.numGroups), grp.groupStart, grp.groupEnd, initNodeIdx, xt, yt, zt, mt, ht, xs, ys,
        zs, ms, hs, childOffsets, internalToLeaf, layout, reinterpret_cast<const float*>(sourceCenter),
        reinterpret_cast<const float*>(multipoles), G, numShells, boxL.data(), p, ax, ay, az, &pot));
    return pot;
}

//! ryoanji::traverse, (Tc,Th,Tm,Ta,Tf) = (d,d,d,d,d) with CartesianQuadrupole<double> (traversal_gpu.cu:635-641)
inline double traverse(GroupView grp, int initNodeIdx, const double* xt, const double* yt, const double* zt,
                       const double* mt, const double* ht, const double* xs, const double* ys, const double* zs,
                       const double* ms, const double* hs, const TreeNodeIndex* childOffsets,
                       const TreeNodeIndex* internalToLeaf, const LocalIndex* layout, const Vec4<double>* sourceCenter,
                       const CartesianQuadrupole<double>* multipoles, double G, int numShells, Vec3<double> boxL,
                       double* p, double* ax, double* ay, double* az, int* /*gmPool*/ = nullptr)
{
    double pot = 0;
    ryoanji_amd::check(ryo_traverse_f64(
        ryoanji_amd::ctx(), int(grp.numGroups), grp.groupStart, grp.groupEnd, initNodeIdx, xt, yt, zt, mt, ht, xs, ys,
        zs, ms, hs, childOffsets, internalToLeaf, layout, reinterpret_cast<const double*>(sourceCenter),
        reinterpret_cast<const double*>(multipoles), G, numShells, boxL.data(), p, ax, ay, az, &pot));
    return pot;
}

//! ryoanji::traverse, (Tc,Th,Tm,Ta,Tf) = (d,f,f,f,d) with CartesianQuadrupole<float> (traversal_gpu.cu:635-641)
inline double traverse(GroupView grp, int initNodeIdx, const double* xt, const double* yt, const double* zt,
                       const float* mt, const float* ht, const double* xs, const double* ys, const double* zs,
                       const float* ms, const float* hs, const TreeNodeIndex* childOffsets,
                       const TreeNodeIndex* internalToLeaf, const LocalIndex* layout, const Vec4<double>* sourceCenter,
                       const CartesianQuadrupole<float>* multipoles, double G, int numShells, Vec3<double> boxL,
                       float* p, float* ax, float* ay, float* az, int* /*gmPool*/ = nullptr)
{
    double pot = 0;
    ryoanji_amd::check(ryo_traverse_f64_f32(
        ryoanji_amd::ctx(), int(grp.numGroups), grp.groupStart, grp.groupEnd, initNodeIdx, xt, yt, zt, mt, ht, xs, ys,
        zs, ms, hs, childOffsets, internalToLeaf, layout, reinterpret_cast<const double*>(sourceCenter),
        reinterpret_cast<const float*>(multipoles), G, numShells, boxL.data(), p, ax, ay, az, &pot));
    return pot;
}

//! ryoanji::traverse with CartesianMDQpole multipoles, for each of the three (Tc,Th,Tm,Ta,Tf) instantiations
inline double traverse(GroupView grp, int initNodeIdx, const float* xt, const float* yt, const float* zt,
                       const float* mt, const float* ht, const float* xs, const float* ys, const float* zs,
                       const float* ms, const float* hs, const TreeNodeIndex* childOffsets,
                       const TreeNodeIndex* internalToLeaf, const LocalIndex* layout, const Vec4<float>* sourceCenter,
                       const CartesianMDQpole<float>* multipoles, float G, int numShells, Vec3<float> boxL,
                       float* p, float* ax, float* ay, float* az, int* /*gmPool*/ = nullptr)
{
    double pot = 0;
    ryoanji_amd::check(ryo_traverse_mdq_f32(
        ryoanji_amd::ctx(), int(grp.numGroups), grp.groupStart, grp.groupEnd, initNodeIdx, xt, yt, zt, mt, ht, xs, ys,
        zs, ms, hs, childOffsets, internalToLeaf, layout, reinterpret_cast<const float*>(sourceCenter),
        reinterpret_cast<const float*>(multipoles), G, numShells, boxL.data(), p, ax, ay, az, &pot));
    return pot;
}

inline double traverse(GroupView grp, int initNodeIdx, const double* xt, const double* yt, const double* zt,
                       const double* mt, const double* ht, const double* xs, const double* ys, const double* zs,
                       const double* ms, const double* hs, const TreeNodeIndex* childOffsets,
                       const TreeNodeIndex* internalToLeaf, const LocalIndex* layout, const Vec4<double>* sourceCenter,
                       const CartesianMDQpole<double>* multipoles, double G, int numShells, Vec3<double> boxL,
                       double* p, double* ax, double* ay, double* az, int* /*gmPool*/ = nullptr)
{
    double pot = 0;
    ryoanji_amd::check(ryo_traverse_mdq_f64(
        ryoanji_amd::ctx(), int(grp.numGroups), grp.groupStart, grp.groupEnd, initNodeIdx, xt, yt, zt, mt, ht, xs, ys,
        zs, ms, hs, childOffsets, internalToLeaf, layout, reinterpret_cast<const double*>(sourceCenter),
        reinterpret_cast<const double*>(multipoles), G, numShells, boxL.data(), p, ax, ay, az, &pot));
    return pot;
}

inline double traverse(GroupView grp, int initNodeIdx, const double* xt, const double* yt, const double* zt,
                       const float* mt, const float* ht, const double* xs, const double* ys, const double* zs,
                       const float* ms, const float* hs, const TreeNodeIndex* childOffsets,
                       const TreeNodeIndex* internalToLeaf, const LocalIndex* layout, const Vec4<double>* sourceCenter,
                       const CartesianMDQpole<float>* multipoles, double G, int numShells, Vec3<double> boxL,
                       float* p, float* ax, float* ay, float* az, int* /*gmPool*/ = nullptr)
{
    double pot = 0;
    ryoanji_amd::check(ryo_traverse_mdq_f64_f32(
        ryoanji_amd::ctx(), int(grp.numGroups), grp.groupStart, grp.groupEnd, initNodeIdx, xt, yt, zt, mt, ht, xs, ys,
        zs, ms, hs, childOffsets, internalToLeaf, layout, reinterpret_cast<const double*>(sourceCenter),
        reinterpret_cast<const float*>(multipoles), G, numShells, boxL.data(), p, ax, ay, az, &pot));
    return pot;
}

//! ryoanji::bhMaxTargetSize -- traversal_gpu.cu:643 (one wavefront of targets)
inline int bhMaxTargetSize() { return 64; }

//! ryoanji::readBhStats -- traversal_gpu.cu:688-700
inline std::array<uint64_t, 5> readBhStats()
{
    std::array<uint64_t, 5> s{};
    ryoanji_amd::check(ryo_read_stats(ryoanji_amd::ctx(), s.data()));
    return s;
}

//! ryoanji::readTraversalRetries -- traversal_gpu.cu:645
inline uint64_t readTraversalRetries() { return ryo_traversal_retries(ryoanji_amd::ctx()); }

//! EwaldSettings -- nbody/ewald.h:24-31
struct EwaldSettings
{
    int numReplicaShells        = 1;
    double lCut                 = 2.6;
    double hCut                 = 2.8;
    double alpha_scale          = 2.0;
    double small_R_scale_factor = 3.0e-3;
};

//! ryoanji::computeGravityEwaldGpu -- interface/ewald.cuh:25-27 (FP64 root multipole and center)
inline void computeGravityEwaldGpu(const Vec3<double>& rootCenter, const CartesianQuadrupole<double>& Mroot,
                                   GroupView grp, const float* x, const float* y, const float* z, const float* m,
                                   const cstone::Box<float>& box, float G, float* ugrav, float* ax, float* ay,
                                   float* az, double* ugravTot, EwaldSettings s = {})
{
    double set[5] = {double(s.numReplicaShells), s.lCut, s.hCut, s.alpha_scale, s.small_R_scale_factor};
    double pot    = 0;
    ryoanji_amd::check(ryo_ewald_f32(ryoanji_amd::ctx(), Mroot.data(), rootCenter.data(), grp.firstBody, grp.lastBody,
                                     x, y, z, m, box.limits(), G, set, ugrav, ax, ay, az, &pot));
    if (ugravTot) *ugravTot += pot;
}

/*! ryoanji::TreeBuilder -- interface/treebuilder.cuh:35-127: keys, sort, permutation, leaves, linked tree and
 *  layout on the device. Unlike the reference (Q1), m and h are permuted together with x, y, z. */
class TreeBuilder
{
public:
    explicit TreeBuilder(unsigned ncrit)
        : bucket_(ncrit)
    {
    }
    ~TreeBuilder()
    {
        if (sys_) ryo_system_destroy(sys_);
    }
    TreeBuilder(const TreeBuilder&) = delete;
    TreeBuilder& operator=(const TreeBuilder&) = delete;

    //! builds the octree of n bodies (device arrays, rewritten in SFC order); returns the number of tree nodes
    int update(float* x, float* y, float* z, float* m, float* h, size_t n, const cstone::Box<float>& box)
    {
        if (!sys_ || cap_ < n)
        {
            if (sys_) ryo_system_destroy(sys_);
            ryoanji_amd::check(ryo_system_create(ryoanji_amd::ctx(), n, &sys_));
            cap_ = n;
        }
        ryoanji_amd::check(ryo_system_load(sys_, n, x, y, z, m, h, box.limits(), 0));
        ryo_params p;
        ryo_default_params(&p);
        p.bucket_size = bucket_;
        ryoanji_amd::check(ryo_system_build(sys_, &p));
        ryoanji_amd::check(ryo_system_store(sys_, x, y, z, m, h));
        ryoanji_amd::check(ryo_system_info(sys_, info_));
        ryoanji_amd::check(ryo_system_tree(sys_, tree_));
        return int(info_[2]);
    }
    const LocalIndex* layout() const { return static_cast<const LocalIndex*>(tree_[3]); }
    const uint64_t* nodeKeys() const { return static_cast<const uint64_t*>(tree_[4]); }
    const TreeNodeIndex* childOffsets() const { return static_cast<const TreeNodeIndex*>(tree_[5]); }
    const TreeNodeIndex* parents() const { return static_cast<const TreeNodeIndex*>(tree_[6]); }
    const TreeNodeIndex* internalToLeaf() const { return static_cast<const TreeNodeIndex*>(tree_[8]); }
    const TreeNodeIndex* d_levelRange() const { return static_cast<const TreeNodeIndex*>(tree_[7]); }
    const TreeNodeIndex* leafToInternal() const
    {
        return static_cast<const TreeNodeIndex*>(tree_[9]) + (info_[1] - 1) / 7;
    }
    TreeNodeIndex numLeafNodes() const { return TreeNodeIndex(info_[1]); }
    unsigned maxTreeLevel() const { return 21; }
    ryo_system* system() const { return sys_; }

private:
    unsigned bucket_;
    ryo_system* sys_ = nullptr;
    size_t cap_      = 0;
    int64_t info_[4] = {0, 0, 0, 0};
    void* tree_[14]  = {};
};
} // namespace ryoanji
