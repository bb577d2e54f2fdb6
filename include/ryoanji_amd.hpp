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

#include <algorithm>
#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>
#if __cplusplus >= 202002L
#include <span>
#endif

#include "ryoanji_amd.h"

namespace ryoanji_amd
{
//! the calling thread's default C-ABI context (stream, scratch, statistics); created on first use on the null
//! stream, or bound to a stream by bindStream
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

/*! the calling thread's context on HIP stream `stream` (a hipStream_t), created on first use: work submitted with an
 *  execution::Gpu runs on that policy's stream, ordered with the caller's other work there. A null stream is the
 *  thread's default context. Each context keeps its own statistics and scratch. */
inline ryo_ctx* ctx(void* stream)
{
    if (!stream) return ctx();
    thread_local std::vector<std::pair<void*, ryo_ctx*>> byStream;
    for (auto& e : byStream)
        if (e.first == stream) return e.second;
    ryo_ctx* c = nullptr;
    check(nullptr, ryo_ctx_create(&c, stream));
    byStream.emplace_back(stream, c);
    return c;
}

//! run this thread's calls that take no execution policy (the ryoanji functions) on `hipStream` (a hipStream_t)
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
template<class T>
using SourceCenterType = Vec4<T>; // cstone/focus/source_center.hpp:25-26

//! maxTreeLevel<KeyType> (cstone/sfc/common.hpp): 21 levels of 64-bit keys, 10 of 32-bit keys
template<class KeyType>
struct maxTreeLevel : std::integral_constant<unsigned, sizeof(KeyType) == 8 ? 21u : 10u>
{
};

//! key range of a node at `level` (cstone/tree/definitions.h nodeRange)
template<class KeyType>
constexpr KeyType nodeRange(unsigned level)
{
    return KeyType(1) << (3u * (maxTreeLevel<KeyType>{} - level));
}

namespace execution
{
//! a HIP stream (cstone/execution.hpp:37-56); work submitted with it runs on that stream's context
struct Gpu
{
    void* stream = nullptr;
    Gpu() = default;
    Gpu(void* s)
        : stream(s)
    {
    }
};
inline const Gpu gpuDefaultStream{};
} // namespace execution

//! the context of a policy's stream
inline ryo_ctx* ctxFor(execution::Gpu exec) { return ryoanji_amd::ctx(exec.stream); }

/*! device vector (cstone/cuda/device_vector.h:33-79) over the C ABI's device allocation: uninitialized on growth
 *  except where the reference's constructors fill it (size + init value, host vector upload) */
template<class T>
class DeviceVector
{
public:
    using value_type = T;

    DeviceVector() = default;
    explicit DeviceVector(std::size_t n) { resize(n); }
    DeviceVector(std::size_t n, T init) { *this = std::vector<T>(n, init); }
    DeviceVector(const std::vector<T>& h) { *this = h; }
    DeviceVector(const T* first, const T* last) { *this = std::vector<T>(first, last); }
    DeviceVector(const DeviceVector& o)
    {
        resize(o.size_);
        copy(ptr_, o.ptr_, size_);
    }
    DeviceVector(DeviceVector&& o) noexcept { swap(o); }
    ~DeviceVector()
    {
        if (ptr_) ryo_free(ryoanji_amd::ctx(), ptr_);
    }
    DeviceVector& operator=(DeviceVector o)
    {
        swap(o);
        return *this;
    }
    DeviceVector& operator=(const std::vector<T>& h)
    {
        resize(h.size());
        copy(ptr_, h.data(), h.size());
        return *this;
    }

    T* data() { return ptr_; }
    const T* data() const { return ptr_; }
    T* begin() { return ptr_; }
    T* end() { return ptr_ + size_; }
    const T* cbegin() const { return ptr_; }
    const T* cend() const { return ptr_ + size_; }
    std::size_t size() const { return size_; }
    std::size_t capacity() const { return cap_; }
    bool empty() const { return size_ == 0; }

    //! keeps the first min(size, n) elements
    void resize(std::size_t n)
    {
        reserve(n);
        size_ = n;
    }
    void reserve(std::size_t n)
    {
        if (n <= cap_) return;
        void* p = nullptr;
        ryoanji_amd::check(ryo_malloc(ryoanji_amd::ctx(), &p, n * sizeof(T)));
        if (ptr_)
        {
            copy(static_cast<T*>(p), ptr_, size_);
            ryo_free(ryoanji_amd::ctx(), ptr_);
        }
        ptr_ = static_cast<T*>(p), cap_ = n;
    }
    void shrink_to_fit() {}
    DeviceVector& swap(DeviceVector& o) noexcept
    {
        std::swap(ptr_, o.ptr_), std::swap(size_, o.size_), std::swap(cap_, o.cap_);
        return *this;
    }
    //! host copy of the contents
    std::vector<T> toHost() const
    {
        std::vector<T> h(size_);
        copy(h.data(), ptr_, size_);
        return h;
    }

private:
    friend void swap(DeviceVector& a, DeviceVector& b) noexcept { a.swap(b); }
    static void copy(T* dst, const T* src, std::size_t n)
    {
        if (n) ryoanji_amd::check(ryo_memcpy(ryoanji_amd::ctx(), dst, src, n * sizeof(T)));
    }

    T* ptr_           = nullptr;
    std::size_t size_ = 0, cap_ = 0;
};

template<class T>
T* rawPtr(DeviceVector<T>& v)
{
    return v.data();
}
template<class T>
const T* rawPtr(const DeviceVector<T>& v)
{
    return v.data();
}
//! number of nodes of a cornerstone leaf array (its length minus the terminator)
template<class V>
TreeNodeIndex nNodes(const V& tree)
{
    return TreeNodeIndex(tree.size()) - 1;
}

enum class BoundaryType : uint8_t
{
    open     = 0,
    periodic = 1
};

//! cstone/sfc/box.hpp:99-214
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

//! cstone/traversal/groups.hpp:37-64 (the GPU form: group boundaries in a device vector)
template<class Exec>
class GroupData
{
public:
    GroupData() = default;
    GroupView view() const { return {firstBody, lastBody, numGroups, groupStart, groupEnd}; }

    DeviceVector<LocalIndex> data;
    LocalIndex firstBody = 0, lastBody = 0;
    LocalIndex numGroups = 0;
    LocalIndex* groupStart = nullptr;
    LocalIndex* groupEnd   = nullptr;
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
inline void computeSfcKeys(execution::Gpu exec, const float* x, const float* y, const float* z, uint64_t* keys,
                           size_t n, const Box<float>& box)
{
    ryoanji_amd::check(ryo_sfc_keys_f32(ctxFor(exec), x, y, z, keys, n, box.limits(), 0, 1));
}

//! cstone::computeSfcKeys with 32-bit Hilbert keys (10 levels) for float and double boxes (sfc_gpu.cu:46-62)
inline void computeSfcKeys(execution::Gpu exec, const float* x, const float* y, const float* z, uint32_t* keys,
                           size_t n, const Box<float>& box)
{
    ryoanji_amd::check(ryo_sfc_keys_u32_f32(ctxFor(exec), x, y, z, keys, n, box.limitsT(), 0, 1));
}
inline void computeSfcKeys(execution::Gpu exec, const double* x, const double* y, const double* z, uint32_t* keys,
                           size_t n, const Box<double>& box)
{
    ryoanji_amd::check(ryo_sfc_keys_u32_f64(ctxFor(exec), x, y, z, keys, n, box.limitsT(), 0, 1));
}

//! cstone::computeSfcKeys for double coordinates and Box<double> (sfc_gpu.cu:46-62)
inline void computeSfcKeys(execution::Gpu exec, const double* x, const double* y, const double* z, uint64_t* keys,
                           size_t n, const Box<double>& box)
{
    ryoanji_amd::check(ryo_sfc_keys_f64(ctxFor(exec), x, y, z, keys, n, box.limitsT(), 0, 1));
}

//! cstone::sortByKeyTempStorage -- cstone/primitives/primitives_gpu.h:120-124
template<class K, class V>
inline uint64_t sortByKeyTempStorage(uint64_t n)
{
    static_assert(sizeof(K) == 8 && sizeof(V) == 4, "(uint64 key, int32 value) pairs");
    return ryo_sort_temp_bytes(n);
}

//! cstone::sortByKey -- cstone/primitives/primitives_gpu.h:126-130
template<class V>
inline void sortByKey(execution::Gpu exec, uint64_t* first, uint64_t* last, V* values, uint64_t* keyBuf, V* valueBuf,
                      void* tmp, uint64_t tmpBytes)
{
    static_assert(sizeof(V) == 4 && std::is_integral_v<V>, "32-bit integer values");
    ryoanji_amd::check(ryo_sort_pairs_u64_i32(ctxFor(exec), first, reinterpret_cast<int32_t*>(values),
                                              size_t(last - first), keyBuf, reinterpret_cast<int32_t*>(valueBuf), tmp,
                                              tmpBytes));
}

//! cstone::sequence -- primitives_gpu.cu:265-273
template<class V>
inline void sequence(execution::Gpu exec, V* out, size_t n, V start)
{
    static_assert(sizeof(V) == 4 && std::is_integral_v<V>, "32-bit integer values");
    ryoanji_amd::check(ryo_sequence_i32(ctxFor(exec), reinterpret_cast<int32_t*>(out), n, int32_t(start)));
}

/*! converged cornerstone leaves + counts (cstone::computeOctree / updateOctreeGpu to convergence,
 *  csarray.hpp:429-440, update_gpu.cuh:43-68) and TreeBuilder's layout; returns the number of leaves */
inline int computeOctreeGpu(execution::Gpu exec, const uint64_t* sortedKeys, size_t n, unsigned bucketSize,
                            uint64_t* leaves, unsigned* counts, LocalIndex* layout, size_t capacity)
{
    int L = 0;
    ryoanji_amd::check(
        ryo_build_leaves_u64(ctxFor(exec), sortedKeys, n, bucketSize, leaves, counts, layout, capacity, &L));
    return L;
}

#if __cplusplus >= 202002L
//! cstone::computeNodeCountsGpu -- cstone/tree/csarray_gpu.h:41-48 (the count guess only shortens the search)
template<class KeyType>
inline void computeNodeCountsGpu(execution::Gpu exec, const KeyType* tree, unsigned* counts, TreeNodeIndex numNodes,
                                 std::span<const KeyType> keys, unsigned maxCount, bool /*useCountsAsGuess*/)
{
    static_assert(sizeof(KeyType) == 8, "64-bit keys");
    ryoanji_amd::check(ryo_node_counts_u64(ctxFor(exec), reinterpret_cast<const uint64_t*>(tree), numNodes,
                                           reinterpret_cast<const uint64_t*>(keys.data()), keys.size(), maxCount,
                                           counts));
}
#endif

//! cstone::computeNodeOpsGpu -- cstone/tree/csarray_gpu.h:59-65: returns the number of nodes after rebalancing
template<class KeyType>
inline TreeNodeIndex computeNodeOpsGpu(execution::Gpu exec, const KeyType* tree, TreeNodeIndex numNodes,
                                       const unsigned* counts, unsigned bucketSize, TreeNodeIndex* nodeOps)
{
    static_assert(sizeof(KeyType) == 8, "64-bit keys");
    int n = 0;
    ryoanji_amd::check(ryo_node_ops_u64(ctxFor(exec), reinterpret_cast<const uint64_t*>(tree), numNodes, counts,
                                        bucketSize, nodeOps, &n));
    return n;
}

//! cstone::rebalanceTreeGpu -- cstone/tree/csarray_gpu.h:67-73: returns true if no node changed (converged)
template<class KeyType>
inline bool rebalanceTreeGpu(execution::Gpu exec, const KeyType* tree, TreeNodeIndex numNodes,
                             TreeNodeIndex newNumNodes, const TreeNodeIndex* nodeOps, KeyType* newTree)
{
    static_assert(sizeof(KeyType) == 8, "64-bit keys");
    int converged = 0;
    ryoanji_amd::check(ryo_rebalance_tree_u64(ctxFor(exec), reinterpret_cast<const uint64_t*>(tree), numNodes,
                                              newNumNodes, nodeOps, reinterpret_cast<uint64_t*>(newTree),
                                              &converged));
    return converged != 0;
}

#if __cplusplus >= 202002L
/*! cstone::updateOctreeGpu -- cstone/tree/update_gpu.cuh:43-68: one rebalance + count step of the cornerstone leaf
 *  array; returns true when the tree did not change (converged). The vectors are device vectors with data(),
 *  size() and resize() (DeviceVector here). */
template<class KeyType, class DevKeyVec, class DevCountVec, class DevIdxVec>
bool updateOctreeGpu(execution::Gpu exec, std::span<const KeyType> keys, unsigned bucketSize, DevKeyVec& tree,
                     DevCountVec& counts, DevKeyVec& tmpTree, DevIdxVec& workArray,
                     unsigned maxCount = 0xffffffffu)
{
    workArray.resize(tree.size());
    TreeNodeIndex newNumNodes =
        computeNodeOpsGpu(exec, tree.data(), nNodes(tree), counts.data(), bucketSize, workArray.data());
    tmpTree.resize(newNumNodes + 1);
    bool converged = rebalanceTreeGpu(exec, tree.data(), nNodes(tree), newNumNodes, workArray.data(), tmpTree.data());
    swap(tree, tmpTree);
    counts.resize(nNodes(tree));
    computeNodeCountsGpu(exec, tree.data(), counts.data(), nNodes(tree), keys, maxCount, true);
    return converged;
}
#endif

//! cstone::buildOctreeGpu(exec, leaves, OctreeView) -- cstone/tree/octree_gpu.h:35-45
inline void buildOctreeGpu(execution::Gpu exec, const uint64_t* leaves, OctreeView<uint64_t> o)
{
    ryoanji_amd::check(ryo_link_octree_u64(ctxFor(exec), leaves, o.numLeafNodes, o.prefixes, o.childOffsets,
                                           o.parents, o.d_levelRange, o.internalToLeaf, o.leafToInternal,
                                           o.levelRange));
}

//! the reference's (Tc, Tm, Tf[, MType]) instantiations as the C ABI's type codes
template<class Tc, class Tm, class Tf>
constexpr int typeCode()
{
    if constexpr (std::is_same_v<Tc, float> && std::is_same_v<Tm, float> && std::is_same_v<Tf, float>) return 0;
    else if constexpr (std::is_same_v<Tc, double> && std::is_same_v<Tm, double> && std::is_same_v<Tf, double>)
        return 1;
    else
    {
        static_assert(std::is_same_v<Tc, double> && std::is_same_v<Tm, float> && std::is_same_v<Tf, double>,
                      "instantiations (f,f,f), (d,d,d), (d,f,d)");
        return 2;
    }
}

/*! cstone::computeLeafSourceCenterGpu -- cstone/focus/source_center_gpu.h:66-75: |m|-weighted leaf centers
 *  {com, sum|m|} at leafToInternal[leaf] */
template<class Tc, class Tm, class Tf>
inline void computeLeafSourceCenterGpu(execution::Gpu exec, const Tc* x, const Tc* y, const Tc* z, const Tm* m,
                                       const TreeNodeIndex* leafToInternal, TreeNodeIndex numLeaves,
                                       const LocalIndex* layout, Vec4<Tf>* centers)
{
    ryoanji_amd::check(ryo_leaf_centers(ctxFor(exec), typeCode<Tc, Tm, Tf>(), x, y, z, m, leafToInternal, numLeaves,
                                        layout, centers->data()));
}

//! cstone::upsweepCentersGpu -- source_center_gpu.h:77-88 (levelRange: host level ranges, as TreeBuilder's)
template<class T>
inline void upsweepCentersGpu(execution::Gpu exec, int numLevels, const TreeNodeIndex* levelRange,
                              const TreeNodeIndex* childOffsets, SourceCenterType<T>* centers)
{
    ryoanji_amd::check(ryo_upsweep_centers(ctxFor(exec), std::is_same_v<T, double> ? 1 : 0, numLevels, levelRange,
                                           childOffsets, centers->data()));
}

//! cstone::setMacGpu -- source_center_gpu.h:107-116: squared vector-MAC radii into centers[i][3]
template<class KeyType, class T>
inline void setMacGpu(execution::Gpu exec, const KeyType* prefixes, TreeNodeIndex numNodes, Vec4<T>* centers,
                      float invTheta, const Box<T>& box)
{
    static_assert(sizeof(KeyType) == 8, "64-bit keys");
    ryoanji_amd::check(ryo_set_mac(ctxFor(exec), std::is_same_v<T, double> ? 1 : 0,
                                   reinterpret_cast<const uint64_t*>(prefixes), numNodes, centers->data(), invTheta,
                                   box.limitsT()));
}

//! cstone::computeFixedGroups -- cstone/traversal/groups_gpu.cu:45-58 (groups has numGroups + 1 entries)
inline GroupView computeFixedGroups(execution::Gpu exec, LocalIndex first, LocalIndex last, unsigned groupSize,
                                    LocalIndex* groups)
{
    int G = 0;
    ryoanji_amd::check(ryo_fixed_groups(ctxFor(exec), first, last, groupSize, groups, &G));
    return GroupView{first, last, LocalIndex(G), groups, groups + 1};
}

//! cstone::computeFixedGroups -- cstone/traversal/groups_gpu.h:33-34, into GroupData's device vector
template<class Exec>
inline void computeFixedGroups(execution::Gpu exec, LocalIndex first, LocalIndex last, unsigned groupSize,
                               GroupData<Exec>& groups)
{
    const LocalIndex G = last > first ? (last - first + groupSize - 1) / groupSize : 0;
    groups.data.resize(G + 1);
    GroupView v       = computeFixedGroups(exec, first, last, groupSize, groups.data.data());
    groups.firstBody  = first;
    groups.lastBody   = last;
    groups.numGroups  = v.numGroups;
    groups.groupStart = groups.data.data();
    groups.groupEnd   = groups.data.data() + 1;
}

/*! cstone::computeGroupSplits -- cstone/traversal/groups_gpu.cu:121-151 for (Tc, T) = (float, float): target
 *  groups of up to groupSize (64) SFC-ordered particles, split where they are not compact. groups must hold
 *  up to last - first + 1 entries (`capacity`); extentRatio as CSTONE_GROUP_EXTENT_SPLIT (default 8). */
inline GroupView computeGroupSplits(execution::Gpu exec, LocalIndex first, LocalIndex last, const float* x,
                                    const float* y, const float* z, const float* h, const uint64_t* leaves,
                                    TreeNodeIndex numLeaves, const LocalIndex* layout, const Box<float>& box,
                                    unsigned groupSize, float tolFactor, LocalIndex* groups, size_t capacity,
                                    float extentRatio = 8.0f)
{
    int G = 0;
    ryoanji_amd::check(ryo_group_splits_f32(ctxFor(exec), first, last, x, y, z, h, leaves, int(numLeaves), layout,
                                            box.limits(), groupSize, tolFactor, extentRatio, groups, capacity, &G));
    return GroupView{first, last, LocalIndex(G), groups, groups + 1};
}

//! the reference's overload (groups_gpu.h:61-78): groups receives the numGroups + 1 boundaries
inline void computeGroupSplits(execution::Gpu exec, LocalIndex first, LocalIndex last, const float* x, const float* y,
                               const float* z, const float* h, const uint64_t* leaves, TreeNodeIndex numLeaves,
                               const LocalIndex* layout, const Box<float> box, unsigned groupSize, float tolFactor,
                               DeviceVector<LocalIndex>& /*numSplitsPerGroup*/, DeviceVector<LocalIndex>& groups)
{
    const size_t cap = size_t(last > first ? last - first : 0) + 2;
    groups.resize(cap);
    GroupView v = computeGroupSplits(exec, first, last, x, y, z, h, leaves, numLeaves, layout, box, groupSize,
                                     tolFactor, groups.data(), cap);
    groups.resize(v.numGroups + 1);
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

/*! ryoanji::computeLeafMultipoles -- nbody/upsweep_gpu.h:25 (P2M of each leaf about its center; the
 *  (f,f,f), (d,d,d) and (d,f,d) instantiations with CartesianQuadrupole or CartesianMDQpole) */
template<class Tc, class Tm, class Tf, class MType>
inline void computeLeafMultipoles(const Tc* x, const Tc* y, const Tc* z, const Tm* m,
                                  const TreeNodeIndex* leafToInternal, TreeNodeIndex numLeaves,
                                  const LocalIndex* layout, const Vec4<Tf>* centers, MType* multipoles)
{
    using Tq = typename MType::value_type;
    constexpr int code = cstone::typeCode<Tc, Tm, Tf>();
    static_assert(std::is_same_v<Tq, std::conditional_t<code == 1, double, float>>, "multipole type of the instantiation");
    static_assert(std::tuple_size<MType>::value == 8 || std::tuple_size<MType>::value == 12, "CQ or MDQ");
    ryoanji_amd::check(ryo_leaf_multipoles(ryoanji_amd::ctx(), code, std::tuple_size<MType>::value == 12, x, y, z, m,
                                           leafToInternal, numLeaves, layout, centers->data(), multipoles->data()));
}

/*! ryoanji::upsweepMultipoles -- nbody/upsweep_gpu.h:40: M2M of cells [firstCell, lastCell) (one tree level)
 *  from their children about the stored centers */
template<class T, class MType>
inline void upsweepMultipoles(TreeNodeIndex firstCell, TreeNodeIndex lastCell, const TreeNodeIndex* childOffsets,
                              const Vec4<T>* centers, MType* multipoles)
{
    using Tq           = typename MType::value_type;
    constexpr int code = std::is_same_v<T, float> ? 0 : (std::is_same_v<Tq, double> ? 1 : 2);
    static_assert(!std::is_same_v<T, float> || std::is_same_v<Tq, float>, "instantiations (f,f), (d,d), (d,f)");
    ryoanji_amd::check(ryo_upsweep_multipoles(ryoanji_amd::ctx(), code, std::tuple_size<MType>::value == 12,
                                              firstCell, lastCell, childOffsets, centers->data(), multipoles->data()));
}

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

//! ryoanji::stackSize -- traversal_gpu.h:62: the traversal keeps its stacks in LDS, so gmPool needs no space
inline LocalIndex stackSize(LocalIndex /*numGroups*/) { return 1; }

/*! ryoanji::splitFailedGroups -- traversal_gpu.h:70-76: the groups failedIds (device indices into `pass`) halved
 *  along the SFC into retryStart / retryEnd; .second counts the single targets that cannot be split */
inline std::pair<GroupView, unsigned> splitFailedGroups(const GroupView& pass, const LocalIndex* failedIds,
                                                        unsigned numFailed, cstone::DeviceVector<LocalIndex>& retryStart,
                                                        cstone::DeviceVector<LocalIndex>& retryEnd)
{
    ryo_ctx* c = ryoanji_amd::ctx();
    std::vector<LocalIndex> ids(numFailed), starts(pass.numGroups), ends(pass.numGroups);
    if (numFailed) ryoanji_amd::check(ryo_memcpy(c, ids.data(), failedIds, numFailed * sizeof(LocalIndex)));
    if (pass.numGroups)
    {
        ryoanji_amd::check(ryo_memcpy(c, starts.data(), pass.groupStart, pass.numGroups * sizeof(LocalIndex)));
        ryoanji_amd::check(ryo_memcpy(c, ends.data(), pass.groupEnd, pass.numGroups * sizeof(LocalIndex)));
    }
    std::vector<LocalIndex> ns, ne;
    unsigned unsplittable = 0;
    for (LocalIndex id : ids)
    {
        const LocalIndex a = starts[id], b = ends[id];
        if (b - a < 2)
        {
            ++unsplittable;
            continue;
        }
        const LocalIndex mid = a + (b - a) / 2;
        ns.push_back(a), ns.push_back(mid);
        ne.push_back(mid), ne.push_back(b);
    }
    retryStart = ns;
    retryEnd   = ne;
    return {GroupView{pass.firstBody, pass.lastBody, LocalIndex(ns.size()), retryStart.data(), retryEnd.data()},
            unsplittable};
}

/*! ryoanji::directSum<T> -- interface/multipole_holder.cuh:92, nbody/direct.cuh:88-107: targets [first, last)
 *  against sources [0, numBodies) and their periodic images, P2P<Exact> (accumulated in FP64 here), ADDED to
 *  p (m phi), ax, ay, az with G = 1 */
template<class T>
inline void directSum(size_t first, size_t last, size_t numBodies, Vec3<T> boxL, int numShells, const T* x,
                      const T* y, const T* z, const T* m, const T* h, T* p, T* ax, T* ay, T* az)
{
    static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>, "float or double");
    ryoanji_amd::check(ryo_direct_sum_range(ryoanji_amd::ctx(), std::is_same_v<T, double> ? 1 : 0, first, last,
                                            numBodies, boxL.data(), numShells, x, y, z, m, h, p, ax, ay, az));
}

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

/*! ryoanji::TreeBuilder<KeyType> -- interface/treebuilder.cuh:35-127: keys, stable sort, permutation, cornerstone
 *  leaves, linked octree and layout on the device, in device vectors the builder owns. update(x, y, z, n, box)
 *  sorts x, y, z into SFC order as the reference does (masses and smoothing lengths are the caller's to permute);
 *  update(x, y, z, m, h, n, box) permutes all five arrays. The leaves come from the one-pass construction of the
 *  converged cornerstone array (equal to iterating updateOctreeGpu from {0, 2^63}). */
template<class KeyType = uint64_t>
class TreeBuilder
{
    static_assert(sizeof(KeyType) == 8, "the octree is built from 64-bit keys");

public:
    explicit TreeBuilder(unsigned ncrit)
        : bucket_(ncrit)
    {
    }

    //! builds the octree of numBodies bodies (device arrays x, y, z sorted in place); returns the number of nodes
    template<class T>
    int update(T* x, T* y, T* z, size_t numBodies, const cstone::Box<T>& box)
    {
        T* arrays[3] = {x, y, z};
        return build<T>(arrays, 3, numBodies, box);
    }

    //! the same with masses and smoothing lengths permuted together with the coordinates
    template<class T>
    int update(T* x, T* y, T* z, T* m, T* h, size_t numBodies, const cstone::Box<T>& box)
    {
        T* arrays[5] = {x, y, z, m, h};
        return build<T>(arrays, 5, numBodies, box);
    }

    const LocalIndex* layout() const { return layout_.data(); }
    const KeyType* nodeKeys() const { return reinterpret_cast<const KeyType*>(prefixes_.data()); }
    const TreeNodeIndex* childOffsets() const { return childOffsets_.data(); }
    const TreeNodeIndex* parents() const { return parents_.data(); }
    const TreeNodeIndex* internalToLeaf() const { return internalToLeaf_.data(); }
    //! the leaf part of the map (numLeafNodes entries), as the reference's leafToInternal(octree)
    const TreeNodeIndex* leafToInternal() const { return leafToInternal_.data() + (numLeaves_ - 1) / 7; }
    //! host-resident octree level cell ranges (maxTreeLevel + 2 entries)
    const TreeNodeIndex* levelRange() const { return levelRange_.data(); }
    //! the same on the device
    const TreeNodeIndex* d_levelRange() const { return dLevelRange_.data(); }
    //! the converged cornerstone leaf keys (numLeafNodes + 1) and their particle counts
    const KeyType* leaves() const { return reinterpret_cast<const KeyType*>(leaves_.data()); }
    const unsigned* leafCounts() const { return counts_.data(); }
    TreeNodeIndex numLeafNodes() const { return numLeaves_; }
    unsigned maxTreeLevel() const { return cstone::maxTreeLevel<KeyType>{}; }

private:
    template<class T>
    int build(T* const* arrays, int numArrays, size_t n, const cstone::Box<T>& box)
    {
        using namespace cstone;
        static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>, "float or double coordinates");
        ryo_ctx* c = ryoanji_amd::ctx();
        keys_.resize(n), keysTmp_.resize(n), order_.resize(n), orderTmp_.resize(n);
        tmp_.resize(n * sizeof(T) * size_t(numArrays));
        const uint64_t tmpBytes = sortByKeyTempStorage<uint64_t, int>(n);
        sortTmp_.resize(tmpBytes);
        // the keys array is written, not read (the reference zero-fills it first, so no key is removeKey)
        if constexpr (std::is_same_v<T, float>)
            ryoanji_amd::check(ryo_sfc_keys_f32(c, arrays[0], arrays[1], arrays[2], keys_.data(), n, box.limits(), 0,
                                                0));
        else
            ryoanji_amd::check(ryo_sfc_keys_f64(c, arrays[0], arrays[1], arrays[2], keys_.data(), n, box.limitsT(), 0,
                                                0));
        ryoanji_amd::check(ryo_sequence_i32(c, order_.data(), n, 0));
        ryoanji_amd::check(ryo_sort_pairs_u64_i32(c, keys_.data(), order_.data(), n, keysTmp_.data(),
                                                  orderTmp_.data(), sortTmp_.data(), tmpBytes));
        T* gathered[5];
        const T* in[5];
        for (int k = 0; k < numArrays; ++k)
            gathered[k] = reinterpret_cast<T*>(tmp_.data()) + size_t(k) * n, in[k] = arrays[k];
        if constexpr (std::is_same_v<T, float>)
            ryoanji_amd::check(ryo_gather_f32(c, order_.data(), n, numArrays, in, gathered));
        else ryoanji_amd::check(ryo_gather_f64(c, order_.data(), n, numArrays, in, gathered));
        for (int k = 0; k < numArrays; ++k)
            ryoanji_amd::check(ryo_memcpy(c, arrays[k], gathered[k], n * sizeof(T)));

        // converged leaves, counts and layout; the capacity grows to the leaf count when it is too small
        size_t cap = std::max<size_t>(64, 2 * n / std::max(1u, bucket_) + 64);
        for (;;)
        {
            leaves_.resize(cap + 1), counts_.resize(cap), layout_.resize(cap + 1);
            int L  = 0;
            int rc = ryo_build_leaves_u64(c, keys_.data(), n, bucket_, leaves_.data(), counts_.data(), layout_.data(),
                                          cap, &L);
            if (rc == RYO_ERR_CAPACITY && size_t(L) > cap)
            {
                cap = size_t(L);
                continue;
            }
            ryoanji_amd::check(c, rc);
            numLeaves_ = L;
            break;
        }
        leaves_.resize(numLeaves_ + 1), counts_.resize(numLeaves_), layout_.resize(numLeaves_ + 1);

        const int nI = (numLeaves_ - 1) / 7, nn = numLeaves_ + nI;
        prefixes_.resize(nn), childOffsets_.resize(nn + 1), parents_.resize(std::max(1, (nn - 1) / 8));
        internalToLeaf_.resize(nn), leafToInternal_.resize(nn), dLevelRange_.resize(cstone::maxTreeLevel<KeyType>{} + 2);
        levelRange_.assign(cstone::maxTreeLevel<KeyType>{} + 2, 0);
        ryoanji_amd::check(ryo_link_octree_u64(c, leaves_.data(), numLeaves_, prefixes_.data(), childOffsets_.data(),
                                               parents_.data(), dLevelRange_.data(), internalToLeaf_.data(),
                                               leafToInternal_.data(), levelRange_.data()));
        return nn;
    }

    unsigned bucket_;
    TreeNodeIndex numLeaves_ = 0;
    cstone::DeviceVector<uint64_t> keys_, keysTmp_, leaves_, prefixes_;
    cstone::DeviceVector<int32_t> order_, orderTmp_;
    cstone::DeviceVector<char> sortTmp_, tmp_;
    cstone::DeviceVector<unsigned> counts_;
    cstone::DeviceVector<LocalIndex> layout_;
    cstone::DeviceVector<TreeNodeIndex> childOffsets_, parents_, internalToLeaf_, leafToInternal_, dLevelRange_;
    std::vector<TreeNodeIndex> levelRange_;
};
} // namespace ryoanji
