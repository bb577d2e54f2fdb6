/*
 * ref_harness.cpp -- C entry points over the reference's own CPU path (test infrastructure).
 *
 * Compiled by oracle/Makefile against the reference headers where they lie; the result,
 * oracle/_ref/libryoref.so, validates the restatement in ryo_oracle.cpp, writes the golden
 * fixtures under tests/golden/ and serves as the "reference" CPU baseline. Nothing here is
 * copied from the reference: this file only calls its functions, with the same C signatures
 * as the orc_* functions of the restatement (prefix ref_ instead of orc_).
 */

#include <vector>

#include "cstone/sfc/box.hpp"
#include "cstone/sfc/sfc.hpp"
#include "cstone/primitives/gather.hpp"
#include "cstone/tree/csarray.hpp"
#include "cstone/tree/octree.hpp"
#include "cstone/focus/source_center.hpp"
#include "ryoanji/nbody/cartesian_qpole.hpp"
#include "ryoanji/nbody/upsweep_cpu.hpp"
#include "ryoanji/nbody/traversal_cpu.hpp"
#include "ryoanji/nbody/ewald.hpp"

using namespace cstone;

namespace
{
template<class T>
Box<T> box6(const T* l)
{
    return Box<T>(l[0], l[1], l[2], l[3], l[4], l[5]);
}

struct Tree
{
    OctreeData<uint64_t, execution::Cpu> o;
    explicit Tree(const uint64_t* leaves, int L)
    {
        o.resize(L);
        updateInternalTree<uint64_t>(std::span<const uint64_t>(leaves, size_t(L) + 1), o.data());
    }
};
} // namespace

extern "C" {

int ref_sfc_keys(const float* x, const float* y, const float* z, uint64_t* keys, size_t n, const float* limits,
                 int hilbert)
{
    auto box = box6(limits);
    if (hilbert) computeSfcKeys(x, y, z, reinterpret_cast<HilbertKey<uint64_t>*>(keys), n, box);
    else computeSfcKeys(x, y, z, reinterpret_cast<MortonKey<uint64_t>*>(keys), n, box);
    return 0;
}

int ref_sfc_keys_f64(const double* x, const double* y, const double* z, uint64_t* keys, size_t n,
                     const double* limits, int hilbert)
{
    auto box = box6(limits);
    if (hilbert) computeSfcKeys(x, y, z, reinterpret_cast<HilbertKey<uint64_t>*>(keys), n, box);
    else computeSfcKeys(x, y, z, reinterpret_cast<MortonKey<uint64_t>*>(keys), n, box);
    return 0;
}

/* 32-bit keys (HilbertKey<uint32_t> / MortonKey<uint32_t>) of float or double coordinates */
int ref_sfc_keys_u32(const float* x, const float* y, const float* z, uint32_t* keys, size_t n, const float* limits,
                     int hilbert)
{
    auto box = box6(limits);
    if (hilbert) computeSfcKeys(x, y, z, reinterpret_cast<HilbertKey<uint32_t>*>(keys), n, box);
    else computeSfcKeys(x, y, z, reinterpret_cast<MortonKey<uint32_t>*>(keys), n, box);
    return 0;
}

int ref_sfc_keys_u32_f64(const double* x, const double* y, const double* z, uint32_t* keys, size_t n,
                         const double* limits, int hilbert)
{
    auto box = box6(limits);
    if (hilbert) computeSfcKeys(x, y, z, reinterpret_cast<HilbertKey<uint32_t>*>(keys), n, box);
    else computeSfcKeys(x, y, z, reinterpret_cast<MortonKey<uint32_t>*>(keys), n, box);
    return 0;
}

void ref_sort_by_key(uint64_t* keys, int32_t* values, size_t n) { sort_by_key(keys, keys + n, values); }

long ref_compute_octree(const uint64_t* sortedKeys, size_t n, unsigned bucket, uint64_t* leaves, unsigned* counts,
                        long capacity)
{
    auto [tree, cnt] = computeOctree<uint64_t>(std::span<const uint64_t>(sortedKeys, n), bucket);
    long L           = long(cnt.size());
    if (L > capacity) return -L;
    std::copy(tree.begin(), tree.end(), leaves);
    std::copy(cnt.begin(), cnt.end(), counts);
    return L;
}

int ref_link_octree(const uint64_t* leaves, int numLeaves, uint64_t* prefixes, int* childOffsets, int* parents,
                    int* levelRange, int* internalToLeaf, int* leafToInternal)
{
    Tree t(leaves, numLeaves);
    auto& o = t.o;
    std::copy(o.prefixes.begin(), o.prefixes.begin() + o.numNodes, prefixes);
    std::copy(o.childOffsets.begin(), o.childOffsets.begin() + o.numNodes + 1, childOffsets);
    std::copy(o.parents.begin(), o.parents.end(), parents);
    std::copy(o.levelRange.begin(), o.levelRange.end(), levelRange);
    std::copy(o.internalToLeaf.begin(), o.internalToLeaf.begin() + o.numNodes, internalToLeaf);
    std::copy(o.leafToInternal.begin(), o.leafToInternal.begin() + o.numNodes, leafToInternal);
    return o.numNodes;
}

} // extern "C"

template<class T, class MType = ryoanji::CartesianQuadrupole<T>>
static void upsweepImpl(const T* x, const T* y, const T* z, const T* m, const uint64_t* leaves, int numLeaves,
                        const unsigned* layout, const T* limits, float invTheta, T* centers, T* mpoles)
{
    Tree t(leaves, numLeaves);
    auto& o  = t.o;
    auto box = box6(limits);
    size_t n = layout[numLeaves];
    auto c   = reinterpret_cast<SourceCenterType<T>*>(centers);
    auto mp  = reinterpret_cast<MType*>(mpoles);
    computeLeafMassCenter<T, T, T>(std::span<const T>(x, n), std::span<const T>(y, n), std::span<const T>(z, n),
                                   std::span<const T>(m, n), leafToInternal(o), layout, c);
    upsweep(std::span<const TreeNodeIndex>(o.levelRange), o.childOffsets.data(), c, CombineSourceCenter<T>{});
    setMac<T, uint64_t>(std::span<const uint64_t>(o.prefixes.data(), o.numNodes),
                        std::span<SourceCenterType<T>>(c, o.numNodes), invTheta, box);
    ryoanji::computeLeafMultipoles(x, y, z, m, leafToInternal(o), layout, c, mp);
    ryoanji::upsweepMultipoles(std::span<const TreeNodeIndex>(o.levelRange), o.childOffsets.data(), c, mp);
}

extern "C" {

int ref_upsweep_f32(const float* x, const float* y, const float* z, const float* m, const uint64_t* leaves,
                    int numLeaves, const unsigned* layout, const float* limits, float invTheta, float* centers,
                    float* mpoles)
{
    upsweepImpl<float>(x, y, z, m, leaves, numLeaves, layout, limits, invTheta, centers, mpoles);
    return 0;
}

int ref_upsweep_f64(const double* x, const double* y, const double* z, const double* m, const uint64_t* leaves,
                    int numLeaves, const unsigned* layout, const double* limits, float invTheta, double* centers,
                    double* mpoles)
{
    upsweepImpl<double>(x, y, z, m, leaves, numLeaves, layout, limits, invTheta, centers, mpoles);
    return 0;
}

/* the same with CartesianMDQpole<T> (monopole, dipole, quadrupole: 12 values per node, cartesian_qpole.hpp:237-380) */
int ref_upsweep_mdq_f32(const float* x, const float* y, const float* z, const float* m, const uint64_t* leaves,
                        int numLeaves, const unsigned* layout, const float* limits, float invTheta, float* centers,
                        float* mpoles)
{
    upsweepImpl<float, ryoanji::CartesianMDQpole<float>>(x, y, z, m, leaves, numLeaves, layout, limits, invTheta,
                                                         centers, mpoles);
    return 0;
}

int ref_upsweep_mdq_f64(const double* x, const double* y, const double* z, const double* m, const uint64_t* leaves,
                        int numLeaves, const unsigned* layout, const double* limits, float invTheta, double* centers,
                        double* mpoles)
{
    upsweepImpl<double, ryoanji::CartesianMDQpole<double>>(x, y, z, m, leaves, numLeaves, layout, limits, invTheta,
                                                           centers, mpoles);
    return 0;
}

/* computeGravity with CartesianMDQpole<float> multipoles, all float */
double ref_compute_gravity_mdq_f32(const uint64_t* leaves, int numLeaves, const unsigned* layout,
                                   const float* centers, const float* mpoles, const float* x, const float* y,
                                   const float* z, const float* h, const float* m, const float* limits, float G,
                                   int numShells, int firstLeaf, int lastLeaf, float* ugrav, float* ax, float* ay,
                                   float* az)
{
    Tree t(leaves, numLeaves);
    auto& o  = t.o;
    auto box = box6(limits);
    float tot = 0;
    ryoanji::computeGravity(o.childOffsets.data(), o.parents.data(), o.internalToLeaf.data(),
                            reinterpret_cast<const SourceCenterType<float>*>(centers),
                            reinterpret_cast<const ryoanji::CartesianMDQpole<float>*>(mpoles), layout, firstLeaf,
                            lastLeaf, x, y, z, h, m, box, G, ugrav, ax, ay, az, &tot, numShells);
    return tot;
}

/* The reference's per-group traversal computeGravityGroup (traversal_cpu.hpp:67-157) with CartesianMDQpole<float>,
 * driven over groups of 64 consecutive targets (the GPU demo's grouping) with the image loop and output
 * accumulation of computeGravity (traversal_cpu.hpp:172-230); n must be a multiple of 64 (the reference CPU reads
 * a partial group's unset targets into its bounding box). Returns 0.5 G sum m phi. */
double ref_compute_gravity_mdq_g64_f32(const uint64_t* leaves, int numLeaves, const unsigned* layout,
                                       const float* centers, const float* mpoles, const float* x, const float* y,
                                       const float* z, const float* h, const float* m, const float* limits, float G,
                                       int numShells, float* ugrav, float* ax, float* ay, float* az)
{
    constexpr unsigned gs = 64;
    Tree t(leaves, numLeaves);
    auto& o            = t.o;
    auto box           = box6(limits);
    const unsigned n   = layout[numLeaves];
    const auto* mp     = reinterpret_cast<const ryoanji::CartesianMDQpole<float>*>(mpoles);
    const auto* cs     = reinterpret_cast<const SourceCenterType<float>*>(centers);
    double tot         = 0;
#pragma omp parallel for reduction(+ : tot)
    for (unsigned i = 0; i < n; i += gs)
    {
        util::array<Vec4<float>, gs> targets, acc;
        for (unsigned k = 0; k < gs; ++k)
            targets[k] = {x[i + k], y[i + k], z[i + k], h[i + k]}, acc[k] = {0, 0, 0, 0};
        for (int iz = -numShells; iz <= numShells; ++iz)
            for (int iy = -numShells; iy <= numShells; ++iy)
                for (int ix = -numShells; ix <= numShells; ++ix)
                {
                    Vec4<float> shift{ix * box.lx(), iy * box.ly(), iz * box.lz(), 0};
                    auto ts = targets;
                    for (auto& q : ts)
                        q -= shift;
                    ryoanji::computeGravityGroup(ts, o.childOffsets.data(), o.parents.data(), o.internalToLeaf.data(),
                                                 cs, const_cast<ryoanji::CartesianMDQpole<float>*>(mp), layout, x, y, z,
                                                 h, m, acc.data());
                }
        for (unsigned k = 0; k < gs; ++k)
        {
            float u = G * m[i + k] * acc[k][0];
            tot += u;
            if (ugrav) ugrav[i + k] += u;
            ax[i + k] += G * acc[k][1];
            ay[i + k] += G * acc[k][2];
            az[i + k] += G * acc[k][3];
        }
    }
    return 0.5 * tot;
}

/* computeGravity (traversal_cpu.hpp:172-230) over the leaf range [firstLeaf, lastLeaf); the reference groups
 * 16 consecutive targets. Returns the reference's ugravTot increment. */
double ref_compute_gravity_f32(const uint64_t* leaves, int numLeaves, const unsigned* layout, const float* centers,
                               const float* mpoles, const float* x, const float* y, const float* z, const float* h,
                               const float* m, const float* limits, float G, int numShells, int firstLeaf,
                               int lastLeaf, float* ugrav, float* ax, float* ay, float* az)
{
    Tree t(leaves, numLeaves);
    auto& o  = t.o;
    auto box = box6(limits);
    float tot = 0;
    ryoanji::computeGravity(o.childOffsets.data(), o.parents.data(), o.internalToLeaf.data(),
                            reinterpret_cast<const SourceCenterType<float>*>(centers),
                            reinterpret_cast<const ryoanji::CartesianQuadrupole<float>*>(mpoles), layout, firstLeaf,
                            lastLeaf, x, y, z, h, m, box, G, ugrav, ax, ay, az, &tot, numShells);
    return tot;
}

/* computeGravity with T1 = double coordinates and centers: (x, y, z, h, m, outputs) all double with
 * CartesianQuadrupole<double>, the CPU counterpart of the GPU (d,d,d,d,d) instantiation */
double ref_compute_gravity_f64(const uint64_t* leaves, int numLeaves, const unsigned* layout, const double* centers,
                               const double* mpoles, const double* x, const double* y, const double* z,
                               const double* h, const double* m, const double* limits, float G, int numShells,
                               int firstLeaf, int lastLeaf, double* ugrav, double* ax, double* ay, double* az)
{
    Tree t(leaves, numLeaves);
    auto& o  = t.o;
    auto box = box6(limits);
    double tot = 0;
    ryoanji::computeGravity(o.childOffsets.data(), o.parents.data(), o.internalToLeaf.data(),
                            reinterpret_cast<const SourceCenterType<double>*>(centers),
                            reinterpret_cast<const ryoanji::CartesianQuadrupole<double>*>(mpoles), layout, firstLeaf,
                            lastLeaf, x, y, z, h, m, box, G, ugrav, ax, ay, az, &tot, numShells);
    return tot;
}

/* computeGravity with double coordinates and centers, float h, m and outputs, CartesianQuadrupole<float>: the CPU
 * counterpart of the GPU (d,f,f,f,d) instantiation */
double ref_compute_gravity_df(const uint64_t* leaves, int numLeaves, const unsigned* layout, const double* centers,
                              const float* mpoles, const double* x, const double* y, const double* z, const float* h,
                              const float* m, const double* limits, float G, int numShells, int firstLeaf,
                              int lastLeaf, float* ugrav, float* ax, float* ay, float* az)
{
    Tree t(leaves, numLeaves);
    auto& o  = t.o;
    auto box = box6(limits);
    double tot = 0;
    ryoanji::computeGravity(o.childOffsets.data(), o.parents.data(), o.internalToLeaf.data(),
                            reinterpret_cast<const SourceCenterType<double>*>(centers),
                            reinterpret_cast<const ryoanji::CartesianQuadrupole<float>*>(mpoles), layout, firstLeaf,
                            lastLeaf, x, y, z, h, m, box, G, ugrav, ax, ay, az, &tot, numShells);
    return tot;
}

/* the two double-coordinate computeGravity instantiations with CartesianMDQpole multipoles: (d,d,d,d,d) with
 * CartesianMDQpole<double>, and (d,f,f,f,d) with CartesianMDQpole<float> */
double ref_compute_gravity_mdq_f64(const uint64_t* leaves, int numLeaves, const unsigned* layout,
                                   const double* centers, const double* mpoles, const double* x, const double* y,
                                   const double* z, const double* h, const double* m, const double* limits, float G,
                                   int numShells, int firstLeaf, int lastLeaf, double* ugrav, double* ax, double* ay,
                                   double* az)
{
    Tree t(leaves, numLeaves);
    auto& o  = t.o;
    auto box = box6(limits);
    double tot = 0;
    ryoanji::computeGravity(o.childOffsets.data(), o.parents.data(), o.internalToLeaf.data(),
                            reinterpret_cast<const SourceCenterType<double>*>(centers),
                            reinterpret_cast<const ryoanji::CartesianMDQpole<double>*>(mpoles), layout, firstLeaf,
                            lastLeaf, x, y, z, h, m, box, G, ugrav, ax, ay, az, &tot, numShells);
    return tot;
}

double ref_compute_gravity_mdq_df(const uint64_t* leaves, int numLeaves, const unsigned* layout,
                                  const double* centers, const float* mpoles, const double* x, const double* y,
                                  const double* z, const float* h, const float* m, const double* limits, float G,
                                  int numShells, int firstLeaf, int lastLeaf, float* ugrav, float* ax, float* ay,
                                  float* az)
{
    Tree t(leaves, numLeaves);
    auto& o  = t.o;
    auto box = box6(limits);
    double tot = 0;
    ryoanji::computeGravity(o.childOffsets.data(), o.parents.data(), o.internalToLeaf.data(),
                            reinterpret_cast<const SourceCenterType<double>*>(centers),
                            reinterpret_cast<const ryoanji::CartesianMDQpole<float>*>(mpoles), layout, firstLeaf,
                            lastLeaf, x, y, z, h, m, box, G, ugrav, ax, ay, az, &tot, numShells);
    return tot;
}

/* directSum (traversal_cpu.hpp:232-268) over all particles, T = double */
int ref_direct_sum_f64(const double* x, const double* y, const double* z, const double* h, const double* m, size_t n,
                       double G, const double* boxL, int numShells, double* ugrav, double* ax, double* ay, double* az)
{
    ryoanji::directSum(x, y, z, h, m, LocalIndex(n), float(G), Vec3<double>{boxL[0], boxL[1], boxL[2]}, numShells, ax,
                       ay, az, ugrav);
    return 0;
}

} // extern "C"

/* computeGravityEwald (ewald.hpp:356-389); settings = {numReplicaShells, lCut, hCut, alphaScale, smallR} */
template<class T>
static double ewaldImpl(const T* Mroot, const T* center, const T* x, const T* y, const T* z, const T* m, size_t n,
                        const T* limits, double G, const double* settings, T* ugrav, T* ax, T* ay, T* az)
{
    ryoanji::CartesianQuadrupole<T> M;
    for (int k = 0; k < 8; ++k)
        M[k] = Mroot[k];
    Vec3<T> c{center[0], center[1], center[2]};
    ryoanji::EwaldSettings s{int(settings[0]), settings[1], settings[2], settings[3], settings[4]};
    T tot = 0;
    ryoanji::computeGravityEwald(c, M, 0, LocalIndex(n), x, y, z, m, box6(limits), float(G), ugrav, ax, ay, az, &tot, s);
    return double(tot);
}

extern "C" {

double ref_ewald_f64(const double* Mroot, const double* center, const double* x, const double* y, const double* z,
                     const double* m, size_t n, const double* limits, double G, const double* settings, double* ugrav,
                     double* ax, double* ay, double* az)
{
    return ewaldImpl<double>(Mroot, center, x, y, z, m, n, limits, G, settings, ugrav, ax, ay, az);
}

double ref_ewald_f32(const float* Mroot, const float* center, const float* x, const float* y, const float* z,
                     const float* m, size_t n, const float* limits, double G, const double* settings, float* ugrav,
                     float* ax, float* ay, float* az)
{
    return ewaldImpl<float>(Mroot, center, x, y, z, m, n, limits, G, settings, ugrav, ax, ay, az);
}

} // extern "C"

/* mixed-dimension Hilbert functions of the reference (hilbert.hpp:117-244, :343-585; box.hpp:160-198) */
extern "C" {

uint64_t ref_hilbert_mixd(unsigned ix, unsigned iy, unsigned iz, const unsigned* bits)
{
    return iHilbert<uint64_t>(ix, iy, iz, bits[0], bits[1], bits[2]);
}

int ref_decode_hilbert_mixd(uint64_t key, const unsigned* bits, unsigned* xyz)
{
    auto [x, y, z] = decodeHilbert<uint64_t>(key, bits[0], bits[1], bits[2]);
    xyz[0] = unsigned(x), xyz[1] = unsigned(y), xyz[2] = unsigned(z);
    return isValidHilbertMixDKey<uint64_t>(key, bits[0], bits[1], bits[2]) ? 1 : 0;
}

void ref_box_bits(const float* limits, unsigned* bits)
{
    auto b = box6(limits).getBoxDimBits(maxTreeLevel<uint64_t>{});
    for (int d = 0; d < 3; ++d)
        bits[d] = b[d];
}

/* integer box {xmin, xmax, ymin, ymax, zmin, zmax} of the cell at `level` with start key `key` */
void ref_hilbert_ibox(uint64_t key, unsigned level, const unsigned* bits, int* ibox)
{
    IBox b = hilbertIBox<uint64_t>(key, level, bits[0], bits[1], bits[2]);
    ibox[0] = b.xmin(), ibox[1] = b.xmax(), ibox[2] = b.ymin(), ibox[3] = b.ymax(), ibox[4] = b.zmin(),
    ibox[5] = b.zmax();
}

} // extern "C"
