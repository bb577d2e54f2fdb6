/*
 * demo_ref.cpp -- a C++ caller with the call sequence of the reference's single-GPU demonstrator
 * (ryoanji/test/demo.cu:55-219) and of TreeBuilder's cornerstone loop (interface/treebuilder.cuh:57-101), compiled
 * against include/ryoanji_amd.hpp: TreeBuilder<uint64_t>::update(x, y, z, n, box); the five-step upsweep
 * computeLeafSourceCenterGpu -> upsweepCentersGpu -> computeLeafMultipoles -> upsweepMultipoles per level ->
 * setMacGpu; computeFixedGroups into GroupData, a gmPool of stackSize(numGroups), traverse from node 1,
 * readBhStats; directSum<T>; and updateOctreeGpu iterated from {0, 2^63} on the sorted keys.
 *
 *   demo_ref <f32|f64> [log2 N = 15] [theta = 0.5] [group size = 64] [dump file]
 * prints one line: N cells p2p/target m2p/target pot err_p50 err_p1 err_max cornerstone_iterations
 * The dump (for checking against the oracle) holds int64 {N, nodes, leaves, sizeof(T), iterations, loop leaves},
 * then T x, y, z as generated, T x, y, z after update, T m, h, uint64 node keys, int32 childOffsets, uint32 layout,
 * T centers [nodes][4], T multipoles [nodes][8], T p, ax, ay, az (traverse), T p, ax, ay, az (directSum),
 * uint64 loop leaves [L' + 1], uint32 loop counts [L'], uint64 stats [5], double potential.
 */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "../../include/ryoanji_amd.hpp"

using namespace ryoanji;
using cstone::DeviceVector;
using cstone::rawPtr;

template<class Tc, class Th, class Tm, class Ta, class Tf, class MType>
static std::array<double, 5> computeAcceleration(size_t firstBody, size_t lastBody, const Tc* x, const Tc* y,
                                                 const Tc* z, const Tm* m, const Th* h, Tc G, int numShells,
                                                 const cstone::Box<Tc>& box, Ta* p, Ta* ax, Ta* ay, Ta* az,
                                                 const TreeNodeIndex* childOffsets, const TreeNodeIndex* internalToLeaf,
                                                 const LocalIndex* layout, const Vec4<Tf>* sourceCenter,
                                                 const MType* Multipole, unsigned groupSize)
{
    auto numBodies = lastBody - firstBody;
    cstone::GroupData<cstone::execution::Gpu> groups;
    cstone::computeFixedGroups(cstone::execution::gpuDefaultStream, firstBody, lastBody, groupSize, groups);
    DeviceVector<int> globalPool(stackSize(groups.numGroups));

    double totalPotential = traverse(groups.view(), 1, x, y, z, m, h, x, y, z, m, h, childOffsets, internalToLeaf,
                                     layout, sourceCenter, Multipole, G, numShells, {box.lx(), box.ly(), box.lz()}, p,
                                     ax, ay, az, rawPtr(globalPool));
    auto stats = readBhStats();
    return {double(stats[0]) / numBodies, double(stats[1]), double(stats[2]) / numBodies, double(stats[3]),
            totalPotential};
}

template<class KeyType, class T, class MType>
static void upsweep(int numSources, int numLeaves, int numLevels, float theta, const TreeNodeIndex* levelRange,
                    const T* x, const T* y, const T* z, const T* m, const cstone::Box<T>& box,
                    const LocalIndex* layout, const KeyType* prefixes, const TreeNodeIndex* childOffsets,
                    const TreeNodeIndex* leafToInternal, Vec4<T>* centers, MType* Multipole)
{
    cstone::computeLeafSourceCenterGpu(cstone::execution::gpuDefaultStream, x, y, z, m, leafToInternal, numLeaves,
                                       layout, centers);
    cstone::upsweepCentersGpu(cstone::execution::gpuDefaultStream, cstone::maxTreeLevel<KeyType>{}, levelRange,
                              childOffsets, centers);

    computeLeafMultipoles(x, y, z, m, leafToInternal, numLeaves, layout, centers, Multipole);
    for (int level = numLevels - 1; level >= 1; level--)
    {
        upsweepMultipoles(levelRange[level], levelRange[level + 1], childOffsets, centers, Multipole);
    }

    cstone::setMacGpu(cstone::execution::gpuDefaultStream, prefixes, numSources, centers, 1.f / theta, box);
}

template<class T>
static void put(std::ofstream& f, const std::vector<T>& v)
{
    f.write(reinterpret_cast<const char*>(v.data()), std::streamsize(v.size() * sizeof(T)));
}

template<class T>
static int run(int power, float theta, unsigned groupSize, const char* dumpPath)
{
    using MultipoleType = CartesianQuadrupole<T>;
    const size_t numBodies = size_t(1) << power;
    const T G              = 1.0;
    const int ncrit        = 64;
    const int numShells    = 0;

    std::mt19937 rng(42);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::vector<T> x(numBodies), y(numBodies), z(numBodies), m(numBodies, T(1.0 / numBodies)), h(numBodies, T(1e-3));
    for (size_t i = 0; i < numBodies; ++i)
        x[i] = T(u(rng)), y[i] = T(u(rng)), z[i] = T(u(rng));

    DeviceVector<T> d_x = x, d_y = y, d_z = z, d_m = m, d_h = h;
    cstone::Box<T> box(-1, 1);

    TreeBuilder<uint64_t> treeBuilder(ncrit);
    int numSources = treeBuilder.update(rawPtr(d_x), rawPtr(d_y), rawPtr(d_z), numBodies, box);

    const TreeNodeIndex* levelRange = treeBuilder.levelRange();
    int highestLevel                = treeBuilder.maxTreeLevel();

    DeviceVector<Vec4<T>> sourceCenter(numSources);
    DeviceVector<MultipoleType> Multipole(numSources);

    upsweep(numSources, treeBuilder.numLeafNodes(), highestLevel, theta, levelRange, rawPtr(d_x), rawPtr(d_y),
            rawPtr(d_z), rawPtr(d_m), box, treeBuilder.layout(), treeBuilder.nodeKeys(), treeBuilder.childOffsets(),
            treeBuilder.leafToInternal(), rawPtr(sourceCenter), rawPtr(Multipole));

    DeviceVector<T> d_p(numBodies, 0), d_ax(numBodies, 0), d_ay(numBodies, 0), d_az(numBodies, 0);
    auto interactions = computeAcceleration(0, numBodies, rawPtr(d_x), rawPtr(d_y), rawPtr(d_z), rawPtr(d_m),
                                            rawPtr(d_h), G, numShells, box, rawPtr(d_p), rawPtr(d_ax), rawPtr(d_ay),
                                            rawPtr(d_az), treeBuilder.childOffsets(), treeBuilder.internalToLeaf(),
                                            treeBuilder.layout(), rawPtr(sourceCenter), rawPtr(Multipole), groupSize);

    DeviceVector<T> refP(numBodies, 0), refAx(numBodies, 0), refAy(numBodies, 0), refAz(numBodies, 0);
    directSum(0, numBodies, numBodies, Vec3<T>{box.lx(), box.ly(), box.lz()}, numShells, rawPtr(d_x), rawPtr(d_y),
              rawPtr(d_z), rawPtr(d_m), rawPtr(d_h), rawPtr(refP), rawPtr(refAx), rawPtr(refAy), rawPtr(refAz));

    // TreeBuilder's cornerstone loop (treebuilder.cuh:81-90) on the sorted keys of the sorted bodies
    DeviceVector<uint64_t> d_keys(numBodies);
    cstone::computeSfcKeys(cstone::execution::gpuDefaultStream, rawPtr(d_x), rawPtr(d_y), rawPtr(d_z),
                           rawPtr(d_keys), numBodies, box);
    DeviceVector<uint64_t> d_tree = std::vector<uint64_t>{0, cstone::nodeRange<uint64_t>(0)}, tmpTree;
    DeviceVector<unsigned> d_counts = std::vector<unsigned>{unsigned(numBodies)};
    DeviceVector<TreeNodeIndex> workArray;
    int iterations = 0;
    while (!cstone::updateOctreeGpu<uint64_t>(cstone::execution::gpuDefaultStream,
                                              {rawPtr(d_keys), d_keys.size()}, ncrit, d_tree, d_counts, tmpTree,
                                              workArray))
    {
        if (++iterations > 64) return std::fprintf(stderr, "cornerstone loop did not converge\n"), 1;
    }

    std::vector<T> h_ax = d_ax.toHost(), h_ay = d_ay.toHost(), h_az = d_az.toHost();
    std::vector<T> r_ax = refAx.toHost(), r_ay = refAy.toHost(), r_az = refAz.toHost();
    std::vector<double> delta(numBodies);
    for (size_t i = 0; i < numBodies; ++i)
    {
        double dx = double(r_ax[i]) - h_ax[i], dy = double(r_ay[i]) - h_ay[i], dz = double(r_az[i]) - h_az[i];
        double r2 = double(r_ax[i]) * r_ax[i] + double(r_ay[i]) * r_ay[i] + double(r_az[i]) * r_az[i];
        delta[i]  = std::sqrt((dx * dx + dy * dy + dz * dz) / r2);
    }
    std::sort(delta.begin(), delta.end());
    std::printf("%zu %d %.1f %.1f %.8f %.3e %.3e %.3e %d\n", numBodies, numSources, interactions[0], interactions[2],
                interactions[4], delta[numBodies / 2], delta[size_t(numBodies * 0.99)], delta[numBodies - 1],
                iterations);

    if (dumpPath)
    {
        std::ofstream f(dumpPath, std::ios::binary);
        const int L = treeBuilder.numLeafNodes();
        std::vector<int64_t> head{int64_t(numBodies), numSources, L, int64_t(sizeof(T)), iterations,
                                  int64_t(d_tree.size()) - 1};
        put(f, head);
        put(f, x), put(f, y), put(f, z);
        put(f, d_x.toHost()), put(f, d_y.toHost()), put(f, d_z.toHost()), put(f, d_m.toHost()), put(f, d_h.toHost());
        std::vector<uint64_t> pre(numSources);
        std::vector<int32_t> child(numSources);
        std::vector<uint32_t> lay(L + 1);
        ryoanji_amd::check(ryo_memcpy(ryoanji_amd::ctx(), pre.data(), treeBuilder.nodeKeys(), pre.size() * 8));
        ryoanji_amd::check(ryo_memcpy(ryoanji_amd::ctx(), child.data(), treeBuilder.childOffsets(), child.size() * 4));
        ryoanji_amd::check(ryo_memcpy(ryoanji_amd::ctx(), lay.data(), treeBuilder.layout(), lay.size() * 4));
        put(f, pre), put(f, child), put(f, lay);
        put(f, sourceCenter.toHost()), put(f, Multipole.toHost());
        put(f, d_p.toHost()), put(f, h_ax), put(f, h_ay), put(f, h_az);
        put(f, refP.toHost()), put(f, r_ax), put(f, r_ay), put(f, r_az);
        put(f, d_tree.toHost()), put(f, d_counts.toHost());
        auto stats = readBhStats();
        put(f, std::vector<uint64_t>(stats.begin(), stats.end()));
        put(f, std::vector<double>{interactions[4]});
    }
    return 0;
}

int main(int argc, char** argv)
{
    if (argc < 2) return std::fprintf(stderr, "usage: demo_ref f32|f64 [log2 N] [theta] [group size] [dump]\n"), 2;
    const std::string type = argv[1];
    const int power        = argc > 2 ? std::atoi(argv[2]) : 15;
    const float theta      = argc > 3 ? float(std::atof(argv[3])) : 0.5f;
    const unsigned gs      = argc > 4 ? unsigned(std::atoi(argv[4])) : unsigned(bhMaxTargetSize());
    const char* dump       = argc > 5 ? argv[5] : nullptr;
    try
    {
        if (type == "f32") return run<float>(power, theta, gs, dump);
        if (type == "f64") return run<double>(power, theta, gs, dump);
    }
    catch (const std::exception& e)
    {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return std::fprintf(stderr, "type must be f32 or f64\n"), 2;
}
