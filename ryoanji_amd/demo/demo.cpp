/*
 * demo.cpp -- a C++ caller written against the reference's entry-point names (include/ryoanji_amd.hpp), the
 * counterpart of ryoanji/test/demo.cu: build the octree of a cube of bodies, upsweep, traverse in 64-target
 * groups, and compare with an FP64 direct sum on sampled targets.
 *
 *   demo [log2 N = 16] [theta = 0.5] [dump file]
 * prints one line: N cells p2p/target m2p/target pot err_p50 err_p1 err_max
 * With a dump file it also writes, for checking against another implementation: int64 {N, nodes, leaves},
 * float32 x, y, z as generated, float32 x, y, z, m, h in SFC order, float32 p, ax, ay, az, uint64 node keys
 * (prefixes), int32 childOffsets, uint32 layout.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <vector>

#include "../../include/ryoanji_amd.hpp"

#define HIPCHECK(x)                                                                                                    \
    do                                                                                                                 \
    {                                                                                                                  \
        if ((x) != hipSuccess)                                                                                         \
        {                                                                                                              \
            std::fprintf(stderr, "HIP error %s at %d\n", hipGetErrorString(x), __LINE__);                           \
            return 1;                                                                                                  \
        }                                                                                                              \
    } while (0)

template<class T>
static T* devAlloc(size_t n)
{
    T* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) std::abort();
    return p;
}

int main(int argc, char** argv)
{
    using namespace ryoanji;
    const int power   = argc > 1 ? std::atoi(argv[1]) : 16;
    const float theta = argc > 2 ? float(std::atof(argv[2])) : 0.5f;
    const size_t n    = size_t(1) << power;
    const float G     = 1.0f;

    std::mt19937 rng(42);
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    std::vector<float> x(n), y(n), z(n), m(n, 1.0f / n), h(n, 1e-3f);
    for (size_t i = 0; i < n; ++i)
        x[i] = u(rng), y[i] = u(rng), z[i] = u(rng);
    const std::vector<float> x0 = x, y0 = y, z0 = z;

    float *dx = devAlloc<float>(n), *dy = devAlloc<float>(n), *dz = devAlloc<float>(n), *dm = devAlloc<float>(n),
          *dh = devAlloc<float>(n);
    for (auto [d, s] : {std::pair{dx, &x}, {dy, &y}, {dz, &z}, {dm, &m}, {dh, &h}})
        HIPCHECK(hipMemcpy(d, s->data(), n * sizeof(float), hipMemcpyHostToDevice));

    cstone::Box<float> box(-1.0f, 1.0f);
    TreeBuilder tb(64);
    int numNodes = tb.update(dx, dy, dz, dm, dh, n, box); // bodies come back in SFC order

    auto* centers    = devAlloc<Vec4<float>>(numNodes);
    auto* multipoles = devAlloc<CartesianQuadrupole<float>>(numNodes);
    upsweep(dx, dy, dz, dm, tb.layout(), tb.numLeafNodes(), tb.nodeKeys(), tb.childOffsets(), tb.parents(),
            tb.leafToInternal(), tb.d_levelRange(), box, theta, centers, multipoles);

    auto* groups = devAlloc<LocalIndex>(n / bhMaxTargetSize() + 2);
    GroupView grp = cstone::computeFixedGroups(cstone::execution::gpuDefaultStream, 0, LocalIndex(n),
                                               bhMaxTargetSize(), groups);
    float *p = devAlloc<float>(n), *ax = devAlloc<float>(n), *ay = devAlloc<float>(n), *az = devAlloc<float>(n);
    for (float* a : {p, ax, ay, az})
        HIPCHECK(hipMemset(a, 0, n * sizeof(float)));
    double pot = traverse(grp, 0, dx, dy, dz, dm, dh, dx, dy, dz, dm, dh, tb.childOffsets(), tb.internalToLeaf(),
                          tb.layout(), centers, multipoles, G, 0, {2.0f, 2.0f, 2.0f}, p, ax, ay, az);
    auto stats = readBhStats();

    // FP64 direct sum on 1024 sampled targets
    const size_t nt = std::min<size_t>(1024, n);
    std::vector<int64_t> tgt(nt);
    for (size_t i = 0; i < nt; ++i)
        tgt[i] = int64_t(i * (n / nt));
    auto* dtgt = devAlloc<int64_t>(nt);
    auto* dref = devAlloc<double>(4 * nt);
    HIPCHECK(hipMemcpy(dtgt, tgt.data(), nt * sizeof(int64_t), hipMemcpyHostToDevice));
    ryoanji_amd::check(
        ryo_direct_sum_f64(ryoanji_amd::ctx(), dx, dy, dz, dh, dm, n, dtgt, nt, G, nullptr, 0, dref));
    std::vector<double> ref(4 * nt);
    std::vector<float> hx(n), hy(n), hz(n);
    HIPCHECK(hipMemcpy(ref.data(), dref, ref.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(hx.data(), ax, n * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(hy.data(), ay, n * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(hz.data(), az, n * sizeof(float), hipMemcpyDeviceToHost));
    std::vector<double> err(nt);
    for (size_t i = 0; i < nt; ++i)
    {
        size_t t  = size_t(tgt[i]);
        double dX = hx[t] - ref[4 * i + 1], dY = hy[t] - ref[4 * i + 2], dZ = hz[t] - ref[4 * i + 3];
        double r  = std::sqrt(ref[4 * i + 1] * ref[4 * i + 1] + ref[4 * i + 2] * ref[4 * i + 2] +
                             ref[4 * i + 3] * ref[4 * i + 3]);
        err[i]    = std::sqrt(dX * dX + dY * dY + dZ * dZ) / r;
    }
    std::sort(err.begin(), err.end());
    if (argc > 3)
    {
        std::ofstream f(argv[3], std::ios::binary);
        const int64_t head[3] = {int64_t(n), int64_t(numNodes), int64_t(tb.numLeafNodes())};
        f.write(reinterpret_cast<const char*>(head), sizeof(head));
        for (const auto* v : {&x0, &y0, &z0})
            f.write(reinterpret_cast<const char*>(v->data()), n * sizeof(float));
        std::vector<float> hbuf(n);
        for (const float* d : {(const float*)dx, (const float*)dy, (const float*)dz, (const float*)dm,
                               (const float*)dh, (const float*)p, (const float*)ax, (const float*)ay,
                               (const float*)az})
        {
            HIPCHECK(hipMemcpy(hbuf.data(), d, n * sizeof(float), hipMemcpyDeviceToHost));
            f.write(reinterpret_cast<const char*>(hbuf.data()), n * sizeof(float));
        }
        std::vector<uint64_t> keys(numNodes);
        std::vector<int32_t> child(numNodes);
        std::vector<uint32_t> layout(tb.numLeafNodes() + 1);
        HIPCHECK(hipMemcpy(keys.data(), tb.nodeKeys(), numNodes * sizeof(uint64_t), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(child.data(), tb.childOffsets(), numNodes * sizeof(int32_t), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(layout.data(), tb.layout(), layout.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
        f.write(reinterpret_cast<const char*>(keys.data()), keys.size() * sizeof(uint64_t));
        f.write(reinterpret_cast<const char*>(child.data()), child.size() * sizeof(int32_t));
        f.write(reinterpret_cast<const char*>(layout.data()), layout.size() * sizeof(uint32_t));
        if (!f) return 2;
    }
    std::printf("%zu %d %.1f %.1f %.9f %.3e %.3e %.3e\n", n, numNodes, double(stats[0]) / n, double(stats[2]) / n, pot,
                err[nt / 2], err[size_t(0.99 * nt)], err[nt - 1]);
    return 0;
}
