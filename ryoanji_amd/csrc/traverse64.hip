// Barnes-Hut traversal for double-precision coordinates (gfx950): the reference's (Tc, Th, Tm, Ta, Tf) =
// (d, d, d, d, d) instantiation with CartesianQuadrupole<double> and (d, f, f, f, d) with
// CartesianQuadrupole<float> (traversal_gpu.cu:635-641). The walk is the float kernel's (one wavefront = one
// target group of up to 64 targets, 8 sibling cells per lane octet, a level-ordered LIFO in LDS, the MAC on the
// group's bounding box with unfused operations, accepted cells and close leaves applied with a wave-uniform
// source); every interaction is evaluated and accumulated in double (Tf), as the reference's acc_i is Vec4<Tf>
// (traversal_gpu.cu:467-470): P2P as kernel.hpp:99-120 (the softening sum h_i + h_j in Th), M2P as
// cartesian_qpole.hpp:152-176. Both take 1/R the way the reference's inverseSquareRoot(double) does
// (kernel.hpp:39-53): in single precision, as 1.0f / sqrtf(float(R^2)) with correctly rounded operations like its
// CPU path. One group per wavefront, no persistent scheduling: this path is not the FP32 hot path and favours the
// plain structure.
#include "common.hpp"

#include <cmath>

namespace ryo
{

void sumDeterministic(hipStream_t s, const double* v, int n, double* partials, double* out);

namespace
{

constexpr int kW64     = 4;       // waves per block
constexpr int kStack64 = 22 * 64; // LIFO entries per wave: 64 per tree level, never exceeded by a valid tree

template<class Tc, class Th, class Tm, class Ta, class Tq>
struct Trav64Args
{
    int numGroups;
    const uint32_t *gs, *ge;
    int initNode;
    const Tc *xt, *yt, *zt;
    const Tm* mt;
    const Th* ht;
    const Tc *xs, *ys, *zs;
    const Tm* ms;
    const Th* hs;
    const int32_t* child;
    const int32_t* i2l;
    const uint32_t* layout;
    const double* centers; // [n][4] {center of mass, mac^2}
    const Tq* mp;          // [n][8] {M, qxx, qxy, qxz, qyy, qyz, qzz, trace}, or [n][12] with px, py, pz, 0 (mdq)
    bool mdq;              // multipoles are CartesianMDQpole
    double G;
    int numShells;
    double boxL[3];
    Ta *p, *ax, *ay, *az;
    double* groupPot;
    unsigned long long* stats; // {sumP2P, maxP2P, sumM2P, maxM2P, maxStack, failed, corrupt}
};

__device__ inline double waveMin64(double v)
{
    for (int o = 32; o >= 1; o >>= 1)
        v = fmin(v, __shfl_xor(v, o));
    return v;
}
__device__ inline double waveMax64(double v)
{
    for (int o = 32; o >= 1; o >>= 1)
        v = fmax(v, __shfl_xor(v, o));
    return v;
}

// the MAC of traversal_gpu.cu:77-84 in double, unfused so that equal inputs give the CPU's decision
__device__ inline bool macClose64(const double* c, double tcx, double tcy, double tcz, double tsx, double tsy,
                                  double tsz)
{
#pragma clang fp contract(off)
    double dx = fabs(tcx - c[0]) - tsx;
    double dy = fabs(tcy - c[1]) - tsy;
    double dz = fabs(tcz - c[2]) - tsz;
    dx = (dx + fabs(dx)) * 0.5;
    dy = (dy + fabs(dy)) * 0.5;
    dz = (dz + fabs(dz)) * 0.5;
    const double R2 = dx * dx + (dy * dy + dz * dz);
    return R2 < fabs(c[3]);
}

struct Acc64
{
    double p, x, y, z;
};

// the reference's inverseSquareRoot for double arguments on its CPU path, 1.0f / std::sqrt(float(x)), with both
// operations correctly rounded: each is evaluated in double and rounded to float once (exact for square roots and
// quotients of floats, whose double results carry more than 2 x 24 + 2 bits); the single-precision v_sqrt_f32 is
// within about 1 ulp only
__device__ inline double invSqrtRef(double x)
{
    const float r = float(sqrt(double(float(x))));
    return double(float(1.0 / double(r)));
}

// softened P2P (kernel.hpp:99-120) with dX, R^2 in Tc and h_ij = h_i + h_j in Th
template<class Tc, class Th, class Tm>
__device__ inline void p2p64(Acc64& a, Tc tx, Tc ty, Tc tz, Th th, Tc sx, Tc sy, Tc sz, Tm sm, Th sh)
{
    const Tc dx = sx - tx, dy = sy - ty, dz = sz - tz;
    const Tc r2 = dx * dx + dy * dy + dz * dz;
    const Th hij  = th + sh;
    const Th hij2 = hij * hij;
    const Tc r2eff  = r2 < hij2 ? Tc(hij2) : r2;
    const Tc invR   = Tc(invSqrtRef(double(r2eff)));
    const Tc invR3m = Tc(sm) * invR * (invR * invR);
    a.p -= invR3m * r2;
    a.x += dx * invR3m;
    a.y += dy * invR3m;
    a.z += dz * invR3m;
}

// quadrupole M2P (cartesian_qpole.hpp:152-176) in double
template<class Tq>
__device__ inline void m2p64(Acc64& a, double tx, double ty, double tz, const double* c, const Tq* q)
{
    const double rx = tx - c[0], ry = ty - c[1], rz = tz - c[2];
    const double r2 = rx * rx + ry * ry + rz * rz;
    const double r1 = invSqrtRef(r2), ri2 = r1 * r1, r5 = ri2 * ri2 * r1;
    const double M = q[0], qxx = q[1], qxy = q[2], qxz = q[3], qyy = q[4], qyz = q[5], qzz = q[6];
    const double Qrx = rx * qxx + ry * qxy + rz * qxz;
    const double Qry = rx * qxy + ry * qyy + rz * qyz;
    const double Qrz = rx * qxz + ry * qyz + rz * qzz;
    const double rQr  = rx * Qrx + ry * Qry + rz * Qrz;
    const double mono = (-2.5 * rQr * r5 - M * r1) * ri2;
    a.p -= M * r1 + 0.5 * r5 * rQr;
    a.x += r5 * Qrx + mono * rx;
    a.y += r5 * Qry + mono * ry;
    a.z += r5 * Qrz + mono * rz;
}

// monopole-dipole-quadrupole M2P (cartesian_qpole.hpp:320-357) in double: the quadrupole terms plus the dipole
// p about the expansion center, potential -(p.r) / r^3 and field (p r^2 - 3 (p.r) r) / r^5
template<class Tq>
__device__ inline void m2pMDQ64(Acc64& a, double tx, double ty, double tz, const double* c, const Tq* q)
{
    const double rx = tx - c[0], ry = ty - c[1], rz = tz - c[2];
    const double r2 = rx * rx + ry * ry + rz * rz;
    const double r1 = invSqrtRef(r2), ri2 = r1 * r1, r5 = ri2 * ri2 * r1;
    const double M = q[0], qxx = q[1], qxy = q[2], qxz = q[3], qyy = q[4], qyz = q[5], qzz = q[6];
    const double px = q[8], py = q[9], pz = q[10];
    const double Qrx = rx * qxx + ry * qxy + rz * qxz;
    const double Qry = rx * qxy + ry * qyy + rz * qyz;
    const double Qrz = rx * qxz + ry * qyz + rz * qzz;
    const double rQr  = rx * Qrx + ry * Qry + rz * Qrz;
    const double mono = (-2.5 * rQr * r5 - M * r1) * ri2;
    const double PR   = rx * px + ry * py + rz * pz;
    a.p += (-M - PR * ri2) * r1 - 0.5 * r5 * rQr;
    a.x += r5 * Qrx + mono * rx + r5 * (px * r2 - 3.0 * PR * rx);
    a.y += r5 * Qry + mono * ry + r5 * (py * r2 - 3.0 * PR * ry);
    a.z += r5 * Qrz + mono * rz + r5 * (pz * r2 - 3.0 * PR * rz);
}

struct Walk64
{
    unsigned m2p, p2p, maxStack;
    bool overflow, corrupt;
};

template<class Tc, class Th, class Tm, class Ta, class Tq>
__device__ Walk64 walk64(const Trav64Args<Tc, Th, Tm, Ta, Tq>& a, int* stack, Tc tx, Tc ty, Tc tz, Th th,
                         double tcx, double tcy, double tcz, double tsx, double tsy, double tsz, Acc64& acc)
{
    const int lane    = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    Walk64 o{0, 0, 0, false, false};

    auto applyCells = [&](uint64_t amask, int cell) {
        o.m2p += __popcll(amask);
        while (amask)
        {
            const int l = __ffsll((unsigned long long)amask) - 1;
            amask &= amask - 1;
            const int c = __builtin_amdgcn_readlane(cell, l);
            if (a.mdq)
                m2pMDQ64(acc, double(tx), double(ty), double(tz), a.centers + 4 * size_t(c), a.mp + 12 * size_t(c));
            else
                m2p64(acc, double(tx), double(ty), double(tz), a.centers + 4 * size_t(c), a.mp + 8 * size_t(c));
        }
    };
    auto applyLeaves = [&](uint64_t dmask, uint32_t rb, uint32_t re) {
        while (dmask)
        {
            const int l = __ffsll((unsigned long long)dmask) - 1;
            dmask &= dmask - 1;
            const uint32_t b = uint32_t(__builtin_amdgcn_readlane(int(rb), l));
            const uint32_t e = uint32_t(__builtin_amdgcn_readlane(int(re), l));
            o.p2p += e - b;
            for (uint32_t j = b; j < e; ++j)
                p2p64(acc, tx, ty, tz, th, a.xs[j], a.ys[j], a.zs[j], a.ms[j], a.hs[j]);
        }
    };

    const int rootChild = a.child[0];
    if (a.initNode == 0 || rootChild == 0)
    {
        const double* c = a.centers;
        if (!macClose64(c, tcx, tcy, tcz, tsx, tsy, tsz))
        {
            if (c[3] != 0.0) applyCells(1ull, 0);
            return o;
        }
        if (rootChild == 0)
        {
            applyLeaves(1ull, a.layout[0], a.layout[1]);
            return o;
        }
    }
    if (lane == 0) stack[0] = rootChild;
    int sp    = 1;
    int steps = 0;
    while (sp > 0)
    {
        if (++steps > (1 << 24))
        {
            o.overflow = o.corrupt = true;
            return o;
        }
        const int k   = sp < 8 ? sp : 8;
        const int e   = lane >> 3;
        const bool ok = e < k;
        __builtin_amdgcn_wave_barrier();
        const int cell = ok ? stack[sp - k + e] + (lane & 7) : 0;
        sp -= k;

        const double* c  = a.centers + 4 * size_t(cell);
        const int ch     = a.child[cell];
        const bool close = ok && macClose64(c, tcx, tcy, tcz, tsx, tsy, tsz);
        const bool node  = ch != 0;
        uint32_t rb = 0, re = 0;
        if (close && !node)
        {
            const int li = a.i2l[cell];
            rb = a.layout[li], re = a.layout[li + 1];
        }
        // a leaf without local particles but with sources elsewhere (remote) is always approximated
        const bool remote = close && !node && rb == re && c[3] != 0.0;
        const bool split  = close && node;
        const bool approx = ok && ((!close && c[3] != 0.0) || remote);
        const bool direct = close && !node && rb != re;

        const uint64_t sm = __ballot(split);
        const int cnt     = __popcll(sm);
        if (sp + cnt > kStack64)
        {
            o.overflow = true;
            return o;
        }
        if (split) stack[sp + __popcll(sm & lt)] = ch;
        sp += cnt;
        o.maxStack = max(o.maxStack, unsigned(sp));
        applyCells(__ballot(approx), cell);
        applyLeaves(__ballot(direct), rb, re);
    }
    return o;
}

template<class Tc, class Th, class Tm, class Ta, class Tq>
__global__ __launch_bounds__(64 * kW64) void traverse64Kernel(Trav64Args<Tc, Th, Tm, Ta, Tq> a)
{
    __shared__ int stackMem[kW64 * kStack64];
    const int w    = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int g    = int(blockIdx.x) * kW64 + w;
    if (g >= a.numGroups) return;
    int* stack = stackMem + w * kStack64;

    const uint32_t b = a.gs[g], e = a.ge[g];
    if (e <= b)
    {
        if (lane == 0) a.groupPot[g] = 0.0;
        return;
    }
    const bool valid  = b + lane < e;
    const uint32_t ti = valid ? b + lane : e - 1;
    const Tc tx = a.xt[ti], ty = a.yt[ti], tz = a.zt[ti];
    const Th th = a.ht[ti];

    // target box in Tf = double (traversal_gpu.cu:455-468)
    const double mnx = waveMin64(double(tx)), mny = waveMin64(double(ty)), mnz = waveMin64(double(tz));
    const double mxx = waveMax64(double(tx)), mxy = waveMax64(double(ty)), mxz = waveMax64(double(tz));
    const double tcx = (mxx + mnx) * 0.5, tcy = (mxy + mny) * 0.5, tcz = (mxz + mnz) * 0.5;
    const double tsx = (mxx - mnx) * 0.5, tsy = (mxy - mny) * 0.5, tsz = (mxz - mnz) * 0.5;

    Acc64 acc{0, 0, 0, 0};
    unsigned nM2P = 0, nP2P = 0, maxStack = 0;
    bool overflow = false, corrupt = false;
    const int S = a.numShells;
    for (int iz = -S; iz <= S && !overflow; ++iz)
        for (int iy = -S; iy <= S && !overflow; ++iy)
            for (int ix = -S; ix <= S && !overflow; ++ix)
            {
                const double sx = ix * a.boxL[0], sy = iy * a.boxL[1], sz = iz * a.boxL[2];
                Walk64 o = walk64(a, stack, Tc(tx + sx), Tc(ty + sy), Tc(tz + sz), th, tcx + sx, tcy + sy, tcz + sz,
                                  tsx, tsy, tsz, acc);
                nM2P += o.m2p;
                nP2P += o.p2p;
                maxStack = max(maxStack, o.maxStack);
                overflow = o.overflow;
                corrupt  = o.corrupt;
            }
    if (overflow)
    {
        if (lane == 0)
        {
            atomicAdd(&a.stats[5], 1ull);
            if (corrupt) atomicAdd(&a.stats[6], 1ull);
        }
        return;
    }
    double myPot = 0;
    if (valid)
    {
        const double m = double(a.mt[ti]);
        if (a.p) a.p[ti] += Ta(a.G * m * acc.p);
        a.ax[ti] += Ta(a.G * acc.x);
        a.ay[ti] += Ta(a.G * acc.y);
        a.az[ti] += Ta(a.G * acc.z);
        myPot = m * acc.p;
    }
    myPot = waveSum(myPot);
    if (lane == 0)
    {
        a.groupPot[g] = myPot;
        const unsigned long long nt = e - b;
        atomicAdd(&a.stats[0], nt * nP2P);
        atomicMax(&a.stats[1], (unsigned long long)nP2P);
        atomicAdd(&a.stats[2], nt * nM2P);
        atomicMax(&a.stats[3], (unsigned long long)nM2P);
        atomicMax(&a.stats[4], (unsigned long long)maxStack);
    }
}

__global__ void zeroStats64Kernel(unsigned long long* s)
{
    if (threadIdx.x < 8) s[threadIdx.x] = 0;
}

template<class Tc, class Th, class Tm, class Ta, class Tq>
int traverse64(Context* ctx, Trav64Args<Tc, Th, Tm, Ta, Tq> a, double* totalPot)
{
    for (int k = 0; k < 6; ++k)
        ctx->hStats[k] = 0;
    ctx->retriedGroups = 0;
    if (a.numGroups <= 0)
    {
        if (totalPot) *totalPot = 0;
        return RYO_OK;
    }
    if (!a.gs || !a.ge || !a.xt || !a.yt || !a.zt || !a.mt || !a.ht || !a.xs || !a.ys || !a.zs || !a.ms || !a.hs ||
        !a.child || !a.i2l || !a.layout || !a.centers || !a.mp || !a.ax || !a.ay || !a.az)
        return setError(ctx, RYO_ERR_ARG, "traverse (double): null pointer");
    hipStream_t s = ctx->stream;
    const size_t G = size_t(a.numGroups);
    const size_t partOff = (8 * G + 255) & ~size_t(255);
    auto* buf = static_cast<char*>(ctx->getScratch(kSlotTraverse, partOff + 256 * sizeof(double)));
    if (!buf) return RYO_ERR_HIP;
    a.groupPot       = reinterpret_cast<double*>(buf);
    double* partials = reinterpret_cast<double*>(buf + partOff);
    a.stats          = ctx->dStats + 8;
    double* potSum   = reinterpret_cast<double*>(ctx->dStats + 16);
    zeroStats64Kernel<<<1, 64, 0, s>>>(a.stats);
    if (ctx->timing) RYO_HIP(ctx, hipEventRecord(ctx->ev[0], s));
    traverse64Kernel<<<unsigned(iceil(G, kW64)), 64 * kW64, 0, s>>>(a);
    if (ctx->timing) RYO_HIP(ctx, hipEventRecord(ctx->ev[1], s));
    RYO_LAUNCH_CHECK(ctx);
    sumDeterministic(s, a.groupPot, a.numGroups, partials, potSum);
    unsigned long long h[7] = {0, 0, 0, 0, 0, 0, 0};
    double pot = 0;
    RYO_HIP(ctx, hipMemcpyAsync(h, a.stats, sizeof(h), hipMemcpyDeviceToHost, s));
    RYO_HIP(ctx, hipMemcpyAsync(&pot, potSum, sizeof(double), hipMemcpyDeviceToHost, s));
    RYO_HIP(ctx, hipStreamSynchronize(s));
    if (ctx->timing)
    {
        float ms = 0;
        RYO_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        ctx->lastTraverseMs = ms;
    }
    if (h[6]) return setError(ctx, RYO_ERR_TRAVERSAL, "traverse (double): step limit exceeded (corrupt tree?)");
    if (h[5]) return setError(ctx, RYO_ERR_TRAVERSAL, "traverse (double): %llu groups exceeded the LIFO", h[5]);
    for (int k = 0; k < 5; ++k)
        ctx->hStats[k] = h[k];
    if (totalPot) *totalPot = 0.5 * a.G * pot;
    return RYO_OK;
}

template<class Tc, class Th, class Tm, class Ta, class Tq>
Trav64Args<Tc, Th, Tm, Ta, Tq> makeArgs(int numGroups, const uint32_t* gs, const uint32_t* ge, int initNode,
                                        const Tc* xt, const Tc* yt, const Tc* zt, const Tm* mt, const Th* ht,
                                        const Tc* xs, const Tc* ys, const Tc* zs, const Tm* ms, const Th* hs,
                                        const int32_t* child, const int32_t* i2l, const uint32_t* layout,
                                        const double* centers, const Tq* mp, double G, int numShells,
                                        const double* boxL, Ta* p, Ta* ax, Ta* ay, Ta* az, bool mdq = false)
{
    Trav64Args<Tc, Th, Tm, Ta, Tq> a{};
    a.numGroups = numGroups, a.gs = gs, a.ge = ge, a.initNode = initNode;
    a.xt = xt, a.yt = yt, a.zt = zt, a.mt = mt, a.ht = ht;
    a.xs = xs, a.ys = ys, a.zs = zs, a.ms = ms, a.hs = hs;
    a.child = child, a.i2l = i2l, a.layout = layout, a.centers = centers, a.mp = mp;
    a.G = G, a.numShells = numShells;
    for (int d = 0; d < 3; ++d)
        a.boxL[d] = boxL ? boxL[d] : 0.0;
    a.p = p, a.ax = ax, a.ay = ay, a.az = az;
    a.mdq = mdq;
    return a;
}

} // namespace

} // namespace ryo

using namespace ryo;

extern "C" {

int ryo_traverse_f64(ryo_ctx* h, int num_groups, const uint32_t* group_start, const uint32_t* group_end,
                     int init_node, const double* xt, const double* yt, const double* zt, const double* mt,
                     const double* ht, const double* xs, const double* ys, const double* zs, const double* ms,
                     const double* hs, const int32_t* child_offsets, const int32_t* internal_to_leaf,
                     const uint32_t* layout, const double* centers, const double* multipoles, double G,
                     int num_shells, const double* box_l, double* p, double* ax, double* ay, double* az,
                     double* total_pot)
{
    if (!h) return RYO_ERR_ARG;
    Context* ctx = &h->c;
    if (init_node != 0 && init_node != 1) return setError(ctx, RYO_ERR_ARG, "traverse: init_node must be 0 or 1");
    if (num_shells < 0 || (num_shells > 0 && !box_l)) return setError(ctx, RYO_ERR_ARG, "traverse: box lengths");
    return traverse64(ctx,
                      makeArgs(num_groups, group_start, group_end, init_node, xt, yt, zt, mt, ht, xs, ys, zs, ms, hs,
                               child_offsets, internal_to_leaf, layout, centers, multipoles, G, num_shells, box_l, p,
                               ax, ay, az),
                      total_pot);
}

int ryo_traverse_f64_f32(ryo_ctx* h, int num_groups, const uint32_t* group_start, const uint32_t* group_end,
                         int init_node, const double* xt, const double* yt, const double* zt, const float* mt,
                         const float* ht, const double* xs, const double* ys, const double* zs, const float* ms,
                         const float* hs, const int32_t* child_offsets, const int32_t* internal_to_leaf,
                         const uint32_t* layout, const double* centers, const float* multipoles, double G,
                         int num_shells, const double* box_l, float* p, float* ax, float* ay, float* az,
                         double* total_pot)
{
    if (!h) return RYO_ERR_ARG;
    Context* ctx = &h->c;
    if (init_node != 0 && init_node != 1) return setError(ctx, RYO_ERR_ARG, "traverse: init_node must be 0 or 1");
    if (num_shells < 0 || (num_shells > 0 && !box_l)) return setError(ctx, RYO_ERR_ARG, "traverse: box lengths");
    return traverse64(ctx,
                      makeArgs(num_groups, group_start, group_end, init_node, xt, yt, zt, mt, ht, xs, ys, zs, ms, hs,
                               child_offsets, internal_to_leaf, layout, centers, multipoles, G, num_shells, box_l, p,
                               ax, ay, az),
                      total_pot);
}

int ryo_traverse_mdq_f64(ryo_ctx* h, int num_groups, const uint32_t* group_start, const uint32_t* group_end,
                         int init_node, const double* xt, const double* yt, const double* zt, const double* mt,
                         const double* ht, const double* xs, const double* ys, const double* zs, const double* ms,
                         const double* hs, const int32_t* child_offsets, const int32_t* internal_to_leaf,
                         const uint32_t* layout, const double* centers, const double* multipoles, double G,
                         int num_shells, const double* box_l, double* p, double* ax, double* ay, double* az,
                         double* total_pot)
{
    if (!h) return RYO_ERR_ARG;
    Context* ctx = &h->c;
    if (init_node != 0 && init_node != 1) return setError(ctx, RYO_ERR_ARG, "traverse: init_node must be 0 or 1");
    if (num_shells < 0 || (num_shells > 0 && !box_l)) return setError(ctx, RYO_ERR_ARG, "traverse: box lengths");
    return traverse64(ctx,
                      makeArgs(num_groups, group_start, group_end, init_node, xt, yt, zt, mt, ht, xs, ys, zs, ms, hs,
                               child_offsets, internal_to_leaf, layout, centers, multipoles, G, num_shells, box_l, p,
                               ax, ay, az, true),
                      total_pot);
}

int ryo_traverse_mdq_f64_f32(ryo_ctx* h, int num_groups, const uint32_t* group_start, const uint32_t* group_end,
                             int init_node, const double* xt, const double* yt, const double* zt, const float* mt,
                             const float* ht, const double* xs, const double* ys, const double* zs, const float* ms,
                             const float* hs, const int32_t* child_offsets, const int32_t* internal_to_leaf,
                             const uint32_t* layout, const double* centers, const float* multipoles, double G,
                             int num_shells, const double* box_l, float* p, float* ax, float* ay, float* az,
                             double* total_pot)
{
    if (!h) return RYO_ERR_ARG;
    Context* ctx = &h->c;
    if (init_node != 0 && init_node != 1) return setError(ctx, RYO_ERR_ARG, "traverse: init_node must be 0 or 1");
    if (num_shells < 0 || (num_shells > 0 && !box_l)) return setError(ctx, RYO_ERR_ARG, "traverse: box lengths");
    return traverse64(ctx,
                      makeArgs(num_groups, group_start, group_end, init_node, xt, yt, zt, mt, ht, xs, ys, zs, ms, hs,
                               child_offsets, internal_to_leaf, layout, centers, multipoles, G, num_shells, box_l, p,
                               ax, ay, az, true),
                      total_pot);
}

} // extern "C"
