// FP64 Ewald periodic correction and the FP64 direct-sum accuracy reference (gfx950).
//
// Ewald: the correction from the root quadrupole of a cubic periodic box (ryoanji/nbody/ewald.hpp:90-327):
// real-space images within lCut (full erfc outside the replica shells, -erf inside them, a series for small
// R) and a k-space sum over the lattice vectors |h| <= hCut. One particle per lane; every term in FP64; the
// k-space table is built on the host (ewaldInitParameters, ewald.hpp:133-198) and staged in LDS.
// Direct sum: P2P<Exact> (kernel.hpp:99-120) over all sources and images, accumulated in FP64, for a
// list of sampled targets (the accuracy oracle; direct.cuh:36-107 is its FP32 reference counterpart).
#include "common.hpp"
#include "erfcx_table.hpp"

#include <cmath>

namespace ryo
{

constexpr double kTwoOverSqrtPi = 1.1283791670955126;
constexpr int kMaxCullShells = 8; // image rows are culled and exponentials factored for up to 8 shells
constexpr int kMaxRows       = (2 * kMaxCullShells + 1) * (2 * kMaxCullShells + 1);
constexpr int kMaxHReps      = 3; // ceil(hCut) <= 3, the reference's maxCeilHcut (ewald.hpp:53)
constexpr int kMaxKPairs     = ((2 * kMaxHReps + 1) * (2 * kMaxHReps + 1) * (2 * kMaxHReps + 1) - 1) / 2;

// the root quadrupole in the form its evaluation uses: M, the traceless components plus trace over 3, and half
// the trace (ewald.hpp:97-108)
struct EwaldQ
{
    double M, qxx, qyy, qzz, qxy, qxz, qyz, Qtr;
};

struct EwaldTable
{
    double Mroot[8];
    EwaldQ q; // derived from Mroot once, on the host
    double center[3];
    int numReplicaShells, numEwaldShells, numKPairs, hReps;
    double L, lCut, alphaScale, smallR, hCut2;
    double kPotSign; // -1: the reference's k-space potential (ewald.hpp:320); +1: consistent with the acceleration
    // k space, one entry per pair {h, -h} of lattice vectors with 0 < |h|^2 <= hCut^2: the coefficient e(h) of
    // the reference's table (ewald.hpp:185-191, hfac_cos) is even in h and the sine coefficient drops out of the
    // pair, so the pair contributes 2 e(h) kPotSign cos(k.r) to the potential and 2 e(h) sin(k.r) k to the
    // acceleration, k = 2 pi h / L. Entries {2 e kPotSign, 2 e k_x, 2 e k_y, 2 e k_z} in the order the kernel
    // visits the half space (hx ascending from 0, hy, hz ascending; hx = 0 only for hy > 0, or hy = 0 and hz > 0)
    double kc[kMaxKPairs][4];
    // exp(-a^2 |r + nL|^2) = exp(-a^2 |r|^2) prod_d exp(-2 a^2 L r_d)^n_d exp(-a^2 L^2 n_d^2): expLat[i] is
    // exp(-a^2 L^2 i^2) and expStep[i + ns] = exp(-a^2 L^2 (2 i + 1)), the ratio of consecutive lattice factors;
    // used when the factors stay far from overflow (productExp)
    int productExp;
    double expLat[kMaxCullShells + 1];
    double expStep[2 * kMaxCullShells + 1];
    // k space: row (hx, hy) holds the vectors hz in [-kRowHz, kRowHz] (none when -1), the origin excluded
    int8_t kRowHz[(2 * kMaxHReps + 1) * (2 * kMaxHReps + 1)];
};

// 1/sqrt(x) for positive normal x: v_rsq_f64 (relative error below 2^-22) and one Newton step, within about
// 1e-13 relative (the Ewald terms are summed in FP64 and stored as float)
__device__ inline double rsqrt64(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-0.5 * x * y, y, 0.5);
    return fma(y, e, y);
}

// gamma-weighted quadrupole evaluation (ewaldEvalMultipoleComplete, ewald.hpp:90-115)
__host__ __device__ inline void ewaldEval(double* pa, const double* r, const double* g, const double* mp)
{
    double M   = mp[0];
    double qxx = (mp[1] + mp[7]) / 3.0, qyy = (mp[4] + mp[7]) / 3.0, qzz = (mp[6] + mp[7]) / 3.0;
    double qxy = mp[2] / 3.0, qxz = mp[3] / 3.0, qyz = mp[5] / 3.0;
    double Qr[3] = {r[0] * qxx + r[1] * qxy + r[2] * qxz, r[0] * qxy + r[1] * qyy + r[2] * qyz,
                    r[0] * qxz + r[1] * qyz + r[2] * qzz};
    double rQr = 0.5 * (r[0] * Qr[0] + r[1] * Qr[1] + r[2] * Qr[2]);
    double Qtr = 0.5 * mp[7];
    pa[0] += -g[0] * M + g[1] * Qtr - g[2] * rQr;
    double f = g[1] * M - g[2] * Qtr + g[3] * rQr;
    for (int d = 0; d < 3; ++d)
        pa[1 + d] += g[2] * Qr[d] - r[d] * f;
}

// erfcx(x) = exp(x^2) erfc(x) for x >= 0, given y = 4 x: on [0, 8] the degree-11 piece of x's interval of width
// 1/4 (erfcx_table.hpp, within 2.5e-16 of erfcx), in the local variable t = 2 fract(y) - 1 in [-1, 1). Each lane
// reads its interval's coefficients from the copy of the table in LDS (c2: pairs of coefficients, constant term
// first); a wave's targets are SFC neighbours that see an image at nearly the same distance, so the reads are
// mostly broadcasts. Beyond 8 the last piece is evaluated at a point of [7.75, 8), so erfc(x) = exp(-x^2)
// erfcx(x) comes out below 1.2e-29 and within the factor x / 8 of its value, far below the FP64 rounding of the
// replica terms -erf(aR) / R (of order 1 / R) that every lane sums with it.
constexpr int kErfcxPairs = kErfcxTerms / 2;
static_assert(kErfcxTerms % 2 == 0, "coefficients are read in pairs");

__device__ inline double erfcxLds(double y, const double2* __restrict__ c2)
{
    const int j      = min(int(y), kErfcxIntervals - 1);
    const double2* c = c2 + j * kErfcxPairs;
    double2 p[kErfcxPairs];
    for (int k = 0; k < kErfcxPairs; ++k)
        p[k] = c[k];
    const double t = fma(2.0, __builtin_amdgcn_fract(y), -1.0);
    double v       = p[kErfcxPairs - 1].y;
    v              = fma(v, t, p[kErfcxPairs - 1].x);
    for (int k = kErfcxPairs - 2; k >= 0; --k)
    {
        v = fma(v, t, p[k].y);
        v = fma(v, t, p[k].x);
    }
    return v;
}

__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 conjIf(double2 a, bool c) { return c ? make_double2(a.x, -a.y) : a; }

// smallest |v + off| over v in [lo, hi] (0 when the interval shifted by off contains 0)
__device__ inline double boxAxisDist(double lo, double hi, double off)
{
    const double a = lo + off, b = hi + off;
    return (a <= 0.0 && 0.0 <= b) ? 0.0 : fmin(fabs(a), fabs(b));
}

// the lattice factors of the image Gaussians; the factorization is used when every factor stays far from overflow
// over the image shells
void ewaldInitFactors(EwaldTable& t)
{
    const int ns    = t.numEwaldShells;
    const double a2 = t.alphaScale * t.alphaScale;
    t.productExp    = ns <= kMaxCullShells && 2.0 * a2 * ns * 1.1 <= 200.0 && a2 * ns * ns <= 600.0;
    for (int i = 0; i <= kMaxCullShells; ++i)
        t.expLat[i] = i <= ns ? std::exp(-a2 * i * i) : 0.0;
    for (int i = -kMaxCullShells; i <= kMaxCullShells; ++i)
        t.expStep[i + kMaxCullShells] = (t.productExp && -ns <= i && i <= ns) ? std::exp(-a2 * (2 * i + 1)) : 0.0;
}

// the visiting order of the k-space pairs: row (hx, hy) of the half space with its largest |hz| (or -1)
template<class F>
void forEachKPair(const EwaldTable& t, F&& f)
{
    const int hr = t.hReps;
    for (int hx = 0; hx <= hr; ++hx)
        for (int hy = -hr; hy <= hr; ++hy)
        {
            if (hx == 0 && hy < 0) continue;
            const int hzm = t.kRowHz[(hx + hr) * (2 * hr + 1) + (hy + hr)];
            for (int hz = -hr; hz <= hr; ++hz)
            {
                if (hx == 0 && hy == 0 && hz <= 0) continue;
                if (std::abs(hz) > hzm) continue;
                f(hx, hy, hz);
            }
        }
}

void ewaldInitTable(EwaldTable& t, const double* Mroot, const double* center, int numReplicaShells, double L,
                    double lCut, double hCut, double alphaScale, double smallR, double kPotSign)
{
    t.q = EwaldQ{Mroot[0],       (Mroot[1] + Mroot[7]) / 3.0, (Mroot[4] + Mroot[7]) / 3.0,
                 (Mroot[6] + Mroot[7]) / 3.0, Mroot[2] / 3.0,  Mroot[3] / 3.0,
                 Mroot[5] / 3.0, 0.5 * Mroot[7]};
    if (lCut == 0 && hCut == 0 && alphaScale == 0) numReplicaShells = 0;
    for (int k = 0; k < 8; ++k)
        t.Mroot[k] = Mroot[k];
    for (int k = 0; k < 3; ++k)
        t.center[k] = center[k];
    t.kPotSign         = kPotSign;
    t.numReplicaShells = numReplicaShells;
    t.numEwaldShells   = std::max(int(std::ceil(lCut)), numReplicaShells);
    t.L = L, t.lCut = lCut, t.alphaScale = alphaScale, t.smallR = smallR;
    t.numKPairs = 0;
    t.hReps     = 0;
    if (t.numEwaldShells == 0) return;
    const int hReps = int(std::ceil(hCut));
    t.hReps         = hReps;
    t.hCut2         = hCut * hCut;
    for (int hx = -hReps; hx <= hReps; hx++)
        for (int hy = -hReps; hy <= hReps; hy++)
        {
            int top = -1;
            for (int hz = 0; hz <= hReps; ++hz)
                if (double(hx * hx + hy * hy + hz * hz) <= hCut * hCut) top = hz;
            t.kRowHz[(hx + hReps) * (2 * hReps + 1) + (hy + hReps)] = int8_t(top);
        }
    const double alpha = alphaScale / L;
    const double k4    = M_PI * M_PI / (alpha * alpha * L * L);
    forEachKPair(t, [&](int hx, int hy, int hz) {
        double hr[3] = {double(hx), double(hy), double(hz)};
        double h2    = hr[0] * hr[0] + hr[1] * hr[1] + hr[2] * hr[2];
        double g0    = std::exp(-k4 * h2) / (M_PI * h2 * L);
        double g1    = 2 * M_PI / L * g0;
        double g2    = -2 * M_PI / L * g1;
        double g4    = 2 * M_PI / L * (-2 * M_PI / L * g2);
        double gc[6] = {g0, 0, g2, 0, g4, 0};
        double a[4]  = {0, 0, 0, 0};
        ewaldEval(a, hr, gc, t.Mroot);
        double* e = t.kc[t.numKPairs++];
        e[0]      = 2 * a[0] * kPotSign;
        for (int d = 0; d < 3; ++d)
            e[1 + d] = 2 * a[0] * (2 * M_PI / L * hr[d]);
    });
}

// k space over the pairs {h, -h} of the table (|h_d| <= HR = hReps): exp(i k.r) as a product of the axis phases
// exp(i 2 pi h_d r_d / L), whose powers 0..HR every lane holds in registers; the loops unroll completely, so
// each phase is a fixed register and the table's row bounds are wave-uniform branches
template<int HR>
__device__ inline void ewaldKSpace(const EwaldTable* __restrict__ tab, const double* r, double* kp)
{
    double2 px[HR + 1], py[HR + 1], pz[HR + 1];
    {
        const double w = 2.0 * M_PI / tab->L;
        double s, c;
        px[0] = py[0] = pz[0] = make_double2(1.0, 0.0);
        sincos(w * r[0], &s, &c);
        px[1] = make_double2(c, s);
        sincos(w * r[1], &s, &c);
        py[1] = make_double2(c, s);
        sincos(w * r[2], &s, &c);
        pz[1] = make_double2(c, s);
#pragma unroll
        for (int j = 2; j <= HR; ++j)
            px[j] = cmul(px[j - 1], px[1]), py[j] = cmul(py[j - 1], py[1]), pz[j] = cmul(pz[j - 1], pz[1]);
    }
    int k = 0;
#pragma unroll
    for (int hx = 0; hx <= HR; ++hx)
#pragma unroll
        for (int hy = -HR; hy <= HR; ++hy)
        {
            if (hx == 0 && hy < 0) continue;
            const int hzm = tab->kRowHz[(hx + HR) * (2 * HR + 1) + (hy + HR)];
            if (hzm < 0) continue;
            const double2 qy  = conjIf(py[hy < 0 ? -hy : hy], hy < 0);
            const double2 exy = hx == 0 ? qy : (hy == 0 ? px[hx] : cmul(px[hx], qy));
#pragma unroll
            for (int hz = -HR; hz <= HR; ++hz)
            {
                if (hx == 0 && hy == 0 && hz <= 0) continue;
                if ((hz < 0 ? -hz : hz) > hzm) continue;
                const double2 ph = hz == 0 ? exy : cmul(exy, conjIf(pz[hz < 0 ? -hz : hz], hz < 0));
                const double* e  = tab->kc[k++];
                kp[0]            = fma(e[0], ph.x, kp[0]);
                kp[1]            = fma(e[1], ph.y, kp[1]);
                kp[2]            = fma(e[2], ph.y, kp[2]);
                kp[3]            = fma(e[3], ph.y, kp[3]);
            }
        }
}

// RowTable: the per-wave row bounds fit the LDS table (at most kMaxCullShells image shells); ProdExp: the image
// Gaussians come from products of per-axis factors (EwaldTable::productExp)
// one block of 256 particles: particle i of the block's thread (valid: i < last); the block's sum of m phi into
// *blockPot
template<bool RowTable, bool ProdExp>
__device__ __forceinline__ void ewaldBlock(const EwaldTable* __restrict__ tab, uint32_t i, uint32_t last,
                                           const float* __restrict__ x, const float* __restrict__ y,
                                           const float* __restrict__ z, const float* __restrict__ m, float G,
                                           float* ugrav, float* ax, float* ay, float* az,
                                           double* __restrict__ blockPot)
{
    // the table is read with wave-uniform addresses (scalar loads), so the quadrupole and the k-space
    // coefficients live in scalar registers
    __shared__ double sPot[256];
    __shared__ double2 sErfcx[kErfcxIntervals * kErfcxPairs];
    __shared__ short2 sRows[4][kMaxRows]; // per wave: the iz range {lo, hi} of each image row (ix, iy)
    for (int k = threadIdx.x; k < kErfcxIntervals * kErfcxPairs; k += blockDim.x)
        sErfcx[k] = make_double2(kErfcx[k / kErfcxPairs][2 * (k % kErfcxPairs)],
                                 kErfcx[k / kErfcxPairs][2 * (k % kErfcxPairs) + 1]);
    __syncthreads();
    const EwaldQ q  = tab->q;

    const double L = tab->L, lCut = tab->lCut;
    const double lCut2  = lCut * lCut * L * L;
    const double alpha  = tab->alphaScale / L, alpha2 = alpha * alpha;
    const double alpha4 = 4.0 * alpha, twoA2 = 2.0 * alpha2;
    const double k1     = M_PI / (alpha2 * L * L * L);
    const double M      = tab->Mroot[0];
    const double ka     = 2.0 * alpha / sqrt(M_PI);
    const double smallR2 = tab->smallR * L * L;
    const int ns = tab->numEwaldShells, nr = tab->numReplicaShells;

    const bool valid = i < last;
    double myPot     = 0;
    // the wave's bounding box of r = x - center over its valid lanes (targets are SFC neighbours, so the box is
    // small): an image row or slab is visited only where some point of the box can be within lCut of it (or it
    // is an inner replica); the per-lane test below still decides every term, so the evaluated set is exactly
    // the reference's
    double r[3] = {0, 0, 0};
    if (valid) r[0] = x[i] - tab->center[0], r[1] = y[i] - tab->center[1], r[2] = z[i] - tab->center[2];
    double blo[3], bhi[3];
    for (int d = 0; d < 3; ++d)
    {
        blo[d] = valid ? r[d] : HUGE_VAL;
        bhi[d] = valid ? r[d] : -HUGE_VAL;
        for (int o = 32; o >= 1; o >>= 1)
        {
            blo[d] = fmin(blo[d], __shfl_xor(blo[d], o));
            bhi[d] = fmax(bhi[d], __shfl_xor(bhi[d], o));
        }
    }
    // the iz range of every image row (ix, iy) for this wave's box, one row per lane (rows of a wave are
    // wave-uniform, so the loop below reads them as scalars); a conservative cut never below any lane's exact
    // lCut^2 test. Waves without valid lanes get empty rows (their box is empty).
    const int w          = 2 * ns + 1;
    const int numRows    = w * w;
    constexpr bool rowTable = RowTable;
    const double cutB    = lCut2 * (1.0 + 1e-9), invL = 1.0 / L;
    const int wave       = threadIdx.x >> 6;
    if (rowTable)
        for (int row = threadIdx.x & 63; row < numRows; row += 64)
        {
            const int ix = row / w - ns, iy = row % w - ns;
            const double dx = boxAxisDist(blo[0], bhi[0], ix * L), dy = boxAxisDist(blo[1], bhi[1], iy * L);
            const double rem = cutB - dx * dx - dy * dy;
            int zlo = ns + 1, zhi = -ns - 1;
            if (rem >= 0)
            {
                const double rad = sqrt(rem);
                zlo = int(ceil((-rad - bhi[2]) * invL - 1e-9));
                zhi = int(floor((rad - blo[2]) * invL + 1e-9));
            }
            if (-nr <= ix && ix <= nr && -nr <= iy && iy <= nr) zlo = min(zlo, -nr), zhi = max(zhi, nr);
            sRows[wave][row] = make_short2(short(max(zlo, -ns)), short(min(zhi, ns)));
        }
    __syncthreads();
    if (valid)
    {
        double pa[4] = {k1 * M, 0, 0, 0};
        constexpr bool prodExp = ProdExp;
        // the exponential factors of the images (productExp): px_ = exp(-2 a^2 L r_x)^ix along the loop
        double e0 = 0, fx = 0, fy = 0, fz = 0, px_ = 0, py0 = 0, pz0 = 0;
        if (prodExp)
        {
            const double c = 2.0 * alpha2 * L;
            e0  = exp(-alpha2 * (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]));
            fx  = exp(-c * r[0]), fy = exp(-c * r[1]), fz = exp(-c * r[2]);
            px_ = exp(c * ns * r[0]), py0 = exp(c * ns * r[1]), pz0 = exp(c * ns * r[2]);
        }
        const double stepK = prodExp ? tab->expLat[1] * tab->expLat[1] : 0.0; // exp(-2 a^2 L^2)
        const double hqzz  = 0.5 * q.qzz;
        for (int ix = -ns; ix <= ns; ix++, px_ *= fx)
        {
            const bool innerX = -nr <= ix && ix <= nr;
            double dx         = 0;
            if (!rowTable)
            {
                dx = boxAxisDist(blo[0], bhi[0], ix * L);
                if (!innerX && dx * dx > cutB) continue;
            }
            const double ex = prodExp ? e0 * px_ * tab->expLat[ix < 0 ? -ix : ix] : 0.0;
            const double Rx = fma(double(ix), L, r[0]);
            double py_      = py0;
            for (int iy = -ns; iy <= ns; iy++, py_ *= fy)
            {
                const bool innerRow = innerX && (-nr <= iy && iy <= nr);
                int zlo, zhi;
                if (rowTable)
                {
                    const short2 b = sRows[wave][(ix + ns) * w + (iy + ns)];
                    zlo = __builtin_amdgcn_readfirstlane(int(b.x));
                    zhi = __builtin_amdgcn_readfirstlane(int(b.y));
                }
                else
                {
                    const double dy  = boxAxisDist(blo[1], bhi[1], iy * L);
                    const double rem = cutB - dx * dx - dy * dy;
                    zlo = ns + 1, zhi = -ns - 1;
                    if (rem >= 0)
                    {
                        const double rad = sqrt(rem);
                        zlo = int(ceil((-rad - bhi[2]) * invL - 1e-9));
                        zhi = int(floor((rad - blo[2]) * invL + 1e-9));
                    }
                    if (innerRow) zlo = min(zlo, -nr), zhi = max(zhi, nr);
                    // wave-uniform bounds (every lane computed them from the same box): scalar loop control
                    zlo = __builtin_amdgcn_readfirstlane(max(zlo, -ns));
                    zhi = __builtin_amdgcn_readfirstlane(min(zhi, ns));
                }
                if (zlo > zhi) continue;
                // per row: R_x, R_y, the xy part of |R|^2, of Q R and of R.Q.R / 2
                const double Ry   = fma(double(iy), L, r[1]);
                const double Rxy2 = fma(Ry, Ry, Rx * Rx);
                const double Qx   = fma(q.qxy, Ry, q.qxx * Rx);
                const double Qy   = fma(q.qyy, Ry, q.qxy * Rx);
                const double Qz   = fma(q.qyz, Ry, q.qxz * Rx);
                const double c0   = 0.5 * fma(Ry, Qy, Rx * Qx);
                // the Gaussian of the row's first image and its ratio to the next: E(iz + 1) = E(iz) cz,
                // cz(iz + 1) = cz(iz) exp(-2 a^2 L^2)
                double gz = 0, cz = 0;
                if (prodExp)
                {
                    double pz_ = pz0;
                    for (int iz = -ns; iz < zlo; ++iz)
                        pz_ *= fz;
                    gz = ex * py_ * tab->expLat[iy < 0 ? -iy : iy] * (pz_ * tab->expLat[zlo < 0 ? -zlo : zlo]);
                    cz = fz * tab->expStep[zlo + kMaxCullShells];
                }
                double Rz = fma(double(zlo), L, r[2]);
                for (int iz = zlo; iz <= zhi; iz++)
                {
                    const bool inner  = innerRow && (-nr <= iz && iz <= nr);
                    const double Rzi  = Rz;
                    const double R2   = fma(Rzi, Rzi, Rxy2);
                    const double e2   = gz;
                    Rz += L;
                    if (prodExp) gz *= cz, cz *= stepK;
                    if (R2 > lCut2 && !inner) continue;
                    double g0, g1, g2, g3;
                    if (R2 < smallR2 && ka > 0)
                    {
                        double c0s = ka, R2a2 = R2 * alpha2;
                        g0 = c0s * (R2a2 / 3.0 - 1.0);
                        c0s *= twoA2;
                        g1 = c0s * (R2a2 / 5.0 - 1.0 / 3.0);
                        c0s *= twoA2;
                        g2 = c0s * (R2a2 / 7.0 - 1.0 / 5.0);
                        c0s *= twoA2;
                        g3 = c0s * (R2a2 / 9.0 - 1.0 / 7.0);
                    }
                    else
                    {
                        const double invR = rsqrt64(R2), Rm = R2 * invR, invR2 = invR * invR;
                        const double eg   = prodExp ? e2 : exp(-R2 * alpha2);
                        const double yv   = alpha4 * Rm; // 4 alpha R
                        const double ex2  = erfcxLds(yv, sErfcx);
                        // erfc(aR) = exp(-a^2 R^2) erfcx(aR) reusing the Gaussian above; -erf = erfc - 1 for the
                        // replicas the traversal already holds
                        const double off = inner ? 1.0 : 0.0;
                        const double fn  = eg * ex2 - off;
                        const double a  = (eg * ka) * invR2;
                        const double a1 = a * twoA2;
                        g0 = fn * invR;
                        g1 = fma(g0, invR2, a);
                        g2 = fma(g1, 3.0 * invR2, a1);
                        g3 = fma(g2, 5.0 * invR2, a1 * twoA2);
                    }
                    // ewaldEvalQ at R = (Rx, Ry, Rz): Q R, R.Q.R / 2 from the row's parts
                    const double Qrx = fma(q.qxz, Rzi, Qx);
                    const double Qry = fma(q.qyz, Rzi, Qy);
                    const double Qrz = fma(q.qzz, Rzi, Qz);
                    const double rQr = fma(Rzi, fma(hqzz, Rzi, Qz), c0);
                    pa[0]            = fma(-g0, q.M, pa[0]);
                    pa[0]            = fma(g1, q.Qtr, pa[0]);
                    pa[0]            = fma(-g2, rQr, pa[0]);
                    const double f   = fma(g3, rQr, fma(g1, q.M, -g2 * q.Qtr));
                    pa[1]            = fma(-Rx, f, fma(g2, Qrx, pa[1]));
                    pa[2]            = fma(-Ry, f, fma(g2, Qry, pa[2]));
                    pa[3]            = fma(-Rzi, f, fma(g2, Qrz, pa[3]));
                }
            }
        }
        double kp[4] = {0, 0, 0, 0};
        switch (tab->hReps)
        {
            case 1: ewaldKSpace<1>(tab, r, kp); break;
            case 2: ewaldKSpace<2>(tab, r, kp); break;
            case 3: ewaldKSpace<3>(tab, r, kp); break;
            default: break;
        }
        double pot = pa[0] + kp[0];
        double mi  = m[i];
        if (ugrav) ugrav[i] += float(double(G) * pot * mi);
        ax[i] += float(double(G) * (pa[1] + kp[1]));
        ay[i] += float(double(G) * (pa[2] + kp[2]));
        az[i] += float(double(G) * (pa[3] + kp[3]));
        myPot = pot * mi;
    }
    sPot[threadIdx.x] = myPot;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if (int(threadIdx.x) < o) sPot[threadIdx.x] += sPot[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *blockPot = sPot[0];
}

template<bool RowTable, bool ProdExp>
__global__ __launch_bounds__(256, 4) void ewaldKernel(const EwaldTable* __restrict__ tab, uint32_t first, uint32_t last,
                                                  const float* __restrict__ x, const float* __restrict__ y,
                                                  const float* __restrict__ z, const float* __restrict__ m, float G,
                                                  float* ugrav, float* ax, float* ay, float* az,
                                                  double* __restrict__ blockPot)
{
    ewaldBlock<RowTable, ProdExp>(tab, first + blockIdx.x * blockDim.x + threadIdx.x, last, x, y, z, m, G, ugrav, ax,
                                  ay, az, blockPot + blockIdx.x);
}

// block b of a batch's Ewald launch: block `block` of the particles [first, last) of one system, with that
// system's table and G (the blocks of a system start at its first particle, so its waves hold the particles
// they hold when the system is corrected alone)
struct EwaldBatchBlock
{
    int table;
    uint32_t first, last, block;
    float G;
};

template<bool RowTable, bool ProdExp>
__global__ __launch_bounds__(256, 4) void ewaldBatchKernel(const EwaldTable* __restrict__ tabs,
                                                       const EwaldBatchBlock* __restrict__ blocks,
                                                       const float* __restrict__ x, const float* __restrict__ y,
                                                       const float* __restrict__ z, const float* __restrict__ m,
                                                       float* ugrav, float* ax, float* ay, float* az,
                                                       double* __restrict__ blockPot)
{
    const EwaldBatchBlock bd = blocks[blockIdx.x];
    ewaldBlock<RowTable, ProdExp>(tabs + bd.table, bd.first + bd.block * blockDim.x + threadIdx.x, bd.last, x, y, z, m,
                                  bd.G, ugrav, ax, ay, az, blockPot + blockIdx.x);
}

// seg[k] = fixed-order sum of v[off[k], off[k+1]) (one block per segment)
__global__ __launch_bounds__(256) void segmentSumKernel(const double* __restrict__ v, const uint32_t* __restrict__ off,
                                                        double* __restrict__ seg)
{
    __shared__ double ss[256];
    const uint32_t b = off[blockIdx.x], e = off[blockIdx.x + 1];
    double s = 0;
    for (uint32_t i = b + threadIdx.x; i < e; i += 256)
        s += v[i];
    ss[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if (int(threadIdx.x) < o) ss[threadIdx.x] += ss[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) seg[blockIdx.x] = ss[0];
}

// block b sums v[b * chunk, min(n, (b + 1) * chunk)) into part[b]: the first stage of a sum in a fixed order
__global__ __launch_bounds__(256) void sumChunksKernel(const double* __restrict__ v, int n, int chunk,
                                                       double* __restrict__ part)
{
    __shared__ double ss[256];
    const int b = blockIdx.x * chunk, e = min(n, b + chunk);
    double s = 0;
    for (int i = b + threadIdx.x; i < e; i += 256)
        s += v[i];
    ss[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if (int(threadIdx.x) < o) ss[threadIdx.x] += ss[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = ss[0];
}

__global__ __launch_bounds__(256) void sumFixedKernel(const double* __restrict__ v, int n, double* __restrict__ out)
{
    __shared__ double ss[256];
    double s = 0;
    for (int i = threadIdx.x; i < n; i += 256)
        s += v[i];
    ss[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if (int(threadIdx.x) < o) ss[threadIdx.x] += ss[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = ss[0];
}

int ewald(Context* ctx, const double* Mroot, const double* center, uint32_t first, uint32_t last, const float* x,
          const float* y, const float* z, const float* m, const float* limits, float G, const double* settings,
          float* ugrav, float* ax, float* ay, float* az, double* totalPot, bool consistentPotential)
{
    float lx = limits[1] - limits[0], ly = limits[3] - limits[2], lz = limits[5] - limits[4];
    if (!(lx == ly && ly == lz)) return setError(ctx, RYO_ERR_UNSUPPORTED, "Ewald gravity requires cubic boxes");
    if (std::ceil(settings[2]) > 3.0)
        return setError(ctx, RYO_ERR_UNSUPPORTED, "Ewald: hCut %g > 3 (the reference's maxCeilHcut, ewald.hpp:53)",
                        settings[2]);
    EwaldTable* t = new EwaldTable();
    ewaldInitTable(*t, Mroot, center, int(settings[0]), double(lx), settings[1], settings[2], settings[3],
                   settings[4], consistentPotential ? 1.0 : -1.0);
    ewaldInitFactors(*t);
    if (t->numEwaldShells == 0 || last <= first)
    {
        delete t;
        if (totalPot) *totalPot = 0;
        return RYO_OK;
    }
    unsigned blocks = iceil(last - first, 256);
    // the per-block potentials are summed in two fixed-order stages: up to 256 chunks, then the chunk sums
    const int chunk = std::max(4096, int(iceil(blocks, 256)));
    const int parts = int(iceil(blocks, unsigned(chunk)));
    size_t need     = sizeof(EwaldTable) + 8 * (size_t(blocks) + 1 + size_t(parts)) + 64;
    auto* buf       = static_cast<char*>(ctx->getScratch(kSlotEwald, need));
    if (!buf)
    {
        delete t;
        return RYO_ERR_HIP;
    }
    auto* dt       = reinterpret_cast<EwaldTable*>(buf);
    auto* blockPot = reinterpret_cast<double*>(buf + ((sizeof(EwaldTable) + 63) & ~size_t(63)));
    double* dsum   = blockPot + blocks;
    hipStream_t s  = ctx->stream;
    const int w     = 2 * t->numEwaldShells + 1;
    const bool rows = w * w <= kMaxRows, prod = t->productExp != 0;
    hipError_t e    = hipMemcpyAsync(dt, t, sizeof(EwaldTable), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s); // the host table is freed below
    delete t;
    if (e != hipSuccess) return setError(ctx, RYO_ERR_HIP, "ewald: %s", hipGetErrorString(e));
    if (rows && prod) ewaldKernel<true, true><<<blocks, 256, 0, s>>>(dt, first, last, x, y, z, m, G, ugrav, ax, ay, az,
                                                                      blockPot);
    else if (rows) ewaldKernel<true, false><<<blocks, 256, 0, s>>>(dt, first, last, x, y, z, m, G, ugrav, ax, ay, az,
                                                                    blockPot);
    else if (prod) ewaldKernel<false, true><<<blocks, 256, 0, s>>>(dt, first, last, x, y, z, m, G, ugrav, ax, ay, az,
                                                                    blockPot);
    else ewaldKernel<false, false><<<blocks, 256, 0, s>>>(dt, first, last, x, y, z, m, G, ugrav, ax, ay, az, blockPot);
    double* part = dsum + 1;
    sumChunksKernel<<<parts, 256, 0, s>>>(blockPot, int(blocks), chunk, part);
    sumFixedKernel<<<1, 256, 0, s>>>(part, parts, dsum);
    RYO_LAUNCH_CHECK(ctx);
    if (totalPot)
    {
        double U = 0;
        RYO_HIP(ctx, hipMemcpyAsync(&U, dsum, sizeof(double), hipMemcpyDeviceToHost, s));
        RYO_HIP(ctx, hipStreamSynchronize(s));
        *totalPot = 0.5 * double(G) * U;
    }
    return RYO_OK;
}

/* The Ewald correction of K periodic systems of a batch (packed arrays x, y, z, m and results): system k owns
 * particles [first[k], last[k]), a cubic box lim[6 k ..], the FP64 root multipole roots[12 k ..] = {Mroot[8],
 * center[3], -}, G[k] and the Ewald mode mode[k] (1: the reference's k-space potential, 2: the consistent one).
 * One launch for all systems (each block within one system, with its table); the block
 * potentials of a system are summed in the same two fixed-order stages as for a system corrected alone, so
 * pot[k] (host) = 0.5 G_k sum(m phi) and every result equal, bit for bit, those of ewald() on that system. */
int ewaldBatch(Context* ctx, int K, const uint32_t* first, const uint32_t* last, const float* lim, const double* roots,
               const float* G, const double* settings, const int* mode, const float* x, const float* y,
               const float* z, const float* m, float* ugrav, float* ax, float* ay, float* az, double* pot)
{
    if (K <= 0) return RYO_OK;
    if (std::ceil(settings[2]) > 3.0)
        return setError(ctx, RYO_ERR_UNSUPPORTED, "Ewald: hCut %g > 3 (the reference's maxCeilHcut, ewald.hpp:53)",
                        settings[2]);
    std::vector<EwaldTable> tabs(K);
    for (int k = 0; k < K; ++k)
    {
        const float* l = lim + 6 * k;
        float lx = l[1] - l[0], ly = l[3] - l[2], lz = l[5] - l[4];
        if (!(lx == ly && ly == lz))
            return setError(ctx, RYO_ERR_UNSUPPORTED, "Ewald gravity requires cubic boxes (periodic system %d)", k);
        ewaldInitTable(tabs[k], roots + 12 * k, roots + 12 * k + 8, int(settings[0]), double(lx), settings[1],
                       settings[2], settings[3], settings[4], mode[k] == 2 ? 1.0 : -1.0);
        ewaldInitFactors(tabs[k]);
        pot[k] = 0;
    }
    if (tabs[0].numEwaldShells == 0) return RYO_OK;
    // blocks of every system, then the chunks of each system's block potentials (ewald()'s first stage) and the
    // ranges of each system's chunk sums (its second stage)
    std::vector<EwaldBatchBlock> blk;
    std::vector<uint32_t> chunkOff{0}, partOff{0};
    for (int k = 0; k < K; ++k)
    {
        const uint32_t nb = last[k] > first[k] ? iceil(last[k] - first[k], 256) : 0;
        const uint32_t b0 = uint32_t(blk.size());
        for (uint32_t j = 0; j < nb; ++j)
            blk.push_back(EwaldBatchBlock{k, first[k], last[k], j, G[k]});
        const uint32_t chunk = std::max(4096u, iceil(nb, 256));
        for (uint32_t c = 0; c < nb; c += chunk)
            chunkOff.push_back(b0 + std::min(nb, c + chunk));
        partOff.push_back(uint32_t(chunkOff.size() - 1));
    }
    const size_t B = blk.size(), C = chunkOff.size() - 1;
    if (B == 0) return RYO_OK;
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    const size_t oTab = 0, oBlk = al(K * sizeof(EwaldTable)), oPot = oBlk + al(B * sizeof(EwaldBatchBlock)),
                 oPart = oPot + al(B * 8), oSeg = oPart + al(C * 8), oCo = oSeg + al(size_t(K) * 8),
                 oPo = oCo + al((C + 1) * 4), need = oPo + al((size_t(K) + 1) * 4);
    auto* buf = static_cast<char*>(ctx->getScratch(kSlotEwald, need));
    if (!buf) return RYO_ERR_HIP;
    auto* dTab  = reinterpret_cast<EwaldTable*>(buf + oTab);
    auto* dBlk  = reinterpret_cast<EwaldBatchBlock*>(buf + oBlk);
    auto* dPot  = reinterpret_cast<double*>(buf + oPot);
    auto* dPart = reinterpret_cast<double*>(buf + oPart);
    auto* dSeg  = reinterpret_cast<double*>(buf + oSeg);
    auto* dCo   = reinterpret_cast<uint32_t*>(buf + oCo);
    auto* dPo   = reinterpret_cast<uint32_t*>(buf + oPo);
    hipStream_t s = ctx->stream;
    RYO_HIP(ctx, hipMemcpyAsync(dTab, tabs.data(), K * sizeof(EwaldTable), hipMemcpyHostToDevice, s));
    RYO_HIP(ctx, hipMemcpyAsync(dBlk, blk.data(), B * sizeof(EwaldBatchBlock), hipMemcpyHostToDevice, s));
    RYO_HIP(ctx, hipMemcpyAsync(dCo, chunkOff.data(), chunkOff.size() * 4, hipMemcpyHostToDevice, s));
    RYO_HIP(ctx, hipMemcpyAsync(dPo, partOff.data(), partOff.size() * 4, hipMemcpyHostToDevice, s));
    const int w     = 2 * tabs[0].numEwaldShells + 1;
    const bool rows = w * w <= kMaxRows, prod = tabs[0].productExp != 0;
    const unsigned nb = unsigned(B);
    if (rows && prod) ewaldBatchKernel<true, true><<<nb, 256, 0, s>>>(dTab, dBlk, x, y, z, m, ugrav, ax, ay, az, dPot);
    else if (rows) ewaldBatchKernel<true, false><<<nb, 256, 0, s>>>(dTab, dBlk, x, y, z, m, ugrav, ax, ay, az, dPot);
    else if (prod) ewaldBatchKernel<false, true><<<nb, 256, 0, s>>>(dTab, dBlk, x, y, z, m, ugrav, ax, ay, az, dPot);
    else ewaldBatchKernel<false, false><<<nb, 256, 0, s>>>(dTab, dBlk, x, y, z, m, ugrav, ax, ay, az, dPot);
    segmentSumKernel<<<unsigned(C), 256, 0, s>>>(dPot, dCo, dPart);
    segmentSumKernel<<<unsigned(K), 256, 0, s>>>(dPart, dPo, dSeg);
    RYO_LAUNCH_CHECK(ctx);
    std::vector<double> U(K);
    RYO_HIP(ctx, hipMemcpyAsync(U.data(), dSeg, K * sizeof(double), hipMemcpyDeviceToHost, s));
    RYO_HIP(ctx, hipStreamSynchronize(s)); // also keeps the host tables alive until the copies are done
    for (int k = 0; k < K; ++k)
        pot[k] = 0.5 * double(G[k]) * U[k];
    return RYO_OK;
}

/* ============================================================================ direct sum */

constexpr int kDirTile = 256;

// partial sums of one source chunk for a block of targets: out[chunk][target][4]
// targets: the listed particles, or first, first + 1, ... when targets is NULL
template<class Tin>
__global__ __launch_bounds__(kDirTile) void directKernel(const Tin* __restrict__ x, const Tin* __restrict__ y,
                                                        const Tin* __restrict__ z, const Tin* __restrict__ h,
                                                        const Tin* __restrict__ m, size_t n,
                                                        const int64_t* __restrict__ targets, size_t first, size_t nt,
                                                        double3 boxL, int numShells, size_t chunkLen,
                                                        double* __restrict__ part)
{
    __shared__ Tin sx[kDirTile], sy[kDirTile], sz[kDirTile], sm[kDirTile], shh[kDirTile];
    const size_t ti = size_t(blockIdx.x) * kDirTile + threadIdx.x;
    const int chunk = blockIdx.y;
    const size_t s0 = size_t(chunk) * chunkLen;
    const size_t s1 = std::min(n, s0 + chunkLen);
    size_t t = 0;
    double px = 0, py = 0, pz = 0, th = 0;
    if (ti < nt)
    {
        t  = targets ? size_t(targets[ti]) : first + ti;
        px = x[t], py = y[t], pz = z[t], th = h[t];
    }
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    for (size_t b = s0; b < s1; b += kDirTile)
    {
        size_t j = b + threadIdx.x;
        __syncthreads();
        if (j < s1) sx[threadIdx.x] = x[j], sy[threadIdx.x] = y[j], sz[threadIdx.x] = z[j], sm[threadIdx.x] = m[j],
                    shh[threadIdx.x] = h[j];
        else sx[threadIdx.x] = 0, sy[threadIdx.x] = 0, sz[threadIdx.x] = 0, sm[threadIdx.x] = 0, shh[threadIdx.x] = 1;
        __syncthreads();
        const int cnt = int(std::min<size_t>(kDirTile, s1 - b));
        for (int iz = -numShells; iz <= numShells; ++iz)
            for (int iy = -numShells; iy <= numShells; ++iy)
                for (int ix = -numShells; ix <= numShells; ++ix)
                {
                    double qx = px + ix * boxL.x, qy = py + iy * boxL.y, qz = pz + iz * boxL.z;
                    for (int k = 0; k < cnt; ++k)
                    {
                        double dx = double(sx[k]) - qx, dy = double(sy[k]) - qy, dz = double(sz[k]) - qz;
                        double R2    = dx * dx + dy * dy + dz * dz;
                        double hij   = th + double(shh[k]);
                        double hij2  = hij * hij;
                        double R2eff = R2 < hij2 ? hij2 : R2;
                        double invR  = 1.0 / sqrt(R2eff);
                        double invR3m = double(sm[k]) * invR * (invR * invR);
                        a0 -= invR3m * R2;
                        a1 += dx * invR3m;
                        a2 += dy * invR3m;
                        a3 += dz * invR3m;
                    }
                }
    }
    if (ti < nt)
    {
        double* o = part + (size_t(chunk) * nt + ti) * 4;
        o[0] = a0, o[1] = a1, o[2] = a2, o[3] = a3;
    }
}

__global__ void directReduceKernel(const double* __restrict__ part, int chunks, size_t nt,
                                   const int64_t* __restrict__ targets, const float* __restrict__ m, double G,
                                   double* __restrict__ out)
{
    size_t ti = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ti >= nt) return;
    double a[4] = {0, 0, 0, 0};
    for (int c = 0; c < chunks; ++c)
        for (int k = 0; k < 4; ++k)
            a[k] += part[(size_t(c) * nt + ti) * 4 + k];
    size_t t = targets ? size_t(targets[ti]) : ti;
    out[4 * ti + 0] = G * double(m[t]) * a[0];
    for (int k = 1; k < 4; ++k)
        out[4 * ti + k] = G * a[k];
}

int directSum(Context* ctx, const float* x, const float* y, const float* z, const float* h, const float* m, size_t n,
              const int64_t* targets, size_t nt, double G, const double* boxL, int numShells, double* out)
{
    if (nt == 0) return RYO_OK;
    unsigned tb = iceil(nt, kDirTile);
    int chunks  = int(std::max<size_t>(1, std::min<size_t>((size_t(ctx->numCUs) * 8 + tb - 1) / tb,
                                                            (n + 4095) / 4096)));
    size_t chunkLen = (n + chunks - 1) / chunks;
    auto* part      = static_cast<double*>(ctx->getScratch(kSlotMisc, size_t(chunks) * nt * 4 * sizeof(double)));
    if (!part) return RYO_ERR_HIP;
    double3 bl = boxL ? make_double3(boxL[0], boxL[1], boxL[2]) : make_double3(0, 0, 0);
    directKernel<float><<<dim3(tb, chunks), kDirTile, 0, ctx->stream>>>(x, y, z, h, m, n, targets, 0, nt, bl,
                                                                         numShells, chunkLen, part);
    directReduceKernel<<<iceil(nt, 256), 256, 0, ctx->stream>>>(part, chunks, nt, targets, m, G, out);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

// the reference's form: p[i] += m_i phi_i, a[i] += acc_i for targets [first, first + nt), G = 1 (direct.cuh:36-85)
template<class T>
__global__ void directAddKernel(const double* __restrict__ part, int chunks, size_t nt, size_t first,
                                const T* __restrict__ m, T* p, T* ax, T* ay, T* az)
{
    size_t ti = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ti >= nt) return;
    double a[4] = {0, 0, 0, 0};
    for (int c = 0; c < chunks; ++c)
        for (int k = 0; k < 4; ++k)
            a[k] += part[(size_t(c) * nt + ti) * 4 + k];
    const size_t t = first + ti;
    p[t] += T(double(m[t]) * a[0]);
    ax[t] += T(a[1]);
    ay[t] += T(a[2]);
    az[t] += T(a[3]);
}

// ryoanji::directSum<T> (direct.cuh:88-107): targets [first, last) against sources [0, numBodies) and their
// (2 numShells + 1)^3 images of boxL, P2P<Exact>, accumulated in FP64 and added to the caller's T outputs
template<class T>
static int directSumRangeT(Context* ctx, size_t first, size_t last, size_t n, const T* boxL, int numShells,
                           const T* x, const T* y, const T* z, const T* m, const T* h, T* p, T* ax, T* ay, T* az)
{
    if (last <= first || n == 0) return RYO_OK;
    const size_t nt = last - first;
    unsigned tb     = iceil(nt, kDirTile);
    int chunks      = int(std::max<size_t>(1, std::min<size_t>((size_t(ctx->numCUs) * 8 + tb - 1) / tb,
                                                                (n + 4095) / 4096)));
    size_t chunkLen = (n + chunks - 1) / chunks;
    auto* part      = static_cast<double*>(ctx->getScratch(kSlotMisc, size_t(chunks) * nt * 4 * sizeof(double)));
    if (!part) return RYO_ERR_HIP;
    double3 bl = make_double3(double(boxL[0]), double(boxL[1]), double(boxL[2]));
    directKernel<T><<<dim3(tb, chunks), kDirTile, 0, ctx->stream>>>(x, y, z, h, m, n, nullptr, first, nt, bl,
                                                                     numShells, chunkLen, part);
    directAddKernel<T><<<iceil(nt, 256), 256, 0, ctx->stream>>>(part, chunks, nt, first, m, p, ax, ay, az);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

int directSumRange(Context* ctx, int types, size_t first, size_t last, size_t n, const void* boxL, int numShells,
                   const void* x, const void* y, const void* z, const void* m, const void* h, void* p, void* ax,
                   void* ay, void* az)
{
    if (last > first && (!boxL || !x || !y || !z || !m || !h || !p || !ax || !ay || !az || last > n))
        return setError(ctx, RYO_ERR_ARG, "direct sum: bad arguments");
    if (numShells < 0) return setError(ctx, RYO_ERR_ARG, "direct sum: negative shells");
    if (types == 0)
        return directSumRangeT<float>(ctx, first, last, n, static_cast<const float*>(boxL), numShells,
                                      static_cast<const float*>(x), static_cast<const float*>(y),
                                      static_cast<const float*>(z), static_cast<const float*>(m),
                                      static_cast<const float*>(h), static_cast<float*>(p), static_cast<float*>(ax),
                                      static_cast<float*>(ay), static_cast<float*>(az));
    if (types == 1)
        return directSumRangeT<double>(ctx, first, last, n, static_cast<const double*>(boxL), numShells,
                                       static_cast<const double*>(x), static_cast<const double*>(y),
                                       static_cast<const double*>(z), static_cast<const double*>(m),
                                       static_cast<const double*>(h), static_cast<double*>(p),
                                       static_cast<double*>(ax), static_cast<double*>(ay), static_cast<double*>(az));
    return setError(ctx, RYO_ERR_ARG, "direct sum: types %d (0 float, 1 double)", types);
}

/* ============================================================================ Ewald pair sum */

// The exact periodic potential of the softened particles, in FP64, on sampled targets: the accuracy reference of
// the periodic pipeline (images + Ewald correction). Every source j contributes m_j psi(r_ij) with the Ewald
// point-mass potential
//   psi(r) = -[ sum_n erfc(a|r+nL|)/|r+nL| + 4 pi / L^3 sum_{k!=0} exp(-k^2 / 4a^2) / k^2 cos(k.r) - pi/(a^2 L^3) ]
// (the convention of ewald.hpp:201-327: neutralizing background, 0 at k = 0), except that a pair image closer
// than h_i + h_j interacts by the softened law of P2P<Exact> (kernel.hpp:99-120: -m R^2/h^3, force m r/h^3),
// as it does in the traversal; the self pair (R = 0) is such a pair. The real-space part runs over the images
// within rc = 5/a (erfc(5) = 1.5e-12), the k-space part over |h| <= ceil(5.5 a L / pi) (weights below 1e-13)
// through structure factors S(k) = sum_j m_j exp(i k.x_j).

struct EwaldK
{
    double kx, ky, kz, w; // k = 2 pi h / L; w = 2 * 4 pi / L^3 * exp(-k^2 / 4a^2) / k^2 (the +-k pair)
};

// part[chunk * K + k] = sum over the chunk's sources of m (cos k.x, sin k.x)
__global__ __launch_bounds__(256) void structureKernel(const float* __restrict__ x, const float* __restrict__ y,
                                                       const float* __restrict__ z, const float* __restrict__ m,
                                                       size_t n, const EwaldK* __restrict__ kv, int K,
                                                       size_t chunkLen, double2* __restrict__ part)
{
    __shared__ double sx[256], sy[256], sz[256], sm[256];
    const int k      = blockIdx.x * 256 + threadIdx.x;
    const size_t s0  = size_t(blockIdx.y) * chunkLen;
    const size_t s1  = std::min(n, s0 + chunkLen);
    double kx = 0, ky = 0, kz = 0;
    if (k < K) kx = kv[k].kx, ky = kv[k].ky, kz = kv[k].kz;
    double c = 0, sn = 0;
    for (size_t b = s0; b < s1; b += 256)
    {
        const size_t j = b + threadIdx.x;
        __syncthreads();
        if (j < s1) sx[threadIdx.x] = x[j], sy[threadIdx.x] = y[j], sz[threadIdx.x] = z[j], sm[threadIdx.x] = m[j];
        __syncthreads();
        const int cnt = int(std::min<size_t>(256, s1 - b));
        for (int q = 0; q < cnt; ++q)
        {
            double sv, cv;
            sincos(kx * sx[q] + ky * sy[q] + kz * sz[q], &sv, &cv);
            c += sm[q] * cv;
            sn += sm[q] * sv;
        }
    }
    if (k < K) part[size_t(blockIdx.y) * K + k] = make_double2(c, sn);
}

__global__ void structureReduceKernel(const double2* __restrict__ part, int chunks, int K, double2* __restrict__ S)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    double c = 0, s = 0;
    for (int q = 0; q < chunks; ++q)
        c += part[size_t(q) * K + k].x, s += part[size_t(q) * K + k].y;
    S[k] = make_double2(c, s);
}

// real-space partial sums of one source chunk for a block of targets: part[chunk][target][4] = {phi, a}
__global__ __launch_bounds__(256) void ewaldPairKernel(const float* __restrict__ x, const float* __restrict__ y,
                                                       const float* __restrict__ z, const float* __restrict__ h,
                                                       const float* __restrict__ m, size_t n,
                                                       const int64_t* __restrict__ targets, size_t nt, double L,
                                                       double alpha, double rc, size_t chunkLen,
                                                       double* __restrict__ part)
{
    __shared__ float sx[256], sy[256], sz[256], sm[256], shh[256];
    const size_t ti = size_t(blockIdx.x) * 256 + threadIdx.x;
    const size_t s0 = size_t(blockIdx.y) * chunkLen;
    const size_t s1 = std::min(n, s0 + chunkLen);
    double px = 0, py = 0, pz = 0, th = 0;
    if (ti < nt)
    {
        const size_t t = targets ? size_t(targets[ti]) : ti;
        px = x[t], py = y[t], pz = z[t], th = h[t];
    }
    const double rc2 = rc * rc, invL = 1.0 / L, a2 = alpha * alpha, a3 = a2 * alpha;
    double phi = 0, ax = 0, ay = 0, az = 0;
    for (size_t b = s0; b < s1; b += 256)
    {
        const size_t j = b + threadIdx.x;
        __syncthreads();
        if (j < s1) sx[threadIdx.x] = x[j], sy[threadIdx.x] = y[j], sz[threadIdx.x] = z[j], sm[threadIdx.x] = m[j],
                    shh[threadIdx.x] = h[j];
        __syncthreads();
        const int cnt = int(std::min<size_t>(256, s1 - b));
        for (int q = 0; q < cnt; ++q)
        {
            // minimum image, then the images one box away along the axes where that can come within rc < L
            double d0[3] = {double(sx[q]) - px, double(sy[q]) - py, double(sz[q]) - pz};
            double sh[3];
            for (int c = 0; c < 3; ++c)
            {
                d0[c] -= L * rint(d0[c] * invL);
                sh[c] = d0[c] > 0 ? -L : L;
            }
            const double hij = th + double(shh[q]), hij2 = hij * hij, mj = double(sm[q]);
            for (int im = 0; im < 8; ++im)
            {
                const double dx = d0[0] + ((im & 1) ? sh[0] : 0.0), dy = d0[1] + ((im & 2) ? sh[1] : 0.0),
                             dz = d0[2] + ((im & 4) ? sh[2] : 0.0);
                const double R2 = dx * dx + dy * dy + dz * dz;
                if (R2 >= rc2) continue;
                double fac;
                if (R2 < hij2)
                {
                    // softened pair: -m R^2/h^3 in place of the point mass's -m/R, plus the smooth part erf(aR)/R
                    // of the Ewald real-space term; g = erf(u)/u, f = ((2/sqrt pi) u exp(-u^2) - erf(u)) / u^3
                    const double R = sqrt(R2), u = alpha * R, u2 = u * u;
                    double g, f;
                    if (u < 0.02)
                    {
                        g = kTwoOverSqrtPi * (1.0 + u2 * (-1.0 / 3.0 + u2 * (0.1 - u2 / 42.0)));
                        f = kTwoOverSqrtPi * (-2.0 / 3.0 + u2 * (0.4 + u2 * (-1.0 / 7.0 + u2 / 27.0)));
                    }
                    else
                    {
                        const double er = erf(u);
                        g = er / u;
                        f = (kTwoOverSqrtPi * u * exp(-u2) - er) / (u2 * u);
                    }
                    const double h3 = hij2 * hij;
                    phi += mj * (alpha * g - R2 / h3);
                    fac = mj * (1.0 / h3 + a3 * f);
                }
                else
                {
                    const double R = sqrt(R2), ec = erfc(alpha * R);
                    phi -= mj * ec / R;
                    fac = mj * (ec / (R2 * R) + kTwoOverSqrtPi * alpha * exp(-a2 * R2) / R2);
                }
                ax += fac * dx;
                ay += fac * dy;
                az += fac * dz;
            }
        }
    }
    if (ti < nt)
    {
        double* o = part + (size_t(blockIdx.y) * nt + ti) * 4;
        o[0] = phi, o[1] = ax, o[2] = ay, o[3] = az;
    }
}

// chunk partials in a fixed order, the k-space sum from the structure factors and the background term; out[4 t] =
// {G m_t phi, G a}
__global__ void ewaldPairReduceKernel(const double* __restrict__ part, int chunks, size_t nt,
                                      const int64_t* __restrict__ targets, const float* __restrict__ x,
                                      const float* __restrict__ y, const float* __restrict__ z,
                                      const float* __restrict__ m, const EwaldK* __restrict__ kv,
                                      const double2* __restrict__ S, int K, double background, double G,
                                      double* __restrict__ out)
{
    const size_t ti = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ti >= nt) return;
    double a[4] = {0, 0, 0, 0};
    for (int c = 0; c < chunks; ++c)
        for (int q = 0; q < 4; ++q)
            a[q] += part[(size_t(c) * nt + ti) * 4 + q];
    const size_t t = targets ? size_t(targets[ti]) : ti;
    const double px = x[t], py = y[t], pz = z[t];
    for (int k = 0; k < K; ++k)
    {
        double sv, cv;
        sincos(kv[k].kx * px + kv[k].ky * py + kv[k].kz * pz, &sv, &cv);
        const double w = kv[k].w;
        a[0] -= w * (S[k].x * cv + S[k].y * sv);
        const double f = w * (S[k].y * cv - S[k].x * sv);
        a[1] += f * kv[k].kx, a[2] += f * kv[k].ky, a[3] += f * kv[k].kz;
    }
    a[0] += background;
    out[4 * ti + 0] = G * double(m[t]) * a[0];
    for (int q = 1; q < 4; ++q)
        out[4 * ti + q] = G * a[q];
}

int ewaldPairSum(Context* ctx, const float* x, const float* y, const float* z, const float* h, const float* m,
                 size_t n, const int64_t* targets, size_t nt, double G, double L, double alphaScale, double* out)
{
    if (nt == 0 || n == 0) return RYO_OK;
    if (!(L > 0)) return setError(ctx, RYO_ERR_ARG, "Ewald pair sum: box length must be positive");
    if (alphaScale == 0) alphaScale = 6.5;
    if (!(alphaScale > 5.0 && alphaScale <= 16.0))
        return setError(ctx, RYO_ERR_ARG, "Ewald pair sum: alpha L in (5, 16]");
    const double alpha = alphaScale / L, rc = 5.0 / alpha;
    const int hmax     = int(std::ceil(5.5 * alphaScale / M_PI));
    std::vector<EwaldK> kv;
    const double c = 8.0 * M_PI / (L * L * L);
    for (int hx = 0; hx <= hmax; ++hx)
        for (int hy = -hmax; hy <= hmax; ++hy)
            for (int hz = -hmax; hz <= hmax; ++hz)
            {
                const bool half = hx > 0 || (hx == 0 && (hy > 0 || (hy == 0 && hz > 0)));
                const int h2    = hx * hx + hy * hy + hz * hz;
                if (!half || h2 > hmax * hmax) continue;
                EwaldK e;
                e.kx = 2 * M_PI * hx / L, e.ky = 2 * M_PI * hy / L, e.kz = 2 * M_PI * hz / L;
                const double k2 = e.kx * e.kx + e.ky * e.ky + e.kz * e.kz;
                e.w = c * std::exp(-k2 / (4 * alpha * alpha)) / k2;
                kv.push_back(e);
            }
    // k = 0 last: its structure factor's real part is the total mass (the background term)
    kv.push_back(EwaldK{0, 0, 0, 0});
    const int K = int(kv.size());
    const unsigned tb = iceil(nt, 256);
    const int pairChunks = int(std::max<size_t>(1, std::min<size_t>((size_t(ctx->numCUs) * 8 + tb - 1) / tb,
                                                                   (n + 4095) / 4096)));
    const size_t pairLen = (n + pairChunks - 1) / pairChunks;
    const unsigned kb    = iceil(K, 256);
    const int sChunks    = int(std::max<size_t>(1, std::min<size_t>((size_t(ctx->numCUs) * 8 + kb - 1) / kb,
                                                                   (n + 4095) / 4096)));
    const size_t sLen    = (n + sChunks - 1) / sChunks;
    const size_t kvBytes = (K * sizeof(EwaldK) + 255) & ~size_t(255);
    const size_t sBytes  = (K * sizeof(double2) + 255) & ~size_t(255);
    const size_t spBytes = (size_t(sChunks) * K * sizeof(double2) + 255) & ~size_t(255);
    const size_t ppBytes = size_t(pairChunks) * nt * 4 * sizeof(double);
    auto* buf = static_cast<char*>(ctx->getScratch(kSlotMisc, kvBytes + sBytes + spBytes + ppBytes));
    if (!buf) return RYO_ERR_HIP;
    auto* dkv   = reinterpret_cast<EwaldK*>(buf);
    auto* dS    = reinterpret_cast<double2*>(buf + kvBytes);
    auto* spart = reinterpret_cast<double2*>(buf + kvBytes + sBytes);
    auto* ppart = reinterpret_cast<double*>(buf + kvBytes + sBytes + spBytes);
    hipStream_t s = ctx->stream;
    RYO_HIP(ctx, hipMemcpyAsync(dkv, kv.data(), K * sizeof(EwaldK), hipMemcpyHostToDevice, s));
    RYO_HIP(ctx, hipStreamSynchronize(s)); // kv is host memory that ends with this call
    structureKernel<<<dim3(kb, sChunks), 256, 0, s>>>(x, y, z, m, n, dkv, K, sLen, spart);
    structureReduceKernel<<<kb, 256, 0, s>>>(spart, sChunks, K, dS);
    ewaldPairKernel<<<dim3(tb, pairChunks), 256, 0, s>>>(x, y, z, h, m, n, targets, nt, L, alpha, rc, pairLen, ppart);
    RYO_LAUNCH_CHECK(ctx);
    // the total mass for the background term: the k = 0 structure factor
    double2 S0;
    RYO_HIP(ctx, hipMemcpyAsync(&S0, dS + (K - 1), sizeof(double2), hipMemcpyDeviceToHost, s));
    RYO_HIP(ctx, hipStreamSynchronize(s));
    const double background = M_PI * S0.x / (alpha * alpha * L * L * L);
    ewaldPairReduceKernel<<<iceil(nt, 256), 256, 0, s>>>(ppart, pairChunks, nt, targets, x, y, z, m, dkv, dS, K - 1,
                                                          background, G, out);
    RYO_LAUNCH_CHECK(ctx);
    return RYO_OK;
}

} // namespace ryo
