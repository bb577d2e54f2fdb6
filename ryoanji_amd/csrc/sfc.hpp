// Hilbert keys for cubic and mixed-dimension boxes (gfx950 device code, also usable on the host).
//
// A box whose axes differ in length by a factor of 4/3 or more gets fewer key bits on the shorter axes
// (cstone/include/cstone/sfc/box.hpp:160-198): bits[d] = 21 - floor(log2(Lmax / L_d) + 0.585). A key then
// starts with 1D levels (only the longest axis has bits there), continues with 2D levels (Hilbert curve in the
// two longest axes, whose orientation also rotates the coordinates of the levels below) and ends with 3D
// Hilbert levels (hilbert.hpp:117-244 encode, :343-515 decode, :517-585 valid keys and cell boxes). With
// bits = {21, 21, 21} everything reduces to the 21-level 3D Hilbert curve.
#pragma once

#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace ryo
{

struct Bits3
{
    unsigned b[3];
};

__host__ __device__ inline bool isCubic(const Bits3& a)
{
    return a.b[0] == kMaxLevel && a.b[1] == kMaxLevel && a.b[2] == kMaxLevel;
}

// 3D Hilbert key of `order` levels (hilbert.hpp:43-94): octant digit through the Morton->Hilbert map
// {0,1,3,2,7,6,4,5}, then the reflections and the rotation/swap that orient the next level
__host__ __device__ constexpr uint64_t hilbert3D(unsigned px, unsigned py, unsigned pz, int order)
{
    constexpr uint32_t toHilbert = (0u << 0) | (1u << 3) | (3u << 6) | (2u << 9) | (7u << 12) | (6u << 15) |
                                   (4u << 18) | (5u << 21); // 3-bit table packed in a word
    uint64_t key = 0;
    for (int level = order - 1; level >= 0; --level)
    {
        unsigned xi = (px >> level) & 1u;
        unsigned yi = (py >> level) & 1u;
        unsigned zi = (pz >> level) & 1u;
        unsigned oct = (xi << 2) | (yi << 1) | zi;
        key = (key << 3) | ((toHilbert >> (3 * oct)) & 7u);
        unsigned nyi = yi ^ 1u, nzi = zi ^ 1u;
        px ^= 0u - (xi & (nyi | zi));
        py ^= 0u - ((xi & (yi | zi)) | (yi & nzi));
        pz ^= 0u - ((xi & nyi & nzi) | (yi & nzi));
        unsigned tx = px, ty = py, tz = pz;
        // zi: cyclic (x,y,z) <- (y,z,x); else !yi: swap x and z
        px = zi ? ty : (nyi ? tz : tx);
        py = zi ? tz : ty;
        pz = zi ? tx : (nyi ? tx : tz);
    }
    return key;
}

// The same 21-level curve as a state machine. hilbert3D complements and permutes the remaining coordinate bits
// after every level; their composition is a signed permutation of the axes, so the coordinates it reads at a
// level are raw bits of the input permuted and complemented by one of 48 states (permutation index * 8 + three
// complement flags). kHilbertLut.two[state][raw bits of two levels] holds both octal digits and the next state,
// so a 21-level key takes one single-level step and ten two-level table reads instead of 21 transform steps.
struct HilbertLut
{
    uint16_t two[48][64]; // index (x1 x0 y1 y0 z1 z0), entry (next state << 6) | (digit1 << 3) | digit0
    uint16_t one[48][8];  // index (x y z), entry (next state << 3) | digit
};

namespace detail
{
constexpr int kPerms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

constexpr int permIndex(int a, int b)
{
    for (int k = 0; k < 6; ++k)
        if (kPerms[k][0] == a && kPerms[k][1] == b) return k;
    return -1;
}

// one level of hilbert3D in state form: raw bits (rx, ry, rz) -> digit and next state
constexpr unsigned hilbertStep(int state, unsigned rx, unsigned ry, unsigned rz, int& next)
{
    constexpr uint32_t toHilbert = (0u << 0) | (1u << 3) | (3u << 6) | (2u << 9) | (7u << 12) | (6u << 15) |
                                   (4u << 18) | (5u << 21);
    const int* p    = kPerms[state >> 3];
    unsigned f[3]   = {unsigned(state >> 2) & 1u, unsigned(state >> 1) & 1u, unsigned(state) & 1u};
    unsigned raw[3] = {rx, ry, rz};
    unsigned xi = raw[p[0]] ^ f[0], yi = raw[p[1]] ^ f[1], zi = raw[p[2]] ^ f[2];
    unsigned digit = (toHilbert >> (3 * ((xi << 2) | (yi << 1) | zi))) & 7u;
    unsigned nyi = yi ^ 1u, nzi = zi ^ 1u;
    f[0] ^= xi & (nyi | zi);
    f[1] ^= (xi & (yi | zi)) | (yi & nzi);
    f[2] ^= (xi & nyi & nzi) | (yi & nzi);
    int q[3]        = {p[0], p[1], p[2]};
    unsigned g[3]   = {f[0], f[1], f[2]};
    if (zi) // coordinates (x, y, z) <- (y, z, x)
        q[0] = p[1], q[1] = p[2], q[2] = p[0], g[0] = f[1], g[1] = f[2], g[2] = f[0];
    else if (nyi) // swap x and z
        q[0] = p[2], q[2] = p[0], g[0] = f[2], g[2] = f[0];
    next = permIndex(q[0], q[1]) * 8 + int((g[0] << 2) | (g[1] << 1) | g[2]);
    return digit;
}

constexpr HilbertLut makeHilbertLut()
{
    HilbertLut t{};
    for (int s = 0; s < 48; ++s)
    {
        for (unsigned o = 0; o < 8; ++o)
        {
            int n = 0;
            unsigned d = hilbertStep(s, o >> 2, (o >> 1) & 1u, o & 1u, n);
            t.one[s][o] = uint16_t((n << 3) | d);
        }
        for (unsigned i = 0; i < 64; ++i)
        {
            int s1 = 0, s2 = 0;
            unsigned d1 = hilbertStep(s, (i >> 5) & 1u, (i >> 3) & 1u, (i >> 1) & 1u, s1);
            unsigned d0 = hilbertStep(s1, (i >> 4) & 1u, (i >> 2) & 1u, i & 1u, s2);
            t.two[s][i] = uint16_t((s2 << 6) | (d1 << 3) | d0);
        }
    }
    return t;
}
} // namespace detail

constexpr HilbertLut kHilbertLut = detail::makeHilbertLut();

// 21-level 3D Hilbert key from the tables (equal to hilbert3D(px, py, pz, 21))
__host__ __device__ constexpr uint64_t hilbert3DLut(unsigned px, unsigned py, unsigned pz, const uint16_t (*two)[64],
                                                    const uint16_t (*one)[8])
{
    unsigned e   = one[0][((px >> 18) & 4u) | ((py >> 19) & 2u) | ((pz >> 20) & 1u)];
    uint64_t key = e & 7u;
    unsigned s   = e >> 3;
    for (int level = 18; level >= 0; level -= 2)
    {
        const unsigned i = (((px >> level) & 3u) << 4) | (((py >> level) & 3u) << 2) | ((pz >> level) & 3u);
        e   = two[s][i];
        key = (key << 6) | (e & 63u);
        s   = e >> 6;
    }
    return key;
}

namespace detail
{
// compile-time check of the tables against hilbert3D on a pseudo-random sample of 21-bit coordinates
constexpr bool hilbertLutMatches()
{
    uint64_t r = 0x9E3779B97F4A7C15ull;
    for (int k = 0; k < 2000; ++k)
    {
        r ^= r << 13, r ^= r >> 7, r ^= r << 17;
        const unsigned x = unsigned(r) & 0x1FFFFFu, y = unsigned(r >> 21) & 0x1FFFFFu, z = unsigned(r >> 42) & 0x1FFFFFu;
        if (hilbert3DLut(x, y, z, kHilbertLut.two, kHilbertLut.one) != hilbert3D(x, y, z, 21))
            return false;
    }
    return true;
}
static_assert(hilbertLutMatches(), "Hilbert tables differ from hilbert3D");
} // namespace detail

// inverse of hilbert3D for `order` levels: corner coordinates of the cell with start key `key`
__host__ __device__ inline void hilbert3DDecode(uint64_t key, int order, unsigned& ox, unsigned& oy, unsigned& oz)
{
    unsigned px = 0, py = 0, pz = 0;
    for (int level = 0; level < order; ++level)
    {
        unsigned oct = unsigned(key >> (3 * level)) & 7u;
        unsigned xi = oct >> 2u, yi = (oct >> 1u) & 1u, zi = oct & 1u;
        unsigned tx = px, ty = py, tz = pz;
        if (yi ^ zi) // inverse cyclic rotation
        {
            px = tz, pz = ty, py = tx;
        }
        else if ((xi | yi | zi) == 0u || (xi & yi & zi) == 1u) // swap x and z
        {
            px = tz, pz = tx;
        }
        unsigned mask = (1u << level) - 1u;
        unsigned nxi = xi ^ 1u, nyi = yi ^ 1u, nzi = zi ^ 1u;
        px ^= mask & (0u - (xi & (yi | zi)));
        py ^= mask & (0u - ((xi & (nyi | nzi)) | (nxi & yi & zi)));
        pz ^= mask & (0u - ((xi & nyi & nzi) | (yi & zi)));
        px |= xi << level;
        py |= (xi ^ yi) << level;
        pz |= (yi ^ zi) << level;
    }
    ox = px, oy = py, oz = pz;
}

// 2D Hilbert key of `order` levels (hilbert.hpp:246-277): two bits per level, 2 xi + (xi ^ yi); when yi = 0 the
// remaining coordinates are swapped and, if xi = 1, complemented
__host__ __device__ inline uint64_t hilbert2D(unsigned px, unsigned py, int order)
{
    uint64_t key = 0;
    for (int level = order - 1; level >= 0; --level)
    {
        unsigned xi = (px >> level) & 1u, yi = (py >> level) & 1u;
        if (yi == 0u)
        {
            unsigned flip = 0u - xi, t = px;
            px = py ^ flip;
            py = t ^ flip;
        }
        key = (key << 2) | (xi << 1) | (xi ^ yi);
    }
    return key;
}

// inverse of hilbert2D for `order` >= 1 levels (hilbert.hpp:317-341)
__host__ __device__ inline void hilbert2DDecode(uint64_t key, int order, unsigned& ox, unsigned& oy)
{
    unsigned x = 0, y = 0;
    for (int level = 0; level < 2 * order; level += 2)
    {
        unsigned sa = unsigned(key >> (level + 1)) & 1u, sb = unsigned(key >> level) & 1u;
        if ((sa ^ sb) == 0u)
        {
            unsigned flip = 0u - sa, t = x;
            x = y ^ flip;
            y = t ^ flip;
        }
        x = (x >> 1) | (sa << 31);
        y = (y >> 1) | ((sa ^ sb) << 31);
    }
    ox = x >> (32 - order);
    oy = y >> (32 - order);
}

// axes ordered by decreasing bit count; ties keep the x, y, z order (the three compare-and-swap steps of the
// reference sort network)
struct AxisOrder
{
    unsigned bits[3];
    int axis[3];
};

__host__ __device__ inline AxisOrder sortAxes(const Bits3& a)
{
    AxisOrder o{{a.b[0], a.b[1], a.b[2]}, {0, 1, 2}};
    auto cas = [&o](int i, int j)
    {
        if (o.bits[i] < o.bits[j])
        {
            unsigned tb = o.bits[i];
            o.bits[i]   = o.bits[j];
            o.bits[j]   = tb;
            int ta      = o.axis[i];
            o.axis[i]   = o.axis[j];
            o.axis[j]   = ta;
        }
    };
    cas(0, 1);
    cas(0, 2);
    cas(1, 2);
    return o;
}

// the 2D levels decide the orientation in which the 3D part of the key is reached: apply (encode) the same swap
// and complement to the low coordinates, from the coarsest 2D level to the finest
__host__ __device__ inline void orient2D(uint64_t key2D, int n, uint64_t mask, uint64_t& u, uint64_t& v, bool forward)
{
    for (int i = 0; i < n; ++i)
    {
        int level      = forward ? n - 1 - i : i;
        unsigned quad  = unsigned(key2D >> (2 * level)) & 3u;
        unsigned xi    = (quad >> 1) & 1u;
        unsigned yi    = xi ^ (quad & 1u);
        if (yi == 0u)
        {
            uint64_t t = u;
            u          = v;
            v          = t;
            if (xi == 1u)
            {
                u = mask - u;
                v = mask - v;
            }
        }
    }
}

// Hilbert key of integer coordinates with per-axis bit counts (hilbert.hpp:117-244)
__host__ __device__ inline uint64_t hilbertMixD(unsigned px, unsigned py, unsigned pz, const Bits3& bits)
{
    if (isCubic(bits)) return hilbert3D(px, py, pz, int(kMaxLevel));
    AxisOrder o    = sortAxes(bits);
    unsigned in[3] = {px, py, pz};
    uint64_t c[3]  = {in[o.axis[0]], in[o.axis[1]], in[o.axis[2]]};
    unsigned b0 = o.bits[0], b1 = o.bits[1], b2 = o.bits[2];
    uint64_t key = 0;
    // 1D levels: the longest axis alone, one bit per 3-bit digit
    for (unsigned bit = b1; bit < b0; ++bit)
        key |= ((c[0] >> bit) & 1u) << (3 * bit);
    c[0] &= (uint64_t(1) << b1) - 1;
    // 2D levels of the two longest axes
    if (b1 > b2)
    {
        const int n         = int(b1 - b2);
        const uint64_t mask = (uint64_t(1) << b2) - 1;
        uint64_t key2D      = hilbert2D(unsigned(c[0] >> b2), unsigned(c[1] >> b2), n);
        uint64_t u = c[0] & mask, v = c[1] & mask;
        orient2D(key2D, n, mask, u, v, true);
        for (int i = 0; i < n; ++i)
            key |= ((key2D >> (2 * (n - 1 - i))) & 3u) << (3 * (b1 - 1 - unsigned(i)));
        c[0] = u, c[1] = v;
    }
    // 3D levels
    return key | hilbert3D(unsigned(c[0]), unsigned(c[1]), unsigned(c[2]), int(b2));
}

// inverse of hilbertMixD (hilbert.hpp:343-515)
__host__ __device__ inline void hilbertMixDDecode(uint64_t key, const Bits3& bits, unsigned& ox, unsigned& oy,
                                                  unsigned& oz)
{
    if (isCubic(bits))
    {
        hilbert3DDecode(key, int(kMaxLevel), ox, oy, oz);
        return;
    }
    AxisOrder o   = sortAxes(bits);
    unsigned b0 = o.bits[0], b1 = o.bits[1], b2 = o.bits[2];
    uint64_t c[3] = {0, 0, 0};
    for (unsigned bit = b1; bit < b0; ++bit)
        c[0] |= ((key >> (3 * bit)) & 1u) << bit;
    key &= (uint64_t(1) << (3 * b1)) - 1;
    uint64_t key2D = 0;
    const int n    = b1 > b2 ? int(b1 - b2) : 0;
    if (n > 0)
    {
        for (int i = 0; i < n; ++i)
            key2D |= ((key >> (3 * (b1 - 1 - unsigned(i)))) & 3u) << (2 * (n - 1 - i));
        unsigned hx, hy;
        hilbert2DDecode(key2D, n, hx, hy);
        const uint64_t nm = (uint64_t(1) << n) - 1;
        c[0] |= (uint64_t(hx) & nm) << b2;
        c[1] |= (uint64_t(hy) & nm) << b2;
        key &= (uint64_t(1) << (3 * b2)) - 1;
    }
    unsigned x3, y3, z3;
    hilbert3DDecode(key, int(b2), x3, y3, z3);
    uint64_t u = x3, v = y3;
    if (n > 0) orient2D(key2D, n, (uint64_t(1) << b2) - 1, u, v, false);
    c[0] |= u, c[1] |= v, c[2] |= z3;
    unsigned out[3];
    out[o.axis[0]] = unsigned(c[0]);
    out[o.axis[1]] = unsigned(c[1]);
    out[o.axis[2]] = unsigned(c[2]);
    ox = out[0], oy = out[1], oz = out[2];
}

// a key reachable by the mixed-dimension encoding: digits of 2D levels are < 4, digits of 1D levels < 2
// (hilbert.hpp:517-548)
__host__ __device__ inline bool validMixDKey(uint64_t key, const Bits3& bits)
{
    unsigned s[3] = {bits.b[0], bits.b[1], bits.b[2]}; // ascending
    auto cas = [&s](int i, int j)
    {
        if (s[i] > s[j])
        {
            unsigned t = s[i];
            s[i]       = s[j];
            s[j]       = t;
        }
    };
    cas(0, 1);
    cas(0, 2);
    cas(1, 2);
    for (unsigned i = s[0] + 1; i <= kMaxLevel; ++i)
    {
        unsigned digit = unsigned(key >> (3 * (i - 1))) & 7u;
        if (i <= s[1])
        {
            if (digit > 3u) return false;
        }
        else if (i <= s[2])
        {
            if (digit > 1u) return false;
        }
    }
    return true;
}

// integer box [lo, lo + len) per axis of the cell at `level` with start key `key` (hilbertIBox,
// hilbert.hpp:563-585); an unreachable key gives the empty box at the origin
struct IBox3
{
    unsigned lo[3], len[3];
};

__host__ __device__ inline IBox3 cellIBox(uint64_t key, unsigned level, const Bits3& bits)
{
    IBox3 r{{0, 0, 0}, {0, 0, 0}};
    const unsigned lfr = kMaxLevel - level;
    if (isCubic(bits))
    {
        // the decode runs from the lowest level up and the cell's lower 3 lfr key bits are zero: those levels
        // leave the coordinates' upper bits as they are (they only reflect and permute bits below `level`, which
        // the box drops), so the top `level` levels decoded on their own give the box corner directly
        unsigned c[3];
        hilbert3DDecode(key >> (3 * lfr), int(level), c[0], c[1], c[2]);
        for (int d = 0; d < 3; ++d)
        {
            r.lo[d]  = c[d] << lfr;
            r.len[d] = 1u << lfr;
        }
        return r;
    }
    if (!validMixDKey(key, bits)) return r;
    unsigned c[3];
    hilbertMixDDecode(key, bits, c[0], c[1], c[2]);
    for (int d = 0; d < 3; ++d)
    {
        unsigned len = 1u << (bits.b[d] < lfr ? bits.b[d] : lfr);
        r.lo[d]      = c[d] & ~(len - 1u);
        r.len[d]     = len;
    }
    return r;
}

} // namespace ryo

namespace ryo
{

// a box as the key kernels use it, in the coordinates' type
template<class T>
struct KeyBoxT
{
    T lo[3];
    T m[3];     // 2^bits[d] / L_d
    Bits3 bits; // key bits per axis (21 on every axis of a cubic box)
};
using KeyBox = KeyBoxT<float>;

// a box as the MAC computation uses it
struct MacBox
{
    double lo[3];
    double half[3]; // half of the smallest cell edge per axis: 0.5 L_d / 2^bits[d]
    Bits3 bits;     // key bits per axis
    float invTheta;
    double softReach = 0; // > 0: cells also open within their radius plus this distance (2 x the largest h)
};

// per-axis key bits of a box given as {xmin, xmax, ymin, ymax, zmin, zmax} (Box<T>::computeBoxDimBits and
// getBoxDimBits, box.hpp:160-198, in the box's type T as the reference's Box<T> computes them)
template<class T>
inline Bits3 boxBits(const T* limits, unsigned maxLevel = kMaxLevel)
{
    T len[3] = {limits[1] - limits[0], limits[3] - limits[2], limits[5] - limits[4]};
    T mx     = std::max(std::max(len[0], len[1]), len[2]);
    Bits3 r;
    for (int d = 0; d < 3; ++d)
    {
        unsigned red = unsigned(std::floor(std::log2(mx / len[d]) + T(0.5849625007211563)));
        r.b[d]       = red >= maxLevel ? 0u : maxLevel - red;
    }
    return r;
}

} // namespace ryo
