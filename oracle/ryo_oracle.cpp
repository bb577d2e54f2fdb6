/*
 * ryo_oracle.cpp -- CPU restatement of the reference's single-GPU Barnes-Hut path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: it is loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the product
 * library (ryoanji_amd/csrc), which fails loudly when its HIP build is missing.
 *
 * Every function follows the reference algorithm it cites (paths relative to the
 * reference checkout). Floating-point operations are written in the reference's
 * operation order and the file is compiled with -ffp-contract=off, so on x86-64 the
 * float results of the serial parts are the reference's bit for bit.
 *
 * Pinning (what checks each function, tests/test_oracle.py):
 *   keys            : reference Hilbert goldens + brute-force inversion/adjacency on
 *                     small orders + equality with oracle/_ref on random inputs
 *   sort            : brute force (std::stable_sort semantics)
 *   octree (leaves) : reference tree examples, canonical-tree invariants, oracle/_ref
 *   linked octree   : reference level counts (1/8/64, 73 nodes), oracle/_ref
 *   centers/mac/P2M/M2M : closed forms, P2M==M2M composite (1e-10), oracle/_ref
 *   P2P / M2P       : reference goldens (kernel.cpp, cartesian_qpole.cpp), oracle/_ref
 *   gravity         : FP64 direct-sum accuracy bounds + oracle/_ref (16-target groups)
 *   ewald           : analytic single-particle energy goldens, disabled==0, oracle/_ref
 *   Ewald pair sum  : Madelung constant of the simple cubic lattice (one particle, eight on a
 *                     sub-lattice), alpha independence, traversal + Ewald correction where the
 *                     root multipole is exact (tests/test_oracle_ewald_pair.py)
 */

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <tuple>
#include <vector>

namespace orc
{

constexpr unsigned kMaxLevel = 21; // 64-bit keys, 63 used bits
constexpr uint64_t kKeyEnd   = 1ull << 63;

inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }
inline int ctz64(uint64_t x) { return x ? __builtin_ctzll(x) : 64; }

// cstone/include/cstone/sfc/common.hpp:89-95
inline uint64_t nodeRange(unsigned level) { return 1ull << (3u * (kMaxLevel - level)); }
// common.hpp:147-151
inline unsigned treeLevel(uint64_t range) { return unsigned(clz64(range - 1) - 1) / 3u; }
// common.hpp:135-138
inline int commonPrefix(uint64_t a, uint64_t b) { return clz64(a ^ b) - 1; }
// common.hpp:163-170
inline uint64_t encodePlaceholderBit(uint64_t key, int prefixLength)
{
    int shifts = 63 - prefixLength;
    return (1ull << prefixLength) | (key >> shifts);
}
// common.hpp:182-186
inline unsigned decodePrefixLength(uint64_t code) { return 63u - unsigned(clz64(code)); }
// common.hpp:196-203
inline uint64_t decodePlaceholderBit(uint64_t code)
{
    unsigned pl = decodePrefixLength(code);
    return (code ^ (1ull << pl)) << (63 - pl);
}
// common.hpp:289-292
inline unsigned octalDigit(uint64_t key, unsigned pos) { return unsigned(key >> (3u * (kMaxLevel - pos))) & 7u; }
// common.hpp:308-312
inline int digitWeight(int d)
{
    int ge4 = -int(d >= 4);
    return ((7 - d) & ge4) - (d & ~ge4);
}
// octree.hpp:55-65
inline int binaryKeyWeight(uint64_t key, unsigned level)
{
    int r = 0;
    for (unsigned l = 1; l <= level + 1; ++l)
        r += digitWeight(int(octalDigit(key, l)));
    return r;
}
// common.hpp:99-106
inline unsigned log8ceil(uint64_t n)
{
    if (n == 0) return 0;
    return kMaxLevel - unsigned(clz64(n - 1) - 1) / 3u;
}

/* ---------------------------------------------------------------- box ---- */

// cstone/include/cstone/sfc/box.hpp:99-214 (float fields computed in T)
template<class T>
struct Box
{
    T lo[3], hi[3], len[3], inv[3];
    int periodic[3];
    unsigned bits[3]; // per-axis key bits, 21 for a cubic box
};

template<class T>
Box<T> makeBox(const T* lim, const int* pbc)
{
    Box<T> b;
    for (int d = 0; d < 3; ++d)
    {
        b.lo[d]       = lim[2 * d];
        b.hi[d]       = lim[2 * d + 1];
        b.len[d]      = b.hi[d] - b.lo[d];
        b.inv[d]      = T(1.) / (b.hi[d] - b.lo[d]);
        b.periodic[d] = pbc ? pbc[d] : 0;
    }
    // box.hpp:193-198: bit reduction floor(log2(maxExtent / l) + bias)
    T maxExt        = std::max(std::max(b.len[0], b.len[1]), b.len[2]);
    constexpr T bias = T(0.5849625007211563);
    for (int d = 0; d < 3; ++d)
        b.bits[d] = kMaxLevel - unsigned(std::floor(std::log2(maxExt / b.len[d]) + bias));
    return b;
}

/* --------------------------------------------------------------- keys ---- */

// cstone/include/cstone/sfc/hilbert.hpp:43-94
inline uint64_t hilbert3D(unsigned px, unsigned py, unsigned pz, int order)
{
    static const unsigned toHilbert[8] = {0, 1, 3, 2, 7, 6, 4, 5};
    uint64_t key = 0;
    for (int level = order - 1; level >= 0; --level)
    {
        unsigned xi = (px >> level) & 1u;
        unsigned yi = (py >> level) & 1u;
        unsigned zi = (pz >> level) & 1u;
        key         = (key << 3) + toHilbert[(xi << 2) | (yi << 1) | zi];

        px ^= 0u - (xi & (unsigned(!yi) | zi));
        py ^= 0u - ((xi & (yi | zi)) | (yi & unsigned(!zi)));
        pz ^= 0u - ((xi & unsigned(!yi) & unsigned(!zi)) | (yi & unsigned(!zi)));
        if (zi)
        {
            unsigned t = px;
            px = py, py = pz, pz = t;
        }
        else if (!yi)
        {
            std::swap(px, pz);
        }
    }
    return key;
}

// hilbert.hpp:281-325 (inverse of hilbert3D)
inline void decodeHilbert3D(uint64_t key, unsigned& ox, unsigned& oy, unsigned& oz)
{
    unsigned px = 0, py = 0, pz = 0;
    for (unsigned level = 0; level < kMaxLevel; ++level)
    {
        unsigned oct = unsigned(key >> (3 * level)) & 7u;
        unsigned xi = oct >> 2u, yi = (oct >> 1u) & 1u, zi = oct & 1u;
        if (yi ^ zi)
        {
            unsigned t = px;
            px = pz, pz = py, py = t;
        }
        else if ((!xi & !yi & !zi) || (xi & yi & zi))
        {
            std::swap(px, pz);
        }
        unsigned mask = (1u << level) - 1u;
        px ^= mask & (0u - (xi & (yi | zi)));
        py ^= mask & (0u - ((xi & (unsigned(!yi) | unsigned(!zi))) | (unsigned(!xi) & yi & zi)));
        pz ^= mask & (0u - ((xi & unsigned(!yi) & unsigned(!zi)) | (yi & zi)));
        px |= (xi << level);
        py |= ((xi ^ yi) << level);
        pz |= ((yi ^ zi) << level);
    }
    ox = px, oy = py, oz = pz;
}

/* Mixed-dimension Hilbert keys (hilbert.hpp:117-244), restated level by level: axes sorted by decreasing bit
 * count (compare-and-swap (0,1), (0,2), (1,2), ties keep x, y, z); levels where only the longest axis has bits
 * carry that bit; levels where two axes have bits follow the 2D Hilbert curve, whose swap/complement is applied
 * to the whole remaining coordinates (so the 3D levels below see the rotated frame); the rest is 3D Hilbert. */
inline void sortAxesDesc(const unsigned* bits, unsigned* sb, int* ax)
{
    for (int d = 0; d < 3; ++d)
        sb[d] = bits[d], ax[d] = d;
    const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (auto& p : pairs)
        if (sb[p[0]] < sb[p[1]])
        {
            std::swap(sb[p[0]], sb[p[1]]);
            std::swap(ax[p[0]], ax[p[1]]);
        }
}

inline uint64_t hilbertMixD(unsigned px, unsigned py, unsigned pz, const unsigned* bits)
{
    if (bits[0] == kMaxLevel && bits[1] == kMaxLevel && bits[2] == kMaxLevel) return hilbert3D(px, py, pz, kMaxLevel);
    unsigned sb[3];
    int ax[3];
    sortAxesDesc(bits, sb, ax);
    unsigned in[3] = {px, py, pz};
    unsigned c0 = in[ax[0]], c1 = in[ax[1]], c2 = in[ax[2]];
    uint64_t key = 0;
    for (int L = int(sb[0]) - 1; L >= 0; --L)
    {
        if (L >= int(sb[1])) // one axis
            key |= uint64_t((c0 >> L) & 1u) << (3 * L);
        else if (L >= int(sb[2])) // two axes: 2D Hilbert digit, then orient the remaining coordinates
        {
            unsigned xi = (c0 >> L) & 1u, yi = (c1 >> L) & 1u;
            key |= uint64_t((xi << 1) | (xi ^ yi)) << (3 * L);
            if (yi == 0u)
            {
                std::swap(c0, c1);
                if (xi) c0 = ~c0, c1 = ~c1;
            }
        }
        else // three axes: the rest is a 3D Hilbert key of L + 1 levels in the current frame
        {
            unsigned m = (L + 1 >= 32) ? ~0u : ((1u << (L + 1)) - 1u);
            key |= hilbert3D(c0 & m, c1 & m, c2 & m, L + 1);
            break;
        }
    }
    return key;
}

/* inverse of hilbertMixD, one coordinate bit per level from the top: at each level the bit combination whose
 * encoding reproduces the key's digit there (digits depend on the coordinate bits at and above their level
 * only). Returns false for a key no point encodes to (hilbert.hpp:517-548). */
inline bool decodeHilbertMixD(uint64_t key, const unsigned* bits, unsigned& ox, unsigned& oy, unsigned& oz)
{
    if (bits[0] == kMaxLevel && bits[1] == kMaxLevel && bits[2] == kMaxLevel)
    {
        decodeHilbert3D(key, ox, oy, oz);
        return true;
    }
    unsigned c[3] = {0, 0, 0};
    for (int L = int(kMaxLevel) - 1; L >= 0; --L)
    {
        unsigned digit = unsigned(key >> (3 * L)) & 7u;
        int act[3], na = 0;
        for (int d = 0; d < 3; ++d)
            if (int(bits[d]) > L) act[na++] = d;
        if (na == 0)
        {
            if (digit != 0) return false;
            continue;
        }
        bool found = false;
        for (unsigned combo = 0; combo < (1u << na) && !found; ++combo)
        {
            unsigned t[3] = {c[0], c[1], c[2]};
            for (int k = 0; k < na; ++k)
                t[act[k]] |= ((combo >> k) & 1u) << L;
            if (((hilbertMixD(t[0], t[1], t[2], bits) >> (3 * L)) & 7u) == digit)
            {
                c[0] = t[0], c[1] = t[1], c[2] = t[2];
                found = true;
            }
        }
        if (!found) return false;
    }
    ox = c[0], oy = c[1], oz = c[2];
    return true;
}

// cstone/include/cstone/sfc/morton.hpp:94-110 (bit interleave, x most significant)
inline uint64_t morton3D(unsigned ix, unsigned iy, unsigned iz)
{
    uint64_t key = 0;
    for (int b = kMaxLevel - 1; b >= 0; --b)
        key = (key << 3) | (uint64_t((ix >> b) & 1u) << 2) | (uint64_t((iy >> b) & 1u) << 1) | ((iz >> b) & 1u);
    return key;
}

// cstone/include/cstone/sfc/sfc.hpp:153-159 (fork patch: split integer / fractional parts)
template<class T>
inline int cellIndex(T x, T xmin, T m)
{
    const T a = x * m, b = xmin * m;
    const T fa = std::floor(a), fb = std::floor(b);
    return int(fa - fb) - int(a - fa < b - fb);
}

// sfc.hpp:161-207 + :297-305 (keys equal to removeKey = 2^63 are left untouched)
template<class T>
void sfcKeys(const T* x, const T* y, const T* z, uint64_t* keys, size_t n, const Box<T>& box, bool hilbert)
{
    T m[3];
    for (int d = 0; d < 3; ++d)
        m[d] = T(1u << box.bits[d]) * box.inv[d];
    const int mc[3] = {int((1u << box.bits[0]) - 1), int((1u << box.bits[1]) - 1), int((1u << box.bits[2]) - 1)};
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i)
    {
        if (keys[i] == kKeyEnd) continue;
        int ix    = std::min(cellIndex(x[i], box.lo[0], m[0]), mc[0]);
        int iy    = std::min(cellIndex(y[i], box.lo[1], m[1]), mc[1]);
        int iz    = std::min(cellIndex(z[i], box.lo[2], m[2]), mc[2]);
        keys[i] = hilbert ? hilbertMixD(ix, iy, iz, box.bits) : morton3D(ix, iy, iz);
    }
}

/* ---------------------------------------------------------------- sort --- */

// cstone/include/cstone/primitives/gather.hpp:43-66 (std::stable_sort on the key)
inline void sortByKey(uint64_t* keys, int32_t* values, size_t n)
{
    std::vector<std::pair<uint64_t, int32_t>> kv(n);
    for (size_t i = 0; i < n; ++i)
        kv[i] = {keys[i], values[i]};
    std::stable_sort(kv.begin(), kv.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (size_t i = 0; i < n; ++i)
        keys[i] = kv[i].first, values[i] = kv[i].second;
}

/* ------------------------------------------------------- leaf octree ----- */

// cstone/include/cstone/tree/csarray.hpp:80-90
inline unsigned nodeCount(uint64_t a, uint64_t b, const uint64_t* k, size_t n)
{
    return unsigned(std::lower_bound(k, k + n, b) - std::lower_bound(k, k + n, a));
}

// csarray.hpp:237-250 (sibling index if all 8 siblings are present at the same level, else -1)
inline int siblingIndex(const uint64_t* tree, long i, unsigned& level)
{
    uint64_t node = tree[i];
    level         = treeLevel(tree[i + 1] - node);
    if (level == 0) return -1;
    int sib = int(octalDigit(node, level));
    bool ok = tree[i - sib + 8] == tree[i - sib] + nodeRange(level - 1);
    return ok ? sib : -1;
}

// csarray.hpp:267-290
inline int nodeOp(const uint64_t* tree, long i, const unsigned* counts, unsigned bucket)
{
    unsigned level;
    int sib = siblingIndex(tree, i, level);
    if (sib > 0)
    {
        const unsigned* g = counts + i - sib;
        size_t parent     = 0;
        for (int c = 0; c < 8; ++c)
            parent += g[c];
        if (parent <= size_t(bucket)) return 0;
    }
    unsigned c = counts[i];
    if (c > bucket * 512 && level + 3 < kMaxLevel) return 4096;
    if (c > bucket * 64 && level + 2 < kMaxLevel) return 512;
    if (c > bucket * 8 && level + 1 < kMaxLevel) return 64;
    if (c > bucket && level < kMaxLevel) return 8;
    return 1;
}

// csarray.hpp:408-440: iterate rebalance (split/merge) + count to the fixed point from {0, 2^63}
inline void computeOctree(const uint64_t* keys, size_t n, unsigned bucket, std::vector<uint64_t>& tree,
                          std::vector<unsigned>& counts)
{
    tree   = {0, kKeyEnd};
    counts = {unsigned(n)};
    for (;;)
    {
        long numNodes = long(tree.size()) - 1;
        std::vector<long> ops(numNodes + 1, 0);
        bool converged = true;
        for (long i = 0; i < numNodes; ++i)
        {
            ops[i] = nodeOp(tree.data(), i, counts.data(), bucket);
            if (ops[i] != 1) converged = false;
        }
        // csarray.hpp:339-389 (exclusive scan of ops, emit new nodes)
        std::exclusive_scan(ops.begin(), ops.end(), ops.begin(), 0L);
        std::vector<uint64_t> next(ops[numNodes] + 1);
        for (long i = 0; i < numNodes; ++i)
        {
            uint64_t node  = tree[i];
            unsigned level = treeLevel(tree[i + 1] - node);
            long op = ops[i + 1] - ops[i], pos = ops[i];
            if (op == 1) next[pos] = node;
            else if (op >= 8)
            {
                unsigned diff = log8ceil(uint64_t(op));
                for (long s = 0; s < op; ++s)
                    next[pos + s] = node + uint64_t(s) * nodeRange(level + diff);
            }
        }
        next.back() = tree.back();
        tree.swap(next);
        counts.resize(tree.size() - 1);
        for (size_t i = 0; i + 1 < tree.size(); ++i)
            counts[i] = nodeCount(tree[i], tree[i + 1], keys, n);
        if (converged) break;
    }
}

/* ----------------------------------------------------- linked octree ----- */

// cstone/include/cstone/tree/octree.hpp:77-196 (unsorted layout, sort, invert, levels, link)
struct Octree
{
    int numLeaves, numInternal, numNodes;
    std::vector<uint64_t> prefixes;
    std::vector<int> childOffsets, parents, levelRange, internalToLeaf, leafToInternal;
};

inline Octree linkOctree(const uint64_t* leaves, int numLeaves)
{
    Octree o;
    o.numLeaves   = numLeaves;
    o.numInternal = (numLeaves - 1) / 7;
    o.numNodes    = o.numLeaves + o.numInternal;
    int nI = o.numInternal, n = o.numNodes;
    std::vector<uint64_t> pre(n);
    std::vector<int> order(n);
    for (int t = 0; t < numLeaves; ++t)
    {
        uint64_t key       = leaves[t];
        unsigned level     = treeLevel(leaves[t + 1] - key);
        pre[t + nI]        = encodePlaceholderBit(key, 3 * level);
        order[t + nI]      = t + nI;
        int pl             = commonPrefix(key, leaves[t + 1]);
        if (pl % 3 == 0 && t < numLeaves - 1)
        {
            int idx    = (t + binaryKeyWeight(key, unsigned(pl / 3))) / 7;
            pre[idx]   = encodePlaceholderBit(key, pl);
            order[idx] = idx;
        }
    }
    std::vector<std::pair<uint64_t, int>> kv(n);
    for (int i = 0; i < n; ++i)
        kv[i] = {pre[i], order[i]};
    std::stable_sort(kv.begin(), kv.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    o.prefixes.resize(n);
    o.internalToLeaf.resize(n);
    o.leafToInternal.resize(n);
    for (int i = 0; i < n; ++i)
        o.prefixes[i] = kv[i].first, o.internalToLeaf[i] = kv[i].second;
    for (int i = 0; i < n; ++i)
        o.leafToInternal[o.internalToLeaf[i]] = i;
    for (int i = 0; i < n; ++i)
        o.internalToLeaf[i] -= nI;
    o.levelRange.resize(kMaxLevel + 2);
    for (unsigned l = 0; l <= kMaxLevel; ++l)
        o.levelRange[l] = int(std::lower_bound(o.prefixes.begin(), o.prefixes.end(), encodePlaceholderBit(0, 3 * l)) -
                              o.prefixes.begin());
    o.levelRange[kMaxLevel + 1] = n;
    o.childOffsets.assign(n + 1, 0);
    o.parents.assign(std::max(1, (n - 1) / 8), 0);
    for (int i = 0; i < nI; ++i)
    {
        int a           = o.leafToInternal[i];
        uint64_t prefix = o.prefixes[a];
        uint64_t key    = decodePlaceholderBit(prefix);
        unsigned pl     = decodePrefixLength(prefix);
        unsigned level  = pl / 3;
        uint64_t child  = encodePlaceholderBit(key, int(pl) + 3);
        int lo = o.levelRange[level + 1], hi = o.levelRange[level + 2];
        int c = int(std::lower_bound(o.prefixes.begin() + lo, o.prefixes.begin() + hi, child) - o.prefixes.begin());
        if (c != hi && o.prefixes[c] == child)
        {
            o.childOffsets[a]        = c;
            o.parents[(c - 1) / 8] = a;
        }
    }
    return o;
}

/* ---------------------------------------------- centers and the MAC ------ */

// cstone/include/cstone/focus/source_center.hpp:46-98
template<class T>
inline void addBody(T* c, T sx, T sy, T sz, T sm)
{
    T w = std::abs(sm);
    c[0] += w * sx;
    c[1] += w * sy;
    c[2] += w * sz;
    c[3] += w;
}
template<class T>
inline void normalizeMass(T* c)
{
    T invM = (c[3] != T(0.0)) ? T(1.0) / c[3] : T(1.0);
    c[0] *= invM;
    c[1] *= invM;
    c[2] *= invM;
}

// source_center.hpp:112-127 then octree.hpp:377-396 with CombineSourceCenter
template<class Tc, class Tm, class Ts>
void computeCenters(const Tc* x, const Tc* y, const Tc* z, const Tm* m, const Octree& o, const unsigned* layout,
                    Ts* centers /* [numNodes][4] */)
{
    for (int leaf = 0; leaf < o.numLeaves; ++leaf)
    {
        int i  = o.leafToInternal[o.numInternal + leaf];
        Ts* c  = centers + 4 * size_t(i);
        c[0] = c[1] = c[2] = c[3] = 0;
        for (unsigned p = layout[leaf]; p < layout[leaf + 1]; ++p)
            addBody(c, Ts(x[p]), Ts(y[p]), Ts(z[p]), Ts(m[p]));
        normalizeMass(c);
    }
    for (int level = int(kMaxLevel); level >= 0; --level)
        for (int i = o.levelRange[level]; i < o.levelRange[level + 1]; ++i)
        {
            int ch = o.childOffsets[i];
            if (!ch) continue;
            Ts c[4] = {0, 0, 0, 0};
            for (int k = ch; k < ch + 8; ++k)
            {
                const Ts* s = centers + 4 * size_t(k);
                addBody(c, s[0], s[1], s[2], s[3]);
            }
            normalizeMass(c);
            std::copy(c, c + 4, centers + 4 * size_t(i));
        }
}

// cstone/include/cstone/sfc/box.hpp:357-381 (cubic and mixed boxes) with hilbert.hpp:549-585
template<class T>
inline void nodeGeometry(uint64_t prefix, const Box<T>& box, T* center, T* size)
{
    uint64_t key   = decodePlaceholderBit(prefix);
    unsigned level = decodePrefixLength(prefix) / 3;
    unsigned ic[3] = {0, 0, 0};
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}; // an unreachable key has the empty box at the origin
    if (decodeHilbertMixD(key, box.bits, ic[0], ic[1], ic[2]))
        for (int d = 0; d < 3; ++d)
        {
            unsigned cube = 1u << std::min(box.bits[d], kMaxLevel - level);
            lo[d]         = int(ic[d] & ~(cube - 1));
            hi[d]         = lo[d] + int(cube);
        }
    constexpr T uL = T(1.) / T(1u << kMaxLevel);
    for (int d = 0; d < 3; ++d)
    {
        T half    = T(0.5) * uL * box.len[d] * T(1u << (kMaxLevel - box.bits[d]));
        center[d] = box.lo[d] + T(hi[d] + lo[d]) * half;
        size[d]   = T(hi[d] - lo[d]) * half;
    }
}

// cstone/include/cstone/traversal/macs.hpp:64-82 via source_center.hpp:130-143
template<class T>
void setMac(const Octree& o, T* centers, float invTheta, const Box<T>& box)
{
    for (int i = 0; i < o.numNodes; ++i)
    {
        T gc[3], gs[3];
        nodeGeometry(o.prefixes[i], box, gc, gs);
        T* c     = centers + 4 * size_t(i);
        T dx[3]  = {c[0] - gc[0], c[1] - gc[1], c[2] - gc[2]};
        T s      = std::sqrt(dx[0] * dx[0] + (dx[1] * dx[1] + dx[2] * dx[2]));
        T l      = T(2.0) * std::max(std::max(gs[0], gs[1]), gs[2]);
        T mac    = l * invTheta + s;
        T macSq  = mac * mac;
        c[3]     = (c[3] != T(0)) ? macSq : T(0);
    }
}

/* -------------------------------------------------------- multipoles ----- */

// ryoanji/src/ryoanji/nbody/cartesian_qpole.hpp:67-117 (order: M, qxx, qxy, qxz, qyy, qyz, qzz, trace)
template<class Tc, class Tm, class Tq>
void P2M(const Tc* x, const Tc* y, const Tc* z, const Tm* m, unsigned b, unsigned e, const Tc* center, Tq* q)
{
    for (int k = 0; k < 8; ++k)
        q[k] = 0;
    for (unsigned i = b; i < e; ++i)
    {
        Tc rx = x[i] - center[0], ry = y[i] - center[1], rz = z[i] - center[2];
        Tc mi = Tc(m[i]);
        q[0] += mi;
        q[1] += rx * rx * mi;
        q[2] += rx * ry * mi;
        q[3] += rx * rz * mi;
        q[4] += ry * ry * mi;
        q[5] += ry * rz * mi;
        q[6] += rz * rz * mi;
    }
    Tq tr = q[1] + q[4] + q[6];
    q[7]  = tr;
    q[1]  = 3 * q[1] - tr;
    q[4]  = 3 * q[4] - tr;
    q[6]  = 3 * q[6] - tr;
    q[2] *= 3;
    q[3] *= 3;
    q[5] *= 3;
}

// cartesian_qpole.hpp:187-210 (parallel axis theorem)
template<class Tq, class Tc>
inline void addQuadrupole(Tq* c, Tc rx, Tc ry, Tc rz, const Tq* a)
{
    Tc rx2 = rx * rx, ry2 = ry * ry, rz2 = rz * rz;
    Tc r2  = (rx2 + ry2 + rz2) * (1.0 / 3.0);
    Tc ml  = a[0] * 3;
    c[7]   = c[7] + a[7] + ml * r2;
    c[0] += a[0];
    c[1] += a[1] + ml * (rx2 - r2);
    c[2] += a[2] + ml * rx * ry;
    c[3] += a[3] + ml * rx * rz;
    c[4] += a[4] + ml * (ry2 - r2);
    c[5] += a[5] + ml * ry * rz;
    c[6] += a[6] + ml * (rz2 - r2);
}

// upsweep_cpu.hpp:42-74 with cartesian_qpole.hpp:223-235 (levels 21..0, root included)
template<class Tc, class Tm, class Tq>
void multipoles(const Tc* x, const Tc* y, const Tc* z, const Tm* m, const Octree& o, const unsigned* layout,
                const Tc* centers, Tq* mp /* [numNodes][8] */)
{
    for (int leaf = 0; leaf < o.numLeaves; ++leaf)
    {
        int i = o.leafToInternal[o.numInternal + leaf];
        P2M(x, y, z, m, layout[leaf], layout[leaf + 1], centers + 4 * size_t(i), mp + 8 * size_t(i));
    }
    for (int level = int(kMaxLevel); level >= 0; --level)
        for (int i = o.levelRange[level]; i < o.levelRange[level + 1]; ++i)
        {
            int ch = o.childOffsets[i];
            if (!ch) continue;
            Tq out[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            const Tc* X = centers + 4 * size_t(i);
            for (int k = ch; k < ch + 8; ++k)
            {
                const Tc* Xi = centers + 4 * size_t(k);
                addQuadrupole(out, X[0] - Xi[0], X[1] - Xi[1], X[2] - Xi[2], mp + 8 * size_t(k));
            }
            std::copy(out, out + 8, mp + 8 * size_t(i));
        }
}

/* ------------------------------------------------------- P2P and M2P ----- */

template<class T>
inline T rsqrtFast(T x);
template<>
inline float rsqrtFast(float x) { return 1.0f / std::sqrt(x); } // kernel.hpp:23-36 host branch
template<>
inline double rsqrtFast(double x) { return double(1.0f / std::sqrt(float(x))); } // kernel.hpp:38-53 host branch
template<class T>
inline T rsqrtExact(T x) { return T(1) / std::sqrt(x); } // kernel.hpp:64-79 host branch

// ryoanji/src/ryoanji/nbody/kernel.hpp:99-120
template<bool Exact, class T>
inline void P2P(T* acc, const T* pi, T xj, T yj, T zj, T mj, T hi, T hj)
{
    T dx = xj - pi[0], dy = yj - pi[1], dz = zj - pi[2];
    T R2    = dx * dx + (dy * dy + dz * dz); // util::dot is a right fold
    T hij   = hi + hj;
    T hij2  = hij * hij;
    T R2eff = (R2 < hij2) ? hij2 : R2;
    T invR   = Exact ? rsqrtExact(R2eff) : rsqrtFast(R2eff);
    T invR2  = invR * invR;
    T invR3m = mj * invR * invR2;
    acc[0] -= invR3m * R2;
    acc[1] += dx * invR3m;
    acc[2] += dy * invR3m;
    acc[3] += dz * invR3m;
}

// cartesian_qpole.hpp:152-176 (Exact: a correctly rounded reciprocal square root, for error measurements)
template<class T, bool Exact = false>
inline void M2P(T* acc, const T* t, const T* c, const T* q)
{
    T rx = t[0] - c[0], ry = t[1] - c[1], rz = t[2] - c[2];
    T r2  = rx * rx + (ry * ry + rz * rz);
    T r1  = Exact ? rsqrtExact(r2) : rsqrtFast(r2);
    T rm2 = r1 * r1;
    T rm5 = rm2 * rm2 * r1;
    T Qrx = rx * q[1] + ry * q[2] + rz * q[3];
    T Qry = rx * q[2] + ry * q[4] + rz * q[5];
    T Qrz = rx * q[3] + ry * q[5] + rz * q[6];
    T rQr = rx * Qrx + ry * Qry + rz * Qrz;
    T mono = (T(-2.5) * rQr * rm5 - q[0] * r1) * rm2;
    acc[0] += -(q[0] * r1 + T(0.5) * rm5 * rQr);
    acc[1] += rm5 * Qrx + mono * rx;
    acc[2] += rm5 * Qry + mono * ry;
    acc[3] += rm5 * Qrz + mono * rz;
}

/* --------------------------------------------------------- traversal ----- */

// cstone/include/cstone/traversal/macs.hpp:93-101
template<class T>
inline bool violatesMac(const T* com, T macSq, const T* tc, const T* ts)
{
    T d[3];
    for (int k = 0; k < 3; ++k)
    {
        d[k] = std::abs(tc[k] - com[k]) - ts[k];
        d[k] += std::abs(d[k]);
        d[k] *= T(0.5);
    }
    T R2 = d[0] * d[0] + (d[1] * d[1] + d[2] * d[2]);
    return R2 < std::abs(macSq);
}

struct TravStats
{
    uint64_t sumP2P = 0, sumM2P = 0;
};

/* ryoanji/src/ryoanji/nbody/traversal_cpu.hpp:67-141 with cstone traversal.hpp:26-69 (DFS from the root).
 * `nt` valid targets; the group bounding box uses the valid targets only. */
template<class T, bool Exact = false>
void gravityGroup(const T (*tgt)[4], int nt, const Octree& o, const T* centers, const T* mp, const unsigned* layout,
                  const T* x, const T* y, const T* z, const T* h, const T* m, T (*acc)[4], TravStats* st)
{
    T tmin[3], tmax[3];
    for (int k = 0; k < 3; ++k)
        tmin[k] = tmax[k] = tgt[0][k];
    for (int i = 1; i < nt; ++i)
        for (int k = 0; k < 3; ++k)
        {
            tmin[k] = std::min(tgt[i][k], tmin[k]);
            tmax[k] = std::max(tgt[i][k], tmax[k]);
        }
    T tc[3], ts[3];
    for (int k = 0; k < 3; ++k)
    {
        tc[k] = (tmax[k] + tmin[k]) * T(0.5);
        ts[k] = (tmax[k] - tmin[k]) * T(0.5);
    }

    auto descendOrM2P = [&](int idx) -> bool
    {
        const T* com = centers + 4 * size_t(idx);
        bool close   = violatesMac(com, com[3], tc, ts);
        int leaf     = o.internalToLeaf[idx];
        bool remote  = leaf >= 0 && layout[leaf] == layout[leaf + 1] && com[3] != 0;
        if (!close || remote)
        {
            for (int k = 0; k < nt; ++k)
                M2P<T, Exact>(acc[k], tgt[k], com, mp + 8 * size_t(idx));
            if (st) st->sumM2P += nt;
        }
        return close;
    };
    auto leafP2P = [&](int idx)
    {
        int leaf = o.internalToLeaf[idx];
        unsigned b = layout[leaf], e = layout[leaf + 1];
        for (int k = 0; k < nt; ++k)
            for (unsigned s = b; s < e; ++s)
                P2P<Exact>(acc[k], tgt[k], x[s], y[s], z[s], m[s], tgt[k][3], h[s]);
        if (st) st->sumP2P += uint64_t(nt) * (e - b);
    };

    // traversal.hpp:26-69
    const int* child   = o.childOffsets.data();
    const int* parents = o.parents.data();
    if (!descendOrM2P(0)) return;
    if (child[0] == 0)
    {
        leafP2P(0);
        return;
    }
    int node       = child[0];
    bool backtrack = false;
    while (node != 0)
    {
        bool isLeaf  = child[node] == 0;
        bool descend = !backtrack && descendOrM2P(node);
        if (isLeaf && descend) leafP2P(node);
        int sib = (node - 1) % 8;
        if (!isLeaf && descend)
        {
            node      = child[node];
            backtrack = false;
        }
        else if (sib < 7)
        {
            node++;
            backtrack = false;
        }
        else
        {
            node      = parents[(node - 1) / 8];
            backtrack = true;
        }
    }
}

/* traversal_cpu.hpp:172-230: groups of `groupSize` consecutive targets (the reference uses 16), all
 * (2s+1)^3 images, outputs added to ugrav/ax/ay/az, returns 0.5 * sum(G m phi). */
template<class T, bool Exact = false>
double computeGravity(const Octree& o, const T* centers, const T* mp, const unsigned* layout, const T* x, const T* y,
                      const T* z, const T* h, const T* m, const Box<T>& box, float G, int numShells, int groupSize,
                      T* ugrav, T* ax, T* ay, T* az, const int* groupList, int numGroupList, TravStats* stats,
                      const unsigned* bounds = nullptr, long numBounded = 0)
{
    // groups: fixed groupSize-target groups of [layout[0], layout[L]), or bounds[g]..bounds[g+1] when given
    const unsigned first = layout[0], last = layout[o.numLeaves];
    long numGroups       = bounds ? numBounded : (long(last - first) + groupSize - 1) / groupSize;
    long work            = groupList ? numGroupList : numGroups;
    double pot           = 0;
    uint64_t p2p = 0, m2p = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : pot, p2p, m2p)
    for (long gi = 0; gi < work; ++gi)
    {
        long g       = groupList ? groupList[gi] : gi;
        unsigned i0  = bounds ? bounds[g] : first + unsigned(g) * groupSize;
        int nt       = bounds ? int(bounds[g + 1] - bounds[g]) : int(std::min<long>(groupSize, long(last) - i0));
        std::vector<T> tbuf(4 * size_t(nt)), sbuf(4 * size_t(nt)), abuf(4 * size_t(nt), T(0));
        auto tgt     = reinterpret_cast<T(*)[4]>(tbuf.data());
        auto shifted = reinterpret_cast<T(*)[4]>(sbuf.data());
        auto acc     = reinterpret_cast<T(*)[4]>(abuf.data());
        for (int k = 0; k < nt; ++k)
        {
            tgt[k][0] = x[i0 + k], tgt[k][1] = y[i0 + k], tgt[k][2] = z[i0 + k], tgt[k][3] = T(h[i0 + k]);
        }
        TravStats st;
        for (int iz = -numShells; iz <= numShells; ++iz)
            for (int iy = -numShells; iy <= numShells; ++iy)
                for (int ix = -numShells; ix <= numShells; ++ix)
                {
                    T sh[3] = {ix * box.len[0], iy * box.len[1], iz * box.len[2]};
                    for (int k = 0; k < nt; ++k)
                    {
                        for (int d = 0; d < 3; ++d)
                            shifted[k][d] = tgt[k][d] - sh[d];
                        shifted[k][3] = tgt[k][3] - T(0);
                    }
                    gravityGroup<T, Exact>(shifted, nt, o, centers, mp, layout, x, y, z, h, m, acc, &st);
                }
        for (int k = 0; k < nt; ++k)
        {
            T u = G * m[i0 + k] * acc[k][0];
            pot += double(u);
            if (ugrav) ugrav[i0 + k] += u;
            ax[i0 + k] += G * acc[k][1];
            ay[i0 + k] += G * acc[k][2];
            az[i0 + k] += G * acc[k][3];
        }
        p2p += st.sumP2P;
        m2p += st.sumM2P;
    }
    if (stats) stats->sumP2P = p2p, stats->sumM2P = m2p;
    return 0.5 * pot;
}

/* -------------------------------------------------------- direct sum ----- */

// traversal_cpu.hpp:232-268 (P2P<Exact>, all images), for a list of targets
template<class Tin, class T>
void directSum(const Tin* x, const Tin* y, const Tin* z, const Tin* h, const Tin* m, size_t n, const int64_t* targets,
               size_t numTargets, double G, const double* boxL, int numShells, T* out /* [numTargets][4] pot,ax,ay,az */)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t ti = 0; ti < numTargets; ++ti)
    {
        size_t t  = targets ? size_t(targets[ti]) : ti;
        T acc[4]  = {0, 0, 0, 0};
        T th      = T(h[t]);
        for (int iz = -numShells; iz <= numShells; ++iz)
            for (int iy = -numShells; iy <= numShells; ++iy)
                for (int ix = -numShells; ix <= numShells; ++ix)
                {
                    T p[3] = {T(x[t]) + T(ix * boxL[0]), T(y[t]) + T(iy * boxL[1]), T(z[t]) + T(iz * boxL[2])};
                    for (size_t s = 0; s < n; ++s)
                        P2P<true>(acc, p, T(x[s]), T(y[s]), T(z[s]), T(m[s]), th, T(h[s]));
                }
        out[4 * ti + 0] = T(G) * T(m[t]) * acc[0];
        out[4 * ti + 1] = T(G) * acc[1];
        out[4 * ti + 2] = T(G) * acc[2];
        out[4 * ti + 3] = T(G) * acc[3];
    }
}

/* ------------------------------------------------------------- Ewald ----- */

// ryoanji/src/ryoanji/nbody/ewald.hpp:42-75
template<class T>
struct EwaldParams
{
    T Mroot[8];
    T center[3];
    int numReplicaShells, numEwaldShells;
    double L, lCut, hCut, alphaScale, smallR;
    int numH;
    T hr[343][3], hrScaled[343][3], hfacCos[343], hfacSin[343];
};

// ewald.hpp:90-115
template<class T>
inline void ewaldEval(T* potAcc, const T* r, const T* g, const T* mp)
{
    T M   = mp[0];
    T qxx = (mp[1] + mp[7]) / T(3), qyy = (mp[4] + mp[7]) / T(3), qzz = (mp[6] + mp[7]) / T(3);
    T qxy = mp[2] / T(3), qxz = mp[3] / T(3), qyz = mp[5] / T(3);
    T Qr[3] = {r[0] * qxx + r[1] * qxy + r[2] * qxz, r[0] * qxy + r[1] * qyy + r[2] * qyz,
               r[0] * qxz + r[1] * qyz + r[2] * qzz};
    T rQr = 0.5 * (r[0] * Qr[0] + (r[1] * Qr[1] + r[2] * Qr[2]));
    T Qtr = 0.5 * mp[7];
    T u   = -g[0] * M + g[1] * Qtr - g[2] * rQr;
    T f   = +g[1] * M - g[2] * Qtr + g[3] * rQr;
    potAcc[0] += u;
    for (int d = 0; d < 3; ++d)
        potAcc[1 + d] += g[2] * Qr[d] - r[d] * f;
}

// ewald.hpp:133-198
template<class T>
void ewaldInit(EwaldParams<T>& p, const T* Mroot, const T* center, int numReplicaShells, double L, double lCut,
               double hCut, double alphaScale, double smallR)
{
    if (lCut == 0 && hCut == 0 && alphaScale == 0) numReplicaShells = 0;
    std::copy(Mroot, Mroot + 8, p.Mroot);
    std::copy(center, center + 3, p.center);
    p.numReplicaShells = numReplicaShells;
    p.numEwaldShells   = std::max(int(std::ceil(lCut)), numReplicaShells);
    p.L = L, p.lCut = lCut, p.hCut = hCut, p.alphaScale = alphaScale, p.smallR = smallR;
    p.numH = 0;
    if (p.numEwaldShells == 0) return;
    int hReps     = int(std::ceil(hCut));
    double alpha  = alphaScale / L;
    double k4     = M_PI * M_PI / (alpha * alpha * L * L);
    double hCut2  = hCut * hCut;
    for (int hx = -hReps; hx <= hReps; hx++)
        for (int hy = -hReps; hy <= hReps; hy++)
            for (int hz = -hReps; hz <= hReps; hz++)
            {
                T hr[3] = {T(hx), T(hy), T(hz)};
                T h2    = hr[0] * hr[0] + (hr[1] * hr[1] + hr[2] * hr[2]);
                if (h2 == 0 || h2 > hCut2) continue;
                T g0 = std::exp(-k4 * h2) / (M_PI * h2 * L);
                T g1 = 2 * M_PI / L * g0;
                T g2 = -2 * M_PI / L * g1;
                T g3 = 2 * M_PI / L * g2;
                T g4 = -2 * M_PI / L * g3;
                T g5 = 2 * M_PI / L * g4;
                T gc[6] = {g0, 0, g2, 0, g4, 0}, gs[6] = {0, g1, 0, g3, 0, g5};
                T a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
                ewaldEval(a, hr, gc, p.Mroot);
                ewaldEval(b, hr, gs, p.Mroot);
                int k = p.numH++;
                for (int d = 0; d < 3; ++d)
                {
                    p.hr[k][d]       = hr[d];
                    p.hrScaled[k][d] = 2 * M_PI / L * hr[d];
                }
                p.hfacCos[k] = a[0];
                p.hfacSin[k] = b[0];
            }
}

// ewald.hpp:201-327
template<class T>
void ewaldParticle(const EwaldParams<T>& p, const T* pos, T* out, int kSign)
{
    const T lCut = p.lCut, L = p.L;
    const T lCut2 = lCut * lCut * L * L;
    const T alpha = p.alphaScale / L, alpha2 = alpha * alpha;
    const T k1 = M_PI / (alpha2 * L * L * L);
    const T ka = 2.0 * alpha / std::sqrt(M_PI);
    const T smallR2 = p.smallR * L * L;
    T potAcc[4] = {k1 * p.Mroot[0], 0, 0, 0};
    T r[3] = {pos[0] - p.center[0], pos[1] - p.center[1], pos[2] - p.center[2]};
    const int ns = p.numEwaldShells, nr = p.numReplicaShells;
    for (int ix = -ns; ix <= ns; ix++)
        for (int iy = -ns; iy <= ns; iy++)
            for (int iz = -ns; iz <= ns; iz++)
            {
                bool inner = (-nr <= ix && ix <= nr) && (-nr <= iy && iy <= nr) && (-nr <= iz && iz <= nr);
                T R[3]  = {r[0] + ix * L, r[1] + iy * L, r[2] + iz * L};
                T R2    = R[0] * R[0] + (R[1] * R[1] + R[2] * R[2]);
                if (R2 > lCut2 && !inner) continue;
                T g[6];
                if (R2 < smallR2 && ka > 0)
                {
                    T c0 = ka, R2a2 = R2 * alpha2;
                    g[0] = c0 * (R2a2 / 3.0 - 1.0);
                    c0 *= 2 * alpha2;
                    g[1] = c0 * (R2a2 / 5.0 - 1.0 / 3.0);
                    c0 *= 2 * alpha2;
                    g[2] = c0 * (R2a2 / 7.0 - 1.0 / 5.0);
                    c0 *= 2 * alpha2;
                    g[3] = c0 * (R2a2 / 9.0 - 1.0 / 7.0);
                    c0 *= 2 * alpha2;
                    g[4] = c0 * (R2a2 / 11.0 - 1.0 / 9.0);
                    c0 *= 2 * alpha2;
                    g[5] = c0 * (R2a2 / 13.0 - 1.0 / 11.0);
                }
                else
                {
                    T Rm = std::sqrt(R2), invR = 1.0 / Rm, invR2 = invR * invR;
                    T a  = std::exp(-R2 * alpha2) * ka * invR2;
                    T an = T(1.0);
                    T fn = inner ? -std::erf(alpha * Rm) : std::erfc(alpha * Rm);
                    g[0] = fn * invR;
                    g[1] = g[0] * invR2 + a;
                    an *= 2 * alpha2;
                    g[2] = 3 * g[1] * invR2 + an * a;
                    an *= 2 * alpha2;
                    g[3] = 5 * g[2] * invR2 + an * a;
                    an *= 2 * alpha2;
                    g[4] = 7 * g[3] * invR2 + an * a;
                    an *= 2 * alpha2;
                    g[5] = 9 * g[4] * invR2 + an * a;
                }
                ewaldEval(potAcc, R, g, p.Mroot);
            }
    T k[4] = {0, 0, 0, 0};
    T dr[3] = {pos[0] - p.center[0], pos[1] - p.center[1], pos[2] - p.center[2]};
    for (int i = 0; i < p.numH; ++i)
    {
        T hx = p.hrScaled[i][0] * dr[0] + (p.hrScaled[i][1] * dr[1] + p.hrScaled[i][2] * dr[2]);
        T c = std::cos(hx), s = std::sin(hx);
        T sum = p.hfacCos[i] * c + p.hfacSin[i] * s;
        T dif = p.hfacCos[i] * s - p.hfacSin[i] * c;
        // kSign -1: the reference's k-space potential (ewald.hpp:320); +1: the potential whose negative gradient
        // is the k-space acceleration below
        k[0] += T(kSign) * sum;
        for (int d = 0; d < 3; ++d)
            k[1 + d] += dif * p.hrScaled[i][d];
    }
    for (int d = 0; d < 4; ++d)
        out[d] = potAcc[d] + k[d];
}

} // namespace orc

/* ============================================================ C ABI ===== */

using namespace orc;

extern "C" {

int orc_sfc_keys(const float* x, const float* y, const float* z, uint64_t* keys, size_t n, const float* limits,
                 int hilbert)
{
    auto box = makeBox<float>(limits, nullptr);
    sfcKeys(x, y, z, keys, n, box, hilbert != 0);
    return 0;
}

int orc_sfc_keys_f64(const double* x, const double* y, const double* z, uint64_t* keys, size_t n,
                     const double* limits, int hilbert)
{
    auto box = makeBox<double>(limits, nullptr);
    sfcKeys(x, y, z, keys, n, box, hilbert != 0);
    return 0;
}

uint64_t orc_hilbert3d(unsigned ix, unsigned iy, unsigned iz, int order) { return hilbert3D(ix, iy, iz, order); }

void orc_decode_hilbert3d(uint64_t key, unsigned* xyz) { decodeHilbert3D(key, xyz[0], xyz[1], xyz[2]); }

uint64_t orc_hilbert_mixd(unsigned ix, unsigned iy, unsigned iz, const unsigned* bits)
{
    return hilbertMixD(ix, iy, iz, bits);
}

/* returns 1 and the coordinates for a reachable key, 0 otherwise */
int orc_decode_hilbert_mixd(uint64_t key, const unsigned* bits, unsigned* xyz)
{
    return decodeHilbertMixD(key, bits, xyz[0], xyz[1], xyz[2]) ? 1 : 0;
}

void orc_box_bits(const float* limits, unsigned* bits)
{
    auto box = makeBox<float>(limits, nullptr);
    for (int d = 0; d < 3; ++d)
        bits[d] = box.bits[d];
}

void orc_sort_by_key(uint64_t* keys, int32_t* values, size_t n) { sortByKey(keys, values, n); }

/* leaves must hold capacity+1 entries; returns the number of leaves, or -(needed) if too small */
long orc_compute_octree(const uint64_t* sortedKeys, size_t n, unsigned bucket, uint64_t* leaves, unsigned* counts,
                        long capacity)
{
    std::vector<uint64_t> tree;
    std::vector<unsigned> cnt;
    computeOctree(sortedKeys, n, bucket, tree, cnt);
    long L = long(cnt.size());
    if (L > capacity) return -L;
    std::copy(tree.begin(), tree.end(), leaves);
    std::copy(cnt.begin(), cnt.end(), counts);
    return L;
}

/* outputs sized: prefixes/i2l/l2i [numNodes], childOffsets [numNodes+1], parents [max(1,(numNodes-1)/8)],
 * levelRange [23] */
int orc_link_octree(const uint64_t* leaves, int numLeaves, uint64_t* prefixes, int* childOffsets, int* parents,
                    int* levelRange, int* internalToLeaf, int* leafToInternal)
{
    Octree o = linkOctree(leaves, numLeaves);
    std::copy(o.prefixes.begin(), o.prefixes.end(), prefixes);
    std::copy(o.childOffsets.begin(), o.childOffsets.end(), childOffsets);
    std::copy(o.parents.begin(), o.parents.end(), parents);
    std::copy(o.levelRange.begin(), o.levelRange.end(), levelRange);
    std::copy(o.internalToLeaf.begin(), o.internalToLeaf.end(), internalToLeaf);
    std::copy(o.leafToInternal.begin(), o.leafToInternal.end(), leafToInternal);
    return o.numNodes;
}

/* The full float pipeline after sorting (the sequence of ryoanji/test/nbody/traversal_cpu.cpp:41-90):
 * centers (float), setMac(invTheta), multipoles (float). */
static Octree rebuild(const uint64_t* leaves, int numLeaves) { return linkOctree(leaves, numLeaves); }

int orc_upsweep_f32(const float* x, const float* y, const float* z, const float* m, const uint64_t* leaves,
                    int numLeaves, const unsigned* layout, const float* limits, float invTheta, float* centers,
                    float* mpoles)
{
    Octree o = rebuild(leaves, numLeaves);
    auto box = makeBox<float>(limits, nullptr);
    computeCenters<float, float, float>(x, y, z, m, o, layout, centers);
    multipoles<float, float, float>(x, y, z, m, o, layout, centers, mpoles);
    setMac<float>(o, centers, invTheta, box);
    return 0;
}

int orc_upsweep_f64(const float* x, const float* y, const float* z, const float* m, const uint64_t* leaves,
                    int numLeaves, const unsigned* layout, const double* limits, float invTheta, double* centers,
                    double* mpoles)
{
    Octree o = rebuild(leaves, numLeaves);
    auto box = makeBox<double>(limits, nullptr);
    std::vector<double> xd(x, x + layout[numLeaves]), yd(y, y + layout[numLeaves]), zd(z, z + layout[numLeaves]),
        md(m, m + layout[numLeaves]);
    computeCenters<double, double, double>(xd.data(), yd.data(), zd.data(), md.data(), o, layout, centers);
    multipoles<double, double, double>(xd.data(), yd.data(), zd.data(), md.data(), o, layout, centers, mpoles);
    setMac<double>(o, centers, invTheta, box);
    return 0;
}

/* computeGravity (float instantiation). numShells > 0 uses the box lengths of limits. groupList, if given,
 * restricts the traversal to those group indices (for a timed sample); stats gets {sumP2P, sumM2P}. */
double orc_compute_gravity_f32(const uint64_t* leaves, int numLeaves, const unsigned* layout, const float* centers,
                               const float* mpoles, const float* x, const float* y, const float* z, const float* h,
                               const float* m, const float* limits, float G, int numShells, int groupSize,
                               float* ugrav, float* ax, float* ay, float* az, const int* groupList, int numGroupList,
                               uint64_t* stats)
{
    Octree o = rebuild(leaves, numLeaves);
    auto box = makeBox<float>(limits, nullptr);
    TravStats st;
    double pot = computeGravity<float>(o, centers, mpoles, layout, x, y, z, h, m, box, G, numShells, groupSize,
                                       ugrav, ax, ay, az, groupList, numGroupList, &st);
    if (stats) stats[0] = st.sumP2P, stats[1] = st.sumM2P;
    return pot;
}

/* computeGravity over arbitrary target groups bounds[g]..bounds[g+1] (g < numGroups), e.g. split groups */
double orc_compute_gravity_groups_f32(const uint64_t* leaves, int numLeaves, const unsigned* layout,
                                      const float* centers, const float* mpoles, const float* x, const float* y,
                                      const float* z, const float* h, const float* m, const float* limits, float G,
                                      int numShells, const unsigned* bounds, long numGroups, float* ugrav, float* ax,
                                      float* ay, float* az, uint64_t* stats)
{
    Octree o = rebuild(leaves, numLeaves);
    auto box = makeBox<float>(limits, nullptr);
    TravStats st;
    double pot = computeGravity<float>(o, centers, mpoles, layout, x, y, z, h, m, box, G, numShells, 64, ugrav, ax,
                                       ay, az, nullptr, 0, &st, bounds, numGroups);
    if (stats) stats[0] = st.sumP2P, stats[1] = st.sumM2P;
    return pot;
}

/* the same traversal evaluated in double with correctly rounded reciprocal square roots (inputs widened from
 * float): the float path's rounding error is measured against it on the same tree, grouping and MAC inputs */
double orc_compute_gravity_f64(const uint64_t* leaves, int numLeaves, const unsigned* layout, const float* centers,
                               const float* mpoles, const float* x, const float* y, const float* z, const float* h,
                               const float* m, const float* limits, double G, int numShells, int groupSize,
                               double* ugrav, double* ax, double* ay, double* az)
{
    Octree o     = rebuild(leaves, numLeaves);
    double lim[6];
    for (int k = 0; k < 6; ++k)
        lim[k] = limits[k];
    auto box = makeBox<double>(lim, nullptr);
    size_t n = layout[numLeaves], nn = size_t(o.numNodes);
    std::vector<double> c(centers, centers + 4 * nn), q(mpoles, mpoles + 8 * nn), xd(x, x + n), yd(y, y + n),
        zd(z, z + n), hd(h, h + n), md(m, m + n);
    return computeGravity<double, true>(o, c.data(), q.data(), layout, xd.data(), yd.data(), zd.data(), hd.data(),
                                  md.data(), box, float(G), numShells, groupSize, ugrav, ax, ay, az, nullptr, 0,
                                  nullptr);
}

/* computeGroupSplits (cstone/traversal/groups_gpu.cu:60-151, groups_gpu.cuh:29-270) for 64-target windows in
 * float: returns the number of groups G and writes G+1 boundaries to groups (capacity entries), or -(G+1) when
 * the capacity is too small. Windows are processed one after the other, target by target. */
long orc_group_splits(unsigned first, unsigned last, const float* x, const float* y, const float* z, const float* h,
                      const uint64_t* leaves, int numLeaves, const unsigned* layout, const float* limits,
                      float tolFactor, float extentRatio, unsigned* groups, long capacity)
{
    const unsigned W = 64;
    auto box         = makeBox<float>(limits, nullptr);
    const float inv  = 1.0f / std::max(std::max(box.len[0], box.len[1]), box.len[2]);
    std::vector<unsigned> out;
    for (unsigned ws = first; ws < last; ws += W)
    {
        float P[64][4], volMin = 1.0f, edge[3] = {1.0f, 1.0f, 1.0f};
        for (unsigned l = 0; l < W; ++l)
        {
            unsigned i = std::min(ws + l, last - 1);
            // the leaf holding particle i: layout[leaf] <= i < layout[leaf + 1]
            int leaf = int(std::upper_bound(layout, layout + numLeaves, i) - layout) - 1;
            uint64_t k0 = leaves[leaf], k1 = leaves[leaf + 1];
            unsigned level = kMaxLevel - unsigned(63 - clz64(k1 - k0)) / 3u;
            // cell edge in the unit box (half-unit length 2^-22 on every axis) from the mixed-dimension cell box
            unsigned c[3];
            float sz[3] = {0, 0, 0};
            if (decodeHilbertMixD(k0, box.bits, c[0], c[1], c[2]))
                for (int d = 0; d < 3; ++d)
                    sz[d] = float(1u << std::min(box.bits[d], kMaxLevel - level)) * (1.0f / float(1u << 22));
            float vol = 8.0f * sz[0] * sz[1] * sz[2];
            if (vol > 0.0f) volMin = std::min(vol, volMin);
            for (int d = 0; d < 3; ++d)
                edge[d] = std::min(edge[d], 2.0f * sz[d]);
            P[l][0] = x[i] * inv, P[l][1] = y[i] * inv, P[l][2] = z[i] * inv;
            P[l][3] = h ? 2.0f * h[i] * inv : 1.0f;
        }
        float crit = std::cbrt(volMin) * tolFactor, crit2 = crit * crit;
        uint64_t mask = 0;
        for (unsigned l = 0; l + 1 < W; ++l)
        {
            float dx = P[l + 1][0] - P[l][0], dy = P[l + 1][1] - P[l][1], dz = P[l + 1][2] - P[l][2];
            float d2 = dx * dx + (dy * dy + dz * dz);
            if (d2 > std::min(crit2, P[l][3] * P[l][3])) mask |= 1ull << l;
        }
        float ratio = 0.0f;
        for (int d = 0; d < 3; ++d)
        {
            float lo = P[0][d], hi = P[0][d];
            for (unsigned l = 1; l < W; ++l)
                lo = std::min(lo, P[l][d]), hi = std::max(hi, P[l][d]);
            float allowed = tolFactor * edge[d];
            if (allowed > 0.0f) ratio = std::max(ratio, (hi - lo) / allowed);
        }
        if (extentRatio > 0.0f && ratio > extentRatio)
        {
            unsigned chunk = W;
            while (chunk > 1 && float(chunk) * ratio > float(W) * extentRatio)
                chunk /= 2;
            unsigned inWindow = std::min(W, last - ws);
            for (unsigned next = chunk; next < inWindow; next += chunk)
                mask |= 1ull << (next - 1);
        }
        out.push_back(ws);
        for (unsigned l = 0; l < W; ++l)
            if (mask >> l & 1u) out.push_back(ws + l + 1);
    }
    out.push_back(last);
    long G = long(out.size()) - 1;
    if (long(out.size()) > capacity) return -long(out.size());
    std::copy(out.begin(), out.end(), groups);
    return G;
}

/* FP64 direct sum over all n sources for the listed targets (or all, if targets == NULL);
 * out[4*i] = G m_t phi, out[4*i+1..3] = G a. */
int orc_direct_sum_f64(const float* x, const float* y, const float* z, const float* h, const float* m, size_t n,
                       const int64_t* targets, size_t numTargets, double G, const double* boxL, int numShells,
                       double* out)
{
    double zero[3] = {0, 0, 0};
    directSum<float, double>(x, y, z, h, m, n, targets, numTargets, G, boxL ? boxL : zero, numShells, out);
    return 0;
}

int orc_direct_sum_f32(const float* x, const float* y, const float* z, const float* h, const float* m, size_t n,
                       const int64_t* targets, size_t numTargets, double G, const double* boxL, int numShells,
                       float* out)
{
    double zero[3] = {0, 0, 0};
    directSum<float, float>(x, y, z, h, m, n, targets, numTargets, G, boxL ? boxL : zero, numShells, out);
    return 0;
}

/* Exact periodic potential and acceleration of softened particles in a cubic box of side L, in FP64, for the
 * listed targets (all, if targets == NULL): the Ewald sum over every source and every image, in the potential
 * convention of ewald.hpp:201-327 (-erfc(aR)/R real space, -(4 pi/L^3) exp(-k^2/4a^2)/k^2 cos(k.r) k-space,
 * +pi M/(a^2 L^3) background), except that a pair image closer than h_i + h_j contributes the softened law of
 * P2P<Exact> (kernel.hpp:99-120) in place of the point mass, as in the traversal. Written as the plain sum:
 * explicit images |n| <= 2 within rc = 5/a, the point-mass real-space term with the point mass swapped for the
 * softened law by difference, the self term as the R -> 0 limit 2a/sqrt(pi), and the k-space sum over the full
 * lattice |h| <= ceil(5.5 a L / pi) through complex structure factors. out[4 i] = {G m phi, G a}.
 * Pinned by the simple-cubic Madelung constant (one particle; eight on a sub-lattice), by independence of
 * alpha, and by the traversal + Ewald correction in the limit where the root multipole is exact. */
int orc_ewald_pair_sum_f64(const float* x, const float* y, const float* z, const float* h, const float* m, size_t n,
                           const int64_t* targets, size_t numTargets, double G, double L, double alphaScale,
                           double* out)
{
    if (!(alphaScale > 5.0)) return -1;
    const double a = alphaScale / L, rc = 5.0 / a, pi = 3.14159265358979323846;
    const int hmax = int(std::ceil(5.5 * alphaScale / pi));
    struct Kv
    {
        double k[3], c;
        std::complex<double> S;
    };
    std::vector<Kv> kv;
    double M = 0;
    for (size_t j = 0; j < n; ++j)
        M += double(m[j]);
    for (int hx = -hmax; hx <= hmax; ++hx)
        for (int hy = -hmax; hy <= hmax; ++hy)
            for (int hz = -hmax; hz <= hmax; ++hz)
            {
                int h2 = hx * hx + hy * hy + hz * hz;
                if (h2 == 0 || h2 > hmax * hmax) continue;
                Kv e;
                e.k[0] = 2 * pi * hx / L, e.k[1] = 2 * pi * hy / L, e.k[2] = 2 * pi * hz / L;
                double k2 = e.k[0] * e.k[0] + e.k[1] * e.k[1] + e.k[2] * e.k[2];
                e.c = 4 * pi / (L * L * L) * std::exp(-k2 / (4 * a * a)) / k2;
                e.S = 0;
                for (size_t j = 0; j < n; ++j)
                    e.S += double(m[j]) * std::exp(std::complex<double>(0, e.k[0] * x[j] + e.k[1] * y[j] + e.k[2] * z[j]));
                kv.push_back(e);
            }
    const size_t nt = targets ? numTargets : n;
#pragma omp parallel for
    for (size_t q = 0; q < nt; ++q)
    {
        const size_t t = targets ? size_t(targets[q]) : q;
        const double p[3] = {x[t], y[t], z[t]};
        double phi = pi * M / (a * a * L * L * L), acc[3] = {0, 0, 0};
        for (size_t j = 0; j < n; ++j)
        {
            const double hij = double(h[t]) + double(h[j]), mj = m[j];
            for (int ix = -2; ix <= 2; ++ix)
                for (int iy = -2; iy <= 2; ++iy)
                    for (int iz = -2; iz <= 2; ++iz)
                    {
                        double r[3] = {x[j] + ix * L - p[0], y[j] + iy * L - p[1], z[j] + iz * L - p[2]};
                        double R2   = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                        if (R2 >= rc * rc) continue;
                        if (R2 == 0)
                        {
                            phi += mj * 2 * a / std::sqrt(pi); // the self term (and exact duplicates)
                            continue;
                        }
                        double R = std::sqrt(R2);
                        double fpot = -std::erfc(a * R) / R;
                        double ffac = std::erfc(a * R) / (R2 * R) + 2 * a / std::sqrt(pi) * std::exp(-a * a * R2) / R2;
                        if (R < hij) // the softened law in place of the point mass: + (-R^2/h^3) - (-1/R)
                        {
                            fpot += -R2 / (hij * hij * hij) + 1.0 / R;
                            ffac += 1.0 / (hij * hij * hij) - 1.0 / (R2 * R);
                        }
                        phi += mj * fpot;
                        for (int d = 0; d < 3; ++d)
                            acc[d] += mj * ffac * r[d];
                    }
        }
        for (const Kv& e : kv)
        {
            std::complex<double> v = e.S * std::exp(std::complex<double>(0, -(e.k[0] * p[0] + e.k[1] * p[1] + e.k[2] * p[2])));
            phi -= e.c * v.real();
            for (int d = 0; d < 3; ++d)
                acc[d] += e.c * e.k[d] * v.imag();
        }
        out[4 * q]     = G * double(m[t]) * phi;
        out[4 * q + 1] = G * acc[0];
        out[4 * q + 2] = G * acc[1];
        out[4 * q + 3] = G * acc[2];
    }
    return 0;
}

/* Ewald correction (ewald.hpp:356-389) in T = double: adds G*m*phi (ugrav), G*a; returns 0.5 G sum(m phi).
 * settings = {numReplicaShells, lCut, hCut, alphaScale, smallR}. */
static double ewaldAll(const double* Mroot, const double* center, const float* x, const float* y, const float* z,
                     const float* m, size_t n, double L, double G, const double* settings, double* ugrav, double* ax,
                     double* ay, double* az, int kSign)
{
    auto* p = new EwaldParams<double>();
    ewaldInit(*p, Mroot, center, int(settings[0]), L, settings[1], settings[2], settings[3], settings[4]);
    double U = 0;
    if (p->numEwaldShells > 0)
    {
#pragma omp parallel for reduction(+ : U)
        for (size_t i = 0; i < n; ++i)
        {
            double pos[3] = {x[i], y[i], z[i]}, out[4];
            ewaldParticle(*p, pos, out, kSign);
            if (ugrav) ugrav[i] += G * out[0] * m[i];
            U += out[0] * m[i];
            ax[i] += G * out[1];
            ay[i] += G * out[2];
            az[i] += G * out[3];
        }
    }
    delete p;
    return 0.5 * G * U;
}

double orc_ewald_f64(const double* Mroot, const double* center, const float* x, const float* y, const float* z,
                     const float* m, size_t n, double L, double G, const double* settings, double* ugrav, double* ax,
                     double* ay, double* az)
{
    return ewaldAll(Mroot, center, x, y, z, m, n, L, G, settings, ugrav, ax, ay, az, -1);
}

/* the same correction with the k-space potential taken consistent with the k-space acceleration (the opposite
 * sign of ewald.hpp:320 for the potential term; accelerations unchanged) */
double orc_ewald_consistent_f64(const double* Mroot, const double* center, const float* x, const float* y,
                                const float* z, const float* m, size_t n, double L, double G, const double* settings,
                                double* ugrav, double* ax, double* ay, double* az)
{
    return ewaldAll(Mroot, center, x, y, z, m, n, L, G, settings, ugrav, ax, ay, az, +1);
}

/* single-interaction kernels, for golden checks */
void orc_p2p_f64(double* acc, const double* pi, const double* pj, double mj, double hi, double hj, int exact)
{
    if (exact) P2P<true>(acc, pi, pj[0], pj[1], pj[2], mj, hi, hj);
    else P2P<false>(acc, pi, pj[0], pj[1], pj[2], mj, hi, hj);
}
void orc_m2p_f64(double* acc, const double* t, const double* c, const double* q) { M2P(acc, t, c, q); }
void orc_p2m_f64(const double* x, const double* y, const double* z, const double* m, unsigned n, double* com4,
                 double* q)
{
    com4[0] = com4[1] = com4[2] = com4[3] = 0;
    for (unsigned i = 0; i < n; ++i)
        addBody(com4, x[i], y[i], z[i], m[i]);
    normalizeMass(com4);
    P2M(x, y, z, m, 0u, n, com4, q);
}
void orc_m2m_f64(const double* Xout, const double* Xsrc /*[k][4]*/, const double* Msrc /*[k][8]*/, int k, double* Mout)
{
    for (int j = 0; j < 8; ++j)
        Mout[j] = 0;
    for (int i = 0; i < k; ++i)
        addQuadrupole(Mout, Xout[0] - Xsrc[4 * i], Xout[1] - Xsrc[4 * i + 1], Xout[2] - Xsrc[4 * i + 2], Msrc + 8 * i);
}

} // extern "C"
