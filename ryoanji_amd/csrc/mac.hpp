// Vector-MAC radius of a tree node (cstone/traversal/macs.hpp:64-82 computeVecMacR2, with the node's geometric box
// from hilbertIBox / centerAndSize, hilbert.hpp:549-585, box.hpp:357-381), evaluated in FP64; shared by the fused
// upsweep and the reference-signature setMac entry point.
#pragma once

#include "common.hpp"
#include "sfc.hpp"

namespace ryo
{

// squared vector-MAC radius (2 max(size)/theta + |com - geometric center|)^2 of the node with `prefix`, the
// geometric box from the node's integer cell box (hilbertIBox / centerAndSize, mixed-dimension aware)
__device__ inline double macSquared(uint64_t prefix, double cx, double cy, double cz, const MacBox& b)
{
    uint64_t key   = decodePlaceholder(prefix);
    unsigned level = prefixLength(prefix) / 3u;
    IBox3 ib       = cellIBox(key, level, b.bits);
    double gc[3], gs[3];
    for (int d = 0; d < 3; ++d)
    {
        gc[d] = b.lo[d] + double(2 * ib.lo[d] + ib.len[d]) * b.half[d];
        gs[d] = double(ib.len[d]) * b.half[d];
    }
    double dx = cx - gc[0], dy = cy - gc[1], dz = cz - gc[2];
    double s   = sqrt(dx * dx + dy * dy + dz * dz);
    double l   = 2.0 * fmax(fmax(gs[0], gs[1]), gs[2]);
    double mac = l * double(b.invTheta) + s;
    // softening-safe variant: every particle of the cell lies within s + the half diagonal of its box from the
    // center of mass, so a cell accepted beyond that radius plus 2 h_max interacts with no target inside the
    // softening length
    if (b.softReach > 0) mac = fmax(mac, s + sqrt(gs[0] * gs[0] + gs[1] * gs[1] + gs[2] * gs[2]) + b.softReach);
    return mac * mac;
}

} // namespace ryo
