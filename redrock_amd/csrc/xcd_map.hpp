// XCD-aware mapping of a 1-D launch onto a 2-D grid of (item tile, z tile)
// workgroups.
//
// MI355X has 8 XCDs, each with its own L2; the dispatcher sends workgroup L
// to XCD L % 8 and, within an XCD, starts workgroups in increasing L.  The
// scan kernels stream the pixel axis: every workgroup reads the R tiles of its
// items and the template rows of its redshifts, pixel block after pixel block.
// A plain 2-D grid scatters workgroups that share those rows over all XCDs,
// so nearly every read misses L2.  This map hands each XCD whole "super-tiles"
// of si item tiles x sz z tiles, sized to the number of workgroups an XCD
// holds at once: those workgroups start together, advance through the pixels
// at the same pace, and read every R tile and template row once per XCD
// instead of once per workgroup.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace rr {

constexpr int kXcds = 8;

struct TileMap {
    int nti, ntz;      // item tiles, z tiles
    int si, sz;        // super-tile shape (tiles)
    int nsi;           // super-tiles along the item axis
    long long blocks;  // workgroups to launch (multiple of kXcds * si * sz)
};

// resident: workgroups one XCD holds at once; bytes_item / bytes_z: bytes one
// item tile / one z tile reads per pixel block (they set the best aspect).
// fit: the super-tile holds at most `resident` workgroups (all of it runs at
// once); otherwise at least `resident`.
inline TileMap make_tile_map(int nti, int ntz, int resident, double bytes_item, double bytes_z,
                             bool fit = false) {
    TileMap m;
    m.nti = nti;
    m.ntz = ntz;
    // minimise bytes per workgroup: bytes_item / sz + bytes_z / si at si*sz = resident
    int si = (int)std::lround(std::sqrt(resident * bytes_z / bytes_item));
    si = si < 1 ? 1 : (si > nti ? nti : si);
    if (fit)   // a divisor of resident: super-tiles then start and run in lockstep batches
        while (si > 1 && resident % si) --si;
    int sz = fit ? resident / si : (resident + si - 1) / si;
    sz = sz < 1 ? 1 : (sz > ntz ? ntz : sz);
    m.si = si;
    m.sz = sz;
    m.nsi = (nti + si - 1) / si;
    const long long nsuper = (long long)m.nsi * ((ntz + sz - 1) / sz);
    const long long per = (long long)si * sz;
    m.blocks = (nsuper + kXcds - 1) / kXcds * kXcds * per;
    return m;
}

// Workgroup L -> (item tile, z tile); false for the padding workgroups.
__host__ __device__ __forceinline__ bool map_tile(const TileMap &m, int L, int &ti, int &tz) {
    const int xcd = L % kXcds, slot = L / kXcds;
    const int per = m.si * m.sz;
    const int S = (slot / per) * kXcds + xcd;
    const int r = slot - (slot / per) * per;
    const int sup_i = S % m.nsi, sup_z = S / m.nsi;
    ti = sup_i * m.si + r % m.si;
    tz = sup_z * m.sz + r / m.si;
    return ti < m.nti && tz < m.ntz;
}

// Workgroups of `kernel` (block threads, dynamic LDS 0) one XCD holds at once.
template <typename K>
inline int resident_per_xcd(K kernel, int threads) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, threads, 0) != hipSuccess || occ < 1)
        occ = 1;
    const int per = cus / kXcds * occ;
    return per < 1 ? 1 : per;
}

}  // namespace rr
