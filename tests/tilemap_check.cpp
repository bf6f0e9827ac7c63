// Host check of the XCD tile map (redrock_amd/csrc/xcd_map.hpp): for every
// shape, the launched workgroups cover each (item tile, z tile) exactly once,
// padding workgroups are at most one super-tile round per XCD, and every
// super-tile's tiles run on one XCD.
#include "../redrock_amd/csrc/xcd_map.hpp"
#include <cstdio>
#include <vector>

int main() {
    long long checked = 0;
    const int nti_list[] = {1, 2, 3, 7, 8, 63, 125, 500, 1667};
    const int ntz_list[] = {1, 2, 5, 17, 64, 124, 248, 3795};
    const int res_list[] = {1, 32, 64, 96, 200};
    for (int nti : nti_list)
        for (int ntz : ntz_list)
            for (int res : res_list)
                for (double ratio : {0.25, 1.0, 4.0}) {
                    const rr::TileMap m = rr::make_tile_map(nti, ntz, res, 1000.0, 1000.0 * ratio);
                    if (m.blocks % rr::kXcds) { printf("blocks not a multiple of %d\n", rr::kXcds); return 1; }
                    std::vector<int> seen((size_t)nti * ntz, 0);
                    std::vector<int> xcd_of_super((size_t)(m.blocks / (m.si * m.sz)) + 8, -1);
                    for (long long L = 0; L < m.blocks; ++L) {
                        int ti, tz;
                        if (!rr::map_tile(m, (int)L, ti, tz)) continue;
                        if (ti < 0 || ti >= nti || tz < 0 || tz >= ntz) { printf("tile out of range\n"); return 1; }
                        if (seen[(size_t)ti * ntz + tz]++) { printf("tile %d,%d twice (nti %d ntz %d res %d)\n", ti, tz, nti, ntz, res); return 1; }
                        const int super = (ti / m.si) + m.nsi * (tz / m.sz);
                        const int xcd = (int)(L % rr::kXcds);
                        if (xcd_of_super[super] < 0) xcd_of_super[super] = xcd;
                        else if (xcd_of_super[super] != xcd) { printf("super-tile split over XCDs\n"); return 1; }
                    }
                    for (size_t k = 0; k < seen.size(); ++k)
                        if (seen[k] != 1) { printf("tile %zu missed (nti %d ntz %d res %d)\n", k, nti, ntz, res); return 1; }
                    const long long tiles = (long long)nti * ntz;
                    const long long per = (long long)m.si * m.sz;
                    if (m.blocks > tiles + (long long)rr::kXcds * per + (long long)(m.nsi + ntz) * per) {
                        printf("too much padding: %lld blocks for %lld tiles\n", m.blocks, tiles); return 1;
                    }
                    ++checked;
                }
    printf("tile map ok: %lld shapes\n", checked);
    return 0;
}
