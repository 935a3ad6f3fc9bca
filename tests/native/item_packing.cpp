// Host-side check of the work-item packing (corruscant_amd/csrc/kd_item.h):
// every tile count 1..127 and every pair of cuts 0..15 survives encode and
// decode, set_cuts keeps the tile count, and set_tiles clears the cuts.
// Built with hipcc (host code only) and run by tests/test_item_packing.py.
#include <cstdio>
#include <cstring>

#include "../../corruscant_amd/csrc/kd_item.h"

int main()
{
    long long checked = 0;
    for (int n = 1; n <= kd::Item::kMaxTiles; n++) {
        kd::Item it;
        std::memset(&it, 0xA5, sizeof(it));  // stale bits must not leak into the fields
        it.q = 7;
        it.t = 123456;
        it.job = kd::kSubtreeItem | 3;
        it.elo = 2;
        it.ehi = 9;
        it.set_tiles(n);
        if (it.tiles() != n || it.cut_lo() != 0 || it.cut_hi() != 0) {
            std::printf("set_tiles(%d) gives tiles %d cuts %d %d\n", n, it.tiles(), it.cut_lo(), it.cut_hi());
            return 1;
        }
        for (int lo = 0; lo <= kd::Item::kMaxCut; lo++) {
            for (int hi = 0; hi <= kd::Item::kMaxCut; hi++) {
                kd::Item w = it;
                w.set_cuts(lo, hi);
                if (w.tiles() != n || w.cut_lo() != lo || w.cut_hi() != hi || w.q != 7 || w.t != 123456 ||
                    w.job != (kd::kSubtreeItem | 3) || w.elo != 2 || w.ehi != 9) {
                    std::printf("n=%d lo=%d hi=%d decoded as %d %d %d\n", n, lo, hi, w.tiles(), w.cut_lo(),
                                w.cut_hi());
                    return 1;
                }
                w.set_tiles(n);
                if (w.cut_lo() != 0 || w.cut_hi() != 0) {
                    std::printf("set_tiles does not clear the cuts\n");
                    return 1;
                }
                checked++;
            }
        }
    }
    static_assert(sizeof(kd::Item) == 16, "Item is 16 bytes");
    std::printf("ok %lld\n", checked);
    return 0;
}
