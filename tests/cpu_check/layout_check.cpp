// Test infrastructure: properties of the host layout (pdx_layout.h) that the
// LDS-staged kernels rely on, checked on the CPU.
//   1. bank_colour: every box of 2x2x4 cells (3-D) or 4x4 cells (2-D), at every
//      offset, holds 16 different colours;
//   2. morton_order on a (jittered) lattice quantises to the lattice sites: in
//      a lattice whose extents are multiples of the box, every aligned run of
//      16 internal indices is one box, so its 16 colours differ;
//   3. sell_pointers: every slice's width is its longest row rounded up to a
//      multiple of PDX_SLOT_ALIGN (the halo kernels load groups of up to that
//      many slots without a bounds test), and the bond total is exact.
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <random>
#include <vector>

#include "pdx_layout.h"

static int fail(const char *m)
{
    fprintf(stderr, "FAIL: %s\n", m);
    return 1;
}

int main()
{
    // 1. colours of shifted boxes
    for (uint64_t ox = 0; ox < 8; ++ox)
        for (uint64_t oy = 0; oy < 8; ++oy)
            for (uint64_t oz = 0; oz < 8; ++oz) {
                int seen3 = 0, seen2 = 0;
                for (uint64_t x = 0; x < 2; ++x)
                    for (uint64_t y = 0; y < 2; ++y)
                        for (uint64_t z = 0; z < 4; ++z) {
                            const uint64_t q[3] = {ox + x, oy + y, oz + z};
                            seen3 |= 1 << pdx_layout::bank_colour(3, q);
                        }
                for (uint64_t x = 0; x < 4; ++x)
                    for (uint64_t y = 0; y < 4; ++y) {
                        const uint64_t q[2] = {ox + x, oy + y};
                        seen2 |= 1 << pdx_layout::bank_colour(2, q);
                    }
                if (seen3 != 0xffff) return fail("3-D box without 16 colours");
                if (seen2 != 0xffff) return fail("2-D box without 16 colours");
            }
    // 2. jittered lattices, two cases (the second offset and with another spacing)
    std::mt19937 rng(7);
    std::uniform_real_distribution<double> jit(-0.2, 0.2);
    for (int D = 2; D <= 3; ++D) {
        const int nx = 8, ny = 12, nz = D == 3 ? 16 : 1;
        const double dxs[2] = {0.37, 1.25};
        const double org[2] = {1.23, -40.0};
        std::vector<double> X, vol;
        std::vector<int64_t> off = {0};
        for (int k = 0; k < 2; ++k) {
            for (int i = 0; i < nx; ++i)
                for (int j = 0; j < ny; ++j)
                    for (int l = 0; l < nz; ++l) {
                        X.push_back(org[k] + dxs[k] * (i + jit(rng)));
                        X.push_back(org[k] + dxs[k] * (j + jit(rng)));
                        if (D == 3) X.push_back(org[k] + dxs[k] * (l + jit(rng)));
                        vol.push_back(D == 3 ? dxs[k] * dxs[k] * dxs[k] : dxs[k] * dxs[k]);
                    }
            off.push_back((int64_t)vol.size());
        }
        const std::vector<int32_t> p0 = pdx_layout::case_starts(off);
        std::vector<int32_t> perm, inv;
        std::vector<uint8_t> colour;
        pdx_layout::morton_order(D, X.data(), vol.data(), off, p0, perm, inv, colour);
        for (int k = 0; k < 2; ++k) {
            const int nk = (int)(off[k + 1] - off[k]);
            for (int r = 0; r + 16 <= nk; r += 16) {
                int seen = 0;
                for (int t = 0; t < 16; ++t) seen |= 1 << colour[p0[k] + r + t];
                if (seen != 0xffff) return fail("aligned run of 16 Morton keys without 16 colours");
            }
        }
        for (size_t r = 0; r < inv.size(); ++r)
            if (perm[inv[r]] != (int32_t)r) return fail("perm / inv not inverse");
    }
    // 3. slice widths
    {
        std::uniform_int_distribution<int> nn(0, 140);
        std::vector<int32_t> cnt(64 * 37);
        int64_t want = 0;
        for (auto &c : cnt) {
            c = nn(rng);
            want += c;
        }
        for (int t = 0; t < 64; ++t) cnt[5 * 64 + t] = 0;     // an empty slice
        for (int t = 0; t < 64; ++t) cnt[6 * 64 + t] = t % 9;  // widest row 8
        want = 0;
        for (int c : cnt) want += c;
        std::vector<int64_t> slptr;
        int64_t total = -1;
        const int64_t slots = pdx_layout::sell_pointers(cnt, slptr, total);
        if (total != want) return fail("sell_pointers bond total");
        if (slots != slptr.back() || slptr[0] != 0) return fail("sell_pointers ends");
        for (size_t s = 0; s + 1 < slptr.size(); ++s) {
            int w = 0;
            for (int t = 0; t < 64; ++t) w = std::max(w, cnt[s * 64 + t]);
            const int64_t width = (slptr[s + 1] - slptr[s]) / 64;
            if ((slptr[s + 1] - slptr[s]) % 64 || width % PDX_SLOT_ALIGN || width < w ||
                width >= w + PDX_SLOT_ALIGN)
                return fail("slice width not the longest row rounded up to PDX_SLOT_ALIGN");
        }
        if (slptr[6] != slptr[5]) return fail("empty slice has slots");
        if (slptr[7] - slptr[6] != 8 * 64) return fail("slice of widest row 8 is not 8 wide");
    }
    printf("layout ok\n");
    return 0;
}
