// Test infrastructure: the halo tables of the LDS-staged hot passes, built by
// the library's host builder (pdx_layout::halo_block) on generated layouts and
// checked under AddressSanitizer / UBSan:
//   1. check_halo_tables: every hl / hs / bl / bk entry within its block's
//      slots, slot counts <= the launch's cap <= PDX_HALO_LIMIT, hs injective
//      per block, bl naming each bond's particle with its broken bit, bk the
//      bond's slot for an intact bond of an active row and the row's own
//      slot otherwise;
//   2. the row loop of the halo kernels (halo_row_loop, pdx_device.h) for
//      every bond group and split, over every row, reading bk arrays sized
//      exactly as the library allocates them (slots + PDX_BK_TAIL): its
//      look-ahead loads stay inside, and every entry it hands a group is a
//      slot of the row's block;
//   3. the lattice kernels' tables (check_lattice below).
// Cases: jittered 3-D lattices, a three-case batch with padding blocks, a
// 2-D plate, a coarse vertebral body (elliptic cylinder) and a wide horizon
// (4.5 spacings) whose colour classes over-fill, each with broken bonds.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "pdx_layout.h"

static int fail(const std::string &m)
{
    fprintf(stderr, "FAIL: %s\n", m.c_str());
    return 1;
}

struct Case {
    const char *name;
    int D;
    double h;
    std::vector<double> X, vol;
    std::vector<int64_t> off;
    bool lattice;  // an exact lattice at a horizon in (3 h, sqrt(10) h]
};

static void lattice(Case &c, int nx, int ny, int nz, double dx, double o, double jit, std::mt19937 &rng)
{
    std::uniform_real_distribution<double> u(-jit, jit);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                c.X.push_back(o + dx * (i + u(rng)));
                c.X.push_back(o + dx * (j + u(rng)));
                if (c.D == 3) c.X.push_back(o + dx * (k + u(rng)));
                c.vol.push_back(c.D == 3 ? dx * dx * dx : dx * dx);
            }
    c.off.push_back((int64_t)c.vol.size());
}

static void cylinder(Case &c, double a, double b, double height, double dx)
{
    for (double x = -a + dx / 2; x < a; x += dx)
        for (double y = -b + dx / 2; y < b; y += dx) {
            if ((x / a) * (x / a) + (y / b) * (y / b) > 1.0) continue;
            for (double z = dx / 2; z < height; z += dx) {
                c.X.push_back(x);
                c.X.push_back(y);
                c.X.push_back(z);
                c.vol.push_back(dx * dx * dx);
            }
        }
    c.off.push_back((int64_t)c.vol.size());
}

// every slot of a row, once over the parts, through the library's row loop
template <int G, int S>
void pdx_layout_check_row(const uint16_t *bp, int width, int part, int nslot, bool &bad, int64_t &visited,
                          std::integral_constant<int, G>, std::integral_constant<int, S>)
{
    halo_row_loop<G, S>(bp, width, part, [&](const uint32_t (&e)[G]) {
        for (int q = 0; q < G; ++q) {
            if ((int)e[q] >= nslot) bad = true;
            ++visited;
        }
    });
}

// 3. the lattice tables (lat_row_one, run for every row over arrays sized as
//    the library allocates them): on an exact lattice no violation, every
//    bks entry a slot of the row's block, the row's own slot where there is
//    no live bond, and for every bond the slot of the particle at h o_k,
//    o_k = lk of the bond; off a lattice (jitter), a violation.
static int check_lattice(const Case &c, const std::vector<int32_t> &perm, int ni, int64_t slots,
                         const std::vector<int64_t> &slptr, std::vector<uint32_t> &bj, std::vector<uint16_t> &bl,
                         const std::vector<int32_t> &ptr, const std::vector<uint16_t> &hs,
                         const std::vector<std::vector<int32_t>> &rows)
{
    // the stencil against an independent enumeration
    int count = 0;
    double mom[3] = {0.0, 0.0, 0.0};
    for (int a = -3; a <= 3; ++a)
        for (int b = -3; b <= 3; ++b)
            for (int e = -3; e <= 3; ++e) {
                const int n = a * a + b * b + e * e;
                const int k = kLat.idx[(a + 3) * 49 + (b + 3) * 7 + (e + 3)];
                if (n < 1 || n > 9) {
                    if (k != -1) return fail("stencil index of an offset outside the stencil");
                    continue;
                }
                if (k < 0 || kLat.o[k][0] != a || kLat.o[k][1] != b || kLat.o[k][2] != e)
                    return fail("stencil index does not map back to its offset");
                ++count;
                const double w = 1.0 - sqrt((double)n) / 3.015;
                mom[0] += w * a;
                mom[1] += w * b;
                mom[2] += w * e;
            }
    if (count != PDX_NST || fabs(mom[0]) > 1e-12 || fabs(mom[1]) > 1e-12 || fabs(mom[2]) > 1e-12)
        return fail("stencil size or first moment");
    double h = 1e300;
    for (int i = 0; i < ni; ++i)
        for (int j : rows[i]) {
            double s2 = 0.0;
            for (int q = 0; q < 3; ++q) {
                const double t = c.X[(size_t)perm[j] * 3 + q] - c.X[(size_t)perm[i] * 3 + q];
                s2 += t * t;
            }
            h = std::min(h, sqrt(s2));
        }
    std::vector<XRec> xr(ni);
    std::vector<KRec> kr(ni);
    std::vector<double> v64(ni, 0.0);
    for (int i = 0; i < ni; ++i) {
        memset(&xr[i], 0, sizeof(XRec));
        memset(&kr[i], 0, sizeof(KRec));
        kr[i].nnb = (uint32_t)rows[i].size();
        if (perm[i] < 0) continue;
        for (int q = 0; q < 3; ++q) xr[i].X[q] = c.X[(size_t)perm[i] * 3 + q];
        kr[i].flags = PDX_FL_ACTIVE;
        v64[i] = c.vol[perm[i]];
    }
    std::vector<uint16_t> bks((size_t)(ni / PDX_SLICE) * PDX_NST * PDX_SLICE, 0xABCD);
    std::vector<uint8_t> lk((size_t)slots, 0xEE);
    Dev d{};
    d.ni = ni;
    d.xu[0] = d.xu[1] = xr.data();
    d.KR = kr.data();
    d.V64 = v64.data();
    d.sl_ptr = const_cast<int64_t *>(slptr.data());
    d.bj = bj.data();
    d.bl = bl.data();
    d.bks = bks.data();
    d.lk = lk.data();
    unsigned viol = 0;
    for (int i = 0; i < ni; ++i) {
        const int b = i / PDX_BLOCK;
        viol |= lat_row_one(d, i, h, 1e-8, c.vol[0], hs[ptr[b] + i % PDX_BLOCK]);
    }
    if (c.lattice != (viol == 0))
        return fail(std::string(c.name) + ": lattice check " + std::to_string(viol) + " on a " +
                    (c.lattice ? "lattice" : "jittered set"));
    if (!c.lattice) {
        printf("%s: not a lattice (violations %u)\n", c.name, viol);
        return 0;
    }
    int64_t live = 0;
    for (int i = 0; i < ni; ++i) {
        const int b = i / PDX_BLOCK, nslot = (ptr[b + 1] - ptr[b] + 15) & ~15;
        const uint16_t own = hs[ptr[b] + i % PDX_BLOCK];
        std::vector<int> at(PDX_NST, -1);
        for (size_t s = 0; s < rows[i].size(); ++s) {
            const int64_t a = slptr[i / PDX_SLICE] + (int64_t)s * PDX_SLICE + i % PDX_SLICE;
            const int k = lk[a];
            if (k >= PDX_NST) return fail(std::string(c.name) + ": lk outside the stencil");
            for (int q = 0; q < 3; ++q) {
                const double t = c.X[(size_t)perm[rows[i][s]] * 3 + q] - c.X[(size_t)perm[i] * 3 + q];
                if (fabs(t - kLat.o[k][q] * h) > 1e-8 * h) return fail(std::string(c.name) + ": lk is not the offset");
            }
            if (at[k] >= 0) return fail(std::string(c.name) + ": two bonds at one offset");
            at[k] = (int)a;
        }
        for (int k = 0; k < PDX_NST; ++k) {
            const uint16_t e = bks[(size_t)lat_at(i, k)];
            if ((int)e >= nslot) return fail(std::string(c.name) + ": bks beyond the block's slots");
            const int64_t a = at[k];
            const bool lv = a >= 0 && !(bj[a] & PDX_BROKEN);
            if (e != (lv ? (uint16_t)(bl[a] & ~PDX_HBROKEN) : own))
                return fail(std::string(c.name) + ": bks neither the bond's nor the row's slot");
            live += lv;
        }
    }
    printf("%s: lattice h %.6g, %lld live stencil entries\n", c.name, h, (long long)live);
    return 0;
}

static int run_case(const Case &c, std::mt19937 &rng)
{
    const int D = c.D;
    const std::vector<int32_t> p0 = pdx_layout::case_starts(c.off);
    std::vector<int32_t> perm, inv;
    std::vector<uint8_t> colour;
    pdx_layout::morton_order(D, c.X.data(), c.vol.data(), c.off, p0, perm, inv, colour);
    const int ni = p0.back(), nbk = ni / PDX_BLOCK, nc = (int)c.off.size() - 1;
    // neighbour rows by brute force within each case (internal indices)
    std::vector<std::vector<int32_t>> rows(ni);
    for (int k = 0; k < nc; ++k)
        for (int i = p0[k]; i < p0[k + 1]; ++i) {
            if (perm[i] < 0) continue;
            for (int j = p0[k]; j < p0[k + 1]; ++j) {
                if (j == i || perm[j] < 0) continue;
                double s2 = 0.0;
                for (int q = 0; q < D; ++q) {
                    const double t = c.X[(size_t)perm[j] * D + q] - c.X[(size_t)perm[i] * D + q];
                    s2 += t * t;
                }
                if (s2 < c.h * c.h) rows[i].push_back(j);
            }
        }
    std::vector<int32_t> cnt(ni);
    for (int i = 0; i < ni; ++i) cnt[i] = (int32_t)rows[i].size();
    std::vector<int64_t> slptr;
    int64_t total = 0;
    const int64_t slots = pdx_layout::sell_pointers(cnt, slptr, total);
    std::vector<uint32_t> bj((size_t)slots, 0xffffffffu);
    std::bernoulli_distribution brk(0.07);
    for (int i = 0; i < ni; ++i)
        for (size_t k = 0; k < rows[i].size(); ++k)
            bj[slptr[i / PDX_SLICE] + (int64_t)k * PDX_SLICE + i % PDX_SLICE] =
                (uint32_t)rows[i][k] | (brk(rng) ? PDX_BROKEN : 0u);
    // the host builder, block by block, then packed as the library packs it
    std::vector<uint16_t> bl((size_t)slots, (uint16_t)PDX_HNONE), bk((size_t)slots + PDX_BK_TAIL, 0);
    std::vector<int32_t> ptr(nbk + 1, 0), hl;
    std::vector<uint16_t> hs;
    int cap = 0;
    int64_t spilled = 0;  // entries of an over-full colour, placed in free slots
    for (int b = 0; b < nbk; ++b) {
        pdx_layout::HaloBlock hbk;
        pdx_layout::halo_block(b, D, bj.data(), slptr, colour, perm, c.X.data(), bl.data(), bk.data(), hbk);
        if (hbk.over) return fail(std::string(c.name) + ": block over the halo limit");
        ptr[b + 1] = ptr[b] + (int32_t)hbk.ent.size();
        hl.insert(hl.end(), hbk.ent.begin(), hbk.ent.end());
        hs.insert(hs.end(), hbk.slot.begin(), hbk.slot.end());
        cap = std::max(cap, hbk.nslot);
        for (size_t q = 0; q < hbk.ent.size(); ++q) spilled += (hbk.slot[q] & 15) != (colour[hbk.ent[q]] & 15);
        // the anchor lies inside the bounding box of the block's active particles
        for (int q = 0; q < D; ++q) {
            double lo = 1e300, hi = -1e300;
            for (int t = 0; t < PDX_BLOCK; ++t)
                if (perm[b * PDX_BLOCK + t] >= 0) {
                    lo = std::min(lo, c.X[(size_t)perm[b * PDX_BLOCK + t] * D + q]);
                    hi = std::max(hi, c.X[(size_t)perm[b * PDX_BLOCK + t] * D + q]);
                }
            if (lo <= hi && !(hbk.anchor[q] >= lo && hbk.anchor[q] <= hi))
                return fail(std::string(c.name) + ": anchor outside the block's box");
        }
    }
    const std::string e =
        pdx_layout::check_halo_tables(nbk, ni, slots, bj.data(), slptr, perm, ptr, hl, hs, bl, bk, cap);
    if (!e.empty()) return fail(std::string(c.name) + ": " + e);
    // the row loop of every halo variant over every row
    int64_t visited = 0;
    auto sweep = [&](auto G_, auto S_) -> int {
        constexpr int G = decltype(G_)::value, S = decltype(S_)::value;
        for (int i = 0; i < ni; ++i) {
            const int b = i / PDX_BLOCK, nslot = (ptr[b + 1] - ptr[b] + 15) & ~15;
            const int64_t s0 = slptr[i / PDX_SLICE];
            const int width = (int)((slptr[i / PDX_SLICE + 1] - s0) / PDX_SLICE);
            const uint16_t *bp = bk.data() + s0 + i % PDX_SLICE;
            for (int part = 0; part < S; ++part) {
                bool bad = false;
                pdx_layout_check_row(bp, width, part, nslot, bad, visited, std::integral_constant<int, G>{},
                                     std::integral_constant<int, S>{});
                if (bad) return 1;
            }
        }
        return 0;
    };
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I4 = std::integral_constant<int, 4>;
    if (sweep(I1{}, I1{}) || sweep(I2{}, I1{}) || sweep(I4{}, I1{}) || sweep(I1{}, I2{}) || sweep(I2{}, I2{}) ||
        sweep(I4{}, I2{}) || sweep(I1{}, I4{}) || sweep(I2{}, I4{}) || sweep(I4{}, I4{}))
        return fail(std::string(c.name) + ": the row loop handed a slot outside the block");
    int64_t want = 0;
    for (int i = 0; i < ni; ++i) want += (slptr[i / PDX_SLICE + 1] - slptr[i / PDX_SLICE]) / PDX_SLICE;
    if (visited != 9 * want) return fail(std::string(c.name) + ": the row loop skipped or repeated slots");
    printf("%s: %d particles, %d blocks, %lld bonds, cap %d slots, %lld entries off their colour\n", c.name, ni,
           nbk, (long long)total, cap, (long long)spilled);
    if (D == 3) return check_lattice(c, perm, ni, slots, slptr, bj, bl, ptr, hs, rows);
    return 0;
}

int main()
{
    std::mt19937 rng(11);
    std::vector<Case> cases;
    {
        Case c{"lattice", 3, 3.015, {}, {}, {0}, false};
        lattice(c, 20, 14, 11, 1.0, 0.0, 0.2, rng);
        cases.push_back(c);
    }
    {
        Case c{"batch", 3, 3.015, {}, {}, {0}, false};
        lattice(c, 9, 8, 7, 1.0, 0.0, 0.0, rng);
        lattice(c, 6, 7, 5, 1.0, 50.0, 0.2, rng);
        cylinder(c, 6.0, 4.5, 7.0, 1.0);
        cases.push_back(c);
    }
    {
        Case c{"plate_2d", 2, 3.015, {}, {}, {0}, false};
        lattice(c, 60, 41, 1, 1.0, 0.0, 0.2, rng);
        cases.push_back(c);
    }
    {
        Case c{"vertebra_coarse", 3, 3.015 * 1.6, {}, {}, {0}, true};
        cylinder(c, 20.0, 15.0, 25.0, 1.6);
        cases.push_back(c);
    }
    {
        Case c{"wide_horizon", 3, 4.5, {}, {}, {0}, false};
        lattice(c, 16, 16, 12, 1.0, 0.0, 0.15, rng);
        cases.push_back(c);
    }
    {
        Case c{"lattice_batch", 3, 3.015 * 0.5, {}, {}, {0}, true};
        lattice(c, 18, 13, 11, 0.5, 0.0, 0.0, rng);
        lattice(c, 7, 9, 6, 0.5, 40.0, 0.0, rng);
        cases.push_back(c);
    }
    for (const Case &c : cases)
        if (run_case(c, rng)) return 1;
    printf("halo ok\n");
    return 0;
}
