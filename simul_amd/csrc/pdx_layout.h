// pdx_layout.h -- host-side layout of a libpdx context: Morton order of each
// case, the padded internal index space, the per-case neighbour-search cell
// grids and the SELL-64 slice pointers.  Shared by the library (pdx_api.hip)
// and the bounds-checked CPU build of the kernels used by the tests.
#pragma once
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <string>
#include <utility>
#include <vector>

#include "pdx_device.h"

namespace pdx_layout {

inline uint64_t spread3(uint64_t x)
{
    x &= 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffULL;
    x = (x | x << 16) & 0x1f0000ff0000ffULL;
    x = (x | x << 8) & 0x100f00f00f00f00fULL;
    x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
    x = (x | x << 2) & 0x1249249249249249ULL;
    return x;
}

inline uint64_t spread2(uint64_t x)
{
    x &= 0x7fffffff;
    x = (x | x << 16) & 0x0000ffff0000ffffULL;
    x = (x | x << 8) & 0x00ff00ff00ff00ffULL;
    x = (x | x << 4) & 0x0f0f0f0f0f0f0f0fULL;
    x = (x | x << 2) & 0x3333333333333333ULL;
    x = (x | x << 1) & 0x5555555555555555ULL;
    return x;
}

// internal start of every case: each case padded to a multiple of PDX_BLOCK
inline std::vector<int32_t> case_starts(const std::vector<int64_t> &off)
{
    const int nc = (int)off.size() - 1;
    std::vector<int32_t> p0(nc + 1);
    int64_t p = 0;
    for (int k = 0; k < nc; ++k) {
        p0[k] = (int32_t)p;
        const int64_t nk = off[k + 1] - off[k];
        p += ((nk + PDX_BLOCK - 1) / PDX_BLOCK) * PDX_BLOCK;
        if (nk == 0) p += PDX_BLOCK;  // an empty case keeps one inactive block
    }
    p0[nc] = (int32_t)p;
    return p0;
}

// Bank colour (0..15) of the integer cell q: (x mod 2, y mod 2, z mod 4) in
// 3-D, (x mod 4, y mod 4) in 2-D.  Any box of 2x2x4 (2-D: 4x4) cells holds
// 16 different colours, and so does every shift of it; an aligned run of 16
// Morton keys is such a box, and a run of 16 lanes that straddles two of
// them still covers each colour once, since the colour is periodic.  So the
// 16 lanes that read an LDS array together, whose neighbours at one stencil
// offset form the shifted run, mostly hit 16 different bank groups of a
// 16-byte-per-entry array.
inline int bank_colour(int D, const uint64_t *q)
{
    if (D == 3) return (int)((q[0] & 1) | ((q[1] & 1) << 1) | ((q[2] & 3) << 2));
    return (int)((q[0] & 3) | ((q[1] & 3) << 2));
}

// Morton order of the caller rows of each case: perm[internal] = caller row
// (-1 for padding), inv[caller row] = internal; colour[internal] = bank
// colour of the particle's cell (0 for padding).  Positions are quantised to
// cells of the case's median particle spacing vol^(1/D), rounded to the
// nearest cell, so a lattice specimen maps one particle to one cell and an
// aligned run of 16 keys is one 2x2x4 (2-D: 4x4) box of lattice sites.
inline void morton_order(int D, const double *X, const double *vol, const std::vector<int64_t> &off,
                         const std::vector<int32_t> &p0, std::vector<int32_t> &perm, std::vector<int32_t> &inv,
                         std::vector<uint8_t> &colour)
{
    const int nc = (int)off.size() - 1;
    perm.assign(p0[nc], -1);
    inv.assign(off[nc], -1);
    colour.assign(p0[nc], 0);
    for (int k = 0; k < nc; ++k) {
        const int64_t r0 = off[k], r1 = off[k + 1], nk = r1 - r0;
        if (nk == 0) continue;
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int64_t r = r0; r < r1; ++r)
            for (int q = 0; q < D; ++q) {
                lo[q] = std::min(lo[q], X[r * D + q]);
                hi[q] = std::max(hi[q], X[r * D + q]);
            }
        std::vector<double> sp(nk);
        for (int64_t r = r0; r < r1; ++r) sp[r - r0] = pow(vol[r], 1.0 / D);
        std::nth_element(sp.begin(), sp.begin() + nk / 2, sp.end());
        double hq = sp[nk / 2];
        for (int q = 0; q < D; ++q) hq = std::max(hq, (hi[q] - lo[q]) / (D == 3 ? 2097151.0 : 2147483647.0));
        if (!(hq > 0.0) || !isfinite(hq)) hq = 1.0;
        std::vector<std::pair<uint64_t, int32_t>> key(nk);
        std::vector<uint8_t> col(nk);
        for (int64_t r = r0; r < r1; ++r) {
            uint64_t qv[3] = {0, 0, 0};
            for (int q = 0; q < D; ++q) qv[q] = (uint64_t)floor((X[r * D + q] - lo[q]) / hq + 0.5);
            const uint64_t kk = D == 3 ? (spread3(qv[0]) << 2 | spread3(qv[1]) << 1 | spread3(qv[2]))
                                       : (spread2(qv[0]) << 1 | spread2(qv[1]));
            key[r - r0] = {kk, (int32_t)r};
            col[r - r0] = (uint8_t)bank_colour(D, qv);
        }
        std::sort(key.begin(), key.end());
        for (int64_t t = 0; t < nk; ++t) {
            perm[p0[k] + t] = key[t].second;
            inv[key[t].second] = (int32_t)(p0[k] + t);
            colour[p0[k] + t] = col[key[t].second - r0];
        }
    }
}

// per-case cell grids of edge 1.01 * horizon (reference neighbor.py:47) over the
// internal particles; cstart holds ncell+1 absolute starts per case
inline bool cell_grids(int D, const double *X, const std::vector<int32_t> &p0, const std::vector<int32_t> &perm,
                       double horizon, std::vector<GridP> &gp, std::vector<int32_t> &cstart,
                       std::vector<int32_t> &cidx, std::string &err)
{
    const int nc = (int)p0.size() - 1;
    const double h = horizon * 1.01;
    gp.assign(nc, GridP{});
    cstart.clear();
    cidx.clear();
    int64_t cell_off = 0;
    for (int k = 0; k < nc; ++k) {
        GridP g{};
        g.h = h;
        g.nc[0] = g.nc[1] = g.nc[2] = 1;
        const int i0 = p0[k], i1 = p0[k + 1];
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        bool any = false;
        for (int i = i0; i < i1; ++i) {
            const int r = perm[i];
            if (r < 0) continue;
            any = true;
            for (int q = 0; q < D; ++q) {
                lo[q] = std::min(lo[q], X[(size_t)r * D + q]);
                hi[q] = std::max(hi[q], X[(size_t)r * D + q]);
            }
        }
        if (!any)
            for (int q = 0; q < 3; ++q) lo[q] = hi[q] = 0.0;
        int64_t ncell = 1;
        for (int q = 0; q < D; ++q) {
            g.lo[q] = lo[q];
            const double span = (hi[q] - lo[q]) / h;
            if (span > 4.0e6) {
                err = "domain too large for the neighbour grid";
                return false;
            }
            g.nc[q] = (int32_t)floor(span) + 1;
            ncell *= g.nc[q];
        }
        if (ncell > ((int64_t)1 << 28)) {
            err = "neighbour grid too large (sparse domain)";
            return false;
        }
        g.cell_off = cell_off;
        std::vector<int32_t> cnt(ncell + 1, 0);
        std::vector<int64_t> cell_of(i1 - i0, -1);
        for (int i = i0; i < i1; ++i) {
            const int r = perm[i];
            if (r < 0) continue;
            int64_t cc[3] = {0, 0, 0};
            for (int q = 0; q < D; ++q) {
                const int64_t t = (int64_t)floor((X[(size_t)r * D + q] - g.lo[q]) / g.h);
                cc[q] = std::max<int64_t>(0, std::min<int64_t>(t, g.nc[q] - 1));
            }
            const int64_t cell = (cc[0] * g.nc[1] + cc[1]) * g.nc[2] + cc[2];
            cell_of[i - i0] = cell;
            cnt[cell + 1]++;
        }
        const int32_t base = (int32_t)cidx.size();
        for (int64_t t = 0; t < ncell; ++t) cnt[t + 1] += cnt[t];
        std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
        cidx.resize(base + cnt[ncell]);
        for (int i = i0; i < i1; ++i)
            if (cell_of[i - i0] >= 0) cidx[base + fill[cell_of[i - i0]]++] = i;
        for (int64_t t = 0; t <= ncell; ++t) cstart.push_back(base + cnt[t]);
        cell_off += ncell + 1;
        gp[k] = g;
    }
    return true;
}

// SELL-64 slice pointers from the per-particle bond counts; returns slots
inline int64_t sell_pointers(const std::vector<int32_t> &cnt, std::vector<int64_t> &slptr, int64_t &total)
{
    const int nsl = (int)cnt.size() / PDX_SLICE;
    slptr.assign(nsl + 1, 0);
    total = 0;
    for (int s = 0; s < nsl; ++s) {
        int w = 0;
        for (int t = 0; t < PDX_SLICE; ++t) {
            w = std::max(w, cnt[s * PDX_SLICE + t]);
            total += cnt[s * PDX_SLICE + t];
        }
        // widths are multiples of PDX_SLOT_ALIGN, so a group of up to that
        // many slots never straddles a slice's end
        w = (w + PDX_SLOT_ALIGN - 1) / PDX_SLOT_ALIGN * PDX_SLOT_ALIGN;
        slptr[s + 1] = slptr[s] + (int64_t)w * PDX_SLICE;
    }
    return slptr[nsl];
}

// block -> case table of a launch of nblk blocks (padding blocks join the last case)
inline void block_cases(const std::vector<int32_t> &p0, int nblk, std::vector<int32_t> &cob, std::vector<int32_t> &cb0)
{
    const int nc = (int)p0.size() - 1;
    cob.assign(nblk, nc - 1);
    cb0.assign(nc + 1, nblk);
    for (int k = 0; k < nc; ++k) {
        cb0[k] = p0[k] / PDX_BLOCK;
        for (int b = p0[k] / PDX_BLOCK; b < p0[k + 1] / PDX_BLOCK; ++b) cob[b] = k;
    }
}

// bonds in caller order: for every caller row the neighbours' caller indices
// sorted ascending, the bj slot of each entry and its position in internal
// CSR order.  Returns false when a row holds an entry that is not a particle
// (a SELL padding word inside a row).
inline bool bond_rows(int n, int ni, const std::vector<int32_t> &perm, const std::vector<int32_t> &inv,
                      const std::vector<int32_t> &nnb, const std::vector<int64_t> &slptr, const uint32_t *bj,
                      std::vector<int64_t> &row, std::vector<int32_t> &col, std::vector<int64_t> &slot,
                      std::vector<int64_t> &csr)
{
    std::vector<int64_t> rowi(ni + 1, 0);
    for (int i = 0; i < ni; ++i) rowi[i + 1] = rowi[i] + nnb[i];
    row.assign(n + 1, 0);
    for (int r = 0; r < n; ++r) row[r + 1] = row[r] + nnb[inv[r]];
    col.resize(row[n]);
    slot.resize(row[n]);
    csr.resize(row[n]);
    std::vector<std::pair<int32_t, int32_t>> tmp;
    for (int r = 0; r < n; ++r) {
        const int i = inv[r];
        const int64_t base = slptr[i / PDX_SLICE] + (i % PDX_SLICE);
        tmp.clear();
        for (int k = 0; k < nnb[i]; ++k) {
            const uint32_t e = bj[base + (int64_t)k * PDX_SLICE] & ~PDX_BROKEN;
            if ((int64_t)e >= ni || perm[e] < 0) return false;
            tmp.push_back({perm[e], k});
        }
        std::sort(tmp.begin(), tmp.end());
        for (size_t t = 0; t < tmp.size(); ++t) {
            col[row[r] + t] = tmp[t].first;
            slot[row[r] + t] = base + (int64_t)tmp[t].second * PDX_SLICE;
            csr[row[r] + t] = rowi[i] + tmp[t].second;
        }
    }
    return true;
}

}  // namespace pdx_layout
