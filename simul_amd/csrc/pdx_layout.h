// pdx_layout.h -- host-side layout of a libpdx context: Morton order of each
// case, the padded internal index space, the per-case neighbour-search cell
// grids and the SELL-64 slice pointers.  Shared by the library (pdx_api.hip)
// and the bounds-checked CPU build of the kernels used by the tests.
#pragma once
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
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
    // cases are independent: host threads take them in turn
    std::atomic<int> next{0};
    auto one_case = [&](int k) {
        const int64_t r0 = off[k], r1 = off[k + 1], nk = r1 - r0;
        if (nk == 0) return;
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
    };
    auto worker = [&]() {
        for (int k = next++; k < nc; k = next++) one_case(k);
    };
    const int nth = std::max(1, std::min(nc, (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
    std::vector<std::thread> th;
    for (int t = 1; t < nth; ++t) th.emplace_back(worker);
    worker();
    for (auto &x : th) x.join();
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

// ---------------------------------------------------------------------------
// halo tables of the LDS-staged hot passes (Dev::hl_ptr, hl, hs, bl, bk,
// anchor), the host builder.  The device builder (k_halo_build in
// pdx_kernels.hip) produces the same tables; this one is kept for contexts it
// cannot take and as the reference its output is compared with.
//
// Per block of PDX_BLOCK rows: the halo is the block's own particles, then
// every other particle its rows reach, in index order (hl).  Each halo entry
// gets an LDS slot (hs) whose low 4 bits are the particle's bank colour where
// the colour classes allow it (slot 16 r + colour for the r-th entry of that
// colour, r < rows = slots / 16); the entries of an over-full colour take the
// remaining free slots in ascending order.  The block's slot count is the
// halo size rounded up to 16.  bl holds each bond slot's LDS slot (PDX_HBROKEN
// for a broken bond, PDX_HNONE for no bond); bk the slot the kernels read:
// bl's slot for an intact bond of an active row, else the row's own slot.
// The anchor is the centre of the bounding box of the block's own active
// particles (0 on an axis without any).
// ---------------------------------------------------------------------------
struct HaloBlock {
    std::vector<int32_t> ent;      // halo entries: own particles, then the others ascending
    std::vector<uint16_t> slot;    // LDS slot of each entry
    int nslot = 0;                 // slots of the block (halo rounded up to 16)
    bool over = false;             // more than PDX_HALO_LIMIT slots: no tables
    double anchor[3] = {0.0, 0.0, 0.0};
};

// one block's halo; writes the block's bl and bk words (bl must hold
// PDX_HNONE, bk 0 in the block's slot range beforehand).  perm maps internal
// to caller rows (-1: padding, an inactive row); X holds the caller-order
// reference positions (D per row).
inline void halo_block(int b, int D, const uint32_t *bj, const std::vector<int64_t> &slptr,
                       const std::vector<uint8_t> &colour, const std::vector<int32_t> &perm, const double *X,
                       uint16_t *bl, uint16_t *bk, HaloBlock &out)
{
    constexpr int SPB = PDX_BLOCK / PDX_SLICE;
    const int i0 = b * PDX_BLOCK;
    const int64_t s_lo = slptr[(size_t)b * SPB], s_hi = slptr[(size_t)(b + 1) * SPB];
    std::vector<int32_t> o;
    for (int64_t s = s_lo; s < s_hi; ++s) {
        const uint32_t e = bj[s];
        if (e == 0xffffffffu) continue;
        const int32_t j = (int32_t)(e & ~PDX_BROKEN);
        if (j - i0 < 0 || j - i0 >= PDX_BLOCK) o.push_back(j);
    }
    std::sort(o.begin(), o.end());
    o.erase(std::unique(o.begin(), o.end()), o.end());
    const int nh = PDX_BLOCK + (int)o.size();
    const int nslot = (nh + 15) & ~15, rows = nslot / 16;
    out.ent.clear();
    out.slot.clear();
    out.nslot = nslot;
    out.over = nslot > PDX_HALO_LIMIT;
    if (out.over) return;
    std::vector<int32_t> &ent = out.ent;
    ent.resize(nh);
    for (int t = 0; t < PDX_BLOCK; ++t) ent[t] = i0 + t;
    std::copy(o.begin(), o.end(), ent.begin() + PDX_BLOCK);
    // slots by colour, then the overflow into the free slots
    std::vector<int32_t> sl(nslot, -1), slot_of(nh, -1), spill;
    int cnt[16] = {0};
    for (int e = 0; e < nh; ++e) {
        const int cl = colour[ent[e]] & 15;
        if (cnt[cl] < rows) {
            const int sidx = 16 * cnt[cl]++ + cl;
            sl[sidx] = ent[e];
            slot_of[e] = sidx;
        } else {
            spill.push_back(e);
        }
    }
    size_t k = 0;
    for (int sidx = 0; sidx < nslot && k < spill.size(); ++sidx)
        if (sl[sidx] < 0) {
            sl[sidx] = ent[spill[k]];
            slot_of[spill[k]] = sidx;
            ++k;
        }
    out.slot.assign(slot_of.begin(), slot_of.end());
    for (int64_t s = s_lo; s < s_hi; ++s) {
        const uint32_t e = bj[s];
        if (e == 0xffffffffu) continue;
        const int32_t j = (int32_t)(e & ~PDX_BROKEN);
        int32_t l = j - i0;
        if (l < 0 || l >= PDX_BLOCK) l = PDX_BLOCK + (int32_t)(std::lower_bound(o.begin(), o.end(), j) - o.begin());
        bl[s] = (uint16_t)(slot_of[l] | ((e & PDX_BROKEN) ? PDX_HBROKEN : 0u));
    }
    for (int q = 0; q < SPB; ++q) {
        const int64_t a0 = slptr[(size_t)b * SPB + q], a1 = slptr[(size_t)b * SPB + q + 1];
        for (int64_t s = a0; s < a1; ++s) {
            const int t = q * PDX_SLICE + (int)((s - a0) % PDX_SLICE);
            const uint16_t e = bl[s];
            const bool live = e != PDX_HNONE && !(e & PDX_HBROKEN) && perm[i0 + t] >= 0;
            bk[s] = live ? e : (uint16_t)slot_of[t];
        }
    }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int t = 0; t < PDX_BLOCK; ++t) {
        const int32_t r = perm[i0 + t];
        if (r < 0) continue;
        for (int q = 0; q < D; ++q) {
            lo[q] = std::min(lo[q], X[(size_t)r * D + q]);
            hi[q] = std::max(hi[q], X[(size_t)r * D + q]);
        }
    }
    for (int q = 0; q < 3; ++q) out.anchor[q] = 0.0;
    for (int q = 0; q < D; ++q) out.anchor[q] = lo[q] <= hi[q] ? 0.5 * (lo[q] + hi[q]) : 0.0;
}

// The invariants of a context's halo tables that the halo kernels rely on
// for their memory accesses; returns "" or the first violation.  nbk blocks,
// slots SELL slots (bl has slots words, bk slots + PDX_BK_TAIL), cap the LDS
// slots a launch provides.
inline std::string check_halo_tables(int nbk, int ni, int64_t slots, const uint32_t *bj,
                                     const std::vector<int64_t> &slptr, const std::vector<int32_t> &perm,
                                     const std::vector<int32_t> &ptr, const std::vector<int32_t> &hl,
                                     const std::vector<uint16_t> &hs, const std::vector<uint16_t> &bl,
                                     const std::vector<uint16_t> &bk, int cap)
{
    constexpr int SPB = PDX_BLOCK / PDX_SLICE;
    auto err = [](int b, const std::string &m) { return "block " + std::to_string(b) + ": " + m; };
    if ((int)ptr.size() != nbk + 1 || ptr[0] != 0) return "hl_ptr size";
    if ((int64_t)bl.size() != slots || (int64_t)bk.size() != slots + PDX_BK_TAIL) return "bl / bk size";
    if (cap % 16 != 0 || cap > PDX_HALO_LIMIT) return "slot cap not a multiple of 16 within the limit";
    std::vector<int32_t> ent_of_slot;
    for (int b = 0; b < nbk; ++b) {
        const int i0 = b * PDX_BLOCK, nh = ptr[b + 1] - ptr[b];
        const int nslot = (nh + 15) & ~15;
        if (nh < PDX_BLOCK) return err(b, "halo smaller than the block");
        if (nslot > cap) return err(b, "more slots than the launch provides");
        ent_of_slot.assign(nslot, -1);
        for (int e = 0; e < nh; ++e) {
            const int32_t j = hl[ptr[b] + e];
            const int sidx = hs[ptr[b] + e];
            if (e < PDX_BLOCK ? j != i0 + e : (j < 0 || j >= ni || (j >= i0 && j < i0 + PDX_BLOCK)))
                return err(b, "halo entry out of place");
            if (e > PDX_BLOCK && j <= hl[ptr[b] + e - 1]) return err(b, "halo not ascending");
            if (sidx >= nslot) return err(b, "LDS slot beyond the block's slots");
            if (ent_of_slot[sidx] >= 0) return err(b, "two entries in one LDS slot");
            ent_of_slot[sidx] = j;
        }
        for (int q = 0; q < SPB; ++q) {
            const int64_t a0 = slptr[(size_t)b * SPB + q], a1 = slptr[(size_t)b * SPB + q + 1];
            for (int64_t s = a0; s < a1; ++s) {
                const int t = q * PDX_SLICE + (int)((s - a0) % PDX_SLICE);
                const uint32_t e = bj[s];
                const uint16_t l = bl[s], k = bk[s];
                const int own = hs[ptr[b] + t];
                if (e == 0xffffffffu) {
                    if (l != PDX_HNONE) return err(b, "bl of an empty slot");
                } else {
                    const int sidx = l & ~PDX_HBROKEN;
                    if (l == PDX_HNONE || sidx >= nslot || ent_of_slot[sidx] != (int32_t)(e & ~PDX_BROKEN))
                        return err(b, "bl does not name the bond's particle");
                    if (((l & PDX_HBROKEN) != 0) != ((e & PDX_BROKEN) != 0)) return err(b, "bl broken bit");
                }
                if (k >= nslot) return err(b, "bk beyond the block's slots");
                const bool live = e != 0xffffffffu && !(e & PDX_BROKEN) && perm[i0 + t] >= 0;
                if (k != (live ? (int)l : own)) return err(b, "bk neither the bond's nor the row's slot");
            }
        }
    }
    for (int64_t s = slots; s < slots + PDX_BK_TAIL; ++s)
        if (bk[s] >= cap) return "bk tail beyond the slots";
    return "";
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
