// pdx_device.h -- device records and the per-particle bodies of every libpdx
// kernel, written as host+device functions.  The HIP kernels in
// pdx_kernels.hip are thin wrappers over these functions; the same functions
// compile with a host C++ compiler for the bounds-checked CPU build of the
// tests (no HIP types are used here).
//
// Particle data lives in Morton order, one case after another, each case
// padded to a multiple of PDX_BLOCK particles so that every block of every
// launch belongs to exactly one case.  Padding particles have no bonds and are
// flagged inactive.
//
// Bonds are stored SELL-64 (sliced ELL, slice = 64 consecutive particles =
// one wavefront): entry k of particle i sits at sl_ptr[i/64] + 64*k + i%64,
// so the k-th bond of the 64 lanes of a wave is one coalesced 256-byte read.
// An entry is the neighbour's internal index with bit 31 = broken.
//
// Reference arithmetic: backend/fea/peridynamics/core/nosb.py:59-94 (K),
// :96-126 (F), :211-227 (P), :229-275 (force); neighbor.py:154-206 (sets);
// bonds.py:116-137,161-180; damage.py:94-142; nosb_solver.py:94-198 (DR).
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define PDX_HD __host__ __device__ __forceinline__
#else
#define PDX_HD inline
#endif

#ifndef PDX_BLOCK
#define PDX_BLOCK 256
#endif
// waves per block that write reduction partials (PDX_BLOCK / 64)
#define PDX_PW 4
#define PDX_SLICE 64
// slice widths (bond slots per row of a slice) are multiples of this
#define PDX_SLOT_ALIGN 4
// bond slots past the end of bk that the halo kernels' look-ahead loads may
// touch (never used): three groups of the largest row split (4 parts x 4)
#define PDX_BK_TAIL (3 * 16 * PDX_SLICE + 4 * PDX_SLICE)
#define PDX_BROKEN 0x80000000u
// halo-local bond entries (16 bits): position in the block's halo, broken flag
#define PDX_HBROKEN 0x8000u
#define PDX_HNONE 0xFFFFu
#define PDX_HALO_LIMIT 0x7FFF
#define PDX_NXCD 8

// flags word of KRec
#define PDX_FL_FIXMASK 0x7u
#define PDX_FL_SING 0x8u
#define PDX_FL_LOADED 0x10u
#define PDX_FL_ACTIVE 0x20u

struct alignas(32) pdx_d4 {
    double x, y, z, w;
};

PDX_HD pdx_d4 d4(double x, double y, double z) { return pdx_d4{x, y, z, 0.0}; }

// gather record of a particle, 64 bytes: reference position, volume,
// stabilisation micromodulus and the displacement iterate.  The record is
// double-buffered on u (ping-pong between passes); X, V and c are identical in
// both buffers.  One 64-byte record per neighbour lets four lanes fetch it
// with one 16-byte load each (wave-cooperative gathers).
struct alignas(64) XRec {
    double X[3];
    float V;  // particle volume
    float c;  // stabilisation micromodulus c_i
    double u[3];
    double w;  // unused, zero
};

// per-particle constants of the hot loop: 64 bytes
struct alignas(16) KRec {
    float Kinv[9];  // K^-1 in FP32 (K itself when singular)
    float m[3];     // sum_j w_ij V_j xi_ij over intact bonds
    float bulk, shear;
    uint32_t flags;
    uint32_t nnb;   // bonds in the row (intact + broken)
};

// correspondence tensor A = P K^-1 of the current iterate: 48 bytes
struct alignas(16) ARec {
    float A[9];
    float pad[3];
};

// record of a particle for the lattice force pass's halo, 64 bytes, written
// by the lattice F-pass: u / h (FP64), A = P K^-1 (FP32), c' = c G_s / 2
struct alignas(64) FRec {
    double u[3];
    float A8;
    float c;
    float A[8];
};

struct alignas(16) J2Par {
    float sy0, hiso, hkin, pad;
};

// committed J2 state (symmetric tensors as xx,yy,zz,xy,yz,xz): 64 bytes
struct alignas(16) J2State {
    float ep[6];
    float epbar;
    float beta[6];
    float pad[3];
};

// per-case device scalars of the iteration
struct DScal {
    double ref, tol, res, rel;
    double c, cmax;
    double ke, prev_ke;
    double num, den;  // adaptive: Rayleigh sums; kinetic: num = kinetic energy of the last pass
    int64_t iter, conv_iter, resets, pad64;
    int32_t converged, sol_buf, first, inc;
    int32_t reset_pending, pad32;
    uint32_t counter, pad_u;
};

// per-case neighbour-search grid
struct GridP {
    double lo[3];
    double h;
    int32_t nc[3];
    int32_t pad;
    int64_t cell_off;  // offset of this case's cells in cell_start
};

// device pointers handed to every kernel (by value)
struct Dev {
    int32_t ni;        // internal particle count (padded)
    int32_t nblk;      // launch blocks (multiple of PDX_NXCD)
    int32_t ncase;
    int32_t dim;
    XRec *xu[2];       // ping-pong gather records (u differs, X/V/c shared)
    KRec *KR;
    ARec *AR;
    J2Par *j2p;
    J2State *j2s;
    pdx_d4 *v;
    pdx_d4 *fp;        // previous force (adaptive DR) / acceleration (kinetic DR)
    pdx_d4 *f;         // force output of force-only passes
    pdx_d4 *fext;
    double *mass;      // physical mass rho V (kinetic DR)
    double *lam;       // adaptive-DR diagonal mass
    double *V64;       // particle volume in FP64 (setup kernels)
    double *K64, *Kinv64;
    float *Fd, *Pd;    // diagnostic F - I and P (9 per particle)
    double *damage;
    int32_t *init_nb;  // bonds at build time
    int64_t *sl_ptr;
    uint32_t *bj;
    // halo tables of the LDS-staged hot passes (null when not built): per
    // block the particles its rows reach (hl_ptr / hl: own particles first,
    // then the others in index order) and each one's LDS slot (hs), the LDS
    // slot of each bond in the SELL layout of bj (bl: PDX_HBROKEN set for a
    // broken bond, 0xFFFF for no bond), the slot the kernels read for each
    // bond (bk: bl's slot for an intact bond of an active row, else the row's
    // own slot, whose bond terms are exact zeros) and the FP64 anchor the
    // staged positions are relative to
    int32_t *hl_ptr;
    int32_t *hl;
    uint16_t *hs;
    uint16_t *bl;
    uint16_t *bk;
    pdx_d4 *anchor;
    // per row, sum_j w V_j xi of the halo F-pass (wx as the pass rounds it,
    // FP64): a function of the reference positions, volumes and intact bonds,
    // kept so that the pass need not sum it (valid while the context's
    // have_m is set)
    pdx_d4 *m64;
    // lattice tables (null unless every bond of the context is an offset of
    // the 122-site stencil on one lattice spacing, with one particle volume):
    // bks[(slice * PDX_NST + k) * PDX_SLICE + lane] is the LDS slot of the
    // row's neighbour at stencil offset k (the row's own slot when there is
    // none, the bond is broken or the row inactive), lk[s] the stencil offset
    // of bond slot s of bj
    uint16_t *bks;
    uint8_t *lk;
    FRec *fr;  // the lattice F-pass's records for the lattice force pass (with bks)
    int32_t halo_cap;  // LDS slots per block: the largest block's slot count (a multiple of 16)
    int32_t halo_pad;
    int32_t *case_of_blk;
    int32_t *case_blk0;  // first block of each case (ncase+1)
    DScal *sc;
    double *part;      // per-wave partial sums (PDX_PW waves x 4 per block)
    // neighbour search
    GridP *grid;
    int32_t *case_of_p;  // case of each internal particle
    int32_t *perm;       // caller index of each internal particle (-1: padding)
    int32_t *cell_start; // per cell start into cell_idx (+1 sentinel per case grid)
    int32_t *cell_idx;
    int32_t *counts;
    uint64_t *nkey;       // neighbour fill only: each row's bond sort keys, in its bond slots
    double h, inv_h, h2;  // horizon, 1/horizon, horizon^2
    double G_s;
    int32_t model, form;
};

PDX_HD int64_t slot0(const Dev &d, int i) { return d.sl_ptr[i >> 6] + (i & 63); }

// ---------------------------------------------------------------------------
// lattice stencil.  A voxel specimen puts every particle on a cubic lattice
// of spacing h; at horizons in (3 h, sqrt(10) h] (the reference's 3.015 h,
// pd_adapter.py:86) a particle's bonds are the lattice offsets o with
// 0 < |o|^2 <= 9 that land on another particle: 122 of them.  With the
// offset known the bond vector is xi = h o exactly, |xi| = h sqrt(|o|^2) and
// the influence weight w = 1 - |xi| / delta depends on |o|^2 alone, so the
// lattice kernels take xi, |xi| and w as constants of the stencil entry
// instead of arithmetic on positions.  Entries are ordered by which
// components are non-zero (x, y, z, xy, xz, yz, xyz), each class in
// lexicographic order of the offsets whose first non-zero component is
// positive, each such o followed by -o: entries 2 m and 2 m + 1 are a pair
// o, -o.  idx maps o + 3 (7 x 7 x 7) to the entry or -1.  The stencil is
// symmetric, so sum_k w(o_k) o_k = 0 over the whole stencil.
// ---------------------------------------------------------------------------
#define PDX_NST 122

struct LatStencil {
    int8_t o[PDX_NST][3];
    int8_t idx[343];
};

constexpr LatStencil lat_make()
{
    LatStencil s{};
    for (int t = 0; t < 343; ++t) s.idx[t] = -1;
    const int order[7] = {1, 2, 4, 3, 5, 6, 7};
    int k = 0;
    for (int pi = 0; pi < 7; ++pi)
        for (int a = -3; a <= 3; ++a)
            for (int b = -3; b <= 3; ++b)
                for (int c = -3; c <= 3; ++c) {
                    const int n = a * a + b * b + c * c;
                    const int p = (a != 0 ? 1 : 0) | (b != 0 ? 2 : 0) | (c != 0 ? 4 : 0);
                    const int lead = a != 0 ? a : (b != 0 ? b : c);
                    if (n < 1 || n > 9 || p != order[pi] || lead < 0) continue;
                    for (int sg = 1; sg >= -1; sg -= 2) {
                        s.o[k][0] = (int8_t)(sg * a);
                        s.o[k][1] = (int8_t)(sg * b);
                        s.o[k][2] = (int8_t)(sg * c);
                        s.idx[(sg * a + 3) * 49 + (sg * b + 3) * 7 + (sg * c + 3)] = (int8_t)k;
                        ++k;
                    }
                }
    return s;
}

static constexpr LatStencil kLat = lat_make();

// per-context constants of the lattice kernels, indexed by n = |o|^2 (1..9),
// w0(n) = 1 - sqrt(n) h / delta
struct LatC {
    double w1[10], w2[10], w3[10];  // w0, 2 w0, 3 w0 (FP64): the F-pass's weights of o_c = 1, 2, 3
    float c1[10], c2[10], c3[10];   // h w0, 2 h w0, 3 h w0: the force pass's correspondence weights
    float wl[10];                   // w0 / sqrt(n): the stabilisation term's weight
    double inv_h;                   // 1 / h
    double fscale;                  // V h: the F-pass sum -> sum_j w V_j du (x) xi
    double V;                       // the particles' common volume
};

PDX_HD int64_t lat_at(int i, int k) { return ((int64_t)(i >> 6) * PDX_NST + k) * PDX_SLICE + (i & 63); }

// lattice tables of row i: lk of each bond slot and the row's PDX_NST bks
// entries.  Returns 0, or bit 1 when a bond is not h o for a stencil offset o
// within tol h per component, bit 2 when two bonds share an offset, bit 4
// when an active row's volume differs from V.  own is the row's LDS slot.
PDX_HD unsigned lat_row_one(const Dev &d, int i, double h, double tol, double V, uint16_t own)
{
    const KRec kr = d.KR[i];
    const bool act = (kr.flags & PDX_FL_ACTIVE) != 0;
    for (int k = 0; k < PDX_NST; ++k) d.bks[lat_at(i, k)] = own;
    unsigned viol = 0;
    if (act && d.V64[i] != V) viol |= 4u;
    const XRec pi = d.xu[0][i];
    const int64_t base = slot0(d, i);
    uint64_t seen0 = 0, seen1 = 0;
    const double ih = 1.0 / h;
    for (int s = 0; s < (int)kr.nnb; ++s) {
        const int64_t a = base + (int64_t)s * PDX_SLICE;
        const uint32_t e = d.bj[a];
        const int j = (int)(e & ~PDX_BROKEN);
        const XRec pj = d.xu[0][j];
        int o[3];
        bool on = true;
        for (int q = 0; q < 3; ++q) {
            const double t = pj.X[q] - pi.X[q];
            const double r = t * ih;
            o[q] = (int)(r < 0.0 ? -floor(-r + 0.5) : floor(r + 0.5));
            if (!(fabs(t - o[q] * h) <= tol * h) || o[q] < -3 || o[q] > 3) on = false;
        }
        const int k = on ? (int)kLat.idx[(o[0] + 3) * 49 + (o[1] + 3) * 7 + (o[2] + 3)] : -1;
        if (k < 0) {
            viol |= 1u;
            d.lk[a] = 0xFF;
            continue;
        }
        const uint64_t bit = 1ull << (k & 63);
        uint64_t &seen = k < 64 ? seen0 : seen1;
        if (seen & bit) viol |= 2u;
        seen |= bit;
        d.lk[a] = (uint8_t)k;
        const uint16_t l = d.bl[a];
        const bool live = act && !(e & PDX_BROKEN) && !(l & PDX_HBROKEN);
        d.bks[lat_at(i, k)] = live ? (uint16_t)(l & ~PDX_HBROKEN) : own;
    }
    return viol;
}

// ---------------------------------------------------------------------------
// small tensor algebra
// ---------------------------------------------------------------------------
template <int D>
PDX_HD void matmul_f(const float *A, const float *B, float *C)
{
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) {
            float s = 0.f;
            for (int k = 0; k < D; ++k) s = fmaf(A[r * D + k], B[k * D + c], s);
            C[r * D + c] = s;
        }
}

template <int D, typename T>
PDX_HD T lame(T Kb, T mu)
{
    return D == 3 ? Kb - mu * (T(2) / T(3)) : Kb - mu;
}

// linear elastic: eps = sym(H), P = lam tr(eps) I + 2 mu eps   (nosb.py:218-227)
template <int D, typename T>
PDX_HD void stress_le(const T *H, T Kb, T mu, T *P)
{
    const T lam = lame<D>(Kb, mu);
    T tr = 0;
    for (int q = 0; q < D; ++q) tr += H[q * D + q];
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) {
            const T e = T(0.5) * (H[r * D + c] + H[c * D + r]);
            P[r * D + c] = T(2) * mu * e + (r == c ? lam * tr : T(0));
        }
}

// Neo-Hookean P = mu (F - F^-T) + lam ln(J_s) F^-T with F = I + H
// (neo_hookean.py:117-127 sigma, :176-180 P = J sigma F^-T), evaluated without
// cancellation: J - 1 from the invariants of H, F^-T = I + M^T with
// M = F^-1 - I = -F^-1 H.
template <int D, typename T>
PDX_HD void stress_nh(const T *H, T Kb, T mu, T *P)
{
    const T lam = lame<D>(Kb, mu);
    T Jm1, adj[D * D];
    if constexpr (D == 3) {
        const T tr = H[0] + H[4] + H[8];
        T trH2 = 0;
        for (int r = 0; r < D; ++r)
            for (int c = 0; c < D; ++c) trH2 += H[r * D + c] * H[c * D + r];
        const T *h = H;
        const T detH = h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
                       h[2] * (h[3] * h[7] - h[4] * h[6]);
        Jm1 = tr + T(0.5) * (tr * tr - trH2) + detH;
        const T F0 = T(1) + h[0], F4 = T(1) + h[4], F8 = T(1) + h[8];
        const T F1 = h[1], F2 = h[2], F3 = h[3], F5 = h[5], F6 = h[6], F7 = h[7];
        T *a = adj;
        a[0] = F4 * F8 - F5 * F7; a[1] = F2 * F7 - F1 * F8; a[2] = F1 * F5 - F2 * F4;
        a[3] = F5 * F6 - F3 * F8; a[4] = F0 * F8 - F2 * F6; a[5] = F2 * F3 - F0 * F5;
        a[6] = F3 * F7 - F4 * F6; a[7] = F1 * F6 - F0 * F7; a[8] = F0 * F4 - F1 * F3;
    } else {
        Jm1 = H[0] + H[3] + (H[0] * H[3] - H[1] * H[2]);
        adj[0] = T(1) + H[3];
        adj[1] = -H[1];
        adj[2] = -H[2];
        adj[3] = T(1) + H[0];
    }
    const T J = T(1) + Jm1;
    const T lnJ = J > T(1e-8) ? log1p(Jm1) : log(T(1e-8));
    const T iJ = T(1) / J;
    T M[D * D];  // M = -F^-1 H
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) {
            T acc = 0;
            for (int k = 0; k < D; ++k) acc += adj[r * D + k] * H[k * D + c];
            M[r * D + c] = -acc * iJ;
        }
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) {
            const T mt = M[c * D + r];
            P[r * D + c] = mu * (H[r * D + c] - mt) + lam * lnJ * ((r == c ? T(1) : T(0)) + mt);
        }
}

// J2 small-strain radial return with linear isotropic + kinematic hardening,
// FP64 (j2_plasticity.py:202-249; h_kin = 0 is the reference).  Symmetric
// tensors packed xx,yy,zz,xy,yz,xz; off-diagonal entries count twice in norms.
PDX_HD void stress_j2(const double *H, double Kb_, double mu_, const J2Par &pp, const J2State &st, double *P,
                      J2State *out)
{
    const double Kb = Kb_, mu = mu_;
    const double lam = Kb - 2.0 * mu / 3.0;
    double eps[6];
    eps[0] = H[0]; eps[1] = H[4]; eps[2] = H[8];
    eps[3] = 0.5 * (H[1] + H[3]);
    eps[4] = 0.5 * (H[5] + H[7]);
    eps[5] = 0.5 * (H[2] + H[6]);
    double ee[6];
    for (int q = 0; q < 6; ++q) ee[q] = eps[q] - (double)st.ep[q];
    const double tre = ee[0] + ee[1] + ee[2];
    double sig[6];
    for (int q = 0; q < 6; ++q) sig[q] = 2.0 * mu * ee[q] + (q < 3 ? lam * tre : 0.0);
    const double pr = (sig[0] + sig[1] + sig[2]) / 3.0;
    double eta[6], nsq = 0.0;
    for (int q = 0; q < 6; ++q) {
        eta[q] = sig[q] - (q < 3 ? pr : 0.0) - (double)st.beta[q];
        nsq += (q < 3 ? 1.0 : 2.0) * eta[q] * eta[q];
    }
    const double nrm = sqrt(nsq > 1e-30 ? nsq : 1e-30);
    const double qv = sqrt(1.5) * nrm;
    const double epbar = st.epbar;
    const double fy = qv - ((double)pp.sy0 + (double)pp.hiso * epbar);
    double sfin[6];
    if (fy <= 0.0) {
        for (int q = 0; q < 6; ++q) sfin[q] = sig[q];
        if (out) *out = st;
    } else {
        const double dg = fy / (3.0 * mu + (double)pp.hiso + (double)pp.hkin);
        const double fac = 1.0 - 3.0 * mu * dg / qv;
        J2State ns = st;
        for (int q = 0; q < 6; ++q) {
            const double nf = eta[q] / nrm;
            ns.ep[q] = (float)((double)st.ep[q] + sqrt(1.5) * dg * nf);
            const double bn = (double)st.beta[q] + sqrt(2.0 / 3.0) * (double)pp.hkin * dg * nf;
            ns.beta[q] = (float)bn;
            // the deviator stays 2 mu dev(eps - eps_p): beta_n + fac (s_trial - beta_n)
            sfin[q] = (double)st.beta[q] + fac * eta[q] + (q < 3 ? pr : 0.0);
        }
        ns.epbar = (float)(epbar + dg);
        if (out) *out = ns;
    }
    P[0] = sfin[0]; P[4] = sfin[1]; P[8] = sfin[2];
    P[1] = P[3] = sfin[3];
    P[5] = P[7] = sfin[4];
    P[2] = P[6] = sfin[5];
}

// Walk the bond slots of a row: slot k's entry is bp[k * PDX_SLICE].  Part p
// of S takes the groups of G slots starting at p G, p G + G S, ...  Two
// groups are processed per trip and the entries of the next two are loaded
// at its start, so an entry load (from HBM; bk is read once per pass) has two
// groups' arithmetic to arrive in.  The slice width is a multiple of
// PDX_SLOT_ALIGN, so a group never straddles it, and the look-ahead loads
// past it land in the next slice or in bk's tail and are never used.  width
// is the same for the 64 rows of a wave, so the test of the second group is
// uniform.
template <int G, int S, typename F>
PDX_HD void halo_row_loop(const uint16_t *__restrict__ bp, int width, int part, F &&group)
{
    static_assert(PDX_SLOT_ALIGN % G == 0 && G * S * 3 + PDX_SLOT_ALIGN <= PDX_BK_TAIL / PDX_SLICE, "group layout");
    constexpr int STEP = G * S;
    const int k0 = part * G;
    auto entry = [&](int k) -> uint32_t { return (uint32_t)bp[(int64_t)k * PDX_SLICE]; };
    uint32_t ea[G], eb[G];
#pragma unroll
    for (int q = 0; q < G; ++q) {
        ea[q] = entry(k0 + q);
        eb[q] = entry(k0 + STEP + q);
    }
    for (int k = k0; k < width; k += 2 * STEP) {
        uint32_t na[G], nb[G];
#pragma unroll
        for (int q = 0; q < G; ++q) {
            na[q] = entry(k + 2 * STEP + q);
            nb[q] = entry(k + 3 * STEP + q);
        }
        group(ea);
        if (k + STEP < width) group(eb);
#pragma unroll
        for (int q = 0; q < G; ++q) {
            ea[q] = na[q];
            eb[q] = nb[q];
        }
    }
}

// square root and reciprocal of the FP32 bond arithmetic: the hardware
// V_SQRT_F32 / V_RCP_F32 (1 ulp) on the GPU, where the IEEE-rounded forms cost
// about a dozen instructions each; the host build of this header (the CPU
// checks) uses the library functions
PDX_HD float bond_sqrt(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}

PDX_HD float bond_rcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// influence weight w = max(0, 1 - |xi|/delta) (bonds.py:131-137) times V_j
PDX_HD float bond_weight(float L, float invh, float Vj)
{
    return (L > 0.f ? fmaxf(0.f, 1.f - L * invh) : 1.f) * Vj;
}

// the same weight without the |xi| = 0 case: the halo kernels' dead bond slots
// (xi = 0, du = 0) contribute zero whatever their weight
PDX_HD float bond_weight_live(float L, float invh, float Vj)
{
    return fmaxf(0.f, 1.f - L * invh) * Vj;
}

// ---------------------------------------------------------------------------
// F-pass: H = [sum_j w V_j du (x) xi] K^-1, P(H), A = P K^-1
// ---------------------------------------------------------------------------
// F-pass term of one bond.  The sum N = sum_j w V_j (u_j - u_i) (x) xi is
// formed as S - u_i (x) m with S = sum_j u_j (x) wx and m = sum_j wx in FP64,
// wx = w V_j xi rounded to FP32 once: both sums see the same rounded wx, so
// N is the FP64 sum without rounding u_j - u_i, and it is linear in u to FP64
// rounding (rounding u_j - u_i per bond made adaptive DR take ~10% more
// iterations to 1e-6).  FP64 sums: a 100+-term FP32 sum carries ~1e-6
// relative error into every A_j and from there into the residual.  A slot
// without a live bond passes the particle's own record (xi = 0), so its terms
// are exactly zero and no select is needed.
// S[r D + c] += u_j[r] wx_c and, with M, S[D D + c] += wx_c (fpass_sum_to_h
// forms N; without M the caller has sum_j wx from a pass that summed it)
template <int D, bool M = true>
PDX_HD void fpass_bond_u(const float *xi, const double *uj, float Vj, float invh, double *S)
{
    float L2 = 0.f;
    for (int r = 0; r < D; ++r) L2 = fmaf(xi[r], xi[r], L2);
    const float w = bond_weight_live(bond_sqrt(L2), invh, Vj);
    for (int c = 0; c < D; ++c) {
        const double wx = (double)(w * xi[c]);
        for (int r = 0; r < D; ++r) S[r * D + c] = fma(uj[r], wx, S[r * D + c]);
        if (M) S[D * D + c] += wx;
    }
}

// gathering kernels' F-pass term: the same sums as the halo F-pass (S of
// fpass_bond_u; fpass_sum_to_h forms the displacement-gradient sum)
template <int D>
PDX_HD void fpass_bond(const XRec &pi, const XRec &pj, float invh, double *S)
{
    float xi[3];
    for (int r = 0; r < 3; ++r) xi[r] = (float)(pj.X[r] - pi.X[r]);
    fpass_bond_u<D>(xi, pj.u, pj.V, invh, S);
}

// sum_j w V_j (u_j - u_i) (x) xi = S - u_i (x) sum_j wx from the sums of
// fpass_bond_u (FP64)
template <int D>
PDX_HD void fpass_sum_to_h(const double *S, const double *ui, double *Hd)
{
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) Hd[r * D + c] = S[r * D + c] - ui[r] * S[D * D + c];
}

// entry of bond slot k of a row, or "no bond" past the row's end
PDX_HD uint32_t bond_entry(const uint32_t *bp, int k, int nnb)
{
    return k < nnb ? bp[(int64_t)k * PDX_SLICE] : PDX_BROKEN;
}

template <int D, int MODEL, bool DIAG, bool COMMIT, bool STORE = true>
PDX_HD ARec fpass_finish_k(const Dev &d, int i, const KRec &kr, const double *Hd, const double *Ki);

template <int D, int MODEL, bool DIAG, bool COMMIT, bool STORE = true>
PDX_HD ARec fpass_finish(const Dev &d, int i, const KRec &kr, const double *Hd)
{
    return fpass_finish_k<D, MODEL, DIAG, COMMIT, STORE>(d, i, kr, Hd, d.Kinv64 + (int64_t)i * 9);
}

template <int D, int MODEL, bool DIAG, bool COMMIT, int G>
PDX_HD void fpass_one(const Dev &d, int i, int ubuf)
{
    const KRec kr = d.KR[i];
    if (!(kr.flags & PDX_FL_ACTIVE)) return;
    const XRec *__restrict__ X = d.xu[ubuf];
    const XRec pi = X[i];
    const float invh = (float)d.inv_h;
    double Sd[D * D + D];
    for (int q = 0; q < D * D + D; ++q) Sd[q] = 0.0;
    // bond loop in groups of G: the entries of the next group are loaded while
    // the current group's neighbour records are gathered
    const uint32_t *__restrict__ bp = d.bj + slot0(d, i);
    const int nnb = (int)kr.nnb;
    uint32_t ec[G], en[G];
    for (int q = 0; q < G; ++q) ec[q] = bond_entry(bp, q, nnb);
    for (int k = 0; k < nnb; k += G) {
        for (int q = 0; q < G; ++q) en[q] = bond_entry(bp, k + G + q, nnb);
        XRec pj[G];
        bool live[G];
        for (int q = 0; q < G; ++q) {
            // broken bonds (bit 31) and anything that is not a particle index
            // gather the particle's own record and get weight 0
            live[q] = ec[q] < (uint32_t)d.ni;
            pj[q] = X[live[q] ? (int)ec[q] : i];
        }
        for (int q = 0; q < G; ++q) fpass_bond<D>(pi, pj[q], invh, Sd);
        for (int q = 0; q < G; ++q) ec[q] = en[q];
    }
    double Hd[D * D];
    fpass_sum_to_h<D>(Sd, pi.u, Hd);
    fpass_finish<D, MODEL, DIAG, COMMIT>(d, i, kr, Hd);
}

// per-particle tensor algebra of the F-pass in FP64 with the FP64 inverse
// Ki (row i's K^-1, from d.Kinv64 or loaded ahead by the caller):
// H = Hd K^-1, P(H), A = P K^-1; A is rounded to FP32 when stored (into
// d.AR with STORE, else only returned)
template <int D, int MODEL, bool DIAG, bool COMMIT, bool STORE>
PDX_HD ARec fpass_finish_k(const Dev &d, int i, const KRec &kr, const double *Hd, const double *Ki)
{
    const bool sing = (kr.flags & PDX_FL_SING) != 0;
    double H[D * D];
    if (sing) {
        // K^-1 fell back to I, so F = N = K + sum w du (x) xi  (nosb.py:92-94, 126)
        const double *Kf = d.K64 + (int64_t)i * 9;
        for (int q = 0; q < D * D; ++q) H[q] = Hd[q] + Kf[q] - ((q % (D + 1)) == 0 ? 1.0 : 0.0);
    } else {
        for (int r = 0; r < D; ++r)
            for (int c = 0; c < D; ++c) {
                double acc = 0.0;
                for (int k = 0; k < D; ++k) acc += Hd[r * D + k] * Ki[k * D + c];
                H[r * D + c] = acc;
            }
    }
    double P[9];
    if constexpr (MODEL == 1) {
        stress_nh<D, double>(H, (double)kr.bulk, (double)kr.shear, P);
    } else if constexpr (MODEL == 2 && D == 3) {
        J2State ns;
        stress_j2(H, kr.bulk, kr.shear, d.j2p[i], d.j2s[i], P, COMMIT ? &ns : nullptr);
        if (COMMIT) d.j2s[i] = ns;
    } else {
        stress_le<D, double>(H, (double)kr.bulk, (double)kr.shear, P);
    }
    ARec ar;
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) {
            double acc = 0.0;
            if (sing) {
                acc = P[r * D + c];
            } else {
                for (int k = 0; k < D; ++k) acc += P[r * D + k] * Ki[k * D + c];
            }
            ar.A[r * D + c] = (float)acc;
        }
    for (int q = D * D; q < 9; ++q) ar.A[q] = 0.f;
    ar.pad[0] = ar.pad[1] = ar.pad[2] = 0.f;
    if (STORE) d.AR[i] = ar;
    if (DIAG) {
        for (int q = 0; q < D * D; ++q) {
            d.Fd[(int64_t)i * 9 + q] = (float)H[q];  // F - I: FP32 keeps the gradient's own precision
            d.Pd[(int64_t)i * 9 + q] = (float)P[q];
        }
    }
    return ar;
}

// ---------------------------------------------------------------------------
// force pass for one particle: f = A_i m_i + sum_j w V_j [A_j xi + g eta]
// + external load, then the DR update of UPD:
//   UPD 0: write f
//   UPD 1: adaptive DR (damping lagged one step), u_out = u_in + V; the damping
//          is 2 w with w^2 = -sum V.dF / sum lam V^2, the Rayleigh quotient of
//          the mass-scaled stiffness along the last increment
//   UPD 2: kinetic DR velocity update (nosb_solver.py:164-182)
// red[0] += |f|^2 over free DOFs; UPD 1: red[1], red[2] = Rayleigh-quotient
// numerator and denominator; UPD 2: red[1] = kinetic energy.
// ---------------------------------------------------------------------------
// force term of one bond, w V_j [(A_j - A_i) xi + g eta] (FP32), summed into
// fa (FP64, or FP32 for a partial sum the caller adds to FP64).  The
// correspondence term is in difference form:
// sum_j w (A_i + A_j) xi = sum_j w (A_j - A_i) xi + 2 A_i m_i, so each FP32
// bond product is rounded relative to |A_j - A_i| rather than |A|.
// HALO: the weight without the |xi| = 0 case, and ci, cj come premultiplied
// by G_s / 2 (NOSB) or 1/2 (bond form).
template <int D, int FORM, bool HALO = false, typename Acc = double>
PDX_HD void force_bond_v(const float *xi, const float *du, float Vj, float ci, float cj, const float *Ai,
                         const float *Aj, float invh, float Gs, Acc *fa)
{
    float L2 = 0.f, xd = 0.f, dd = 0.f;
    for (int r = 0; r < D; ++r) {
        L2 = fmaf(xi[r], xi[r], L2);
        xd = fmaf(xi[r], du[r], xd);
        dd = fmaf(du[r], du[r], dd);
    }
    const float L = bond_sqrt(L2);
    const float w = HALO ? bond_weight_live(L, invh, Vj) : bond_weight(L, invh, Vj);
    float eta[3];
    for (int r = 0; r < D; ++r) eta[r] = xi[r] + du[r];
    // |eta|^2 = |xi|^2 + (2 xi.du + du.du); the stretch uses the
    // cancellation-free |eta| - |xi| = (2 xi.du + du.du) / (|eta| + |xi|)
    const float dn = fmaf(2.f, xd, dd);
    const float el = bond_sqrt(fmaxf(L2 + dn, 0.f));
    float tb[3] = {0.f, 0.f, 0.f};
    float g = 0.f;
    if (FORM == 0) {
        for (int r = 0; r < D; ++r) {
            float t = 0.f;
            for (int c = 0; c < D; ++c) t = fmaf(Aj[r * D + c] - Ai[r * D + c], xi[c], t);
            tb[r] = t;
        }
        if (el > 1e-10f && L > 1e-10f)
            g = (HALO ? ci + cj : Gs * 0.5f * (ci + cj)) * dn * bond_rcp((el + L) * L * el);
    } else if (el > 1e-10f) {
        g = (HALO ? ci + cj : 0.5f * (ci + cj)) * dn * bond_rcp((el + L) * L * el);
    }
    for (int r = 0; r < D; ++r) {
        if constexpr (sizeof(Acc) == sizeof(float)) {
            fa[r] = fmaf(w, fmaf(g, eta[r], tb[r]), fa[r]);
        } else {
            const float contrib = w * fmaf(g, eta[r], tb[r]);
            fa[r] += (Acc)contrib;
        }
    }
}

template <int D, int FORM>
PDX_HD void force_bond(const XRec &pi, const XRec &pj, const float *Ai, const float *Aj, float invh, float Gs,
                       double *fa)
{
    float xi[3], du[3];
    for (int r = 0; r < 3; ++r) {
        xi[r] = (float)(pj.X[r] - pi.X[r]);
        du[r] = (float)(pj.u[r] - pi.u[r]);
    }
    force_bond_v<D, FORM>(xi, du, pj.V, pi.c, pj.c, Ai, Aj, invh, Gs, fa);
}

// per-particle inputs of force_finish (DR mass, velocity, previous force or
// acceleration, external load), which a kernel may load ahead of its use
struct FinishIn {
    double m;
    pdx_d4 v, fp, fe;
};

template <int UPD>
PDX_HD FinishIn finish_in(const Dev &d, int i, const KRec &kr)
{
    FinishIn fi{};
    if (kr.flags & PDX_FL_LOADED) fi.fe = d.fext[i];
    if (UPD == 1) fi.m = d.lam[i];
    if (UPD == 2) fi.m = d.mass[i];
    if (UPD != 0) {
        fi.v = d.v[i];
        fi.fp = d.fp[i];
    }
    return fi;
}

template <int D, int FORM, int UPD>
PDX_HD void force_finish(const Dev &d, int i, const KRec &kr, const XRec &pi, const ARec &ai, int ubuf, int cs,
                         double dt, double visc, double *fa, double *red, const FinishIn &fi);

template <int D, int FORM, int UPD, int G>
PDX_HD void force_one(const Dev &d, int i, int ubuf, int cs, double dt, double visc, double *red)
{
    const KRec kr = d.KR[i];
    if (!(kr.flags & PDX_FL_ACTIVE)) return;
    const XRec *__restrict__ X = d.xu[ubuf];
    const XRec pi = X[i];
    const float invh = (float)d.inv_h;
    const float Gs = (float)d.G_s;
    double fa[3] = {0.0, 0.0, 0.0};
    ARec ai{};
    if (FORM == 0) ai = d.AR[i];
    const uint32_t *__restrict__ bp = d.bj + slot0(d, i);
    const int nnb = (int)kr.nnb;
    uint32_t ec[G], en[G];
    for (int q = 0; q < G; ++q) ec[q] = bond_entry(bp, q, nnb);
    for (int k = 0; k < nnb; k += G) {
        for (int q = 0; q < G; ++q) en[q] = bond_entry(bp, k + G + q, nnb);
        XRec pj[G];
        ARec aj[G]{};
        bool live[G];
        for (int q = 0; q < G; ++q) {
            live[q] = ec[q] < (uint32_t)d.ni;
            const int j = live[q] ? (int)ec[q] : i;
            pj[q] = X[j];
            if (FORM == 0) aj[q] = d.AR[j];
        }
        for (int q = 0; q < G; ++q) force_bond<D, FORM>(pi, pj[q], ai.A, aj[q].A, invh, Gs, fa);
        for (int q = 0; q < G; ++q) ec[q] = en[q];
    }
    force_finish<D, FORM, UPD>(d, i, kr, pi, ai, ubuf, cs, dt, visc, fa, red, finish_in<UPD>(d, i, kr));
}

// per-particle end of the force pass: + 2 A_i m_i, + external load, |f|^2
// over free DOFs into red[0], then the update of UPD
template <int D, int FORM, int UPD>
PDX_HD void force_finish(const Dev &d, int i, const KRec &kr, const XRec &pi, const ARec &ai, int ubuf, int cs,
                         double dt, double visc, double *fa, double *red, const FinishIn &fi)
{
    if (FORM == 0) {
        for (int r = 0; r < D; ++r) {
            double t = 0.0;
            for (int c = 0; c < D; ++c) t += (double)ai.A[r * D + c] * (double)kr.m[c];
            fa[r] += 2.0 * t;
        }
    }
    if (kr.flags & PDX_FL_LOADED) {
        const pdx_d4 fe = fi.fe;
        fa[0] += fe.x;
        fa[1] += fe.y;
        if (D == 3) fa[2] += fe.z;
    }
    const uint32_t fm = kr.flags & PDX_FL_FIXMASK;
    for (int r = 0; r < D; ++r)
        if (!((fm >> r) & 1u)) red[0] += fa[r] * fa[r];
    if (UPD == 0) {
        d.f[i] = d4(fa[0], fa[1], D == 3 ? fa[2] : 0.0);
    } else if (UPD == 1) {
        const DScal &S = d.sc[cs];
        const double c = S.c;
        const int first = S.first;
        const double lam = fi.m;
        const pdx_d4 vo4 = fi.v, fo4 = fi.fp;
        const double vo[3] = {vo4.x, vo4.y, vo4.z}, fo[3] = {fo4.x, fo4.y, fo4.z};
        const double uo[3] = {pi.u[0], pi.u[1], pi.u[2]};
        double vn[3] = {0.0, 0.0, 0.0}, un[3] = {uo[0], uo[1], uo[2]}, fn[3] = {0.0, 0.0, 0.0};
        // two divisions per particle instead of two per DOF
        const double rl = 1.0 / lam, rc = 1.0 / (2.0 + c);
        for (int r = 0; r < D; ++r) {
            if ((fm >> r) & 1u) continue;
            const double F = fa[r];
            if (first) {
                vn[r] = 0.5 * F * rl;
            } else {
                red[1] -= vo[r] * (F - fo[r]);
                red[2] += lam * vo[r] * vo[r];
                vn[r] = ((2.0 - c) * vo[r] + 2.0 * F * rl) * rc;
            }
            fn[r] = F;
            un[r] = uo[r] + vn[r];
        }
        d.v[i] = d4(vn[0], vn[1], vn[2]);
        d.fp[i] = d4(fn[0], fn[1], fn[2]);
        double *uw = d.xu[ubuf ^ 1][i].u;
        uw[0] = un[0];
        uw[1] = un[1];
        uw[2] = un[2];
    } else {
        const double m = fi.m;
        const pdx_d4 vo4 = fi.v, ao4 = fi.fp;
        double vv[3] = {vo4.x, vo4.y, vo4.z}, aa[3] = {ao4.x, ao4.y, ao4.z};
        const double damp = 1.0 - visc;
        for (int r = 0; r < D; ++r) {
            if (!((fm >> r) & 1u) && m > 1e-20) {
                const double an = fa[r] / m;
                vv[r] += 0.5 * (aa[r] + an) * dt;
                vv[r] *= damp;
                aa[r] = an;
                red[1] += 0.5 * m * vv[r] * vv[r];
            } else {
                vv[r] = 0.0;
                aa[r] = 0.0;
            }
        }
        d.v[i] = d4(vv[0], vv[1], vv[2]);
        d.fp[i] = d4(aa[0], aa[1], aa[2]);
        d.f[i] = d4(fa[0], fa[1], D == 3 ? fa[2] : 0.0);
    }
}

// per-case scalar update done once per force pass by the case's last block
template <int UPD>
PDX_HD void case_update(DScal &S, const double *tot, int ubuf)
{
    const double res = sqrt(tot[0]);
    S.res = res;
    S.rel = res / S.ref;
    if (UPD == 1) {
        if (S.rel < S.tol && !S.converged) {
            S.converged = 1;
            S.conv_iter = S.iter;
            S.sol_buf = ubuf;
        }
        S.num = tot[1];
        S.den = tot[2];
        if (!S.first) {
            const double cn = (tot[2] > 0.0 && tot[1] > 0.0) ? 2.0 * sqrt(tot[1] / tot[2]) : 0.0;
            S.c = cn < S.cmax ? cn : S.cmax;
        }
        S.first = 0;
        S.iter += 1;
    } else if (UPD == 2) {
        double ke = tot[1];
        S.num = ke;  // the pass's kinetic energy, kept when a reset zeroes S.ke
        if (ke < S.prev_ke && S.inc) {
            S.reset_pending = 1;
            ke = 0.0;
            S.inc = 0;
            S.resets += 1;
        } else {
            S.reset_pending = 0;
            if (ke > S.prev_ke) S.inc = 1;
        }
        S.prev_ke = ke;
        S.ke = ke;
        S.iter += 1;
        if (S.rel < S.tol && !S.converged) {
            S.converged = 1;
            S.conv_iter = S.iter;
            S.sol_buf = ubuf;
        }
    }
}

// kinetic-damping position update x += v dt + a dt^2/2 on free DOFs
// (nosb_solver.py:156-162), after a pending velocity reset (:239-243)
template <int D>
PDX_HD void kd_move_one(const Dev &d, int i, int ubuf_in, double dt, bool reset)
{
    const KRec kr = d.KR[i];
    const double *ui = d.xu[ubuf_in][i].u;
    const pdx_d4 u = d4(ui[0], ui[1], ui[2]);
    pdx_d4 v = d.v[i];
    if (reset) {
        v = d4(0.0, 0.0, 0.0);
        d.v[i] = v;
    }
    const pdx_d4 a = d.fp[i];
    const uint32_t fm = kr.flags & PDX_FL_FIXMASK;
    pdx_d4 o = u;
    if (kr.flags & PDX_FL_ACTIVE) {
        if (!(fm & 1u)) o.x = u.x + v.x * dt + 0.5 * a.x * dt * dt;
        if (!(fm & 2u)) o.y = u.y + v.y * dt + 0.5 * a.y * dt * dt;
        if (D == 3 && !(fm & 4u)) o.z = u.z + v.z * dt + 0.5 * a.z * dt * dt;
    }
    double *uo = d.xu[ubuf_in ^ 1][i].u;
    uo[0] = o.x;
    uo[1] = o.y;
    uo[2] = o.z;
}

// ---------------------------------------------------------------------------
// setup (FP64)
// ---------------------------------------------------------------------------
// sort key of a bond: quantised (z, y, x) offset in 10 bits each, then the index
template <int D>
PDX_HD uint64_t stencil_key(const XRec &pi, const XRec &pj, uint32_t j, double hk)
{
    uint64_t key = 0;
    for (int q = D - 1; q >= 0; --q) {
        long long t = llround((pj.X[q] - pi.X[q]) * hk) + 512;
        t = t < 0 ? 0 : (t > 1023 ? 1023 : t);
        key = (key << 10) | (uint64_t)t;
    }
    return (key << 31) | (uint64_t)j;
}

// neighbour scan over the 27 (9) cells around i; distance summed as
// (dx*dx + dy*dy) + dz*dz without contraction, the reference's set test
// (neighbor.py:179-183, 199-203); FILL writes the row sorted by index
template <int D, bool FILL>
PDX_HD void neighbors_one(const Dev &d, int i)
{
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const int cs = d.case_of_p[i];
    if (cs < 0) {
        if (!FILL) d.counts[i] = 0;
        return;
    }
    const GridP g = d.grid[cs];
    const XRec pi = d.xu[0][i];
    int c[3] = {0, 0, 0};
    for (int q = 0; q < D; ++q) {
        int t = (int)floor((pi.X[q] - g.lo[q]) / g.h);
        c[q] = t < 0 ? 0 : (t >= g.nc[q] ? g.nc[q] - 1 : t);
    }
    const int rz = D == 3 ? 1 : 0;
    int cnt = 0;
    int64_t base = 0;
    if (FILL) base = slot0(d, i);
    const double hk = 256.0 / sqrt(d.h2);
    for (int a = c[0] - 1; a <= c[0] + 1; ++a) {
        if (a < 0 || a >= g.nc[0]) continue;
        for (int b = c[1] - 1; b <= c[1] + 1; ++b) {
            if (b < 0 || b >= g.nc[1]) continue;
            for (int e = c[2] - rz; e <= c[2] + rz; ++e) {
                if (e < 0 || e >= g.nc[2]) continue;
                const int64_t cell = g.cell_off + ((int64_t)a * g.nc[1] + b) * g.nc[2] + e;
                const int k0 = d.cell_start[cell], k1 = d.cell_start[cell + 1];
                for (int k = k0; k < k1; ++k) {
                    const int j = d.cell_idx[k];
                    if (j == i) continue;
                    const XRec pj = d.xu[0][j];
                    double s = 0.0;
                    for (int q = 0; q < D; ++q) {
                        const double t = pj.X[q] - pi.X[q];
                        s = s + t * t;
                    }
                    if (s < d.h2) {
                        if (FILL) d.nkey[base + (int64_t)cnt * PDX_SLICE] = stencil_key<D>(pi, pj, (uint32_t)j, hk);
                        cnt++;
                    }
                }
            }
        }
    }
    if (!FILL) {
        d.counts[i] = cnt;
        return;
    }
    // rows in stencil order: by the bond vector quantised to horizon/256
    // (z, then y, then x), ties by index -- on a lattice every interior
    // particle lists its bonds in the same order, so the k-th gathers of a
    // wave land on one shifted Morton brick.  The keys (index in the low 31
    // bits, so no two are equal) were stored as the scan found the bonds;
    // insertion sort, then the indices into the row
    uint64_t *__restrict__ kp = d.nkey + base;
    for (int p = 1; p < cnt; ++p) {
        const uint64_t key = kp[(int64_t)p * PDX_SLICE];
        int q = p - 1;
        while (q >= 0) {
            const uint64_t kq = kp[(int64_t)q * PDX_SLICE];
            if (kq <= key) break;
            kp[(int64_t)(q + 1) * PDX_SLICE] = kq;
            --q;
        }
        kp[(int64_t)(q + 1) * PDX_SLICE] = key;
    }
    for (int p = 0; p < cnt; ++p) d.bj[base + (int64_t)p * PDX_SLICE] = (uint32_t)(kp[(int64_t)p * PDX_SLICE] & 0x7fffffffu);
    d.init_nb[i] = cnt;
}

PDX_HD double det3(const double *M)
{
    return M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
           M[2] * (M[3] * M[7] - M[4] * M[6]);
}

// bonds whose data the FP64 bond loops of the set-up passes (shape tensor,
// adaptive-DR mass) gather at a time
#ifndef PDX_MASS_BATCH
#define PDX_MASS_BATCH 8
#endif

// shape tensor K = sum_intact w xi (x) xi V_j in FP64 with its inverse
// (nosb.py:59-94); mode 0: |det| > tol, mode 1: |det| > tol (tr K/dim)^dim;
// also m_i = sum_intact w V_j xi for the force pass
template <int D>
PDX_HD void shape_one(const Dev &d, int i, int mode, double tol)
{
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    KRec kr = d.KR[i];
    if (!(kr.flags & PDX_FL_ACTIVE)) return;
    const XRec pi = d.xu[0][i];
    double K[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double m[3] = {0, 0, 0};
    const int64_t base = slot0(d, i);
    const int nnb = (int)kr.nnb;
    // PDX_MASS_BATCH bonds per step, their slots, positions and volumes
    // requested before any is used (a broken bond or one past the row reads
    // the particle's own record and is skipped); the sums keep the bond order
    for (int k0 = 0; k0 < nnb; k0 += PDX_MASS_BATCH) {
        uint32_t e[PDX_MASS_BATCH];
        for (int b = 0; b < PDX_MASS_BATCH; ++b)
            e[b] = k0 + b < nnb ? d.bj[base + (int64_t)(k0 + b) * PDX_SLICE] : PDX_BROKEN;
        double X[PDX_MASS_BATCH][3], Vb[PDX_MASS_BATCH];
        for (int b = 0; b < PDX_MASS_BATCH; ++b) {
            const int j = (e[b] & PDX_BROKEN) ? i : (int)e[b];
            for (int q = 0; q < 3; ++q) X[b][q] = d.xu[0][j].X[q];
            Vb[b] = d.V64[j];
        }
        for (int b = 0; b < PDX_MASS_BATCH; ++b) {
            if (e[b] & PDX_BROKEN) continue;
            double xi[3] = {0, 0, 0}, s = 0.0;
            for (int q = 0; q < D; ++q) {
                xi[q] = X[b][q] - pi.X[q];
                s = s + xi[q] * xi[q];
            }
            const double L = sqrt(s);
            double w = 1.0;
            if (L > 0) {
                w = 1.0 - L / d.h;
                w = w > 0.0 ? w : 0.0;
            }
            const double Vj = Vb[b];
            for (int r = 0; r < D; ++r) {
                for (int c = 0; c < D; ++c) K[r * D + c] += w * xi[r] * xi[c] * Vj;
                m[r] += w * Vj * xi[r];
            }
        }
    }
    double tr = 0.0;
    for (int q = 0; q < D; ++q) tr += K[q * D + q];
    tr /= D;
    const double det = D == 3 ? det3(K) : K[0] * K[3] - K[1] * K[2];
    const double thr = mode == 1 ? tol * (D == 3 ? tr * tr * tr : tr * tr) : tol;
    const bool sing = !(fabs(det) > thr && det != 0.0);
    double Ki[9];
    if (!sing) {
        if (D == 3) {
            Ki[0] = (K[4] * K[8] - K[5] * K[7]) / det;
            Ki[1] = (K[2] * K[7] - K[1] * K[8]) / det;
            Ki[2] = (K[1] * K[5] - K[2] * K[4]) / det;
            Ki[3] = (K[5] * K[6] - K[3] * K[8]) / det;
            Ki[4] = (K[0] * K[8] - K[2] * K[6]) / det;
            Ki[5] = (K[2] * K[3] - K[0] * K[5]) / det;
            Ki[6] = (K[3] * K[7] - K[4] * K[6]) / det;
            Ki[7] = (K[1] * K[6] - K[0] * K[7]) / det;
            Ki[8] = (K[0] * K[4] - K[1] * K[3]) / det;
        } else {
            Ki[0] = K[3] / det; Ki[1] = -K[1] / det; Ki[2] = -K[2] / det; Ki[3] = K[0] / det;
        }
    } else {
        for (int q = 0; q < D * D; ++q) Ki[q] = (q % (D + 1)) == 0 ? 1.0 : 0.0;
    }
    for (int q = 0; q < D * D; ++q) {
        d.K64[(int64_t)i * 9 + q] = K[q];
        d.Kinv64[(int64_t)i * 9 + q] = Ki[q];
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            kr.Kinv[r * 3 + c] = (r < D && c < D) ? (float)(sing ? K[r * D + c] : Ki[r * D + c]) : 0.f;
    for (int q = 0; q < 3; ++q) kr.m[q] = (float)m[q];
    kr.flags = sing ? (kr.flags | PDX_FL_SING) : (kr.flags & ~PDX_FL_SING);
    d.KR[i] = kr;
}

// adaptive-DR mass (a Gershgorin bound of the force-density stiffness):
//   g_ik = w V_k Kinv_i xi_ik, G_i = sum_k g_ik, S_i = sum_k |g_ik| + |G_i|,
//   C_i = max(3 K_i, 2 mu_i), k_corr_i = sqrt3 (C_i S_i |G_i| + sum_j w V_j C_j S_j |Kinv_j xi_ji|),
//   k_stab_i = 2 sqrt3 G_s sum_j c_ij w V_j / |xi|,  lam_i = beta/4 (k_corr_i + k_stab_i)
// pass A writes C_i S_i and |G_i|
template <int D>
PDX_HD void adr_mass_a_one(const Dev &d, int i, double *cs_out, double *g_out)
{
    const KRec kr = d.KR[i];
    cs_out[i] = 0.0;
    g_out[i] = 0.0;
    if (!(kr.flags & PDX_FL_ACTIVE) || d.form != 0) return;
    const XRec pi = d.xu[0][i];
    const double *Ki = d.Kinv64 + (int64_t)i * 9;
    double G[3] = {0, 0, 0}, sabs = 0.0;
    const int64_t base = slot0(d, i);
    const int nnb = (int)kr.nnb;
    // PDX_MASS_BATCH bonds per step: their slots, then their positions and
    // volumes, are all requested before any is used (a broken bond or one
    // past the row reads the particle's own record and is skipped); the
    // sums keep the bond order
    for (int k0 = 0; k0 < nnb; k0 += PDX_MASS_BATCH) {
        uint32_t e[PDX_MASS_BATCH];
        for (int b = 0; b < PDX_MASS_BATCH; ++b)
            e[b] = k0 + b < nnb ? d.bj[base + (int64_t)(k0 + b) * PDX_SLICE] : PDX_BROKEN;
        double X[PDX_MASS_BATCH][3], Vj[PDX_MASS_BATCH];
        for (int b = 0; b < PDX_MASS_BATCH; ++b) {
            const int j = (e[b] & PDX_BROKEN) ? i : (int)e[b];
            for (int q = 0; q < 3; ++q) X[b][q] = d.xu[0][j].X[q];
            Vj[b] = d.V64[j];
        }
        for (int b = 0; b < PDX_MASS_BATCH; ++b) {
            if (e[b] & PDX_BROKEN) continue;
            double xi[3] = {0, 0, 0}, s = 0.0;
            for (int q = 0; q < D; ++q) {
                xi[q] = X[b][q] - pi.X[q];
                s += xi[q] * xi[q];
            }
            const double L = sqrt(s);
            const double w = L > 0 ? fmax(0.0, 1.0 - L / d.h) : 1.0;
            double gn = 0.0;
            for (int r = 0; r < D; ++r) {
                double t = 0.0;
                for (int c = 0; c < D; ++c) t += Ki[r * D + c] * xi[c];
                t *= w * Vj[b];
                G[r] += t;
                gn += t * t;
            }
            sabs += sqrt(gn);
        }
    }
    double gl = 0.0;
    for (int r = 0; r < D; ++r) gl += G[r] * G[r];
    gl = sqrt(gl);
    const double C = fmax(3.0 * (double)kr.bulk, 2.0 * (double)kr.shear);
    cs_out[i] = C * (sabs + gl);
    g_out[i] = gl;
}

template <int D>
PDX_HD void adr_mass_b_one(const Dev &d, int i, const double *cs_in, const double *g_in, double beta,
                           double *raw = nullptr)
{
    const KRec kr = d.KR[i];
    if (!(kr.flags & PDX_FL_ACTIVE)) {
        if (raw)
            raw[i] = NAN;  // the scaled mass of an inactive row is 1 whatever beta
        else
            d.lam[i] = 1.0;
        return;
    }
    const XRec pi = d.xu[0][i];
    const bool nosb = d.form == 0;
    double kc = nosb ? cs_in[i] * g_in[i] : 0.0, ks = 0.0;
    const int64_t base = slot0(d, i);
    const int nnb = (int)kr.nnb;
    // PDX_MASS_BATCH bonds per step, their data requested before any is used
    // (as in adr_mass_a_one); the sums keep the bond order
    for (int k0 = 0; k0 < nnb; k0 += PDX_MASS_BATCH) {
        uint32_t e[PDX_MASS_BATCH];
        for (int b = 0; b < PDX_MASS_BATCH; ++b)
            e[b] = k0 + b < nnb ? d.bj[base + (int64_t)(k0 + b) * PDX_SLICE] : PDX_BROKEN;
        double X[PDX_MASS_BATCH][3], Vj[PDX_MASS_BATCH], csj[PDX_MASS_BATCH], Kj[PDX_MASS_BATCH][9];
        float cj[PDX_MASS_BATCH];
        for (int b = 0; b < PDX_MASS_BATCH; ++b) {
            const int j = (e[b] & PDX_BROKEN) ? i : (int)e[b];
            for (int q = 0; q < 3; ++q) X[b][q] = d.xu[0][j].X[q];
            cj[b] = d.xu[0][j].c;
            Vj[b] = d.V64[j];
            csj[b] = nosb ? cs_in[j] : 0.0;
            for (int q = 0; q < D * D; ++q) Kj[b][q] = nosb ? d.Kinv64[(int64_t)j * 9 + q] : 0.0;
        }
        for (int b = 0; b < PDX_MASS_BATCH; ++b) {
            if (e[b] & PDX_BROKEN) continue;
            double xi[3] = {0, 0, 0}, s = 0.0;
            for (int q = 0; q < D; ++q) {
                xi[q] = X[b][q] - pi.X[q];
                s += xi[q] * xi[q];
            }
            const double L = sqrt(s);
            const double w = L > 0 ? fmax(0.0, 1.0 - L / d.h) : 1.0;
            if (nosb) {
                double gn = 0.0;
                for (int r = 0; r < D; ++r) {
                    double t = 0.0;
                    for (int c = 0; c < D; ++c) t -= Kj[b][r * D + c] * xi[c];
                    gn += t * t;
                }
                kc += w * Vj[b] * csj[b] * sqrt(gn);
            }
            if (L > 1e-10) ks += 0.5 * ((double)pi.c + (double)cj[b]) * w * Vj[b] / L;
        }
    }
    const double gs = nosb ? d.G_s : 1.0;
    const double r3 = 1.7320508075688772;
    if (raw) {
        // the Gershgorin mass before beta: beta * raw rounds as 0.25 beta (...) does
        raw[i] = 0.25 * (r3 * kc + 2.0 * r3 * gs * ks);
        return;
    }
    double lam = 0.25 * beta * (r3 * kc + 2.0 * r3 * gs * ks);
    if (!(lam > 0.0)) lam = 1.0;
    d.lam[i] = lam;
}

// the adaptive-DR mass beta x the cached Gershgorin mass (adr_mass_b_one's raw
// values), 1 where that is not positive, as adr_mass_b_one computes it
PDX_HD void mass_scale_one(const Dev &d, int i, const double *raw, double beta)
{
    double lam = beta * raw[i];
    if (!(lam > 0.0)) lam = 1.0;
    d.lam[i] = lam;
}

// dt estimate per particle (nosb_solver.py:121-143): lam_i = k_eff / m_i
template <int D>
PDX_HD void dt_est_one(const Dev &d, int i, double modulus, double *out)
{
    const KRec kr = d.KR[i];
    out[i] = 0.0;
    if (!(kr.flags & PDX_FL_ACTIVE)) return;
    const double mi = d.mass[i];
    if (mi < 1e-30) return;
    const XRec pi = d.xu[0][i];
    const double *Ki = d.Kinv64 + (int64_t)i * 9;
    double ds[3] = {0, 0, 0}, dsq = 0.0;
    const int64_t base = slot0(d, i);
    for (int k = 0; k < (int)kr.nnb; ++k) {
        const uint32_t e = d.bj[base + (int64_t)k * PDX_SLICE];
        if (e & PDX_BROKEN) continue;
        const XRec pj = d.xu[0][(int)e];
        double xi[3] = {0, 0, 0}, s = 0.0;
        for (int q = 0; q < D; ++q) {
            xi[q] = pj.X[q] - pi.X[q];
            s += xi[q] * xi[q];
        }
        const double L = sqrt(s);
        const double w = L > 0 ? fmax(0.0, 1.0 - L / d.h) : 1.0;
        for (int r = 0; r < D; ++r) {
            double t = 0.0;
            for (int c = 0; c < D; ++c) t += Ki[r * D + c] * xi[c];
            t *= w;
            ds[r] += t;
            dsq += t * t;
        }
    }
    double s2 = 0.0;
    for (int r = 0; r < D; ++r) s2 += ds[r] * ds[r];
    out[i] = modulus * d.V64[i] * (s2 + dsq) / mi;
}

// critical-stretch bond breaking with the FP64 stretch of bonds.py:161-180:
// x = X + u, eta = x_j - x_i, s = (|eta| - |xi|)/|xi| if |xi| > 1e-10, broken if
// s > s_c strictly (damage.py:94-116); returns the bonds broken now
template <int D>
PDX_HD unsigned damage_one(const Dev &d, int i, int ubuf, double s_c)
{
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const KRec kr = d.KR[i];
    if (!(kr.flags & PDX_FL_ACTIVE)) return 0;
    const XRec *X = d.xu[ubuf];
    const XRec pi = X[i];
    const double xi0[3] = {pi.X[0] + pi.u[0], pi.X[1] + pi.u[1], pi.X[2] + pi.u[2]};
    const int64_t base = slot0(d, i);
    unsigned nb = 0;
    const int nnb = (int)kr.nnb;
    // PDX_MASS_BATCH bonds per step, their slots and the neighbours' X and u
    // requested before any is used (a broken bond or one past the row reads
    // the particle's own record and is skipped)
    for (int k0 = 0; k0 < nnb; k0 += PDX_MASS_BATCH) {
        uint32_t e[PDX_MASS_BATCH];
        for (int b = 0; b < PDX_MASS_BATCH; ++b)
            e[b] = k0 + b < nnb ? d.bj[base + (int64_t)(k0 + b) * PDX_SLICE] : PDX_BROKEN;
        double Xj[PDX_MASS_BATCH][3], Uj[PDX_MASS_BATCH][3];
        for (int b = 0; b < PDX_MASS_BATCH; ++b) {
            const int j = (e[b] & PDX_BROKEN) ? i : (int)e[b];
            for (int q = 0; q < 3; ++q) {
                Xj[b][q] = X[j].X[q];
                Uj[b][q] = X[j].u[q];
            }
        }
        for (int b = 0; b < PDX_MASS_BATCH; ++b) {
            if (e[b] & PDX_BROKEN) continue;
            const int64_t a = base + (int64_t)(k0 + b) * PDX_SLICE;
            const double xj[3] = {Xj[b][0] + Uj[b][0], Xj[b][1] + Uj[b][1], Xj[b][2] + Uj[b][2]};
            double se = 0.0, sl = 0.0;
            for (int q = 0; q < D; ++q) {
                const double t = xj[q] - xi0[q];
                se = se + t * t;
                const double r = Xj[b][q] - pi.X[q];
                sl = sl + r * r;
            }
            const double el = sqrt(se), L = sqrt(sl);
            double st = 0.0;
            if (L > 1e-10) st = (el - L) / L;
            if (st > s_c) {
                d.bj[a] = e[b] | PDX_BROKEN;
                if (d.bl) {
                    d.bl[a] |= PDX_HBROKEN;
                    const uint16_t own = d.hs[d.hl_ptr[i / PDX_BLOCK] + i % PDX_BLOCK];
                    d.bk[a] = own;
                    if (d.bks) d.bks[lat_at(i, d.lk[a])] = own;
                }
                nb++;
            }
        }
    }
    return nb;
}

// phi_i = broken / initial bonds (damage.py:118-142)
PDX_HD void local_damage_one(const Dev &d, int i)
{
    const KRec kr = d.KR[i];
    const int init = d.init_nb[i];
    if (!(kr.flags & PDX_FL_ACTIVE) || init <= 0) {
        d.damage[i] = 0.0;
        return;
    }
    const int64_t base = slot0(d, i);
    int nbk = 0;
    for (int k = 0; k < (int)kr.nnb; ++k) nbk += (d.bj[base + (int64_t)k * PDX_SLICE] & PDX_BROKEN) ? 1 : 0;
    d.damage[i] = (double)nbk / (double)init;
}

// per-bond reference data in internal CSR order (bonds.py:116-137)
template <int D>
PDX_HD void bond_data_one(const Dev &d, int i, double *xi_o, double *len_o, double *w_o, const int64_t *rowi)
{
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const KRec kr = d.KR[i];
    if (!(kr.flags & PDX_FL_ACTIVE)) return;
    const XRec pi = d.xu[0][i];
    const int64_t base = slot0(d, i);
    for (int k = 0; k < (int)kr.nnb; ++k) {
        const uint32_t e = d.bj[base + (int64_t)k * PDX_SLICE];
        const XRec pj = d.xu[0][(int)(e & ~PDX_BROKEN)];
        const int64_t b = rowi[i] + k;
        double s = 0.0;
        for (int q = 0; q < D; ++q) {
            const double t = pj.X[q] - pi.X[q];
            if (xi_o) xi_o[b * D + q] = t;
            s = s + t * t;
        }
        const double L = sqrt(s);
        if (len_o) len_o[b] = L;
        if (w_o) w_o[b] = L > 0 ? fmax(0.0, 1.0 - L / d.h) : 1.0;
    }
}

// weighted neighbour mean of a per-particle feature vector (width <= 64) over
// the intact bonds: out_i = sum_j w_ij V_j in_j / sum_j w_ij V_j (FP64), the
// message-passing step of the learned warm start; a particle without intact
// bonds keeps its own value.  Columns [q0, q0 + CH) per call: CH is a
// compile-time count so that the accumulators stay in registers.
template <int CH>
PDX_HD void nbr_mean_one(const Dev &d, int i, const double *in, double *out, int width, int q0)
{
    const KRec kr = d.KR[i];
    const XRec pi = d.xu[0][i];
    double acc[CH];
    for (int q = 0; q < CH; ++q) acc[q] = 0.0;
    double ws = 0.0;
    const int64_t base = slot0(d, i);
    if (kr.flags & PDX_FL_ACTIVE) {
        for (int k = 0; k < (int)kr.nnb; ++k) {
            const uint32_t e = d.bj[base + (int64_t)k * PDX_SLICE];
            if (e >= (uint32_t)d.ni) continue;  // broken bond or no bond
            const XRec pj = d.xu[0][(int)e];
            double s = 0.0;
            for (int q = 0; q < 3; ++q) s += (pj.X[q] - pi.X[q]) * (pj.X[q] - pi.X[q]);
            const double L = sqrt(s);
            const double w = (L > 0 ? fmax(0.0, 1.0 - L / d.h) : 1.0) * d.V64[(int)e];
            const double *row = in + (int64_t)e * width + q0;
            if constexpr (CH == 20 || CH == 10) {
                // launched for width CH only: the 160- or 80-byte row in 16-byte loads
                struct alignas(16) Pair {
                    double a, b;
                };
                for (int q = 0; q < CH; q += 2) {
                    const Pair t = *(const Pair *)(row + q);
                    acc[q] += w * t.a;
                    acc[q + 1] += w * t.b;
                }
            } else {
                for (int q = 0; q < CH; ++q)
                    if (q0 + q < width) acc[q] += w * row[q];
            }
            ws += w;
        }
    }
    for (int q = 0; q < CH; ++q)
        if (q0 + q < width)
            out[(int64_t)i * width + q0 + q] = ws > 0.0 ? acc[q] / ws : in[(int64_t)i * width + q0 + q];
}

// |f|^2 over free DOFs from the stored force
template <int D>
PDX_HD double resid_one(const Dev &d, int i)
{
    const KRec kr = d.KR[i];
    if (!(kr.flags & PDX_FL_ACTIVE)) return 0.0;
    const pdx_d4 f = d.f[i];
    const double fv[3] = {f.x, f.y, f.z};
    const uint32_t fm = kr.flags & PDX_FL_FIXMASK;
    double s = 0.0;
    for (int r = 0; r < D; ++r)
        if (!((fm >> r) & 1u)) s += fv[r] * fv[r];
    return s;
}

// free DOFs of u set to zero (cold-start state)
PDX_HD void zero_free_one(const Dev &d, int i, int ubuf)
{
    const uint32_t fm = d.KR[i].flags & PDX_FL_FIXMASK;
    double *u = d.xu[ubuf][i].u;
    for (int r = 0; r < 3; ++r)
        if (!((fm >> r) & 1u)) u[r] = 0.0;
}
