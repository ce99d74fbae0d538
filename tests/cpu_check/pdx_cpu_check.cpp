// pdx_cpu_check.cpp -- bounds-checked CPU run of the libpdx kernel bodies.
//
// Test infrastructure: compiles simul_amd/csrc/pdx_device.h (the per-particle
// functions every HIP kernel wraps) and pdx_layout.h (the library's host-side
// layout) with a host compiler under AddressSanitizer / UBSan, builds the same
// internal layout the library builds, and sweeps every kernel body over every
// particle with each array allocated at exactly the size the library
// allocates.  Any out-of-bounds access aborts with a sanitizer report.
//
// Input (binary, little endian, written by tests/test_cpu_check.py):
//   int32 n, dim, model, form, ncase; int64 offsets[ncase+1];
//   double horizon, G_s; double X[n*dim], vol[n], mass[n], bulk[n], shear[n],
//   cst[n], fext[n*dim], u0[n*dim]; uint8 fixed[n];
//   if model == 2: double sy[n], hiso[n], hkin[n]
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "pdx_layout.h"

template <typename T>
static T *alloc(size_t n)
{
    if (n == 0) n = 1;
    T *p = (T *)malloc(sizeof(T) * n);
    memset((void *)p, 0, sizeof(T) * n);
    return p;
}

template <typename T>
static void rd(FILE *f, T *p, size_t n)
{
    if (fread(p, sizeof(T), n, f) != n) {
        fprintf(stderr, "short read\n");
        exit(2);
    }
}

struct Run {
    Dev d{};
    int D = 3, model = 0, form = 0, ncase = 1;
    std::vector<int32_t> p0, perm, inv;
    int64_t slots = 0, nb = 0;
};

template <int D, int MODEL, int FORM, int G>
static double sweep(Run &R, int iters)
{
    Dev &d = R.d;
    const int ni = d.ni;
    double check = 0.0;
    for (int i = 0; i < ni; ++i) shape_one<D>(d, i, 1, 1e-12);
    for (int i = 0; i < ni; ++i) zero_free_one(d, i, 1);
    for (int i = 0; i < ni; ++i) fpass_one<D, MODEL, true, false, 1>(d, i, 0);
    std::vector<double> red(3 * ni, 0.0);
    for (int i = 0; i < ni; ++i) {
        const int cs = d.case_of_blk[i / PDX_BLOCK];
        force_one<D, FORM, 0, 1>(d, i, 0, cs, 0.0, 0.0, &red[3 * i]);
    }
    for (int i = 0; i < ni; ++i) check += red[3 * i] + resid_one<D>(d, i);
    // adaptive-DR mass, then a few adaptive iterations with the case update
    double *cs_buf = alloc<double>(ni), *g_buf = alloc<double>(ni);
    for (int i = 0; i < ni; ++i) adr_mass_a_one<D>(d, i, cs_buf, g_buf);
    for (int i = 0; i < ni; ++i) adr_mass_b_one<D>(d, i, cs_buf, g_buf, 1.0);
    free(cs_buf);
    free(g_buf);
    int ub = 0;
    for (int it = 0; it < iters; ++it) {
        for (int i = 0; i < ni; ++i) fpass_one<D, MODEL, false, false, G>(d, i, ub);
        std::vector<double> tot(3 * R.ncase, 0.0);
        for (int i = 0; i < ni; ++i) {
            const int cs = d.case_of_blk[i / PDX_BLOCK];
            double r[3] = {0, 0, 0};
            force_one<D, FORM, 1, G>(d, i, ub, cs, 0.0, 0.0, r);
            for (int q = 0; q < 3; ++q) tot[3 * cs + q] += r[q];
        }
        for (int k = 0; k < R.ncase; ++k) case_update<1>(d.sc[k], &tot[3 * k], ub);
        ub ^= 1;
    }
    // kinetic-damping step
    for (int i = 0; i < ni; ++i) kd_move_one<D>(d, i, ub, 1e-3, (i & 1) != 0);
    ub ^= 1;
    for (int i = 0; i < ni; ++i) fpass_one<D, MODEL, false, false, G>(d, i, ub);
    for (int i = 0; i < ni; ++i) {
        double r[3] = {0, 0, 0};
        force_one<D, FORM, 2, G>(d, i, ub, d.case_of_blk[i / PDX_BLOCK], 1e-3, 0.0, r);
        check += r[0];
    }
    if (MODEL == 2)
        for (int i = 0; i < ni; ++i) fpass_one<D, MODEL, false, true, 1>(d, i, ub);
    // setup and damage bodies
    double *dtb = alloc<double>(ni);
    for (int i = 0; i < ni; ++i) dt_est_one<D>(d, i, 1.0, dtb);
    free(dtb);
    int64_t *rowi = alloc<int64_t>(ni + 1);
    for (int i = 0; i < ni; ++i) rowi[i + 1] = rowi[i] + d.KR[i].nnb;
    double *xi = alloc<double>((size_t)rowi[ni] * D), *len = alloc<double>(rowi[ni]), *w = alloc<double>(rowi[ni]);
    for (int i = 0; i < ni; ++i) bond_data_one<D>(d, i, xi, len, w, rowi);
    for (int b = 0; b < rowi[ni]; ++b) check += w[b];
    free(xi);
    free(len);
    free(w);
    free(rowi);
    unsigned nbk = 0;
    for (int i = 0; i < ni; ++i) nbk += damage_one<D>(d, i, ub, 1e-4);
    for (int i = 0; i < ni; ++i) local_damage_one(d, i);
    for (int i = 0; i < ni; ++i) shape_one<D>(d, i, 0, 1e-12);
    printf("broken now %u\n", nbk);
    return check;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s input.bin [iters] [group]\n", argv[0]);
        return 2;
    }
    const int iters = argc > 2 ? atoi(argv[2]) : 3;
    const int gsel = argc > 3 ? atoi(argv[3]) : 0;  // bond-group size to sweep (0 = all)
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t hdr[5];
    rd(f, hdr, 5);
    const int n = hdr[0], D = hdr[1];
    Run R;
    R.D = D;
    R.model = hdr[2];
    R.form = hdr[3];
    R.ncase = hdr[4];
    std::vector<int64_t> off(R.ncase + 1);
    rd(f, off.data(), off.size());
    double hg[2];
    rd(f, hg, 2);
    std::vector<double> X((size_t)n * D), vol(n), mass(n), bulk(n), shear(n), cst(n), fext((size_t)n * D),
        u0((size_t)n * D), sy(n), hiso(n), hkin(n);
    std::vector<uint8_t> fixed(n);
    rd(f, X.data(), X.size());
    rd(f, vol.data(), n);
    rd(f, mass.data(), n);
    rd(f, bulk.data(), n);
    rd(f, shear.data(), n);
    rd(f, cst.data(), n);
    rd(f, fext.data(), fext.size());
    rd(f, u0.data(), u0.size());
    rd(f, fixed.data(), n);
    if (R.model == 2) {
        rd(f, sy.data(), n);
        rd(f, hiso.data(), n);
        rd(f, hkin.data(), n);
    }
    fclose(f);

    R.p0 = pdx_layout::case_starts(off);
    const int ni = R.p0[R.ncase];
    const int nblk = ((ni / PDX_BLOCK + PDX_NXCD - 1) / PDX_NXCD) * PDX_NXCD;
    std::vector<uint8_t> colour;
    pdx_layout::morton_order(D, X.data(), vol.data(), off, R.p0, R.perm, R.inv, colour);
    Dev &d = R.d;
    d.ni = ni;
    d.nblk = nblk;
    d.ncase = R.ncase;
    d.dim = D;
    d.xu[0] = alloc<XRec>(ni);
    d.xu[1] = alloc<XRec>(ni);
    d.KR = alloc<KRec>(ni);
    d.AR = alloc<ARec>(ni);
    d.j2p = alloc<J2Par>(ni);
    d.j2s = alloc<J2State>(ni);
    d.v = alloc<pdx_d4>(ni);
    d.fp = alloc<pdx_d4>(ni);
    d.f = alloc<pdx_d4>(ni);
    d.fext = alloc<pdx_d4>(ni);
    d.mass = alloc<double>(ni);
    d.lam = alloc<double>(ni);
    d.V64 = alloc<double>(ni);
    d.K64 = alloc<double>((size_t)ni * 9);
    d.Kinv64 = alloc<double>((size_t)ni * 9);
    d.Fd = alloc<float>((size_t)ni * 9);
    d.Pd = alloc<float>((size_t)ni * 9);
    d.damage = alloc<double>(ni);
    d.init_nb = alloc<int32_t>(ni);
    d.case_of_p = alloc<int32_t>(ni);
    d.counts = alloc<int32_t>(ni);
    d.case_of_blk = alloc<int32_t>(nblk);
    d.case_blk0 = alloc<int32_t>(R.ncase + 1);
    d.sc = alloc<DScal>(R.ncase);
    d.part = alloc<double>((size_t)nblk * 4);
    d.grid = alloc<GridP>(R.ncase);
    d.G_s = hg[1];
    d.model = R.model;
    d.form = R.form;
    std::vector<int32_t> cob, cb0;
    pdx_layout::block_cases(R.p0, nblk, cob, cb0);
    memcpy(d.case_of_blk, cob.data(), sizeof(int32_t) * nblk);
    memcpy(d.case_blk0, cb0.data(), sizeof(int32_t) * cb0.size());
    for (int k = 0; k < R.ncase; ++k) {
        d.sc[k].ref = 1.0;
        d.sc[k].cmax = 1.9;
        d.sc[k].first = 1;
        d.sc[k].inc = 1;
        d.sc[k].tol = -1.0;
        for (int i = R.p0[k]; i < R.p0[k + 1]; ++i)
            if (R.perm[i] >= 0) d.case_of_p[i] = k;
            else d.case_of_p[i] = -1;
    }
    const uint8_t full = D == 3 ? 7 : 3;
    for (int i = 0; i < ni; ++i) {
        const int r = R.perm[i];
        if (r < 0) continue;
        XRec p{};
        for (int q = 0; q < D; ++q) p.X[q] = X[(size_t)r * D + q];
        p.V = (float)vol[r];
        p.c = (float)cst[r];
        d.xu[1][i] = p;
        for (int q = 0; q < D; ++q) p.u[q] = u0[(size_t)r * D + q];
        d.xu[0][i] = p;
        KRec &k = d.KR[i];
        k.bulk = (float)bulk[r];
        k.shear = (float)shear[r];
        k.flags = PDX_FL_ACTIVE | (fixed[r] & full);
        bool loaded = false;
        for (int q = 0; q < D; ++q) loaded |= fext[(size_t)r * D + q] != 0.0;
        if (loaded) k.flags |= PDX_FL_LOADED;
        d.fext[i] = d4(fext[(size_t)r * D], fext[(size_t)r * D + 1], D == 3 ? fext[(size_t)r * D + 2] : 0.0);
        d.mass[i] = mass[r];
        d.V64[i] = vol[r];
        d.j2p[i] = J2Par{(float)sy[r], (float)hiso[r], (float)hkin[r], 0.f};
    }
    // neighbour grid, count, SELL, fill (pdx_build_bonds)
    std::vector<GridP> gp;
    std::vector<int32_t> cstart, cidx;
    std::string err;
    if (!pdx_layout::cell_grids(D, X.data(), R.p0, R.perm, hg[0], gp, cstart, cidx, err)) {
        fprintf(stderr, "grid: %s\n", err.c_str());
        return 3;
    }
    memcpy(d.grid, gp.data(), sizeof(GridP) * gp.size());
    d.cell_start = alloc<int32_t>(cstart.size());
    memcpy(d.cell_start, cstart.data(), sizeof(int32_t) * cstart.size());
    d.cell_idx = alloc<int32_t>(cidx.size());
    memcpy(d.cell_idx, cidx.data(), sizeof(int32_t) * cidx.size());
    d.h = hg[0];
    d.inv_h = 1.0 / hg[0];
    d.h2 = hg[0] * hg[0];
    for (int i = 0; i < ni; ++i) {
        if (D == 3) neighbors_one<3, false>(d, i);
        else neighbors_one<2, false>(d, i);
    }
    std::vector<int32_t> cnt(d.counts, d.counts + ni);
    std::vector<int64_t> slptr;
    R.slots = pdx_layout::sell_pointers(cnt, slptr, R.nb);
    d.sl_ptr = alloc<int64_t>(slptr.size());
    memcpy(d.sl_ptr, slptr.data(), sizeof(int64_t) * slptr.size());
    d.bj = alloc<uint32_t>(R.slots);
    memset(d.bj, 0xff, sizeof(uint32_t) * (R.slots > 0 ? R.slots : 1));
    d.nkey = alloc<uint64_t>(R.slots);
    for (int i = 0; i < ni; ++i) {
        if (D == 3) neighbors_one<3, true>(d, i);
        else neighbors_one<2, true>(d, i);
    }
    for (int i = 0; i < ni; ++i) {
        d.KR[i].nnb = (uint32_t)cnt[i];
        if (d.init_nb[i] != cnt[i]) {
            fprintf(stderr, "count/fill mismatch at %d\n", i);
            return 4;
        }
    }
    {
        std::vector<int64_t> row, slot, csr;
        std::vector<int32_t> col;
        if (!pdx_layout::bond_rows(n, ni, R.perm, R.inv, cnt, slptr, d.bj, row, col, slot, csr)) {
            fprintf(stderr, "bond_rows: entry outside the particles\n");
            return 5;
        }
        if (row[n] != R.nb) {
            fprintf(stderr, "bond_rows: %lld bonds, expected %lld\n", (long long)row[n], (long long)R.nb);
            return 6;
        }
    }
    double chk = 0.0;
    for (int G : {1, 2, 4}) {
        if (gsel && G != gsel) continue;
        if (D == 3) {
            if (R.form == 1) chk = G == 1 ? sweep<3, 0, 1, 1>(R, iters) : G == 2 ? sweep<3, 0, 1, 2>(R, iters) : sweep<3, 0, 1, 4>(R, iters);
            else if (R.model == 1) chk = G == 1 ? sweep<3, 1, 0, 1>(R, iters) : G == 2 ? sweep<3, 1, 0, 2>(R, iters) : sweep<3, 1, 0, 4>(R, iters);
            else if (R.model == 2) chk = G == 1 ? sweep<3, 2, 0, 1>(R, iters) : G == 2 ? sweep<3, 2, 0, 2>(R, iters) : sweep<3, 2, 0, 4>(R, iters);
            else chk = G == 1 ? sweep<3, 0, 0, 1>(R, iters) : G == 2 ? sweep<3, 0, 0, 2>(R, iters) : sweep<3, 0, 0, 4>(R, iters);
        } else {
            if (R.form == 1) chk = G == 1 ? sweep<2, 0, 1, 1>(R, iters) : G == 2 ? sweep<2, 0, 1, 2>(R, iters) : sweep<2, 0, 1, 4>(R, iters);
            else if (R.model == 1) chk = G == 1 ? sweep<2, 1, 0, 1>(R, iters) : G == 2 ? sweep<2, 1, 0, 2>(R, iters) : sweep<2, 1, 0, 4>(R, iters);
            else chk = G == 1 ? sweep<2, 0, 0, 1>(R, iters) : G == 2 ? sweep<2, 0, 0, 2>(R, iters) : sweep<2, 0, 0, 4>(R, iters);
        }
    }
    printf("ok n=%d ni=%d bonds=%lld slots=%lld check=%.6e\n", n, ni, (long long)R.nb, (long long)R.slots, chk);
    return 0;
}
