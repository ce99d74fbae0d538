// pdx_internal.h -- host context and kernel launchers of libpdx (device
// records and per-particle bodies live in pdx_device.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/pdx.h"

#include "pdx_device.h"

struct pdx_ctx {
    int32_t n = 0, dim = 3, ncase = 1;
    int32_t ni = 0, nblk = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<int64_t> case_off;     // caller rows per case (ncase+1)
    std::vector<int32_t> case_p0;      // internal start per case (ncase+1)
    std::vector<int32_t> perm;         // internal -> caller row (-1 padding)
    std::vector<int32_t> inv;          // caller row -> internal
    std::vector<uint8_t> colour;       // LDS bank colour per internal particle
    std::vector<double> hX;            // caller-order reference positions
    bool have_particles = false, have_bonds = false, have_K = false, have_mass = false;
    bool have_A = false;
    // the Gershgorin adaptive-DR mass before beta (adr_mass_b_one's raw
    // values): a function of positions, volumes, moduli, K^-1 and the intact
    // bonds, kept until one of them changes (wherever have_mass is cleared)
    double *lam_g = nullptr;
    bool have_lam_g = false;
    bool have_m = false;  // d.m64 holds the halo F-pass's sum_j wx for the current bonds
    bool layout_host = false;  // build the halo tables on the host (pdx_set_layout_builder)
    int32_t model = PDX_MODEL_LINEAR_ELASTIC, form = PDX_FORM_NOSB;
    double horizon = 0.0, G_s = 0.0;
    int64_t nb = 0, slots = 0;
    int ucur = 0;
    // hot-pass variants of the force pass (tune_g) and the F-pass (tune_f):
    // bond group 1, 2, 4, PDX_VARIANT_COOP | 1, 2, a PDX_HALO_VARIANTS code or
    // a PDX_VARIANT_LATTICE code (default: the lattice kernels, which run the
    // LDS-staged halo kernels of the same split where the context is not a
    // lattice, and those the cooperative group-2 kernels where a halo does not
    // fit)
    int tune_g = PDX_VARIANT_LATTICE | PDX_VARIANT_SPLIT4, tune_f = PDX_VARIANT_LATTICE;
    int64_t lds_dyn_max = 0;  // dynamic LDS bytes a block may use on this device
    // the lattice kernels' constants; have_lat: d.bks / d.lk hold this
    // context's lattice tables (PDX_VARIANT_LATTICE runs)
    LatC lat{};
    bool have_lat = false;
    // buffer whose d.fr records the last F-pass (a lattice F-pass followed by
    // a lattice force pass) wrote, or -1: the F-pass then wrote d.AR instead
    int fr_ubuf = -1;
    Dev d{};
    // host mirror of slice pointers
    std::vector<int64_t> h_slptr;
    // graph cache
    hipGraphExec_t gexec = nullptr;
    int64_t gkey = -1;
    // pinned host scalars
    DScal *h_sc = nullptr;
    // device address of this device's copy of the allocation-free entry
    // point's scratch array (hipGetSymbolAddress at creation; null if that failed)
    void *scratch = nullptr;
};

// error handling -------------------------------------------------------------
void pdx_set_error(const std::string &msg);

#define PDX_HIP(expr)                                                                    \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess) {                                                          \
            pdx_set_error(std::string(#expr) + ": " + hipGetErrorString(_e));            \
            return PDX_EDEVICE;                                                          \
        }                                                                                \
    } while (0)

// kernel launchers (pdx_kernels.hip) ---------------------------------------------
int pdx_launch_count(pdx_ctx *c, int32_t *d_counts);
// the row fill; d_keys: one 8-byte sort key per bond slot (c->slots)
int pdx_launch_fill(pdx_ctx *c, uint64_t *d_keys);
int pdx_launch_shape(pdx_ctx *c, int mode, double tol);
// red_ubuf >= 0: the adaptive-DR case reduction of the force pass that read buffer red_ubuf
// runs first (fused into the halo F-pass where it runs)
int pdx_launch_fpass(pdx_ctx *c, int ubuf, bool diag, bool commit, bool solve_mode, int red_ubuf = -1);
// force pass followed by the per-case reduction and update (k_case_reduce)
int pdx_launch_force(pdx_ctx *c, int ubuf, int upd, bool solve_mode, double dt, double visc);
// the two launches separately (pdx_time_kernels times the force pass alone)
int pdx_launch_force_pass(pdx_ctx *c, int ubuf, int upd, bool solve_mode, double dt, double visc);
int pdx_launch_case_reduce(pdx_ctx *c, int ubuf, int upd, bool solve_mode);
// d.m64 from the halo F-pass's own sums (sets have_m; no-op without a halo)
int pdx_build_m64(pdx_ctx *c);
int pdx_launch_kd_move(pdx_ctx *c, int ubuf_in, bool solve_mode, double dt, bool apply_reset);
int pdx_launch_halo_bits(pdx_ctx *c);
// lattice tables: the smallest squared bond length (as FP64 bits, atomicMin into
// d_out) and every row's lk / bks entries (violations ORed into d_flag)
int pdx_launch_lat_minlen(pdx_ctx *c, unsigned long long *d_out);
int pdx_launch_lat_rows(pdx_ctx *c, double h, double tol, double V, unsigned *d_flag);
// device builder of the halo tables: mode 0 counts each block's halo into d_nh
// (-1: too many neighbours for the device builder), mode 1 writes the tables
// (hl / hs at d_ptr[b], bl, bk, anchor)
int pdx_launch_halo_build(pdx_ctx *c, int mode, const uint8_t *d_colour, int32_t *d_nh, const int32_t *d_ptr);
int pdx_halo_lds(const pdx_ctx *c, int entry_bytes, int split, int nv);

// the halo-staged variants: bond group 1 or 2, rows split over 1, 2 or 4 waves
#define PDX_HALO_VARIANTS(X)                                                                   \
    X(PDX_VARIANT_HALO | 1) X(PDX_VARIANT_HALO | 2) X(PDX_VARIANT_HALO | PDX_VARIANT_SPLIT2 | 1) \
    X(PDX_VARIANT_HALO | PDX_VARIANT_SPLIT2 | 2) X(PDX_VARIANT_HALO | PDX_VARIANT_SPLIT4 | 1)     \
    X(PDX_VARIANT_HALO | PDX_VARIANT_SPLIT4 | 2) X(PDX_VARIANT_HALO | 4) X(PDX_VARIANT_HALO | PDX_VARIANT_SPLIT2 | 4) \
    X(PDX_VARIANT_HALO | PDX_VARIANT_SPLIT4 | 4)
int pdx_launch_resid(pdx_ctx *c);
int pdx_launch_zero_free(pdx_ctx *c, int ubuf);
int pdx_launch_kr_put(pdx_ctx *c, const float *d_src, int off, int cnt);
// KRec word `off` of every particle: bits in keep kept, the rest from d_src
int pdx_launch_kr_merge(pdx_ctx *c, const uint32_t *d_src, int off, uint32_t keep);
// the stabilisation modulus c of both gather buffers from d_src (internal order)
int pdx_launch_xu_c(pdx_ctx *c, const float *d_src);
int pdx_launch_xu_geom(pdx_ctx *c, const XRec *d_src);
int pdx_launch_nbr_mean(pdx_ctx *c, const double *d_in, double *d_out, int width);
// power-iteration steps of pdx_adr_mass_auto: what 0 init v, 1 perturb u, 2 apply
// K and the block partials, 3 per-case totals into out, 4 scale lam by scale[case]
int pdx_launch_pow(pdx_ctx *c, int what, int ubuf, pdx_d4 *v, const pdx_d4 *f0, const double *scale, double *out);
int pdx_launch_u_device(pdx_ctx *c, int ubuf, double *out, const double *in);
int pdx_launch_rows_internal(pdx_ctx *c, const double *d_in, double *d_out, int width);
// specimen arrays of the warm-start raw features (device, caller order; mat may be null)
struct PdxRawIn {
    const double *X, *E, *fext;
    const uint8_t *fixed;
    const int32_t *mat;
};
int pdx_launch_ctx_rows(pdx_ctx *c, double *d_X, double *d_fext, uint8_t *d_fixed);
int pdx_launch_raw_features(pdx_ctx *c, const PdxRawIn &in, const int64_t *d_chunk, int nchunk,
                            const int32_t *d_chunk_first, const int64_t *d_case_off, double dx, double *d_part,
                            double *d_cc, double *d_out);
int pdx_launch_feat_pack(pdx_ctx *c, const double *d_hv, const int32_t *d_perm, float *d_out, int width, int hops,
                         int ldo);
int pdx_launch_adr_mass(pdx_ctx *c, double beta, double *raw = nullptr);
int pdx_launch_mass_scale(pdx_ctx *c, const double *raw, double beta);
int pdx_launch_dt_est(pdx_ctx *c, double modulus, double *d_out);
int pdx_launch_damage(pdx_ctx *c, double s_c, unsigned long long *d_count);
int pdx_launch_local_damage(pdx_ctx *c);
int pdx_launch_bond_data(pdx_ctx *c, double *d_xi, double *d_len, double *d_w, const int64_t *d_rowi);
