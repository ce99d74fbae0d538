// pdx_internal.h -- device records, context and launch helpers of libpdx.
//
// Particle data lives on the device in Morton order, one case after another,
// each case padded to a multiple of PDX_BLOCK particles so that every block of
// every launch belongs to exactly one case.  Padding particles carry no bonds
// and are flagged inactive.
//
// Bonds are stored SELL-64 (sliced ELL, slice = 64 consecutive particles =
// one wavefront): entry k of particle i sits at sl_ptr[i/64] + 64*k + i%64,
// so the k-th bond of the 64 lanes of a wave is one coalesced 256-byte read.
// An entry is the neighbour's internal index with bit 31 = broken.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/pdx.h"

#define PDX_BLOCK 256
#define PDX_SLICE 64
#define PDX_BROKEN 0x80000000u
#define PDX_NXCD 8

// flags word of KRec
#define PDX_FL_FIXMASK 0x7u
#define PDX_FL_SING 0x8u
#define PDX_FL_LOADED 0x10u
#define PDX_FL_ACTIVE 0x20u

// gather record of the reference geometry: 32 bytes
struct __align__(32) PRec {
    double X[3];
    float V;  // particle volume
    float c;  // stabilisation micromodulus c_i
};

// per-particle constants of the hot loop: 64 bytes
struct __align__(16) KRec {
    float Kinv[9];  // K^-1 in FP32 (K itself when singular)
    float m[3];     // sum_j w_ij V_j xi_ij over intact bonds
    float bulk, shear;
    uint32_t flags;
    uint32_t nnb;   // bonds in the row (intact + broken)
};

// correspondence tensor A = P K^-1 of the current iterate: 48 bytes
struct __align__(16) ARec {
    float A[9];
    float pad[3];
};

struct __align__(16) J2Par {
    float sy0, hiso, hkin, pad;
};

// committed J2 state (symmetric tensors as xx,yy,zz,xy,yz,xz): 64 bytes
struct __align__(16) J2State {
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
    double num, den;
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
    PRec *P;
    KRec *KR;
    ARec *AR;
    J2Par *j2p;
    J2State *j2s;
    double4 *u[2];
    double4 *v;
    double4 *fp;       // previous force (adaptive DR) / acceleration (kinetic DR)
    double4 *f;        // force output of force-only passes
    double4 *fext;
    double *mass;      // physical mass rho V (kinetic DR)
    double *lam;       // adaptive-DR diagonal mass
    double *V64;       // particle volume in FP64 (setup kernels)
    double *K64, *Kinv64;
    float *Fd, *Pd;    // diagnostic F and P (9 per particle)
    double *damage;
    int32_t *init_nb;  // bonds at build time
    int64_t *sl_ptr;
    uint32_t *bj;
    int32_t *case_of_blk;
    int32_t *case_blk0;  // first block of each case (ncase+1)
    DScal *sc;
    double *part;      // per-block partial sums (4 per block)
    // neighbour search
    GridP *grid;
    int32_t *case_of_p;  // case of each internal particle
    int32_t *cell_start; // per cell start into cell_idx (+1 sentinel per case grid)
    int32_t *cell_idx;
    int32_t *counts;
    double h, inv_h, h2;  // horizon, 1/horizon, horizon^2
    double G_s;
    int32_t model, form;
};

struct pdx_ctx {
    int32_t n = 0, dim = 3, ncase = 1;
    int32_t ni = 0, nblk = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<int64_t> case_off;     // caller rows per case (ncase+1)
    std::vector<int32_t> case_p0;      // internal start per case (ncase+1)
    std::vector<int32_t> perm;         // internal -> caller row (-1 padding)
    std::vector<int32_t> inv;          // caller row -> internal
    std::vector<double> hX;            // caller-order reference positions
    bool have_particles = false, have_bonds = false, have_K = false, have_mass = false;
    bool have_A = false;
    int32_t model = PDX_MODEL_LINEAR_ELASTIC, form = PDX_FORM_NOSB;
    double horizon = 0.0, G_s = 0.0;
    int64_t nb = 0, slots = 0;
    int ucur = 0;
    Dev d{};
    // host mirror of slice pointers
    std::vector<int64_t> h_slptr;
    // graph cache
    hipGraphExec_t gexec = nullptr;
    int64_t gkey = -1;
    // pinned host scalars
    DScal *h_sc = nullptr;
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
int pdx_launch_fill(pdx_ctx *c);
int pdx_launch_shape(pdx_ctx *c, int mode, double tol);
int pdx_launch_fpass(pdx_ctx *c, int ubuf, bool diag, bool commit, bool solve_mode);
int pdx_launch_force(pdx_ctx *c, int ubuf, int upd, bool solve_mode, double dt, double visc);
int pdx_launch_kd_move(pdx_ctx *c, int ubuf_in, bool solve_mode, double dt, bool apply_reset);
int pdx_launch_resid(pdx_ctx *c);
int pdx_launch_zero_free(pdx_ctx *c, int ubuf);
int pdx_launch_kr_put(pdx_ctx *c, const float *d_src, int off, int cnt);
int pdx_launch_prec_c(pdx_ctx *c, const float *d_c);
int pdx_launch_adr_mass(pdx_ctx *c, double beta);
int pdx_launch_dt_est(pdx_ctx *c, double modulus, double *d_out);
int pdx_launch_damage(pdx_ctx *c, double s_c, unsigned long long *d_count);
int pdx_launch_local_damage(pdx_ctx *c);
int pdx_launch_bond_data(pdx_ctx *c, double *d_xi, double *d_len, double *d_w, const int64_t *d_rowi);
