/*
 * pdx.h -- C ABI of libpdx, the MI355X (gfx950) NOSB peridynamics solver step.
 *
 * The library owns all device memory of a context; the caller owns its host
 * arrays and data crosses only at the explicit set/get calls.  Every call
 * returns PDX_OK (0) or an error code; pdx_last_error() gives the message of
 * the last failing call on the calling thread.  One context per host thread;
 * calls are synchronous on the context's own HIP stream.
 *
 * Particle arrays are in the caller's order (n rows); vectors are n*dim
 * doubles, tensors n*dim*dim doubles row-major per particle.  Internally the
 * particles are Morton-reordered; the reordering never shows at the ABI.
 *
 * A context holds one or more cases (independent bodies that share the
 * launches).  pdx_create makes a single case; pdx_create_batch packs several
 * cases of uneven size into the same arrays, with one residual, damping and
 * convergence state per case.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository root, monostylegc/simul).
 */
#ifndef PDX_H
#define PDX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PDX_OK 0
#define PDX_EINVAL 1   /* invalid argument: maps to ValueError / FEAValidationError */
#define PDX_ESTATE 2   /* call out of order (e.g. forces before bonds): RuntimeError */
#define PDX_EDEVICE 3  /* HIP runtime error or no usable GPU */
#define PDX_ENOMEM 4   /* device allocation failed */
#define PDX_EUNSUP 5   /* unsupported combination (e.g. J2 in 2-D) */

/* constitutive models (pdx_set_materials) */
#define PDX_MODEL_LINEAR_ELASTIC 0 /* backend/fea/peridynamics/core/nosb.py:211-227 */
#define PDX_MODEL_NEO_HOOKEAN 1    /* backend/fea/fem/material/neo_hookean.py:117-127,176-180 */
#define PDX_MODEL_J2 2             /* backend/fea/fem/material/j2_plasticity.py:202-249 */

/* force formulations (pdx_set_formulation) */
#define PDX_FORM_NOSB 0 /* nosb.py:192-275 correspondence + stabilisation */
#define PDX_FORM_BOND 1 /* backend/fea/peridynamics/solver/quasi_static.py:119-143 PMB */

/* dynamic-relaxation methods (pdx_dr_params.method) */
#define PDX_DR_ADAPTIVE 0 /* adaptive DR (Kilic & Madenci 2010); no reference counterpart */
#define PDX_DR_KINETIC 1  /* velocity Verlet + kinetic damping, nosb_solver.py:156-332 */

/* singular shape-tensor test (pdx_shape_tensor) */
#define PDX_SING_ABSOLUTE 0 /* |det K| > tol, the reference test nosb.py:89-94 */
#define PDX_SING_RELATIVE 1 /* |det K| > tol * (tr K / dim)^dim, unit-independent */

/* field ids for pdx_get / pdx_set (f64 unless noted) */
enum {
    PDX_FIELD_X = 0,      /* n*dim reference positions (get)                particles.py:37 */
    PDX_FIELD_x = 1,      /* n*dim current positions X+u (get/set)         particles.py:38 */
    PDX_FIELD_u = 2,      /* n*dim displacement (get/set)                   particles.py:225-227 */
    PDX_FIELD_v = 3,      /* n*dim velocity (get/set)                       particles.py:39 */
    PDX_FIELD_a = 4,      /* n*dim acceleration (kinetic DR) or previous force (adaptive DR) */
    PDX_FIELD_f = 5,      /* n*dim force density of the last force evaluation particles.py:41 */
    PDX_FIELD_VOLUME = 6, /* n (get/set)                                    particles.py:44 */
    PDX_FIELD_MASS = 7,   /* n (get/set)                                    particles.py:46 */
    PDX_FIELD_BULK = 8,   /* n (get/set)                                    particles.py:52 */
    PDX_FIELD_SHEAR = 9,  /* n (get/set)                                    particles.py:53 */
    PDX_FIELD_CSTAB = 10, /* n stabilisation micromodulus c_i (get/set)     nosb.py:335-346 */
    PDX_FIELD_K = 11,     /* n*dim*dim shape tensor (get)                   particles.py:56 */
    PDX_FIELD_KINV = 12,  /* n*dim*dim inverse (get)                        particles.py:57 */
    PDX_FIELD_F = 13,     /* n*dim*dim deformation gradient (get)           particles.py:58 */
    PDX_FIELD_P = 14,     /* n*dim*dim first Piola-Kirchhoff stress (get)   particles.py:59 */
    PDX_FIELD_DAMAGE = 15,/* n (get/set)                                    particles.py:49 */
    PDX_FIELD_FIXED = 16, /* n uint8 DOF bitmask (get/set)                  particles.py:62 */
    PDX_FIELD_FEXT = 17,  /* n*dim external force density (get/set)         quasi_static.py:344-377 */
    PDX_FIELD_EP = 18,    /* n*9 committed plastic strain (get/set)         j2_plasticity.py:117 */
    PDX_FIELD_EPBAR = 19, /* n committed equivalent plastic strain (get/set) j2_plasticity.py:118 */
    PDX_FIELD_BACK = 20,  /* n*9 committed back stress (get/set)            extension (0 in the reference) */
    PDX_FIELD_NNB = 21,   /* n int32 bonds per particle (get)               bonds.py:72 */
    PDX_FIELD_SING = 22,  /* n uint8 singular-K flag (get)                  nosb.py:92-94 */
    PDX_FIELD_DRMASS = 23,/* n adaptive-DR diagonal mass (get/set) */
    PDX_FIELD_DENSITY = 24,/* n (get/set)                                   particles.py:45 */
    PDX_FIELD_COUNT_
};

typedef struct pdx_ctx pdx_ctx;

typedef struct {
    int32_t method;       /* PDX_DR_ADAPTIVE or PDX_DR_KINETIC */
    int32_t use_graph;    /* 1: replay iteration chunks as a captured hipGraph */
    int32_t chunk;        /* iterations per host convergence check (even; 0 = default) */
    int32_t resume;       /* 1: continue from the current velocity / damping state */
    int64_t max_iterations;
    double tol;           /* stop when residual / ref < tol (strict, nosb_solver.py:318) */
    double ref;           /* > 0: fixed reference residual; <= 0: residual of the start state (nosb_solver.py:288-293) */
    double dt;            /* kinetic DR time step (nosb_solver.py:81-83) */
    double viscous_damping;/* kinetic DR viscous factor (nosb_solver.py:168-174) */
    double cmax;          /* adaptive DR damping cap (0 = default 1.9) */
    const double *refs;   /* per-case reference residuals (ncase values) or NULL for `ref` */
} pdx_dr_params;

typedef struct {
    int32_t converged;
    int32_t reserved;
    int64_t iterations;
    double residual;
    double relative_residual;
    double reference_residual;
    double damping;        /* last adaptive damping coefficient */
    int64_t resets;        /* kinetic DR velocity resets */
    double elapsed_s;      /* wall time of the iteration loop */
} pdx_dr_result;

/* ---- lifetime -------------------------------------------------------- */
/* ParticleSystem(n, dim), particles.py:26-62 */
int pdx_create(int32_t n, int32_t dim, pdx_ctx **out);
/* several independent bodies in one context: case c owns caller rows
 * [offsets[c], offsets[c+1]); offsets has ncase+1 entries, offsets[0] = 0 */
int pdx_create_batch(int32_t ncase, const int64_t *offsets, int32_t dim, pdx_ctx **out);
int pdx_destroy(pdx_ctx *ctx);
const char *pdx_last_error(void);
int pdx_num_cases(pdx_ctx *ctx, int32_t *ncase);

/* ---- inputs ------------------------------------------------------------ */
/* ParticleSystem.initialize_from_arrays / _init_fields, particles.py:121-173.
 * X n*dim, vol n, mass n (mass may be NULL: then mass = vol) */
int pdx_set_particles(pdx_ctx *ctx, const double *X, const double *vol, const double *mass);
/* per-particle constants: set_material_constants_per_particle, particles.py:211-219;
 * yield_stress/h_iso/h_kin only for PDX_MODEL_J2 (may be NULL otherwise) */
int pdx_set_materials(pdx_ctx *ctx, int32_t model, const double *bulk, const double *shear,
                      const double *yield_stress, const double *h_iso, const double *h_kin);
/* NOSBCompute(stabilization) + bond constant, nosb.py:35-57, 193-209; c per particle
 * (bond modulus (c_i+c_j)/2), or c == NULL with c_scalar for the reference scalar */
int pdx_set_stabilization(pdx_ctx *ctx, double G_s, const double *c, double c_scalar);
int pdx_set_formulation(pdx_ctx *ctx, int32_t formulation);
/* fixed DOF bitmask per particle (set_fixed_particles, particles.py:201-209) and
 * external force density (LoadControl, quasi_static.py:344-377); fext may be NULL */
int pdx_set_bc(pdx_ctx *ctx, const uint8_t *fixed, const double *fext);

/* ---- bonds --------------------------------------------------------------- */
/* NeighborSearch.count_neighbors_only, neighbor.py:239-319 (counts in caller order) */
int pdx_count_neighbors(pdx_ctx *ctx, double horizon, int32_t *counts);
/* NeighborSearch.build + BondSystem.build_from_neighbor_search,
 * neighbor.py:208-218, bonds.py:82-137.  Neighbour sets {j != i : |X_j-X_i|^2 < h^2}. */
int pdx_build_bonds(pdx_ctx *ctx, double horizon, int64_t *n_bonds);
/* bonds in caller order: row n+1, col/broken n_bonds, each row sorted by j
 * (BondSystem.neighbors / broken, bonds.py:68-77); any pointer may be NULL */
int pdx_get_bonds(pdx_ctx *ctx, int64_t *row, int32_t *col, uint8_t *broken);
/* per-bond reference data xi (nb*dim), |xi| (nb), omega (nb) in pdx_get_bonds order,
 * bonds.py:116-137; any pointer may be NULL */
int pdx_get_bond_data(pdx_ctx *ctx, double *xi, double *xi_len, double *omega);
/* overwrite broken flags (nb, pdx_get_bonds order), BondSystem.broken, bonds.py:76 */
int pdx_set_broken(pdx_ctx *ctx, const uint8_t *broken);

/* bond layout builder (no reference counterpart; the reference builds its
 * neighbour lists on the device, neighbor.py:119-319, bonds.py:101-137): the halo
 * tables of the LDS-staged hot passes are built on the device (PDX_LAYOUT_DEVICE,
 * the default) or by host threads (PDX_LAYOUT_HOST, the same tables; also the
 * fallback for blocks with more than 6 144 distinct out-of-block neighbours).
 * Takes effect at the next pdx_build_bonds. */
#define PDX_LAYOUT_DEVICE 0
#define PDX_LAYOUT_HOST 1
int pdx_set_layout_builder(pdx_ctx *ctx, int32_t where);
/* the halo tables (internal order; diagnostic, no reference counterpart).
 * sizes[5] = blocks, halo entries, bl words, bk words (bl words + tail), LDS slot
 * cap; all 0 when no tables are built.  hl_ptr (blocks + 1), hl and hs (entries),
 * bl, bk, anchor (3 per block) may each be NULL. */
int pdx_get_halo_tables(pdx_ctx *ctx, int64_t *sizes, int32_t *hl_ptr, int32_t *hl, uint16_t *hs,
                        uint16_t *bl, uint16_t *bk, double *anchor);

/* ---- kernels ------------------------------------------------------------- */
/* NOSBCompute.compute_shape_tensor, nosb.py:59-94 (FP64 K and inverse) */
int pdx_shape_tensor(pdx_ctx *ctx, int32_t singular_mode, double tol);
/* compute_deformation_gradient + stress loop of compute_force_state_with_stabilization,
 * nosb.py:96-126, 211-227; keeps F and P for pdx_get */
int pdx_compute_F_stress(pdx_ctx *ctx);
/* NOSBSolver.compute_forces, nosb_solver.py:200-216 (+ LoadControl.apply) */
int pdx_compute_forces(pdx_ctx *ctx);
/* _compute_residual_norm, nosb_solver.py:190-198, per case (ncase values) */
int pdx_residual_norm(pdx_ctx *ctx, double *out);
/* residual at the cold-start state (free DOFs at u = 0, prescribed DOFs as set),
 * per case; the state is left unchanged */
int pdx_cold_residual(pdx_ctx *ctx, double *out);
/* No counterpart in the reference: pdx_cold_residual for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same results, bit for bit, without any allocation: the records it
 * overwrites are kept in a library-owned device array of PDX_NOALLOC_MAX_PARTICLES
 * records (32 MiB per device, held until the process ends).  Contexts with more
 * internal particles (cases padded to multiples of 256) get PDX_EINVAL.  All calls on
 * one device share that array: calls that overlap on one device race. */
#define PDX_NOALLOC_MAX_PARTICLES 524288
int pdx_cold_residual_noalloc(pdx_ctx *ctx, double *out);
/* DamageModel.update_damage + compute_local_damage, damage.py:94-152 (FP64 stretch) */
int pdx_break_bonds(pdx_ctx *ctx, double critical_stretch, int64_t *n_broken_now);
int pdx_local_damage(pdx_ctx *ctx);
/* J2: commit the trial plastic state of the current displacement (once per
 * converged load increment; the reference commits per call, j2_plasticity.py:242-243) */
int pdx_commit_state(pdx_ctx *ctx);

/* weighted neighbour mean over the intact bonds of a per-particle feature vector
 * (in/out n*width doubles, width 1..64): out_i = sum_j w_ij V_j in_j / sum_j w_ij V_j.
 * No reference counterpart: the message-passing step of the learned warm start. */
int pdx_neighbor_mean(pdx_ctx *ctx, int32_t width, const double *in, double *out);
/* feature input of the learned warm start (no reference counterpart): `hops` successive
 * neighbour means of `in` (n x width, caller order, host) computed on the device, and
 * [in | hop 1 | ... | hop `hops`] written as FP32 into the DEVICE array d_out (n rows of
 * ldo floats, caller order; columns from width * (hops + 1) on are not written).  width
 * 1..64, hops 0..4, ldo >= width * (hops + 1).  One host-to-device copy in all. */
int pdx_neighbor_features(pdx_ctx *ctx, int32_t width, int32_t hops, const double *in, float *d_out, int32_t ldo);
/* the same with the input rows already in device memory (d_in, n x width, caller order;
 * no reference counterpart): no host copy at all */
int pdx_neighbor_features_device(pdx_ctx *ctx, int32_t width, int32_t hops, const double *d_in, float *d_out,
                                 int32_t ldo);
/* raw inputs of the learned warm start (no reference counterpart): 20 FP64 columns per
 * particle (n x 20, caller order) into the DEVICE array d_out, from the host arrays of the
 * specimen: X (n x 3), E (n), fext (n x 3 force density), fixed (n DOF masks), material_id
 * (n, 0 = cortical; NULL: none cortical) and the particle spacing dx.  X, fext or fixed
 * NULL: the context's own positions, external forces or DOF masks (as pdx_set_particles and
 * pdx_set_bc stored them), gathered on the device instead of copied from the host.
 * Per-case box, load resultant and modulus statistics over the context's cases.  3-D
 * contexts only. */
int pdx_warm_raw_features(pdx_ctx *ctx, const double *X, const double *E, const double *fext, const uint8_t *fixed,
                          const int32_t *material_id, double dx, double *d_out);

/* ---- dynamic relaxation ---------------------------------------------------- */
/* NOSBSolver._estimate_stable_dt, nosb_solver.py:94-154 */
int pdx_estimate_dt(pdx_ctx *ctx, double bulk, double shear, double rho_mean, double *dt);
/* adaptive-DR diagonal mass (Gershgorin bound of the force stiffness) x beta */
int pdx_adr_mass(pdx_ctx *ctx, double beta);
/* adaptive-DR mass from the spectrum (no reference counterpart): the Gershgorin mass of
 * pdx_adr_mass(ctx, 1) scaled per case by beta = safety * rho / 4 (at most 1), rho the
 * largest eigenvalue of M^-1 K estimated by `iters` power iterations (force differences
 * at a perturbation of 1e-4 horizons around the current u; 2 force passes each).  safety
 * is the margin to the explicit update's stability limit (the Python layer uses 1.4); DR
 * needs about sqrt(beta) of the Gershgorin iterations.  beta_out: ncase values or NULL.
 * The solver state is unchanged except the second displacement buffer and the forces. */
int pdx_adr_mass_auto(pdx_ctx *ctx, int32_t iters, double safety, double *beta_out);
/* NOSBSolver.step, nosb_solver.py:218-256: one kinetic-damping step with the
 * reset decision taken on the host state passed in/out (prev_ke, ke_increasing) */
int pdx_kd_step(pdx_ctx *ctx, double dt, double viscous_damping, double *prev_ke,
                int32_t *ke_increasing, double *ke, double *residual, int32_t *reset);
/* NOSBSolver.solve, nosb_solver.py:262-332, iterations on the device; results per case
 * (res has ncase entries) */
int pdx_dr_solve(pdx_ctx *ctx, const pdx_dr_params *params, pdx_dr_result *res);

/* ---- explicit dynamics ------------------------------------------------------------ */
/* ExplicitSolver.run, backend/fea/peridynamics/solver/explicit.py:62-158, with
 * DamageModel.step (core/damage.py:144-152) after every step when critical_stretch > 0:
 * n_steps of velocity-Verlet dynamics from the context's current u, v and a
 * (PDX_FIELD_a): x += v dt + a dt^2 / 2, forces, a = f / m, v = (1 - damping)(v +
 * (a_old + a) dt / 2) on free DOFs (fixed DOFs keep their u with v = a = 0), then every
 * intact bond with stretch > critical_stretch breaks.  The force is the context's
 * formulation: PDX_FORM_BOND is the reference's explicit PMB force c s w V_j eta/|eta|
 * with c the per-bond stabilisation modulus (pdx_set_stabilization, G_s = 1), plus the
 * external force density.  The shape tensor is not recomputed during the run.  Local
 * damage is current at return.  n_broken: bonds broken during the run (may be NULL);
 * ke: per case, sum m v^2 / 2 after the last step (explicit.py:226-233; may be NULL). */
int pdx_explicit_run(pdx_ctx *ctx, double dt, double damping, int64_t n_steps, double critical_stretch,
                     int64_t *n_broken, double *ke);

/* ---- field access -------------------------------------------------------------- */
/* field.to_numpy() / field.from_numpy() of the reference's Taichi fields
 * (particles.py:37-62, 221-231); layout per field id above, caller order */
int pdx_get(pdx_ctx *ctx, int32_t field, void *out);
int pdx_set(pdx_ctx *ctx, int32_t field, const void *in);
/* the displacement (PDX_FIELD_u) between the context and caller-order device memory
 * (n*dim doubles on the context's device); no reference counterpart: for callers that
 * keep displacement fields on the GPU (load-stepping drivers, warm starts), where the
 * host round trip of pdx_get/pdx_set costs more than a DR increment.  Both return after
 * the copy has completed; d_in must be complete when pdx_set_u_device is called. */
int pdx_get_u_device(pdx_ctx *ctx, double *d_out);
int pdx_set_u_device(pdx_ctx *ctx, const double *d_in);

/* ---- measurement ------------------------------------------------------------ */
/* times `iters` adaptive-DR iterations with HIP events around every launch on the
 * context stream; ms[0] = mean F-pass, ms[1] = mean force pass (the force kernel; the
 * per-case reduction that follows it is a separate small launch), ms[2] = total */
int pdx_time_kernels(pdx_ctx *ctx, int32_t iters, double *ms);
/* hot-pass variants: per-lane neighbour gathers with bond groups 1, 2, 4, or
 * wave-cooperative gathers (four lanes fetch one 64-byte neighbour record, staged
 * through LDS) with bond groups 1, 2 as PDX_VARIANT_COOP | group.  These variants
 * give the same bits.  PDX_VARIANT_HALO | group (1, 2, 4): each block stages the
 * records of every particle its rows reach into LDS once, positions relative to a
 * per-block anchor in FP32, and reads its bonds from LDS; the NOSB force pass sums each
 * group's terms in FP32 before adding them to its FP64 sum.  Results agree with the
 * others to FP32 rounding (about 1e-7 relative).  A pass whose halo does not fit the
 * device's LDS runs the PDX_VARIANT_COOP | 2 kernel instead. */
#define PDX_VARIANT_COOP 16
#define PDX_VARIANT_HALO 32
/* with PDX_VARIANT_HALO: each row's bond slots split over 2 or 4 waves that share the
 * block's halo (partial sums added in a fixed order) */
#define PDX_VARIANT_SPLIT2 64
#define PDX_VARIANT_SPLIT4 128
/* PDX_VARIANT_LATTICE, optionally | PDX_VARIANT_SPLIT2 or PDX_VARIANT_SPLIT4 (no bond
 * group): the halo staging of PDX_VARIANT_HALO, with each row walking the 122 offsets o
 * of the lattice stencil (0 < |o|^2 <= 9) instead of its bond list, the bond vector
 * xi = h o, |xi| and the influence weight being constants of the offset.  Available in
 * 3-D NOSB contexts whose every bond is h o for one lattice spacing h (to 1e-8 h) and
 * whose particles share one volume (voxel specimens at the reference's horizon
 * 3.015 h); elsewhere, and for the diagnostic and J2 commit passes, the pass runs
 * PDX_VARIANT_HALO with the same split.  Results agree with the halo variants to
 * FP32 rounding.  pdx_lattice_spacing reports whether a context qualifies. */
#define PDX_VARIANT_LATTICE 256
/* h of the lattice kernels (0 when the context does not qualify for them); no
 * reference counterpart */
int pdx_lattice_spacing(pdx_ctx *ctx, double *h);
/* launch tuning (no reference counterpart): times `iters` adaptive-DR iterations of
 * every hot-pass variant on this context's own data, keeps the fastest for each pass
 * and leaves the solver state as it was; group_out = the force pass's variant */
int pdx_autotune(pdx_ctx *ctx, int32_t iters, int32_t *group_out);
/* sets the hot-pass variant (1, 2, 4, PDX_VARIANT_COOP | 1 or 2, PDX_VARIANT_HALO | 1, 2 or 4
 * optionally | PDX_VARIANT_SPLIT2 or PDX_VARIANT_SPLIT4; the halo variants need the bonds
 * built) */
int pdx_set_group(pdx_ctx *ctx, int32_t group);
/* launch tuning (no reference counterpart): the F-pass and the force pass are separate
 * launches and may use different variants; pdx_autotune keeps the fastest of each and
 * reports the force pass's in group_out */
int pdx_set_variants(pdx_ctx *ctx, int32_t fpass_variant, int32_t force_variant);
int pdx_get_variants(pdx_ctx *ctx, int32_t *fpass_variant, int32_t *force_variant);
/* n_particles, n_bonds (intact+broken), n_intact, sell_slots (padded bond slots) */
int pdx_stats(pdx_ctx *ctx, int64_t *out4);

#ifdef __cplusplus
}
#endif
#endif /* PDX_H */
