// pdx_kernels.hip -- gfx950 kernels of the NOSB peridynamics solver step.
//
// Every kernel is a thin wrapper over a per-particle function of pdx_device.h
// (one thread per particle, wave64).  Hot loop per dynamic-relaxation
// iteration:
//   k_fpass  : H = [sum_j w V_j du (x) xi] K^-1 (FP32 terms, FP64 sums), P(H)
//              (LE/NH in FP32, J2 radial return in FP64), A = P K^-1
//   k_force  : f_i = A_i m_i + sum_j w V_j [A_j xi + G_s c_ij s eta/|eta|]
//              (FP32 bond terms, FP64 per-particle sums), + external load,
//              fused DR update, FP64 block partials and the case's last-block
//              reduction of residual / damping / kinetic energy
//   k_kd_move: kinetic-damping position update (reference-faithful mode)
// Setup kernels (FP64): neighbour count/fill, shape tensor and inverse,
// adaptive-DR mass, dt estimate, damage.
#include "pdx_internal.h"

// blocks are dispatched round-robin over the 8 XCDs; give every XCD one
// contiguous range of Morton blocks so the neighbour halo of a block mostly
// hits the same L2
__device__ __forceinline__ int logical_block()
{
    const int b = blockIdx.x, nb = gridDim.x;
    return (b % PDX_NXCD) * (nb / PDX_NXCD) + b / PDX_NXCD;
}

__device__ __forceinline__ double wave_sum(double x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

__device__ __forceinline__ double ld_coherent(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block sum of NV doubles into partials[lb]; returns true in the block that
// arrives last for its case, with tot[] = case totals summed in block order
// (a fixed order: the result does not depend on block scheduling).
template <int NV>
__device__ bool reduce_case(const Dev &d, int lb, int cs, double (&val)[NV], double (&tot)[NV])
{
    __shared__ double sred[PDX_BLOCK / 64][NV];
    __shared__ int s_last;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double t = wave_sum(val[q]);
        if (lane == 0) sred[wid][q] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double t = 0.0;
            for (int w = 0; w < PDX_BLOCK / 64; ++w) t += sred[w][q];
            d.part[(int64_t)lb * 4 + q] = t;
        }
        __threadfence();
        const int nbc = d.case_blk0[cs + 1] - d.case_blk0[cs];
        unsigned prev = atomicAdd(&d.sc[cs].counter, 1u);
        s_last = (prev == (unsigned)(nbc - 1));
    }
    __syncthreads();
    if (!s_last) return false;
    __threadfence();
    const int b0 = d.case_blk0[cs], b1 = d.case_blk0[cs + 1];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double t = 0.0;
        for (int b = b0 + threadIdx.x; b < b1; b += PDX_BLOCK) t += ld_coherent(&d.part[(int64_t)b * 4 + q]);
        t = wave_sum(t);
        __syncthreads();
        if (lane == 0) sred[wid][q] = t;
        __syncthreads();
        double s = 0.0;
        for (int w = 0; w < PDX_BLOCK / 64; ++w) s += sred[w][q];
        tot[q] = s;
    }
    if (threadIdx.x == 0) d.sc[cs].counter = 0u;
    return true;
}

// ---------------------------------------------------------------------------
// hot kernels
// ---------------------------------------------------------------------------
template <int D, int MODEL, bool DIAG, bool COMMIT, int G>
__global__ __launch_bounds__(PDX_BLOCK) void k_fpass(Dev d, int ubuf, int solve_mode)
{
    const int lb = logical_block();
    if (solve_mode && d.sc[d.case_of_blk[lb]].converged) return;
    const int i = lb * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    fpass_one<D, MODEL, DIAG, COMMIT, G>(d, i, ubuf);
}

template <int D, int FORM, int UPD, int G>
__global__ __launch_bounds__(PDX_BLOCK) void k_force(Dev d, int ubuf, int solve_mode, double dt, double visc)
{
    const int lb = logical_block();
    const int cs = d.case_of_blk[lb];
    if (solve_mode && d.sc[cs].converged) return;
    const int i = lb * PDX_BLOCK + threadIdx.x;
    double red[3] = {0.0, 0.0, 0.0};
    if (i < d.ni) force_one<D, FORM, UPD, G>(d, i, ubuf, cs, dt, visc, red);
    double tot[3];
    if (!reduce_case<3>(d, lb, cs, red, tot)) return;
    if (threadIdx.x == 0) case_update<UPD>(d.sc[cs], tot, ubuf);
}

template <int D>
__global__ __launch_bounds__(PDX_BLOCK) void k_kd_move(Dev d, int ubuf_in, int solve_mode, double dt, int apply_reset)
{
    const int lb = logical_block();
    const int cs = d.case_of_blk[lb];
    if (solve_mode && d.sc[cs].converged) return;
    const int i = lb * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    kd_move_one<D>(d, i, ubuf_in, dt, apply_reset && d.sc[cs].reset_pending);
}

template <int D>
__global__ __launch_bounds__(PDX_BLOCK) void k_resid(Dev d)
{
    const int lb = logical_block();
    const int cs = d.case_of_blk[lb];
    const int i = lb * PDX_BLOCK + threadIdx.x;
    double red[1] = {i < d.ni ? resid_one<D>(d, i) : 0.0};
    double tot[1];
    if (!reduce_case<1>(d, lb, cs, red, tot)) return;
    if (threadIdx.x == 0) d.sc[cs].res = sqrt(tot[0]);
}

// ---------------------------------------------------------------------------
// setup kernels
// ---------------------------------------------------------------------------
#define PDX_PER_PARTICLE(name, call)                                   \
    __global__ __launch_bounds__(PDX_BLOCK) void name                  \
    {                                                                  \
        const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;            \
        if (i >= d.ni) return;                                         \
        call;                                                          \
    }

template <int D, bool FILL> PDX_PER_PARTICLE(k_neighbors(Dev d), (neighbors_one<D, FILL>(d, i)))
template <int D> PDX_PER_PARTICLE(k_shape(Dev d, int mode, double tol), (shape_one<D>(d, i, mode, tol)))
template <int D> PDX_PER_PARTICLE(k_adr_mass_a(Dev d, double *cs, double *g), (adr_mass_a_one<D>(d, i, cs, g)))
template <int D> PDX_PER_PARTICLE(k_adr_mass_b(Dev d, const double *cs, const double *g, double beta),
                                  (adr_mass_b_one<D>(d, i, cs, g, beta)))
template <int D> PDX_PER_PARTICLE(k_dt_est(Dev d, double modulus, double *out), (dt_est_one<D>(d, i, modulus, out)))
template <int D> PDX_PER_PARTICLE(k_bond_data(Dev d, double *xi, double *len, double *w, const int64_t *rowi),
                                  (bond_data_one<D>(d, i, xi, len, w, rowi)))
PDX_PER_PARTICLE(k_local_damage(Dev d), (local_damage_one(d, i)))
PDX_PER_PARTICLE(k_zero_free(Dev d, int ubuf), (zero_free_one(d, i, ubuf)))

template <int D>
__global__ __launch_bounds__(PDX_BLOCK) void k_damage(Dev d, int ubuf, double s_c, unsigned long long *count)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const unsigned nb = damage_one<D>(d, i, ubuf, s_c);
    if (nb) atomicAdd(count, (unsigned long long)nb);
}

// scatter cnt 32-bit words per particle into KRec word offset off
__global__ __launch_bounds__(PDX_BLOCK) void k_kr_put(Dev d, const float *src, int off, int cnt)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    float *w = reinterpret_cast<float *>(&d.KR[i]);
    for (int q = 0; q < cnt; ++q) w[off + q] = src[(int64_t)i * cnt + q];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline int grid1(int n) { return (n + PDX_BLOCK - 1) / PDX_BLOCK; }

#define PDX_LAUNCH_CHECK()                                                   \
    do {                                                                     \
        hipError_t _e = hipGetLastError();                                   \
        if (_e != hipSuccess) {                                              \
            pdx_set_error(std::string("kernel launch: ") + hipGetErrorString(_e)); \
            return PDX_EDEVICE;                                              \
        }                                                                    \
    } while (0)

int pdx_launch_count(pdx_ctx *c, int32_t *d_counts)
{
    Dev d = c->d;
    d.counts = d_counts;
    if (c->dim == 3)
        hipLaunchKernelGGL((k_neighbors<3, false>), dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, d);
    else
        hipLaunchKernelGGL((k_neighbors<2, false>), dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, d);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

int pdx_launch_fill(pdx_ctx *c)
{
    if (c->dim == 3)
        hipLaunchKernelGGL((k_neighbors<3, true>), dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d);
    else
        hipLaunchKernelGGL((k_neighbors<2, true>), dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

int pdx_launch_shape(pdx_ctx *c, int mode, double tol)
{
    if (c->dim == 3)
        hipLaunchKernelGGL((k_shape<3>), dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d, mode, tol);
    else
        hipLaunchKernelGGL((k_shape<2>), dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d, mode, tol);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

// the hot variants come in bond-group sizes 1, 2, 4 (chosen per context by
// pdx_autotune); diagnostic and commit passes use group 1
template <int D, int MODEL, int G>
static void launch_fpass_g(pdx_ctx *c, int ubuf, bool diag, bool commit, int sm)
{
    dim3 g(c->nblk), b(PDX_BLOCK);
    if (commit)
        hipLaunchKernelGGL((k_fpass<D, MODEL, false, true, 1>), g, b, 0, c->stream, c->d, ubuf, sm);
    else if (diag)
        hipLaunchKernelGGL((k_fpass<D, MODEL, true, false, 1>), g, b, 0, c->stream, c->d, ubuf, sm);
    else
        hipLaunchKernelGGL((k_fpass<D, MODEL, false, false, G>), g, b, 0, c->stream, c->d, ubuf, sm);
}

template <int D, int MODEL>
static void launch_fpass_t(pdx_ctx *c, int ubuf, bool diag, bool commit, int sm)
{
    switch (c->tune_g) {
    case 4: launch_fpass_g<D, MODEL, 4>(c, ubuf, diag, commit, sm); break;
    case 2: launch_fpass_g<D, MODEL, 2>(c, ubuf, diag, commit, sm); break;
    default: launch_fpass_g<D, MODEL, 1>(c, ubuf, diag, commit, sm); break;
    }
}

int pdx_launch_fpass(pdx_ctx *c, int ubuf, bool diag, bool commit, bool solve_mode)
{
    const int sm = solve_mode ? 1 : 0;
    if (c->dim == 3) {
        switch (c->model) {
        case PDX_MODEL_NEO_HOOKEAN: launch_fpass_t<3, PDX_MODEL_NEO_HOOKEAN>(c, ubuf, diag, commit, sm); break;
        case PDX_MODEL_J2: launch_fpass_t<3, PDX_MODEL_J2>(c, ubuf, diag, commit, sm); break;
        default: launch_fpass_t<3, PDX_MODEL_LINEAR_ELASTIC>(c, ubuf, diag, commit, sm); break;
        }
    } else {
        switch (c->model) {
        case PDX_MODEL_NEO_HOOKEAN: launch_fpass_t<2, PDX_MODEL_NEO_HOOKEAN>(c, ubuf, diag, commit, sm); break;
        default: launch_fpass_t<2, PDX_MODEL_LINEAR_ELASTIC>(c, ubuf, diag, commit, sm); break;
        }
    }
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

template <int D, int FORM, int G>
static void launch_force_g(pdx_ctx *c, int ubuf, int upd, int sm, double dt, double visc)
{
    dim3 g(c->nblk), b(PDX_BLOCK);
    if (upd == 1)
        hipLaunchKernelGGL((k_force<D, FORM, 1, G>), g, b, 0, c->stream, c->d, ubuf, sm, dt, visc);
    else if (upd == 2)
        hipLaunchKernelGGL((k_force<D, FORM, 2, G>), g, b, 0, c->stream, c->d, ubuf, sm, dt, visc);
    else
        hipLaunchKernelGGL((k_force<D, FORM, 0, 1>), g, b, 0, c->stream, c->d, ubuf, sm, dt, visc);
}

template <int D, int FORM>
static void launch_force_t(pdx_ctx *c, int ubuf, int upd, int sm, double dt, double visc)
{
    switch (c->tune_g) {
    case 4: launch_force_g<D, FORM, 4>(c, ubuf, upd, sm, dt, visc); break;
    case 2: launch_force_g<D, FORM, 2>(c, ubuf, upd, sm, dt, visc); break;
    default: launch_force_g<D, FORM, 1>(c, ubuf, upd, sm, dt, visc); break;
    }
}

int pdx_launch_force(pdx_ctx *c, int ubuf, int upd, bool solve_mode, double dt, double visc)
{
    const int sm = solve_mode ? 1 : 0;
    if (c->dim == 3) {
        if (c->form == PDX_FORM_BOND) launch_force_t<3, PDX_FORM_BOND>(c, ubuf, upd, sm, dt, visc);
        else launch_force_t<3, PDX_FORM_NOSB>(c, ubuf, upd, sm, dt, visc);
    } else {
        if (c->form == PDX_FORM_BOND) launch_force_t<2, PDX_FORM_BOND>(c, ubuf, upd, sm, dt, visc);
        else launch_force_t<2, PDX_FORM_NOSB>(c, ubuf, upd, sm, dt, visc);
    }
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

int pdx_launch_kd_move(pdx_ctx *c, int ubuf_in, bool solve_mode, double dt, bool apply_reset)
{
    dim3 g(c->nblk), b(PDX_BLOCK);
    if (c->dim == 3)
        hipLaunchKernelGGL((k_kd_move<3>), g, b, 0, c->stream, c->d, ubuf_in, solve_mode ? 1 : 0, dt, apply_reset ? 1 : 0);
    else
        hipLaunchKernelGGL((k_kd_move<2>), g, b, 0, c->stream, c->d, ubuf_in, solve_mode ? 1 : 0, dt, apply_reset ? 1 : 0);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

int pdx_launch_adr_mass(pdx_ctx *c, double beta)
{
    double *tmp = nullptr;
    PDX_HIP(hipMallocAsync((void **)&tmp, sizeof(double) * 2 * (size_t)c->ni, c->stream));
    dim3 g(grid1(c->ni)), b(PDX_BLOCK);
    if (c->dim == 3) {
        hipLaunchKernelGGL((k_adr_mass_a<3>), g, b, 0, c->stream, c->d, tmp, tmp + c->ni);
        hipLaunchKernelGGL((k_adr_mass_b<3>), g, b, 0, c->stream, c->d, tmp, tmp + c->ni, beta);
    } else {
        hipLaunchKernelGGL((k_adr_mass_a<2>), g, b, 0, c->stream, c->d, tmp, tmp + c->ni);
        hipLaunchKernelGGL((k_adr_mass_b<2>), g, b, 0, c->stream, c->d, tmp, tmp + c->ni, beta);
    }
    PDX_LAUNCH_CHECK();
    PDX_HIP(hipFreeAsync(tmp, c->stream));
    return PDX_OK;
}

int pdx_launch_dt_est(pdx_ctx *c, double modulus, double *d_out)
{
    dim3 g(grid1(c->ni)), b(PDX_BLOCK);
    if (c->dim == 3)
        hipLaunchKernelGGL((k_dt_est<3>), g, b, 0, c->stream, c->d, modulus, d_out);
    else
        hipLaunchKernelGGL((k_dt_est<2>), g, b, 0, c->stream, c->d, modulus, d_out);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

int pdx_launch_damage(pdx_ctx *c, double s_c, unsigned long long *d_count)
{
    dim3 g(grid1(c->ni)), b(PDX_BLOCK);
    if (c->dim == 3)
        hipLaunchKernelGGL((k_damage<3>), g, b, 0, c->stream, c->d, c->ucur, s_c, d_count);
    else
        hipLaunchKernelGGL((k_damage<2>), g, b, 0, c->stream, c->d, c->ucur, s_c, d_count);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

int pdx_launch_local_damage(pdx_ctx *c)
{
    hipLaunchKernelGGL(k_local_damage, dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

int pdx_launch_bond_data(pdx_ctx *c, double *d_xi, double *d_len, double *d_w, const int64_t *d_rowi)
{
    dim3 g(grid1(c->ni)), b(PDX_BLOCK);
    if (c->dim == 3)
        hipLaunchKernelGGL((k_bond_data<3>), g, b, 0, c->stream, c->d, d_xi, d_len, d_w, d_rowi);
    else
        hipLaunchKernelGGL((k_bond_data<2>), g, b, 0, c->stream, c->d, d_xi, d_len, d_w, d_rowi);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}


int pdx_launch_resid(pdx_ctx *c)
{
    if (c->dim == 3)
        hipLaunchKernelGGL((k_resid<3>), dim3(c->nblk), dim3(PDX_BLOCK), 0, c->stream, c->d);
    else
        hipLaunchKernelGGL((k_resid<2>), dim3(c->nblk), dim3(PDX_BLOCK), 0, c->stream, c->d);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

int pdx_launch_zero_free(pdx_ctx *c, int ubuf)
{
    hipLaunchKernelGGL(k_zero_free, dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d, ubuf);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

int pdx_launch_kr_put(pdx_ctx *c, const float *d_src, int off, int cnt)
{
    hipLaunchKernelGGL(k_kr_put, dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d, d_src, off, cnt);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

