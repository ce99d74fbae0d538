// pdx_kernels.hip -- gfx950 kernels of the NOSB peridynamics solver step.
//
// Hot loop, per dynamic-relaxation iteration (one thread per particle, SELL-64
// bonds, wave64):
//   k_fpass  : H = [sum_j w V_j du (x) xi] K^-1 (FP32), P(H) (LE/NH in FP32,
//              J2 radial return in FP64), A = P K^-1 -> ARec
//   k_force  : f_i = A_i m_i + sum_j w V_j [A_j xi + G_s c_ij s eta/|eta|]
//              (FP32 bond terms, FP64 per-particle sums), + external load,
//              fused DR update, FP64 block partials, last-block reduction of the
//              case's residual / damping / kinetic energy
//   k_kd_move: kinetic-damping position update (reference-faithful mode)
// Reference arithmetic: backend/fea/peridynamics/core/nosb.py:96-126 (F),
// :211-227 (P), :229-275 (force); backend/fea/peridynamics/solver/nosb_solver.py
// :156-198 (kinetic DR).
//
// Setup kernels (FP64): neighbour count/fill over a cell grid, shape tensor
// and inverse, adaptive-DR mass, dt estimate, damage.
#include <float.h>

#include "pdx_internal.h"

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
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

// Block sum of NV doubles; partials[lb] written; returns true in the block that
// arrives last for its case, with tot[] = case totals summed in block order.
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

template <int D>
__device__ __forceinline__ void matmul_f(const float *A, const float *B, float *C)
{
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            float s = 0.f;
#pragma unroll
            for (int k = 0; k < D; ++k) s = fmaf(A[r * D + k], B[k * D + c], s);
            C[r * D + c] = s;
        }
}

template <int D>
__device__ __forceinline__ float lame_f(float Kb, float mu)
{
    return D == 3 ? Kb - 2.0f * mu / 3.0f : Kb - mu;
}

// linear elastic: eps = sym(H), P = lam tr(eps) I + 2 mu eps   (nosb.py:218-227)
template <int D>
__device__ __forceinline__ void stress_le(const float *H, float Kb, float mu, float *P)
{
    const float lam = lame_f<D>(Kb, mu);
    float tr = 0.f;
#pragma unroll
    for (int q = 0; q < D; ++q) tr += H[q * D + q];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const float e = 0.5f * (H[r * D + c] + H[c * D + r]);
            P[r * D + c] = 2.0f * mu * e + (r == c ? lam * tr : 0.0f);
        }
}

// Neo-Hookean P = mu (F - F^-T) + lam ln(J_s) F^-T with F = I + H
// (neo_hookean.py:117-127 sigma, :176-180 P = J sigma F^-T), evaluated without
// cancellation: J - 1 from the invariants of H, F^-T = I + M^T with
// M = F^-1 - I = -F^-1 H.
template <int D>
__device__ __forceinline__ void stress_nh(const float *H, float Kb, float mu, float *P)
{
    const float lam = lame_f<D>(Kb, mu);
    float Jm1, adj[D * D];
    if (D == 3) {
        const float tr = H[0] + H[4] + H[8];
        float trH2 = 0.f;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) trH2 = fmaf(H[r * 3 + c], H[c * 3 + r], trH2);
        const float detH = H[0] * (H[4] * H[8] - H[5] * H[7]) - H[1] * (H[3] * H[8] - H[5] * H[6]) +
                           H[2] * (H[3] * H[7] - H[4] * H[6]);
        Jm1 = tr + 0.5f * (tr * tr - trH2) + detH;
        const float F0 = 1.f + H[0], F4 = 1.f + H[4], F8 = 1.f + H[8];
        const float F1 = H[1], F2 = H[2], F3 = H[3], F5 = H[5], F6 = H[6], F7 = H[7];
        adj[0] = F4 * F8 - F5 * F7; adj[1] = F2 * F7 - F1 * F8; adj[2] = F1 * F5 - F2 * F4;
        adj[3] = F5 * F6 - F3 * F8; adj[4] = F0 * F8 - F2 * F6; adj[5] = F2 * F3 - F0 * F5;
        adj[6] = F3 * F7 - F4 * F6; adj[7] = F1 * F6 - F0 * F7; adj[8] = F0 * F4 - F1 * F3;
    } else {
        Jm1 = H[0] + H[3] + (H[0] * H[3] - H[1] * H[2]);
        adj[0] = 1.f + H[3]; adj[1] = -H[1]; adj[2] = -H[2]; adj[3] = 1.f + H[0];
    }
    const float J = 1.f + Jm1;
    const float lnJ = J > 1e-8f ? log1pf(Jm1) : logf(1e-8f);
    const float iJ = 1.0f / J;
    float M[D * D];  // M = -F^-1 H
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            float s = 0.f;
#pragma unroll
            for (int k = 0; k < D; ++k) s = fmaf(adj[r * D + k], H[k * D + c], s);
            M[r * D + c] = -s * iJ;
        }
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const float mt = M[c * D + r];
            P[r * D + c] = mu * (H[r * D + c] - mt) + lam * lnJ * ((r == c ? 1.f : 0.f) + mt);
        }
}

// J2 small-strain radial return with linear isotropic + kinematic hardening,
// FP64 (j2_plasticity.py:202-249; h_kin = 0 is the reference).  Symmetric
// tensors packed xx,yy,zz,xy,yz,xz; off-diagonal entries count twice in norms.
__device__ __forceinline__ void stress_j2(const float *H, float Kb_, float mu_, const J2Par &pp,
                                          const J2State &st, float *P, J2State *out)
{
    const double Kb = Kb_, mu = mu_;
    const double lam = Kb - 2.0 * mu / 3.0;
    double eps[6];
    eps[0] = H[0]; eps[1] = H[4]; eps[2] = H[8];
    eps[3] = 0.5 * ((double)H[1] + (double)H[3]);
    eps[4] = 0.5 * ((double)H[5] + (double)H[7]);
    eps[5] = 0.5 * ((double)H[2] + (double)H[6]);
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
            sfin[q] = bn + fac * eta[q] + (q < 3 ? pr : 0.0);
        }
        ns.epbar = (float)(epbar + dg);
        if (out) *out = ns;
    }
    P[0] = (float)sfin[0]; P[4] = (float)sfin[1]; P[8] = (float)sfin[2];
    P[1] = P[3] = (float)sfin[3];
    P[5] = P[7] = (float)sfin[4];
    P[2] = P[6] = (float)sfin[5];
}

// ---------------------------------------------------------------------------
// k_fpass: deformation gradient, stress, A = P K^-1
// ---------------------------------------------------------------------------
template <int D, int MODEL, bool DIAG, bool COMMIT>
__global__ __launch_bounds__(PDX_BLOCK) void k_fpass(Dev d, int ubuf, int solve_mode)
{
    const int lb = logical_block();
    if (solve_mode && d.sc[d.case_of_blk[lb]].converged) return;
    const int i = lb * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const KRec kr = d.KR[i];
    if (!(kr.flags & PDX_FL_ACTIVE)) return;
    const double4 *__restrict__ U = d.u[ubuf];
    const PRec pi = d.P[i];
    const double4 ui = U[i];
    const float invh = (float)d.inv_h;
    float Hs[D * D];
#pragma unroll
    for (int q = 0; q < D * D; ++q) Hs[q] = 0.f;
    const int64_t base = d.sl_ptr[i >> 6] + (i & 63);
    const int nnb = (int)kr.nnb;
    for (int k = 0; k < nnb; ++k) {
        const uint32_t e = d.bj[base + (int64_t)k * PDX_SLICE];
        if (e & PDX_BROKEN) continue;
        const int j = (int)e;
        const PRec pj = d.P[j];
        const double4 uj = U[j];
        float xi[3], du[3];
        xi[0] = (float)(pj.X[0] - pi.X[0]);
        xi[1] = (float)(pj.X[1] - pi.X[1]);
        xi[2] = (float)(pj.X[2] - pi.X[2]);
        du[0] = (float)(uj.x - ui.x);
        du[1] = (float)(uj.y - ui.y);
        du[2] = (float)(uj.z - ui.z);
        float L2 = 0.f;
#pragma unroll
        for (int q = 0; q < D; ++q) L2 = fmaf(xi[q], xi[q], L2);
        const float L = sqrtf(L2);
        const float w = (L > 0.f ? fmaxf(0.f, 1.f - L * invh) : 1.f) * pj.V;
#pragma unroll
        for (int r = 0; r < D; ++r) {
            const float wd = w * du[r];
#pragma unroll
            for (int c = 0; c < D; ++c) Hs[r * D + c] = fmaf(wd, xi[c], Hs[r * D + c]);
        }
    }
    float Km[D * D];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) Km[r * D + c] = kr.Kinv[r * 3 + c];
    float H[D * D];
    const bool sing = (kr.flags & PDX_FL_SING) != 0;
    if (sing) {
        // K^-1 fell back to I, so F = N = K + sum w du (x) xi  (nosb.py:92-94, 126)
#pragma unroll
        for (int q = 0; q < D * D; ++q) H[q] = Hs[q] + Km[q] - ((q % (D + 1)) == 0 ? 1.f : 0.f);
    } else {
        matmul_f<D>(Hs, Km, H);
    }
    float P[D * D];
    if (MODEL == PDX_MODEL_NEO_HOOKEAN) {
        stress_nh<D>(H, kr.bulk, kr.shear, P);
    } else if (MODEL == PDX_MODEL_J2) {
        J2State ns;
        stress_j2(H, kr.bulk, kr.shear, d.j2p[i], d.j2s[i], P, COMMIT ? &ns : nullptr);
        if (COMMIT) d.j2s[i] = ns;
    } else {
        stress_le<D>(H, kr.bulk, kr.shear, P);
    }
    ARec ar;
    if (sing) {
#pragma unroll
        for (int q = 0; q < D * D; ++q) ar.A[q] = P[q];
    } else {
        matmul_f<D>(P, Km, ar.A);
    }
#pragma unroll
    for (int q = D * D; q < 9; ++q) ar.A[q] = 0.f;
    ar.pad[0] = ar.pad[1] = ar.pad[2] = 0.f;
    d.AR[i] = ar;
    if (DIAG) {
#pragma unroll
        for (int q = 0; q < D * D; ++q) {
            d.Fd[(int64_t)i * 9 + q] = H[q] + ((q % (D + 1)) == 0 ? 1.f : 0.f);
            d.Pd[(int64_t)i * 9 + q] = P[q];
        }
    }
}

// ---------------------------------------------------------------------------
// k_force: force gather + external load + fused DR update + reductions
//   UPD 0: write f, reduce |f|^2 over free DOFs (force-only evaluation)
//   UPD 1: adaptive DR (damping lagged one step), u_out = u_in + V
//   UPD 2: kinetic DR velocity update and kinetic energy (nosb_solver.py:164-182)
// ---------------------------------------------------------------------------
template <int D, int FORM, int UPD>
__global__ __launch_bounds__(PDX_BLOCK) void k_force(Dev d, int ubuf, int solve_mode, double dt, double visc)
{
    const int lb = logical_block();
    const int cs = d.case_of_blk[lb];
    if (solve_mode && d.sc[cs].converged) return;
    const int i = lb * PDX_BLOCK + threadIdx.x;
    double red[3] = {0.0, 0.0, 0.0};
    bool act = i < d.ni;
    KRec kr;
    if (act) {
        kr = d.KR[i];
        act = (kr.flags & PDX_FL_ACTIVE) != 0;
    }
    if (act) {
        const double4 *__restrict__ U = d.u[ubuf];
        const PRec pi = d.P[i];
        const double4 ui = U[i];
        const float invh = (float)d.inv_h;
        const float Gs = (float)d.G_s;
        const float ci = pi.c;
        double fa[3] = {0.0, 0.0, 0.0};
        const int64_t base = d.sl_ptr[i >> 6] + (i & 63);
        const int nnb = (int)kr.nnb;
        for (int k = 0; k < nnb; ++k) {
            const uint32_t e = d.bj[base + (int64_t)k * PDX_SLICE];
            if (e & PDX_BROKEN) continue;
            const int j = (int)e;
            const PRec pj = d.P[j];
            const double4 uj = U[j];
            float xi[3], du[3];
            xi[0] = (float)(pj.X[0] - pi.X[0]);
            xi[1] = (float)(pj.X[1] - pi.X[1]);
            xi[2] = (float)(pj.X[2] - pi.X[2]);
            du[0] = (float)(uj.x - ui.x);
            du[1] = (float)(uj.y - ui.y);
            du[2] = (float)(uj.z - ui.z);
            float L2 = 0.f, xd = 0.f, dd = 0.f;
#pragma unroll
            for (int q = 0; q < D; ++q) {
                L2 = fmaf(xi[q], xi[q], L2);
                xd = fmaf(xi[q], du[q], xd);
                dd = fmaf(du[q], du[q], dd);
            }
            const float L = sqrtf(L2);
            const float w = (L > 0.f ? fmaxf(0.f, 1.f - L * invh) : 1.f) * pj.V;
            float eta[3];
#pragma unroll
            for (int q = 0; q < D; ++q) eta[q] = xi[q] + du[q];
            // |eta|^2 = |xi|^2 + (2 xi.du + du.du); the stretch uses the
            // cancellation-free difference |eta| - |xi| = (2 xi.du + du.du)/(|eta|+|xi|)
            const float dn = fmaf(2.f, xd, dd);
            const float E2 = L2 + dn;
            const float el = sqrtf(fmaxf(E2, 0.f));
            float tb[3] = {0.f, 0.f, 0.f};
            if (FORM == PDX_FORM_NOSB) {
                const ARec aj = d.AR[j];
#pragma unroll
                for (int r = 0; r < D; ++r) {
                    float t = 0.f;
#pragma unroll
                    for (int c = 0; c < D; ++c) t = fmaf(aj.A[r * D + c], xi[c], t);
                    tb[r] = t;
                }
                if (el > 1e-10f && L > 1e-10f) {
                    const float s = dn / ((el + L) * L);
                    const float g = Gs * 0.5f * (ci + pj.c) * s / el;
#pragma unroll
                    for (int r = 0; r < D; ++r) tb[r] = fmaf(g, eta[r], tb[r]);
                }
            } else {
                if (el > 1e-10f) {
                    const float s = dn / ((el + L) * L);
                    const float g = 0.5f * (ci + pj.c) * s / el;
#pragma unroll
                    for (int r = 0; r < D; ++r) tb[r] = g * eta[r];
                }
            }
#pragma unroll
            for (int r = 0; r < D; ++r) fa[r] += (double)(w * tb[r]);
        }
        if (FORM == PDX_FORM_NOSB) {
            const ARec ai = d.AR[i];
#pragma unroll
            for (int r = 0; r < D; ++r) {
                float t = 0.f;
#pragma unroll
                for (int c = 0; c < D; ++c) t = fmaf(ai.A[r * D + c], kr.m[c], t);
                fa[r] += (double)t;
            }
        }
        if (kr.flags & PDX_FL_LOADED) {
            const double4 fe = d.fext[i];
            fa[0] += fe.x;
            fa[1] += fe.y;
            if (D == 3) fa[2] += fe.z;
        }
        const uint32_t fm = kr.flags & PDX_FL_FIXMASK;
#pragma unroll
        for (int r = 0; r < D; ++r)
            if (!((fm >> r) & 1u)) red[0] += fa[r] * fa[r];
        if (UPD == 0) {
            d.f[i] = make_double4(fa[0], fa[1], D == 3 ? fa[2] : 0.0, 0.0);
        } else if (UPD == 1) {
            const DScal &S = d.sc[cs];
            const double c = S.c;
            const int first = S.first;
            const double lam = d.lam[i];
            const double4 vo4 = d.v[i], fo4 = d.fp[i];
            const double vo[3] = {vo4.x, vo4.y, vo4.z}, fo[3] = {fo4.x, fo4.y, fo4.z};
            const double uo[3] = {ui.x, ui.y, ui.z};
            double vn[3] = {0.0, 0.0, 0.0}, un[3] = {uo[0], uo[1], uo[2]}, fn[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int r = 0; r < D; ++r) {
                if ((fm >> r) & 1u) continue;
                const double F = fa[r];
                if (first) {
                    vn[r] = F / (2.0 * lam);
                } else {
                    if (vo[r] != 0.0) {
                        const double kk = -(F - fo[r]) / (lam * vo[r]);
                        red[1] += uo[r] * kk * uo[r];
                    }
                    red[2] += uo[r] * uo[r];
                    vn[r] = ((2.0 - c) * vo[r] + 2.0 * F / lam) / (2.0 + c);
                }
                fn[r] = F;
                un[r] = uo[r] + vn[r];
            }
            d.v[i] = make_double4(vn[0], vn[1], vn[2], 0.0);
            d.fp[i] = make_double4(fn[0], fn[1], fn[2], 0.0);
            d.u[ubuf ^ 1][i] = make_double4(un[0], un[1], un[2], 0.0);
        } else {
            const double m = d.mass[i];
            const double4 vo4 = d.v[i], ao4 = d.fp[i];
            double vv[3] = {vo4.x, vo4.y, vo4.z}, aa[3] = {ao4.x, ao4.y, ao4.z};
            const double damp = 1.0 - visc;
#pragma unroll
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
            d.v[i] = make_double4(vv[0], vv[1], vv[2], 0.0);
            d.fp[i] = make_double4(aa[0], aa[1], aa[2], 0.0);
            d.f[i] = make_double4(fa[0], fa[1], D == 3 ? fa[2] : 0.0, 0.0);
        }
    }
    double tot[3];
    if (!reduce_case<3>(d, lb, cs, red, tot)) return;
    if (threadIdx.x != 0) return;
    DScal &S = d.sc[cs];
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
// (nosb_solver.py:156-162), after applying a pending velocity reset (:239-243)
template <int D>
__global__ __launch_bounds__(PDX_BLOCK) void k_kd_move(Dev d, int ubuf_in, int solve_mode, double dt, int apply_reset)
{
    const int lb = logical_block();
    const int cs = d.case_of_blk[lb];
    if (solve_mode && d.sc[cs].converged) return;
    const int i = lb * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const KRec kr = d.KR[i];
    const double4 u = d.u[ubuf_in][i];
    double4 v = d.v[i];
    if (apply_reset && d.sc[cs].reset_pending) {
        v = make_double4(0.0, 0.0, 0.0, 0.0);
        d.v[i] = v;
    }
    const double4 a = d.fp[i];
    const uint32_t fm = kr.flags & PDX_FL_FIXMASK;
    double4 o = u;
    if (kr.flags & PDX_FL_ACTIVE) {
        if (!(fm & 1u)) o.x = u.x + v.x * dt + 0.5 * a.x * dt * dt;
        if (!(fm & 2u)) o.y = u.y + v.y * dt + 0.5 * a.y * dt * dt;
        if (D == 3 && !(fm & 4u)) o.z = u.z + v.z * dt + 0.5 * a.z * dt * dt;
    }
    d.u[ubuf_in ^ 1][i] = o;
}

// ---------------------------------------------------------------------------
// setup kernels (FP64)
// ---------------------------------------------------------------------------
// neighbour scan over the 27 (9) cells around i; distance summed as
// (dx*dx + dy*dy) + dz*dz without contraction, the reference's set test
// (neighbor.py:179-183, 199-203)
template <int D, bool FILL>
__global__ __launch_bounds__(PDX_BLOCK) void k_neighbors(Dev d)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const int cs = d.case_of_p[i];
    if (cs < 0) {
        if (!FILL) d.counts[i] = 0;
        return;
    }
    const GridP g = d.grid[cs];
    const PRec pi = d.P[i];
    int c[3] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < D; ++q) {
        int t = (int)floor((pi.X[q] - g.lo[q]) / g.h);
        c[q] = t < 0 ? 0 : (t >= g.nc[q] ? g.nc[q] - 1 : t);
    }
    const int rz = D == 3 ? 1 : 0;
    int cnt = 0;
    int64_t base = 0;
    if (FILL) base = d.sl_ptr[i >> 6] + (i & 63);
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
                    const PRec pj = d.P[j];
                    double s = 0.0;
#pragma unroll
                    for (int q = 0; q < D; ++q) {
                        const double t = pj.X[q] - pi.X[q];
                        s = s + t * t;
                    }
                    if (s < d.h2) {
                        if (FILL) d.bj[base + (int64_t)cnt * PDX_SLICE] = (uint32_t)j;
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
    // rows sorted by ascending neighbour index (deterministic order)
    for (int p = 1; p < cnt; ++p) {
        const uint32_t key = d.bj[base + (int64_t)p * PDX_SLICE];
        int q = p - 1;
        while (q >= 0 && d.bj[base + (int64_t)q * PDX_SLICE] > key) {
            d.bj[base + (int64_t)(q + 1) * PDX_SLICE] = d.bj[base + (int64_t)q * PDX_SLICE];
            --q;
        }
        d.bj[base + (int64_t)(q + 1) * PDX_SLICE] = key;
    }
    d.init_nb[i] = cnt;
}

__device__ __forceinline__ double det3(const double *M)
{
    return M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
           M[2] * (M[3] * M[7] - M[4] * M[6]);
}

// shape tensor K = sum_intact w xi (x) xi V_j, FP64, with its inverse
// (nosb.py:59-94); also m_i = sum_intact w V_j xi for the force pass
template <int D>
__global__ __launch_bounds__(PDX_BLOCK) void k_shape(Dev d, int mode, double tol)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    KRec kr = d.KR[i];
    if (!(kr.flags & PDX_FL_ACTIVE)) return;
    const PRec pi = d.P[i];
    double K[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double m[3] = {0, 0, 0};
    const int64_t base = d.sl_ptr[i >> 6] + (i & 63);
    for (int k = 0; k < (int)kr.nnb; ++k) {
        const uint32_t e = d.bj[base + (int64_t)k * PDX_SLICE];
        if (e & PDX_BROKEN) continue;
        const PRec pj = d.P[(int)e];
        double xi[3] = {0, 0, 0}, s = 0.0;
        for (int q = 0; q < D; ++q) {
            xi[q] = pj.X[q] - pi.X[q];
            s = s + xi[q] * xi[q];
        }
        const double L = sqrt(s);
        double w = 1.0;
        if (L > 0) {
            w = 1.0 - L / d.h;
            w = w > 0.0 ? w : 0.0;
        }
        const double Vj = d.V64[(int)e & 0x7fffffff];
        for (int r = 0; r < D; ++r) {
            for (int c = 0; c < D; ++c) K[r * D + c] += w * xi[r] * xi[c] * Vj;
            m[r] += w * Vj * xi[r];
        }
    }
    double det, tr = 0.0;
    for (int q = 0; q < D; ++q) tr += K[q * D + q];
    tr /= D;
    double Ki[9];
    if (D == 3) {
        det = det3(K);
    } else {
        det = K[0] * K[3] - K[1] * K[2];
    }
    const double thr = mode == PDX_SING_RELATIVE ? tol * (D == 3 ? tr * tr * tr : tr * tr) : tol;
    bool sing = !(fabs(det) > thr && det != 0.0);
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

// adaptive-DR mass, pass A: CS_i = max(3K,2mu) S_i and |G_i| (see pdx_adr_mass)
template <int D>
__global__ __launch_bounds__(PDX_BLOCK) void k_adr_mass_a(Dev d, double *cs_out, double *g_out)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const KRec kr = d.KR[i];
    cs_out[i] = 0.0;
    g_out[i] = 0.0;
    if (!(kr.flags & PDX_FL_ACTIVE) || d.form != PDX_FORM_NOSB) return;
    const PRec pi = d.P[i];
    const double *Ki = d.Kinv64 + (int64_t)i * 9;
    double G[3] = {0, 0, 0}, sabs = 0.0;
    const int64_t base = d.sl_ptr[i >> 6] + (i & 63);
    for (int k = 0; k < (int)kr.nnb; ++k) {
        const uint32_t e = d.bj[base + (int64_t)k * PDX_SLICE];
        if (e & PDX_BROKEN) continue;
        const PRec pj = d.P[(int)e];
        double xi[3] = {0, 0, 0}, s = 0.0;
        for (int q = 0; q < D; ++q) {
            xi[q] = pj.X[q] - pi.X[q];
            s += xi[q] * xi[q];
        }
        const double L = sqrt(s);
        const double w = L > 0 ? fmax(0.0, 1.0 - L / d.h) : 1.0;
        double gn = 0.0;
        for (int r = 0; r < D; ++r) {
            double t = 0.0;
            for (int c = 0; c < D; ++c) t += Ki[r * D + c] * xi[c];
            t *= w * d.V64[(int)e];
            G[r] += t;
            gn += t * t;
        }
        sabs += sqrt(gn);
    }
    double gl = 0.0;
    for (int r = 0; r < D; ++r) gl += G[r] * G[r];
    gl = sqrt(gl);
    const double C = fmax(3.0 * (double)kr.bulk, 2.0 * (double)kr.shear);
    cs_out[i] = C * (sabs + gl);
    g_out[i] = gl;
}

// pass B: lam_i = beta/4 (sqrt3 k_corr + 2 sqrt3 G_s k_stab)
template <int D>
__global__ __launch_bounds__(PDX_BLOCK) void k_adr_mass_b(Dev d, const double *cs_in, const double *g_in, double beta)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const KRec kr = d.KR[i];
    if (!(kr.flags & PDX_FL_ACTIVE)) {
        d.lam[i] = 1.0;
        return;
    }
    const PRec pi = d.P[i];
    const bool nosb = d.form == PDX_FORM_NOSB;
    double kc = nosb ? cs_in[i] * g_in[i] : 0.0, ks = 0.0;
    const int64_t base = d.sl_ptr[i >> 6] + (i & 63);
    for (int k = 0; k < (int)kr.nnb; ++k) {
        const uint32_t e = d.bj[base + (int64_t)k * PDX_SLICE];
        if (e & PDX_BROKEN) continue;
        const int j = (int)e;
        const PRec pj = d.P[j];
        double xi[3] = {0, 0, 0}, s = 0.0;
        for (int q = 0; q < D; ++q) {
            xi[q] = pj.X[q] - pi.X[q];
            s += xi[q] * xi[q];
        }
        const double L = sqrt(s);
        const double w = L > 0 ? fmax(0.0, 1.0 - L / d.h) : 1.0;
        if (nosb) {
            const double *Kj = d.Kinv64 + (int64_t)j * 9;
            double gn = 0.0;
            for (int r = 0; r < D; ++r) {
                double t = 0.0;
                for (int c = 0; c < D; ++c) t -= Kj[r * D + c] * xi[c];
                gn += t * t;
            }
            kc += w * d.V64[j] * cs_in[j] * sqrt(gn);
        }
        if (L > 1e-10) ks += 0.5 * ((double)pi.c + (double)pj.c) * w * d.V64[j] / L;
    }
    const double gs = nosb ? d.G_s : 1.0;
    const double r3 = 1.7320508075688772;
    double lam = 0.25 * beta * (r3 * kc + 2.0 * r3 * gs * ks);
    if (!(lam > 0.0)) lam = 1.0;
    d.lam[i] = lam;
}

// dt estimate per particle (nosb_solver.py:121-143): lam_i = k_eff / m_i
template <int D>
__global__ __launch_bounds__(PDX_BLOCK) void k_dt_est(Dev d, double modulus, double *out)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const KRec kr = d.KR[i];
    out[i] = 0.0;
    if (!(kr.flags & PDX_FL_ACTIVE)) return;
    const double mi = d.mass[i];
    if (mi < 1e-30) return;
    const PRec pi = d.P[i];
    const double *Ki = d.Kinv64 + (int64_t)i * 9;
    double ds[3] = {0, 0, 0}, dsq = 0.0;
    const int64_t base = d.sl_ptr[i >> 6] + (i & 63);
    for (int k = 0; k < (int)kr.nnb; ++k) {
        const uint32_t e = d.bj[base + (int64_t)k * PDX_SLICE];
        if (e & PDX_BROKEN) continue;
        const PRec pj = d.P[(int)e];
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
// s > s_c strictly (damage.py:94-116)
template <int D>
__global__ __launch_bounds__(PDX_BLOCK) void k_damage(Dev d, int ubuf, double s_c, unsigned long long *count)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const KRec kr = d.KR[i];
    if (!(kr.flags & PDX_FL_ACTIVE)) return;
    const double4 *U = d.u[ubuf];
    const PRec pi = d.P[i];
    const double4 ui = U[i];
    const double xi0[3] = {pi.X[0] + ui.x, pi.X[1] + ui.y, pi.X[2] + ui.z};
    const int64_t base = d.sl_ptr[i >> 6] + (i & 63);
    unsigned long long nb = 0;
    for (int k = 0; k < (int)kr.nnb; ++k) {
        const int64_t a = base + (int64_t)k * PDX_SLICE;
        const uint32_t e = d.bj[a];
        if (e & PDX_BROKEN) continue;
        const int j = (int)e;
        const PRec pj = d.P[j];
        const double4 uj = U[j];
        const double xj[3] = {pj.X[0] + uj.x, pj.X[1] + uj.y, pj.X[2] + uj.z};
        double se = 0.0, sl = 0.0;
        for (int q = 0; q < D; ++q) {
            const double t = xj[q] - xi0[q];
            se = se + t * t;
            const double r = pj.X[q] - pi.X[q];
            sl = sl + r * r;
        }
        const double el = sqrt(se), L = sqrt(sl);
        double st = 0.0;
        if (L > 1e-10) st = (el - L) / L;
        if (st > s_c) {
            d.bj[a] = e | PDX_BROKEN;
            nb++;
        }
    }
    if (nb) atomicAdd(count, nb);
}

// phi_i = broken / initial bonds (damage.py:118-142)
__global__ __launch_bounds__(PDX_BLOCK) void k_local_damage(Dev d)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const KRec kr = d.KR[i];
    const int init = d.init_nb[i];
    if (!(kr.flags & PDX_FL_ACTIVE) || init <= 0) {
        d.damage[i] = 0.0;
        return;
    }
    const int64_t base = d.sl_ptr[i >> 6] + (i & 63);
    int nbk = 0;
    for (int k = 0; k < (int)kr.nnb; ++k) nbk += (d.bj[base + (int64_t)k * PDX_SLICE] & PDX_BROKEN) ? 1 : 0;
    d.damage[i] = (double)nbk / (double)init;
}

// per-bond reference data in internal CSR order (bonds.py:116-137)
template <int D>
__global__ __launch_bounds__(PDX_BLOCK) void k_bond_data(Dev d, double *xi_o, double *len_o, double *w_o,
                                                         const int64_t *rowi)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const KRec kr = d.KR[i];
    if (!(kr.flags & PDX_FL_ACTIVE)) return;
    const PRec pi = d.P[i];
    const int64_t base = d.sl_ptr[i >> 6] + (i & 63);
    for (int k = 0; k < (int)kr.nnb; ++k) {
        const uint32_t e = d.bj[base + (int64_t)k * PDX_SLICE];
        const PRec pj = d.P[(int)(e & ~PDX_BROKEN)];
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

template <int D, int MODEL>
static void launch_fpass_t(pdx_ctx *c, int ubuf, bool diag, bool commit, int sm)
{
    dim3 g(c->nblk), b(PDX_BLOCK);
    if (commit)
        hipLaunchKernelGGL((k_fpass<D, MODEL, false, true>), g, b, 0, c->stream, c->d, ubuf, sm);
    else if (diag)
        hipLaunchKernelGGL((k_fpass<D, MODEL, true, false>), g, b, 0, c->stream, c->d, ubuf, sm);
    else
        hipLaunchKernelGGL((k_fpass<D, MODEL, false, false>), g, b, 0, c->stream, c->d, ubuf, sm);
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

template <int D, int FORM>
static void launch_force_t(pdx_ctx *c, int ubuf, int upd, int sm, double dt, double visc)
{
    dim3 g(c->nblk), b(PDX_BLOCK);
    if (upd == 1)
        hipLaunchKernelGGL((k_force<D, FORM, 1>), g, b, 0, c->stream, c->d, ubuf, sm, dt, visc);
    else if (upd == 2)
        hipLaunchKernelGGL((k_force<D, FORM, 2>), g, b, 0, c->stream, c->d, ubuf, sm, dt, visc);
    else
        hipLaunchKernelGGL((k_force<D, FORM, 0>), g, b, 0, c->stream, c->d, ubuf, sm, dt, visc);
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


// residual |f|_free per case from the stored force (nosb_solver.py:190-198)
template <int D>
__global__ __launch_bounds__(PDX_BLOCK) void k_resid(Dev d)
{
    const int lb = logical_block();
    const int cs = d.case_of_blk[lb];
    const int i = lb * PDX_BLOCK + threadIdx.x;
    double red[1] = {0.0};
    if (i < d.ni) {
        const KRec kr = d.KR[i];
        if (kr.flags & PDX_FL_ACTIVE) {
            const double4 f = d.f[i];
            const double fv[3] = {f.x, f.y, f.z};
            const uint32_t fm = kr.flags & PDX_FL_FIXMASK;
            for (int r = 0; r < D; ++r)
                if (!((fm >> r) & 1u)) red[0] += fv[r] * fv[r];
        }
    }
    double tot[1];
    if (!reduce_case<1>(d, lb, cs, red, tot)) return;
    if (threadIdx.x == 0) d.sc[cs].res = sqrt(tot[0]);
}

// free DOFs of u set to zero (cold-start state)
__global__ __launch_bounds__(PDX_BLOCK) void k_zero_free(Dev d, int ubuf)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const uint32_t fm = d.KR[i].flags & PDX_FL_FIXMASK;
    double4 u = d.u[ubuf][i];
    if (!(fm & 1u)) u.x = 0.0;
    if (!(fm & 2u)) u.y = 0.0;
    if (!(fm & 4u)) u.z = 0.0;
    d.u[ubuf][i] = u;
}

// scatter cnt 32-bit words per particle into KRec word offset off
__global__ __launch_bounds__(PDX_BLOCK) void k_kr_put(Dev d, const float *src, int off, int cnt)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    float *w = reinterpret_cast<float *>(&d.KR[i]);
    for (int q = 0; q < cnt; ++q) w[off + q] = src[(int64_t)i * cnt + q];
}

__global__ __launch_bounds__(PDX_BLOCK) void k_prec_c(Dev d, const float *c)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    d.P[i].c = c[i];
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

int pdx_launch_prec_c(pdx_ctx *c, const float *d_c)
{
    hipLaunchKernelGGL(k_prec_c, dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d, d_c);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}
