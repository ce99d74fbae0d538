// pdx_kernels.hip -- gfx950 kernels of the NOSB peridynamics solver step.
//
// The per-particle and per-bond arithmetic lives in pdx_device.h (one thread
// per particle, wave64).  Hot loop per dynamic-relaxation iteration, on a
// voxel lattice k_fpass_l / k_force_l (each row walks the 122-offset lattice
// stencil, the bond vectors, lengths and weights compile-time constants of
// the offset), elsewhere:
//   F-pass (k_fpass_h; gathering variants k_fpass, k_fpass_w):
//              H = [sum_j w V_j du (x) xi] K^-1 (FP32 terms, FP64 sums), P(H)
//              (LE, NH and the J2 radial return in FP64), A = P K^-1 stored in FP32
//   force pass (k_force_h; k_force, k_force_w):
//              f_i = A_i m_i + sum_j w V_j [A_j xi + G_s c_ij s eta/|eta|]
//              (FP32 bond terms, FP64 per-particle sums), + external load,
//              fused DR update, FP64 block partials of residual / damping /
//              kinetic energy
//   k_case_reduce: the case totals of those partials and the per-case update
//   k_kd_move: kinetic-damping position update (reference-faithful mode)
// Setup kernels (FP64): neighbour count/fill, shape tensor and inverse,
// adaptive-DR mass and its spectral scaling, dt estimate, damage.
#include "pdx_internal.h"

#include <utility>

// blocks are dispatched round-robin over the 8 XCDs; give every XCD one
// contiguous range of Morton blocks so the neighbour halo of a block mostly
// hits the same L2
__device__ __forceinline__ int logical_block_n(int nb)
{
    const int b = blockIdx.x;
    return (b % PDX_NXCD) * (nb / PDX_NXCD) + b / PDX_NXCD;
}

__device__ __forceinline__ int logical_block() { return logical_block_n(gridDim.x); }

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

// Wave sums of NV doubles into d.part[(lb * PDX_PW + w) * 4 + q], one entry
// per wave w < NA, written by each wave as it ends (no block barrier).  The
// case totals are summed by k_case_reduce, a separate launch: a last-block
// reduction inside the pass would need an agent-scope release per block,
// which writes back the XCD's L2 (every block's stores) on gfx950.  Only the
// first NA waves hold values (the others' are zero and are left out, which
// does not change the sum).
template <int NV, int NT = PDX_BLOCK, int NA = NT / 64>
__device__ void block_partials(const Dev &d, int lb, double (&val)[NV])
{
    static_assert(NA == PDX_PW, "one partial per wave slot");
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (wid < NA) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            const double t = wave_sum(val[q]);
            if (lane == 0) d.part[((int64_t)lb * PDX_PW + wid) * 4 + q] = t;
        }
    }
}

// case totals of the wave partials (strided per thread, one 32-byte entry per
// load, then waves in order: a fixed order, independent of block scheduling)
// and the per-case update of case cs, by one block of NT threads; MODE 0..2 =
// the force pass's UPD, MODE 3 = residual only
template <int MODE, int NT>
__device__ void case_reduce_block(const Dev &d, int cs, int ubuf, int solve_mode)
{
    constexpr int NV = MODE == 3 ? 1 : 3;
    __shared__ double sred[NT / 64][NV];
    if (solve_mode && d.sc[cs].converged) return;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int b0 = d.case_blk0[cs], b1 = d.case_blk0[cs + 1];
    const pdx_d4 *pp = reinterpret_cast<const pdx_d4 *>(d.part);
    double t[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) t[q] = 0.0;
    for (int64_t e = (int64_t)b0 * PDX_PW + threadIdx.x; e < (int64_t)b1 * PDX_PW; e += NT) {
        const pdx_d4 v = pp[e];
        t[0] += v.x;
        if constexpr (NV == 3) {
            t[1] += v.y;
            t[2] += v.z;
        }
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        const double w = wave_sum(t[q]);
        if (lane == 0) sred[wid][q] = w;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double tot[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s = 0.0;
        for (int w = 0; w < NT / 64; ++w) s += sred[w][q];
        tot[q] = s;
    }
    if (MODE == 3)
        d.sc[cs].res = sqrt(tot[0]);
    else
        case_update<MODE>(d.sc[cs], tot, ubuf);
}

// one block of PDX_RED_THREADS per case
#define PDX_RED_THREADS 1024
template <int MODE>
__global__ __launch_bounds__(PDX_RED_THREADS) void k_case_reduce(Dev d, int ubuf, int solve_mode)
{
    case_reduce_block<MODE, PDX_RED_THREADS>(d, blockIdx.x, ubuf, solve_mode);
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
    block_partials<3>(d, lb, red);
}

// ---------------------------------------------------------------------------
// wave-cooperative hot kernels
// ---------------------------------------------------------------------------
// One wave = one SELL slice = 64 particles; the wave walks the slice's bond
// slots k = 0 .. width-1 in lockstep.  At each slot the 64 neighbour records
// are fetched cooperatively: lane L loads the 16-byte quarter (L & 3) of the
// 64-byte XRec of bond 16h + (L >> 2), h = 0..3, so one load instruction
// covers 16 whole records instead of 64 scattered 16-byte pieces (the gathers
// of the per-lane kernels are bound by L1 tag lookups, about one per lane and
// instruction).  The quarters pass through a per-wave LDS tile and every lane
// reads back its own bond's record.  The arithmetic is fpass_bond /
// force_bond in the same bond order as the per-lane kernels, so both give the
// same bits.
#define PDX_WAVES (PDX_BLOCK / 64)

// 16-byte quarter of a record as a native vector (register-resident arrays)
typedef float pdx_q4 __attribute__((ext_vector_type(4)));

// LDS slot (16-byte units) of quarter q of record b: rotating the quarter by
// b >> 2 spreads 16 consecutive lanes' reads of one quarter over all 64 banks
// (four 64-byte records fill one 256-byte bank row)
__device__ __forceinline__ int xq_slot(int b, int q) { return b * 4 + ((q + (b >> 2)) & 3); }

// orders a wave's LDS writes before its reads of other lanes' data
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// XRec / ARec from the 16-byte quarters read back from LDS, in registers
struct PdxD2 {
    double a, b;
};

__device__ __forceinline__ XRec xrec_of(const pdx_q4 (&q)[4])
{
    const PdxD2 x01 = __builtin_bit_cast(PdxD2, q[0]);
    const PdxD2 u01 = __builtin_bit_cast(PdxD2, q[2]);
    const PdxD2 u2w = __builtin_bit_cast(PdxD2, q[3]);
    XRec r;
    r.X[0] = x01.a;
    r.X[1] = x01.b;
    r.X[2] = __builtin_bit_cast(double, q[1].xy);
    r.V = q[1].z;
    r.c = q[1].w;
    r.u[0] = u01.a;
    r.u[1] = u01.b;
    r.u[2] = u2w.a;
    r.w = u2w.b;
    return r;
}

__device__ __forceinline__ void arec_of(const pdx_q4 (&q)[3], float (&A)[9])
{
    A[0] = q[0].x; A[1] = q[0].y; A[2] = q[0].z; A[3] = q[0].w;
    A[4] = q[1].x; A[5] = q[1].y; A[6] = q[1].z; A[7] = q[1].w;
    A[8] = q[2].x;
}

// bond slot q of a wave's row walk: entry, liveness and the index to gather
// (own record for a dead slot)
struct WSlot {
    bool live;
    int j;
};

__device__ __forceinline__ WSlot wslot(const Dev &d, uint32_t e, int k, int nnb, bool act, int i)
{
    WSlot s;
    s.live = act && k < nnb && e < (uint32_t)d.ni;
    s.j = s.live ? (int)e : i;
    return s;
}

template <int D, int MODEL, int G>
__global__ __launch_bounds__(PDX_BLOCK) void k_fpass_w(Dev d, int ubuf, int solve_mode)
{
    __shared__ pdx_q4 tile[PDX_WAVES][64 * 4];
    const int lb = logical_block();
    if (solve_mode && d.sc[d.case_of_blk[lb]].converged) return;
    if (lb * PDX_BLOCK >= d.ni) return;  // grid padding blocks (whole block)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = lb * PDX_BLOCK + threadIdx.x;
    const KRec kr = d.KR[i];
    const bool act = (kr.flags & PDX_FL_ACTIVE) != 0;
    const int nnb = (int)kr.nnb;
    const XRec *__restrict__ X = d.xu[ubuf];
    const pdx_q4 *__restrict__ Xq = reinterpret_cast<const pdx_q4 *>(X);
    const XRec pi = X[i];
    const float invh = (float)d.inv_h;
    double Sd[D * D + D];
#pragma unroll
    for (int q = 0; q < D * D + D; ++q) Sd[q] = 0.0;
    const int64_t s0 = d.sl_ptr[i >> 6];
    const int width = (int)((d.sl_ptr[(i >> 6) + 1] - s0) / PDX_SLICE);  // same in all lanes
    const uint32_t *__restrict__ bp = d.bj + s0 + lane;
    pdx_q4 *tw = tile[wv];
    uint32_t ec[G], en[G];
#pragma unroll
    for (int q = 0; q < G; ++q) ec[q] = q < width ? bp[(int64_t)q * PDX_SLICE] : PDX_BROKEN;
    for (int k = 0; k < width; k += G) {
#pragma unroll
        for (int q = 0; q < G; ++q)
            en[q] = (k + G + q) < width ? bp[(int64_t)(k + G + q) * PDX_SLICE] : PDX_BROKEN;
        // all source indices first, then all loads, so the group's gathers
        // are in flight together
        WSlot sl[G];
        int jb[G][4];
        pdx_q4 g[G][4];
#pragma unroll
        for (int q = 0; q < G; ++q) {
            sl[q] = wslot(d, ec[q], k + q, nnb, act, i);
#pragma unroll
            for (int h = 0; h < 4; ++h) jb[q][h] = __shfl(sl[q].j, 16 * h + (lane >> 2), 64);
        }
#pragma unroll
        for (int q = 0; q < G; ++q)
#pragma unroll
            for (int h = 0; h < 4; ++h) g[q][h] = Xq[(int64_t)jb[q][h] * 4 + (lane & 3)];
        // keep the scheduler from sinking the loads between the LDS writes
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < G; ++q) {
#pragma unroll
            for (int h = 0; h < 4; ++h) tw[xq_slot(16 * h + (lane >> 2), lane & 3)] = g[q][h];
            wave_lds_sync();
            pdx_q4 rq[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) rq[p] = tw[xq_slot(lane, p)];
            wave_lds_sync();
            fpass_bond<D>(pi, xrec_of(rq), invh, Sd);
        }
#pragma unroll
        for (int q = 0; q < G; ++q) ec[q] = en[q];
    }
    if (act) {
        double Hd[D * D];
        fpass_sum_to_h<D>(Sd, pi.u, Hd);
        fpass_finish<D, MODEL, false, false>(d, i, kr, Hd);
    }
}

template <int D, int FORM, int UPD, int G>
__global__ __launch_bounds__(PDX_BLOCK) void k_force_w(Dev d, int ubuf, int solve_mode, double dt, double visc)
{
    __shared__ pdx_q4 tile[PDX_WAVES][64 * 4];
    __shared__ pdx_q4 atile[PDX_WAVES][FORM == 0 ? 64 * 4 : 1];
    const int lb = logical_block();
    const int cs = d.case_of_blk[lb];
    if (solve_mode && d.sc[cs].converged) return;
    double red[3] = {0.0, 0.0, 0.0};
    if (lb * PDX_BLOCK < d.ni) {  // grid padding blocks only join the reduction
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        const int i = lb * PDX_BLOCK + threadIdx.x;
        const KRec kr = d.KR[i];
        const bool act = (kr.flags & PDX_FL_ACTIVE) != 0;
        const int nnb = (int)kr.nnb;
        const XRec *__restrict__ X = d.xu[ubuf];
        const pdx_q4 *__restrict__ Xq = reinterpret_cast<const pdx_q4 *>(X);
        const pdx_q4 *__restrict__ Aq = reinterpret_cast<const pdx_q4 *>(d.AR);
        const XRec pi = X[i];
        const float invh = (float)d.inv_h;
        const float Gs = (float)d.G_s;
        ARec ai{};
        if (FORM == 0) ai = d.AR[i];
        double fa[3] = {0.0, 0.0, 0.0};
        const int64_t s0 = d.sl_ptr[i >> 6];
        const int width = (int)((d.sl_ptr[(i >> 6) + 1] - s0) / PDX_SLICE);
        const uint32_t *__restrict__ bp = d.bj + s0 + lane;
        pdx_q4 *tw = tile[wv];
        pdx_q4 *ta = atile[wv];
        const int qa = lane & 3;  // quarter this lane fetches
        // A records have 3 quarters; lanes of quarter 3 fetch quarter 2 again
        // (same cache line, keeps the loads unconditional)
        const int qa3 = qa < 3 ? qa : 2;
        uint32_t ec[G], en[G];
#pragma unroll
        for (int q = 0; q < G; ++q) ec[q] = q < width ? bp[(int64_t)q * PDX_SLICE] : PDX_BROKEN;
        for (int k = 0; k < width; k += G) {
#pragma unroll
            for (int q = 0; q < G; ++q)
                en[q] = (k + G + q) < width ? bp[(int64_t)(k + G + q) * PDX_SLICE] : PDX_BROKEN;
            WSlot sl[G];
            int jb[G][4];
            pdx_q4 g[G][4];
            pdx_q4 ga[G][4];
#pragma unroll
            for (int q = 0; q < G; ++q) {
                sl[q] = wslot(d, ec[q], k + q, nnb, act, i);
#pragma unroll
                for (int h = 0; h < 4; ++h) jb[q][h] = __shfl(sl[q].j, 16 * h + (lane >> 2), 64);
            }
#pragma unroll
            for (int q = 0; q < G; ++q) {
#pragma unroll
                for (int h = 0; h < 4; ++h) g[q][h] = Xq[(int64_t)jb[q][h] * 4 + qa];
                if (FORM == 0)
#pragma unroll
                    for (int h = 0; h < 4; ++h) ga[q][h] = Aq[(int64_t)jb[q][h] * 3 + qa3];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < G; ++q) {
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    const int b = 16 * h + (lane >> 2);
                    tw[xq_slot(b, qa)] = g[q][h];
                    if (FORM == 0) ta[xq_slot(b, qa)] = ga[q][h];
                }
                wave_lds_sync();
                pdx_q4 rq[4], ra[3];
                float aj[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int p = 0; p < 4; ++p) rq[p] = tw[xq_slot(lane, p)];
                if (FORM == 0)
#pragma unroll
                    for (int p = 0; p < 3; ++p) ra[p] = ta[xq_slot(lane, p)];
                wave_lds_sync();
                if (FORM == 0) arec_of(ra, aj);
                force_bond<D, FORM>(pi, xrec_of(rq), ai.A, aj, invh, Gs, fa);
            }
#pragma unroll
            for (int q = 0; q < G; ++q) ec[q] = en[q];
        }
        if (act) force_finish<D, FORM, UPD>(d, i, kr, pi, ai, ubuf, cs, dt, visc, fa, red, finish_in<UPD>(d, i, kr));
    }
    block_partials<3>(d, lb, red);
}

// ---------------------------------------------------------------------------
// halo-staged hot kernels
// ---------------------------------------------------------------------------
// The 256 rows of a block reach about six times as many distinct particles
// (the block's halo), and the gathering kernels fetch each of those records
// once per bond, ~116 times per pass over all blocks, through an L1 that
// holds a few hundred records.  Here a block first stages its halo into LDS
// (one 64-byte record per entry), then every lane walks its row reading LDS
// at the bond's slot (Dev::bk, which already points dead bonds at the row's
// own slot, so the loops carry no broken-bond test).  Slots follow the
// particles' bank colours, so the 16 lanes that read together mostly hit 16
// different bank groups.  Staged per slot (the F-pass in one LDS array per
// field, the force pass in one record per slot, see hslot):
//   xv  : X - anchor in FP32 (3), V                              16 B
//   u01 : u_x, u_y in FP64                                       16 B
//   F-pass: u2 = u_z in FP64                                      8 B
//   force pass: u2c = u_z in FP64, A_zz (NOSB), c'               16 B
//   force pass, NOSB: a0 = A (0..3), a1 = A (4..7)               32 B
// where c' = c G_s / 2 (NOSB) or c / 2 (bond form).  The anchor is an FP64
// point at the block's centre, so |X - anchor| is bounded by the block's
// extent plus the horizon and the staged bond vector xv_j - xv_i differs from
// the FP64 difference by that extent's FP32 rounding (about 1e-7 of the
// horizon).  Displacement differences are taken in FP64 and rounded per
// bond: an FP32 u - u_anchor would round once per particle, by the span of
// u over the halo rather than over a bond, and that per-particle error put a
// floor of a few 1e-6 under the relative residual of adaptive DR.
typedef float pdx_q2 __attribute__((ext_vector_type(2)));

// field offsets (bytes) and bytes per slot of the staged halo
template <int FORCE, int FORM>
struct HaloOff {
    static constexpr int xv = 0, u01 = 16, u2 = 32, a0 = 48, a1 = 64;
    static constexpr int bytes = !FORCE ? 40 : (FORM == 0 ? 80 : 48);
};

// start of field `off`'s array in the F-pass layout (one array per field)
template <typename T>
__device__ __forceinline__ T *lds_at(pdx_q4 *base, int off, int cap)
{
    return reinterpret_cast<T *>(reinterpret_cast<char *>(base) + (size_t)off * cap);
}

// byte address of field f (its offset above) of LDS slot s.  The force pass
// keeps one record of HaloOff::bytes per slot, so a bond needs one address and
// the fields are immediate offsets; 80 and 48 bytes are 20 and 12 dwords, odd
// multiples of 4, so 16 consecutive slots start in 16 different 4-bank groups
// and the colour-slot layout stays conflict-free.  The F-pass keeps one array
// per field (40 bytes per slot in all), which leaves room for two blocks per
// CU.
template <int FORCE, int FORM>
__device__ __forceinline__ char *hslot(pdx_q4 *base, int s, int f, int cap)
{
    char *b = reinterpret_cast<char *>(base);
    if (FORCE) return b + s * HaloOff<1, FORM>::bytes + f;
    return b + f * cap + s * (f == HaloOff<0, 0>::u2 ? 8 : 16);
}

// element at a 32-bit byte offset from a base that is the same for the whole
// wave: the load takes the base from SGPRs and the offset from one VGPR, with
// no 64-bit address arithmetic per lane.  The halo tables are built only when
// every offset used this way fits in 32 bits (build_halo: particle records
// and halo lists).  The row loops keep a per-lane pointer instead, whose
// slot loads differ by immediate offsets.
template <typename T, typename B>
__device__ __forceinline__ T at32(const B *base, uint32_t byte_off)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byte_off);
}

// x of lane P_k of this lane's quad (k = lane & 3), by DPP quad permutation;
// every lane of the quad must be active
template <int P0, int P1, int P2, int P3>
__device__ __forceinline__ float dpp_quad(float x)
{
    constexpr int ctrl = P0 | (P1 << 2) | (P2 << 4) | (P3 << 6);
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), ctrl, 0xF, 0xF, false));
}

// stage the halo of block lb with NT threads (ends with a block barrier).
// Four lanes per entry: lane L of a wave loads the 16-byte quarter L & 3 of
// entry 16 k + (L >> 2), so one load instruction covers 16 consecutive halo
// entries (mostly consecutive records, as the halo is in index order); each
// lane converts its own quarter and writes it to the entry's LDS slot.
template <int FORCE, int FORM, int NT>
__device__ __forceinline__ void halo_stage(const Dev &d, int ubuf, int lb, pdx_q4 *lds)
{
    using O = HaloOff<FORCE, FORM>;
    constexpr int NW = NT / 64;
    // 16-entry groups per wave in flight: the force pass's 112 bytes per entry
    // are fetched best two groups at a time, the F-pass's 64 four at a time
    // with 1024 threads per block and eight with fewer
    constexpr int U = FORCE ? 2 : (NT >= 1024 ? 4 : 8);
    const int cap = d.halo_cap;
    const int h0 = d.hl_ptr[lb], nh = d.hl_ptr[lb + 1] - h0;
    const pdx_d4 an = d.anchor[lb];
    const pdx_q4 *__restrict__ Xq = reinterpret_cast<const pdx_q4 *>(d.xu[ubuf]);
    const pdx_q4 *__restrict__ Aq = reinterpret_cast<const pdx_q4 *>(d.AR);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int q = lane & 3, sub = lane >> 2;
    const int qa = q < 3 ? q : 2;  // A records have 3 quarters
    const float csc = FORM == 0 ? (float)(0.5 * d.G_s) : 0.5f;
    for (int g0 = wv * 16; g0 < nh; g0 += U * NW * 16) {
        int j[U], sl[U];
        bool ok[U];
        pdx_q4 r[U], a[U];
#pragma unroll
        for (int v = 0; v < U; ++v) {
            const int e = g0 + v * NW * 16 + sub;
            ok[v] = e < nh;
            const uint32_t he = (uint32_t)(h0 + (ok[v] ? e : 0));
            j[v] = at32<int32_t>(d.hl, he * 4u);
            sl[v] = ok[v] ? (int)at32<uint16_t>(d.hs, he * 2u) : 0;
        }
#pragma unroll
        for (int v = 0; v < U; ++v) {
            r[v] = at32<pdx_q4>(Xq, (uint32_t)j[v] * 64u + (uint32_t)q * 16u);
            if (FORCE && FORM == 0) a[v] = at32<pdx_q4>(Aq, (uint32_t)j[v] * 48u + (uint32_t)qa * 16u);
        }
        if constexpr (FORCE && FORM == 0) {
            // NOSB record: the quad's lanes trade three values with DPP and
            // each lane writes one 16-byte field (lane 0 also a0): xv from
            // lane 0 (its x, y and lane 1's z, V), a1 from lane 1, u01 from
            // lane 2, u2 from lane 3 (its u_z and lane 2's A_zz, lane 1's c')
#pragma unroll
            for (int v = 0; v < U; ++v) {
                const PdxD2 x = __builtin_bit_cast(PdxD2, r[v]);
                const float t0 = (float)(x.a - (q == 0 ? an.x : an.z));  // lane 0: x - a_x; lane 1: z - a_z
                const float t1 = (float)(x.b - an.y);                     // lane 0: y - a_y
                const float m1 = dpp_quad<1, 1, 1, 2>(q == 1 ? t0 : a[v].x);  // lane 0: z; lane 3: A_zz
                const float mv = dpp_quad<1, 1, 1, 1>(r[v].z);                 // lane 0: V
                const float mc = dpp_quad<1, 1, 1, 1>(r[v].w * csc);           // lane 3: c'
                pdx_q4 out;
                if (q == 0) out = pdx_q4{t0, t1, m1, mv};
                else if (q == 1) out = a[v];
                else if (q == 2) out = r[v];
                else out = pdx_q4{r[v].x, r[v].y, m1, mc};
                const int off = q == 0 ? O::xv : (q == 1 ? O::a1 : (q == 2 ? O::u01 : O::u2));
                if (ok[v]) {
                    char *b = hslot<FORCE, FORM>(lds, sl[v], 0, cap);
                    *reinterpret_cast<pdx_q4 *>(b + off) = out;
                    if (q == 0) *reinterpret_cast<pdx_q4 *>(b + O::a0) = a[v];
                }
            }
            continue;
        }
        if constexpr (!FORCE) {
            // F-pass arrays: lane 0 of the quad writes xv (its x, y and lane
            // 1's z, V), lane 2 u01 (its u_x, u_y), lane 3 u2 (its u_z)
#pragma unroll
            for (int v = 0; v < U; ++v) {
                const PdxD2 x = __builtin_bit_cast(PdxD2, r[v]);
                const float t0 = (float)(x.a - (q == 0 ? an.x : an.z));  // lane 0: x - a_x; lane 1: z - a_z
                const float t1 = (float)(x.b - an.y);                     // lane 0: y - a_y
                const float mz = dpp_quad<1, 1, 1, 1>(t0);
                const float mv = dpp_quad<1, 1, 1, 1>(r[v].z);
                if (ok[v] && q != 1) {
                    const int sv = sl[v];
                    if (q == 3) {
                        *reinterpret_cast<pdx_q2 *>(hslot<FORCE, FORM>(lds, sv, O::u2, cap)) = r[v].xy;
                    } else {
                        const pdx_q4 out = q == 0 ? pdx_q4{t0, t1, mz, mv} : r[v];
                        *reinterpret_cast<pdx_q4 *>(hslot<FORCE, FORM>(lds, sv, q == 0 ? O::xv : O::u01, cap)) = out;
                    }
                }
            }
            continue;
        }
#pragma unroll
        for (int v = 0; v < U; ++v) {
            if (!ok[v]) continue;
            const int sv = sl[v];
            if (q == 0) {  // X_x, X_y
                const PdxD2 x = __builtin_bit_cast(PdxD2, r[v]);
                *reinterpret_cast<pdx_q2 *>(hslot<FORCE, FORM>(lds, sv, O::xv, cap)) =
                    pdx_q2{(float)(x.a - an.x), (float)(x.b - an.y)};
                if (FORCE && FORM == 0) *reinterpret_cast<pdx_q4 *>(hslot<FORCE, FORM>(lds, sv, O::a0, cap)) = a[v];
            } else if (q == 1) {  // X_z, V, c
                const double x2 = __builtin_bit_cast(double, r[v].xy);
                *reinterpret_cast<pdx_q2 *>(hslot<FORCE, FORM>(lds, sv, O::xv, cap) + 8) =
                    pdx_q2{(float)(x2 - an.z), r[v].z};
                if (FORCE) *reinterpret_cast<float *>(hslot<FORCE, FORM>(lds, sv, O::u2, cap) + 12) = r[v].w * csc;
                if (FORCE && FORM == 0) *reinterpret_cast<pdx_q4 *>(hslot<FORCE, FORM>(lds, sv, O::a1, cap)) = a[v];
            } else if (q == 2) {  // u_x, u_y
                *reinterpret_cast<pdx_q4 *>(hslot<FORCE, FORM>(lds, sv, O::u01, cap)) = r[v];
                if (FORCE && FORM == 0)  // A_zz
                    *reinterpret_cast<float *>(hslot<FORCE, FORM>(lds, sv, O::u2, cap) + 8) = a[v].x;
            } else {  // u_z
                *reinterpret_cast<pdx_q2 *>(hslot<FORCE, FORM>(lds, sv, O::u2, cap)) = r[v].xy;
            }
        }
    }
    __syncthreads();
}

// bond vector and displacement difference of halo entry (xj, uj01, uj2)
// against the own entry (xo, uo01, uo2)
__device__ __forceinline__ void halo_bond(const pdx_q4 &xj, const pdx_q4 &uj01, const pdx_q2 &uj2, const pdx_q4 &xo,
                                          const PdxD2 &uo01, double uo2, float (&xi)[3], float (&du)[3])
{
    xi[0] = xj.x - xo.x;
    xi[1] = xj.y - xo.y;
    xi[2] = xj.z - xo.z;
    const PdxD2 a = __builtin_bit_cast(PdxD2, uj01);
    const double b = __builtin_bit_cast(double, uj2);
    du[0] = (float)(a.a - uo01.a);
    du[1] = (float)(a.b - uo01.b);
    du[2] = (float)(b - uo2);
}

// Split rows: S waves share a block's halo and rows; part p (threadIdx.x /
// PDX_BLOCK) takes the bond slots whose group index is p mod S, and the
// partial sums meet in LDS (reusing the halo area) in part order before the
// per-particle finish.  More waves per halo hide the LDS and FP64 latencies
// when the halo leaves room for one block per CU.
template <int NV, int S>
__device__ __forceinline__ void split_sum(pdx_q4 *lds, double (&acc)[NV])
{
    if constexpr (S > 1) {
        double *sp = reinterpret_cast<double *>(lds);  // [S - 1][NV][PDX_BLOCK]
        const int t = threadIdx.x % PDX_BLOCK, part = threadIdx.x / PDX_BLOCK;
        __syncthreads();  // every part is done reading the halo
        if (part > 0)
#pragma unroll
            for (int q = 0; q < NV; ++q) sp[((part - 1) * NV + q) * PDX_BLOCK + t] = acc[q];
        __syncthreads();
        if (part == 0)
            for (int p = 1; p < S; ++p)
#pragma unroll
                for (int q = 0; q < NV; ++q) acc[q] += sp[((p - 1) * NV + q) * PDX_BLOCK + t];
    }
}

// RED: the grid has ncase more blocks, which run the adaptive-DR case
// reduction of the force pass before this F-pass (its input buffer red_ubuf)
// instead of a launch of their own; the F-pass blocks may then read a case's
// converged flag before or after that update (a converged case's force pass
// is skipped either way).  MM 0: sum_j wx summed here; 1: read from d.m64;
// 2: summed here and written to d.m64, nothing else (pdx_build_m64)
template <int D, int MODEL, int G, int S, bool RED = false, int MM = 0>
__global__ __launch_bounds__(PDX_BLOCK * S) void k_fpass_h(Dev d, int ubuf, int solve_mode, int red_ubuf)
{
    extern __shared__ pdx_q4 hlds[];
    using O = HaloOff<0, 0>;
    if constexpr (RED) {
        if ((int)blockIdx.x >= d.nblk) {
            case_reduce_block<1, PDX_BLOCK * S>(d, (int)blockIdx.x - d.nblk, red_ubuf, solve_mode);
            return;
        }
    }
    const int lb = logical_block_n(d.nblk);
    if (solve_mode &&
        __hip_atomic_load(&d.sc[d.case_of_blk[lb]].converged, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        return;
    if (lb * PDX_BLOCK >= d.ni) return;  // grid padding blocks (whole block)
    halo_stage<0, 0, PDX_BLOCK * S>(d, ubuf, lb, hlds);
    const int cap = d.halo_cap;
    const pdx_q4 *xv = lds_at<pdx_q4>(hlds, O::xv, cap);
    const pdx_q4 *u01 = lds_at<pdx_q4>(hlds, O::u01, cap);
    const pdx_q2 *u2 = lds_at<pdx_q2>(hlds, O::u2, cap);
    const int t = threadIdx.x % PDX_BLOCK, lane = t & 63, part = threadIdx.x / PDX_BLOCK;
    const int i = lb * PDX_BLOCK + t;
    const KRec kr = d.KR[i];
    const bool act = (kr.flags & PDX_FL_ACTIVE) != 0;
    const int so = d.hs[d.hl_ptr[lb] + t];
    const pdx_q4 xo = xv[so];
    const PdxD2 uo01 = __builtin_bit_cast(PdxD2, u01[so]);
    const double uo2 = __builtin_bit_cast(double, u2[so]);
    const float invh = (float)d.inv_h;
    // sum_j w V_j (u_j - u_i) (x) xi as sum_j u_j (x) wx_j - u_i (x) sum_j wx_j
    // with wx = w V_j xi rounded to FP32 once and both sums in FP64: the
    // rounding of wx is shared by the two sums, so their difference is the
    // FP64 sum of (u_j - u_i) (x) wx without rounding u_j - u_i (Sd[D * D + c]
    // holds sum_j wx_c)
    constexpr int NS = MM == 1 ? D * D : D * D + D;
    double Sd[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) Sd[q] = 0.0;
    const int64_t s0 = d.sl_ptr[i >> 6];
    const int width = (int)((d.sl_ptr[(i >> 6) + 1] - s0) / PDX_SLICE);
    const uint16_t *__restrict__ bp = d.bk + s0 + lane;
    auto group = [&](const uint32_t (&e)[G]) {
        pdx_q4 xj[G], uj01[G];
        pdx_q2 uj2[G];
#pragma unroll
        for (int q = 0; q < G; ++q) {
            xj[q] = xv[e[q]];
            uj01[q] = u01[e[q]];
            uj2[q] = u2[e[q]];
        }
#pragma unroll
        for (int q = 0; q < G; ++q) {
            const float xi[3] = {xj[q].x - xo.x, xj[q].y - xo.y, xj[q].z - xo.z};
            const PdxD2 a = __builtin_bit_cast(PdxD2, uj01[q]);
            const double uj[3] = {a.a, a.b, __builtin_bit_cast(double, uj2[q])};
            fpass_bond_u<D, MM != 1>(xi, uj, xj[q].w, invh, Sd);
        }
    };
    halo_row_loop<G, S>(bp, width, part, group);
    pdx_d4 mi{};
    if (MM == 1 && part == 0 && act) mi = d.m64[i];
    split_sum<NS, S>(hlds, Sd);
    if (part == 0 && act) {
        if constexpr (MM == 2) {
            d.m64[i] = d4(Sd[NS - D], Sd[NS - D + 1], D == 3 ? Sd[NS - 1] : 0.0);
        } else {
            double Sf[D * D + D];
#pragma unroll
            for (int q = 0; q < D * D; ++q) Sf[q] = Sd[q];
            const double mv[3] = {mi.x, mi.y, mi.z};
#pragma unroll
            for (int c = 0; c < D; ++c) Sf[D * D + c] = MM == 1 ? mv[c] : Sd[(D * D + c) % NS];
            const double ui[3] = {uo01.a, uo01.b, uo2};
            double Hd[D * D];
            fpass_sum_to_h<D>(Sf, ui, Hd);
            fpass_finish<D, MODEL, false, false>(d, i, kr, Hd);
        }
    }
}

template <int D, int FORM, int UPD, int G, int S>
__global__ __launch_bounds__(PDX_BLOCK * S) void k_force_h(Dev d, int ubuf, int solve_mode, double dt, double visc)
{
    extern __shared__ pdx_q4 hlds[];
    using O = HaloOff<1, FORM>;
    const int lb = logical_block();
    const int cs = d.case_of_blk[lb];
    if (solve_mode && d.sc[cs].converged) return;
    double red[3] = {0.0, 0.0, 0.0};
    if (lb * PDX_BLOCK < d.ni) {  // grid padding blocks only join the reduction
        halo_stage<1, FORM, PDX_BLOCK * S>(d, ubuf, lb, hlds);
        // slot l's record: xv at 0, u01 at 16, u2c at 32, a0 at 48, a1 at 64
        const char *hb = reinterpret_cast<const char *>(hlds);
        // (slots are below 2^15, so the 24-bit multiply gives the offset)
        auto rec = [&](int l, int f) {
            return *reinterpret_cast<const pdx_q4 *>(hb + __umul24((unsigned)l, (unsigned)O::bytes) + f);
        };
        const int t = threadIdx.x % PDX_BLOCK, lane = t & 63, part = threadIdx.x / PDX_BLOCK;
        const int i = lb * PDX_BLOCK + t;
        const KRec kr = d.KR[i];
        const bool act = (kr.flags & PDX_FL_ACTIVE) != 0;
        // the own record and A_i come from the staged halo: force_finish
        // reads u of pi, not X
        const int so = d.hs[d.hl_ptr[lb] + t];
        const pdx_q4 xo = rec(so, O::xv);
        const PdxD2 uo01 = __builtin_bit_cast(PdxD2, rec(so, O::u01));
        const pdx_q4 uco = rec(so, O::u2);
        const double uo2 = __builtin_bit_cast(double, uco.xy);
        XRec pi{};
        pi.u[0] = uo01.a;
        pi.u[1] = uo01.b;
        pi.u[2] = uo2;
        ARec ai{};
        if (FORM == 0) {
            const pdx_q4 p0 = rec(so, O::a0), p1 = rec(so, O::a1);
            ai.A[0] = p0.x; ai.A[1] = p0.y; ai.A[2] = p0.z; ai.A[3] = p0.w;
            ai.A[4] = p1.x; ai.A[5] = p1.y; ai.A[6] = p1.z; ai.A[7] = p1.w;
            ai.A[8] = uco.z;
        }
        const float invh = (float)d.inv_h;
        double fa[3] = {0.0, 0.0, 0.0};
        const int64_t s0 = d.sl_ptr[i >> 6];
        const int width = (int)((d.sl_ptr[(i >> 6) + 1] - s0) / PDX_SLICE);
        const uint16_t *__restrict__ bp = d.bk + s0 + lane;
        auto group = [&](const uint32_t (&e)[G]) {
            pdx_q4 xj[G], uj01[G], ujc[G], aj0[G], aj1[G];
#pragma unroll
            for (int q = 0; q < G; ++q) {
                const int l = (int)e[q];
                xj[q] = rec(l, O::xv);
                uj01[q] = rec(l, O::u01);
                ujc[q] = rec(l, O::u2);
                if (FORM == 0) {
                    aj0[q] = rec(l, O::a0);
                    aj1[q] = rec(l, O::a1);
                }
            }
            // NOSB: the group's terms are summed in FP32 (each is a difference
            // term, small against the particle's sum) and the group sums in
            // FP64; the bond form adds every term in FP64
            float fg[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int q = 0; q < G; ++q) {
                float xi[3], du[3];
                halo_bond(xj[q], uj01[q], ujc[q].xy, xo, uo01, uo2, xi, du);
                if constexpr (FORM == 0) {
                    const float aj[9] = {aj0[q].x, aj0[q].y, aj0[q].z, aj0[q].w, aj1[q].x,
                                         aj1[q].y, aj1[q].z, aj1[q].w, ujc[q].z};
                    force_bond_v<D, FORM, true, float>(xi, du, xj[q].w, uco.w, ujc[q].w, ai.A, aj, invh, 1.f, fg);
                } else {
                    force_bond_v<D, FORM, true, double>(xi, du, xj[q].w, uco.w, ujc[q].w, ai.A, nullptr, invh, 1.f,
                                                        fa);
                }
            }
            if constexpr (FORM == 0) {
#pragma unroll
                for (int r = 0; r < 3; ++r) fa[r] += (double)fg[r];
            }
        };
        halo_row_loop<G, S>(bp, width, part, group);
        // part 0 finishes its rows first (a SIMD issues to its oldest wave
        // first) and loads the update's inputs while the other parts end
        FinishIn fin{};
        if (part == 0 && act) fin = finish_in<UPD>(d, i, kr);
        split_sum<3, S>(hlds, fa);
        if (part == 0 && act) force_finish<D, FORM, UPD>(d, i, kr, pi, ai, ubuf, cs, dt, visc, fa, red, fin);
    }
    block_partials<3, PDX_BLOCK * S, PDX_BLOCK / 64>(d, lb, red);  // part 0's waves
}

// ---------------------------------------------------------------------------
// lattice kernels (3-D NOSB on the 122-site stencil, one particle volume V).
// The same halo staging and split rows as k_fpass_h / k_force_h, but a row
// walks the PDX_NST stencil entries of bks instead of its bond list, and
// every entry k is code of its own, compiled with o_k: the bond vector, its
// length and the influence weight are constants, the components of o_k that
// are zero cost nothing, and the per-context weights come from the LatC
// kernel argument (scalar registers).  Part p of S takes the entries
// k = p, p + S, ... (each part gets a share of every class).  An entry
// without a live bond reads the row's own slot: its F-pass term u_i (x) w o
// cancels against the opposite entry's over the symmetric stencil, and its
// force term is exactly zero (u_j - u_i = 0, A_j - A_i = 0).
// ---------------------------------------------------------------------------
// |o| for |o|^2 = n (the stencil's lengths in lattice units)
constexpr float lat_len(int n)
{
    return n == 1 ? 1.0f : n == 2 ? 1.41421356f : n == 3 ? 1.73205081f : n == 4 ? 2.0f : n == 5 ? 2.23606798f
         : n == 6 ? 2.44948974f : n == 8 ? 2.82842712f : 3.0f;
}

// stage u of block lb's halo for the lattice F-pass into the arrays u01
// (u_x, u_y: 16 B per slot) and u2 (u_z: 8 B per slot): lane L of a wave
// takes entry 64 g + L, loading the record's bytes 32..47 and 48..55
template <int NT>
__device__ __forceinline__ void lat_stage_f(const Dev &d, int ubuf, int lb, pdx_q4 *lds)
{
    constexpr int NW = NT / 64, U = 4;
    const int cap = d.halo_cap;
    const int h0 = d.hl_ptr[lb], nh = d.hl_ptr[lb + 1] - h0;
    const pdx_q4 *__restrict__ Xq = reinterpret_cast<const pdx_q4 *>(d.xu[ubuf]);
    pdx_q4 *u01 = lds;
    double *u2 = reinterpret_cast<double *>(lds + cap);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int g0 = wv * 64; g0 < nh; g0 += U * NW * 64) {
        int j[U], sl[U];
        bool ok[U];
        pdx_q4 a[U];
        pdx_q2 b[U];
#pragma unroll
        for (int v = 0; v < U; ++v) {
            const int e = g0 + v * NW * 64 + lane;
            ok[v] = e < nh;
            const uint32_t he = (uint32_t)(h0 + (ok[v] ? e : 0));
            j[v] = at32<int32_t>(d.hl, he * 4u);
            sl[v] = ok[v] ? (int)at32<uint16_t>(d.hs, he * 2u) : 0;
        }
#pragma unroll
        for (int v = 0; v < U; ++v) {
            a[v] = at32<pdx_q4>(Xq, (uint32_t)j[v] * 64u + 32u);
            b[v] = at32<pdx_q2>(Xq, (uint32_t)j[v] * 64u + 48u);
        }
#pragma unroll
        for (int v = 0; v < U; ++v) {
            if (!ok[v]) continue;
            u01[sl[v]] = a[v];
            u2[sl[v]] = __builtin_bit_cast(double, b[v]);
        }
    }
    __syncthreads();
}

// F-pass terms of stencil pair M (entries 2 M: o, 2 M + 1: -o):
// S[r][c] += (u_+ - u_-)[r] w0 o_c (FP64), the two entries' u_j (x) w0 (+-o)
template <int OC, int N, int C>
__device__ __forceinline__ void lat_facc(const double (&du)[3], const LatC &lc, double (&S)[9])
{
    if constexpr (OC != 0) {
        constexpr int m = OC < 0 ? -OC : OC;
        const double w = m == 1 ? lc.w1[N] : (m == 2 ? lc.w2[N] : lc.w3[N]);
        const double ws = OC < 0 ? -w : w;
#pragma unroll
        for (int r = 0; r < 3; ++r) S[r * 3 + C] = fma(du[r], ws, S[r * 3 + C]);
    }
}

template <int M>
__device__ __forceinline__ void lat_fpair(const uint16_t *__restrict__ bp, const pdx_q4 *u01, const double *u2,
                                          const LatC &lc, double (&S)[9])
{
    constexpr int K = 2 * M;
    constexpr int o0 = kLat.o[K][0], o1 = kLat.o[K][1], o2 = kLat.o[K][2];
    static_assert(kLat.o[K + 1][0] == -o0 && kLat.o[K + 1][1] == -o1 && kLat.o[K + 1][2] == -o2, "stencil pair");
    constexpr int n = o0 * o0 + o1 * o1 + o2 * o2;
    const uint32_t ep = bp[K * PDX_SLICE], em = bp[(K + 1) * PDX_SLICE];
    const PdxD2 a = __builtin_bit_cast(PdxD2, u01[ep]), b = __builtin_bit_cast(PdxD2, u01[em]);
    const double du[3] = {a.a - b.a, a.b - b.b, u2[ep] - u2[em]};
    lat_facc<o0, n, 0>(du, lc, S);
    lat_facc<o1, n, 1>(du, lc, S);
    lat_facc<o2, n, 2>(du, lc, S);
}

template <int S, int P, size_t... I>
__device__ __forceinline__ void lat_fpart(const uint16_t *__restrict__ bp, const pdx_q4 *u01, const double *u2,
                                          const LatC &lc, double (&Sd)[9], std::index_sequence<I...>)
{
    (lat_fpair<P + (int)I * S>(bp, u01, u2, lc, Sd), ...);
}

// the pairs of part P of S: P, P + S, ...
template <int S, int P>
__device__ __forceinline__ void lat_fpart_p(const uint16_t *bp, const pdx_q4 *u01, const double *u2, const LatC &lc,
                                            double (&Sd)[9])
{
    lat_fpart<S, P>(bp, u01, u2, lc, Sd, std::make_index_sequence<(PDX_NST / 2 - P + S - 1) / S>{});
}

// RED as in k_fpass_h.  FR: A goes into the lattice force pass's records
// d.fr (with u / h and c') instead of d.AR.
template <int MODEL, int S, bool RED = false, bool FR = false>
__global__ __launch_bounds__(PDX_BLOCK * S) void k_fpass_l(Dev d, LatC lc, int ubuf, int solve_mode, int red_ubuf)
{
    extern __shared__ pdx_q4 hlds[];
    if constexpr (RED) {
        if ((int)blockIdx.x >= d.nblk) {
            case_reduce_block<1, PDX_BLOCK * S>(d, (int)blockIdx.x - d.nblk, red_ubuf, solve_mode);
            return;
        }
    }
    const int lb = logical_block_n(d.nblk);
    if (solve_mode &&
        __hip_atomic_load(&d.sc[d.case_of_blk[lb]].converged, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        return;
    if (lb * PDX_BLOCK >= d.ni) return;  // grid padding blocks (whole block)
    const int t = threadIdx.x % PDX_BLOCK, part = threadIdx.x / PDX_BLOCK;
    const int i = lb * PDX_BLOCK + t;
    // the finish's per-row inputs, loaded by part 0 while the block stages
    KRec kr{};
    double Kp[9];
    if (part == 0) {
        kr = d.KR[i];
#pragma unroll
        for (int q = 0; q < 9; ++q) Kp[q] = d.Kinv64[(int64_t)i * 9 + q];
    }
    lat_stage_f<PDX_BLOCK * S>(d, ubuf, lb, hlds);
    const int cap = d.halo_cap;
    const pdx_q4 *u01 = hlds;
    const double *u2 = reinterpret_cast<const double *>(hlds + cap);
    const uint16_t *__restrict__ bp = d.bks + lat_at(i, 0);
    double Sd[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) Sd[q] = 0.0;
    if (part == 0) lat_fpart_p<S, 0>(bp, u01, u2, lc, Sd);
    if constexpr (S > 1) {
        if (part == 1) lat_fpart_p<S, 1>(bp, u01, u2, lc, Sd);
    }
    if constexpr (S > 2) {
        if (part == 2) lat_fpart_p<S, 2>(bp, u01, u2, lc, Sd);
        if (part == 3) lat_fpart_p<S, 3>(bp, u01, u2, lc, Sd);
    }
    // the own u for the force pass's record, read before the split sums
    // reuse the halo's LDS
    PdxD2 uo{0.0, 0.0};
    double uz = 0.0;
    if (FR && part == 0) {
        const int so = d.hs[d.hl_ptr[lb] + t];
        uo = __builtin_bit_cast(PdxD2, u01[so]);
        uz = u2[so];
    }
    split_sum<9, S>(hlds, Sd);
    if (part != 0) return;
    if (!(kr.flags & PDX_FL_ACTIVE)) return;
    // sum_j w V_j (u_j - u_i) (x) xi = V h sum_k u_j(k) (x) w0 o_k: the own
    // slot's u_i (x) w0 o_k of the entries without a live bond and the live
    // entries' u_i (x) w0 o_k add up to u_i (x) sum_k w0 o_k = 0 (the sum runs
    // over the pairs, as (u_+ - u_-) (x) w0 o)
    double Hd[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) Hd[q] = Sd[q] * lc.fscale;
    if constexpr (!FR) {
        fpass_finish_k<3, MODEL, false, false>(d, i, kr, Hd, Kp);
    } else {
        const float cc = d.xu[ubuf][i].c * (float)(0.5 * d.G_s);
        const ARec ar = fpass_finish_k<3, MODEL, false, false, false>(d, i, kr, Hd, Kp);
        pdx_q4 *o = reinterpret_cast<pdx_q4 *>(d.fr + i);
        const double ih = lc.inv_h;
        o[0] = __builtin_bit_cast(pdx_q4, PdxD2{uo.a * ih, uo.b * ih});
        const pdx_q2 z = __builtin_bit_cast(pdx_q2, uz * ih);
        o[1] = pdx_q4{z.x, z.y, ar.A[8], cc};
        o[2] = pdx_q4{ar.A[0], ar.A[1], ar.A[2], ar.A[3]};
        o[3] = pdx_q4{ar.A[4], ar.A[5], ar.A[6], ar.A[7]};
    }
}

// stage block lb's halo for the lattice force pass, one 80-byte LDS record
// per slot: u / h in FP64 (x, y at 0; z at 16), A_zz at 24, c' = c G_s / 2
// at 28, A(0..3) at 32, A(4..7) at 48 -- the FRec layout.  Four lanes per
// entry as in halo_stage.  FR: lane q copies quarter q of the entry's FRec.
// Else lane q loads quarter q of the XRec and quarter min(q, 2) of the ARec;
// lane 3 takes A_zz from lane 2 and c' from lane 1 by DPP.
template <int NT, bool FR>
__device__ __forceinline__ void lat_stage_g(const Dev &d, int ubuf, int lb, pdx_q4 *lds, double ih, int sb)
{
    constexpr int NW = NT / 64, U = FR ? 8 : 2;
    const int h0 = d.hl_ptr[lb], nh = d.hl_ptr[lb + 1] - h0;
    const pdx_q4 *__restrict__ Xq = reinterpret_cast<const pdx_q4 *>(d.xu[ubuf]);
    const pdx_q4 *__restrict__ Aq = reinterpret_cast<const pdx_q4 *>(d.AR);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int q = lane & 3, sub = lane >> 2;
    const int qa = q < 3 ? q : 2;
    const float csc = (float)(0.5 * d.G_s);
    char *hb = reinterpret_cast<char *>(lds);
    for (int g0 = wv * 16; g0 < nh; g0 += U * NW * 16) {
        int j[U], sl[U];
        bool ok[U];
        pdx_q4 r[U], a[U];
#pragma unroll
        for (int v = 0; v < U; ++v) {
            const int e = g0 + v * NW * 16 + sub;
            ok[v] = e < nh;
            const uint32_t he = (uint32_t)(h0 + (ok[v] ? e : 0));
            j[v] = at32<int32_t>(d.hl, he * 4u);
            sl[v] = ok[v] ? (int)at32<uint16_t>(d.hs, he * 2u) : 0;
        }
        if constexpr (FR) {
            const pdx_q4 *__restrict__ Fq = reinterpret_cast<const pdx_q4 *>(d.fr);
#pragma unroll
            for (int v = 0; v < U; ++v) r[v] = at32<pdx_q4>(Fq, (uint32_t)j[v] * 64u + (uint32_t)q * 16u);
#pragma unroll
            for (int v = 0; v < U; ++v)
                if (ok[v]) *reinterpret_cast<pdx_q4 *>(hb + sl[v] * sb + q * 16) = r[v];
            continue;
        }
#pragma unroll
        for (int v = 0; v < U; ++v) {
            r[v] = at32<pdx_q4>(Xq, (uint32_t)j[v] * 64u + (uint32_t)q * 16u);
            a[v] = at32<pdx_q4>(Aq, (uint32_t)j[v] * 48u + (uint32_t)qa * 16u);
        }
#pragma unroll
        for (int v = 0; v < U; ++v) {
            const float azz = dpp_quad<0, 1, 2, 2>(a[v].x);      // lane 3: lane 2's A_zz
            const float cc = dpp_quad<0, 1, 2, 1>(r[v].w * csc);  // lane 3: lane 1's c'
            const PdxD2 x = __builtin_bit_cast(PdxD2, r[v]);
            pdx_q4 out;
            int off;
            if (q == 0) {
                out = a[v];
                off = 32;
            } else if (q == 1) {
                out = a[v];
                off = 48;
            } else if (q == 2) {
                out = __builtin_bit_cast(pdx_q4, PdxD2{x.a * ih, x.b * ih});
                off = 0;
            } else {
                const pdx_q2 z = __builtin_bit_cast(pdx_q2, x.a * ih);
                out = pdx_q4{z.x, z.y, azz, cc};
                off = 16;
            }
            if (ok[v]) *reinterpret_cast<pdx_q4 *>(hb + sl[v] * sb + off) = out;
        }
    }
    __syncthreads();
}

// force terms of stencil pair M (entries 2 M: o and 2 M + 1: -o, the
// neighbours at x_i + h o and x_i - h o) into the FP32 group sum fg.
// Stabilisation, per entry, in lattice units (eta = +-o + (u_j - u_i) / h,
// dn = |eta|^2 - |o|^2): w0 / |o| (c'_i + c'_j) dn / ((|eta| + |o|) |eta|) eta.
// Correspondence: the two entries' h w0 (A_j - A_i) (+-o) add up to
// h w0 (A_+ - A_-) o, A_i cancelling (still a difference of neighbouring
// tensors, so its FP32 products are rounded relative to that difference).
// The caller multiplies the particle's sum by V.  |eta|^2 is floored at 1e-12
// (a bond squeezed to 1e-6 h has an undefined direction; the halo kernels
// drop its stabilisation term instead).
template <int O0, int O1, int O2, int N>
__device__ __forceinline__ void lat_stab(const pdx_q4 &f0, const pdx_q4 &f1, const LatC &lc, const double (&ui)[3],
                                         const float (&ciw)[10], float (&fg)[3])
{
    constexpr int o[3] = {O0, O1, O2};
    constexpr float L = lat_len(N);
    const PdxD2 u01 = __builtin_bit_cast(PdxD2, f0);
    const double u2 = __builtin_bit_cast(double, f1.xy);
    const float du[3] = {(float)(u01.a - ui[0]), (float)(u01.b - ui[1]), (float)(u2 - ui[2])};
    float dn = 0.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) dn = fmaf(du[c], o[c] != 0 ? du[c] + (float)(2 * o[c]) : du[c], dn);
    const float E2 = fmaxf(dn + (float)N, 1e-12f);
    const float el = bond_sqrt(E2);
    const float gam = dn * bond_rcp(fmaf(L, el, E2));          // dn / ((|eta| + L) |eta|)
    const float wg = fmaf(f1.w, lc.wl[N], ciw[N]) * gam;       // (c'_i + c'_j) w0 / sqrt(n)
#pragma unroll
    for (int r = 0; r < 3; ++r) fg[r] = fmaf(wg, o[r] != 0 ? du[r] + (float)o[r] : du[r], fg[r]);
}

template <int M>
__device__ __forceinline__ void lat_gpair(const uint16_t *__restrict__ bp, const char *hb, int sb, const LatC &lc,
                                          const double (&ui)[3], const float (&ciw)[10], float (&fg)[3])
{
    constexpr int K = 2 * M;
    constexpr int o0 = kLat.o[K][0], o1 = kLat.o[K][1], o2 = kLat.o[K][2];
    static_assert(kLat.o[K + 1][0] == -o0 && kLat.o[K + 1][1] == -o1 && kLat.o[K + 1][2] == -o2, "stencil pair");
    constexpr int o[3] = {o0, o1, o2};
    constexpr int n = o0 * o0 + o1 * o1 + o2 * o2;
    const unsigned ep = bp[K * PDX_SLICE], em = bp[(K + 1) * PDX_SLICE];
    const char *rp = hb + __umul24(ep, (unsigned)sb), *rm = hb + __umul24(em, (unsigned)sb);
    const pdx_q4 p0 = *reinterpret_cast<const pdx_q4 *>(rp), p1 = *reinterpret_cast<const pdx_q4 *>(rp + 16);
    const pdx_q4 p2 = *reinterpret_cast<const pdx_q4 *>(rp + 32), p3 = *reinterpret_cast<const pdx_q4 *>(rp + 48);
    const pdx_q4 m0 = *reinterpret_cast<const pdx_q4 *>(rm), m1 = *reinterpret_cast<const pdx_q4 *>(rm + 16);
    const pdx_q4 m2 = *reinterpret_cast<const pdx_q4 *>(rm + 32), m3 = *reinterpret_cast<const pdx_q4 *>(rm + 48);
    lat_stab<o0, o1, o2, n>(p0, p1, lc, ui, ciw, fg);
    lat_stab<-o0, -o1, -o2, n>(m0, m1, lc, ui, ciw, fg);
    const float Ap[9] = {p2.x, p2.y, p2.z, p2.w, p3.x, p3.y, p3.z, p3.w, p1.z};
    const float Am[9] = {m2.x, m2.y, m2.z, m2.w, m3.x, m3.y, m3.z, m3.w, m1.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (o[c] == 0) continue;
        const int m = o[c] < 0 ? -o[c] : o[c];
        const float cw = m == 1 ? lc.c1[n] : (m == 2 ? lc.c2[n] : lc.c3[n]);
        const float cs = o[c] < 0 ? -cw : cw;
#pragma unroll
        for (int r = 0; r < 3; ++r) fg[r] = fmaf(Ap[r * 3 + c] - Am[r * 3 + c], cs, fg[r]);
    }
}

template <int M>
__device__ __forceinline__ void lat_gpair_if(const uint16_t *bp, const char *hb, int sb, const LatC &lc,
                                             const double (&ui)[3], const float (&ciw)[10], float (&fg)[3])
{
    if constexpr (M < PDX_NST / 2) lat_gpair<M>(bp, hb, sb, lc, ui, ciw, fg);
}

// group g of part P: pairs P + (2 g + m) S, m < 2 (four bonds), summed in
// FP32, the group sum added to the FP64 particle sum
template <int S, int P, int GI, size_t... Q>
__device__ __forceinline__ void lat_ggroup(const uint16_t *bp, const char *hb, int sb, const LatC &lc,
                                           const double (&ui)[3], const float (&ciw)[10], double (&fa)[3],
                                           std::index_sequence<Q...>)
{
    float fg[3] = {0.f, 0.f, 0.f};
    (lat_gpair_if<P + (GI * 2 + (int)Q) * S>(bp, hb, sb, lc, ui, ciw, fg), ...);
#pragma unroll
    for (int r = 0; r < 3; ++r) fa[r] += (double)fg[r];
}

template <int S, int P, size_t... GI>
__device__ __forceinline__ void lat_gpart(const uint16_t *bp, const char *hb, int sb, const LatC &lc,
                                          const double (&ui)[3], const float (&ciw)[10], double (&fa)[3],
                                          std::index_sequence<GI...>)
{
    (lat_ggroup<S, P, (int)GI>(bp, hb, sb, lc, ui, ciw, fa, std::make_index_sequence<2>{}), ...);
}

// the pairs of part P of S: P, P + S, ...
template <int S, int P>
__device__ __forceinline__ void lat_gpart_p(const uint16_t *bp, const char *hb, int sb, const LatC &lc,
                                            const double (&ui)[3], const float (&ciw)[10], double (&fa)[3])
{
    constexpr int np = (PDX_NST / 2 - P + S - 1) / S;
    lat_gpart<S, P>(bp, hb, sb, lc, ui, ciw, fa, std::make_index_sequence<(np + 1) / 2>{});
}

template <int UPD, int S, bool FR>
__global__ __launch_bounds__(PDX_BLOCK * S) void k_force_l(Dev d, LatC lc, int ubuf, int solve_mode, double dt,
                                                           double visc, int lds_bytes)
{
    extern __shared__ pdx_q4 hlds[];
    const int lb = logical_block();
    const int cs = d.case_of_blk[lb];
    if (solve_mode && d.sc[cs].converged) return;
    double red[3] = {0.0, 0.0, 0.0};
    if (lb * PDX_BLOCK < d.ni) {  // grid padding blocks only join the reduction
        // the staged records' stride: 80 bytes (bank-coloured slots on distinct
        // banks), 64 for a block whose halo takes more LDS than that
        const int nslot = ((d.hl_ptr[lb + 1] - d.hl_ptr[lb]) + 15) & ~15;
        const int sb = nslot * 80 <= lds_bytes ? 80 : 64;
        lat_stage_g<PDX_BLOCK * S, FR>(d, ubuf, lb, hlds, lc.inv_h, sb);
        const char *hb = reinterpret_cast<const char *>(hlds);
        const int t = threadIdx.x % PDX_BLOCK, part = threadIdx.x / PDX_BLOCK;
        const int i = lb * PDX_BLOCK + t;
        const int so = d.hs[d.hl_ptr[lb] + t];
        const char *rp = hb + so * sb;
        const pdx_q4 f0 = *reinterpret_cast<const pdx_q4 *>(rp);
        const pdx_q4 f1 = *reinterpret_cast<const pdx_q4 *>(rp + 16);
        const pdx_q4 f2 = *reinterpret_cast<const pdx_q4 *>(rp + 32);
        const pdx_q4 f3 = *reinterpret_cast<const pdx_q4 *>(rp + 48);
        const PdxD2 u01 = __builtin_bit_cast(PdxD2, f0);
        const double ui[3] = {u01.a, u01.b, __builtin_bit_cast(double, f1.xy)};
        const float Ai[9] = {f2.x, f2.y, f2.z, f2.w, f3.x, f3.y, f3.z, f3.w, f1.z};
        // c'_i w0(n) / sqrt(n) of each stencil shell n
        float ciw[10];
#pragma unroll
        for (int n = 0; n < 10; ++n) ciw[n] = f1.w * lc.wl[n];
        const uint16_t *__restrict__ bp = d.bks + lat_at(i, 0);
        double fa[3] = {0.0, 0.0, 0.0};
        if (part == 0) lat_gpart_p<S, 0>(bp, hb, sb, lc, ui, ciw, fa);
        if constexpr (S > 1) {
            if (part == 1) lat_gpart_p<S, 1>(bp, hb, sb, lc, ui, ciw, fa);
        }
        if constexpr (S > 2) {
            if (part == 2) lat_gpart_p<S, 2>(bp, hb, sb, lc, ui, ciw, fa);
            if (part == 3) lat_gpart_p<S, 3>(bp, hb, sb, lc, ui, ciw, fa);
        }
        // part 0 loads the update's inputs while the other parts end
        const KRec kr = d.KR[i];
        const bool act = (kr.flags & PDX_FL_ACTIVE) != 0;
        FinishIn fin{};
        XRec pi{};
        if (part == 0 && act) {
            fin = finish_in<UPD>(d, i, kr);
            const XRec &xr = d.xu[ubuf][i];
            pi.u[0] = xr.u[0];
            pi.u[1] = xr.u[1];
            pi.u[2] = xr.u[2];
        }
        split_sum<3, S>(hlds, fa);
        if (part == 0 && act) {
            ARec ai{};
#pragma unroll
            for (int q = 0; q < 9; ++q) ai.A[q] = Ai[q];
#pragma unroll
            for (int r = 0; r < 3; ++r) fa[r] *= lc.V;
            force_finish<3, 0, UPD>(d, i, kr, pi, ai, ubuf, cs, dt, visc, fa, red, fin);
        }
    }
    block_partials<3, PDX_BLOCK * S, PDX_BLOCK / 64>(d, lb, red);  // part 0's waves
}

// lattice tables of every row (lat_row_one); violations ORed into *flag
__global__ __launch_bounds__(PDX_BLOCK) void k_lat_rows(Dev d, double h, double tol, double V, unsigned *flag)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const uint16_t own = d.hs[d.hl_ptr[i / PDX_BLOCK] + i % PDX_BLOCK];
    const unsigned v = lat_row_one(d, i, h, tol, V, own);
    if (v) atomicOr(flag, v);
}

// smallest squared bond length of the context (FP64 bits: positive doubles
// order as unsigned integers)
__global__ __launch_bounds__(PDX_BLOCK) void k_lat_minlen(Dev d, unsigned long long *out)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    double m = 1e300;
    if (i < d.ni) {
        const KRec kr = d.KR[i];
        const XRec pi = d.xu[0][i];
        const int64_t base = slot0(d, i);
        for (int s = 0; s < (int)kr.nnb; ++s) {
            const int j = (int)(d.bj[base + (int64_t)s * PDX_SLICE] & ~PDX_BROKEN);
            const XRec pj = d.xu[0][j];
            double l2 = 0.0;
            for (int q = 0; q < 3; ++q) {
                const double t = pj.X[q] - pi.X[q];
                l2 += t * t;
            }
            m = l2 < m ? l2 : m;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double t = __shfl_xor(m, o, 64);
        m = t < m ? t : m;
    }
    if ((threadIdx.x & 63) == 0 && m < 1e300) atomicMin(out, (unsigned long long)__double_as_longlong(m));
}

// ---------------------------------------------------------------------------
// device builder of the halo tables: the tables of the host builder
// (pdx_layout::halo_block), one workgroup of PDX_BLOCK threads per block of
// rows.  The block's distinct out-of-block neighbours go into an LDS hash set
// (linear probing, compare-and-swap), are sorted by a bitonic sort and ranked
// per bank colour with wave ballots in entry order; the rare entries of an
// over-full colour then take the free slots in ascending order (one thread,
// in entry order).  MODE 0 only counts the halo (nh[b]: own + distinct
// neighbours, -1 when more than PDX_HB_MAXOUT neighbours: the host builder
// takes over); MODE 1 writes hl / hs at ptr[b], bl / bk of the block's slots
// and its anchor.
// ---------------------------------------------------------------------------
#define PDX_HB_HASH 8192
#define PDX_HB_LOG 13
#define PDX_HB_MAXOUT 6144

__device__ __forceinline__ void hb_insert(int32_t *tab, int32_t j, int *count)
{
    uint32_t h = ((uint32_t)j * 2654435761u) >> (32 - PDX_HB_LOG);
    for (int probe = 0; probe < PDX_HB_HASH; ++probe) {
        const int32_t old = atomicCAS(&tab[h], -1, j);
        if (old == -1) {
            atomicAdd(count, 1);
            return;
        }
        if (old == j) return;
        h = (h + 1) & (PDX_HB_HASH - 1);
    }
}

template <int D, int MODE>
__global__ __launch_bounds__(PDX_BLOCK) void k_halo_build(Dev d, const uint8_t *__restrict__ colour,
                                                          int32_t *__restrict__ nh_out, const int32_t *__restrict__ ptr)
{
    extern __shared__ int32_t hb[];
    int32_t *tab = hb, *arr = hb + PDX_HB_HASH;
    __shared__ int cnt_s, ext_s;
    __shared__ int run[16], wtot[PDX_BLOCK / 64][16];
    __shared__ uint8_t act_s[PDX_BLOCK];
    __shared__ double rlo[PDX_BLOCK / 64][3], rhi[PDX_BLOCK / 64][3];
    constexpr int SPB = PDX_BLOCK / PDX_SLICE, NW = PDX_BLOCK / 64;
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int i0 = b * PDX_BLOCK;
    const int64_t s_lo = d.sl_ptr[(int64_t)b * SPB], s_hi = d.sl_ptr[(int64_t)(b + 1) * SPB];
    for (int k = tid; k < PDX_HB_HASH; k += PDX_BLOCK) tab[k] = -1;
    if (tid == 0) {
        cnt_s = 0;
        ext_s = 0;
    }
    __syncthreads();
    for (int64_t s = s_lo + tid; s < s_hi; s += PDX_BLOCK) {
        const uint32_t e = d.bj[s];
        if (e == 0xffffffffu) continue;
        const int32_t j = (int32_t)(e & ~PDX_BROKEN);
        if ((unsigned)(j - i0) >= (unsigned)PDX_BLOCK && __hip_atomic_load(&cnt_s, __ATOMIC_RELAXED,
                                                                             __HIP_MEMORY_SCOPE_WORKGROUP) <= PDX_HB_MAXOUT)
            hb_insert(tab, j, &cnt_s);
    }
    __syncthreads();
    const int nu = cnt_s;
    if constexpr (MODE == 0) {
        if (tid == 0) nh_out[b] = nu > PDX_HB_MAXOUT ? -1 : PDX_BLOCK + nu;
        return;
    } else {
        if (nu > PDX_HB_MAXOUT) return;  // the host does not launch this mode then
        // the distinct neighbours in ascending order
        int P = 1;
        while (P < nu) P <<= 1;
        for (int k = tid; k < PDX_HB_HASH; k += PDX_BLOCK) {
            const int32_t v = tab[k];
            if (v >= 0) arr[atomicAdd(&ext_s, 1)] = v;
        }
        __syncthreads();
        for (int k = nu + tid; k < P; k += PDX_BLOCK) arr[k] = 0x7fffffff;
        __syncthreads();
        for (int k = 2; k <= P; k <<= 1)
            for (int jj = k >> 1; jj > 0; jj >>= 1) {
                for (int i = tid; i < P; i += PDX_BLOCK) {
                    const int ixj = i ^ jj;
                    if (ixj > i) {
                        const int32_t x = arr[i], y = arr[ixj];
                        if ((x > y) == ((i & k) == 0)) {
                            arr[i] = y;
                            arr[ixj] = x;
                        }
                    }
                }
                __syncthreads();
            }
        const int nh = PDX_BLOCK + nu, nslot = (nh + 15) & ~15, rows = nslot >> 4;
        uint16_t *slot_of = reinterpret_cast<uint16_t *>(tab);  // the hash set is no longer needed
        auto ent = [&](int e) { return e < PDX_BLOCK ? i0 + e : arr[e - PDX_BLOCK]; };
        if (tid < 16) run[tid] = 0;
        act_s[tid] = (d.KR[i0 + tid].flags & PDX_FL_ACTIVE) ? 1 : 0;
        __syncthreads();
        // colour ranks in entry order: slot 16 r + colour for r < rows
        for (int c0 = 0; c0 < nh; c0 += PDX_BLOCK) {
            const int e = c0 + tid;
            const bool ok = e < nh;
            const int cl = ok ? (colour[ent(e)] & 15) : 16;
            uint64_t mine = 0;
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const uint64_t m = __ballot(cl == k);
                if (cl == k) mine = m;
                if (lane == k) wtot[wv][k] = __popcll(m);
            }
            __syncthreads();
            if (ok) {
                int r = run[cl] + __popcll(mine & ((1ull << lane) - 1ull));
                for (int w = 0; w < wv; ++w) r += wtot[w][cl];
                slot_of[e] = r < rows ? (uint16_t)(16 * r + cl) : (uint16_t)0xFFFFu;
            }
            __syncthreads();
            if (tid < 16) {
                int t = 0;
                for (int w = 0; w < NW; ++w) t += wtot[w][tid];
                run[tid] += t;
            }
            __syncthreads();
        }
        // entries of an over-full colour: the free slots, ascending, in entry order
        if (tid == 0) {
            bool any = false;
            for (int c = 0; c < 16; ++c) any |= run[c] > rows;
            if (any) {
                int pe = 0;
                for (int sidx = 0; sidx < nslot; ++sidx) {
                    const int c = sidx & 15;
                    if ((sidx >> 4) < min(run[c], rows)) continue;  // taken by its colour
                    while (pe < nh && slot_of[pe] != 0xFFFFu) ++pe;
                    if (pe >= nh) break;
                    slot_of[pe++] = (uint16_t)sidx;
                }
            }
        }
        __syncthreads();
        const int32_t p0 = ptr[b];
        for (int e = tid; e < nh; e += PDX_BLOCK) {
            d.hl[p0 + e] = ent(e);
            d.hs[p0 + e] = slot_of[e];
        }
        // each bond slot: its particle's LDS slot (bl) and the slot the kernels read (bk)
        const int64_t a1 = d.sl_ptr[(int64_t)b * SPB + 1], a2 = d.sl_ptr[(int64_t)b * SPB + 2],
                      a3 = d.sl_ptr[(int64_t)b * SPB + 3];
        for (int64_t s = s_lo + tid; s < s_hi; s += PDX_BLOCK) {
            const int q = s < a1 ? 0 : (s < a2 ? 1 : (s < a3 ? 2 : 3));
            const int64_t aq = q == 0 ? s_lo : (q == 1 ? a1 : (q == 2 ? a2 : a3));
            const int t = q * PDX_SLICE + (int)((s - aq) & (PDX_SLICE - 1));
            const uint32_t e = d.bj[s];
            uint16_t blv = PDX_HNONE, bkv = slot_of[t];
            if (e != 0xffffffffu) {
                const int32_t j = (int32_t)(e & ~PDX_BROKEN);
                int l = j - i0;
                if ((unsigned)l >= (unsigned)PDX_BLOCK) {
                    int lo = 0, hi = nu;  // lower bound of j in arr[0, nu)
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (arr[mid] < j) lo = mid + 1;
                        else hi = mid;
                    }
                    l = PDX_BLOCK + lo;
                }
                blv = (uint16_t)(slot_of[l] | ((e & PDX_BROKEN) ? PDX_HBROKEN : 0u));
                if (!(e & PDX_BROKEN) && act_s[t]) bkv = blv;
            }
            d.bl[s] = blv;
            d.bk[s] = bkv;
        }
        // anchor: centre of the bounding box of the block's active particles
        double lo[3], hi[3];
        const XRec xr = d.xu[0][i0 + tid];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            lo[q] = act_s[tid] ? xr.X[q] : 1e300;
            hi[q] = act_s[tid] ? xr.X[q] : -1e300;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                lo[q] = fmin(lo[q], __shfl_xor(lo[q], o, 64));
                hi[q] = fmax(hi[q], __shfl_xor(hi[q], o, 64));
            }
        if (lane == 0)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                rlo[wv][q] = lo[q];
                rhi[wv][q] = hi[q];
            }
        __syncthreads();
        if (tid == 0) {
            double m[3] = {0.0, 0.0, 0.0};
            for (int q = 0; q < D; ++q) {
                double l = rlo[0][q], h = rhi[0][q];
                for (int w = 1; w < NW; ++w) {
                    l = fmin(l, rlo[w][q]);
                    h = fmax(h, rhi[w][q]);
                }
                m[q] = l <= h ? 0.5 * (l + h) : 0.0;
            }
            d.anchor[b] = d4(m[0], m[1], m[2]);
        }
    }
}

// bl's broken flags from bj (after bj was rewritten) and the slots the halo
// kernels read (bk), one row per thread
__global__ __launch_bounds__(PDX_BLOCK) void k_halo_bits(Dev d)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const uint16_t own = d.hs[d.hl_ptr[i / PDX_BLOCK] + i % PDX_BLOCK];
    const bool act = (d.KR[i].flags & PDX_FL_ACTIVE) != 0;
    const int64_t s0 = d.sl_ptr[i >> 6];
    const int width = (int)((d.sl_ptr[(i >> 6) + 1] - s0) / PDX_SLICE);
    for (int k = 0; k < width; ++k) {
        const int64_t s = s0 + (int64_t)k * PDX_SLICE + (i & (PDX_SLICE - 1));
        const uint16_t e = d.bl[s];
        if (e == PDX_HNONE) {
            d.bk[s] = own;
            continue;
        }
        const bool br = (d.bj[s] & PDX_BROKEN) != 0;
        const uint16_t l = (uint16_t)(e & ~PDX_HBROKEN);
        d.bl[s] = (uint16_t)(l | (br ? PDX_HBROKEN : 0u));
        d.bk[s] = (act && !br) ? l : own;
    }
    if (!d.bks) return;
    // the lattice slots follow the same rule, at each bond's stencil offset
    for (int k = 0; k < PDX_NST; ++k) d.bks[lat_at(i, k)] = own;
    for (int k = 0; k < width; ++k) {
        const int64_t s = s0 + (int64_t)k * PDX_SLICE + (i & (PDX_SLICE - 1));
        if (d.bl[s] != PDX_HNONE) d.bks[lat_at(i, d.lk[s])] = d.bk[s];
    }
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
    const int i = lb * PDX_BLOCK + threadIdx.x;
    double red[1] = {i < d.ni ? resid_one<D>(d, i) : 0.0};
    block_partials<1>(d, lb, red);
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
template <int D> PDX_PER_PARTICLE(k_adr_mass_b(Dev d, const double *cs, const double *g, double beta, double *raw),
                                  (adr_mass_b_one<D>(d, i, cs, g, beta, raw)))
PDX_PER_PARTICLE(k_mass_scale(Dev d, const double *raw, double beta), (mass_scale_one(d, i, raw, beta)))
template <int D> PDX_PER_PARTICLE(k_dt_est(Dev d, double modulus, double *out), (dt_est_one<D>(d, i, modulus, out)))
template <int D> PDX_PER_PARTICLE(k_bond_data(Dev d, double *xi, double *len, double *w, const int64_t *rowi),
                                  (bond_data_one<D>(d, i, xi, len, w, rowi)))
template <int CH>
__device__ __forceinline__ void nbr_mean_cols(const Dev &d, int i, const double *in, double *out, int width)
{
    for (int q0 = 0; q0 < width; q0 += CH) nbr_mean_one<CH>(d, i, in, out, width, q0);
}

PDX_PER_PARTICLE(k_local_damage(Dev d), (local_damage_one(d, i)))
PDX_PER_PARTICLE(k_zero_free(Dev d, int ubuf), (zero_free_one(d, i, ubuf)))
template <int CH>
PDX_PER_PARTICLE(k_nbr_mean(Dev d, const double *in, double *out, int width),
                 (nbr_mean_cols<CH>(d, i, in, out, width)))

template <int D>
__global__ __launch_bounds__(PDX_BLOCK) void k_damage(Dev d, int ubuf, double s_c, unsigned long long *count)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const unsigned nb = damage_one<D>(d, i, ubuf, s_c);
    if (nb) atomicAdd(count, (unsigned long long)nb);
}

// X, V and c of both gather buffers from a staged copy (u left as it is)
__global__ __launch_bounds__(PDX_BLOCK) void k_xu_geom(Dev d, const XRec *src)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const XRec s = src[i];
    for (int b = 0; b < 2; ++b) {
        XRec &r = d.xu[b][i];
        for (int q = 0; q < 3; ++q) r.X[q] = s.X[q];
        r.V = s.V;
        r.c = s.c;
        r.w = 0.0;
    }
}

// scatter cnt 32-bit words per particle into KRec word offset off
__global__ __launch_bounds__(PDX_BLOCK) void k_kr_put(Dev d, const float *src, int off, int cnt)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    float *w = reinterpret_cast<float *>(&d.KR[i]);
    for (int q = 0; q < cnt; ++q) w[off + q] = src[(int64_t)i * cnt + q];
}

// one 32-bit KRec word per particle: the bits in keep stay, the others come from src
__global__ __launch_bounds__(PDX_BLOCK) void k_kr_merge(Dev d, const uint32_t *src, int off, uint32_t keep)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    uint32_t *w = reinterpret_cast<uint32_t *>(&d.KR[i]);
    w[off] = (w[off] & keep) | (src[i] & ~keep);
}

// the stabilisation modulus word of both gather buffers
__global__ __launch_bounds__(PDX_BLOCK) void k_xu_c(Dev d, const float *src)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    d.xu[0][i].c = src[i];
    d.xu[1][i].c = src[i];
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

int pdx_launch_fill(pdx_ctx *c, uint64_t *d_keys)
{
    Dev d = c->d;
    d.nkey = d_keys;
    if (c->dim == 3)
        hipLaunchKernelGGL((k_neighbors<3, true>), dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, d);
    else
        hipLaunchKernelGGL((k_neighbors<2, true>), dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, d);
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

// the hot passes come in variants chosen per context by pdx_autotune:
// per-lane gathers with bond groups 1, 2, 4 and wave-cooperative gathers with
// groups 1, 2 (PDX_VARIANT_COOP | group); diagnostic and commit passes use
// the per-lane kernel with group 1
// LDS bytes of a halo-staged pass (the halo, or the split partial sums
// where larger), or 0 when it does not fit (then the pass runs the
// cooperative group-2 kernel)
int pdx_halo_lds(const pdx_ctx *c, int entry_bytes, int split, int nv)
{
    if (!c->d.hl_ptr || c->d.halo_cap <= 0) return 0;
    const int64_t halo = (int64_t)c->d.halo_cap * entry_bytes;
    const int64_t part = (int64_t)(split - 1) * nv * PDX_BLOCK * (int64_t)sizeof(double);
    const int64_t bytes = std::max(halo, part);
    return bytes <= c->lds_dyn_max ? (int)bytes : 0;
}

// variant code -> halo kernel bond group and split
constexpr int halo_group(int v) { return v & 15; }
constexpr int halo_split(int v) { return (v & PDX_VARIANT_SPLIT4) ? 4 : ((v & PDX_VARIANT_SPLIT2) ? 2 : 1); }

template <int D, int MODEL, int G>
static void launch_fpass_g(pdx_ctx *c, int ubuf, bool diag, bool commit, int sm, int red_ubuf)
{
    dim3 g(c->nblk), b(PDX_BLOCK);
    // red_ubuf >= 0: the adaptive-DR case reduction of the preceding force
    // pass goes first, fused into the halo F-pass, else as its own launch
    if (red_ubuf >= 0 && ((G & PDX_VARIANT_HALO) == 0 || diag || commit))
        pdx_launch_case_reduce(c, red_ubuf, 1, sm != 0);
    if (commit) {
        hipLaunchKernelGGL((k_fpass<D, MODEL, false, true, 1>), g, b, 0, c->stream, c->d, ubuf, sm);
    } else if (diag) {
        hipLaunchKernelGGL((k_fpass<D, MODEL, true, false, 1>), g, b, 0, c->stream, c->d, ubuf, sm);
    } else {
        if constexpr ((G & PDX_VARIANT_HALO) != 0) {
            constexpr int GH = halo_group(G), S = halo_split(G);
            const int lds = pdx_halo_lds(c, HaloOff<0, 0>::bytes, S, D * D + D);
            const dim3 bh(PDX_BLOCK * S), gr(c->nblk + c->ncase);
            if (lds && red_ubuf >= 0 && c->have_m)
                hipLaunchKernelGGL((k_fpass_h<D, MODEL, GH, S, true, 1>), gr, bh, lds, c->stream, c->d, ubuf, sm,
                                   red_ubuf);
            else if (lds && red_ubuf >= 0)
                hipLaunchKernelGGL((k_fpass_h<D, MODEL, GH, S, true>), gr, bh, lds, c->stream, c->d, ubuf, sm,
                                   red_ubuf);
            else if (lds && c->have_m)
                hipLaunchKernelGGL((k_fpass_h<D, MODEL, GH, S, false, 1>), g, bh, lds, c->stream, c->d, ubuf, sm, -1);
            else if (lds)
                hipLaunchKernelGGL((k_fpass_h<D, MODEL, GH, S>), g, bh, lds, c->stream, c->d, ubuf, sm, -1);
            else {
                if (red_ubuf >= 0) pdx_launch_case_reduce(c, red_ubuf, 1, sm != 0);
                hipLaunchKernelGGL((k_fpass_w<D, MODEL, 2>), g, b, 0, c->stream, c->d, ubuf, sm);
            }
        } else if constexpr (G > PDX_VARIANT_COOP) {
            hipLaunchKernelGGL((k_fpass_w<D, MODEL, G - PDX_VARIANT_COOP>), g, b, 0, c->stream, c->d, ubuf, sm);
        } else {
            hipLaunchKernelGGL((k_fpass<D, MODEL, false, false, G>), g, b, 0, c->stream, c->d, ubuf, sm);
        }
    }
}

// LDS bytes of the lattice force pass with split S, or 0 when it cannot run
// dynamic LDS of the lattice force pass (0: it cannot run).  Its staged
// records take 64 bytes; a block lays them 80 bytes apart (the 16 lanes of a
// read then meet distinct banks, with the bank-coloured slots) when its halo
// fits that way, else 64 apart (4-way bank conflicts; a block whose rows
// reach more than about 2 000 particles), so the launch takes the 80-byte
// size of the widest halo, or the most the device gives when only 64 bytes
// per record fit it.
static int lat_force_lds(const pdx_ctx *c, int split)
{
    if (!c->have_lat || !c->d.bks || c->dim != 3 || c->form != PDX_FORM_NOSB) return 0;
    if (const int lds = pdx_halo_lds(c, 80, split, 3)) return lds;
    return pdx_halo_lds(c, 64, split, 3) ? c->lds_dyn_max : 0;
}

// the force pass's variant is a lattice variant that runs
static bool lat_force_runs(const pdx_ctx *c)
{
    if (!(c->tune_g & PDX_VARIANT_LATTICE)) return false;
    const int split = (c->tune_g & PDX_VARIANT_SPLIT4) ? 4 : ((c->tune_g & PDX_VARIANT_SPLIT2) ? 2 : 1);
    return lat_force_lds(c, split) > 0;
}

// the lattice F-pass with rows split over S waves; false when it cannot run
// (no lattice tables, 2-D, a diagnostic or commit pass, or a halo the
// device's LDS does not take)
template <int D, int MODEL, int S>
static bool launch_fpass_lat(pdx_ctx *c, int ubuf, bool diag, bool commit, int sm, int red_ubuf)
{
    if constexpr (D != 3) {
        return false;
    } else {
        if (diag || commit || !c->have_lat || !c->d.bks || c->form != PDX_FORM_NOSB) return false;
        const int lds = pdx_halo_lds(c, 24, S, 9);
        if (!lds) return false;
        const dim3 bh(PDX_BLOCK * S), g(c->nblk), gr(c->nblk + c->ncase);
        // with a lattice force pass to follow, A goes into its records
        const bool fr = lat_force_runs(c);
        if (red_ubuf >= 0 && fr)
            hipLaunchKernelGGL((k_fpass_l<MODEL, S, true, true>), gr, bh, lds, c->stream, c->d, c->lat, ubuf, sm,
                               red_ubuf);
        else if (red_ubuf >= 0)
            hipLaunchKernelGGL((k_fpass_l<MODEL, S, true>), gr, bh, lds, c->stream, c->d, c->lat, ubuf, sm, red_ubuf);
        else if (fr)
            hipLaunchKernelGGL((k_fpass_l<MODEL, S, false, true>), g, bh, lds, c->stream, c->d, c->lat, ubuf, sm, -1);
        else
            hipLaunchKernelGGL((k_fpass_l<MODEL, S>), g, bh, lds, c->stream, c->d, c->lat, ubuf, sm, -1);
        c->fr_ubuf = fr ? ubuf : -1;
        return true;
    }
}

template <int D, int MODEL>
static void launch_fpass_t(pdx_ctx *c, int ubuf, bool diag, bool commit, int sm, int red)
{
    switch (c->tune_f) {
    case PDX_VARIANT_LATTICE:
        if (!launch_fpass_lat<D, MODEL, 1>(c, ubuf, diag, commit, sm, red))
            launch_fpass_g<D, MODEL, PDX_VARIANT_HALO | 1>(c, ubuf, diag, commit, sm, red);
        break;
    case PDX_VARIANT_LATTICE | PDX_VARIANT_SPLIT2:
        if (!launch_fpass_lat<D, MODEL, 2>(c, ubuf, diag, commit, sm, red))
            launch_fpass_g<D, MODEL, PDX_VARIANT_HALO | PDX_VARIANT_SPLIT2 | 1>(c, ubuf, diag, commit, sm, red);
        break;
    case PDX_VARIANT_LATTICE | PDX_VARIANT_SPLIT4:
        if (!launch_fpass_lat<D, MODEL, 4>(c, ubuf, diag, commit, sm, red))
            launch_fpass_g<D, MODEL, PDX_VARIANT_HALO | PDX_VARIANT_SPLIT4 | 1>(c, ubuf, diag, commit, sm, red);
        break;
#define PDX_HALO_CASE(v) \
    case v: launch_fpass_g<D, MODEL, v>(c, ubuf, diag, commit, sm, red); break;
    PDX_HALO_VARIANTS(PDX_HALO_CASE)
#undef PDX_HALO_CASE
    case PDX_VARIANT_COOP | 2: launch_fpass_g<D, MODEL, PDX_VARIANT_COOP | 2>(c, ubuf, diag, commit, sm, red); break;
    case PDX_VARIANT_COOP | 1: launch_fpass_g<D, MODEL, PDX_VARIANT_COOP | 1>(c, ubuf, diag, commit, sm, red); break;
    case 4: launch_fpass_g<D, MODEL, 4>(c, ubuf, diag, commit, sm, red); break;
    case 2: launch_fpass_g<D, MODEL, 2>(c, ubuf, diag, commit, sm, red); break;
    default: launch_fpass_g<D, MODEL, 1>(c, ubuf, diag, commit, sm, red); break;
    }
}

int pdx_launch_fpass(pdx_ctx *c, int ubuf, bool diag, bool commit, bool solve_mode, int red_ubuf)
{
    c->fr_ubuf = -1;  // set again by a lattice F-pass that writes the force pass's records
    const int sm = solve_mode ? 1 : 0;
    const int red = red_ubuf;
    if (c->dim == 3) {
        switch (c->model) {
        case PDX_MODEL_NEO_HOOKEAN: launch_fpass_t<3, PDX_MODEL_NEO_HOOKEAN>(c, ubuf, diag, commit, sm, red); break;
        case PDX_MODEL_J2: launch_fpass_t<3, PDX_MODEL_J2>(c, ubuf, diag, commit, sm, red); break;
        default: launch_fpass_t<3, PDX_MODEL_LINEAR_ELASTIC>(c, ubuf, diag, commit, sm, red); break;
        }
    } else {
        switch (c->model) {
        case PDX_MODEL_NEO_HOOKEAN: launch_fpass_t<2, PDX_MODEL_NEO_HOOKEAN>(c, ubuf, diag, commit, sm, red); break;
        default: launch_fpass_t<2, PDX_MODEL_LINEAR_ELASTIC>(c, ubuf, diag, commit, sm, red); break;
        }
    }
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

// d.m64 by the halo F-pass of the tuned variant, summed in the order that
// variant sums it: with rows split over S waves, part p takes the slot groups
// p G, p G + G S, ..., so the FP64 order depends on the bond group G as well
// as on S.  Its model does not enter the sum.  m64 is valid only for the
// variant it was built with: every change of the F-pass variant clears have_m.
template <int D, int V>
static int build_m64_v(pdx_ctx *c)
{
    constexpr int GH = halo_group(V), S = halo_split(V);
    const int lds = pdx_halo_lds(c, HaloOff<0, 0>::bytes, S, D * D + D);
    if (!lds) return PDX_OK;
    hipLaunchKernelGGL((k_fpass_h<D, PDX_MODEL_LINEAR_ELASTIC, GH, S, false, 2>), dim3(c->nblk),
                       dim3(PDX_BLOCK * S), lds, c->stream, c->d, c->ucur, 0, -1);
    PDX_LAUNCH_CHECK();
    c->have_m = true;
    return PDX_OK;
}

int pdx_build_m64(pdx_ctx *c)
{
    c->have_m = false;
    // only the halo F-pass reads m64; a lattice variant runs the halo F-pass
    // of its split (group 1) where the lattice F-pass cannot run
    int tf = c->tune_f;
    if (tf & PDX_VARIANT_LATTICE) {
        const int split = (tf & PDX_VARIANT_SPLIT4) ? 4 : ((tf & PDX_VARIANT_SPLIT2) ? 2 : 1);
        if (c->dim == 3 && c->have_lat && c->d.bks && pdx_halo_lds(c, 24, split, 9)) return PDX_OK;
        tf = PDX_VARIANT_HALO | (tf & (PDX_VARIANT_SPLIT2 | PDX_VARIANT_SPLIT4)) | 1;
    }
    if (!c->d.m64 || !c->d.hl_ptr || c->form != PDX_FORM_NOSB || !(tf & PDX_VARIANT_HALO)) return PDX_OK;
#define PDX_M64_CASE(v)                                                          \
    if (tf == (v)) return c->dim == 3 ? build_m64_v<3, (v)>(c) : build_m64_v<2, (v)>(c);
    PDX_HALO_VARIANTS(PDX_M64_CASE)
#undef PDX_M64_CASE
    return PDX_OK;
}

template <int D, int FORM, int G>
static void launch_force_g(pdx_ctx *c, int ubuf, int upd, int sm, double dt, double visc)
{
    dim3 g(c->nblk), b(PDX_BLOCK);
    if constexpr ((G & PDX_VARIANT_HALO) != 0) {
        constexpr int GH = halo_group(G), S = halo_split(G);
        const int lds = pdx_halo_lds(c, HaloOff<1, FORM>::bytes, S, 3);
        const dim3 bh(PDX_BLOCK * S);
        if (!lds)
            launch_force_g<D, FORM, PDX_VARIANT_COOP | 2>(c, ubuf, upd, sm, dt, visc);
        else if (upd == 1)
            hipLaunchKernelGGL((k_force_h<D, FORM, 1, GH, S>), g, bh, lds, c->stream, c->d, ubuf, sm, dt, visc);
        else if (upd == 2)
            hipLaunchKernelGGL((k_force_h<D, FORM, 2, GH, S>), g, bh, lds, c->stream, c->d, ubuf, sm, dt, visc);
        else
            hipLaunchKernelGGL((k_force_h<D, FORM, 0, GH, S>), g, bh, lds, c->stream, c->d, ubuf, sm, dt, visc);
    } else if constexpr (G > PDX_VARIANT_COOP) {
        constexpr int GW = G - PDX_VARIANT_COOP;
        if (upd == 1)
            hipLaunchKernelGGL((k_force_w<D, FORM, 1, GW>), g, b, 0, c->stream, c->d, ubuf, sm, dt, visc);
        else if (upd == 2)
            hipLaunchKernelGGL((k_force_w<D, FORM, 2, GW>), g, b, 0, c->stream, c->d, ubuf, sm, dt, visc);
        else
            hipLaunchKernelGGL((k_force_w<D, FORM, 0, GW>), g, b, 0, c->stream, c->d, ubuf, sm, dt, visc);
    } else {
        if (upd == 1)
            hipLaunchKernelGGL((k_force<D, FORM, 1, G>), g, b, 0, c->stream, c->d, ubuf, sm, dt, visc);
        else if (upd == 2)
            hipLaunchKernelGGL((k_force<D, FORM, 2, G>), g, b, 0, c->stream, c->d, ubuf, sm, dt, visc);
        else
            hipLaunchKernelGGL((k_force<D, FORM, 0, 1>), g, b, 0, c->stream, c->d, ubuf, sm, dt, visc);
    }
}

template <int S, bool FR>
static void launch_force_l(pdx_ctx *c, int lds, int ubuf, int upd, int sm, double dt, double visc)
{
    const dim3 g(c->nblk), bh(PDX_BLOCK * S);
    if (upd == 1)
        hipLaunchKernelGGL((k_force_l<1, S, FR>), g, bh, lds, c->stream, c->d, c->lat, ubuf, sm, dt, visc, lds);
    else if (upd == 2)
        hipLaunchKernelGGL((k_force_l<2, S, FR>), g, bh, lds, c->stream, c->d, c->lat, ubuf, sm, dt, visc, lds);
    else
        hipLaunchKernelGGL((k_force_l<0, S, FR>), g, bh, lds, c->stream, c->d, c->lat, ubuf, sm, dt, visc, lds);
}

// the lattice force pass with rows split over S waves; false when it cannot
// run (no lattice tables, 2-D, the bond form, or a halo the device's LDS does
// not take)
template <int D, int FORM, int S>
static bool launch_force_lat(pdx_ctx *c, int ubuf, int upd, int sm, double dt, double visc)
{
    if constexpr (D != 3 || FORM != PDX_FORM_NOSB) {
        return false;
    } else {
        const int lds = lat_force_lds(c, S);
        if (!lds) return false;
        // the records of the lattice F-pass that ran last, on this buffer
        if (c->fr_ubuf == ubuf)
            launch_force_l<S, true>(c, lds, ubuf, upd, sm, dt, visc);
        else
            launch_force_l<S, false>(c, lds, ubuf, upd, sm, dt, visc);
        return true;
    }
}

template <int D, int FORM>
static void launch_force_t(pdx_ctx *c, int ubuf, int upd, int sm, double dt, double visc)
{
    switch (c->tune_g) {
    case PDX_VARIANT_LATTICE:
        if (!launch_force_lat<D, FORM, 1>(c, ubuf, upd, sm, dt, visc))
            launch_force_g<D, FORM, PDX_VARIANT_HALO | 4>(c, ubuf, upd, sm, dt, visc);
        break;
    case PDX_VARIANT_LATTICE | PDX_VARIANT_SPLIT2:
        if (!launch_force_lat<D, FORM, 2>(c, ubuf, upd, sm, dt, visc))
            launch_force_g<D, FORM, PDX_VARIANT_HALO | PDX_VARIANT_SPLIT2 | 4>(c, ubuf, upd, sm, dt, visc);
        break;
    case PDX_VARIANT_LATTICE | PDX_VARIANT_SPLIT4:
        if (!launch_force_lat<D, FORM, 4>(c, ubuf, upd, sm, dt, visc))
            launch_force_g<D, FORM, PDX_VARIANT_HALO | PDX_VARIANT_SPLIT4 | 4>(c, ubuf, upd, sm, dt, visc);
        break;
#define PDX_HALO_CASE(v) \
    case v: launch_force_g<D, FORM, v>(c, ubuf, upd, sm, dt, visc); break;
    PDX_HALO_VARIANTS(PDX_HALO_CASE)
#undef PDX_HALO_CASE
    case PDX_VARIANT_COOP | 2: launch_force_g<D, FORM, PDX_VARIANT_COOP | 2>(c, ubuf, upd, sm, dt, visc); break;
    case PDX_VARIANT_COOP | 1: launch_force_g<D, FORM, PDX_VARIANT_COOP | 1>(c, ubuf, upd, sm, dt, visc); break;
    case 4: launch_force_g<D, FORM, 4>(c, ubuf, upd, sm, dt, visc); break;
    case 2: launch_force_g<D, FORM, 2>(c, ubuf, upd, sm, dt, visc); break;
    default: launch_force_g<D, FORM, 1>(c, ubuf, upd, sm, dt, visc); break;
    }
}

int pdx_launch_force(pdx_ctx *c, int ubuf, int upd, bool solve_mode, double dt, double visc)
{
    const int r = pdx_launch_force_pass(c, ubuf, upd, solve_mode, dt, visc);
    if (r != PDX_OK) return r;
    return pdx_launch_case_reduce(c, ubuf, upd, solve_mode);
}

int pdx_launch_force_pass(pdx_ctx *c, int ubuf, int upd, bool solve_mode, double dt, double visc)
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

int pdx_launch_lat_minlen(pdx_ctx *c, unsigned long long *d_out)
{
    hipLaunchKernelGGL(k_lat_minlen, dim3((unsigned)grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d, d_out);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

int pdx_launch_lat_rows(pdx_ctx *c, double h, double tol, double V, unsigned *d_flag)
{
    hipLaunchKernelGGL(k_lat_rows, dim3((unsigned)grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d, h, tol, V,
                       d_flag);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

int pdx_launch_halo_bits(pdx_ctx *c)
{
    c->have_m = false;  // the intact bonds change
    if (!c->d.bl || c->slots <= 0) return PDX_OK;
    hipLaunchKernelGGL(k_halo_bits, dim3((unsigned)((c->ni + PDX_BLOCK - 1) / PDX_BLOCK)), dim3(PDX_BLOCK), 0,
                       c->stream, c->d);
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

// the Gershgorin mass times beta into d.lam, or (raw != null) its values
// before beta into raw
int pdx_launch_adr_mass(pdx_ctx *c, double beta, double *raw)
{
    double *tmp = nullptr;
    PDX_HIP(hipMallocAsync((void **)&tmp, sizeof(double) * 2 * (size_t)c->ni, c->stream));
    dim3 g(grid1(c->ni)), b(PDX_BLOCK);
    if (c->dim == 3) {
        hipLaunchKernelGGL((k_adr_mass_a<3>), g, b, 0, c->stream, c->d, tmp, tmp + c->ni);
        hipLaunchKernelGGL((k_adr_mass_b<3>), g, b, 0, c->stream, c->d, tmp, tmp + c->ni, beta, raw);
    } else {
        hipLaunchKernelGGL((k_adr_mass_a<2>), g, b, 0, c->stream, c->d, tmp, tmp + c->ni);
        hipLaunchKernelGGL((k_adr_mass_b<2>), g, b, 0, c->stream, c->d, tmp, tmp + c->ni, beta, raw);
    }
    PDX_LAUNCH_CHECK();
    PDX_HIP(hipFreeAsync(tmp, c->stream));
    return PDX_OK;
}

int pdx_launch_mass_scale(pdx_ctx *c, const double *raw, double beta)
{
    hipLaunchKernelGGL(k_mass_scale, dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d, raw, beta);
    PDX_LAUNCH_CHECK();
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
    c->have_m = false;  // bonds may break
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
    hipLaunchKernelGGL((k_case_reduce<3>), dim3(c->ncase), dim3(PDX_RED_THREADS), 0, c->stream, c->d, 0, 0);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

int pdx_launch_case_reduce(pdx_ctx *c, int ubuf, int upd, bool solve_mode)
{
    const int sm = solve_mode ? 1 : 0;
    if (upd == 1)
        hipLaunchKernelGGL((k_case_reduce<1>), dim3(c->ncase), dim3(PDX_RED_THREADS), 0, c->stream, c->d, ubuf, sm);
    else if (upd == 2)
        hipLaunchKernelGGL((k_case_reduce<2>), dim3(c->ncase), dim3(PDX_RED_THREADS), 0, c->stream, c->d, ubuf, sm);
    else
        hipLaunchKernelGGL((k_case_reduce<0>), dim3(c->ncase), dim3(PDX_RED_THREADS), 0, c->stream, c->d, ubuf, sm);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

int pdx_launch_zero_free(pdx_ctx *c, int ubuf)
{
    hipLaunchKernelGGL(k_zero_free, dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d, ubuf);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

// weighted neighbour means of W FP64 columns on a lattice context: the rows
// of the block's halo are staged in LDS (W 8 bytes per slot), each stencil
// offset k weighs w0(|o_k|) V, and a bond counts when bks names another slot
// than the row's own (an intact bond of an active row; the own slot stands
// for no bond, a broken one or an inactive row).  A row without bonds keeps
// its value.  The slot numbers are read NG offsets at a time, the next
// group's while the current one is summed.
template <int W>
__global__ __launch_bounds__(PDX_BLOCK) void k_nbr_mean_l(Dev d, LatC lc, const double *__restrict__ in,
                                                         double *__restrict__ out)
{
    static_assert(W % 2 == 0, "16-byte row pieces");
    constexpr int NG = 16, NK = (PDX_NST + NG - 1) / NG * NG;
    extern __shared__ pdx_q4 hlds[];
    __shared__ double wk[NK];
    const int lb = logical_block();
    if (lb * PDX_BLOCK >= d.ni) return;  // grid padding: the whole block leaves
    const int t = threadIdx.x, i = lb * PDX_BLOCK + t;
    const uint16_t *__restrict__ bp = d.bks + lat_at(i, 0);
    const int h0 = d.hl_ptr[lb], nh = d.hl_ptr[lb + 1] - h0;
    const int so = d.hs[h0 + t];
    int cur[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) cur[g] = bp[g * PDX_SLICE];
    for (int e = t; e < nh; e += PDX_BLOCK) {
        const pdx_q4 *src = reinterpret_cast<const pdx_q4 *>(in + (int64_t)d.hl[h0 + e] * W);
        pdx_q4 *dst = hlds + (int)d.hs[h0 + e] * (W / 2);
#pragma unroll
        for (int q = 0; q < W / 2; ++q) dst[q] = src[q];
    }
    if (t < NK) {
        double w = 0.0;  // offsets past the stencil weigh nothing
        if (t < PDX_NST) {
            const int8_t *o = kLat.o[t];
            w = lc.w1[o[0] * o[0] + o[1] * o[1] + o[2] * o[2]] * lc.V;
        }
        wk[t] = w;
    }
    __syncthreads();
    const double *hb = reinterpret_cast<const double *>(hlds);
    double acc[W], ws = 0.0;
#pragma unroll
    for (int q = 0; q < W; ++q) acc[q] = 0.0;
    for (int k0 = 0; k0 < NK; k0 += NG) {
        int nxt[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const int k = k0 + NG + g;
            nxt[g] = k < PDX_NST ? bp[k * PDX_SLICE] : so;
        }
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const int sl = k0 + g < PDX_NST ? cur[g] : so;
            const double w = sl == so ? 0.0 : wk[k0 + g];
            const double *row = hb + sl * W;
#pragma unroll
            for (int q = 0; q < W; ++q) acc[q] += w * row[q];
            ws += w;
        }
#pragma unroll
        for (int g = 0; g < NG; ++g) cur[g] = nxt[g];
    }
#pragma unroll
    for (int q = 0; q < W; ++q) out[(int64_t)i * W + q] = ws > 0.0 ? acc[q] / ws : hb[so * W + q];
}

int pdx_launch_nbr_mean(pdx_ctx *c, const double *d_in, double *d_out, int width)
{
    // the warm-start features' varying columns (10) on a lattice context:
    // the lattice kernel, when its halo fits the LDS
    if (width == 10 && c->have_lat && c->d.bks && c->dim == 3) {
        const int lds = pdx_halo_lds(c, 80, 1, 0);
        if (lds) {
            hipLaunchKernelGGL(k_nbr_mean_l<10>, dim3(c->nblk), dim3(PDX_BLOCK), lds, c->stream, c->d, c->lat, d_in,
                               d_out);
            PDX_LAUNCH_CHECK();
            return PDX_OK;
        }
    }
    // the warm-start features (20 columns, or the 10 that vary within a case)
    // in one pass over the bonds; other widths four columns per pass
    if (width == 20)
        hipLaunchKernelGGL(k_nbr_mean<20>, dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d, d_in, d_out,
                           width);
    else if (width == 10)
        hipLaunchKernelGGL(k_nbr_mean<10>, dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d, d_in, d_out,
                           width);
    else
        hipLaunchKernelGGL(k_nbr_mean<4>, dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d, d_in, d_out,
                           width);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

// caller-order rows (n x width FP64) into internal order (padding rows 0)
// one thread per value, consecutive threads along a row
__global__ __launch_bounds__(PDX_BLOCK) void k_rows_internal(Dev d, const double *in, double *out, int width)
{
    const int64_t e = (int64_t)blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (e >= (int64_t)d.ni * width) return;
    const int i = (int)(e / width), q = (int)(e - (int64_t)i * width);
    const int r = d.perm[i];
    out[e] = r >= 0 ? in[(int64_t)r * width + q] : 0.0;
}

int pdx_launch_rows_internal(pdx_ctx *c, const double *d_in, double *d_out, int width)
{
    const int64_t total = (int64_t)c->ni * width;
    hipLaunchKernelGGL(k_rows_internal, dim3((unsigned)((total + PDX_BLOCK - 1) / PDX_BLOCK)), dim3(PDX_BLOCK), 0,
                       c->stream, c->d, d_in, d_out, width);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

// u of buffer ubuf between the internal records and a caller-order device
// array (n x dim FP64): out != null reads into it, else in is written
__global__ __launch_bounds__(PDX_BLOCK) void k_u_device(Dev d, int ubuf, double *out, const double *in)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const int r = d.perm[i];
    if (r < 0) return;
    double *u = d.xu[ubuf][i].u;
    if (out) {
        for (int q = 0; q < d.dim; ++q) out[(int64_t)r * d.dim + q] = u[q];
    } else {
        for (int q = 0; q < d.dim; ++q) u[q] = in[(int64_t)r * d.dim + q];
    }
}

int pdx_launch_u_device(pdx_ctx *c, int ubuf, double *out, const double *in)
{
    hipLaunchKernelGGL(k_u_device, dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d, ubuf, out, in);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

// [hop 0 | hop 1 | ...] of internal-order FP64 blocks (ni x width each) into
// the caller-order FP32 rows of out (stride ldo)
// one thread per packed value, consecutive threads along a row's columns
// (width x (hops + 1) of them), so the reads and writes of a wave cover whole
// row pieces
__global__ __launch_bounds__(PDX_BLOCK) void k_feat_pack(Dev d, const double *hv, const int32_t *perm, float *out,
                                                        int width, int hops, int ldo)
{
    const int nc = width * (hops + 1);
    const int64_t e = (int64_t)blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (e >= (int64_t)d.ni * nc) return;
    const int i = (int)(e / nc), col = (int)(e - (int64_t)i * nc);
    const int r = perm[i];
    if (r < 0) return;
    const int k = col / width, q = col - k * width;
    out[(int64_t)r * ldo + col] = (float)hv[((int64_t)k * d.ni + i) * width + q];
}

int pdx_launch_feat_pack(pdx_ctx *c, const double *d_hv, const int32_t *d_perm, float *d_out, int width, int hops,
                         int ldo)
{
    const int64_t total = (int64_t)c->ni * width * (hops + 1);
    hipLaunchKernelGGL(k_feat_pack, dim3((unsigned)((total + PDX_BLOCK - 1) / PDX_BLOCK)), dim3(PDX_BLOCK), 0,
                       c->stream, c->d, d_hv, d_perm, d_out, width, hops, ldo);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

// ---------------------------------------------------------------------------
// raw inputs of the learned warm start: 20 FP64 columns per particle in caller
// order, from the specimen's positions, moduli, loads, DOF masks and phases.
// Per-case sums over row chunks (k_raw_partials), per-case constants summed
// in chunk order (k_raw_cases, deterministic), then one thread per row
// (k_raw_fill).  Columns:
//   0-2   (x - cx) / a, (y - cy) / b, (z - z_min) / H   (bounding box of the case)
//   3     log10(E / 1000)          4  1 for phase 0 (cortical)
//   5-7   fixed-DOF bits           8  1 on loaded rows (any load component != 0)
//   9-11  unit resultant load      12 log10 of the resultant pressure |sum f| dx / n_loaded
//   13-15 H / 25, a / 20, b / 20   16 case mean of column 3, 17 case mean of column 4
//   18-19 E-weighted centroid minus the box centre, over a and b
// ---------------------------------------------------------------------------
enum { RS_MIN = 0, RS_MAX = 3, RS_F = 6, RS_NLOAD = 9, RS_LOGE = 10, RS_CORT = 11, RS_E = 12, RS_EX = 13,
       RS_N = 15 };
enum { RC_C = 0, RC_AX = 3, RC_BY = 4, RC_H = 5, RC_Z0 = 6, RC_DIR = 7, RC_LOGP = 10, RC_LOGE = 11, RC_CORT = 12,
       RC_ECC = 13, RC_N = 16 };

__global__ __launch_bounds__(PDX_BLOCK) void k_raw_partials(PdxRawIn in, const int64_t *chunk, double *part)
{
    __shared__ double sh[RS_N][PDX_BLOCK / 64];
    const int64_t r0 = chunk[2 * blockIdx.x], r1 = chunk[2 * blockIdx.x + 1];
    double v[RS_N];
    for (int q = 0; q < 3; ++q) {
        v[RS_MIN + q] = INFINITY;
        v[RS_MAX + q] = -INFINITY;
    }
    for (int q = RS_F; q < RS_N; ++q) v[q] = 0.0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += PDX_BLOCK) {
        const double *x = in.X + 3 * r, *f = in.fext + 3 * r;
        const double E = in.E[r];
        for (int q = 0; q < 3; ++q) {
            v[RS_MIN + q] = fmin(v[RS_MIN + q], x[q]);
            v[RS_MAX + q] = fmax(v[RS_MAX + q], x[q]);
            v[RS_F + q] += f[q];
        }
        v[RS_NLOAD] += (fabs(f[0]) + fabs(f[1]) + fabs(f[2]) > 0.0) ? 1.0 : 0.0;
        v[RS_LOGE] += log10(E / 1000.0);
        v[RS_CORT] += (in.mat && in.mat[r] == 0) ? 1.0 : 0.0;
        v[RS_E] += E;
        v[RS_EX] += E * x[0];
        v[RS_EX + 1] += E * x[1];
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int q = 0; q < RS_N; ++q) {
        double t = v[q];
        for (int o = 32; o > 0; o >>= 1) {
            const double u = __shfl_xor(t, o, 64);
            t = q < RS_MAX ? fmin(t, u) : q < RS_F ? fmax(t, u) : t + u;
        }
        if (lane == 0) sh[q][wv] = t;
    }
    __syncthreads();
    if (threadIdx.x < RS_N) {
        const int q = threadIdx.x;
        double t = sh[q][0];
        for (int w = 1; w < PDX_BLOCK / 64; ++w)
            t = q < RS_MAX ? fmin(t, sh[q][w]) : q < RS_F ? fmax(t, sh[q][w]) : t + sh[q][w];
        part[(int64_t)blockIdx.x * RS_N + q] = t;
    }
}

__global__ void k_raw_cases(const double *part, const int32_t *chunk_first, const int64_t *case_off, int ncase,
                            double dx, double *cc)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncase) return;
    double v[RS_N];
    for (int q = 0; q < 3; ++q) {
        v[RS_MIN + q] = INFINITY;
        v[RS_MAX + q] = -INFINITY;
    }
    for (int q = RS_F; q < RS_N; ++q) v[q] = 0.0;
    for (int b = chunk_first[k]; b < chunk_first[k + 1]; ++b)
        for (int q = 0; q < RS_N; ++q) {
            const double t = part[(int64_t)b * RS_N + q];
            v[q] = q < RS_MAX ? fmin(v[q], t) : q < RS_F ? fmax(v[q], t) : v[q] + t;
        }
    const double cnt = (double)(case_off[k + 1] - case_off[k]);
    double *o = cc + (int64_t)k * RC_N;
    for (int q = 0; q < 3; ++q) o[RC_C + q] = 0.5 * (v[RS_MIN + q] + v[RS_MAX + q]);
    o[RC_AX] = fmax(0.5 * (v[RS_MAX] - v[RS_MIN]), 1e-9);
    o[RC_BY] = fmax(0.5 * (v[RS_MAX + 1] - v[RS_MIN + 1]), 1e-9);
    o[RC_H] = fmax(v[RS_MAX + 2] - v[RS_MIN + 2], 1e-9);
    o[RC_Z0] = v[RS_MIN + 2];
    const double nrm = sqrt(v[RS_F] * v[RS_F] + v[RS_F + 1] * v[RS_F + 1] + v[RS_F + 2] * v[RS_F + 2]);
    for (int q = 0; q < 3; ++q) o[RC_DIR + q] = nrm > 0.0 ? v[RS_F + q] / nrm : 0.0;
    // resultant pressure: |sum f| V / loaded area, area = n_loaded dx^2
    o[RC_LOGP] = nrm > 0.0 ? log10(fmax(nrm * dx * dx * dx / fmax(v[RS_NLOAD] * dx * dx, 1e-30), 1e-12)) : 0.0;
    o[RC_LOGE] = v[RS_LOGE] / cnt;
    o[RC_CORT] = v[RS_CORT] / cnt;
    const double es = fmax(v[RS_E], 1e-300);
    o[RC_ECC] = (v[RS_EX] / es - o[RC_C]) / o[RC_AX];
    o[RC_ECC + 1] = (v[RS_EX + 1] / es - o[RC_C + 1]) / o[RC_BY];
}

__global__ __launch_bounds__(PDX_BLOCK) void k_raw_fill(PdxRawIn in, const int64_t *case_off, int ncase,
                                                       const double *cc, double *out)
{
    const int64_t r = (int64_t)blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (r >= case_off[ncase]) return;
    int lo = 0, hi = ncase;  // case k with case_off[k] <= r < case_off[k + 1]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (case_off[mid] <= r) lo = mid;
        else hi = mid;
    }
    const double *o = cc + (int64_t)lo * RC_N;
    const double *x = in.X + 3 * r, *f = in.fext + 3 * r;
    double *y = out + r * 20;
    y[0] = (x[0] - o[RC_C]) / o[RC_AX];
    y[1] = (x[1] - o[RC_C + 1]) / o[RC_BY];
    y[2] = (x[2] - o[RC_Z0]) / o[RC_H];
    y[3] = log10(in.E[r] / 1000.0);
    y[4] = (in.mat && in.mat[r] == 0) ? 1.0 : 0.0;
    for (int q = 0; q < 3; ++q) y[5 + q] = (double)((in.fixed[r] >> q) & 1);
    y[8] = (fabs(f[0]) + fabs(f[1]) + fabs(f[2]) > 0.0) ? 1.0 : 0.0;
    for (int q = 0; q < 3; ++q) y[9 + q] = o[RC_DIR + q];
    y[12] = o[RC_LOGP];
    y[13] = o[RC_H] / 25.0;
    y[14] = o[RC_AX] / 20.0;
    y[15] = o[RC_BY] / 20.0;
    y[16] = o[RC_LOGE];
    y[17] = o[RC_CORT];
    y[18] = o[RC_ECC];
    y[19] = o[RC_ECC + 1];
}

// the context's positions, external forces and fixed-DOF bits in caller
// order (the arrays pdx_set_particles / pdx_set_bc stored), for the raw
// features of a caller that passes none of its own
__global__ __launch_bounds__(PDX_BLOCK) void k_ctx_rows(Dev d, double *X, double *fext, uint8_t *fixed)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const int r = d.perm[i];
    if (r < 0) return;
    const XRec &x = d.xu[0][i];
    const pdx_d4 f = d.fext[i];
    if (X)
        for (int q = 0; q < 3; ++q) X[(int64_t)r * 3 + q] = x.X[q];
    if (fext) {
        fext[(int64_t)r * 3] = f.x;
        fext[(int64_t)r * 3 + 1] = f.y;
        fext[(int64_t)r * 3 + 2] = f.z;
    }
    if (fixed) fixed[r] = (uint8_t)(d.KR[i].flags & PDX_FL_FIXMASK);
}

int pdx_launch_ctx_rows(pdx_ctx *c, double *d_X, double *d_fext, uint8_t *d_fixed)
{
    hipLaunchKernelGGL(k_ctx_rows, dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d, d_X, d_fext, d_fixed);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

int pdx_launch_raw_features(pdx_ctx *c, const PdxRawIn &in, const int64_t *d_chunk, int nchunk,
                            const int32_t *d_chunk_first, const int64_t *d_case_off, double dx, double *d_part,
                            double *d_cc, double *d_out)
{
    if (nchunk > 0) {
        hipLaunchKernelGGL(k_raw_partials, dim3(nchunk), dim3(PDX_BLOCK), 0, c->stream, in, d_chunk, d_part);
        PDX_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_raw_cases, dim3((c->ncase + 63) / 64), dim3(64), 0, c->stream, d_part, d_chunk_first,
                       d_case_off, c->ncase, dx, d_cc);
    PDX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_raw_fill, dim3((unsigned)((c->n + PDX_BLOCK - 1) / PDX_BLOCK)), dim3(PDX_BLOCK), 0,
                       c->stream, in, d_case_off, c->ncase, d_cc, d_out);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

// ---------------------------------------------------------------------------
// power iteration for the largest eigenvalue of M^-1 K (adaptive-DR mass)
// ---------------------------------------------------------------------------
// start vector: +-1 on every free DOF of an active particle (a hash of the
// DOF index), 0 elsewhere
__global__ __launch_bounds__(PDX_BLOCK) void k_pow_init(Dev d, pdx_d4 *v)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const KRec kr = d.KR[i];
    const uint32_t fm = kr.flags & PDX_FL_FIXMASK;
    const bool act = (kr.flags & PDX_FL_ACTIVE) != 0;
    double e[3] = {0.0, 0.0, 0.0};
    for (int r = 0; r < d.dim; ++r) {
        uint32_t h = (uint32_t)(3 * i + r) * 2654435761u;
        h ^= h >> 15;
        h *= 2246822519u;
        h ^= h >> 13;
        e[r] = (act && !((fm >> r) & 1u)) ? ((h & 1u) ? 1.0 : -1.0) : 0.0;
    }
    v[i] = d4(e[0], e[1], e[2]);
}

// u of buffer ubuf ^ 1 = u of ubuf + s_case v
__global__ __launch_bounds__(PDX_BLOCK) void k_pow_perturb(Dev d, int ubuf, const pdx_d4 *v, const double *scale)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const double s = scale[d.case_of_p[i] < 0 ? 0 : d.case_of_p[i]];
    const pdx_d4 vi = v[i];
    const double *ui = d.xu[ubuf][i].u;
    double *uo = d.xu[ubuf ^ 1][i].u;
    uo[0] = ui[0] + s * vi.x;
    uo[1] = ui[1] + s * vi.y;
    uo[2] = ui[2] + s * vi.z;
}

// K v = -(f - f0) / s_case on the free DOFs; block partials of v.Kv, lam v.v,
// |v_next|^2 and the free-DOF count; v_next = K v / lam
__global__ __launch_bounds__(PDX_BLOCK) void k_pow_apply(Dev d, const pdx_d4 *f0, pdx_d4 *v, const double *scale)
{
    const int lb = blockIdx.x;
    const int i = lb * PDX_BLOCK + threadIdx.x;
    double red[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < d.ni) {
        const KRec kr = d.KR[i];
        const uint32_t fm = kr.flags & PDX_FL_FIXMASK;
        if (kr.flags & PDX_FL_ACTIVE) {
            const double s = scale[d.case_of_blk[lb]];
            const double lam = d.lam[i];
            const pdx_d4 vi = v[i], fi = d.f[i], gi = f0[i];
            const double vv[3] = {vi.x, vi.y, vi.z}, ff[3] = {fi.x, fi.y, fi.z}, gg[3] = {gi.x, gi.y, gi.z};
            double vn[3] = {0.0, 0.0, 0.0};
            for (int r = 0; r < d.dim; ++r) {
                if ((fm >> r) & 1u) continue;
                const double kv = -(ff[r] - gg[r]) / s;
                red[0] += vv[r] * kv;
                red[1] += lam * vv[r] * vv[r];
                vn[r] = kv / lam;
                red[2] += vn[r] * vn[r];
                red[3] += 1.0;
            }
            v[i] = d4(vn[0], vn[1], vn[2]);
        }
    }
    block_partials<4>(d, lb, red);
}

// per-case totals of the 4 block partials into out[4 case + q]
__global__ __launch_bounds__(PDX_BLOCK) void k_pow_reduce(Dev d, double *out)
{
    __shared__ double sred[PDX_BLOCK / 64][4];
    const int cs = blockIdx.x;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int b0 = d.case_blk0[cs], b1 = d.case_blk0[cs + 1];
    for (int q = 0; q < 4; ++q) {
        double t = 0.0;
        for (int64_t e = (int64_t)b0 * PDX_PW + threadIdx.x; e < (int64_t)b1 * PDX_PW; e += PDX_BLOCK)
            t += d.part[e * 4 + q];
        t = wave_sum(t);
        if (lane == 0) sred[wid][q] = t;
        __syncthreads();
        if (threadIdx.x == 0) {
            double sum = 0.0;
            for (int w = 0; w < PDX_BLOCK / 64; ++w) sum += sred[w][q];
            out[4 * cs + q] = sum;
        }
        __syncthreads();
    }
}

// lam *= beta of the particle's case
__global__ __launch_bounds__(PDX_BLOCK) void k_scale_lam(Dev d, const double *beta)
{
    const int i = blockIdx.x * PDX_BLOCK + threadIdx.x;
    if (i >= d.ni) return;
    const int cs = d.case_of_p[i];
    if (cs >= 0) d.lam[i] *= beta[cs];
}

int pdx_launch_pow(pdx_ctx *c, int what, int ubuf, pdx_d4 *v, const pdx_d4 *f0, const double *scale, double *out)
{
    const dim3 g(grid1(c->ni)), b(PDX_BLOCK);
    switch (what) {
    case 0: hipLaunchKernelGGL(k_pow_init, g, b, 0, c->stream, c->d, v); break;
    case 1: hipLaunchKernelGGL(k_pow_perturb, g, b, 0, c->stream, c->d, ubuf, v, scale); break;
    case 2: hipLaunchKernelGGL(k_pow_apply, dim3(c->nblk), b, 0, c->stream, c->d, f0, v, scale); break;
    case 3: hipLaunchKernelGGL(k_pow_reduce, dim3(c->ncase), b, 0, c->stream, c->d, out); break;
    default: hipLaunchKernelGGL(k_scale_lam, g, b, 0, c->stream, c->d, scale); break;
    }
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

int pdx_launch_xu_geom(pdx_ctx *c, const XRec *d_src)
{
    hipLaunchKernelGGL(k_xu_geom, dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d, d_src);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

int pdx_launch_kr_put(pdx_ctx *c, const float *d_src, int off, int cnt)
{
    hipLaunchKernelGGL(k_kr_put, dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d, d_src, off, cnt);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

int pdx_launch_kr_merge(pdx_ctx *c, const uint32_t *d_src, int off, uint32_t keep)
{
    hipLaunchKernelGGL(k_kr_merge, dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d, d_src, off, keep);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

int pdx_launch_xu_c(pdx_ctx *c, const float *d_src)
{
    hipLaunchKernelGGL(k_xu_c, dim3(grid1(c->ni)), dim3(PDX_BLOCK), 0, c->stream, c->d, d_src);
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}

int pdx_launch_halo_build(pdx_ctx *c, int mode, const uint8_t *d_colour, int32_t *d_nh, const int32_t *d_ptr)
{
    const int nbk = c->ni / PDX_BLOCK;
    const size_t lds = sizeof(int32_t) * 2 * PDX_HB_HASH;
    if (c->dim == 3) {
        if (mode == 0)
            hipLaunchKernelGGL((k_halo_build<3, 0>), dim3(nbk), dim3(PDX_BLOCK), lds, c->stream, c->d, d_colour, d_nh, d_ptr);
        else
            hipLaunchKernelGGL((k_halo_build<3, 1>), dim3(nbk), dim3(PDX_BLOCK), lds, c->stream, c->d, d_colour, d_nh, d_ptr);
    } else {
        if (mode == 0)
            hipLaunchKernelGGL((k_halo_build<2, 0>), dim3(nbk), dim3(PDX_BLOCK), lds, c->stream, c->d, d_colour, d_nh, d_ptr);
        else
            hipLaunchKernelGGL((k_halo_build<2, 1>), dim3(nbk), dim3(PDX_BLOCK), lds, c->stream, c->d, d_colour, d_nh, d_ptr);
    }
    PDX_LAUNCH_CHECK();
    return PDX_OK;
}
