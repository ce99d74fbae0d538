// pdx_nn.hip -- BF16 MFMA kernels of the learned warm start (a per-particle
// MLP, 4 linear layers of width 256 with ReLU).  Training and inference run
// on these kernels; host code (Python) owns the device buffers and calls the
// C ABI below with device pointers on a given HIP stream.
//
// Matrices are row-major; transposed operands (W^T in the data gradient, X^T in
// the weight gradient) are read through strides or transpose reads, never
// copied.  The layers' shapes run on gfx950's v_mfma_f32_32x32x16_bf16 with
// FP32 accumulation:
//   k_nn_gemm_lds  forward and data gradient, the weight resident in LDS;
//   k_nn_wgrad     weight gradient, the reduction over the particle rows with
//                  LDS transpose reads, in fixed-size row slices whose FP32
//                  partials are summed in a fixed order (k_nn_sum_slices4):
//                  training is deterministic.
// k_nn_gemm, one 32x32 tile per wave on v_mfma_f32_32x32x8_bf16 with operands
// loaded element by element, serves every other shape or alignment.
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef short pdx_bf4 __attribute__((ext_vector_type(4)));
typedef float pdx_f16v __attribute__((ext_vector_type(16)));

#define NN_TM 128  // rows per block (4 waves x 32)
#define NN_TN 32   // columns per block

// FP32 -> BF16, round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

enum { NN_EPI_FWD = 0, NN_EPI_DGRAD = 1, NN_EPI_WGRAD = 2 };

// C[M x N] = A[M x K] B[K x N]; A(i, k) = A[i * sa0 + k * sa1], B(k, j) = B[k * sb0 + j * sb1].
// EPI_FWD:   y = acc + bias[j], ReLU if relu; writes ybf (ldy) and/or yf (ldy)
// EPI_DGRAD: y = acc, zeroed where act(i, j) <= 0 (ReLU derivative of the layer input) if act
// EPI_WGRAD: rows of the reduction over [z * kslice, (z + 1) * kslice); FP32 partial to yf + z * M * N
template <int EPI>
__global__ __launch_bounds__(256) void k_nn_gemm(const uint16_t *__restrict__ A, int64_t sa0, int64_t sa1,
                                                 const uint16_t *__restrict__ B, int64_t sb0, int64_t sb1, int M,
                                                 int N, int K, int kslice, const float *__restrict__ bias, int relu,
                                                 const uint16_t *__restrict__ act, int64_t lda_act,
                                                 uint16_t *__restrict__ ybf, float *__restrict__ yf, int64_t ldy)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int row0 = blockIdx.x * NN_TM + wv * 32, col0 = blockIdx.y * NN_TN;
    int k_lo = 0, k_hi = K;
    if (EPI == NN_EPI_WGRAD) {
        k_lo = blockIdx.z * kslice;
        k_hi = min(K, k_lo + kslice);
    }
    const int r = row0 + (lane & 31), c = col0 + (lane & 31), kq = (lane >> 5) * 4;
    pdx_f16v acc;
    for (int v = 0; v < 16; ++v) acc[v] = 0.f;
    for (int k0 = k_lo; k0 < k_hi; k0 += 8) {
        pdx_bf4 a, b;
        for (int q = 0; q < 4; ++q) {
            const int k = k0 + kq + q;
            const bool kin = k < k_hi;
            a[q] = (short)((kin && r < M) ? A[(int64_t)r * sa0 + (int64_t)k * sa1] : 0);
            b[q] = (short)((kin && c < N) ? B[(int64_t)k * sb0 + (int64_t)c * sb1] : 0);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    // D layout of the 32x32 MFMA: lane l, register v -> row 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32
    const int j = col0 + (lane & 31);
    if (j >= N) return;
    for (int v = 0; v < 16; ++v) {
        const int i = row0 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (i >= M) continue;
        float y = acc[v];
        if (EPI == NN_EPI_FWD) {
            if (bias) y += bias[j];
            if (relu) y = fmaxf(y, 0.f);
        } else if (EPI == NN_EPI_DGRAD) {
            if (act && !(bf2f(act[(int64_t)i * lda_act + j]) > 0.f)) y = 0.f;
        }
        if (EPI == NN_EPI_WGRAD) {
            yf[(int64_t)blockIdx.z * M * N + (int64_t)i * N + j] = y;
        } else {
            if (ybf) ybf[(int64_t)i * ldy + j] = f2bf(y);
            if (yf) yf[(int64_t)i * ldy + j] = y;
        }
    }
}

// ---------------------------------------------------------------------------
// The layer GEMMs of the forward pass and the data gradient on the gfx950
// 32x32x16 BF16 MFMA, with the whole weight operand resident in LDS.
//
// C (M x NC) = A (M x KR) B (KR x NC), KR <= 256 the reduction (a layer's input
// or output width), NC <= 256 the output columns.  A is row-major with
// 16-byte aligned rows; B(k, j) = B[k * sb0 + j * sb1] (W, or W^T for the data
// gradient).  A persistent grid of at most one block per CU stages B once,
// transposed to column-major (the 16-byte k chunks of each column XOR-swizzled
// so that staging stores and operand reads both spread over the LDS banks),
// then walks tiles of 128 rows.  Each of the 4 waves (one per SIMD) owns RT
// row tiles x CT column tiles of 32 x 32 (WC waves across the columns).
//
// The tile loop is software-pipelined so that one wave per SIMD keeps the
// MFMA unit, the loads and the stores busy at once.  During tile t's k steps a
// lane
//   * reloads each A operand register with tile t + PF G's value (G the grid;
//     PF sets of operands are in flight) right after its MFMAs are issued (8
//     consecutive values of a row per 16-wide k step, 16-byte loads, every row
//     read once),
//   * stores the previous tile's results, packed in the store layout at its
//     end (below), a few 16-byte row pieces per k step,
//   * (data gradient) loads tile t's ReLU mask in the first half of the steps,
//   * reads the next k step's B operands from LDS.
// Every global access goes through a raw buffer descriptor of the tile's rows
// (rows [128 t, min(M, 128 t + 128))): loads past the end return zero and
// stores past it are dropped, so the loop issues the same memory instructions
// on every tile, without branches, and the wait counts the compiler derives
// for the MFMAs are those of the steady state (the prologue ends with the
// first tile's empty-descriptor stores for the same reason).
// MFMA operand layout (32x32x16): lane l holds row / column l % 32 and k
// 8 (l / 32) .. 8 (l / 32) + 7; result register v of lane l is row
// 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32.  The operands are passed
// swapped (C^T = B^T A^T), so a lane's registers 4 g .. 4 g + 3 of a tile are
// columns 8 g + 4 (l / 32) .. + 3 of its row l % 32: 8- or 16-byte writes
// into the staging rows.
// ---------------------------------------------------------------------------
typedef __bf16 pdx_bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 pdx_bf4v __attribute__((ext_vector_type(4)));
typedef uint32_t pdx_u4v __attribute__((ext_vector_type(4)));

enum { NN_OUT_BF16 = 0, NN_OUT_F32 = 1 };

// raw buffer descriptor of `bytes` bytes at p (word 3: the gfx9 family's
// untyped 32-bit data format); loads past the end return zero, stores past it
// are dropped
__device__ __forceinline__ __amdgpu_buffer_rsrc_t nn_rsrc(const void *p, int bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, bytes, 0x00020000);
}

template <int EPI, int OUT, int RELU, int KR, int NC, int RT, int CT, int WC, int PF>
__global__ __launch_bounds__(256) void k_nn_gemm_lds(const uint16_t *__restrict__ A, int64_t lda,
                                                     const uint16_t *__restrict__ B, int64_t sb0, int64_t sb1, int M,
                                                     const float *__restrict__ bias,
                                                     const uint16_t *__restrict__ act, int64_t lda_act,
                                                     void *__restrict__ Y, int64_t ldy)
{
    static_assert(KR % 16 == 0 && KR <= 256 && (4 / WC) * RT * 32 == 128 && WC * CT * 32 == NC && PF >= 1,
                  "tile shape");
    constexpr int KP = KR;                                      // LDS row of one B column (bf16 values)
    // the 16-byte k chunks of column j sit XOR-swizzled: chunk c at position
    // c ^ swz(j), which spreads both the staging stores (8 columns apart) and the
    // operand reads (consecutive columns) over all LDS banks
    constexpr int FM = KR / 8 - 1 < 7 ? KR / 8 - 1 : 7;
    auto swz = [](int j) { return (j + (j >> 3)) & FM; };
    constexpr int KS = KR / 16;                                 // k steps
    constexpr int ESZ = OUT == NN_OUT_F32 ? 4 : 2;              // output element size
    // The results leave through store units of 32 rows x UC columns (128 bytes
    // of a row, or the wave's width): written to the wave's LDS staging rows in
    // the MFMA layout, read back 16 bytes per lane, LPR lanes per row, and
    // stored as whole row segments.
    constexpr int UC = 128 / ESZ < CT * 32 ? 128 / ESZ : CT * 32;  // columns of a store unit
    constexpr int UPR = CT * 32 / UC;                           // units per row tile
    constexpr int NU = RT * UPR;                                // units per tile
    constexpr int LPR = UC * ESZ / 16;                          // lanes per row of a store
    constexpr int RPI = 64 / LPR;                               // rows per store instruction
    constexpr int NP = 32 / RPI;                                // store instructions per unit
    constexpr int NST = NU * NP;                                // 16-byte stores per lane per tile
    constexpr int RS = UC * ESZ + 16;                           // staging row stride (bytes)
    constexpr int GPS = (NST + KS - 1) / KS;                    // stores per k step
    constexpr int MH = (KS + 1) / 2;                            // k steps that load the mask
    constexpr int MPS = (NST + MH - 1) / MH;                    // mask loads per such step
    constexpr int NMK = EPI == NN_EPI_DGRAD ? NST : 1;
    static_assert(UC % 32 == 0 && (UC * ESZ) % 16 == 0 && 64 % LPR == 0 && 32 % RPI == 0, "store units");
    // B operands held in registers for the whole kernel when they fit (64
    // VGPRs); else read from LDS one k step ahead
    constexpr bool BREG = KS * CT <= 16;
    extern __shared__ uint16_t bt[];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = wv / WC, wc = wv % WC;
    const int l32 = lane & 31, kh = (lane >> 5) * 8;
    const int ntile = (M + 127) / 128;
    const int lrow = wr * RT * 32 + l32;         // the lane's row in row tile 0 of a tile
    const int jl = wc * CT * 32 + 4 * (lane >> 5);  // first column of the lane's groups
    // byte offsets within a tile (the host keeps 128 rows of A, Y and the mask under 2^31 bytes)
    int offa[RT];
#pragma unroll
    for (int a = 0; a < RT; ++a) offa[a] = (lrow + 32 * a) * (int)lda * 2 + kh * 2;
    // store q (unit u = q / NP, piece p = q % NP): row (wr RT + a) 32 + p RPI + lane / LPR of the
    // tile, columns wc CT 32 + ub UC + (lane % LPR) 16 / ESZ .. + 16 / ESZ - 1 (a = u / UPR, ub = u % UPR)
    const int srow = wr * RT * 32 + lane / LPR, scol = wc * CT * 32 + (lane % LPR) * (16 / ESZ);
    const int offy0 = srow * (int)ldy * ESZ + scol * ESZ, offm0 = srow * (int)lda_act * 2 + scol * 2;
    auto qrow = [](int q) { return (q / NP / UPR) * 32 + (q % NP) * RPI; };
    auto qcol = [](int q) { return (q / NP % UPR) * UC; };
    auto rows = [&](int t) { return t < ntile ? min(128, M - 128 * t) : 0; };
    auto a_rsrc = [&](int t) {
        return nn_rsrc(A + (int64_t)128 * (t < ntile ? t : 0) * lda, rows(t) * (int)lda * 2);
    };

    // prologue: the A operands of the block's first PF tiles, in the loop's order
    // (set p holds the tiles p, p + PF, p + 2 PF, ... of the block)
    uint4 af[PF][RT][KS];
    int tile = blockIdx.x;
#pragma unroll
    for (int p = 0; p < PF; ++p) {
        const __amdgpu_buffer_rsrc_t ra = a_rsrc(tile + p * (int)gridDim.x);
#pragma unroll
        for (int s = 0; s < KS; ++s)
#pragma unroll
            for (int a = 0; a < RT; ++a)
                af[p][a][s] =
                    __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(ra, offa[a] + 32 * s, 0, 0));
    }
    // B -> LDS, column-major: B(k, j) at bt[j * KP + 8 ((k / 8) ^ swz(j)) + k % 8], 16-byte loads along
    // the contiguous index
    // (a thread's loads go out in batches of 8 before their LDS stores)
    constexpr int NL = (KR * NC / 8 + 255) / 256;
    static_assert(KR % 8 == 0 && NC % 8 == 0, "8 x 8 staging blocks");
    if (sb1 == 1 && (sb0 & 7) == 0 && !(reinterpret_cast<uintptr_t>(B) & 15)) {
        // row-major B: a thread takes 8 x 8 sub-blocks (8 rows of 8 values, 16-byte
        // loads), transposes them in registers and writes 8 columns of 8 k values
        // (16-byte LDS stores)
        constexpr int NB = (KR / 8) * (NC / 8);
        for (int blk = tid; blk < NB; blk += 256) {
            const int kb = blk / (NC / 8), jb = blk % (NC / 8);
            uint32_t w[8][4];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const uint4 r = *reinterpret_cast<const uint4 *>(B + (int64_t)(8 * kb + i) * sb0 + 8 * jb);
                w[i][0] = r.x;
                w[i][1] = r.y;
                w[i][2] = r.z;
                w[i][3] = r.w;
            }
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                uint32_t o[4];
#pragma unroll
                for (int p2 = 0; p2 < 4; ++p2) {
                    const uint32_t lo = w[2 * p2][c >> 1], hi = w[2 * p2 + 1][c >> 1];
                    o[p2] = (c & 1) ? ((lo >> 16) | (hi & 0xffff0000u)) : ((lo & 0xffffu) | (hi << 16));
                }
                *reinterpret_cast<uint4 *>(&bt[(8 * jb + c) * KP + 8 * (kb ^ swz(8 * jb + c))]) =
                    make_uint4(o[0], o[1], o[2], o[3]);
            }
        }
    } else if (sb0 == 1 && (sb1 & 7) == 0 && !(reinterpret_cast<uintptr_t>(B) & 15)) {
        for (int t0 = 0; t0 < NL; t0 += 8) {
            uint4 v[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int e = tid + 256 * (t0 + t);
                v[t] = e < NC * (KR / 8)
                           ? *reinterpret_cast<const uint4 *>(B + (int64_t)(e / (KR / 8)) * sb1 + (e % (KR / 8)) * 8)
                           : make_uint4(0, 0, 0, 0);
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int e = tid + 256 * (t0 + t);
                const int j = e / (KR / 8);
                if (e < NC * (KR / 8)) *reinterpret_cast<uint4 *>(&bt[j * KP + 8 * ((e % (KR / 8)) ^ swz(j))]) = v[t];
            }
        }
    } else {
        for (int e = tid; e < KR * NC; e += 256) {
            const int k = e / NC, j = e % NC;
            bt[j * KP + 8 * ((k >> 3) ^ swz(j)) + (k & 7)] = B[(int64_t)k * sb0 + (int64_t)j * sb1];
        }
    }
    // the bias in LDS too, read in every tile's epilogue
    float *bl = reinterpret_cast<float *>(bt + NC * KP);
    if (EPI == NN_EPI_FWD)
        for (int j = tid; j < NC; j += 256) bl[j] = bias ? bias[j] : 0.f;
    __syncthreads();

    // the previous tile's packed results and their descriptor: before the first
    // tile, zeros through an empty descriptor (the stores are dropped)
    pdx_u4v pv[NST];
#pragma unroll
    for (int q = 0; q < NST; ++q) pv[q] = pdx_u4v{};
    __amdgpu_buffer_rsrc_t ry = nn_rsrc(Y, 0);
    const int ldyb = (int)ldy * ESZ;
    // results are written once and not read again by this launch: non-temporal
    // stores (cache-policy bit NT), which measured faster for a single layer and
    // for the chained network
    auto store_group = [&](int q) {
        __builtin_amdgcn_raw_buffer_store_b128(pv[q], ry, offy0 + qrow(q) * ldyb + qcol(q) * ESZ, 0, 2);
    };
#pragma unroll
    for (int q = 0; q < NST; ++q) store_group(q);
    // this wave's staging rows
    char *stg = reinterpret_cast<char *>(bt + NC * KP) + sizeof(float) * NC + wv * 32 * RS;
    const bool use_mask = EPI == NN_EPI_DGRAD && act != nullptr;
    // the bias of the lane's columns, in registers for the whole kernel (read
    // from LDS in the epilogue where the FP32 results leave no room)
    constexpr bool BIASREG = EPI == NN_EPI_FWD && CT <= 4 && (OUT == NN_OUT_BF16 || CT == 1);
    float4 bv[BIASREG ? CT * 4 : 1];
    if constexpr (BIASREG)
#pragma unroll
        for (int q = 0; q < CT * 4; ++q) bv[q] = *reinterpret_cast<const float4 *>(bl + jl + (q / 4) * 32 + 8 * (q % 4));
    // the B operands of k step s (the same for every tile): column j_b = (wc CT + b) 32 + l % 32, chunk
    // 2 s + l / 32, i.e. element offset boff + bsw[b][s % 4] + 64 (s / 4) (the swizzle acts on the low
    // 3 bits of the chunk)
    int boff = 0, bsw[CT][4];
#pragma unroll
    for (int b = 0; b < CT; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = (wc * CT + b) * 32 + l32;
            bsw[b][q] = j * KP + 8 * ((2 * q + (kh >> 3)) ^ swz(j));
        }
    auto read_b = [&](int s, pdx_bf8 (&bf)[CT]) {
#pragma unroll
        for (int b = 0; b < CT; ++b)
            bf[b] = __builtin_bit_cast(pdx_bf8,
                                       *reinterpret_cast<const uint4 *>(&bt[boff + bsw[b][s & 3] + 64 * (s >> 2)]));
    };
    pdx_bf8 bres[BREG ? KS : 1][CT], bfr[2][CT];
    if constexpr (BREG) {
#pragma unroll
        for (int s = 0; s < KS; ++s) read_b(s, bres[s]);
    } else {
        read_b(0, bfr[0]);
    }
    // one tile t with operand set p: its MFMAs, its operands reloaded with tile t + PF G's, the
    // previous tile's stores; then its results packed for the stores
    auto process = [&](const int t, uint4 (&afc)[RT][KS]) {
        // an opaque LDS offset per tile keeps the per-step reads inside the loop
        if constexpr (!BREG) asm volatile("" : "+v"(boff));
        const __amdgpu_buffer_rsrc_t ra = a_rsrc(t + PF * (int)gridDim.x);
        const __amdgpu_buffer_rsrc_t rm =
            nn_rsrc(use_mask ? act + (int64_t)128 * t * lda_act : act, use_mask ? rows(t) * (int)lda_act * 2 : 0);
        pdx_u4v mk[NMK];
        pdx_f16v acc[RT][CT];
#pragma unroll
        for (int a = 0; a < RT; ++a)
#pragma unroll
            for (int b = 0; b < CT; ++b)
#pragma unroll
                for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            // the next step's B operands (step 0 of the next tile after the last)
            if constexpr (!BREG) read_b((s + 1) % KS, bfr[(s + 1) & 1]);
#pragma unroll
            for (int a = 0; a < RT; ++a) {
                const pdx_bf8 av = __builtin_bit_cast(pdx_bf8, afc[a][s]);
#pragma unroll
                for (int b = 0; b < CT; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(BREG ? bres[BREG ? s : 0][b] : bfr[s & 1][b],
                                                                        av, acc[a][b], 0, 0, 0);
                afc[a][s] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(ra, offa[a] + 32 * s, 0, 0));
            }
            if (EPI == NN_EPI_DGRAD)
#pragma unroll
                for (int q = s * MPS; q < s * MPS + MPS; ++q)
                    if (q < NST)
                        mk[q] = __builtin_amdgcn_raw_buffer_load_b128(rm, offm0 + qrow(q) * (int)lda_act * 2 + qcol(q) * 2,
                                                                      0, 0);
#pragma unroll
            for (int q = s * GPS; q < s * GPS + GPS; ++q)
                if (q < NST) store_group(q);
            // each step's loads and stores stay among its MFMAs
            __builtin_amdgcn_sched_barrier(0);
        }
        if (!BREG && (KS & 1)) {  // the step-0 operands were read into bfr[KS & 1]; keep them in bfr[0]
#pragma unroll
            for (int b = 0; b < CT; ++b) bfr[0][b] = bfr[1][b];
        }
        // epilogue of this tile: bias and ReLU (forward), rounding, and through
        // the staging rows into the store layout, one unit at a time (a wave's
        // LDS operations complete in order, so its own lanes read back what they
        // wrote without a barrier)
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            const int a = u / UPR, ub = u % UPR;
            // the unit's bias, read together ahead of its use unless held in registers
            // (or, for FP32 results, read at its use: they leave no registers for it)
            constexpr bool BIASU = EPI == NN_EPI_FWD && !BIASREG && OUT == NN_OUT_BF16;
            float4 bu[UC / 32][4];
            if constexpr (BIASU)
#pragma unroll
                for (int b = 0; b < UC / 32; ++b)
#pragma unroll
                    for (int g = 0; g < 4; ++g)
                        bu[b][g] = *reinterpret_cast<const float4 *>(bl + jl + (ub * UC / 32 + b) * 32 + 8 * g);
#pragma unroll
            for (int b = ub * UC / 32; b < (ub + 1) * UC / 32; ++b)
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    float y[4] = {acc[a][b][4 * g], acc[a][b][4 * g + 1], acc[a][b][4 * g + 2], acc[a][b][4 * g + 3]};
                    if (EPI == NN_EPI_FWD) {
                        const float4 bb = BIASREG ? bv[BIASREG ? b * 4 + g : 0]
                                          : BIASU ? bu[b - ub * UC / 32][g]
                                                  : *reinterpret_cast<const float4 *>(bl + jl + b * 32 + 8 * g);
                        y[0] += bb.x;
                        y[1] += bb.y;
                        y[2] += bb.z;
                        y[3] += bb.w;
                        if (RELU)
#pragma unroll
                            for (int r = 0; r < 4; ++r) y[r] = fmaxf(y[r], 0.f);
                    }
                    char *dst = stg + l32 * RS + (b * 32 - ub * UC + 8 * g + 4 * (lane >> 5)) * ESZ;
                    if constexpr (OUT == NN_OUT_F32) {
                        *reinterpret_cast<float4 *>(dst) = make_float4(y[0], y[1], y[2], y[3]);
                    } else {
                        pdx_bf4v o;
#pragma unroll
                        for (int r = 0; r < 4; ++r) o[r] = (__bf16)y[r];  // round to nearest even
                        *reinterpret_cast<pdx_bf4v *>(dst) = o;
                    }
                }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int p = 0; p < NP; ++p)
                pv[u * NP + p] = *reinterpret_cast<const pdx_u4v *>(stg + (p * RPI + lane / LPR) * RS + (lane % LPR) * 16);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        // the data gradient's ReLU derivative: zero where the layer input is not > 0
        // (a BF16 h is > 0 exactly when h - 1 < 0x7f80 as unsigned: sign clear, not
        // zero, not NaN)
        if (use_mask) {
#pragma unroll
            for (int q = 0; q < NST; ++q)
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    const uint32_t m = mk[EPI == NN_EPI_DGRAD ? q : 0][w];
                    const uint32_t keep = (((m & 0xffffu) - 1u) < 0x7f80u ? 0x0000ffffu : 0u) |
                                          (((m >> 16) - 1u) < 0x7f80u ? 0xffff0000u : 0u);
                    pv[q][w] &= keep;
                }
        }
        ry = nn_rsrc(static_cast<char *>(Y) + (int64_t)128 * t * ldy * ESZ, rows(t) * (int)ldy * ESZ);
    };
    for (; tile < ntile; tile += PF * (int)gridDim.x) {
#pragma unroll
        for (int p = 0; p < PF; ++p) {
            if (p > 0 && tile + p * (int)gridDim.x >= ntile) break;
            process(tile + p * (int)gridDim.x, af[p]);
        }
    }
    // the last tile's results
#pragma unroll
    for (int q = 0; q < NST; ++q) store_group(q);
}

// persistent blocks: at most one per CU (the LDS-resident B allows one)
static int nn_persistent_blocks(int ntile)
{
    static int ncu = 0;
    if (ncu == 0) {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                     hipSuccess || n <= 0)
            n = 256;
        ncu = n;
    }
    return ntile < ncu ? ntile : ncu;
}

// bytes of a store unit's row: 128, or a wave's columns when fewer
static int nn_unit_bytes(int esz, int nc_wave) { return 128 < nc_wave * esz ? 128 : nc_wave * esz; }

// the LDS-resident-B kernel for the shapes of the network's layers, one
// output (BF16 or FP32) per launch; false: not one of them (the generic
// kernel runs)
template <int EPI>
static bool nn_gemm_lds(const uint16_t *A, int64_t lda, const uint16_t *B, int64_t sb0, int64_t sb1, int M, int KR,
                        int NC, const float *bias, int relu, const uint16_t *act, int64_t lda_act, uint16_t *ybf,
                        float *yf, int64_t ldy, hipStream_t st)
{
    if ((ybf != nullptr) == (yf != nullptr) || (EPI == NN_EPI_DGRAD && !ybf)) return false;
    if ((reinterpret_cast<uintptr_t>(A) & 15) || (lda % 8) || M <= 0 || (ldy % 4) ||
        (reinterpret_cast<uintptr_t>(ybf) & 7) || (reinterpret_cast<uintptr_t>(yf) & 15) ||
        (reinterpret_cast<uintptr_t>(bias) & 15) || (reinterpret_cast<uintptr_t>(act) & 7) || (lda_act % 4))
        return false;
    // byte offsets within 128 rows of A, Y and the mask stay below 2^31
    if (lda > (1 << 20) || ldy > (1 << 20) || lda_act > (1 << 20)) return false;
    const dim3 g(nn_persistent_blocks((M + 127) / 128));
    void *y = yf ? (void *)yf : (void *)ybf;
    // one launch: the block shape (RT, CT, WC) and the prefetch depth PF of a variant
    auto launch = [&](auto kern, int kr, int nc, int wc) {
        const size_t lds = sizeof(uint16_t) * (size_t)nc * kr + sizeof(float) * nc +
                           4 * 32 * (size_t)(nn_unit_bytes(yf ? 4 : 2, nc / wc) + 16);
        hipLaunchKernelGGL(kern, g, dim3(256), lds, st, A, lda, B, sb0, sb1, M, bias, act, lda_act, y, ldy);
        return true;
    };
#define NN_LDS_VARIANTS(kr, nc, f32, bf16, dgrad)                                                                \
    if (KR == kr && NC == nc) {                                                                                  \
        if constexpr (EPI == NN_EPI_FWD) {                                                                       \
            if (yf && relu) return launch(k_nn_gemm_lds<EPI, NN_OUT_F32, 1, kr, nc, NN_UNPACK f32>, kr, nc, NN_WC f32); \
            if (yf) return launch(k_nn_gemm_lds<EPI, NN_OUT_F32, 0, kr, nc, NN_UNPACK f32>, kr, nc, NN_WC f32);     \
            if (relu) return launch(k_nn_gemm_lds<EPI, NN_OUT_BF16, 1, kr, nc, NN_UNPACK bf16>, kr, nc, NN_WC bf16); \
            return launch(k_nn_gemm_lds<EPI, NN_OUT_BF16, 0, kr, nc, NN_UNPACK bf16>, kr, nc, NN_WC bf16);        \
        } else {                                                                                                 \
            return launch(k_nn_gemm_lds<EPI, NN_OUT_BF16, 0, kr, nc, NN_UNPACK dgrad>, kr, nc, NN_WC dgrad);      \
        }                                                                                                        \
    }
#define NN_UNPACK(rt, ct, wc, pf) rt, ct, wc, pf
#define NN_WC(rt, ct, wc, pf) wc
    // (RT, CT, WC, PF) of the FP32-output forward, the BF16 forward and the data gradient: one wave
    // across all columns where its registers allow, so that every A row is loaded once per block
    NN_LDS_VARIANTS(256, 256, (1, 8, 1, 1), (1, 8, 1, 2), (1, 8, 1, 1))
    NN_LDS_VARIANTS(64, 256, (2, 4, 2, 1), (2, 4, 2, 1), (2, 4, 2, 1))
    NN_LDS_VARIANTS(32, 256, (2, 4, 2, 1), (2, 4, 2, 1), (2, 4, 2, 1))
    NN_LDS_VARIANTS(256, 64, (2, 1, 2, 1), (2, 1, 2, 1), (2, 1, 2, 1))
    NN_LDS_VARIANTS(256, 32, (1, 1, 1, 1), (1, 1, 1, 1), (1, 1, 1, 1))
#undef NN_UNPACK
#undef NN_WC
#undef NN_LDS_VARIANTS
    return false;
}

// ---------------------------------------------------------------------------
// The weight gradient dW (K x N) = X^T dY over a slice of rows, on the 32x32x16
// BF16 MFMA with the reduction over the rows m.  The operands need 8 consecutive
// rows of one column per lane; the gfx950 LDS transpose read (ds_read_b64_tr_b16)
// gives exactly that from rows staged as they lie in memory: within each group
// of 16 lanes, lane i receives element i % 4 of the 8-byte pieces addressed by
// lanes 4 q + i / 4, q = 0..3.  Lane j = 4 q + t of group g (lanes 16 g' ..) of
// half h (lanes 32 h ..) addresses row 8 h + 4 p + q, columns c0 + 16 (g' % 2) +
// 4 t .. + 3 of the staged rows; two such reads (p = 0, 1) give lane l the rows
// 8 (l / 32) .. + 7 of column c0 + l % 32, the operand layout of the MFMA with
// the rows as its k.
// A block takes one slice of rows and the whole K x N product: WI x WJ waves
// over the rows x columns of dW, each TI x TJ tiles of 32 x 32.  Stages
// of 32 rows of X and dY pass through two LDS buffers and two register slots:
// a stage is loaded (16-byte buffer loads, zero past the slice) three stages
// ahead, written to LDS one stage ahead, one barrier per stage.  The
// 16-byte chunks of row r sit XOR-swizzled by 2 (r % 4), so that the four rows
// a transpose read touches fall on distinct banks.  The FP32 partial of the
// slice goes to part; the slices are summed in a fixed order by k_nn_sum_slices4.
// ---------------------------------------------------------------------------
typedef short pdx_s4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) pdx_s4 *pdx_lds_s4p;

template <int K, int N, int WI, int WJ>
__global__ __launch_bounds__(64 * WI * WJ) void k_nn_wgrad(const uint16_t *__restrict__ X, const uint16_t *__restrict__ dY,
                                                  int M, int R, float *__restrict__ part)
{
    constexpr int TI = K / WI / 32, TJ = N / WJ / 32;
    constexpr int NT = 64 * WI * WJ;              // threads
    static_assert(TI >= 1 && TJ >= 1 && TI * WI * 32 == K && TJ * WJ * 32 == N, "wgrad shape");
    constexpr int S = 32;                          // rows per stage
    constexpr int CX = K / 8, CY = N / 8;          // 16-byte chunks per row
    constexpr int XB = S * K * 2, YB = S * N * 2;  // stage bytes
    constexpr int NX = (S * CX + NT - 1) / NT, NY = (S * CY + NT - 1) / NT;  // loads per thread per stage
    constexpr int SB = XB + YB;                    // one LDS buffer
    extern __shared__ __attribute__((aligned(16))) char wsm[];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wi = wv / WJ, wj = wv % WJ;
    const int z = blockIdx.x, row0 = z * R;
    const int nrows = min(R, M - row0);
    const int ns = (nrows + S - 1) / S;
    const __amdgpu_buffer_rsrc_t rx = nn_rsrc(X + (int64_t)row0 * K, nrows * K * 2);
    const __amdgpu_buffer_rsrc_t ryy = nn_rsrc(dY + (int64_t)row0 * N, nrows * N * 2);
    auto swz = [](int r, int ch) { return (2 * (r & 3)) & (ch - 1); };
    uint4 sx[2][NX], sy[2][NY];
    auto load = [&](int st, uint4 (&vx)[NX], uint4 (&vy)[NY]) {
#pragma unroll
        for (int u = 0; u < NX; ++u) {
            const int e = tid + NT * u;  // chunk e of the stage: row e / CX, chunk e % CX
            vx[u] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(
                                                  rx, (st * S + e / CX) * K * 2 + (e % CX) * 16, 0, 0));
        }
#pragma unroll
        for (int u = 0; u < NY; ++u) {
            const int e = tid + NT * u;
            vy[u] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(
                                                  ryy, (st * S + e / CY) * N * 2 + (e % CY) * 16, 0, 0));
        }
    };
    auto put = [&](int buf, const uint4 (&vx)[NX], const uint4 (&vy)[NY]) {
        char *b = wsm + buf * SB;
#pragma unroll
        for (int u = 0; u < NX; ++u) {
            const int e = tid + NT * u, r = e / CX, c = e % CX;
            if (S * CX % NT == 0 || e < S * CX)
                *reinterpret_cast<uint4 *>(b + r * K * 2 + (c ^ swz(r, CX)) * 16) = vx[u];
        }
#pragma unroll
        for (int u = 0; u < NY; ++u) {
            const int e = tid + NT * u, r = e / CY, c = e % CY;
            if (S * CY % NT == 0 || e < S * CY)
                *reinterpret_cast<uint4 *>(b + XB + r * N * 2 + (c ^ swz(r, CY)) * 16) = vy[u];
        }
    };
    // the transpose-read address of this lane within a stage, for k step ks, part p and the 32 columns from c0
    const int h = lane >> 5, gq = (lane >> 4) & 1, q = (lane >> 2) & 3, t = lane & 3;
    auto tr_off = [&](int ks, int p, int c0, int rowbytes, int ch) {
        const int r = 16 * ks + 8 * h + 4 * p + q, col = c0 + 16 * gq + 4 * t;
        return r * rowbytes + (((col >> 3) ^ swz(r, ch)) << 4) + ((col >> 2) & 1) * 8;
    };
    pdx_f16v acc[TI][TJ];
#pragma unroll
    for (int a = 0; a < TI; ++a)
#pragma unroll
        for (int b = 0; b < TJ; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;
    auto compute = [&](int buf) {
        const char *bx = wsm + buf * SB, *by = bx + XB;
#pragma unroll
        for (int ks = 0; ks < S / 16; ++ks) {
            pdx_bf8 fa[TI], fb[TJ];
#pragma unroll
            for (int a = 0; a < TI; ++a) {
                const int c0 = (wi * TI + a) * 32;
                const pdx_s4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((pdx_lds_s4p)(bx + tr_off(ks, 0, c0, K * 2, CX)));
                const pdx_s4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((pdx_lds_s4p)(bx + tr_off(ks, 1, c0, K * 2, CX)));
                fa[a] = __builtin_bit_cast(pdx_bf8, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
            }
#pragma unroll
            for (int b = 0; b < TJ; ++b) {
                const int c0 = (wj * TJ + b) * 32;
                const pdx_s4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((pdx_lds_s4p)(by + tr_off(ks, 0, c0, N * 2, CY)));
                const pdx_s4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((pdx_lds_s4p)(by + tr_off(ks, 1, c0, N * 2, CY)));
                fb[b] = __builtin_bit_cast(pdx_bf8, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
            }
#pragma unroll
            for (int a = 0; a < TI; ++a)
#pragma unroll
                for (int b = 0; b < TJ; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
        }
    };
    // prologue: stages 0 and 1 into the register slots, stage 0 into buffer 0, stage 2 into slot 0
    load(0, sx[0], sy[0]);
    load(1, sx[1], sy[1]);
    put(0, sx[0], sy[0]);
    load(2, sx[0], sy[0]);
    // stage s: slot s % 2, buffer s % 2; at stage it (after the barrier that ends
    // everyone's stage it - 1), stage it + 1 goes to LDS and its slot is reloaded
    // with stage it + 3 (loads past the slice return zero)
    for (int it = 0; it < ns; it += 2) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            if (it + u >= ns) break;
            __syncthreads();
            const int nx = (u + 1) % 2;
            put(nx, sx[nx], sy[nx]);
            load(it + u + 3, sx[nx], sy[nx]);
            compute(u);
        }
    }
    // the slice's partial: lane l, register v of tile (a, b) is row 8 (v / 4) + 4 (l / 32) + v % 4 and
    // column l % 32 of the tile
    float *pz = part + (int64_t)z * K * N;
#pragma unroll
    for (int a = 0; a < TI; ++a)
#pragma unroll
        for (int b = 0; b < TJ; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int i = (wi * TI + a) * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                const int j = (wj * TJ + b) * 32 + (lane & 31);
                pz[i * N + j] = acc[a][b][v];
            }
}

// the transpose-read weight gradient for the layers' shapes; false: not one of them
static bool nn_wgrad_tr(const uint16_t *X, const uint16_t *dY, int M, int K, int N, int R, float *part,
                        hipStream_t st)
{
    if ((reinterpret_cast<uintptr_t>(X) & 15) || (reinterpret_cast<uintptr_t>(dY) & 15) || R % 32 || R <= 0 ||
        (int64_t)R * 512 >= (1ll << 31))
        return false;
    const int ns = (M + R - 1) / R;
    // (K, N, WI, WJ): two LDS buffers of 32 rows of X and dY
#define NN_WGRAD_CASE(k, n, wi, wj)                                                                              \
    if (K == k && N == n) {                                                                                      \
        hipLaunchKernelGGL((k_nn_wgrad<k, n, wi, wj>), dim3(ns), dim3(64 * wi * wj), 2 * 32 * (k + n) * 2, st, X,   \
                           dY, M, R, part);                                                                      \
        return true;                                                                                             \
    }
    NN_WGRAD_CASE(256, 256, 4, 2)
    NN_WGRAD_CASE(64, 256, 1, 4)
    NN_WGRAD_CASE(256, 64, 2, 2)
    NN_WGRAD_CASE(256, 32, 4, 1)
#undef NN_WGRAD_CASE
    return false;
}

// out[i] = sum over slices s of part[s * n + i], in slice order
__global__ __launch_bounds__(256) void k_nn_sum_slices(const float *__restrict__ part, int nslice, int64_t n,
                                                       float *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int z = 0; z < nslice; ++z) s += part[(int64_t)z * n + i];
    out[i] = s;
}

// column sums of a BF16 matrix (M x N) in row slices: part[z * N + j]
__global__ __launch_bounds__(256) void k_nn_colsum(const uint16_t *__restrict__ Y, int M, int N, int rslice,
                                                   float *__restrict__ part)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= N) return;
    const int r0 = blockIdx.y * rslice, r1 = min(M, r0 + rslice);
    float s = 0.f;
    for (int i = r0; i < r1; ++i) s += bf2f(Y[(int64_t)i * N + j]);
    part[(int64_t)blockIdx.y * N + j] = s;
}

// out[i] = sum over slices z of part[z * n + i], n % 4 == 0, 16-byte aligned: a
// block takes 256 consecutive outputs (64 float4 columns) x 4 slice groups;
// group g sums the slices g, g + 4, g + 8, ... in that order (8 loads in
// flight), and the four group sums are added in group order -- a fixed order,
// the same on every run
__global__ __launch_bounds__(256) void k_nn_sum_slices4(const float *__restrict__ part, int nslice, int64_t n,
                                                        float *__restrict__ out)
{
    __shared__ float4 red[4][64];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int64_t i4 = (int64_t)blockIdx.x * 64 + c, n4 = n / 4;
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i4 < n4) {
        const float4 *p = reinterpret_cast<const float4 *>(part) + i4;
        int z = g;
        for (; z + 28 < nslice; z += 32) {
            float4 v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = p[(int64_t)(z + 4 * u) * n4];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                s.x += v[u].x;
                s.y += v[u].y;
                s.z += v[u].z;
                s.w += v[u].w;
            }
        }
        for (; z < nslice; z += 4) {
            const float4 v = p[(int64_t)z * n4];
            s.x += v.x;
            s.y += v.y;
            s.z += v.z;
            s.w += v.w;
        }
    }
    red[g][c] = s;
    __syncthreads();
    if (g == 0 && i4 < n4) {
        float4 o = red[0][c];
#pragma unroll
        for (int q = 1; q < 4; ++q) {
            o.x += red[q][c].x;
            o.y += red[q][c].y;
            o.z += red[q][c].z;
            o.w += red[q][c].w;
        }
        reinterpret_cast<float4 *>(out)[i4] = o;
    }
}

// column sums of a BF16 matrix (M x N, N even, 256 % (N / 2) == 0) in row
// slices: part[z * N + j].  Block z takes slice z: thread (column pair cp, row
// group rg) sums rows rg, rg + RG, ... of the slice (8 loads in flight), and the
// RG row-group sums are added in group order.
__global__ __launch_bounds__(256) void k_nn_colsum2(const uint16_t *__restrict__ Y, int M, int N, int rslice,
                                                    float *__restrict__ part)
{
    __shared__ float2 red[256];
    const int cpn = N / 2, RG = 256 / cpn;
    const int cp = threadIdx.x % cpn, rg = threadIdx.x / cpn;
    const int r0 = blockIdx.x * rslice, r1 = min(M, r0 + rslice);
    const uint32_t *y = reinterpret_cast<const uint32_t *>(Y) + cp;
    float2 s = make_float2(0.f, 0.f);
    int r = r0 + rg;
    for (; r + 7 * RG < r1; r += 8 * RG) {
        uint32_t v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = y[(int64_t)(r + u * RG) * cpn];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            s.x += __uint_as_float(v[u] << 16);
            s.y += __uint_as_float(v[u] & 0xffff0000u);
        }
    }
    for (; r < r1; r += RG) {
        const uint32_t v = y[(int64_t)r * cpn];
        s.x += __uint_as_float(v << 16);
        s.y += __uint_as_float(v & 0xffff0000u);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    if (rg == 0) {
        float2 o = red[cp];
        for (int q = 1; q < RG; ++q) {
            o.x += red[q * cpn + cp].x;
            o.y += red[q * cpn + cp].y;
        }
        part[(int64_t)blockIdx.x * N + 2 * cp] = o.x;
        part[(int64_t)blockIdx.x * N + 2 * cp + 1] = o.y;
    }
}

// sum over the slices of part into out (n values): the fixed-order float4 kernel when n allows it
static void nn_sum_slices(const float *part, int ns, int64_t n, float *out, hipStream_t st)
{
    if (n % 4 == 0 && !(reinterpret_cast<uintptr_t>(part) & 15) && !(reinterpret_cast<uintptr_t>(out) & 15))
        hipLaunchKernelGGL(k_nn_sum_slices4, dim3((unsigned)((n / 4 + 63) / 64)), dim3(256), 0, st, part, ns, n, out);
    else
        hipLaunchKernelGGL(k_nn_sum_slices, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, part, ns, n, out);
}

// masked mean-squared error of Y (M x ldy, first 3 columns, FP32) against T
// (M x 3): per-slice loss partials and the BF16 gradient 2 (y - t) w_i / W
__global__ __launch_bounds__(256) void k_nn_mse(const float *__restrict__ Y, int64_t ldy, const float *__restrict__ T,
                                                const float *__restrict__ w, int M, float inv_wsum,
                                                uint16_t *__restrict__ dY, float *__restrict__ loss_part)
{
    __shared__ float red[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    float l = 0.f;
    if (i < M) {
        const float wi = w ? w[i] : 1.f;
        for (int q = 0; q < 3; ++q) {
            const float d = Y[(int64_t)i * ldy + q] - T[(int64_t)i * 3 + q];
            l += wi * d * d;
            dY[(int64_t)i * ldy + q] = f2bf(2.f * wi * d * inv_wsum);
        }
        for (int q = 3; q < ldy; ++q) dY[(int64_t)i * ldy + q] = 0;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0] * inv_wsum;
}

// Adam on FP32 master weights; refreshes the BF16 copy used by the GEMMs
__global__ __launch_bounds__(256) void k_nn_adam(float *__restrict__ p, const float *__restrict__ g,
                                                 float *__restrict__ m, float *__restrict__ v,
                                                 uint16_t *__restrict__ pbf, int64_t n, float lr, float b1, float b2,
                                                 float eps, float c1, float c2, float wd)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * p[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    if (pbf) pbf[i] = f2bf(pi);
}

__global__ __launch_bounds__(256) void k_nn_to_bf16(const float *__restrict__ x, int64_t n, uint16_t *__restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// ---------------------------------------------------------------------------
// C ABI (declared in include/pdx_nn.h)
// ---------------------------------------------------------------------------
static int nn_check()
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : 3;
}

static int nn_grid_rows(int M) { return (M + NN_TM - 1) / NN_TM; }

extern "C" int pdx_nn_linear_fwd(const uint16_t *X, const uint16_t *W, const float *bias, int M, int K, int N,
                                 int relu, uint16_t *Ybf, float *Yf, int ldy, void *stream)
{
    if (!X || !W || M < 0 || K <= 0 || N <= 0 || ldy < N || (!Ybf && !Yf)) return 1;
    if (M == 0) return 0;
    if (!nn_gemm_lds<NN_EPI_FWD>(X, K, W, N, 1, M, K, N, bias, relu, nullptr, 0, Ybf, Yf, ldy, (hipStream_t)stream)) {
        dim3 g(nn_grid_rows(M), (N + NN_TN - 1) / NN_TN);
        hipLaunchKernelGGL(k_nn_gemm<NN_EPI_FWD>, g, dim3(256), 0, (hipStream_t)stream, X, (int64_t)K, (int64_t)1, W,
                           (int64_t)N, (int64_t)1, M, N, K, K, bias, relu, nullptr, (int64_t)0, Ybf, Yf, (int64_t)ldy);
    }
    return nn_check();
}

extern "C" int pdx_nn_linear_dgrad(const uint16_t *dY, const uint16_t *W, int M, int K, int N, int ldy,
                                   const uint16_t *act, uint16_t *dX, void *stream)
{
    if (!dY || !W || !dX || M < 0 || K <= 0 || N <= 0 || ldy < N) return 1;
    if (M == 0) return 0;
    // dX (M x K) = dY (M x N) W^T; B(k = n, j = k_in) = W[k_in * N + n]
    if (!nn_gemm_lds<NN_EPI_DGRAD>(dY, ldy, W, 1, N, M, N, K, nullptr, 0, act, K, dX, nullptr, K,
                                   (hipStream_t)stream)) {
        dim3 g(nn_grid_rows(M), (K + NN_TN - 1) / NN_TN);
        hipLaunchKernelGGL(k_nn_gemm<NN_EPI_DGRAD>, g, dim3(256), 0, (hipStream_t)stream, dY, (int64_t)ldy,
                           (int64_t)1, W, (int64_t)1, (int64_t)N, M, K, N, N, nullptr, 0, act, (int64_t)K, dX, nullptr,
                           (int64_t)K);
    }
    return nn_check();
}

extern "C" int pdx_nn_linear_wgrad(const uint16_t *X, const uint16_t *dY, int M, int K, int N, int ldy,
                                   int rows_per_slice, float *part, float *dW, float *db, void *stream)
{
    if (!X || !dY || !part || !dW || M <= 0 || K <= 0 || N <= 0 || ldy != N || rows_per_slice <= 0 ||
        rows_per_slice % 8)
        return 1;
    const int ns = (M + rows_per_slice - 1) / rows_per_slice;
    // dW (K x N) = X^T (K x M) dY (M x N): A(i = k_in, k = m) = X[m * K + k_in], B(k = m, j) = dY[m * ldy + j];
    // the transpose-read MFMA kernel for the layers' shapes, else the generic one
    if (!nn_wgrad_tr(X, dY, M, K, N, rows_per_slice, part, (hipStream_t)stream)) {
        dim3 g(nn_grid_rows(K), (N + NN_TN - 1) / NN_TN, ns);
        hipLaunchKernelGGL(k_nn_gemm<NN_EPI_WGRAD>, g, dim3(256), 0, (hipStream_t)stream, X, (int64_t)1, (int64_t)K,
                           dY, (int64_t)ldy, (int64_t)1, K, N, M, rows_per_slice, nullptr, 0, nullptr, (int64_t)0,
                           nullptr, part, (int64_t)N);
    }
    const int64_t nw = (int64_t)K * N;
    nn_sum_slices(part, ns, nw, dW, (hipStream_t)stream);
    if (db) {
        // bias gradient: column sums of dY in the same row slices
        if (N % 2 == 0 && N <= 512 && 256 % (N / 2) == 0 && !(reinterpret_cast<uintptr_t>(dY) & 3))
            hipLaunchKernelGGL(k_nn_colsum2, dim3(ns), dim3(256), 0, (hipStream_t)stream, dY, M, N, rows_per_slice,
                               part);
        else
            hipLaunchKernelGGL(k_nn_colsum, dim3((N + 255) / 256, ns), dim3(256), 0, (hipStream_t)stream, dY, M, ldy,
                               rows_per_slice, part);
        nn_sum_slices(part, ns, (int64_t)N, db, (hipStream_t)stream);
    }
    return nn_check();
}

extern "C" int pdx_nn_mse(const float *Y, int ldy, const float *T, const float *w, int M, float wsum, uint16_t *dY,
                          float *loss_part, void *stream)
{
    if (!Y || !T || !dY || !loss_part || M <= 0 || ldy < 3 || !(wsum > 0.f)) return 1;
    hipLaunchKernelGGL(k_nn_mse, dim3((M + 255) / 256), dim3(256), 0, (hipStream_t)stream, Y, (int64_t)ldy, T, w, M,
                       1.f / wsum, dY, loss_part);
    return nn_check();
}

extern "C" int pdx_nn_adam(float *p, const float *g, float *m, float *v, uint16_t *pbf, int64_t n, float lr,
                           float beta1, float beta2, float eps, int64_t step, float weight_decay, void *stream)
{
    if (!p || !g || !m || !v || n <= 0 || step < 1) return 1;
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_nn_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, p, g, m, v,
                       pbf, n, lr, beta1, beta2, eps, c1, c2, weight_decay);
    return nn_check();
}

extern "C" int pdx_nn_to_bf16(const float *x, int64_t n, uint16_t *y, void *stream)
{
    if (!x || !y || n < 0) return 1;
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_nn_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, x, n, y);
    return nn_check();
}
