// vh_nrc.hip - neural radiance cache: encoding, fused BF16-MFMA inference, fused forward/backward
// training, gradient reduction and Adam, and the NRC integrator of the C ABI.
//
// No reference code exists for this (docs/implementation/M7-nrc-adaptive-integrator/README.md
// plans it, rows 2-5 "not started"). The network is the workload's: 12-frequency position
// encoding + degree-3 real SH of the outgoing direction + normal + roughness (96 inputs), five
// hidden ReLU layers of 64, 16 linear outputs of which 3 are RGB; relative L2 loss; Adam with FP32
// master weights. Matrix products run on v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
//
// MFMA operand convention used throughout (16x16 output tile, 32-deep k step, wave64):
//   lane l holds row r = l % 16 of A and column r of B, and k-slice g = l / 16 (8 values);
//   both operands are read as 8 contiguous bf16 along k, so A is stored [row][k] and B is stored
//   [col][k]; the accumulator's element t of lane l is C[4 g + t][r].
#include <algorithm>
#include <cmath>
#include <memory>
#include <mutex>
#include <vector>

#include "vh_internal.h"

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

namespace vh {

namespace {

constexpr int kIn = 96, kHid = 64, kOut = 16, kNL = 6;
constexpr int kLdX = 104;   // padded leading dimension of 96-wide rows in LDS
constexpr int kLdH = 72;    // padded leading dimension of 64-wide rows in LDS

__host__ __device__ constexpr int layerK(int l) { return l == 0 ? kIn : kHid; }
__host__ __device__ constexpr int layerN(int l) { return l == kNL - 1 ? kOut : kHid; }
__host__ __device__ constexpr int layerNpad(int l) { return l == kNL - 1 ? 32 : kHid; } // k-depth of the backward product
__host__ __device__ constexpr int wOff(int l) { return l == 0 ? 0 : wOff(l - 1) + layerK(l - 1) * layerN(l - 1) + layerN(l - 1); }
__host__ __device__ constexpr int bOff(int l) { return wOff(l) + layerK(l) * layerN(l); }
__host__ __device__ constexpr int wtOff(int l) { return l == 0 ? 0 : wtOff(l - 1) + layerK(l - 1) * layerN(l - 1); }
__host__ __device__ constexpr int wbOff(int l) { return l == 0 ? 0 : wbOff(l - 1) + layerK(l - 1) * layerNpad(l - 1); }
constexpr int kParams = bOff(kNL - 1) + kOut; // 23888
constexpr int kWt = wtOff(kNL);               // 23552 bf16
constexpr int kWb = wbOff(kNL);               // 24576 bf16
constexpr int kBias = kParams - kWt;          // 336
static_assert(kParams == 23888, "parameter layout");

struct NrcNet {
    const float* theta;     // fp32 master parameters
    const __bf16* wt;       // per layer [n][k] (forward B operand)
    const __bf16* wb;       // per layer [k][npad] (backward B operand)
    float lo[3], inv[3];    // position normalisation to [0, 1]^3
};

__device__ __forceinline__ f32x4_t mma16(const __bf16* a, int lda, const __bf16* b, int ldb, int K, f32x4_t c)
{
    int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const __bf16* pa = a + r * lda + g * 8;
    const __bf16* pb = b + r * ldb + g * 8;
    for (int k0 = 0; k0 < K; k0 += 32)
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const bf16x8_t*>(pa + k0),
                                                     *reinterpret_cast<const bf16x8_t*>(pb + k0), c, 0, 0, 0);
    return c;
}

// 24 of the 96 encoded inputs of one query: part 0..2 the frequency features of one position axis,
// part 3 the 16 SH terms, the normal, the roughness and 4 zeros. Layout of oracle/nrc.py `encode`.
__device__ void encodePart(const NrcNet& net, float4 q0, float4 q1, float4 q2, int part, __bf16* dst)
{
    if (part < 3) {
        float p = part == 0 ? q0.x : (part == 1 ? q0.y : q0.z);
        p = fminf(fmaxf((p - net.lo[part]) * net.inv[part], 0.0f), 1.0f);
        // sin, cos (2^k pi p) = sin, cos (2 pi frac(2^(k-1) p)): the scaling by a power of two and
        // the fraction are exact, so the only rounding is the hardware sine's (v_sin_f32 and
        // v_cos_f32 take their argument in revolutions), far below the bf16 rounding that follows
        float x = 0.5f * p;
#pragma unroll
        for (int k = 0; k < 12; k++) {
            const float f = __builtin_amdgcn_fractf(x);
            dst[2 * k] = (__bf16)__builtin_amdgcn_sinf(f);
            dst[2 * k + 1] = (__bf16)__builtin_amdgcn_cosf(f);
            x *= 2.0f;
        }
    } else {
        float x = q1.x, y = q1.y, z = q1.z;
        float v[24] = {0.28209479177387814f,
                       -0.4886025119029199f * y, 0.4886025119029199f * z, -0.4886025119029199f * x,
                       1.0925484305920792f * x * y, -1.0925484305920792f * y * z, 0.31539156525252005f * (3.0f * z * z - 1.0f),
                       -1.0925484305920792f * x * z, 0.5462742152960396f * (x * x - y * y),
                       -0.5900435899266435f * y * (3.0f * x * x - y * y), 2.890611442640554f * x * y * z,
                       -0.4570457994644658f * y * (5.0f * z * z - 1.0f), 0.3731763325901154f * z * (5.0f * z * z - 3.0f),
                       -0.4570457994644658f * x * (5.0f * z * z - 1.0f), 1.445305721320277f * z * (x * x - y * y),
                       -0.5900435899266435f * x * (x * x - 3.0f * y * y),
                       q2.x, q2.y, q2.z, q0.w, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int k = 0; k < 24; k++)
            dst[k] = (__bf16)v[k];
    }
}

// Forward weights in the inference's LDS: layer l as [n][layerK(l) + 8] (rows padded by 16 bytes so
// the 16 rows an MFMA operand read touches fall in different banks).
__host__ __device__ constexpr int wpLd(int l) { return layerK(l) + 8; }
__host__ __device__ constexpr int wpOff(int l) { return l == 0 ? 0 : wpOff(l - 1) + layerN(l - 1) * wpLd(l - 1); }
constexpr int kWp = wpOff(kNL); // 26240 bf16

// MODE 0: out[3 q + c] = prediction of raw query q (3 float4 per record).
// MODE 1: render: query records of the terminating depth (5 float4: inputs, pixel, throughput,
//         total); film[pixel] += total + throughput * max(prediction, 0).
// Each wave runs 16 queries at a time through the network on its own: out^T = W^T act^T per layer,
// the weights as the MFMA A operand (LDS) and the activations, row-major per query, as B, so a
// lane's result is 4 consecutive features of one query, stored with one 8-byte write. One block of
// kInferWaves waves per CU shares one copy of the weights (52 KB of the CU's 160 KB LDS; the
// waves' activations take the rest), for 4 waves per SIMD.
constexpr int kInferWaves = 16;
template <int MODE>
__global__ __launch_bounds__(64 * kInferWaves) void k_nrc_infer(NrcNet net, const float4* rec, const uint32_t* countPtr,
                                                                uint32_t countImm, float* out, double* film)
{
    __shared__ __attribute__((aligned(16))) __bf16 s_w[kWp];
    __shared__ float s_b[kBias];
    __shared__ __attribute__((aligned(16))) __bf16 s_act[kInferWaves][2][16 * kLdX];
    for (int l = 0; l < kNL; l++) {
        const int K = layerK(l), N = layerN(l), chunks = K / 8;
        for (int i = threadIdx.x; i < N * chunks; i += 64 * kInferWaves) {
            const int nn = i / chunks, c = i % chunks;
            *reinterpret_cast<bf16x8_t*>(s_w + wpOff(l) + nn * wpLd(l) + 8 * c) =
                *reinterpret_cast<const bf16x8_t*>(net.wt + wtOff(l) + nn * K + 8 * c);
        }
    }
    for (int l = 0, o = 0; l < kNL; l++) {
        for (int i = threadIdx.x; i < layerN(l); i += 64 * kInferWaves)
            s_b[o + i] = net.theta[bOff(l) + i];
        o += layerN(l);
    }
    __syncthreads();
    const uint32_t n = countPtr ? *countPtr : countImm;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
    const int stride = MODE == 0 ? 3 : (int)kNrcQueryWords;
    // wave-uniform loop over this wave's 16-query tiles
    for (uint32_t tile = (blockIdx.x * kInferWaves + wave) * 16u; tile < n; tile += gridDim.x * kInferWaves * 16u) {
        const uint32_t q = tile + r; // this lane's query in the B operand and the result
        // the whole record at once (no load waits on another), and in mode 1 the pixel's film
        // while the network runs (each pixel has at most one query, so nothing else writes it)
        float4 q0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), q1 = q0, q2 = q0, q3 = q0, q4 = q0;
        if (q < n) {
            const float4* qr = rec + (size_t)q * stride;
            q0 = qr[0];
            q1 = qr[1];
            q2 = qr[2];
            if (MODE == 1) {
                q3 = qr[3];
                q4 = qr[4];
            }
        }
        const bool valid = q < n && (MODE == 0 || q4.w != 0.0f);
        double fp[3] = {0.0, 0.0, 0.0};
        if (MODE == 1 && g == 0 && valid) {
            const double* F = film + 3 * (size_t)asu(q1.w);
            fp[0] = F[0];
            fp[1] = F[1];
            fp[2] = F[2];
        }
        __bf16* a0 = s_act[wave][0] + r * kLdX + g * 24;
        if (valid) {
            encodePart(net, q0, q1, q2, g, a0);
        } else {
            for (int k = 0; k < 24; k++)
                a0[k] = (__bf16)0.0f;
        }
        waveSync(); // a wave's queries are its own from the encoding to the output: no block barrier
        int bo = 0;
#pragma unroll
        for (int l = 0; l < kNL; l++) {
            const int K = layerK(l), N = layerN(l);
            const __bf16* in = s_act[wave][l & 1];
            __bf16* nx = s_act[wave][(l + 1) & 1];
#pragma unroll
            for (int nt = 0; nt < N / 16; nt++) {
                f32x4_t c = {0.0f, 0.0f, 0.0f, 0.0f};
                c = mma16(s_w + wpOff(l) + nt * 16 * wpLd(l), wpLd(l), in, kLdX, K, c);
                const int f0 = nt * 16 + g * 4; // the lane's 4 features
                if (l < kNL - 1) {
                    bf16x4_t v;
#pragma unroll
                    for (int t = 0; t < 4; t++)
                        v[t] = (__bf16)fmaxf(c[t] + s_b[bo + f0 + t], 0.0f);
                    *reinterpret_cast<bf16x4_t*>(nx + r * kLdX + f0) = v;
                } else if (g == 0 && q < n) { // features 0-2: the RGB prediction
                    if (MODE == 0) {
#pragma unroll
                        for (int ch = 0; ch < 3; ch++)
                            out[(size_t)q * 3 + ch] = c[ch] + s_b[bo + ch];
                    } else if (valid) {
                        double* F = film + 3 * (size_t)asu(q1.w);
                        const float tv[3] = {q3.x, q3.y, q3.z}, tot[3] = {q4.x, q4.y, q4.z};
#pragma unroll
                        for (int ch = 0; ch < 3; ch++)
                            F[ch] = fp[ch] + (double)(tot[ch] + tv[ch] * fmaxf(c[ch] + s_b[bo + ch], 0.0f));
                    }
                }
            }
            bo += N;
            waveSync();
        }
    }
}

// The training kernel works on 32-row tiles (two 16-row MFMA tiles) with 8 waves per block: each
// wave computes one 16 x 16 tile of every product of a layer, so a layer's chain of dependent MFMAs
// is one tile long. Every wave keeps the weight operands of its tiles (forward and backward, 88
// VGPRs) in registers for the whole launch: no global load waits inside the tile loop. The
// activations, their transposed copies and the gradients of one tile take 71 KB of LDS; one block
// per CU (2 waves per SIMD, 256 registers each).
constexpr int kTrainRows = 32;
constexpr int kTrainWaves = 8;
constexpr int kLdS = 40; // padded leading dimension of transposed training tiles (32 samples) in LDS

// c += A B over K (A: 16 rows at a with leading dimension lda, read from LDS; B: this lane's operand
// fragments of KS 32-deep steps, held in registers)
template <int KS>
__device__ __forceinline__ f32x4_t mmaReg(const __bf16* a, int lda, const bf16x8_t (&b)[KS], f32x4_t c)
{
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const __bf16* pa = a + r * lda + g * 8;
#pragma unroll
    for (int k = 0; k < KS; k++)
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const bf16x8_t*>(pa + 32 * k), b[k], c, 0, 0, 0);
    return c;
}

// this lane's B fragments of the column tile of a matrix stored [col][k] at b (leading dimension ldb)
template <int KS>
__device__ __forceinline__ void loadFrag(const __bf16* b, int ldb, bf16x8_t (&f)[KS])
{
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int k = 0; k < KS; k++)
        f[k] = *reinterpret_cast<const bf16x8_t*>(b + r * ldb + g * 8 + 32 * k);
}

// Forward layer L of a tile: wave w computes row tile w & 1, column tile w >> 1 (waves past the
// layer's column tiles idle); ReLU outputs go row-major to hout and transposed to hT, the output
// layer's RGB to y.
template <int L, int KS>
__device__ __forceinline__ void fwdLayer(const __bf16* in, int ldin, const bf16x8_t (&w)[KS], float bias, __bf16* hout, __bf16* hT,
                                         float (*y)[4])
{
    constexpr int N = layerN(L);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
    const int rt = wave & 1, nt = wave >> 1;
    if (nt < N / 16) {
        f32x4_t c = {0.0f, 0.0f, 0.0f, 0.0f};
        c = mmaReg<KS>(in + rt * 16 * ldin, ldin, w, c);
        const int col = nt * 16 + r;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int rw = rt * 16 + g * 4 + t;
            if (L < kNL - 1) {
                __bf16 v = (__bf16)fmaxf(c[t] + bias, 0.0f);
                hout[rw * kLdH + col] = v;
                hT[col * kLdS + rw] = v;
            } else if (col < 3) {
                y[rw][col] = c[t] + bias;
            }
        }
    }
    __syncthreads();
}

// One layer of the backward pass for a tile: dW (and db) accumulate into this wave's tiles
// (t = wave + 8 j), then (L > 0) dZ of the layer below = (dZ W^T) * relu'(input), one tile per wave
// (row tile w & 1, column tile w >> 1) with W^T's fragments from registers. hinT: the layer's input
// transposed ([k][row]); dz and dzT: its output gradient, row-major and transposed; writes the
// input's gradient in both layouts. The forward pass stores every activation transposed as well,
// so the weight gradient hinT x dz reads MFMA operands straight from LDS: one 32-deep MFMA per
// 16 x 16 output tile.
template <int L, int NT, int KS>
__device__ __forceinline__ void backLayer(const __bf16* hinT, const __bf16* dz, const __bf16* dzT, __bf16* dzNext, __bf16* dzNextT,
                                          const bf16x8_t (&wT)[KS], f32x4_t (&acc)[NT], float& db)
{
    constexpr int K = layerK(L), N = layerN(L);
    constexpr int NTILE = N / 16, TILES = (K / 16) * NTILE;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int j = 0; j < NT; j++) {
        const int t = wave + kTrainWaves * j;
        if (t < TILES) {
            const int kt = t / NTILE, nt = t % NTILE;
            acc[j] = mma16(hinT + kt * 16 * kLdS, kLdS, dzT + nt * 16 * kLdS, kLdS, kTrainRows, acc[j]);
        }
    }
    if ((int)threadIdx.x < N) { // the bias gradient: one contiguous row of dzT
        const bf16x8_t* row = reinterpret_cast<const bf16x8_t*>(dzT + threadIdx.x * kLdS);
        float s = 0.0f;
#pragma unroll
        for (int q = 0; q < kTrainRows / 8; q++) {
            bf16x8_t v = row[q];
#pragma unroll
            for (int e = 0; e < 8; e++)
                s += (float)v[e];
        }
        db += s;
    }
    if constexpr (L > 0) {
        static_assert(K / 16 * 2 == kTrainWaves, "one input-gradient tile per wave");
        const int rt = wave & 1, kt = wave >> 1;
        f32x4_t c = {0.0f, 0.0f, 0.0f, 0.0f};
        c = mmaReg<KS>(dz + rt * 16 * kLdH, kLdH, wT, c);
        const int col = kt * 16 + r;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int row = rt * 16 + g * 4 + t;
            float h = (float)hinT[col * kLdS + row];
            __bf16 v = (__bf16)(h > 0.0f ? c[t] : 0.0f);
            dzNext[row * kLdH + col] = v;
            dzNextT[col * kLdS + row] = v;
        }
    }
    __syncthreads();
}

template <int L, int NT>
__device__ __forceinline__ void storeLayer(float* part, const f32x4_t (&acc)[NT], float db)
{
    constexpr int K = layerK(L), N = layerN(L);
    constexpr int NTILE = N / 16, TILES = (K / 16) * NTILE;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int j = 0; j < NT; j++) {
        const int t = wave + kTrainWaves * j;
        if (t < TILES) {
            const int kt = t / NTILE, nt = t % NTILE;
#pragma unroll
            for (int e = 0; e < 4; e++)
                part[wOff(L) + (kt * 16 + g * 4 + e) * N + nt * 16 + r] = acc[j][e];
        }
    }
    if ((int)threadIdx.x < N)
        part[bOff(L) + threadIdx.x] = db;
}

// The samples lo .. hi - 1 of slice `slice` of `slices` equal parts of n samples.
__host__ __device__ __forceinline__ void sliceRange(uint32_t n, uint32_t slice, uint32_t slices, uint32_t& lo, uint32_t& hi)
{
    lo = (uint32_t)((uint64_t)n * slice / slices);
    hi = (uint32_t)((uint64_t)n * (slice + 1) / slices);
}

// Fused forward + backward over 32-row tiles of training samples (4 float4 each: 3 of inputs, the
// RGB target), over slice `slice` of `slices` of them. Writes this block's gradient partial sums
// (kParams floats) and loss sum. Every product is split over the 8 waves by 16-row tile and
// column tile, with a block barrier between layers.
__global__ __launch_bounds__(64 * kTrainWaves, 1) void k_nrc_train(NrcNet net, const float4* samples, const uint32_t* countPtr,
                                                                   uint32_t countImm, uint32_t cap, uint32_t slice, uint32_t slices,
                                                                   float* partials, float* lossPartials)
{
    __shared__ __attribute__((aligned(16))) __bf16 s_x[kTrainRows * kLdX];     // encoded inputs, row-major
    __shared__ __attribute__((aligned(16))) __bf16 s_xT[kIn * kLdS];           // and transposed
    __shared__ __attribute__((aligned(16))) __bf16 s_hr[2][kTrainRows * kLdH]; // hidden activations, row-major (ping-pong)
    __shared__ __attribute__((aligned(16))) __bf16 s_hT[5][kHid * kLdS];       // every hidden activation, transposed
    __shared__ float s_y[kTrainRows][4];                                        // RGB outputs
    __shared__ float s_t[kTrainRows][4];                                        // RGB targets, weight
    __shared__ __attribute__((aligned(16))) __bf16 s_dz[2][kTrainRows * kLdH];
    __shared__ __attribute__((aligned(16))) __bf16 s_dzT[2][kHid * kLdS];
    __shared__ float s_loss[64 * kTrainWaves];
    uint32_t total = countPtr ? *countPtr : countImm;
    if (total > cap)
        total = cap;
    uint32_t first, n;
    sliceRange(total, slice, slices, first, n); // n: one past the slice's last sample
    const float invB = 1.0f / (float)(n > first ? n - first : 1);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
    const int nt = wave >> 1; // this wave's column tile of the forward products, and (kt) of the input gradients
    // this wave's weight fragments: forward B = W^T [n][k] (column tile nt of each layer), backward
    // B = W [k][npad] (column tile kt = nt of each input gradient); biases of column nt 16 + r
    bf16x8_t f0[3], f1[2], f2[2], f3[2], f4[2], f5[2] = {}, b1[2], b2[2], b3[2], b4[2], b5[1];
    loadFrag<3>(net.wt + wtOff(0) + nt * 16 * layerK(0), layerK(0), f0);
    loadFrag<2>(net.wt + wtOff(1) + nt * 16 * kHid, kHid, f1);
    loadFrag<2>(net.wt + wtOff(2) + nt * 16 * kHid, kHid, f2);
    loadFrag<2>(net.wt + wtOff(3) + nt * 16 * kHid, kHid, f3);
    loadFrag<2>(net.wt + wtOff(4) + nt * 16 * kHid, kHid, f4);
    if (nt == 0)
        loadFrag<2>(net.wt + wtOff(5), kHid, f5);
    loadFrag<2>(net.wb + wbOff(1) + nt * 16 * layerNpad(1), layerNpad(1), b1);
    loadFrag<2>(net.wb + wbOff(2) + nt * 16 * layerNpad(2), layerNpad(2), b2);
    loadFrag<2>(net.wb + wbOff(3) + nt * 16 * layerNpad(3), layerNpad(3), b3);
    loadFrag<2>(net.wb + wbOff(4) + nt * 16 * layerNpad(4), layerNpad(4), b4);
    loadFrag<1>(net.wb + wbOff(5) + nt * 16 * layerNpad(5), layerNpad(5), b5);
    float bias[kNL];
#pragma unroll
    for (int l = 0; l < kNL; l++)
        bias[l] = nt * 16 < layerN(l) ? net.theta[bOff(l) + nt * 16 + r] : 0.0f;
    f32x4_t a0[3], a1[2], a2[2], a3[2], a4[2], a5[1];
    float d0 = 0.0f, d1 = 0.0f, d2 = 0.0f, d3 = 0.0f, d4 = 0.0f, d5 = 0.0f;
    for (auto& v : a0) v = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
    for (auto& v : a1) v = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
    for (auto& v : a2) v = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
    for (auto& v : a3) v = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
    for (auto& v : a4) v = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
    a5[0] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
    float lossAcc = 0.0f;
    for (uint32_t tile = first + blockIdx.x * kTrainRows; tile < n; tile += gridDim.x * kTrainRows) {
        // encode (waves 0-1: one row and a quarter of its inputs per lane), row-major and transposed
        if (wave < 2) {
            const int row = wave * 16 + r;
            uint32_t q = tile + row;
            __bf16* xr = s_x + row * kLdX + g * 24;
            if (q < n) {
                const float4* sp = samples + (size_t)q * 4;
                encodePart(net, sp[0], sp[1], sp[2], g, xr);
                if (g == 0) {
                    float4 t = sp[3];
                    s_t[row][0] = t.x;
                    s_t[row][1] = t.y;
                    s_t[row][2] = t.z;
                    s_t[row][3] = 1.0f;
                }
            } else {
                for (int k = 0; k < 24; k++)
                    xr[k] = (__bf16)0.0f;
                if (g == 0)
                    s_t[row][3] = 0.0f;
            }
            for (int k = 0; k < 24; k++)
                s_xT[(g * 24 + k) * kLdS + row] = xr[k];
        }
        __syncthreads();
        fwdLayer<0, 3>(s_x, kLdX, f0, bias[0], s_hr[0], s_hT[0], s_y);
        fwdLayer<1, 2>(s_hr[0], kLdH, f1, bias[1], s_hr[1], s_hT[1], s_y);
        fwdLayer<2, 2>(s_hr[1], kLdH, f2, bias[2], s_hr[0], s_hT[2], s_y);
        fwdLayer<3, 2>(s_hr[0], kLdH, f3, bias[3], s_hr[1], s_hT[3], s_y);
        fwdLayer<4, 2>(s_hr[1], kLdH, f4, bias[4], s_hr[0], s_hT[4], s_y);
        fwdLayer<5, 2>(s_hr[0], kLdH, f5, bias[5], nullptr, nullptr, s_y);
        // relative L2 loss gradient, d(loss)/dy = 2 (y - t) / (y^2 + 0.01) / B for RGB
        if (threadIdx.x < kTrainRows) {
            int i = threadIdx.x;
            bool ok = s_t[i][3] != 0.0f;
            for (int c = 0; c < 32; c++) {
                float gv = 0.0f;
                if (ok && c < 3) {
                    float y = s_y[i][c], t = s_t[i][c];
                    float den = y * y + 0.01f;
                    gv = 2.0f * (y - t) / den * invB;
                    lossAcc += (y - t) * (y - t) / den;
                }
                s_dz[0][i * kLdH + c] = (__bf16)gv;
                if (c < kOut)
                    s_dzT[0][c * kLdS + i] = (__bf16)gv;
            }
        }
        __syncthreads();
        backLayer<5, 1, 1>(s_hT[4], s_dz[0], s_dzT[0], s_dz[1], s_dzT[1], b5, a5, d5);
        backLayer<4, 2, 2>(s_hT[3], s_dz[1], s_dzT[1], s_dz[0], s_dzT[0], b4, a4, d4);
        backLayer<3, 2, 2>(s_hT[2], s_dz[0], s_dzT[0], s_dz[1], s_dzT[1], b3, a3, d3);
        backLayer<2, 2, 2>(s_hT[1], s_dz[1], s_dzT[1], s_dz[0], s_dzT[0], b2, a2, d2);
        backLayer<1, 2, 2>(s_hT[0], s_dz[0], s_dzT[0], s_dz[1], s_dzT[1], b1, a1, d1);
        backLayer<0, 3, 1>(s_xT, s_dz[1], s_dzT[1], s_dz[0], s_dzT[0], b5, a0, d0); // (no input gradient: b5 unused)
    }
    float* part = partials + (size_t)blockIdx.x * kParams;
    storeLayer<0, 3>(part, a0, d0);
    storeLayer<1, 2>(part, a1, d1);
    storeLayer<2, 2>(part, a2, d2);
    storeLayer<3, 2>(part, a3, d3);
    storeLayer<4, 2>(part, a4, d4);
    storeLayer<5, 1>(part, a5, d5);
    s_loss[threadIdx.x] = lossAcc;
    __syncthreads();
    if (threadIdx.x < 64) { // fixed order: kTrainWaves per lane, then a butterfly (lane 0's is kept)
        float sl = 0.0f;
#pragma unroll
        for (int w = 0; w < kTrainWaves; w++)
            sl += s_loss[threadIdx.x + 64 * w];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            sl += __shfl_xor(sl, off, 64);
        if (threadIdx.x == 0)
            lossPartials[blockIdx.x] = sl;
    }
}

// The gradient is the sum of the per-block partials in a fixed order (deterministic for a given
// sample order): first groups of kReduceGroup consecutive blocks, in block order, into one sum per
// group (enough threads to stream the partials at memory bandwidth), then the group sums in
// group order.
constexpr uint32_t kReduceGroup = 32;

// Four parameters per thread (16-byte loads; kParams is a multiple of 4), each summed in block order.
static_assert(kParams % 4 == 0, "float4 rows of the partials");
__global__ void k_nrc_reduce_groups(const float* partials, uint32_t blocks, float* groupSums)
{
    const uint32_t p4 = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t g = blockIdx.y;
    if (p4 >= (uint32_t)kParams / 4)
        return;
    const uint32_t b0 = g * kReduceGroup;
    const float4* src = reinterpret_cast<const float4*>(partials);
    float4 part[kReduceGroup];
#pragma unroll
    for (uint32_t k = 0; k < kReduceGroup; k++) // every load in flight at once
        part[k] = b0 + k < blocks ? src[(size_t)(b0 + k) * (kParams / 4) + p4] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (uint32_t k = 0; k < kReduceGroup; k++)
        if (b0 + k < blocks) {
            s.x += part[k].x;
            s.y += part[k].y;
            s.z += part[k].z;
            s.w += part[k].w;
        }
    reinterpret_cast<float4*>(groupSums)[(size_t)g * (kParams / 4) + p4] = s;
}

// Parameter p -> its bf16 copies (forward [n][k] and, when wb is given, backward [k][npad]);
// biases have none.
__device__ __forceinline__ void writeCopies(uint32_t p, float v, __bf16* wt, __bf16* wb)
{
#pragma unroll
    for (int l = 0; l < kNL; l++) {
        const int K = layerK(l), N = layerN(l);
        if (p >= (uint32_t)wOff(l) && p < (uint32_t)bOff(l)) {
            int e = (int)p - wOff(l);
            int k = e / N, nn = e % N;
            wt[wtOff(l) + nn * K + k] = (__bf16)v;
            if (wb)
                wb[wbOff(l) + k * layerNpad(l) + nn] = (__bf16)v;
        }
    }
}

// The moving average the integrator infers with: avg = decay avg + (1 - decay) theta over Adam
// steps from avg = 0, read as avg / (1 - decay^step) (bias-corrected like Adam's moments);
// decay 0 makes it theta itself.
struct NrcAverage {
    float* avg;    // running sum
    float* eval;   // corrected average, fp32 (biases are read from here)
    __bf16* ewt;   // its forward bf16 copy
    float decay;
    float correction; // 1 - decay^step
};

#ifndef VH_NRC_TRAIN_BLOCKS
#define VH_NRC_TRAIN_BLOCKS 256
#endif
constexpr uint32_t kTrainBlocks = VH_NRC_TRAIN_BLOCKS; // one 8-wave block per CU
constexpr uint32_t kReduceGroups = (kTrainBlocks + kReduceGroup - 1) / kReduceGroup;

struct AdamStep {
    float lr, b1, b2, c1, c2, eps; // c1, c2: the bias corrections 1 - b1^step, 1 - b2^step
};

// The second stage of the gradient sum and (ADAM) the Adam step, one thread per parameter:
// grad[p] = the group sums added in group order (a fixed count, so all its loads are in flight at
// once; the order of a loop over them, hence its bits); the loss = the blocks' loss partials in a
// fixed order; then Adam (beta 0.9 / 0.99, bias-corrected) on the FP32 master parameter, its bf16
// copies and the moving average the integrator infers with.
template <bool ADAM>
__global__ void k_nrc_reduce_step(const float* groupSums, float* grad, const float* lossPartials, float* loss, const uint32_t* countPtr,
                                  uint32_t countImm, uint32_t cap, uint32_t slice, uint32_t slices, float* theta, float* m, float* v,
                                  AdamStep a, __bf16* wt, __bf16* wb, NrcAverage A)
{
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < (uint32_t)kParams) {
        float part[kReduceGroups];
#pragma unroll
        for (uint32_t g = 0; g < kReduceGroups; g++)
            part[g] = groupSums[(size_t)g * kParams + p];
        float gsum = 0.0f;
#pragma unroll
        for (uint32_t g = 0; g < kReduceGroups; g++)
            gsum += part[g];
        grad[p] = gsum;
        if (ADAM) {
            float mm = a.b1 * m[p] + (1.0f - a.b1) * gsum;
            float vv = a.b2 * v[p] + (1.0f - a.b2) * gsum * gsum;
            m[p] = mm;
            v[p] = vv;
            float th = theta[p] - a.lr * (mm / a.c1) / (sqrtf(vv / a.c2) + a.eps);
            theta[p] = th;
            writeCopies(p, th, wt, wb);
            float av = A.decay * A.avg[p] + (1.0f - A.decay) * th;
            A.avg[p] = av;
            float e = A.decay > 0.0f ? av / A.correction : th;
            A.eval[p] = e;
            writeCopies(p, e, A.ewt, nullptr);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < 64) {
        // the loss partials by one wave in a fixed order: lane sums b = lane, lane + 64, ...;
        // then a fixed butterfly over the lanes (lane 0's result is deterministic)
        float ls = 0.0f;
        for (uint32_t b = threadIdx.x; b < kTrainBlocks; b += 64)
            ls += lossPartials[b];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            ls += __shfl_xor(ls, off, 64);
        if (threadIdx.x == 0) {
            uint32_t total = countPtr ? *countPtr : countImm;
            if (total > cap)
                total = cap;
            uint32_t lo, hi;
            sliceRange(total, slice, slices, lo, hi);
            uint32_t n = hi - lo;
            loss[0] = ls / (float)(n > 0 ? n : 1);
            loss[1] = (float)n;
        }
    }
}

// theta -> its bf16 copies; the average restarts at theta (its running sum at zero).
__global__ void k_nrc_copies(const float* theta, __bf16* wt, __bf16* wb, NrcAverage A)
{
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < (uint32_t)kParams) {
        float th = theta[p];
        writeCopies(p, th, wt, wb);
        A.avg[p] = 0.0f;
        A.eval[p] = th;
        writeCopies(p, th, A.ewt, nullptr);
    }
}

__global__ void k_zero_bf16(__bf16* p, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = (__bf16)0.0f;
}

// Sample records (4 float4: position + roughness, direction, normal, target + weight) of
// vh_nrc_gradient_noalloc, staged in a file-scope device array of kNoAllocSamples records (4 MiB)
// that every device running the library holds until the process ends. Host code must not pass the
// array's name to a kernel (on the host it is a stand-in, not a device address): it uses the
// address hipGetSymbolAddress returns for the current device, kept per device by scratch_address.
constexpr uint32_t kNoAllocSamples = 1u << 16;
static __device__ __align__(256) float4 scratch[4 * kNoAllocSamples];

constexpr int kMaxDevices = 64;
std::mutex g_scratchLock;
float4* g_scratchAddr[kMaxDevices] = {};

// The record layout of upload_batch, from the caller's device arrays (12 and 3 floats per sample).
__global__ void k_nrc_pack(const float* __restrict__ raw12, const float* __restrict__ target3, uint32_t n, float4* __restrict__ rec)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float* r = raw12 + 12 * (size_t)i;
    const float* t = target3 + 3 * (size_t)i;
    float4* o = rec + 4 * (size_t)i;
    o[0] = make_float4(r[0], r[1], r[2], r[3]);
    o[1] = make_float4(r[4], r[5], r[6], r[7]);
    o[2] = make_float4(r[8], r[9], r[10], r[11]);
    o[3] = make_float4(t[0], t[1], t[2], 1.0f);
}

} // namespace

} // namespace vh

using namespace vh;

struct vh_nrc {
    vh_renderer* r = nullptr;
    vh_nrc_config cfg{};
    DevBuf theta, m, v, grad, wt, wb, partials, groupSums, lossPart, loss;
    DevBuf avg, evalTheta, ewt; // moving average of theta and its forward copy (inference)
    DevBuf queries;
    // Training paths run beside the render, two training frames ahead, on two side streams with a
    // wavefront each (65,536 paths are too few to fill the GPU: a frame of them is bound by
    // traversal latency, so two run at once). samples[tf % 3] receives training frame tf;
    // pathsDone[b] / adamDone[b] order its writes and reads between the streams.
    struct TrainSet {
        DevBuf trainPix, vertices, state0, state1, hits, counts;
        // deferred shadow rays of the training paths (kMaxShadow records per path)
        DevBuf shRec, shVis, shQueue, shNRec, shOut, shPixel, shTotal;
        hipStream_t stream = nullptr;
    };
    static constexpr uint32_t kSets = 2, kBufs = 3;
    TrainSet sets[kSets];
    DevBuf samples[kBufs], sampleCount[kBufs];
    hipEvent_t pathsDone[kBufs] = {}, adamDone[kBufs] = {};
    bool adamIssued[kBufs] = {};
    // The Adam steps run on a stream of their own, beside the render's path tracing: only the
    // render's cache inference reads the weights, so it alone waits for the frame's steps, and the
    // next frame's steps wait for that inference (inferDone) before they overwrite the weights.
    hipStream_t adamStream = nullptr;
    hipEvent_t inferDone = nullptr;
    uint32_t step = 0;
    uint32_t frames = 0;
    uint32_t trainFrames = 0; // frames of training paths drawn so far (warm-up included)
    double trainSeconds = 0.0, renderSeconds = 0.0;
    float lastLoss = 0.0f;
    uint32_t lastSamples = 0;
    NrcAverage average() const
    {
        float corr = 1.0f - powf(cfg.ema_decay, (float)step);
        return NrcAverage{avg.as<float>(), evalTheta.as<float>(), ewt.as<__bf16>(), cfg.ema_decay, corr > 0.0f ? corr : 1.0f};
    }
    // the training network (master parameters) or the inference network (their moving average)
    NrcNet net(bool unitBox, bool inference = false) const
    {
        NrcNet n;
        n.theta = inference ? evalTheta.as<float>() : theta.as<float>();
        n.wt = inference ? ewt.as<__bf16>() : wt.as<__bf16>();
        n.wb = inference ? nullptr : wb.as<__bf16>();
        for (int k = 0; k < 3; k++) {
            float lo = unitBox ? 0.0f : r->bvh.lo[k], hi = unitBox ? 1.0f : r->bvh.hi[k];
            n.lo[k] = lo;
            n.inv[k] = 1.0f / fmaxf(hi - lo, 1.0e-20f);
        }
        return n;
    }
    ~vh_nrc()
    {
        for (auto& t : sets)
            if (t.stream)
                (void)hipStreamSynchronize(t.stream);
        if (adamStream)
            (void)hipStreamSynchronize(adamStream);
        for (uint32_t b = 0; b < kBufs; b++) {
            if (pathsDone[b])
                (void)hipEventDestroy(pathsDone[b]);
            if (adamDone[b])
                (void)hipEventDestroy(adamDone[b]);
        }
        if (inferDone)
            (void)hipEventDestroy(inferDone);
        for (auto& t : sets)
            if (t.stream)
                (void)hipStreamDestroy(t.stream);
        if (adamStream)
            (void)hipStreamDestroy(adamStream);
    }
};

namespace {

#define TRY(x)            \
    do {                  \
        int rc_ = (x);    \
        if (rc_ != VH_OK) \
            return rc_;   \
    } while (0)
#define HIPTRY(x) TRY(check_hip((x), #x))

// `steps` Adam steps (or one gradient when !adam) on the samples; with `slices` > 1 step k takes
// slice k % slices of them (minibatches), else every step takes all of them.
int train_steps(vh_nrc* nn, const float4* samples, const uint32_t* countPtr, uint32_t countImm, uint32_t cap, uint32_t steps,
                bool adam, hipStream_t s, uint32_t slices = 1)
{
    NrcNet net = nn->net(countPtr == nullptr);
    for (uint32_t k = 0; k < steps; k++) {
        const uint32_t slice = slices > 1 ? k % slices : 0u, nsl = slices > 1 ? slices : 1u;
        k_nrc_train<<<kTrainBlocks, 64 * kTrainWaves, 0, s>>>(net, samples, countPtr, countImm, cap, slice, nsl, nn->partials.as<float>(),
                                                 nn->lossPart.as<float>());
        k_nrc_reduce_groups<<<dim3((kParams / 4 + 255) / 256, kReduceGroups), 256, 0, s>>>(nn->partials.as<float>(), kTrainBlocks,
                                                                                       nn->groupSums.as<float>());
        AdamStep st{nn->cfg.learning_rate, 0.9f, 0.99f, 1.0f, 1.0f, 1.0e-8f};
        if (adam) {
            nn->step++;
            st.c1 = 1.0f - powf(0.9f, (float)nn->step);
            st.c2 = 1.0f - powf(0.99f, (float)nn->step);
            k_nrc_reduce_step<true><<<(kParams + 255) / 256, 256, 0, s>>>(
                nn->groupSums.as<float>(), nn->grad.as<float>(), nn->lossPart.as<float>(), nn->loss.as<float>(), countPtr, countImm, cap,
                slice, nsl, nn->theta.as<float>(), nn->m.as<float>(), nn->v.as<float>(), st, nn->wt.as<__bf16>(), nn->wb.as<__bf16>(),
                nn->average());
        } else {
            k_nrc_reduce_step<false><<<(kParams + 255) / 256, 256, 0, s>>>(
                nn->groupSums.as<float>(), nn->grad.as<float>(), nn->lossPart.as<float>(), nn->loss.as<float>(), countPtr, countImm, cap,
                slice, nsl, nullptr, nullptr, nullptr, st, nullptr, nullptr, NrcAverage{});
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return fail(VH_ERR_KERNEL, std::string("NRC training: ") + hipGetErrorString(e));
    }
    return VH_OK;
}

// He-uniform initialisation (He et al., ICCV 2015: weights uniform in +-sqrt(6 / fan_in)) with
// zero biases; the output layer is scaled by 0.1 and its 13 unused outputs start at zero. Weight
// p (its index in the flat parameter vector) takes u = (pcg(pcg(seed ^ 0xA511E9B3) + p) >> 8) 2^-24,
// the stream of Rng::next (pt_core.slang:79-83) keyed by the parameter, and w = lim (2u - 1).
void init_params(uint32_t seed, float* theta)
{
    const uint32_t key = pcg(seed ^ 0xA511E9B3u);
    for (int p = 0; p < kParams; p++)
        theta[p] = 0.0f;
    for (int l = 0; l < kNL; l++) {
        const int K = layerK(l), N = layerN(l);
        const bool last = l == kNL - 1;
        const float lim = sqrtf(6.0f / (float)K) * (last ? 0.1f : 1.0f);
        for (int e = 0; e < K * N; e++) {
            const uint32_t p = (uint32_t)(wOff(l) + e);
            if (last && e % N >= 3)
                continue;
            const float u = (float)(pcg(key + p) >> 8) * (1.0f / 16777216.0f);
            theta[p] = lim * (2.0f * u - 1.0f);
        }
    }
}

int upload_batch(vh_nrc* nn, uint64_t n, const float* raw12, const float* target3, DevBuf& buf, hipStream_t s)
{
    // sample records: (p, rough) (wo, 0) (n, 0) (target, 1)
    std::vector<float> host(16 * n, 0.0f);
    for (uint64_t i = 0; i < n; i++) {
        for (int k = 0; k < 12; k++)
            host[16 * i + k] = raw12[12 * i + k];
        for (int c = 0; c < 3; c++)
            host[16 * i + 12 + c] = target3 ? target3[3 * i + c] : 0.0f;
        host[16 * i + 15] = 1.0f;
    }
    HIPTRY(buf.alloc(sizeof(float) * host.size()));
    HIPTRY(hipMemcpyAsync(buf.p, host.data(), sizeof(float) * host.size(), hipMemcpyHostToDevice, s));
    (void)nn;
    return VH_OK;
}

// The device address of `scratch` on `device`, which must be current: looked up once per device
// (hipGetSymbolAddress, with its status checked) and kept for the life of the process.
int scratch_address(int device, float4** out)
{
    std::lock_guard<std::mutex> lock(g_scratchLock);
    if (device < 0 || device >= kMaxDevices)
        return fail(VH_ERR_UNSUPPORTED, "device index beyond the no-allocation staging table");
    if (!g_scratchAddr[device]) {
        int current = -1;
        HIPTRY(hipGetDevice(&current));
        if (current != device)
            return fail(VH_ERR_INVALID_ARGUMENT, "the cache's device is not the current device");
        void* p = nullptr;
        HIPTRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p)
            return fail(VH_ERR_KERNEL, "no device address for the no-allocation staging array");
        g_scratchAddr[device] = static_cast<float4*>(p);
    }
    *out = g_scratchAddr[device];
    return VH_OK;
}

} // namespace

extern "C" {

int vh_nrc_create(vh_renderer* r, const vh_nrc_config* cfg, vh_nrc** out)
{
    if (!r || !out)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    std::unique_ptr<vh_nrc> nn(new vh_nrc());
    nn->r = r;
    nn->cfg = cfg ? *cfg : vh_nrc_config{65536, 8, 1.0e-2f, 1, 0.95f, 16, 0};
    if (nn->cfg.train_paths == 0 || nn->cfg.train_steps == 0)
        return fail(VH_ERR_INVALID_ARGUMENT, "train_paths and train_steps must be positive");
    if (nn->cfg.train_paths > (1u << 24))
        return fail(VH_ERR_INVALID_ARGUMENT, "train_paths is at most 2^24");
    if (nn->cfg.terminate_depth < 1 || nn->cfg.terminate_depth >= kMaxDepthNrc)
        return fail(VH_ERR_INVALID_ARGUMENT, "terminate_depth must be in [1, 7]");
    if (!(nn->cfg.ema_decay >= 0.0f && nn->cfg.ema_decay < 1.0f))
        return fail(VH_ERR_INVALID_ARGUMENT, "ema_decay must be in [0, 1)");
    if (!(nn->cfg.learning_rate > 0.0f))
        return fail(VH_ERR_INVALID_ARGUMENT, "learning_rate must be positive");
    hipStream_t s = r->ctx->stream;
    uint64_t tp = nn->cfg.train_paths;
    uint64_t bytes = 4ull * kParams * 6 + 2ull * (2 * kWt + kWb) + 4ull * kParams * kTrainBlocks + tp * (2 * (4 + 16 * 5 * kMaxDepthNrc + 2 * 4 * 17 + 16) + 3 * 64 * kMaxDepthNrc);
    TRY(check_fits(bytes));
    HIPTRY(nn->theta.alloc(4ull * kParams));
    HIPTRY(nn->m.alloc(4ull * kParams));
    HIPTRY(nn->v.alloc(4ull * kParams));
    HIPTRY(nn->grad.alloc(4ull * kParams));
    HIPTRY(nn->avg.alloc(4ull * kParams));
    HIPTRY(nn->evalTheta.alloc(4ull * kParams));
    HIPTRY(nn->wt.alloc(2ull * kWt));
    HIPTRY(nn->ewt.alloc(2ull * kWt));
    HIPTRY(nn->wb.alloc(2ull * kWb));
    HIPTRY(nn->partials.alloc(4ull * kParams * kTrainBlocks));
    HIPTRY(nn->groupSums.alloc(4ull * kParams * ((kTrainBlocks + kReduceGroup - 1) / kReduceGroup)));
    HIPTRY(nn->lossPart.alloc(4ull * kTrainBlocks));
    HIPTRY(nn->loss.alloc(16));
    std::vector<float> theta0(kParams);
    init_params(0u, theta0.data()); // seed 0; vh_nrc_init draws another
    HIPTRY(hipMemcpyAsync(nn->theta.p, theta0.data(), 4ull * kParams, hipMemcpyHostToDevice, s));
    HIPTRY(hipMemsetAsync(nn->m.p, 0, 4ull * kParams, s));
    HIPTRY(hipMemsetAsync(nn->v.p, 0, 4ull * kParams, s));
    k_zero_bf16<<<(kWb + 255) / 256, 256, 0, s>>>(nn->wb.as<__bf16>(), kWb); // padding columns stay zero
    k_nrc_copies<<<(kParams + 255) / 256, 256, 0, s>>>(nn->theta.as<float>(), nn->wt.as<__bf16>(), nn->wb.as<__bf16>(), nn->average());
    for (auto& t : nn->sets) {
        HIPTRY(t.trainPix.alloc(4 * tp));
        HIPTRY(t.vertices.alloc(16ull * kNrcVertexWords * kMaxDepthNrc * tp));
        HIPTRY(t.state0.alloc(4ull * 17 * tp));
        HIPTRY(t.state1.alloc(4ull * 17 * tp));
        HIPTRY(t.hits.alloc(16 * tp));
        HIPTRY(t.counts.alloc(4 * kCountWords));
        HIPTRY(t.shRec.alloc(4ull * kShadowWords * kMaxShadow * tp));
        HIPTRY(t.shVis.alloc(4ull * kMaxShadow * tp));
        HIPTRY(t.shQueue.alloc(4ull * kMaxShadow * tp));
        HIPTRY(t.shNRec.alloc(4 * tp));
        HIPTRY(t.shOut.alloc(4 * tp));
        HIPTRY(t.shPixel.alloc(4 * tp));
        HIPTRY(t.shTotal.alloc(12 * tp));
        HIPTRY(hipStreamCreateWithFlags(&t.stream, hipStreamNonBlocking));
    }
    for (uint32_t b = 0; b < vh_nrc::kBufs; b++) {
        HIPTRY(nn->samples[b].alloc(64ull * kMaxDepthNrc * tp));
        HIPTRY(nn->sampleCount[b].alloc(16));
        HIPTRY(hipEventCreateWithFlags(&nn->pathsDone[b], hipEventDisableTiming));
        HIPTRY(hipEventCreateWithFlags(&nn->adamDone[b], hipEventDisableTiming));
    }
    HIPTRY(hipStreamCreateWithFlags(&nn->adamStream, hipStreamNonBlocking));
    HIPTRY(hipEventCreateWithFlags(&nn->inferDone, hipEventDisableTiming));
    HIPTRY(hipStreamSynchronize(s));
    float4* staging = nullptr; // the no-allocation gradient's staging array on this device (set-up)
    TRY(scratch_address(r->ctx->device, &staging));
    *out = nn.release();
    return VH_OK;
}

void vh_nrc_destroy(vh_nrc* nn)
{
    if (!nn)
        return;
    (void)hipStreamSynchronize(nn->r->ctx->stream);
    delete nn;
}

int vh_nrc_set_params(vh_nrc* nn, const float* theta, uint64_t n)
{
    if (!nn || !theta)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    if (n != (uint64_t)kParams)
        return fail(VH_ERR_INVALID_ARGUMENT, "expected " + std::to_string(kParams) + " parameters");
    hipStream_t s = nn->r->ctx->stream;
    HIPTRY(hipMemcpyAsync(nn->theta.p, theta, 4ull * kParams, hipMemcpyHostToDevice, s));
    HIPTRY(hipMemsetAsync(nn->m.p, 0, 4ull * kParams, s));
    HIPTRY(hipMemsetAsync(nn->v.p, 0, 4ull * kParams, s));
    nn->step = 0;
    k_nrc_copies<<<(kParams + 255) / 256, 256, 0, s>>>(nn->theta.as<float>(), nn->wt.as<__bf16>(), nn->wb.as<__bf16>(), nn->average());
    HIPTRY(hipGetLastError());
    HIPTRY(hipStreamSynchronize(s));
    return VH_OK;
}

int vh_nrc_init_params(uint32_t seed, float* params, uint64_t n)
{
    if (!params || n != (uint64_t)kParams)
        return fail(VH_ERR_INVALID_ARGUMENT, "expected " + std::to_string(kParams) + " parameters");
    init_params(seed, params);
    return VH_OK;
}

int vh_nrc_init(vh_nrc* nn, uint32_t seed)
{
    if (!nn)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    std::vector<float> theta(kParams);
    init_params(seed, theta.data());
    return vh_nrc_set_params(nn, theta.data(), theta.size()); // synchronises before theta goes
}

int vh_nrc_get_params(vh_nrc* nn, float* theta, uint64_t n)
{
    if (!nn || !theta || n != (uint64_t)kParams)
        return fail(VH_ERR_INVALID_ARGUMENT, "bad argument");
    HIPTRY(hipStreamSynchronize(nn->r->ctx->stream));
    HIPTRY(hipMemcpy(theta, nn->theta.p, 4ull * kParams, hipMemcpyDeviceToHost));
    return VH_OK;
}

int vh_nrc_infer(vh_nrc* nn, uint64_t n, const float* raw12, float* out3)
{
    if (!nn || (n && (!raw12 || !out3)))
        return fail(VH_ERR_INVALID_ARGUMENT, "bad argument");
    if (n == 0)
        return VH_OK;
    hipStream_t s = nn->r->ctx->stream;
    std::vector<float> host(12 * n);
    std::copy(raw12, raw12 + 12 * n, host.begin());
    DevBuf in, out;
    HIPTRY(in.alloc(4 * host.size()));
    HIPTRY(out.alloc(12 * n));
    HIPTRY(hipMemcpyAsync(in.p, host.data(), 4 * host.size(), hipMemcpyHostToDevice, s));
    uint32_t grid = (uint32_t)std::min<uint64_t>((n + 16 * kInferWaves - 1) / (16 * kInferWaves), nn->r->ctx->cus);
    k_nrc_infer<0><<<grid, 64 * kInferWaves, 0, s>>>(nn->net(true, true), in.as<float4>(), nullptr, (uint32_t)n, out.as<float>(), nullptr);
    HIPTRY(hipGetLastError());
    HIPTRY(hipStreamSynchronize(s));
    HIPTRY(hipMemcpy(out3, out.p, 12 * n, hipMemcpyDeviceToHost));
    return VH_OK;
}

int vh_nrc_gradient(vh_nrc* nn, uint64_t n, const float* raw12, const float* target3, float* grad, float* loss)
{
    if (!nn || !raw12 || !target3 || !grad || n == 0 || n > (1ull << 24))
        return fail(VH_ERR_INVALID_ARGUMENT, "bad argument");
    hipStream_t s = nn->r->ctx->stream;
    DevBuf batch;
    TRY(upload_batch(nn, n, raw12, target3, batch, s));
    TRY(train_steps(nn, batch.as<float4>(), nullptr, (uint32_t)n, (uint32_t)n, 1, false, s));
    HIPTRY(hipStreamSynchronize(s));
    HIPTRY(hipMemcpy(grad, nn->grad.p, 4ull * kParams, hipMemcpyDeviceToHost));
    if (loss)
        HIPTRY(hipMemcpy(loss, nn->loss.p, 4, hipMemcpyDeviceToHost));
    return VH_OK;
}

int vh_nrc_gradient_noalloc(vh_nrc* nn, uint64_t n, const float* raw12, const float* target3, float* grad, float* loss)
{
    if (!nn || !raw12 || !target3 || !grad || n == 0)
        return fail(VH_ERR_INVALID_ARGUMENT, "bad argument");
    if (n > kNoAllocSamples)
        return fail(VH_ERR_INVALID_ARGUMENT, "vh_nrc_gradient_noalloc takes at most 65536 samples");
    float4* rec = nullptr;
    TRY(scratch_address(nn->r->ctx->device, &rec));
    hipStream_t s = nn->r->ctx->stream;
    k_nrc_pack<<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(raw12, target3, (uint32_t)n, rec);
    HIPTRY(hipGetLastError());
    TRY(train_steps(nn, rec, nullptr, (uint32_t)n, (uint32_t)n, 1, false, s));
    HIPTRY(hipMemcpyAsync(grad, nn->grad.p, 4ull * kParams, hipMemcpyDeviceToDevice, s));
    if (loss)
        HIPTRY(hipMemcpyAsync(loss, nn->loss.p, 4, hipMemcpyDeviceToDevice, s));
    HIPTRY(hipStreamSynchronize(s));
    return VH_OK;
}

int vh_nrc_train(vh_nrc* nn, uint64_t n, const float* raw12, const float* target3, uint32_t steps, float* loss)
{
    if (!nn || !raw12 || !target3 || n == 0 || n > (1ull << 24))
        return fail(VH_ERR_INVALID_ARGUMENT, "bad argument");
    hipStream_t s = nn->r->ctx->stream;
    DevBuf batch;
    TRY(upload_batch(nn, n, raw12, target3, batch, s));
    TRY(train_steps(nn, batch.as<float4>(), nullptr, (uint32_t)n, (uint32_t)n, steps, true, s));
    HIPTRY(hipStreamSynchronize(s));
    if (loss)
        HIPTRY(hipMemcpy(loss, nn->loss.p, 4, hipMemcpyDeviceToHost));
    return VH_OK;
}

int vh_nrc_accumulate(vh_nrc* nn, vh_accumulation* a, uint32_t frames)
{
    if (!nn || !a || a->r != nn->r)
        return fail(VH_ERR_INVALID_ARGUMENT, "accumulation does not belong to the cache's renderer");
    vh_renderer* r = nn->r;
    if (a->lighting || r->adaptiveOn || r->restirOn || a->adaptiveOn || a->restirOn)
        return fail(VH_ERR_UNSUPPORTED, "the radiance cache runs without the lighting split, adaptive sampling and ReSTIR");
    if (frames == 0)
        return VH_OK;
    hipStream_t s = r->ctx->stream;
    TRY(ensure_wavefront(a));
    uint64_t npix = a->pixels();
    if (!nn->queries.p || nn->queries.bytes < 16ull * kNrcQueryWords * npix)
        HIPTRY(nn->queries.alloc(16ull * kNrcQueryWords * npix));
    Scene S = r->scene();
    Frame F{a->film.as<double>(), nullptr, nullptr, nullptr};
    Wavefront W = a->wavefront(false);
    LaunchCfg cfg{s, r->wavefrontBlocks()};
    Params base = base_params(r, a->width, a->height, false, false, vh_adaptive{}, false, vh_restir{});
    const uint32_t termDepth = nn->cfg.terminate_depth;
    const uint32_t maxDepth = base.w[7];
    const uint32_t sampleCap = nn->cfg.train_paths * kMaxDepthNrc;
    // Training frame tf: plain paths on a random pixel subset (their own sample stream) on side
    // stream tf % 2 into samples[tf % 3], after the Adam steps that last read that buffer; then
    // Adam steps on the render stream once those paths are done.
    auto issuePaths = [&](uint32_t tf) -> int {
        vh_nrc::TrainSet& ts = nn->sets[tf % vh_nrc::kSets];
        const uint32_t b = tf % vh_nrc::kBufs;
        if (nn->adamIssued[b])
            HIPTRY(hipStreamWaitEvent(ts.stream, nn->adamDone[b], 0));
        Wavefront T;
        T.state[0] = ts.state0.as<uint32_t>();
        T.state[1] = ts.state1.as<uint32_t>();
        T.hits = ts.hits.as<uint4>();
        T.queue0 = nullptr;
        T.counts = ts.counts.as<uint32_t>();
        T.n = nn->cfg.train_paths;
        T.words = 17;
        T.traceWaves = nn->r->ctx->cus * kTraceWavesPerCu; // short queues spread over the resident waves
        if (r->info.light_count + 1 <= kMaxShadow) { // shadow rays deferred to the cooperative shadow kernel
            T.shK = kMaxShadow;
            T.shRec = ts.shRec.as<float>();
            T.shVis = ts.shVis.as<uint32_t>();
            T.shQueue = ts.shQueue.as<uint32_t>();
            T.shNRec = ts.shNRec.as<uint32_t>();
            T.shOut = ts.shOut.as<uint32_t>();
            T.shPixel = ts.shPixel.as<uint32_t>();
            T.shTotal = ts.shTotal.as<float>();
        }
        Params PT = base;
        PT.w[6] = 0x40000000u + tf;
        TRY(launch_nrc_train_paths(S, PT, T, ts.trainPix.as<uint32_t>(), nn->cfg.train_paths, (uint32_t)npix, tf,
                                   ts.vertices.as<float4>(), nn->samples[b].as<float4>(), nn->sampleCount[b].as<uint32_t>(), sampleCap,
                                   LaunchCfg{ts.stream, r->wavefrontBlocks()}));
        HIPTRY(hipEventRecord(nn->pathsDone[b], ts.stream));
        return VH_OK;
    };
    bool inferIssued = false; // an inference of this call that the next Adam steps must wait for
    auto adam = [&](uint32_t tf) -> int {
        const uint32_t b = tf % vh_nrc::kBufs;
        hipStream_t as = nn->adamStream;
        HIPTRY(hipStreamWaitEvent(as, nn->pathsDone[b], 0));
        if (inferIssued)
            HIPTRY(hipStreamWaitEvent(as, nn->inferDone, 0));
        TRY(train_steps(nn, nn->samples[b].as<float4>(), nn->sampleCount[b].as<uint32_t>(), 0, sampleCap, nn->cfg.train_steps, true, as,
                        nn->cfg.minibatch ? nn->cfg.train_steps : 1u));
        HIPTRY(hipEventRecord(nn->adamDone[b], as));
        nn->adamIssued[b] = true;
        return VH_OK;
    };
    // one sample of every pixel, ending in the cache at its first opaque vertex at termDepth: the
    // paths up to the cache queries (no weights read), then, once the frame's Adam steps are done
    // (`waitAdam`, which records the timing events around the wait), the inference and the rest
    const bool merge = termDepth >= 1 && termDepth < maxDepth && W.shK > 0 && !S.translucent;
    auto renderPaths = [&](uint32_t f) -> int {
        Params P = base;
        P.w[6] = a->samples + f;
        P.w[31] = (uint32_t)npix;
        TRY(launch_wf_reset(P, W, s));
        // opaque scenes: the cache depth's closest rays go with the previous depth's shadow rays
        for (uint32_t d = 0; d < std::min(termDepth, maxDepth); d++)
            TRY(launch_wf_depth(S, F, P, W, d, true, cfg, merge && d + 1 == termDepth));
        if (termDepth >= maxDepth)
            return VH_OK;
        if (merge)
            TRY(launch_wf_trace_resolve(S, F, P, W, termDepth, cfg));
        else
            TRY(launch_wf_depth(S, F, P, W, termDepth, false, cfg));
        return launch_nrc_query_shade(S, F, P, W, nn->queries.as<float4>(), termDepth, cfg);
    };
    auto renderCache = [&](uint32_t f) -> int {
        if (termDepth >= maxDepth)
            return VH_OK;
        Params P = base;
        P.w[6] = a->samples + f;
        P.w[31] = (uint32_t)npix;
        const NrcNet net = nn->net(false, true);
        uint32_t grid = std::min<uint32_t>((uint32_t)((npix + 16 * kInferWaves - 1) / (16 * kInferWaves)), r->ctx->cus);
        k_nrc_infer<1><<<grid, 64 * kInferWaves, 0, s>>>(net, nn->queries.as<float4>(), W.counts + termDepth, 0u, nullptr, F.film);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return fail(VH_ERR_KERNEL, std::string("NRC inference: ") + hipGetErrorString(e));
        HIPTRY(hipEventRecord(nn->inferDone, s));
        inferIssued = true;
        // the depths past the cache carry only the paths it could not end: those that passed a
        // surface of opacity < 1 (none in a scene without such materials, where every path that
        // reaches the cache depth alive ends there); small grids, no spreading, so their near-empty
        // launches cost little
        if (!S.translucent)
            return VH_OK;
        Wavefront WT = W;
        WT.traceWaves = 0;
        const LaunchCfg tail{s, std::min<uint32_t>(cfg.wfGrid ? cfg.wfGrid : 1024u, r->ctx->cus)};
        for (uint32_t d = termDepth + 1; d < maxDepth; d++)
            TRY(launch_wf_depth(S, F, P, WT, d, true, tail));
        return VH_OK;
    };
    // Training frames of this call (the warm-up on the cache's first call, then one per frame); the
    // paths of training frame j + 2 run beside the Adam steps and render of frame j. Timing events
    // on the render stream: ev[0] at the start, then per training frame three: before and after
    // the render stream waits for the frame's Adam steps, and at the frame's end. The wait is the
    // training time the render stream sees; the rest is render time.
    const uint32_t warm = nn->frames == 0 ? nn->cfg.warmup_frames : 0;
    const uint32_t total = warm + frames;
    std::vector<hipEvent_t> ev(1 + 3 * (size_t)total, nullptr);
    for (auto& e : ev)
        HIPTRY(hipEventCreate(&e));
    int rc = check_hip(hipEventRecord(ev[0], s), "hipEventRecord");
    // the side streams start after everything issued before this call (scene uploads, parameters)
    for (auto& t : nn->sets)
        if (rc == VH_OK)
            rc = check_hip(hipStreamWaitEvent(t.stream, ev[0], 0), "hipStreamWaitEvent");
    if (rc == VH_OK)
        rc = check_hip(hipStreamWaitEvent(nn->adamStream, ev[0], 0), "hipStreamWaitEvent");
    const uint32_t tf0 = nn->trainFrames;
    for (uint32_t j = 0; j < std::min(total, 2u) && rc == VH_OK; j++)
        rc = issuePaths(tf0 + j);
    for (uint32_t j = 0; j < total && rc == VH_OK; j++) {
        rc = adam(tf0 + j);
        if (rc == VH_OK && j + 2 < total)
            rc = issuePaths(tf0 + j + 2);
        if (rc == VH_OK && j >= warm)
            rc = renderPaths(j - warm);
        if (rc == VH_OK)
            rc = check_hip(hipEventRecord(ev[1 + 3 * j], s), "hipEventRecord");
        if (rc == VH_OK)
            rc = check_hip(hipStreamWaitEvent(s, nn->adamDone[(tf0 + j) % vh_nrc::kBufs], 0), "hipStreamWaitEvent");
        if (rc == VH_OK)
            rc = check_hip(hipEventRecord(ev[2 + 3 * j], s), "hipEventRecord");
        if (rc == VH_OK && j >= warm)
            rc = renderCache(j - warm);
        if (rc == VH_OK)
            rc = check_hip(hipEventRecord(ev[3 + 3 * j], s), "hipEventRecord");
    }
    if (rc == VH_OK && hipEventSynchronize(ev[3 * total]) != hipSuccess)
        rc = fail(VH_ERR_KERNEL, "NRC frames failed on the device");
    if (rc != VH_OK) {
        for (auto& t : nn->sets)
            (void)hipStreamSynchronize(t.stream);
        (void)hipStreamSynchronize(nn->adamStream);
        (void)hipStreamSynchronize(s);
    }
    float trainMs = 0.0f, renderMs = 0.0f;
    for (uint32_t j = 0; j < total && rc == VH_OK; j++) {
        float t1 = 0.0f, t2 = 0.0f, t3 = 0.0f;
        (void)hipEventElapsedTime(&t1, ev[3 * j], ev[1 + 3 * j]);
        (void)hipEventElapsedTime(&t2, ev[1 + 3 * j], ev[2 + 3 * j]);
        (void)hipEventElapsedTime(&t3, ev[2 + 3 * j], ev[3 + 3 * j]);
        if (j < warm)
            trainMs += t1 + t2 + t3;
        else {
            trainMs += t2;
            renderMs += t1 + t3;
        }
    }
    for (auto e : ev)
        (void)hipEventDestroy(e);
    nn->trainFrames = tf0 + total;
    TRY(rc);
    float lossHost[2];
    HIPTRY(hipMemcpy(lossHost, nn->loss.p, 8, hipMemcpyDeviceToHost));
    nn->lastLoss = lossHost[0];
    nn->lastSamples = (uint32_t)lossHost[1];
    nn->frames += frames;
    nn->trainSeconds += (double)trainMs * 1.0e-3;
    nn->renderSeconds += (double)renderMs * 1.0e-3;
    a->samples += frames;
    a->gpuSeconds += (double)(trainMs + renderMs) * 1.0e-3;
    return VH_OK;
}

int vh_nrc_stats(const vh_nrc* nn, double out[4])
{
    if (!nn || !out)
        return fail(VH_ERR_INVALID_ARGUMENT, "null argument");
    out[0] = (double)nn->lastLoss;
    out[1] = nn->lastSamples;
    out[2] = nn->trainSeconds;
    out[3] = nn->renderSeconds;
    return VH_OK;
}

} // extern "C"
