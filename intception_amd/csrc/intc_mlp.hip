// intc_mlp.hip -- the quartet-magnitude network: 12 -> 64 -> 64 -> 1, ReLU.
//
// Predicts log10(max |integral| / (Q_bra Q_ket)) of a contracted quartet (the
// decay below the Schwarz bound, clamped at 0) from pair features,
// to tighten Schwarz screening (skip) and to route quartets between the FP64
// and FP32-recurrence kernels.  Weights are stored as BF16 for inference
// (FP32 master copies for training); every product accumulates in FP32.
//
//   intc_mlp_route   features (from the pair table) -> 3 layers on
//                    v_mfma_f32_16x16x16_bf16 -> route code per quartet
//   intc_mlp_train   one minibatch: forward + backward per sample (VALU FP32
//                    with the BF16-rounded weights), per-block partial
//                    gradients (deterministic, no atomics)
//   intc_mlp_adam    sums the partials and applies Adam to the master weights,
//                    refreshing the BF16 copies
//
// Pair feature table (8 floats per shell pair): log10 Q (Schwarz factor),
// log10 zeta_min, |AB|, la + lb, P_x, P_y, P_z (centre of the most diffuse
// primitive pair), lowest multipole order m (|la - lb| for a one-centre pair,
// 0 otherwise).  Quartet features (12, standardised with the model's
// mean/scale): log10 Q_bra, log10 Q_ket, log10 zeta_bra, log10 zeta_ket,
// |AB|, |CD|, L_bra, L_ket, R_PQ, log10 rho, s-type Coulomb decay
// log10[F0(rho R_PQ^2) (4 zeta_bra zeta_ket)^(1/4) / sqrt(zeta_bra + zeta_ket)],
// multipole decay -(m_bra + m_ket) log10(1 + R_PQ sqrt(rho)) (an order-m
// distribution of width ~rho^(-1/2) interacts ~(rho^(-1/2) / R)^m more weakly
// than a monopole once R exceeds the width).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>
#include <cmath>

#define INTC_API extern "C" __attribute__((visibility("default")))

namespace {

constexpr int NF = 12, KF = 16, NH = 64;
// FP32 master parameter layout
constexpr int OW1 = 0, OB1 = OW1 + NH * KF, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH, OB3 = OW3 + NH;
constexpr int NPARAM = OB3 + 1;
// BF16 copy layout
constexpr int BW1 = 0, BW2 = BW1 + NH * KF, BW3 = BW2 + NH * NH, NB16 = BW3 + NH;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// 12 unstandardised features of a quartet from its two pair-table rows a, c
__device__ __forceinline__ void features_from_rows(const float *a, const float *c, float *f)
{
    float dx = a[4] - c[4], dy = a[5] - c[5], dz = a[6] - c[6];
    float r2 = dx * dx + dy * dy + dz * dz;
    float za = exp10f(a[1]), zc = exp10f(c[1]);
    float rho = za * zc / (za + zc);
    f[0] = a[0];
    f[1] = c[0];
    f[2] = a[1];
    f[3] = c[1];
    f[4] = a[2];
    f[5] = c[2];
    f[6] = a[3];
    f[7] = c[3];
    const float R = sqrtf(r2);
    f[8] = R;
    f[9] = log10f(rho);
    // s-type Coulomb decay below the Schwarz bound for the diffuse pairs:
    // (ab|cd) / sqrt((ab|ab)(cd|cd)) = F0(T) (2 za)^(1/4) (2 zc)^(1/4) / sqrt(za + zc)
    const float T = rho * r2;
    const float f0 = T < 1e-6f ? 1.0f : 0.886226925f * erff(sqrtf(T)) / sqrtf(T);
    f[10] = log10f(f0) + 0.25f * log10f(4.0f * za * zc) - 0.5f * log10f(za + zc);
    f[11] = -(a[7] + c[7]) * log10f(1.0f + R * sqrtf(rho));
}

// 12 unstandardised features of quartet (pb | pk) into f[0..11]
__device__ __forceinline__ void quartet_features(const float *__restrict__ pf, int pb, int pk, float *f)
{
    features_from_rows(pf + 8 * (size_t)pb, pf + 8 * (size_t)pk, f);
}

// the two pair-table rows of quartet q (clamped to n - 1), as float4s; the
// table is 16-byte aligned (checked by the host)
struct QuartetRows {
    float4 a0, a1, c0, c1;
};

__device__ __forceinline__ int2 quartet_pairs_of(const int *__restrict__ qp, long long q, long long n)
{
    const long long qc = q < n ? q : n - 1;
    return *reinterpret_cast<const int2 *>(qp + 2 * qc);
}

__device__ __forceinline__ QuartetRows rows_of(const float *__restrict__ pf, int2 pr)
{
    const float4 *a = reinterpret_cast<const float4 *>(pf + 8 * (size_t)pr.x);
    const float4 *c = reinterpret_cast<const float4 *>(pf + 8 * (size_t)pr.y);
    return QuartetRows{a[0], a[1], c[0], c[1]};
}

}  // namespace

// ---------------------------------------------------------------------------
// features for training data: X[n][16] standardised, zero padded
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) intc_mlp_features_k(long long n, const float *__restrict__ pf,
                                                           const int *__restrict__ qp, const float *__restrict__ norm,
                                                           float *__restrict__ X)
{
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    float f[NF];
    quartet_features(pf, qp[2 * q], qp[2 * q + 1], f);
    for (int k = 0; k < NF; ++k) X[KF * q + k] = (f[k] - norm[k]) * norm[NF + k];
    for (int k = NF; k < KF; ++k) X[KF * q + k] = 0.0f;
}

// ---------------------------------------------------------------------------
// inference + routing on MFMA.  One wave per super-tile of 64 quartets: lane
// l computes the standardised features of quartet 64 t + l into LDS, then the
// wave runs the network on the four 16-quartet tiles, transposed (hidden
// units as rows, quartets as columns) so that each layer's accumulator is the
// next layer's B operand in registers:
//   H1^T = relu(W1 X^T + b1)      v_mfma_f32_16x16x16_bf16, 4 row tiles
//   H2^T = relu(W2 H1^T + b2)     v_mfma_f32_16x16x32_bf16, 4 row tiles x 2 K chunks
//   y    = w3 . bf16(H2^T) + b3   v_mfma_f32_16x16x32_bf16, 2 K chunks, with
//                                 w3 in every row of A (each row of D is y)
// Operands (lane l, r = l % 16, g = l / 16, slot s):
//   A[i][k]: i = r; B[k][n]: n = r; D[i][n]: n = r, i = 4 g + s (s < 4).
// The K index of a slot is the same function of (g, s) in A and B, so the
// K chunks of layer 2 may order the hidden-1 units freely: chunk c, slot s
// <-> unit 16 (2c + s / 4) + 4 g + s % 4, the unit layer 1's D holds there.
// route: 2 if pred + m_skip < log_skip or log10 Q_bra + log10 Q_ket <
// log_skip (the Schwarz bound alone), else 1 if pred + m_32 < log_32 and
// log10 Q_bra + log10 Q_ket < log_qq32, else 0.
// ---------------------------------------------------------------------------
constexpr int SF = KF + 1;  // LDS row: 16 standardised features (zero padded), padded to 17 words

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// max(x, 0) on the bit pattern: one integer max, no NaN canonicalisation
__device__ __forceinline__ float relu_bits(float x) { return __int_as_float(max(__float_as_int(x), 0)); }

__global__ void __launch_bounds__(256) intc_mlp_route_k(long long n, const float *__restrict__ pf,
                                                        const int *__restrict__ qp, const float *__restrict__ norm,
                                                        const uint16_t *__restrict__ wb,
                                                        const float *__restrict__ master, float m_skip,
                                                        float log_skip, float m_32, float log_32, float log_qq32,
                                                        float *__restrict__ pred, signed char *__restrict__ route)
{
    __shared__ float sF[4][64][SF];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, g4 = lane >> 4;
    // A operands of layers 1 and 2 (rows 16 nt + r16), and the biases and
    // W3 entries of the hidden units this lane's accumulator slots hold
    bf16x4 a1w[4];
    bf16x8 a2w[4][2], a3w[2];
    f32x4 b1v[4], b2v[4];
    for (int nt = 0; nt < 4; ++nt) {
        const int row = 16 * nt + r16;
        for (int j = 0; j < 4; ++j) a1w[nt][j] = (short)wb[BW1 + row * KF + 4 * g4 + j];
        for (int c = 0; c < 2; ++c)
            for (int sl = 0; sl < 8; ++sl)
                a2w[nt][c][sl] = __builtin_bit_cast(__bf16, wb[BW2 + row * NH + 16 * (2 * c + sl / 4) + 4 * g4 + sl % 4]);
        for (int j = 0; j < 4; ++j) {
            const int h = 16 * nt + 4 * g4 + j;
            b1v[nt][j] = master[OB1 + h];
            b2v[nt][j] = master[OB2 + h];
        }
    }
    // A of layer 3: w3 in every row, K chunk c slot s = hidden-2 unit
    // 16 (2c + s/4) + 4 g4 + s%4 (the units layer 2's D holds there)
    for (int c = 0; c < 2; ++c)
        for (int sl = 0; sl < 8; ++sl)
            a3w[c][sl] = __builtin_bit_cast(__bf16, wb[BW3 + 16 * (2 * c + sl / 4) + 4 * g4 + sl % 4]);
    const float b3 = master[OB3];
    const long long nsuper = (n + 63) / 64;
    const long long stride = (long long)gridDim.x * (blockDim.x >> 6);
    // software pipeline over this wave's super-tiles: the pair ids of tile
    // t + 2 stride and the pair-table rows of tile t + stride are in flight
    // while tile t is evaluated
    long long t = (long long)blockIdx.x * (blockDim.x >> 6) + wave;
    int2 pr_next = quartet_pairs_of(qp, (t + stride) * 64 + lane, n);
    QuartetRows rows = rows_of(pf, quartet_pairs_of(qp, t * 64 + lane, n));
    for (; t < nsuper; t += stride) {
        const long long q = t * 64 + lane;
        const QuartetRows cur = rows;
        if (t + stride < nsuper) {
            rows = rows_of(pf, pr_next);
            pr_next = quartet_pairs_of(qp, (t + 2 * stride) * 64 + lane, n);
        }
        float f[NF];
        {
            const float a[8] = {cur.a0.x, cur.a0.y, cur.a0.z, cur.a0.w, cur.a1.x, cur.a1.y, cur.a1.z, cur.a1.w};
            const float c[8] = {cur.c0.x, cur.c0.y, cur.c0.z, cur.c0.w, cur.c1.x, cur.c1.y, cur.c1.z, cur.c1.w};
            features_from_rows(a, c, f);
        }
        for (int k = 0; k < NF; ++k) sF[wave][lane][k] = (f[k] - norm[k]) * norm[NF + k];
        for (int k = NF; k < KF; ++k) sF[wave][lane][k] = 0.0f;
        const float lq = f[0] + f[1];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        float mine = 0.0f;  // network output of quartet 64 t + lane
        // all four tiles always: rows past n hold the features of quartet
        // n - 1 and are not stored
        for (int tt = 0; tt < 4; ++tt) {
            // B of layer 1: X^T, features 4 g4 + j of quartet 16 tt + r16
            bf16x4 bx;
            for (int j = 0; j < 4; ++j) bx[j] = __builtin_bit_cast(short, (__bf16)sF[wave][16 * tt + r16][4 * g4 + j]);
            f32x4 d1[4];
            for (int nt = 0; nt < 4; ++nt)
                d1[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1w[nt], bx, b1v[nt], 0, 0, 0);
            // B of layer 2: relu(H1) in BF16, K chunk c slot s = unit 16 (2c + s/4) + 4 g4 + s%4
            bf16x8 h1[2];
            for (int c = 0; c < 2; ++c)
                for (int sl = 0; sl < 8; ++sl) {
                    const float h = d1[2 * c + sl / 4][sl % 4];
                    h1[c][sl] = (__bf16)relu_bits(h);
                }
            f32x4 d2[4];
            for (int nt = 0; nt < 4; ++nt) {
                d2[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2w[nt][0], h1[0], b2v[nt], 0, 0, 0);
                d2[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2w[nt][1], h1[1], d2[nt], 0, 0, 0);
            }
            // B of layer 3: relu(H2) in BF16, in the K order of layer 2's B
            bf16x8 h2[2];
            for (int c = 0; c < 2; ++c)
                for (int sl = 0; sl < 8; ++sl)
                    h2[c][sl] = (__bf16)relu_bits(d2[2 * c + sl / 4][sl % 4]);
            const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
            f32x4 d3 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3w[0], h2[0], zero, 0, 0, 0);
            d3 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3w[1], h2[1], d3, 0, 0, 0);
            if (g4 == tt) mine = d3[0];  // lane 16 tt + r16 keeps quartet 16 tt + r16
        }
        if (q < n) {
            // the network predicts log10(max|integral| / (Q_bra Q_ket)) <= 0
            const float y = lq + fminf(mine + b3, 0.0f);
            if (pred) pred[q] = y;
            if (route) {
                signed char r = 0;
                // skip on the network's calibrated prediction, or on the
                // Schwarz bound itself (|(ab|cd)| <= Q_bra Q_ket; 1e-4 decade
                // guard for the FP32 logarithms)
                if (y + m_skip < log_skip || lq < log_skip - 1e-4f) r = 2;
                else if (y + m_32 < log_32 && lq < log_qq32) r = 1;
                route[q] = r;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// ---------------------------------------------------------------------------
// training: one block of 64 threads handles 32 samples of the minibatch
// ---------------------------------------------------------------------------
constexpr int TS = 32;

__global__ void __launch_bounds__(64) intc_mlp_train_k(int nb, const float *__restrict__ X,
                                                       const float *__restrict__ T, const int *__restrict__ idx,
                                                       const uint16_t *__restrict__ wb,
                                                       const float *__restrict__ master, float *__restrict__ partial,
                                                       float *__restrict__ loss, float under_weight)
{
    __shared__ float sX[TS][KF];
    __shared__ float sH1[TS][NH + 1];
    __shared__ float sH2[TS][NH + 1];
    __shared__ float sD2[TS][NH + 1];
    __shared__ float sD1[TS][NH + 1];
    __shared__ float sDy[TS];
    __shared__ float sW2[NH][NH + 1];
    const int tid = threadIdx.x;
    const int s0 = blockIdx.x * TS;
    for (int i = tid; i < NH * NH; i += 64) sW2[i / NH][i % NH] = bf2f(wb[BW2 + i]);
    for (int i = tid; i < TS * KF; i += 64) {
        int s = i / KF, k = i % KF;
        sX[s][k] = (s0 + s < nb) ? X[(size_t)KF * idx[s0 + s] + k] : 0.0f;
    }
    __syncthreads();
    if (tid < TS) {
        const int s = tid;
        const bool valid = s0 + s < nb;
        // forward with the BF16 weights
        for (int h = 0; h < NH; ++h) {
            float z = master[OB1 + h];
            for (int k = 0; k < KF; ++k) z += bf2f(wb[BW1 + h * KF + k]) * bf2f(f2bf(sX[s][k]));
            sH1[s][h] = z > 0.f ? z : 0.f;
        }
        float y = master[OB3];
        for (int h = 0; h < NH; ++h) {
            float z = master[OB2 + h];
            for (int k = 0; k < NH; ++k) z += sW2[h][k] * bf2f(f2bf(sH1[s][k]));
            z = z > 0.f ? z : 0.f;
            sH2[s][h] = z;
            y += bf2f(f2bf(z)) * bf2f(wb[BW3 + h]);
        }
        const float t = valid ? T[idx[s0 + s]] : 0.0f;
        const float r = valid ? (y - t) : 0.0f;
        // weighted squared error: under-predictions (y < t) weigh under_weight
        const float w = r < 0.0f ? under_weight : 1.0f;
        sDy[s] = 2.0f * w * r / (float)nb;
        if (valid && loss) atomicAdd(loss, w * r * r / (float)nb);
        // backward through layer 3 and the second ReLU
        for (int h = 0; h < NH; ++h) sD2[s][h] = sH2[s][h] > 0.f ? sDy[s] * bf2f(wb[BW3 + h]) : 0.0f;
        // through W2 and the first ReLU
        for (int k = 0; k < NH; ++k) {
            float g = 0.f;
            for (int h = 0; h < NH; ++h) g += sD2[s][h] * sW2[h][k];
            sD1[s][k] = sH1[s][k] > 0.f ? g : 0.0f;
        }
    }
    __syncthreads();
    float *pg = partial + (size_t)blockIdx.x * (NPARAM + 1);
    // dW2[h][k] = sum_s d2[s][h] * bf16(h1[s][k]); db2[h] = sum_s d2[s][h]
    for (int i = tid; i < NH * NH; i += 64) {
        int h = i / NH, k = i % NH;
        float g = 0.f;
        for (int s = 0; s < TS; ++s) g += sD2[s][h] * bf2f(f2bf(sH1[s][k]));
        pg[OW2 + i] = g;
    }
    for (int i = tid; i < NH * KF; i += 64) {
        int h = i / KF, k = i % KF;
        float g = 0.f;
        for (int s = 0; s < TS; ++s) g += sD1[s][h] * bf2f(f2bf(sX[s][k]));
        pg[OW1 + i] = g;
    }
    {
        const int h = tid;
        float g1 = 0.f, g2 = 0.f, g3 = 0.f;
        for (int s = 0; s < TS; ++s) {
            g1 += sD1[s][h];
            g2 += sD2[s][h];
            g3 += sDy[s] * bf2f(f2bf(sH2[s][h]));
        }
        pg[OB1 + h] = g1;
        pg[OB2 + h] = g2;
        pg[OW3 + h] = g3;
        if (h == 0) {
            float g = 0.f;
            for (int s = 0; s < TS; ++s) g += sDy[s];
            pg[OB3] = g;
        }
    }
}

// sum of per-block partial gradients + Adam; refreshes the BF16 copies
__global__ void __launch_bounds__(256) intc_mlp_adam_k(int nblk, const float *__restrict__ partial,
                                                       float *__restrict__ master, float *__restrict__ m1,
                                                       float *__restrict__ m2, uint16_t *__restrict__ wb, float lr,
                                                       float b1, float b2, float eps, float c1, float c2,
                                                       float *__restrict__ grad_out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    float g = 0.f;
    for (int b = 0; b < nblk; ++b) g += partial[(size_t)b * (NPARAM + 1) + i];
    // padding columns of W1 stay zero
    if (i < OB1 && (i % KF) >= NF) g = 0.f;
    if (grad_out) grad_out[i] = g;
    if (lr > 0.f) {
        float a = b1 * m1[i] + (1.f - b1) * g;
        float v = b2 * m2[i] + (1.f - b2) * g * g;
        m1[i] = a;
        m2[i] = v;
        master[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
    }
    float w = master[i];
    if (i < OB1) wb[BW1 + i] = f2bf(w);
    else if (i >= OW2 && i < OB2) wb[BW2 + (i - OW2)] = f2bf(w);
    else if (i >= OW3 && i < OB3) wb[BW3 + (i - OW3)] = f2bf(w);
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
INTC_API int intc_mlp_num_params(void) { return NPARAM; }
INTC_API int intc_mlp_num_bf16(void) { return NB16; }
INTC_API int intc_mlp_train_block(void) { return TS; }

INTC_API int intc_mlp_features(long long n, const float *pair_feat, const int *quartet_pairs, const float *norm,
                               float *X, void *stream)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(intc_mlp_features_k, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n,
                       pair_feat, quartet_pairs, norm, X);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

INTC_API int intc_mlp_route(long long n, const float *pair_feat, const int *quartet_pairs, const float *norm,
                            const uint16_t *wb, const float *master, float m_skip, float log_skip, float m_32,
                            float log_32, float log_qq32, float *pred, signed char *route, void *stream)
{
    if (n <= 0) return 0;
    // the kernel reads pair-table rows as float4 and pair ids as int2
    if (((uintptr_t)pair_feat & 15) || ((uintptr_t)quartet_pairs & 7)) return -1;
    long long supers = (n + 63) / 64;
    int grid = (int)std::min<long long>((supers + 3) / 4, 256 * 12);
    hipLaunchKernelGGL(intc_mlp_route_k, dim3(grid), dim3(256), 0, (hipStream_t)stream, n, pair_feat,
                       quartet_pairs, norm, wb, master, m_skip, log_skip, m_32, log_32, log_qq32, pred, route);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

// one minibatch of nb samples (indices idx into X/T); partial: nblk*(NPARAM+1)
// floats with nblk = ceil(nb / train_block); loss accumulates the batch's
// weighted MSE (under-predictions weigh under_weight)
INTC_API int intc_mlp_train_step(int nb, const float *X, const float *T, const int *idx, const uint16_t *wb,
                                 const float *master, float *partial, float *loss, float under_weight, void *stream)
{
    if (nb <= 0) return 0;
    int nblk = (nb + TS - 1) / TS;
    hipLaunchKernelGGL(intc_mlp_train_k, dim3(nblk), dim3(64), 0, (hipStream_t)stream, nb, X, T, idx, wb, master,
                       partial, loss, under_weight);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

INTC_API int intc_mlp_adam(int nblk, const float *partial, float *master, float *m1, float *m2, uint16_t *wb,
                           float lr, float beta1, float beta2, float eps, int step, float *grad_out, void *stream)
{
    float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    if (lr <= 0.f) c1 = c2 = 1.f;
    hipLaunchKernelGGL(intc_mlp_adam_k, dim3((NPARAM + 255) / 256), dim3(256), 0, (hipStream_t)stream, nblk, partial,
                       master, m1, m2, wb, lr, beta1, beta2, eps, c1, c2, grad_out);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}
