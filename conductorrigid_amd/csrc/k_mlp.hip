// k_mlp.hip -- learned warm start for the GMRES solve: a per-bead MLP that predicts
// induced charges from local geometry, with BF16 MFMA forward, backward and
// weight-gradient kernels accumulating in FP32, and an Adam update on FP32 master
// weights.  No reference counterpart: the reference starts GMRES from the previous
// step's solution (cppmodule/ConductorRigid.cu:1125-1127).
//
// At fixed geometry the solved charges are linear in the field and the body charges,
// q = E0 . a + c, so the network outputs the four geometric responses (a_x, a_y, a_z, c)
// and the field enters only as a gate:
//   trunk: x[0:kMlpTrunk] -> z1 = x W1 + b1 -> h1 = silu(z1) (128) -> z2 = h1 W2 + b2
//          -> h2 = silu(z2) (128) -> u = h2 W3 + b3 (4)
//   y = g . u with the row's gate g = (E0x, E0y, E0z, 1), E0 read from the row's gate
//   features (each component stored as a bf16 pair hi + lo).
// Rounding points (shared with the CPU reference in oracle/mlp_ref.py): the features,
// W1, W2 and W3 are bf16; h1, h2 are rounded to bf16 where they feed the next layer;
// dz1, dz2 are rounded to bf16 where they feed a matrix product; everything else,
// including every accumulation, is FP32.
//
// MFMA: v_mfma_f32_16x16x16_bf16 (wave64).  For D = A B with A 16x16 and B 16x16, lane l
// holds A[l % 16][4 (l / 16) + j] and B[4 (l / 16) + j][l % 16], j = 0..3, and receives
// D[4 (l / 16) + r][l % 16], r = 0..3.
#include "cr_kernels.h"

#include <algorithm>
#include <cmath>

namespace cr {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int KIN = kMlpIn;        // 176 feature columns
constexpr int TRUNK = kMlpTrunk;   // 168 columns feed the trunk; the rest are gates
constexpr int H = kMlpHidden;      // 128
constexpr int NO = kMlpOut;        // 4 heads
constexpr int KT1 = KIN / 16;      // 11 k-chunks of layer 1
constexpr int NT = H / 16;         // 8 column tiles
constexpr int kWaves = 4;

__device__ __forceinline__ float bf2f(unsigned short u) { return __uint_as_float((unsigned)u << 16); }
__device__ __forceinline__ unsigned short f2bf(float f)
{
    const unsigned u = __float_as_uint(f);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);  // round to nearest even
}
__device__ __forceinline__ float silu(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = __builtin_amdgcn_rcpf(1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ bf16x4 load4(const unsigned short* p)
{
    const uint2 v = *reinterpret_cast<const uint2*>(p);
    bf16x4 r;
    r[0] = (short)(v.x & 0xffffu);
    r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xffffu);
    r[3] = (short)(v.y >> 16);
    return r;
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Sum over the 16 lanes that share l / 16 (xor over lane bits 0..3).
__device__ __forceinline__ float sum16(float v)
{
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// gate of a row: (E0x, E0y, E0z, 1), each field component a bf16 pair hi + lo
__device__ __forceinline__ void row_gate(const unsigned short* __restrict__ xr, float g[NO])
{
#pragma unroll
    for (int d = 0; d < 3; ++d) g[d] = bf2f(xr[kMlpGate + d]) + bf2f(xr[kMlpGate + 3 + d]);
    g[3] = 1.0f;
}

// Minimum-image difference and squared distance without contraction into fused
// multiply-adds, so that the CPU reference reproduces the neighbour order bit for bit.
// The image count is rint(d * (1 / L)), which differs from rint(d / L) only where d / L
// lies within rounding of a half-integer; a pair closer than half the box (every
// neighbour-list pair, every bead and its body's centre) never does, so the result is
// bit-identical to d - L rint(d / L) there without a double division per component.
__device__ __forceinline__ double mi_inv(double d, double L, double invL)
{
    return __dadd_rn(d, -__dmul_rn(L, rint(d * invL)));
}
__device__ __forceinline__ float dist2(double dx, double dy, double dz)
{
    return (float)__dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
}

// ---- features -------------------------------------------------------------
// Per bead, columns: the kMlpNeigh nearest neighbours with r < kMlpRadius (minimum
// image), by increasing distance, each as d = (dx, dy, dz) / kMlpRadius, a same-body
// flag (+1 same body, -1 other) and |d| (empty slots zero) [0, 160); the body's net
// charge over its bead count [160]; the offset o from its rigid body's centre of mass
// / 4 and |o| (zero for isolated beads) [161, 165); zero [165, 168); the gate: E0 as
// bf16 pairs hi [168, 171) and lo = E0 - hi [171, 174); zero [174, 176).
// One thread per bead, 64 beads per workgroup.  The K nearest stay sorted in registers
// (an unrolled insertion: entries after the first one farther than the candidate shift
// down one slot, so equal distances keep list order).  Neighbours are read four at a
// time so that their loads are in flight together; they are inserted in list order.
// Rows are assembled in LDS in two halves of kFeatHalf columns, and each half of the
// workgroup's 64 consecutive rows leaves as dword stores.
constexpr int kFeatBlock = 64;
constexpr int kFeatHalf = KIN / 2;         // 88 columns per half
constexpr int kFeatPitch = kFeatHalf + 2;  // ushorts per LDS row: 45 dwords, odd, so rows fall on distinct banks
constexpr int kFeatBatch = 4;

// column col of the row is written when it lies in the half that starts at lo
__device__ __forceinline__ void put(unsigned short* x, int lo, int col, unsigned short v)
{
    if (col >= lo && col < lo + kFeatHalf) x[col - lo] = v;
}

__global__ __launch_bounds__(kFeatBlock) void k_mlp_features(SysView s, double Lx, double Ly, double Lz, double E0x,
                                                             double E0y, double E0z,
                                                             const unsigned* __restrict__ body_count,
                                                             unsigned short* __restrict__ X)
{
    __shared__ unsigned short rows[kFeatBlock * kFeatPitch];
    constexpr int K = kMlpNeigh;
    const int t = threadIdx.x;
    const double iLx = 1.0 / Lx, iLy = 1.0 / Ly, iLz = 1.0 / Lz;
    const long b0 = (long)blockIdx.x * kFeatBlock;
    const long i = b0 + t;
    const bool live = i < s.N;
    unsigned short* x = rows + t * kFeatPitch;
    float dist[K];
    unsigned idx[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        dist[k] = INFINITY;
        idx[k] = 0;
    }
    int cnt = 0;
    double4 ri = make_double4(0, 0, 0, 0);
    unsigned bi = 0;
    if (live) {
        ri = s.pos[i];
        bi = s.iso[i];
        const unsigned head = s.head_list[i], nn = s.n_neigh[i];
        const float r2max = (float)(kMlpRadius * kMlpRadius);
        for (unsigned n0 = 0; n0 < nn; n0 += kFeatBatch) {
            unsigned jb[kFeatBatch];
            double4 rb[kFeatBatch];
#pragma unroll
            for (int u = 0; u < kFeatBatch; ++u) jb[u] = s.nlist[head + (n0 + u < nn ? n0 + u : n0)];
#pragma unroll
            for (int u = 0; u < kFeatBatch; ++u) rb[u] = s.pos[jb[u]];
#pragma unroll
            for (int u = 0; u < kFeatBatch; ++u) {
                if (n0 + u >= nn) break;
                const double4 rj = rb[u];
                const float d2 = dist2(mi_inv(ri.x - rj.x, Lx, iLx), mi_inv(ri.y - rj.y, Ly, iLy), mi_inv(ri.z - rj.z, Lz, iLz));
                // in range, and closer than the current K-th (an empty slot holds +inf)
                if (!(d2 < r2max) || !(d2 < dist[K - 1])) continue;
                float cd = d2;
                unsigned cj = jb[u];
                bool shift = false;
#pragma unroll
                for (int p = 0; p < K; ++p) {
                    shift = shift || (cd < dist[p]);
                    const float od = dist[p];
                    const unsigned oj = idx[p];
                    dist[p] = shift ? cd : od;
                    idx[p] = shift ? cj : oj;
                    cd = shift ? od : cd;
                    cj = shift ? oj : cj;
                }
                cnt += cnt < K;
            }
        }
    }
    const long left = (long)s.N - b0;
    const int nr = left < kFeatBlock ? (int)left : kFeatBlock;
    constexpr int DW = KIN / 2, HDW = kFeatHalf / 2;
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        const int lo = half * kFeatHalf;
        if (live) {
            const float inv = 1.0f / (float)kMlpRadius;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (5 * k + 4 < lo || 5 * k >= lo + kFeatHalf) continue;
                float f0 = 0, f1 = 0, f2 = 0, f3 = 0, f4 = 0;
                if (k < cnt) {
                    const double4 rj = s.pos[idx[k]];
                    f0 = __fmul_rn((float)mi_inv(rj.x - ri.x, Lx, iLx), inv);
                    f1 = __fmul_rn((float)mi_inv(rj.y - ri.y, Ly, iLy), inv);
                    f2 = __fmul_rn((float)mi_inv(rj.z - ri.z, Lz, iLz), inv);
                    f3 = (s.iso[idx[k]] == bi) ? 1.0f : -1.0f;
                    f4 = __fmul_rn(__fsqrt_rn(dist[k]), inv);
                }
                put(x, lo, 5 * k + 0, f2bf(f0));
                put(x, lo, 5 * k + 1, f2bf(f1));
                put(x, lo, 5 * k + 2, f2bf(f2));
                put(x, lo, 5 * k + 3, f2bf(f3));
                put(x, lo, 5 * k + 4, f2bf(f4));
            }
            if (half == 1) {
                int o = 5 * K;
                const unsigned nb = body_count[bi];
                put(x, lo, o++, f2bf((float)(s.body_charge[bi] / (double)(nb > 0 ? nb : 1))));
                float cx = 0, cy = 0, cz = 0;
                const unsigned r = s.rigid_membership[i];
                if (r != NOT_MEMBER) {
                    const double4 c = s.rigid_com[r];
                    cx = __fmul_rn((float)mi_inv(ri.x - c.x, Lx, iLx), 0.25f);
                    cy = __fmul_rn((float)mi_inv(ri.y - c.y, Ly, iLy), 0.25f);
                    cz = __fmul_rn((float)mi_inv(ri.z - c.z, Lz, iLz), 0.25f);
                }
                put(x, lo, o++, f2bf(cx));
                put(x, lo, o++, f2bf(cy));
                put(x, lo, o++, f2bf(cz));
                put(x, lo, o++,
                    f2bf(__fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(cx, cx), __fmul_rn(cy, cy)), __fmul_rn(cz, cz)))));
                for (; o < kMlpGate; ++o) put(x, lo, o, 0);
                const float e[3] = {(float)E0x, (float)E0y, (float)E0z};
                for (int d = 0; d < 3; ++d) {
                    const unsigned short hi = f2bf(e[d]);
                    put(x, lo, kMlpGate + d, hi);
                    put(x, lo, kMlpGate + 3 + d, f2bf(e[d] - bf2f(hi)));
                }
                for (o = kMlpGate + 6; o < KIN; ++o) put(x, lo, o, 0);
            }
        }
        __syncthreads();
        // this half of rows b0 .. b0 + nr - 1: HDW dwords of each KIN / 2-dword row
        const unsigned* src = reinterpret_cast<const unsigned*>(rows);
        unsigned* dst = reinterpret_cast<unsigned*>(X + b0 * KIN) + half * HDW;
        for (int d = t; d < nr * HDW; d += kFeatBlock) {
            const int row = d / HDW, col = d - row * HDW;
            dst[row * DW + col] = src[row * (kFeatPitch / 2) + col];
        }
        __syncthreads();
    }
}

// ---- forward ----------------------------------------------------------------
// One wavefront per T tiles of 16 rows: each weight fragment is loaded once and feeds
// the T tiles' MFMAs (every row still sees the same products in the same order).  TRAIN
// additionally stores z1, z2 (FP32) and the bf16 h1, h2 for the backward pass.
constexpr int kFwdTilesTrain = 2, kFwdTilesInfer = 4;
template <bool TRAIN, int T>
__global__ __launch_bounds__(64 * kWaves) void k_mlp_forward(int M, const unsigned short* __restrict__ X,
                                                             MlpDevWeights w, float* __restrict__ y,
                                                             float* __restrict__ z1g, float* __restrict__ z2g,
                                                             unsigned short* __restrict__ h1g,
                                                             unsigned short* __restrict__ h2g)
{
    __shared__ unsigned short hs[kWaves][16 * T][H + 8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = (blockIdx.x * kWaves + wave) * 16 * T;
    if (m0 >= M) return;  // whole wavefront leaves together; no block barrier follows
    const int lr = lane & 15, lq = lane >> 4;
    // layer 1: A = trunk columns of the rows, kept in registers; gate columns masked.
    // Rows past M repeat the last one and are never stored.
    bf16x4 a1[T][KT1];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int arow = min(m0 + 16 * t + lr, M - 1);
#pragma unroll
        for (int kt = 0; kt < KT1; ++kt) {
            a1[t][kt] = load4(X + (long)arow * KIN + kt * 16 + 4 * lq);
            if (kt * 16 + 4 * lq >= TRUNK) a1[t][kt] = bf16x4{0, 0, 0, 0};
        }
    }
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        bf16x4 bw[KT1];
#pragma unroll
        for (int kt = 0; kt < KT1; ++kt) bw[kt] = load4(w.W1t + (long)(nt * 16 + lr) * KIN + kt * 16 + 4 * lq);
        f32x4 acc[T];
#pragma unroll
        for (int t = 0; t < T; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < KT1; ++kt)
#pragma unroll
            for (int t = 0; t < T; ++t) acc[t] = mfma(a1[t][kt], bw[kt], acc[t]);
        const int n = nt * 16 + lr;
        const float b = w.b1[n];
#pragma unroll
        for (int t = 0; t < T; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * t + 4 * lq + r;
                const float z = acc[t][r] + b;
                const unsigned short h = f2bf(silu(z));
                hs[wave][row][n] = h;
                if (TRAIN && m0 + row < M) {
                    z1g[(long)(m0 + row) * H + n] = z;
                    h1g[(long)(m0 + row) * H + n] = h;
                }
            }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    // layer 2: A = h1 rows from LDS
    bf16x4 a2[T][NT];
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
        for (int kt = 0; kt < NT; ++kt) {
            bf16x4 v;
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] = (short)hs[wave][16 * t + lr][kt * 16 + 4 * lq + j];
            a2[t][kt] = v;
        }
    float part[T][4][NO];
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int o = 0; o < NO; ++o) part[t][r][o] = 0.f;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        bf16x4 bw[NT];
#pragma unroll
        for (int kt = 0; kt < NT; ++kt) bw[kt] = load4(w.W2t + (long)(nt * 16 + lr) * H + kt * 16 + 4 * lq);
        f32x4 acc[T];
#pragma unroll
        for (int t = 0; t < T; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < NT; ++kt)
#pragma unroll
            for (int t = 0; t < T; ++t) acc[t] = mfma(a2[t][kt], bw[kt], acc[t]);
        const int n = nt * 16 + lr;
        const float b = w.b2[n];
        float w3[NO];
#pragma unroll
        for (int o = 0; o < NO; ++o) w3[o] = bf2f(w.W3[n * NO + o]);
#pragma unroll
        for (int t = 0; t < T; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * t + 4 * lq + r;
                const float z = acc[t][r] + b;
                const unsigned short h = f2bf(silu(z));
                const float hf = bf2f(h);
#pragma unroll
                for (int o = 0; o < NO; ++o) part[t][r][o] += hf * w3[o];
                if (TRAIN && m0 + row < M) {
                    z2g[(long)(m0 + row) * H + n] = z;
                    h2g[(long)(m0 + row) * H + n] = h;
                }
            }
    }
    // layer 3: u = h2 W3 + b3 reduced over the 16 lanes holding one row's columns; y = g . u
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float u[NO];
#pragma unroll
            for (int o = 0; o < NO; ++o) u[o] = sum16(part[t][r][o]) + w.b3[o];
            const int m = m0 + 16 * t + 4 * lq + r;
            if (lr == 0 && m < M) {
                float g[NO];
                row_gate(X + (long)m * KIN, g);
                y[m] = ((g[0] * u[0] + g[1] * u[1]) + g[2] * u[2]) + g[3] * u[3];
            }
        }
}

// ---- inference (persistent, gfx950 MFMA 16x16x32, chained layers) -------------
// Every layer is computed transposed, D[n][m] = sum_k W[n][k] act[m][k], with the weights
// as the A operand (from LDS) and the activations as the B operand (registers).  In
// v_mfma_f32_16x16x32_bf16 lane l holds 8 values of A row l % 16 and of B column l % 16
// (slots j = 0..7 of lane group l / 16; A and B pair the same slots, so any assignment
// of k to (group, slot) that both operands share gives the full sum), and receives
// D[4 (l / 16) + r][l % 16], r = 0..3.  A lane therefore leaves a layer holding row
// m = l % 16 at columns n = 16 nt + 4 (l / 16) + r: for the 32-wide k chunk kc of the next
// layer, slots 0..3 come from nt = 2 kc and slots 4..7 from nt = 2 kc + 1, i.e.
//   k(kc, g, j) = 32 kc + 16 (j / 4) + 4 g + j % 4,
// and the next layer's weights are staged in LDS with their columns in that order, so
// activations never leave registers.  Layer 3 (4 heads) is one more MFMA chain with rows
// 4..15 of its A operand zero; its D rows 0..3 are the heads of row m in lanes 0..15.
// Each workgroup stages the weights once (78 KB) and its wavefronts loop over tiles of
// 32 rows.  Rounding points as k_mlp_forward: bf16
// operands and activations, FP32 sums (the bias starts the sum here).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float fl2 __attribute__((ext_vector_type(2)));
constexpr int kInferWaves = 12;
constexpr int kInferT = 2;              // 16-row tiles per wavefront pass
constexpr int kW1Pitch = TRUNK;         // ushorts per LDS row of W1 (336 B; conflict-free 16-byte reads)
constexpr int kW2Pitch = H + 8;         // ushorts per LDS row of the permuted W2 (272 B)
constexpr int KC1 = (TRUNK + 31) / 32;  // 6 chunks of 32 columns (trunk 168, zero above)
constexpr int KC2 = H / 32;             // 4
constexpr size_t kInferLdsW1 = sizeof(unsigned short) * H * kW1Pitch;
constexpr size_t kInferLdsW2 = sizeof(unsigned short) * H * kW2Pitch;
constexpr size_t kInferLdsW3 = sizeof(unsigned short) * NO * H;
constexpr size_t kInferLds = kInferLdsW1 + kInferLdsW2 + kInferLdsW3 + 2 * sizeof(float) * H;
static_assert(kInferLds <= 160 * 1024, "inference weights fit one CU's LDS");
static_assert(NO == 4 && H % 32 == 0 && TRUNK % 8 == 0, "inference tiling");

__device__ __forceinline__ bf16x8 ld8(const unsigned short* p)
{
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ unsigned pack_bf16(float a, float b)
{
    const fl2 v = {a, b};
    return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2));  // round to nearest even
}

// column of the permuted weight row that holds original column n (the k order above)
__device__ __forceinline__ int perm_col(int n)
{
    const int kc = n >> 5, w = n & 31, hi = w >> 4, g = (w >> 2) & 3, j = w & 3;
    return 32 * kc + 8 * g + 4 * hi + j;
}

// activations of one 16-row tile for chunk kc from the accumulators of nt = 2 kc, 2 kc + 1
__device__ __forceinline__ bf16x8 act8(f32x4 lo, f32x4 hi)
{
    const uint4 u = make_uint4(pack_bf16(silu(lo[0]), silu(lo[1])), pack_bf16(silu(lo[2]), silu(lo[3])),
                               pack_bf16(silu(hi[0]), silu(hi[1])), pack_bf16(silu(hi[2]), silu(hi[3])));
    return __builtin_bit_cast(bf16x8, u);
}

__global__ __launch_bounds__(64 * kInferWaves) void k_mlp_infer(int M, const unsigned short* __restrict__ X,
                                                                   MlpDevWeights w, float* __restrict__ y)
{
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned short* W1s = reinterpret_cast<unsigned short*>(smem);
    unsigned short* W2s = reinterpret_cast<unsigned short*>(smem + kInferLdsW1);
    unsigned short* W3s = reinterpret_cast<unsigned short*>(smem + kInferLdsW1 + kInferLdsW2);
    float* b1s = reinterpret_cast<float*>(smem + kInferLdsW1 + kInferLdsW2 + kInferLdsW3);
    float* b2s = b1s + H;
    // stage: W1 trunk columns as they are; W2 and W3^T with their columns permuted
    for (int k = threadIdx.x; k < H * (TRUNK / 8); k += blockDim.x) {
        const int n = k / (TRUNK / 8), c = k - n * (TRUNK / 8);
        *reinterpret_cast<uint4*>(W1s + n * kW1Pitch + 8 * c) = *reinterpret_cast<const uint4*>(w.W1t + (long)n * KIN + 8 * c);
    }
    for (int k = threadIdx.x; k < H * (H / 4); k += blockDim.x) {
        const int n = k / (H / 4), c = 4 * (k - n * (H / 4));
        *reinterpret_cast<uint2*>(W2s + n * kW2Pitch + perm_col(c)) = *reinterpret_cast<const uint2*>(w.W2t + (long)n * H + c);
    }
    for (int k = threadIdx.x; k < NO * H; k += blockDim.x) {
        const int o = k / H, n = k - o * H;
        W3s[o * H + perm_col(n)] = w.W3[n * NO + o];
    }
    for (int k = threadIdx.x; k < H; k += blockDim.x) {
        b1s[k] = w.b1[k];
        b2s[k] = w.b2[k];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    const bf16x8 zero8 = __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0));
    const float4 b3 = make_float4(w.b3[0], w.b3[1], w.b3[2], w.b3[3]);
    const long ngroups = ((long)M + 16 * kInferT - 1) / (16 * kInferT);
    for (long gi = (long)blockIdx.x * kInferWaves + wave; gi < ngroups; gi += (long)gridDim.x * kInferWaves) {
        const long m0 = gi * 16 * kInferT;
        // an offset the compiler cannot see through, so the weight fragments are read from
        // LDS on every pass instead of being hoisted out of the loop (336 registers)
        int wo = 0;
        asm volatile("" : "+v"(wo));
        // layer 1, k chunk by k chunk: B = trunk columns of the tile rows (rows past M repeat
        // the last one and are never stored), all 8 column tiles accumulated at once
        f32x4 acc1[NT][kInferT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const f32x4 bias = *reinterpret_cast<const f32x4*>(b1s + 16 * nt + 4 * lq);
#pragma unroll
            for (int t = 0; t < kInferT; ++t) acc1[nt][t] = bias;
        }
#pragma unroll
        for (int kc = 0; kc < KC1; ++kc) {
            const int col = kc * 32 + 8 * lq;
            bf16x8 xb[kInferT];
#pragma unroll
            for (int t = 0; t < kInferT; ++t) {
                const long row = min(m0 + 16 * t + lr, (long)M - 1);
                xb[t] = col < TRUNK ? ld8(X + row * KIN + col) : zero8;
            }
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const bf16x8 a = col < TRUNK ? ld8(W1s + wo + (16 * nt + lr) * kW1Pitch + col) : zero8;
#pragma unroll
                for (int t = 0; t < kInferT; ++t) acc1[nt][t] = mfma32(a, xb[t], acc1[nt][t]);
            }
        }
        // h1 in B-operand form
        bf16x8 h1[kInferT][KC2];
#pragma unroll
        for (int kc = 0; kc < KC2; ++kc)
#pragma unroll
            for (int t = 0; t < kInferT; ++t) h1[t][kc] = act8(acc1[2 * kc][t], acc1[2 * kc + 1][t]);
        // layer 2 -> h2 chunks, each fed at once to layer 3
        f32x4 acc3[kInferT];
#pragma unroll
        for (int t = 0; t < kInferT; ++t) acc3[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int nt = 0; nt < NT; nt += 2) {
            f32x4 acc[2][kInferT];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const f32x4 bias = *reinterpret_cast<const f32x4*>(b2s + 16 * (nt + u) + 4 * lq);
#pragma unroll
                for (int t = 0; t < kInferT; ++t) acc[u][t] = bias;
#pragma unroll
                for (int kc = 0; kc < KC2; ++kc) {
                    const bf16x8 a = ld8(W2s + wo + (16 * (nt + u) + lr) * kW2Pitch + kc * 32 + 8 * lq);
#pragma unroll
                    for (int t = 0; t < kInferT; ++t) acc[u][t] = mfma32(a, h1[t][kc], acc[u][t]);
                }
            }
            const bf16x8 a3 = lr < NO ? ld8(W3s + wo + lr * H + (nt / 2) * 32 + 8 * lq) : zero8;
#pragma unroll
            for (int t = 0; t < kInferT; ++t) acc3[t] = mfma32(a3, act8(acc[0][t], acc[1][t]), acc3[t]);
        }
        // heads of row m0 + 16 t + lr in lanes 0..15: y = g . (u + b3)
        if (lq == 0) {
#pragma unroll
            for (int t = 0; t < kInferT; ++t) {
                const long m = m0 + 16 * t + lr;
                if (m < M) {
                    float g[NO];
                    row_gate(X + m * KIN, g);
                    y[m] = ((g[0] * (acc3[t][0] + b3.x) + g[1] * (acc3[t][1] + b3.y)) + g[2] * (acc3[t][2] + b3.z)) +
                           g[3] * (acc3[t][3] + b3.w);
                }
            }
        }
    }
}

// ---- backward, per row ------------------------------------------------------
// loss = mean over the batch of (y - t)^2.  dy = 2 (y - t) / M; du = dy g;
// dz2 = (W3 du) silu'(z2); dh1 = dz2 W2^T; dz1 = dh1 silu'(z1).  Stores du (FP32) and
// bf16 dz1, dz2.
__global__ __launch_bounds__(64 * kWaves) void k_mlp_backward_rows(int M, const unsigned short* __restrict__ X,
                                                                   const float* __restrict__ y,
                                                                   const float* __restrict__ t, MlpDevWeights w,
                                                                   const float* __restrict__ z1g,
                                                                   const float* __restrict__ z2g,
                                                                   float* __restrict__ dug,
                                                                   unsigned short* __restrict__ dz1g,
                                                                   unsigned short* __restrict__ dz2g)
{
    __shared__ unsigned short ds[kWaves][16][H + 8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = (blockIdx.x * kWaves + wave) * 16;
    if (m0 >= M) return;
    const int lr = lane & 15, lq = lane >> 4;
    const float scale = 2.0f / (float)M;
    float du[4][NO];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int m = m0 + 4 * lq + r;
        float g[NO] = {0.f, 0.f, 0.f, 0.f};
        float dy = 0.f;
        if (m < M) {
            row_gate(X + (long)m * KIN, g);
            dy = scale * (y[m] - t[m]);
        }
#pragma unroll
        for (int o = 0; o < NO; ++o) {
            du[r][o] = dy * g[o];
            if (lr == 0 && m < M) dug[(long)m * NO + o] = du[r][o];
        }
    }
    // dz2 in the MFMA output layout: rows 4 lq + r, column n = nt*16 + lr
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int n = nt * 16 + lr;
        float w3[NO];
#pragma unroll
        for (int o = 0; o < NO; ++o) w3[o] = bf2f(w.W3[n * NO + o]);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 4 * lq + r;
            const int m = m0 + row;
            const float dh = ((du[r][0] * w3[0] + du[r][1] * w3[1]) + du[r][2] * w3[2]) + du[r][3] * w3[3];
            const float g = (m < M) ? dh * dsilu(z2g[(long)m * H + n]) : 0.0f;
            const unsigned short gb = f2bf(g);
            ds[wave][row][n] = gb;
            if (m < M) dz2g[(long)m * H + n] = gb;
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    bf16x4 a[NT];
#pragma unroll
    for (int kt = 0; kt < NT; ++kt) {
        bf16x4 v;
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = (short)ds[wave][lr][kt * 16 + 4 * lq + j];
        a[kt] = v;
    }
    // dh1[m][k] = sum_n dz2[m][n] W2[k][n]: B[n][k] = W2[k][n], read from W2 stored [k][n]
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < NT; ++kt)
            acc = mfma(a[kt], load4(w.W2n + (long)(ct * 16 + lr) * H + kt * 16 + 4 * lq), acc);
        const int k = ct * 16 + lr;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int m = m0 + 4 * lq + r;
            if (m < M) dz1g[(long)m * H + k] = f2bf(acc[r] * dsilu(z1g[(long)m * H + k]));
        }
    }
}

// ---- weight gradients -------------------------------------------------------
// part[s][k][n] = sum over rows m of slice s of A[m][k] D[m][n] (A: M rows of lda bf16,
// the first KA columns used; D: M x H bf16).  One wavefront per (16 x 16 output tile,
// batch slice); A columns >= kmax read as zero (the gate columns of the features).
__global__ __launch_bounds__(64) void k_mlp_wgrad(int M, int KA, int lda, int kmax, const unsigned short* __restrict__ A,
                                                  const unsigned short* __restrict__ D, int slices,
                                                  float* __restrict__ part)
{
    const int lane = threadIdx.x, lr = lane & 15, lq = lane >> 4;
    const int ntn = H / 16;
    const int tile = blockIdx.x, s = blockIdx.y;
    const int k0 = (tile / ntn) * 16, n0 = (tile % ntn) * 16;
    const int per = ((M + slices - 1) / slices + 15) / 16 * 16;
    const int mb = s * per, me = min(M, mb + per);
    const bool kin = k0 + lr < kmax;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int m = mb; m < me; m += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int mm = m + 4 * lq + j;
            const bool ok = mm < me;
            a[j] = (ok && kin) ? (short)A[(long)mm * lda + k0 + lr] : (short)0;
            b[j] = ok ? (short)D[(long)mm * H + n0 + lr] : (short)0;
        }
        acc = mfma(a, b, acc);  // D[k][n]: row k = k0 + 4 lq + r, column n0 + lr
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) part[((long)s * KA + k0 + 4 * lq + r) * H + n0 + lr] = acc[r];
}

// grad[k][n] = sum_s part[s][k][n] (fixed order)
__global__ void k_mlp_reduce_slices(int KA, int slices, const float* __restrict__ part, float* __restrict__ grad)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= KA * H) return;
    float v = 0.f;
    for (int s = 0; s < slices; ++s) v += part[(long)s * KA * H + e];
    grad[e] = v;
}

// Bias and output-layer gradients, one workgroup per output element, fixed-order tree:
// blocks [0, H) db1 = sum dz1; [H, 2H) db2 = sum dz2; [2H, 2H + H NO) dW3[n][o] =
// sum h2[n] du[o]; [2H + H NO, 2H + H NO + NO) db3[o] = sum du[o].
__global__ __launch_bounds__(256) void k_mlp_small_grads(int M, const unsigned short* __restrict__ dz1,
                                                         const unsigned short* __restrict__ dz2,
                                                         const unsigned short* __restrict__ h2,
                                                         const float* __restrict__ du, float* __restrict__ db1,
                                                         float* __restrict__ db2, float* __restrict__ dW3,
                                                         float* __restrict__ db3)
{
    const int blk = blockIdx.x;
    int which, n, o = 0;
    if (blk < H) {
        which = 0;
        n = blk;
    } else if (blk < 2 * H) {
        which = 1;
        n = blk - H;
    } else if (blk < 2 * H + H * NO) {
        which = 2;
        n = (blk - 2 * H) / NO;
        o = (blk - 2 * H) % NO;
    } else {
        which = 3;
        n = 0;
        o = blk - 2 * H - H * NO;
    }
    float v = 0.f;
    for (int m = threadIdx.x; m < M; m += blockDim.x) {
        if (which == 0) v += bf2f(dz1[(long)m * H + n]);
        else if (which == 1) v += bf2f(dz2[(long)m * H + n]);
        else if (which == 2) v += bf2f(h2[(long)m * H + n]) * du[(long)m * NO + o];
        else v += du[(long)m * NO + o];
    }
    __shared__ float red[256];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (which == 0) db1[n] = red[0];
        else if (which == 1) db2[n] = red[0];
        else if (which == 2) dW3[n * NO + o] = red[0];
        else db3[o] = red[0];
    }
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters.
__global__ void k_mlp_adam(int P, float* __restrict__ theta, const float* __restrict__ g, float* __restrict__ m1,
                           float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    theta[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// bf16 compute copies of the master weights: W1t [n][k], W2t [n][k], W2n [k][n], W3 [n][o].
__global__ void k_mlp_pack(const float* __restrict__ theta, unsigned short* __restrict__ W1t,
                           unsigned short* __restrict__ W2t, unsigned short* __restrict__ W2n,
                           unsigned short* __restrict__ W3)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const float* W1 = theta + kMlpOffW1;
    const float* W2 = theta + kMlpOffW2;
    if (e < KIN * H) {
        const int k = e / H, n = e % H;
        W1t[(long)n * KIN + k] = f2bf(W1[e]);
    }
    if (e < H * H) {
        const int k = e / H, n = e % H;
        const unsigned short v = f2bf(W2[e]);
        W2t[(long)n * H + k] = v;
        W2n[e] = v;
    }
    if (e < H * NO) W3[e] = f2bf(theta[kMlpOffW3 + e]);
}

// t_i = (q_i - Q_b / n_b) / kMlpChargeUnit (to_charges = 0) or
// q_i = Q_b / n_b + kMlpChargeUnit y_i (to_charges = 1)
__global__ void k_mlp_share(SysView s, const unsigned* __restrict__ counts, const double* __restrict__ q,
                            const float* __restrict__ y, float* __restrict__ t, double* __restrict__ q_out,
                            int to_charges)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.N) return;
    const unsigned b = s.iso[i];
    const double share = s.body_charge[b] / (double)(counts[b] ? counts[b] : 1u);
    if (to_charges) q_out[i] = share + kMlpChargeUnit * (double)y[i];
    else t[i] = (float)((q[i] - share) / kMlpChargeUnit);
}

inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b > 0 ? (n + b - 1) / b : 1); }

}  // namespace

hipError_t launch_mlp_features(const SysView& s, const double L[3], const double E0[3], const unsigned* body_count,
                               unsigned short* X, hipStream_t st)
{
    if (s.N <= 0) return hipSuccess;
    static_assert(KIN % 4 == 0 && kMlpGate >= 5 * kMlpNeigh, "feature rows leave as two halves of dwords");
    if (reinterpret_cast<uintptr_t>(X) & 3u) return hipErrorInvalidValue;  // rows leave as dword stores
    hipLaunchKernelGGL(k_mlp_features, dim3(nblk(s.N, kFeatBlock)), dim3(kFeatBlock), 0, st, s, L[0], L[1], L[2],
                       E0[0], E0[1], E0[2], body_count, X);
    return hipGetLastError();
}

hipError_t launch_mlp_share(const SysView& s, const unsigned* counts, const double* q, const float* y, float* t,
                            double* q_out, int to_charges, hipStream_t st)
{
    if (s.N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_share, dim3(nblk(s.N, 256)), dim3(256), 0, st, s, counts, q, y, t, q_out, to_charges);
    return hipGetLastError();
}

hipError_t launch_mlp_pack(const float* theta, const MlpDevWeights& w, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_pack, dim3(nblk((long)KIN * H, 256)), dim3(256), 0, st, theta,
                       const_cast<unsigned short*>(w.W1t), const_cast<unsigned short*>(w.W2t),
                       const_cast<unsigned short*>(w.W2n), const_cast<unsigned short*>(w.W3));
    return hipGetLastError();
}

// Persistent inference: one workgroup of kInferWaves wavefronts per CU (3 per SIMD at the
// kernel's 138 registers; its LDS holds the weights), capped by the row tiles there are.
hipError_t launch_mlp_infer(int M, const unsigned short* X, const MlpDevWeights& w, float* y, hipStream_t st)
{
    if (M <= 0) return hipSuccess;
    static int cus[64] = {};
    static bool lds_set[64] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    if (!cus[dev]) {
        e = hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
    }
    if (!lds_set[dev]) {
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mlp_infer), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)kInferLds);
        if (e != hipSuccess) return e;
        lds_set[dev] = true;
    }
    const long groups = ((long)M + 16 * kInferT * kInferWaves - 1) / (16 * kInferT * kInferWaves);
    const unsigned blocks = (unsigned)std::max(1L, std::min<long>(groups, (long)cus[dev]));
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(64 * kInferWaves), kInferLds, st, M, X, w, y);
    return hipGetLastError();
}

hipError_t launch_mlp_forward(int M, const unsigned short* X, const MlpDevWeights& w, float* y, const MlpTrainBufs* tb,
                              hipStream_t st)
{
    if (M <= 0) return hipSuccess;
    if (tb)
        hipLaunchKernelGGL((k_mlp_forward<true, kFwdTilesTrain>), dim3(nblk(M, 16 * kFwdTilesTrain * kWaves)),
                           dim3(64 * kWaves), 0, st, M, X, w, y, tb->z1, tb->z2, tb->h1, tb->h2);
    else
        return launch_mlp_infer(M, X, w, y, st);
    return hipGetLastError();
}

hipError_t launch_mlp_gradient(int M, const unsigned short* X, const float* t, const MlpDevWeights& w,
                               const MlpTrainBufs& tb, float* y, float* grad, hipStream_t st)
{
    if (M <= 0) return hipErrorInvalidValue;
    hipError_t e = launch_mlp_forward(M, X, w, y, &tb, st);
    if (e != hipSuccess) return e;
    const unsigned blocks = nblk(M, 16 * kWaves);
    hipLaunchKernelGGL(k_mlp_backward_rows, dim3(blocks), dim3(64 * kWaves), 0, st, M, X, (const float*)y, t, w,
                       (const float*)tb.z1, (const float*)tb.z2, tb.du, tb.dz1, tb.dz2);
    const int slices = tb.slices;
    hipLaunchKernelGGL(k_mlp_wgrad, dim3((KIN / 16) * (H / 16), slices), dim3(64), 0, st, M, KIN, KIN, TRUNK, X,
                       (const unsigned short*)tb.dz1, slices, tb.part);
    hipLaunchKernelGGL(k_mlp_reduce_slices, dim3(nblk((long)KIN * H, 256)), dim3(256), 0, st, KIN, slices,
                       (const float*)tb.part, grad + kMlpOffW1);
    hipLaunchKernelGGL(k_mlp_wgrad, dim3((H / 16) * (H / 16), slices), dim3(64), 0, st, M, H, H, H,
                       (const unsigned short*)tb.h1, (const unsigned short*)tb.dz2, slices, tb.part);
    hipLaunchKernelGGL(k_mlp_reduce_slices, dim3(nblk((long)H * H, 256)), dim3(256), 0, st, H, slices,
                       (const float*)tb.part, grad + kMlpOffW2);
    hipLaunchKernelGGL(k_mlp_small_grads, dim3(2 * H + H * NO + NO), dim3(256), 0, st, M,
                       (const unsigned short*)tb.dz1, (const unsigned short*)tb.dz2, (const unsigned short*)tb.h2,
                       (const float*)tb.du, grad + kMlpOffB1, grad + kMlpOffB2, grad + kMlpOffW3, grad + kMlpOffB3);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float* theta, const float* g, float* m1, float* m2, float lr, int step, hipStream_t st)
{
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
    hipLaunchKernelGGL(k_mlp_adam, dim3(nblk(kMlpParams, 256)), dim3(256), 0, st, kMlpParams, theta, g, m1, m2, lr,
                       b1, b2, eps, c1, c2);
    return hipGetLastError();
}

}  // namespace cr
