// gk_mlp.hpp -- the flux-correction MLP (16 -> 64 -> 64 -> 4, ReLU) on CDNA4 matrix cores.
//
// No counterpart in the reference.  BF16 operands, FP32 accumulation on the matrix cores.  The
// network runs in transposed form, H^T = W X^T, 16 samples per tile.  With v_mfma_f32_16x16x16_bf16:
//   A operand (16 x 16, rows x k): lane l holds A[l % 16][4 (l / 16) + i], i = 0..3
//   B operand (16 x 16, k x cols): lane l holds B[4 (l / 16) + i][l % 16]
//   D result  (16 x 16):           lane l holds D[4 (l / 16) + i][l % 16]
// so the D tile of one layer (rows = hidden units, columns = samples) is already the B operand of
// the next: no data moves between layers.  The contractions over 64 hidden units (layers 2 and 3
// of the forward pass) use gfx950's v_mfma_f32_16x16x32_bf16, half the instructions: a lane's 8
// operand values (slot i of lane group g pairs A with B) are the 4 values it holds of two
// consecutive 16-unit blocks 2c and 2c+1, so chunk c sums units 16(2c) + 4g + i and
// 16(2c+1) + 4g + i over g and i -- the same 32 units in another order on both operands, built
// from registers each lane already holds.  Weights are the A operands, read once per wave.  The
// weight gradients contract over samples, which needs each tile transposed (through LDS).
//
// Parameter vector (FP32 master copy, 5508 values): W1[64][16], b1[64], W2[64][64], b2[64],
// W3[4][64], b3[4], row-major (W[out][in]).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gk {
namespace mlp {

constexpr int kIn = 16, kHid = 64, kOut = 4;
constexpr int oW1 = 0, ob1 = 1024, oW2 = 1088, ob2 = 5184, oW3 = 5248, ob3 = 5504, kParams = 5508;
// LDS row pitches in BF16 elements: rows 40 / 136 bytes apart, so the 16 rows an A-operand read
// touches fall in distinct bank pairs.
constexpr int kP16 = 20;
constexpr int kP64 = 68;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16v8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16, round to nearest, ties to even (finite inputs).
__host__ __device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned u = __builtin_bit_cast(unsigned, f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __builtin_bit_cast(float, (unsigned)h << 16); }

__device__ __forceinline__ bf16x4 pack4(f32x4 v) {
    bf16x4 r;
    r[0] = (short)f2bf(v[0]); r[1] = (short)f2bf(v[1]); r[2] = (short)f2bf(v[2]); r[3] = (short)f2bf(v[3]);
    return r;
}
__device__ __forceinline__ f32x4 relu(f32x4 v) {
    f32x4 r;
#pragma unroll
    for (int i = 0; i < 4; i++) r[i] = v[i] > 0.f ? v[i] : 0.f;
    return r;
}
__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// K = 32: the operands of two consecutive 16-wide blocks, concatenated per lane
__device__ __forceinline__ bf16x8 cat(bf16x4 lo, bf16x4 hi) {
    return bf16x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16v8, a), __builtin_bit_cast(bf16v8, b), c,
                                                   0, 0, 0);
}
__device__ __forceinline__ bf16x4 lds4(const unsigned short* p) {
    return *reinterpret_cast<const bf16x4*>(p);
}

// Network weights in LDS: BF16 matrices (padded rows), FP32 biases and the fixed input/output
// scales (network input = feature * in_scale; correction = output * out_scale).
// The same struct in device memory is the prebuilt image a kernel copies into LDS.
struct alignas(16) Smem {
    unsigned short w1[kHid * kP16];
    unsigned short w2[kHid * kP64];
    unsigned short w3[kOut * kP64];
    float b1[kHid], b2[kHid], b3[kOut];
    float in_scale[kIn], out_scale[kOut];
};
static_assert(sizeof(Smem) % 16 == 0, "the image is copied in 16-byte pieces");
constexpr int kImageVec = (int)(sizeof(Smem) / 16);

// A prebuilt image is copied into LDS in 16-byte pieces, NPER per thread of an NT-thread block:
// load_image issues the loads, store_image (later, so other work hides the latency) writes LDS.
template <int NT>
struct ImageCopy {
    static constexpr int kPer = (kImageVec + NT - 1) / NT;
    uint4 v[kPer];
    // Unconditional (clamped) loads: predicated ones are paired with their stores by the
    // compiler and each waits out a full memory latency.  Four 4-byte loads per piece: measured
    // on the corrected stage kernel at 256^2, 16-byte loads of the image took the load phase
    // from 1.6 us to 8.5 us.
    __device__ __forceinline__ void load(const Smem* img, int tid) {
        const uint4* src = reinterpret_cast<const uint4*>(img);
#pragma unroll
        for (int q = 0; q < kPer; q++) {
            const unsigned* w = reinterpret_cast<const unsigned*>(src + min(tid + q * NT, kImageVec - 1));
            v[q].x = w[0];
            v[q].y = w[1];
            v[q].z = w[2];
            v[q].w = w[3];
        }
    }
    __device__ __forceinline__ void store(Smem& s, int tid) const {
        uint4* dst = reinterpret_cast<uint4*>(&s);
#pragma unroll
        for (int q = 0; q < kPer; q++) {
            const int k = tid + q * NT;
            if (k < kImageVec) dst[k] = v[q];
        }
    }
};
// Backward operands: W2^T and W3^T (rows padded to 16 outputs with zeros).
struct SmemT {
    unsigned short w2t[kHid * kP64];
    unsigned short w3t[kHid * kP16];
};

__device__ __forceinline__ void stage(Smem& s, const float* p, const float* in_scale, const float* out_scale,
                                      int tid, int nt) {
    for (int k = tid; k < kHid * kIn; k += nt) s.w1[(k / kIn) * kP16 + k % kIn] = f2bf(p[oW1 + k]);
    for (int k = tid; k < kHid * kHid; k += nt) s.w2[(k / kHid) * kP64 + k % kHid] = f2bf(p[oW2 + k]);
    for (int k = tid; k < kOut * kHid; k += nt) s.w3[(k / kHid) * kP64 + k % kHid] = f2bf(p[oW3 + k]);
    for (int k = tid; k < kHid; k += nt) { s.b1[k] = p[ob1 + k]; s.b2[k] = p[ob2 + k]; }
    for (int k = tid; k < kOut; k += nt) { s.b3[k] = p[ob3 + k]; s.out_scale[k] = out_scale[k]; }
    for (int k = tid; k < kIn; k += nt) s.in_scale[k] = in_scale[k];
}
__device__ __forceinline__ void stage_t(SmemT& s, const float* p, int tid, int nt) {
    for (int k = tid; k < kHid * kHid; k += nt) {
        const int o = k / kHid, i = k % kHid;   // W2[o][i] -> w2t[i][o]
        s.w2t[i * kP64 + o] = f2bf(p[oW2 + k]);
    }
    for (int k = tid; k < kHid * 16; k += nt) {
        const int h = k / 16, o = k % 16;       // W3[o][h] -> w3t[h][o], o >= 4 zero
        s.w3t[h * kP16 + o] = o < kOut ? f2bf(p[oW3 + o * kHid + h]) : (unsigned short)0;
    }
}

// A operands and biases of one wave's lane, loaded once; layers 2 and 3 in K = 32 chunks c.
struct Regs {
    bf16x4 w1[4];      // [rt]: W1[16 rt + r][4 g .. 4 g + 3]
    bf16x8 w2[4][2];   // [ro][c]: W2[16 ro + r][32 c + 4 g .. + 3], then W2[16 ro + r][32 c + 16 + 4 g .. + 3]
    bf16x8 w3[2];      // [c]: r < 4 ? (W3[r][32 c + 4 g .. + 3], W3[r][32 c + 16 + 4 g .. + 3]) : 0
    f32x4 b1[4], b2[4], b3;
};

__device__ __forceinline__ void load_regs(Regs& w, const Smem& s, int lane) {
    const int r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        w.w1[t] = lds4(&s.w1[(16 * t + r) * kP16 + 4 * g]);
#pragma unroll
        for (int c = 0; c < 2; c++)
            w.w2[t][c] = cat(lds4(&s.w2[(16 * t + r) * kP64 + 32 * c + 4 * g]),
                             lds4(&s.w2[(16 * t + r) * kP64 + 32 * c + 16 + 4 * g]));
#pragma unroll
        for (int i = 0; i < 4; i++) {
            w.b1[t][i] = s.b1[16 * t + 4 * g + i];
            w.b2[t][i] = s.b2[16 * t + 4 * g + i];
        }
    }
    const bf16x4 z = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 2; c++)
        w.w3[c] = r < kOut ? cat(lds4(&s.w3[r * kP64 + 32 * c + 4 * g]), lds4(&s.w3[r * kP64 + 32 * c + 16 + 4 * g]))
                           : cat(z, z);
#pragma unroll
    for (int i = 0; i < 4; i++) w.b3[i] = g == 0 ? s.b3[i] : 0.f;
}

// Forward pass of one 16-sample tile.  x: lane (g, s) holds the scaled inputs 4g..4g+3 of sample
// s as BF16.  Returns the D tile of the output layer: lanes 0..15 hold outputs 0..3 of sample
// `lane`; other lanes hold zeros.
// The same with the pre-activations and activations kept (training needs the ReLU masks and the
// operands of the weight gradients).
__device__ __forceinline__ f32x4 forward_tile_keep(const Regs& w, bf16x4 x, f32x4 z1[4], f32x4 z2[4], bf16x4 h1[4],
                                                   bf16x4 h2[4]) {
#pragma unroll
    for (int t = 0; t < 4; t++) { z1[t] = mfma(w.w1[t], x, w.b1[t]); h1[t] = pack4(relu(z1[t])); }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        f32x4 z = w.b2[t];
#pragma unroll
        for (int c = 0; c < 2; c++) z = mfma32(w.w2[t][c], cat(h1[2 * c], h1[2 * c + 1]), z);
        z2[t] = z;
        h2[t] = pack4(relu(z));
    }
    f32x4 y = w.b3;
#pragma unroll
    for (int c = 0; c < 2; c++) y = mfma32(w.w3[c], cat(h2[2 * c], h2[2 * c + 1]), y);
    return y;
}
__device__ __forceinline__ f32x4 forward_tile(const Regs& w, bf16x4 x) {
    f32x4 z1[4], z2[4];
    bf16x4 h1[4], h2[4];
    return forward_tile_keep(w, x, z1, z2, h1, h2);
}

}  // namespace mlp
}  // namespace gk
