// Device side of the learned WCNS sensor: parameter layout, features,
// training label and the wavefront-wide BF16 MFMA evaluation used inside
// the WCNS kernels.  No counterpart in the reference.
#pragma once
#include "common.hpp"

namespace acfd {

constexpr int kSensorNF = 12, kSensorH = 32;
constexpr int kSensorOffW1 = 0;
constexpr int kSensorOffB1 = kSensorOffW1 + kSensorNF * kSensorH;
constexpr int kSensorOffW2 = kSensorOffB1 + kSensorH;
constexpr int kSensorOffB2 = kSensorOffW2 + kSensorH * kSensorH;
constexpr int kSensorOffW3 = kSensorOffB2 + kSensorH;
constexpr int kSensorOffB3 = kSensorOffW3 + kSensorH;
constexpr int kSensorNParam = kSensorOffB3 + 1;  // 1505
constexpr double kSensorLabelBound = 1.0e-3;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// inference copy of the parameters (device memory)
struct SensorNet {
    __bf16 w1[16 * kSensorH];        // [k][j], rows 12..15 zero
    __bf16 w2[kSensorH * kSensorH];  // [c][j]
    float b1[kSensorH], b2[kSensorH], w3[kSensorH], b3;
};

// LDS of one workgroup (up to 4 wavefronts): the weights transposed for
// the MFMA B operand, and per-wavefront activations
struct SensorLds {
    alignas(16) __bf16 w1t[kSensorH][16];        // [j][k]
    alignas(16) __bf16 w2t[kSensorH][kSensorH];  // [j][c]
    alignas(16) float b1[kSensorH];
    float b2[kSensorH], w3[kSensorH], b3;
    struct alignas(16) Wave {
        alignas(16) __bf16 h1[64][kSensorH];
        union alignas(16) {
            __bf16 xs[64][16];
            float h2[64][kSensorH + 1];
        };
    } wv[4];
};

// all threads of the block; a barrier must follow
__device__ inline void sensor_load_net(const SensorNet *__restrict__ net, SensorLds &L)
{
    const int t = threadIdx.x + blockDim.x * threadIdx.y;
    const int nt = blockDim.x * blockDim.y;
    for (int e = t; e < 16 * kSensorH; e += nt) L.w1t[e % kSensorH][e / kSensorH] = net->w1[e];
    for (int e = t; e < kSensorH * kSensorH; e += nt) L.w2t[e % kSensorH][e / kSensorH] = net->w2[e];
    for (int e = t; e < kSensorH; e += nt) {
        L.b1[e] = net->b1[e];
        L.b2[e] = net->b2[e];
        L.w3[e] = net->w3[e];
    }
    if (t == 0) L.b3 = net->b3;
}

__device__ inline void wave_lds_sync()
{
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_wave_barrier();
}

// Probability that face of this lane needs the nonlinear weights.  Every
// lane of the wavefront must call it (MFMA is wavefront-wide); `w` is the
// wavefront's index in the workgroup.  MFMA v_mfma_f32_32x32x16_bf16
// operand layout: lane l holds A[l % 32][8 (l / 32) + t] and
// B[8 (l / 32) + t][l % 32], t = 0..7; result register r of lane l is
// D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
__device__ inline float sensor_eval_wave(SensorLds &L, int w, const float *f)
{
    const int l = __lane_id();
    SensorLds::Wave &V = L.wv[w];
    const int col = l & 31, half = l >> 5;
#pragma unroll
    for (int c = 0; c < 16; c++) V.xs[l][c] = (__bf16)(c < kSensorNF ? f[c] : 0.f);
    wave_lds_sync();
    // layer 1: [64 x 16] x [16 x 32]
#pragma unroll
    for (int T = 0; T < 2; T++) {
        bf16x8 a = *reinterpret_cast<const bf16x8 *>(&V.xs[T * 32 + col][8 * half]);
        bf16x8 b = *reinterpret_cast<const bf16x8 *>(&L.w1t[col][8 * half]);
        floatx16 acc = {};
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int row = 8 * (r >> 2) + 4 * half + (r & 3);
            V.h1[T * 32 + row][col] = (__bf16)fmaxf(acc[r] + L.b1[col], 0.f);
        }
    }
    wave_lds_sync();
    // layer 2: [64 x 32] x [32 x 32]
    floatx16 acc2[2];
#pragma unroll
    for (int T = 0; T < 2; T++) {
        floatx16 acc = {};
#pragma unroll
        for (int kc = 0; kc < 2; kc++) {
            bf16x8 a = *reinterpret_cast<const bf16x8 *>(&V.h1[T * 32 + col][16 * kc + 8 * half]);
            bf16x8 b = *reinterpret_cast<const bf16x8 *>(&L.w2t[col][16 * kc + 8 * half]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        acc2[T] = acc;
    }
    wave_lds_sync();  // every lane is done reading xs (h2 shares its storage)
#pragma unroll
    for (int T = 0; T < 2; T++)
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int row = 8 * (r >> 2) + 4 * half + (r & 3);
            V.h2[T * 32 + row][col] = fmaxf(acc2[T][r] + L.b2[col], 0.f);
        }
    wave_lds_sync();
    // layer 3: per lane
    float z = L.b3;
#pragma unroll
    for (int j = 0; j < kSensorH; j++) z = fmaf(V.h2[l][j], L.w3[j], z);
    wave_lds_sync();  // the wavefront's buffers are free again on return
    return 1.f / (1.f + __expf(-z));
}

// Features of the face a+1/2 from the conservative variables at a-2..a+3.
template <typename R>
__device__ inline void sensor_features(const R (*W)[5], double gamma, float *f)
{
    float v[2][6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const float rho = (float)W[k][0];
        const float m2 = (float)W[k][1] * (float)W[k][1] + (float)W[k][2] * (float)W[k][2] +
                         (float)W[k][3] * (float)W[k][3];
        v[0][k] = rho;
        v[1][k] = ((float)gamma - 1.f) * ((float)W[k][4] - 0.5f * m2 / rho);
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
        float b[4], S = 0.f;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float c2 = v[q][k] - 2.f * v[q][k + 1] + v[q][k + 2];
            const float c1 = v[q][k] - v[q][k + 2];
            b[k] = 13.f / 12.f * c2 * c2 + 0.25f * c1 * c1;
            S += b[k];
        }
        const float rs = 1.f / (S + 1e-30f);
#pragma unroll
        for (int k = 0; k < 4; k++) f[6 * q + k] = b[k] * rs;
        f[6 * q + 4] = fabsf(v[q][3] - v[q][2]) / (fabsf(v[q][2]) + fabsf(v[q][3]) + 1e-30f);
        f[6 * q + 5] = (log10f(S + 1e-30f) + 12.f) / 10.f;
    }
}

// max over the ten interpolations of the face of max_k |w_k - d_k|
__device__ inline double wcns_weight_deviation(const double (*W)[5])
{
    const double d[3] = {1.0 / 16.0, 10.0 / 16.0, 5.0 / 16.0};
    double dev = 0.0;
#pragma unroll
    for (int x = 0; x < 5; x++)
#pragma unroll
        for (int side = 0; side < 2; side++) {
            double v[5];
#pragma unroll
            for (int o = 0; o < 5; o++) v[o] = side == 0 ? W[o][x] : W[5 - o][x];
            const double k = 13.0 / 12.0;
            const double b0 = k * (v[0] - 2 * v[1] + v[2]) * (v[0] - 2 * v[1] + v[2]) +
                              0.25 * (v[0] - 4 * v[1] + 3 * v[2]) * (v[0] - 4 * v[1] + 3 * v[2]);
            const double b1 = k * (v[1] - 2 * v[2] + v[3]) * (v[1] - 2 * v[2] + v[3]) + 0.25 * (v[1] - v[3]) * (v[1] - v[3]);
            const double b2 = k * (v[2] - 2 * v[3] + v[4]) * (v[2] - 2 * v[3] + v[4]) +
                              0.25 * (3 * v[2] - 4 * v[3] + v[4]) * (3 * v[2] - 4 * v[3] + v[4]);
            const double a0 = d[0] / ((1e-6 + b0) * (1e-6 + b0)), a1 = d[1] / ((1e-6 + b1) * (1e-6 + b1)),
                         a2 = d[2] / ((1e-6 + b2) * (1e-6 + b2));
            const double s = a0 + a1 + a2;
            dev = fmax(dev, fmax(fabs(a0 / s - d[0]), fmax(fabs(a1 / s - d[1]), fabs(a2 / s - d[2]))));
        }
    return dev;
}

}  // namespace acfd
