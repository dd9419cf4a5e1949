// Learned WCNS sensor (no counterpart in the reference).
//
// Per interface a+1/2 of a grid line a small network reads 12 features of
// density and pressure at the six stencil nodes a-2..a+3 and predicts
// whether the WENO5 nonlinear weights of the WCNS midpoint interpolation
// differ from the linear ones by more than 1e-3 (the training label).  Where
// it predicts "smooth" the reconstruction uses the linear weights and skips
// the smoothness indicators and the weight division.
//
//   features  per quantity (rho, p): the four 3-point smoothness indicators
//             beta_k = 13/12 (v_k - 2 v_k+1 + v_k+2)^2 + 1/4 (v_k - v_k+2)^2
//             normalised by their sum S, the face jump |v3 - v2|/(|v2|+|v3|)
//             and (log10 S + 12)/10.
//   network   12 -> 32 (ReLU) -> 32 (ReLU) -> 1 (sigmoid); parameters W1[12][32],
//             b1[32], W2[32][32], b2[32], w3[32], b3 (1505 floats).
//   training  hand-written FP32 kernels: per-block forward/backward with the
//             per-sample activations in LDS and a deterministic reduction of
//             the gradient, then Adam; binary cross-entropy.
//   inference BF16 MFMA (v_mfma_f32_32x32x16_bf16) with FP32 accumulate, one
//             wavefront = 64 interfaces = two 32-row tiles, inside the WCNS
//             kernels; the last layer is a per-lane FP32 dot product.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"
#include "sensor.hpp"

namespace acfd {

// ----------------------------------------------------------- training
// Per block of kMlpB samples: forward, loss, backward; the block's gradient
// (sum over its samples) goes to part[block][1505], its loss sum to
// lpart[block].  Samples are taken in a fixed pseudo-random order:
// index(k) = (k * stride + offset) mod n with stride coprime to n.
constexpr int kMlpB = 128;

__global__ void __launch_bounds__(kMlpB) k_mlp_grad(const float *__restrict__ w, const float *__restrict__ X,
                                                    const float *__restrict__ Y, long n, long first, long count,
                                                    long stride, long offset, float inv_count,
                                                    float *__restrict__ part, float *__restrict__ lpart)
{
    __shared__ float a0[kMlpB][kSensorNF];
    __shared__ float a1[kMlpB][kSensorH + 1];
    __shared__ float a2[kMlpB][kSensorH + 1];
    __shared__ float d1[kMlpB][kSensorH + 1];
    __shared__ float d2[kMlpB][kSensorH + 1];
    __shared__ float d3[kMlpB];
    __shared__ float ls[kMlpB];
    const int t = threadIdx.x;
    const long k = first + (long)blockIdx.x * kMlpB + t;
    const bool ok = k < first + count;
    const long s = ok ? (k * stride + offset) % n : 0;
    const float *W1 = w + kSensorOffW1, *b1 = w + kSensorOffB1, *W2 = w + kSensorOffW2, *b2 = w + kSensorOffB2;
    const float *w3 = w + kSensorOffW3;
    const float b3 = w[kSensorOffB3];
    float x[kSensorNF];
#pragma unroll
    for (int f = 0; f < kSensorNF; f++) {
        x[f] = ok ? X[s * kSensorNF + f] : 0.f;
        a0[t][f] = x[f];
    }
    float z1[kSensorH], h1[kSensorH];
#pragma unroll
    for (int j = 0; j < kSensorH; j++) {
        float acc = b1[j];
#pragma unroll
        for (int f = 0; f < kSensorNF; f++) acc = fmaf(x[f], W1[f * kSensorH + j], acc);
        z1[j] = acc;
        h1[j] = fmaxf(acc, 0.f);
        a1[t][j] = h1[j];
    }
    float z2[kSensorH];
    float z3 = b3;
#pragma unroll
    for (int j = 0; j < kSensorH; j++) {
        float acc = b2[j];
#pragma unroll
        for (int c = 0; c < kSensorH; c++) acc = fmaf(h1[c], W2[c * kSensorH + j], acc);
        z2[j] = acc;
        const float h2 = fmaxf(acc, 0.f);
        a2[t][j] = h2;
        z3 = fmaf(h2, w3[j], z3);
    }
    const float y = ok ? Y[s] : 0.f;
    // BCE with logits: softplus(z) - y z;  dL/dz = sigmoid(z) - y
    const float sp = z3 > 0.f ? z3 + log1pf(expf(-z3)) : log1pf(expf(z3));
    ls[t] = ok ? sp - y * z3 : 0.f;
    const float p = 1.f / (1.f + expf(-z3));
    const float g3 = ok ? (p - y) * inv_count : 0.f;
    d3[t] = g3;
    float g2[kSensorH];
#pragma unroll
    for (int j = 0; j < kSensorH; j++) {
        g2[j] = z2[j] > 0.f ? g3 * w3[j] : 0.f;
        d2[t][j] = g2[j];
    }
#pragma unroll
    for (int c = 0; c < kSensorH; c++) {
        float acc = 0.f;
#pragma unroll
        for (int j = 0; j < kSensorH; j++) acc = fmaf(g2[j], W2[c * kSensorH + j], acc);
        d1[t][c] = z1[c] > 0.f ? acc : 0.f;
    }
    __syncthreads();
    // gradient entries, each summed over the block's samples in order
    for (int e = t; e < kSensorNParam; e += kMlpB) {
        float acc = 0.f;
        if (e < kSensorOffB1) {
            const int f = e / kSensorH, j = e % kSensorH;
            for (int r = 0; r < kMlpB; r++) acc = fmaf(a0[r][f], d1[r][j], acc);
        } else if (e < kSensorOffW2) {
            const int j = e - kSensorOffB1;
            for (int r = 0; r < kMlpB; r++) acc += d1[r][j];
        } else if (e < kSensorOffB2) {
            const int c = (e - kSensorOffW2) / kSensorH, j = (e - kSensorOffW2) % kSensorH;
            for (int r = 0; r < kMlpB; r++) acc = fmaf(a1[r][c], d2[r][j], acc);
        } else if (e < kSensorOffW3) {
            const int j = e - kSensorOffB2;
            for (int r = 0; r < kMlpB; r++) acc += d2[r][j];
        } else if (e < kSensorOffB3) {
            const int j = e - kSensorOffW3;
            for (int r = 0; r < kMlpB; r++) acc = fmaf(a2[r][j], d3[r], acc);
        } else {
            for (int r = 0; r < kMlpB; r++) acc += d3[r];
        }
        part[(long)blockIdx.x * kSensorNParam + e] = acc;
    }
    if (t == 0) {
        float acc = 0.f;
        for (int r = 0; r < kMlpB; r++) acc += ls[r];
        lpart[blockIdx.x] = acc;
    }
}

// sum of the block partials in block order (deterministic)
__global__ void k_mlp_reduce(const float *__restrict__ part, const float *__restrict__ lpart, int nblocks,
                             float *__restrict__ grad, float *__restrict__ loss, float inv_count)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < kSensorNParam) {
        float acc = 0.f;
        for (int b = 0; b < nblocks; b++) acc += part[(long)b * kSensorNParam + e];
        grad[e] = acc;
    }
    if (e == kSensorNParam) {
        float acc = 0.f;
        for (int b = 0; b < nblocks; b++) acc += lpart[b];
        *loss = acc * inv_count;
    }
}

__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= kSensorNParam) return;
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    w[e] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

int launch_mlp_grad(hipStream_t s, const float *w, const float *X, const float *Y, long n, long first, long count,
                    long stride, long offset, float *part, float *lpart, float *grad, float *loss)
{
    if (count <= 0) return 0;
    const int nb = (int)((count + kMlpB - 1) / kMlpB);
    const float inv = 1.f / (float)count;
    hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(kMlpB), 0, s, w, X, Y, n, first, count, stride, offset, inv,
                       part, lpart);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kSensorNParam + 1 + 255) / 256), dim3(256), 0, s, part, lpart, nb,
                       grad, loss, inv);
    ACFD_CHECK_LAUNCH();
    return 0;
}

int launch_adam(hipStream_t s, float *w, const float *g, float *m, float *v, float lr, int t)
{
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f / (1.f - powf(b1, (float)t)), c2 = 1.f / (1.f - powf(b2, (float)t));
    hipLaunchKernelGGL(k_adam, dim3((kSensorNParam + 255) / 256), dim3(256), 0, s, w, g, m, v, lr, b1, b2, eps, c1,
                       c2);
    ACFD_CHECK_LAUNCH();
    return 0;
}

long mlp_partials_needed(long count) { return ((count + kMlpB - 1) / kMlpB) * (long)kSensorNParam; }

// ---------------------------------------------------------- training data
// Faces with the full six-node stencil (a = lo+2 .. hi-3 of the extended
// range) on every stride-th line of the computed range: features and the
// label (nonlinear WCNS weights off the linear ones by more than 1e-3).
// Sample s*L + (a - lo - 2) of sampled line s, so the order is fixed.
template <int AXIS>
__global__ void __launch_bounds__(256) k_sensor_collect(Ext e, const double *__restrict__ q, double gamma,
                                                        int stride, long nlines, float *__restrict__ X,
                                                        float *__restrict__ Y)
{
    const int lo = e.lo(AXIS), hi = e.hi(AXIS);
    const int L = hi - lo - 4;
    const int pos = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const long s = (long)blockIdx.y + 65535L * blockIdx.z;
    const long line = s * stride;
    if (pos >= L || line >= nlines) return;
    // the other two axes over their computed ranges
    const int o1 = AXIS == 0 ? 1 : 0, o2 = AXIS == 2 ? 1 : 2;
    const int n1 = e.hi1(o1) - e.lo1(o1) + 1;
    const int c1 = e.lo1(o1) + (int)(line % n1), c2 = e.lo1(o2) + (int)(line / n1);
    const int a = lo + 2 + pos;
    int ijk[3];
    ijk[AXIS] = a;
    ijk[o1] = c1;
    ijk[o2] = c2;
    const long N = e.n, st = e.stride(AXIS);
    const long n = e.ix(ijk[0], ijk[1], ijk[2]);
    double W[6][5];
#pragma unroll
    for (int o = 0; o < 6; o++)
#pragma unroll
        for (int x = 0; x < 5; x++) W[o][x] = q[x * N + n + (long)(o - 2) * st];
    float f[kSensorNF];
    sensor_features<double>(W, gamma, f);
    const long out = s * L + pos;
#pragma unroll
    for (int c = 0; c < kSensorNF; c++) X[out * kSensorNF + c] = f[c];
    Y[out] = wcns_weight_deviation(W) > kSensorLabelBound ? 1.f : 0.f;
}

long sensor_samples(const Ext &e, int axis, int stride)
{
    const int o1 = axis == 0 ? 1 : 0, o2 = axis == 2 ? 1 : 2;
    const long nl = (long)(e.hi1(o1) - e.lo1(o1) + 1) * (e.hi1(o2) - e.lo1(o2) + 1);
    const long L = e.hi(axis) - e.lo(axis) - 4;
    if (nl <= 0 || L <= 0) return 0;
    return ((nl + stride - 1) / stride) * L;
}

int launch_sensor_collect(hipStream_t s, const Ext &e, const Fields &f, double gamma, int axis, int stride,
                          float *X, float *Y)
{
    const int o1 = axis == 0 ? 1 : 0, o2 = axis == 2 ? 1 : 2;
    const long nl = (long)(e.hi1(o1) - e.lo1(o1) + 1) * (e.hi1(o2) - e.lo1(o2) + 1);
    const int L = e.hi(axis) - e.lo(axis) - 4;
    if (nl <= 0 || L <= 0) return 0;
    const long ns = (nl + stride - 1) / stride;
    dim3 g((L + 255) / 256, (unsigned)std::min(ns, 65535L), (unsigned)((ns + 65534) / 65535));
    if (axis == 0) hipLaunchKernelGGL(k_sensor_collect<0>, g, dim3(256), 0, s, e, f.q, gamma, stride, nl, X, Y);
    else if (axis == 1) hipLaunchKernelGGL(k_sensor_collect<1>, g, dim3(256), 0, s, e, f.q, gamma, stride, nl, X, Y);
    else hipLaunchKernelGGL(k_sensor_collect<2>, g, dim3(256), 0, s, e, f.q, gamma, stride, nl, X, Y);
    ACFD_CHECK_LAUNCH();
    return 0;
}

// ------------------------------------------------------------- audits
// per-workgroup sums of (a - b)^2 and b^2 over the computed range (5
// components); the host adds the partials in order
__global__ void __launch_bounds__(256) k_diff_partial(Ext e, Box bx, const double *__restrict__ a,
                                                      const double *__restrict__ b, double *__restrict__ part)
{
    __shared__ double s0[256], s1[256];
    int i, j, k;
    double d2 = 0.0, b2 = 0.0;
    if (box_node(bx, i, j, k)) {
        const long N = e.n, n = e.ix(i, j, k);
        for (int x = 0; x < 5; x++) {
            const double d = a[x * N + n] - b[x * N + n];
            d2 += d * d;
            b2 += b[x * N + n] * b[x * N + n];
        }
    }
    const int t = threadIdx.x + blockDim.x * threadIdx.y;
    s0[t] = d2;
    s1[t] = b2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            s0[t] += s0[t + w];
            s1[t] += s1[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        const long blk = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
        part[2 * blk] = s0[0];
        part[2 * blk + 1] = s1[0];
    }
}

long diff_partials(const Ext &e)
{
    Box b{e.is1, e.ie1, e.js1, e.je1, e.ks1, e.ke1};
    dim3 g = box_grid(b);
    return (long)g.x * g.y * g.z;
}

int launch_diff_partial(hipStream_t s, const Ext &e, const double *a, const double *b, double *part)
{
    Box bx{e.is1, e.ie1, e.js1, e.je1, e.ks1, e.ke1};
    dim3 g = box_grid(bx);
    if (g.x == 0) return 0;
    hipLaunchKernelGGL(k_diff_partial, g, dim3(64, 4, 1), 0, s, e, bx, a, b, part);
    ACFD_CHECK_LAUNCH();
    return 0;
}

// ------------------------------------------------------ BF16 inference
// Pack the FP32 parameters for inference: W1 as BF16 [16][32] (rows 12..15
// zero), W2 as BF16 [32][32], the rest FP32.
__global__ void k_sensor_pack(const float *__restrict__ w, SensorNet *net)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < 16 * kSensorH) {
        const int f = e / kSensorH, j = e % kSensorH;
        net->w1[e] = (__bf16)(f < kSensorNF ? w[kSensorOffW1 + f * kSensorH + j] : 0.f);
    }
    if (e < kSensorH * kSensorH) net->w2[e] = (__bf16)w[kSensorOffW2 + e];
    if (e < kSensorH) {
        net->b1[e] = w[kSensorOffB1 + e];
        net->b2[e] = w[kSensorOffB2 + e];
        net->w3[e] = w[kSensorOffW3 + e];
    }
    if (e == 0) net->b3 = w[kSensorOffB3];
}

size_t sensor_net_bytes() { return sizeof(SensorNet); }

int launch_sensor_pack(hipStream_t s, const float *w, SensorNet *net)
{
    hipLaunchKernelGGL(k_sensor_pack, dim3(4), dim3(256), 0, s, w, net);
    ACFD_CHECK_LAUNCH();
    return 0;
}

// Batch inference through the fused wavefront routine (the path the WCNS
// kernels use), for testing and for measuring the sensor's accuracy.
__global__ void __launch_bounds__(256) k_sensor_infer(const SensorNet *__restrict__ net, const float *__restrict__ X,
                                                      long n, float *__restrict__ prob)
{
    __shared__ SensorLds lds;
    sensor_load_net(net, lds);
    __syncthreads();
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float f[kSensorNF];
#pragma unroll
    for (int c = 0; c < kSensorNF; c++) f[c] = k < n ? X[k * kSensorNF + c] : 0.f;
    const float p = sensor_eval_wave(lds, threadIdx.x >> 6, f);
    if (k < n) prob[k] = p;
}

int launch_sensor_infer(hipStream_t s, const SensorNet *net, const float *X, long n, float *prob)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_sensor_infer, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, net, X, n, prob);
    ACFD_CHECK_LAUNCH();
    return 0;
}

}  // namespace acfd
