// Small multilayer perceptron on MFMA (BF16 operands, FP32 accumulation) with its
// backward pass and Adam, for the learned preconditioner.  No reference counterpart.
//
// Layer l: Z_l = H_l W_l^T + b_l, H_{l+1} = relu(Z_l) (no activation on the last
// layer).  Master weights, activations and gradients are FP32 in device memory;
// every GEMM rounds both operands to BF16 (round to nearest even) on their way into
// LDS and accumulates in FP32 with v_mfma_f32_16x16x16_bf16.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "tg_internal.h"
#include "tg_mlp.h"

namespace tg {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short to_bf16(float f) {
    unsigned int u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// C[m, n] (+)= sum_k A(m, k) B(k, n) with A(m, k) = A[m*sam + k*sak], B(k, n) =
// B[k*sbk + n*sbn], C[m*ldc + n].  64 x 64 tile per workgroup of 4 waves, each wave a
// 32 x 32 quadrant of 2 x 2 MFMA 16x16x16 tiles; K advances 16 at a time through LDS.
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const float *A, int64_t sam, int64_t sak,
                                                   const float *B, int64_t sbk, int64_t sbn, float *C, int64_t ldc,
                                                   int accumulate) {
    __shared__ unsigned short As[64][16 + 4];
    __shared__ unsigned short Bs[16][64 + 4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        for (int t = tid; t < 64 * 16; t += 256) {
            const int mm = t >> 4, kk = t & 15;
            const int gm = m0 + mm, gk = k0 + kk;
            As[mm][kk] = (gm < M && gk < K) ? to_bf16(A[gm * sam + gk * sak]) : (unsigned short)0;
        }
        for (int t = tid; t < 16 * 64; t += 256) {
            const int kk = t >> 6, nn = t & 63;
            const int gk = k0 + kk, gn = n0 + nn;
            Bs[kk][nn] = (gk < K && gn < N) ? to_bf16(B[gk * sbk + gn * sbn]) : (unsigned short)0;
        }
        __syncthreads();
        const int r = lane & 15, kq = (lane >> 4) * 4;
        for (int a = 0; a < 2; ++a) {
            bf16x4 av;
            for (int i = 0; i < 4; ++i) av[i] = (short)As[wm + a * 16 + r][kq + i];
            for (int b = 0; b < 2; ++b) {
                bf16x4 bv;
                for (int i = 0; i < 4; ++i) bv[i] = (short)Bs[kq + i][wn + b * 16 + r];
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[a][b], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    const int col = lane & 15, rq = (lane >> 4) * 4;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int i = 0; i < 4; ++i) {
                const int gm = m0 + wm + a * 16 + rq + i, gn = n0 + wn + b * 16 + col;
                if (gm < M && gn < N) {
                    float *dst = C + (int64_t)gm * ldc + gn;
                    *dst = accumulate ? *dst + acc[a][b][i] : acc[a][b][i];
                }
            }
}

static int gemm(hipStream_t s, int M, int N, int K, const float *A, int64_t sam, int64_t sak, const float *B,
                int64_t sbk, int64_t sbn, float *C, int64_t ldc, int accumulate) {
    dim3 grid((N + 63) / 64, (M + 63) / 64);
    k_gemm_bf16<<<grid, 256, 0, s>>>(M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, accumulate);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

// H = act(Z + b) row-wise; Z kept.
__global__ void k_bias_act(const float *Z, const float *bias, float *Hout, int64_t B, int n, int relu) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * n) return;
    float v = Z[t] + bias[t % n];
    Hout[t] = relu ? fmaxf(v, 0.f) : v;
}

// Z += b in place (pre-activation with bias, kept for the backward pass)
__global__ void k_add_bias(float *Z, const float *bias, int64_t B, int n) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * n) return;
    Z[t] += bias[t % n];
}

// db[j] += sum_r dZ[r, j] (fixed row order per column)
__global__ void k_bias_grad(const float *dZ, int64_t B, int n, float *db) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float s = 0.f;
    for (int64_t r = 0; r < B; ++r) s += dZ[r * n + j];
    db[j] += s;
}

// dZ = dH * (Z > 0)
__global__ void k_relu_back(const float *dH, const float *Z, float *dZ, int64_t count) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    dZ[t] = Z[t] > 0.f ? dH[t] : 0.f;
}

__global__ void k_adam(float *p, const float *g, float *m, float *v, int64_t count, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const float gg = g[t];
    const float mm = b1 * m[t] + (1.f - b1) * gg;
    const float vv = b2 * v[t] + (1.f - b2) * gg * gg;
    m[t] = mm;
    v[t] = vv;
    p[t] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

static int grid_of(int64_t count) { return (int)((count + 255) / 256); }

int mlp_ensure_batch(tg_mlp *m, int64_t batch) {
    if (batch <= m->capacity) return TG_OK;
    int64_t per = 0;
    for (int l = 0; l <= m->layers; ++l) per += m->dims[l];   // H_0..H_L
    for (int l = 0; l < m->layers; ++l) per += m->dims[l + 1]; // Z_0..Z_{L-1}
    per += 2 * m->width_max;                                   // two gradient scratch rows
    TG_TRY(m->acts.ensure(sizeof(float) * per * batch));
    m->capacity = batch;
    return TG_OK;
}

// Offsets of H_l and Z_l inside acts for the current capacity.
static float *act_H(tg_mlp *m, int l) {
    int64_t off = 0;
    for (int i = 0; i < l; ++i) off += m->dims[i];
    return m->acts.as<float>() + off * m->capacity;
}
static float *act_Z(tg_mlp *m, int l) {
    int64_t off = 0;
    for (int i = 0; i <= m->layers; ++i) off += m->dims[i];
    for (int i = 0; i < l; ++i) off += m->dims[i + 1];
    return m->acts.as<float>() + off * m->capacity;
}
static float *act_scratch(tg_mlp *m, int which) {
    int64_t off = 0;
    for (int i = 0; i <= m->layers; ++i) off += m->dims[i];
    for (int i = 0; i < m->layers; ++i) off += m->dims[i + 1];
    return m->acts.as<float>() + (off + which * m->width_max) * m->capacity;
}

float *mlp_W(tg_mlp *m, DevBuf &buf, int l) { return buf.as<float>() + m->w_off[l]; }
float *mlp_b(tg_mlp *m, DevBuf &buf, int l) { return buf.as<float>() + m->b_off[l]; }

int mlp_forward(tg_mlp *m, const float *x, int64_t batch, float *y) {
    TG_TRY(mlp_ensure_batch(m, batch));
    hipStream_t s = m->stream;
    TG_HIP(hipMemcpyAsync(act_H(m, 0), x, sizeof(float) * batch * m->dims[0], hipMemcpyDeviceToDevice, s));
    for (int l = 0; l < m->layers; ++l) {
        const int din = m->dims[l], dout = m->dims[l + 1];
        float *Z = act_Z(m, l);
        TG_TRY(gemm(s, (int)batch, dout, din, act_H(m, l), din, 1, mlp_W(m, m->params, l), 1, din, Z, dout, 0));
        k_add_bias<<<grid_of(batch * dout), 256, 0, s>>>(Z, mlp_b(m, m->params, l), batch, dout);
        TG_HIP(hipGetLastError());
        const bool last = l == m->layers - 1;
        k_bias_act<<<grid_of(batch * dout), 256, 0, s>>>(Z, mlp_b(m, m->zeros, l), last ? y : act_H(m, l + 1), batch,
                                                         dout, last ? 0 : 1);
        TG_HIP(hipGetLastError());
    }
    m->last_batch = batch;
    return TG_OK;
}

int mlp_backward(tg_mlp *m, const float *dy) {
    const int64_t batch = m->last_batch;
    hipStream_t s = m->stream;
    float *dZ = act_scratch(m, 0), *dH = act_scratch(m, 1);
    const int L = m->layers;
    TG_HIP(hipMemcpyAsync(dZ, dy, sizeof(float) * batch * m->dims[L], hipMemcpyDeviceToDevice, s));
    for (int l = L - 1; l >= 0; --l) {
        const int din = m->dims[l], dout = m->dims[l + 1];
        // dW_l += dZ^T H_l  (dout x din, K = batch)
        TG_TRY(gemm(s, dout, din, (int)batch, dZ, 1, dout, act_H(m, l), din, 1, mlp_W(m, m->grads, l), din, 1));
        k_bias_grad<<<(dout + 255) / 256, 256, 0, s>>>(dZ, batch, dout, mlp_b(m, m->grads, l));
        TG_HIP(hipGetLastError());
        if (l > 0) {
            // dH_l = dZ W_l (batch x din), then dZ_{l-1} = dH_l * (Z_{l-1} > 0)
            TG_TRY(gemm(s, (int)batch, din, dout, dZ, dout, 1, mlp_W(m, m->params, l), din, 1, dH, din, 0));
            k_relu_back<<<grid_of(batch * din), 256, 0, s>>>(dH, act_Z(m, l - 1), dZ, batch * din);
            TG_HIP(hipGetLastError());
        }
    }
    return TG_OK;
}

static uint64_t splitmix(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace tg

using namespace tg;

extern "C" {

int tg_mlp_create(int n_in, int width, int n_out, int hidden_layers, uint64_t seed, void *stream, tg_mlp **out) {
    if (!out || n_in < 1 || width < 1 || n_out < 1 || hidden_layers < 1 || hidden_layers > 8) {
        set_error("invalid MLP shape");
        return TG_INVALID_INPUT;
    }
    tg_mlp *m = new tg_mlp();
    m->stream = (hipStream_t)stream;
    m->layers = hidden_layers + 1;
    m->dims.push_back(n_in);
    for (int i = 0; i < hidden_layers; ++i) m->dims.push_back(width);
    m->dims.push_back(n_out);
    m->width_max = 0;
    for (int d : m->dims) m->width_max = std::max(m->width_max, d);
    int64_t off = 0;
    for (int l = 0; l < m->layers; ++l) {
        m->w_off.push_back(off);
        off += (int64_t)m->dims[l] * m->dims[l + 1];
        m->b_off.push_back(off);
        off += m->dims[l + 1];
    }
    m->count = off;
    std::vector<float> host(off, 0.f);
    uint64_t state = seed;
    for (int l = 0; l < m->layers; ++l) {
        const bool last = l == m->layers - 1;
        const float bound = 1.f / std::sqrt((float)m->dims[l]);
        for (int64_t i = 0; i < (int64_t)m->dims[l] * m->dims[l + 1]; ++i) {
            double u = (splitmix(state) >> 11) * (1.0 / 9007199254740992.0);
            host[m->w_off[l] + i] = last ? 0.f : (float)((2.0 * u - 1.0) * bound);
        }
    }
    int st = TG_OK;
    for (DevBuf *b : {&m->params, &m->grads, &m->adam_m, &m->adam_v, &m->zeros})
        if (!st) st = b->ensure(sizeof(float) * off);
    if (!st && hipMemcpy(m->params.ptr, host.data(), sizeof(float) * off, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("MLP upload failed");
        st = TG_DEVICE_ERROR;
    }
    for (DevBuf *b : {&m->grads, &m->adam_m, &m->adam_v, &m->zeros})
        if (!st && hipMemset(b->ptr, 0, sizeof(float) * off) != hipSuccess) st = TG_DEVICE_ERROR;
    if (st) {
        tg_mlp_destroy(m);
        return st;
    }
    *out = m;
    return TG_OK;
}

int tg_mlp_destroy(tg_mlp *m) {
    if (!m) return TG_OK;
    for (DevBuf *b : {&m->params, &m->grads, &m->adam_m, &m->adam_v, &m->zeros, &m->acts}) b->release();
    delete m;
    return TG_OK;
}

int64_t tg_mlp_param_count(const tg_mlp *m) { return m ? m->count : 0; }

int tg_mlp_get_params(tg_mlp *m, float *host) {
    TG_HIP(hipMemcpy(host, m->params.ptr, sizeof(float) * m->count, hipMemcpyDeviceToHost));
    return TG_OK;
}

int tg_mlp_set_params(tg_mlp *m, const float *host) {
    TG_HIP(hipMemcpy(m->params.ptr, host, sizeof(float) * m->count, hipMemcpyHostToDevice));
    return TG_OK;
}

int tg_mlp_get_grads(tg_mlp *m, float *host) {
    TG_HIP(hipStreamSynchronize(m->stream));
    TG_HIP(hipMemcpy(host, m->grads.ptr, sizeof(float) * m->count, hipMemcpyDeviceToHost));
    return TG_OK;
}

int tg_mlp_zero_grad(tg_mlp *m) {
    TG_HIP(hipMemsetAsync(m->grads.ptr, 0, sizeof(float) * m->count, m->stream));
    return TG_OK;
}

int tg_mlp_forward(tg_mlp *m, const float *x, int64_t batch, float *y) {
    if (!m || batch < 1) {
        set_error("invalid MLP call");
        return TG_INVALID_INPUT;
    }
    TG_TRY(mlp_forward(m, x, batch, y));
    TG_HIP(hipStreamSynchronize(m->stream));
    return TG_OK;
}

int tg_mlp_backward(tg_mlp *m, const float *dy) {
    if (!m || m->last_batch < 1) {
        set_error("MLP backward needs a forward pass first");
        return TG_INVALID_INPUT;
    }
    TG_TRY(mlp_backward(m, dy));
    TG_HIP(hipStreamSynchronize(m->stream));
    return TG_OK;
}

int tg_mlp_adam_step(tg_mlp *m, float lr, float beta1, float beta2, float eps, int step) {
    if (!m || step < 1) {
        set_error("Adam needs step >= 1");
        return TG_INVALID_INPUT;
    }
    const float c1 = 1.f - std::pow(beta1, (float)step), c2 = 1.f - std::pow(beta2, (float)step);
    k_adam<<<grid_of(m->count), 256, 0, m->stream>>>(m->params.as<float>(), m->grads.as<float>(),
                                                      m->adam_m.as<float>(), m->adam_v.as<float>(), m->count, lr,
                                                      beta1, beta2, eps, c1, c2);
    TG_HIP(hipGetLastError());
    TG_HIP(hipStreamSynchronize(m->stream));
    return TG_OK;
}

}  // extern "C"
