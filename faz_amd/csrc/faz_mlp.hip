// MLP surrogate of the learned tuner on gfx950: in -> H -> ... -> H -> out,
// ReLU hidden layers, linear output.  Matrix products run on the matrix
// cores (v_mfma_f32_32x32x16_bf16: BF16 operands, FP32 accumulate); weights
// keep an FP32 master copy, Adam moments in FP32, and a BF16 copy the
// products read.  Dimensions are padded: inputs and hidden widths to
// multiples of 32 (K steps of 16), outputs to 32, the batch to 32.
//
// GEMM: one wave computes a 32x32 tile of C = A * B; a lane feeds the MFMA
// with 8 consecutive K values of one A row (lanes 0-31: K 0-7, lanes 32-63:
// K 8-15) and of one B column, and receives 16 results of its column
// (rows 8*(i/4) + 4*(lane/32) + i%4).  A and B are read through strides,
// so the backward products (A^T * dZ, dZ * W^T) use the same kernel.
#include "faz_gpu.h"
#include "faz_ctx.h"
#include "faz_internal.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace faz {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

__host__ __device__ inline uint32_t pad32(uint32_t x) { return (x + 31) / 32 * 32; }

// C[M x N] (+)= A[M x K] * B[K x N]; A(i,k) = A[i*sam + k*sak], B(k,j) = B[k*sbk + j*sbn].
// M, N multiples of 32, K a multiple of 16.  4 waves per workgroup, 2 x 2 tiles.
__global__ void __launch_bounds__(256) gemm_bf16_kernel(int M, int N, int K, const __bf16 *__restrict__ A,
                                                        int64_t sam, int64_t sak, const __bf16 *__restrict__ B,
                                                        int64_t sbk, int64_t sbn, float *__restrict__ C, int64_t scm,
                                                        int accumulate) {
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int m0 = (blockIdx.y * 2 + wave / 2) * 32, n0 = (blockIdx.x * 2 + wave % 2) * 32;
    if (m0 >= M || n0 >= N) return;
    const int r = lane % 32, kh = lane / 32;
    f32x16 acc;
    for (int i = 0; i < 16; i++) acc[i] = 0.0f;
    const __bf16 *pa = A + (int64_t) (m0 + r) * sam;
    const __bf16 *pb = B + (int64_t) (n0 + r) * sbn;
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x8 a, b;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int64_t k = k0 + 8 * kh + j;
            a[j] = pa[k * sak];
            b[j] = pb[k * sbk];
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int row = m0 + 8 * (i / 4) + 4 * kh + (i % 4);
        float *c = C + (int64_t) row * scm + n0 + r;
        *c = accumulate ? *c + acc[i] : acc[i];
    }
}

void gemm(int M, int N, int K, const __bf16 *A, int64_t sam, int64_t sak, const __bf16 *B, int64_t sbk, int64_t sbn,
          float *C, int64_t scm, bool accumulate, hipStream_t st) {
    dim3 grid((N + 63) / 64, (M + 63) / 64);
    hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, st, M, N, K, A, sam, sak, B, sbk, sbn, C, scm,
                       accumulate ? 1 : 0);
}

// Z = C + bias (kept for the backward pass), A = bf16(relu(Z)) or bf16(Z).
__global__ void bias_act_kernel(int M, int N, float *__restrict__ Z, const float *__restrict__ bias, int relu,
                                __bf16 *__restrict__ Aout) {
    const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t) M * N) return;
    const float z = Z[t] + bias[t % N];
    Z[t] = z;
    if (Aout) Aout[t] = (__bf16) (relu ? fmaxf(z, 0.0f) : z);
}

// Packs fp32 rows [batch x din] into bf16 [Mp x dinp], zero padding.
__global__ void pack_input_kernel(int batch, int din, int Mp, int dinp, const float *__restrict__ x,
                                  __bf16 *__restrict__ X) {
    const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t) Mp * dinp) return;
    const int i = (int) (t / dinp), j = (int) (t % dinp);
    X[t] = (__bf16) (i < batch && j < din ? x[(int64_t) i * din + j] : 0.0f);
}

__global__ void unpack_output_kernel(int batch, int dout, int ldz, const float *__restrict__ Z, float *__restrict__ y) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= batch * dout) return;
    y[t] = Z[(int64_t) (t / dout) * ldz + t % dout];
}

// Mean squared error over the first dout columns of the batch rows and its
// gradient dZ = 2 (Z - T) / (batch * dout) (bf16 for the products, zero on
// padding); per-row sums of squares in rowloss.
__global__ void mse_grad_kernel(int batch, int dout, int Mp, int ldz, const float *__restrict__ Z,
                                const float *__restrict__ T, __bf16 *__restrict__ dZ, float *__restrict__ rowloss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Mp) return;
    float s = 0.0f;
    for (int j = 0; j < ldz; j++) {
        float g = 0.0f;
        if (i < batch && j < dout) {
            const float d = Z[(int64_t) i * ldz + j] - T[(int64_t) i * dout + j];
            s += d * d;
            g = 2.0f * d / (float) (batch * dout);
        }
        dZ[(int64_t) i * ldz + j] = (__bf16) g;
    }
    rowloss[i] = s;
}

// dZ = dA * relu'(Z) as bf16 (dA in fp32).
__global__ void relu_grad_kernel(int64_t n, const float *__restrict__ dA, const float *__restrict__ Z,
                                 __bf16 *__restrict__ dZ) {
    const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dZ[t] = (__bf16) (Z[t] > 0.0f ? dA[t] : 0.0f);
}

// db[j] = sum over the batch rows of dZ(i, j), in row order.
__global__ void colsum_kernel(int M, int N, const __bf16 *__restrict__ dZ, float *__restrict__ db) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    float s = 0.0f;
    for (int i = 0; i < M; i++) s += (float) dZ[(int64_t) i * N + j];
    db[j] = s;
}

// Adam (beta1, beta2, eps) on fp32 masters; refreshes the bf16 copy.
__global__ void adam_kernel(int64_t n, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, __bf16 *__restrict__ wb, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
    const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.0f - b1) * gt;
    const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    const float wt = w[t] - lr * (mt / c1) / (sqrtf(vt / c2) + eps);
    w[t] = wt;
    if (wb) wb[t] = (__bf16) wt;
}

__global__ void to_bf16_kernel(int64_t n, const float *__restrict__ w, __bf16 *__restrict__ wb) {
    const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) wb[t] = (__bf16) w[t];
}

inline dim3 grid_for(int64_t n) { return dim3((unsigned) ((n + 255) / 256)); }

}  // namespace
}  // namespace faz

using namespace faz;

struct faz_mlp {
    int din = 0, hidden = 0, hlayers = 0, dout = 0;
    int device = -1;
    std::vector<int> K, N;  // padded layer shapes (K_l x N_l)
    std::vector<faz::DevBuf<float>> w, b, gw, gb, mw, vw, mb, vb;
    std::vector<faz::DevBuf<__bf16>> wb;
    int step = 0;
    int cap = 0;  // padded batch the activation buffers hold
    faz::DevBuf<__bf16> X;
    std::vector<faz::DevBuf<float>> Z, dA;      // per layer: pre-activation, gradient of the output
    std::vector<faz::DevBuf<__bf16>> Ab, dZ;    // per layer: bf16 activation, bf16 dZ
    faz::DevBuf<float> rowloss;
    std::vector<float> h_rowloss;
};

namespace {

int ensure_batch(faz_gpu_ctx *ctx, faz_mlp *net, int Mp) {
    if (net->cap >= Mp) return FAZ_OK;
    const int L = (int) net->K.size();
    HIPCHECK(net->X.ensure((size_t) Mp * net->K[0]));
    net->Z.resize(L);
    net->dA.resize(L);
    net->Ab.resize(L);
    net->dZ.resize(L);
    for (int l = 0; l < L; l++) {
        HIPCHECK(net->Z[l].ensure((size_t) Mp * net->N[l]));
        HIPCHECK(net->dA[l].ensure((size_t) Mp * net->N[l]));
        HIPCHECK(net->Ab[l].ensure((size_t) Mp * net->N[l]));
        HIPCHECK(net->dZ[l].ensure((size_t) Mp * net->N[l]));
    }
    HIPCHECK(net->rowloss.ensure(Mp));
    net->cap = Mp;
    return FAZ_OK;
}

// Forward pass over the packed input in net->X (Mp rows).
void forward(faz_mlp *net, int Mp, hipStream_t st) {
    const int L = (int) net->K.size();
    const __bf16 *in = net->X.p;
    for (int l = 0; l < L; l++) {
        const int K = net->K[l], N = net->N[l];
        gemm(Mp, N, K, in, K, 1, net->wb[l].p, N, 1, net->Z[l].p, N, false, st);
        const bool relu = l + 1 < L;
        hipLaunchKernelGGL(bias_act_kernel, grid_for((int64_t) Mp * N), dim3(256), 0, st, Mp, N, net->Z[l].p,
                           net->b[l].p, relu ? 1 : 0, net->Ab[l].p);
        in = net->Ab[l].p;
    }
}

// Backward pass from the loss on the first dout outputs; leaves dW, db.
void backward(faz_mlp *net, int batch, int Mp, const float *d_t, hipStream_t st) {
    const int L = (int) net->K.size();
    hipLaunchKernelGGL(mse_grad_kernel, grid_for(Mp), dim3(256), 0, st, batch, net->dout, Mp, net->N[L - 1],
                       net->Z[L - 1].p, d_t, net->dZ[L - 1].p, net->rowloss.p);
    for (int l = L - 1; l >= 0; l--) {
        const int K = net->K[l], N = net->N[l];
        if (l + 1 < L)
            hipLaunchKernelGGL(relu_grad_kernel, grid_for((int64_t) Mp * N), dim3(256), 0, st, (int64_t) Mp * N,
                               net->dA[l].p, net->Z[l].p, net->dZ[l].p);
        const __bf16 *in = l ? net->Ab[l - 1].p : net->X.p;
        // dW = in^T dZ : [K x Mp] [Mp x N]
        gemm(K, N, Mp, in, 1, K, net->dZ[l].p, N, 1, net->gw[l].p, N, false, st);
        hipLaunchKernelGGL(colsum_kernel, grid_for(N), dim3(256), 0, st, Mp, N, net->dZ[l].p, net->gb[l].p);
        // dA(l-1) = dZ W^T : [Mp x N] [N x K]
        if (l) gemm(Mp, K, N, net->dZ[l].p, N, 1, net->wb[l].p, 1, N, net->dA[l - 1].p, K, false, st);
    }
}

int pack(faz_gpu_ctx *ctx, faz_mlp *net, const float *d_x, int batch, int &Mp) {
    Mp = (int) pad32((uint32_t) batch);
    int rc = ensure_batch(ctx, net, Mp);
    if (rc) return rc;
    hipLaunchKernelGGL(pack_input_kernel, grid_for((int64_t) Mp * net->K[0]), dim3(256), 0, ctx->stream, batch,
                       net->din, Mp, net->K[0], d_x, net->X.p);
    return FAZ_OK;
}

}  // namespace

extern "C" {

faz_status faz_mlp_create(faz_gpu_ctx *ctx, int in_dim, int hidden, int hidden_layers, int out_dim, uint64_t seed,
                          faz_mlp **out) {
    if (!ctx || !out || in_dim < 1 || hidden < 1 || hidden_layers < 1 || out_dim < 1 || in_dim > 4096 ||
        hidden > 4096 || out_dim > 32 || hidden_layers > 16)
        return FAZ_ERR_INVALID;
    HIPCHECK(hipSetDevice(ctx->device));
    faz_mlp *net = new faz_mlp();
    net->din = in_dim, net->hidden = hidden, net->hlayers = hidden_layers, net->dout = out_dim;
    net->device = ctx->device;
    const int hp = (int) pad32((uint32_t) hidden);
    net->K.push_back((int) pad32((uint32_t) in_dim));
    net->N.push_back(hp);
    for (int l = 1; l < hidden_layers; l++) {
        net->K.push_back(hp);
        net->N.push_back(hp);
    }
    net->K.push_back(hp);
    net->N.push_back(32);
    const int L = (int) net->K.size();
    net->w.resize(L), net->b.resize(L), net->gw.resize(L), net->gb.resize(L), net->mw.resize(L), net->vw.resize(L),
            net->mb.resize(L), net->vb.resize(L), net->wb.resize(L);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < L; l++) {
        const int K = net->K[l], N = net->N[l];
        const int kin = l == 0 ? in_dim : hidden, nout = l + 1 == L ? out_dim : hidden;
        std::vector<float> w((size_t) K * N, 0.0f), z((size_t) K * N, 0.0f), zb(N, 0.0f);
        // He-uniform on the real (unpadded) block, zero on padding
        const float lim = std::sqrt(6.0f / (float) kin);
        std::uniform_real_distribution<float> U(-lim, lim);
        for (int i = 0; i < kin; i++)
            for (int j = 0; j < nout; j++) w[(size_t) i * N + j] = U(rng);
        bool ok = net->w[l].ensure(w.size()) == hipSuccess && net->gw[l].ensure(w.size()) == hipSuccess &&
                  net->mw[l].ensure(w.size()) == hipSuccess && net->vw[l].ensure(w.size()) == hipSuccess &&
                  net->wb[l].ensure(w.size()) == hipSuccess && net->b[l].ensure(N) == hipSuccess &&
                  net->gb[l].ensure(N) == hipSuccess && net->mb[l].ensure(N) == hipSuccess &&
                  net->vb[l].ensure(N) == hipSuccess;
        ok = ok && hipMemcpy(net->w[l].p, w.data(), w.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(net->mw[l].p, z.data(), z.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(net->vw[l].p, z.data(), z.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(net->b[l].p, zb.data(), N * 4, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(net->mb[l].p, zb.data(), N * 4, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(net->vb[l].p, zb.data(), N * 4, hipMemcpyHostToDevice) == hipSuccess;
        if (!ok) {
            faz_mlp_destroy(net);
            CTX_FAIL(FAZ_ERR_HIP, "MLP allocation failed");
        }
        hipLaunchKernelGGL(to_bf16_kernel, grid_for((int64_t) w.size()), dim3(256), 0, ctx->stream,
                           (int64_t) w.size(), net->w[l].p, net->wb[l].p);
    }
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    *out = net;
    return FAZ_OK;
}

void faz_mlp_destroy(faz_mlp *net) {
    if (!net) return;
    (void) hipSetDevice(net->device);
    auto rel = [](auto &v) {
        for (auto &x: v) x.release();
    };
    rel(net->w), rel(net->b), rel(net->gw), rel(net->gb), rel(net->mw), rel(net->vw), rel(net->mb), rel(net->vb);
    rel(net->wb), rel(net->Z), rel(net->dA), rel(net->Ab), rel(net->dZ);
    net->X.release();
    net->rowloss.release();
    delete net;
}

faz_status faz_mlp_layout(const faz_mlp *net, int *n_layers, int *K, int *N) {
    if (!net || !n_layers) return FAZ_ERR_INVALID;
    *n_layers = (int) net->K.size();
    for (size_t l = 0; l < net->K.size(); l++) {
        if (K) K[l] = net->K[l];
        if (N) N[l] = net->N[l];
    }
    return FAZ_OK;
}

faz_status faz_mlp_forward(faz_gpu_ctx *ctx, faz_mlp *net, const float *d_x, int batch, float *d_y) {
    if (!ctx || !net || !d_x || !d_y || batch < 1) return FAZ_ERR_INVALID;
    HIPCHECK(hipSetDevice(ctx->device));
    int Mp;
    int rc = pack(ctx, net, d_x, batch, Mp);
    if (rc) return rc;
    forward(net, Mp, ctx->stream);
    const int L = (int) net->K.size();
    hipLaunchKernelGGL(unpack_output_kernel, grid_for((int64_t) batch * net->dout), dim3(256), 0, ctx->stream, batch,
                       net->dout, net->N[L - 1], net->Z[L - 1].p, d_y);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    return FAZ_OK;
}

faz_status faz_mlp_gradients(faz_gpu_ctx *ctx, faz_mlp *net, const float *d_x, const float *d_t, int batch,
                             float *h_grad, size_t cap, double *loss) {
    if (!ctx || !net || !d_x || !d_t || batch < 1) return FAZ_ERR_INVALID;
    HIPCHECK(hipSetDevice(ctx->device));
    int Mp;
    int rc = pack(ctx, net, d_x, batch, Mp);
    if (rc) return rc;
    forward(net, Mp, ctx->stream);
    backward(net, batch, Mp, d_t, ctx->stream);
    HIPCHECK(hipGetLastError());
    size_t need = 0;
    for (size_t l = 0; l < net->K.size(); l++) need += (size_t) net->K[l] * net->N[l] + net->N[l];
    if (h_grad) {
        if (cap < need) CTX_FAIL(FAZ_ERR_CAPACITY, "gradient buffer needs %zu floats", need);
        size_t at = 0;
        for (size_t l = 0; l < net->K.size(); l++) {
            const size_t nw = (size_t) net->K[l] * net->N[l];
            HIPCHECK(hipMemcpyAsync(h_grad + at, net->gw[l].p, nw * 4, hipMemcpyDeviceToHost, ctx->stream));
            at += nw;
            HIPCHECK(hipMemcpyAsync(h_grad + at, net->gb[l].p, net->N[l] * 4, hipMemcpyDeviceToHost, ctx->stream));
            at += net->N[l];
        }
    }
    net->h_rowloss.resize(Mp);
    HIPCHECK(hipMemcpyAsync(net->h_rowloss.data(), net->rowloss.p, Mp * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    if (loss) {
        double s = 0;
        for (int i = 0; i < batch; i++) s += net->h_rowloss[i];
        *loss = s / ((double) batch * net->dout);
    }
    return FAZ_OK;
}

faz_status faz_mlp_train_step(faz_gpu_ctx *ctx, faz_mlp *net, const float *d_x, const float *d_t, int batch,
                              float lr, double *loss) {
    int rc = faz_mlp_gradients(ctx, net, d_x, d_t, batch, nullptr, 0, loss);
    if (rc) return rc;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    net->step++;
    const float c1 = 1.0f - std::pow(b1, (float) net->step), c2 = 1.0f - std::pow(b2, (float) net->step);
    for (size_t l = 0; l < net->K.size(); l++) {
        const int64_t nw = (int64_t) net->K[l] * net->N[l];
        hipLaunchKernelGGL(adam_kernel, grid_for(nw), dim3(256), 0, ctx->stream, nw, net->w[l].p, net->gw[l].p,
                           net->mw[l].p, net->vw[l].p, net->wb[l].p, lr, b1, b2, eps, c1, c2);
        hipLaunchKernelGGL(adam_kernel, grid_for(net->N[l]), dim3(256), 0, ctx->stream, (int64_t) net->N[l],
                           net->b[l].p, net->gb[l].p, net->mb[l].p, net->vb[l].p, (__bf16 *) nullptr, lr, b1, b2,
                           eps, c1, c2);
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    return FAZ_OK;
}

faz_status faz_mlp_get_params(faz_gpu_ctx *ctx, const faz_mlp *net, float *h, size_t cap, size_t *n) {
    if (!ctx || !net || !n) return FAZ_ERR_INVALID;
    size_t need = 0;
    for (size_t l = 0; l < net->K.size(); l++) need += (size_t) net->K[l] * net->N[l] + net->N[l];
    *n = need;
    if (!h || cap < need) return FAZ_ERR_CAPACITY;
    HIPCHECK(hipSetDevice(ctx->device));
    size_t at = 0;
    for (size_t l = 0; l < net->K.size(); l++) {
        const size_t nw = (size_t) net->K[l] * net->N[l];
        HIPCHECK(hipMemcpy(h + at, net->w[l].p, nw * 4, hipMemcpyDeviceToHost));
        at += nw;
        HIPCHECK(hipMemcpy(h + at, net->b[l].p, net->N[l] * 4, hipMemcpyDeviceToHost));
        at += net->N[l];
    }
    return FAZ_OK;
}

faz_status faz_mlp_set_params(faz_gpu_ctx *ctx, faz_mlp *net, const float *h, size_t n) {
    if (!ctx || !net || !h) return FAZ_ERR_INVALID;
    size_t need = 0;
    for (size_t l = 0; l < net->K.size(); l++) need += (size_t) net->K[l] * net->N[l] + net->N[l];
    if (n != need) CTX_FAIL(FAZ_ERR_INVALID, "parameter count %zu, network holds %zu", n, need);
    HIPCHECK(hipSetDevice(ctx->device));
    size_t at = 0;
    for (size_t l = 0; l < net->K.size(); l++) {
        const size_t nw = (size_t) net->K[l] * net->N[l];
        HIPCHECK(hipMemcpy(net->w[l].p, h + at, nw * 4, hipMemcpyHostToDevice));
        at += nw;
        HIPCHECK(hipMemcpy(net->b[l].p, h + at, net->N[l] * 4, hipMemcpyHostToDevice));
        at += net->N[l];
        hipLaunchKernelGGL(to_bf16_kernel, grid_for((int64_t) nw), dim3(256), 0, ctx->stream, (int64_t) nw,
                           net->w[l].p, net->wb[l].p);
    }
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    return FAZ_OK;
}

}  // extern "C"
