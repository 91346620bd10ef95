// Surrogate network: a fully connected MLP (GELU between layers) with BF16
// weights and activations and FP32 accumulation on the matrix cores, trained
// by hand-written forward, backward and AdamW kernels.  FP32 master weights,
// AdamW moments and gradients stay on the device.
//
// GEMM: one wavefront computes one 16x16 output tile with
// v_mfma_f32_16x16x16_bf16 (K stepped by 16); a workgroup holds 2x2 tiles.
// Operand lane layout of that instruction (wave64): lane l holds
//   A[i = l%16][k = 4*(l/16) .. +3],  B[k = 4*(l/16) .. +3][j = l%16],
//   D[i = 4*(l/16) .. +3][j = l%16].
// Operands are addressed through (row, column) strides, so transposed
// products (X^T dZ, dZ W^T) need no copies.  The matrices here are at most a
// few thousand by 256, L2-resident: the kernel reads operands straight from
// global memory.
#include "cosenu_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <utility>
#include <cstring>
#include <string>
#include <vector>

// Scratch of cosenu_mlp_forward_noalloc: the BF16 copy of the input and the
// hidden activations.  One copy per GPU in the code object, held until the
// process ends; host code reaches it only through the device address that
// hipGetSymbolAddress returns for the current device (resolved once per
// device by cosenu_mlp_create, kept in g_scratch_ptr).
constexpr size_t kNoallocScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::vector<void *> g_scratch_ptr; // device address of `scratch`, by device ordinal

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;

thread_local std::string g_mlp_err;

int mfail(int code, const std::string &m)
{
    g_mlp_err = m;
    return code;
}

#define MHIP(expr)                                                                                 \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return mfail(COSENU_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));         \
    } while (0)

__host__ __device__ inline bf16_t f2bf(float f)
{
    // round to nearest even on the upper 16 bits
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) // inf / nan: truncate (keep nan quiet)
        return (bf16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}

__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }

__device__ inline float gelu_grad(float x)
{
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) +
           x * 0.39894228040143268f * expf(-0.5f * x * x);
}

enum { EPI_NONE = 0, EPI_GELU = 1 };

// Z(m,n) = sum_k A(m,k) B(k,n) [+ bias(n)], A(m,k) = A[m*am + k*ak],
// B(k,n) = B[k*bk + n*bn].  Writes Z as FP32 (if z) and act(Z) as BF16 (if o).
__global__ __launch_bounds__(256) void k_gemm(const bf16_t *__restrict__ A, long am, long ak,
                                              const bf16_t *__restrict__ B, long bk, long bn,
                                              int M, int N, int K, const float *__restrict__ bias,
                                              float *__restrict__ z, long ldz,
                                              bf16_t *__restrict__ o, long ldo, int epi)
{
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int tm = blockIdx.y * 32 + (w >> 1) * 16;
    const int tn = blockIdx.x * 32 + (w & 1) * 16;
    const int r16 = l & 15, q = l >> 4;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const int ia = tm + r16, jb = tn + r16;
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int kk = k0 + 4 * q + r;
            a[r] = (ia < M && kk < K) ? (short)A[(long)ia * am + (long)kk * ak] : (short)0;
            b[r] = (jb < N && kk < K) ? (short)B[(long)kk * bk + (long)jb * bn] : (short)0;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const int j = tn + r16;
    if (j >= N)
        return;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int i = tm + 4 * q + r;
        if (i >= M)
            continue;
        float v = acc[r] + (bias ? bias[j] : 0.0f);
        if (z)
            z[(long)i * ldz + j] = v;
        if (o)
            o[(long)i * ldo + j] = f2bf(epi == EPI_GELU ? gelu(v) : v);
    }
}

// dZ = dA * gelu'(Z) -> BF16 (operand of the next products).
__global__ void k_gelu_bwd(const float *dA, const float *Z, long n, bf16_t *dZ16, float *dZ32)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float g = dA[i] * gelu_grad(Z[i]);
    dZ16[i] = f2bf(g);
    if (dZ32)
        dZ32[i] = g;
}

// db(n) = sum_m dZ(m, n), fixed order per column.
__global__ void k_colsum(const float *dZ, int M, int N, float *db)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N)
        return;
    float s = 0.f;
    for (int i = 0; i < M; i++)
        s += dZ[(long)i * N + j];
    db[j] = s;
}

// Mean squared error over M x N outputs: dY = 2 (Y - T) / (M N) (FP32 and
// BF16); per-block partial sums of (Y - T)^2.
__global__ __launch_bounds__(256) void k_mse(const float *Y, const float *T, long n, float scale,
                                             float *dY32, bf16_t *dY16, float *part)
{
    __shared__ float red[256];
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float e = 0.f;
    if (i < n) {
        float d = Y[i] - T[i];
        float g = 2.0f * d * scale;
        dY32[i] = g;
        dY16[i] = f2bf(g);
        e = d * d;
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[blockIdx.x] = red[0];
}

__global__ void k_sum(const float *part, int n, float scale, float *out)
{
    if (threadIdx.x != 0)
        return;
    float s = 0.f;
    for (int i = 0; i < n; i++)
        s += part[i];
    *out = s * scale;
}

// AdamW (decoupled weight decay) on FP32 master weights; refreshes the BF16
// copy used by the matrix cores.
__global__ void k_adamw(float *w, const float *g, float *m, float *v, long n, float lr, float b1,
                        float b2, float eps, float wd, float bc1, float bc2, bf16_t *w16)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.f - b1) * gi;
    float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float wi = w[i] * (1.f - lr * wd);
    wi -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    w[i] = wi;
    if (w16)
        w16[i] = f2bf(wi);
}

__global__ void k_to_bf16(const float *x, long n, bf16_t *y)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        y[i] = f2bf(x[i]);
}

unsigned blocks(long n, int t) { return (unsigned)((n + t - 1) / t); }

} // namespace

struct cosenu_mlp {
    int nl = 0, maxB = 0, device = 0;
    std::vector<int> dims;
    std::vector<float *> W, b, gW, gb, mW, vW, mb, vb; // FP32, per layer
    std::vector<bf16_t *> W16;
    bf16_t *X16 = nullptr;
    std::vector<float *> Z;      // pre-activations, per layer (maxB x dims[l+1])
    std::vector<bf16_t *> A16;   // activations, per layer
    float *Xf = nullptr, *T = nullptr, *dY32 = nullptr, *dA = nullptr, *part = nullptr;
    float *loss = nullptr;
    bf16_t *dZ16a = nullptr, *dZ16b = nullptr;
    hipStream_t s = nullptr;
    long long t = 0; // AdamW step count
    size_t nparams = 0;
    std::vector<void *> owned;
};

namespace {

template <class T>
int malloc_own(cosenu_mlp *m, T **p, size_t n)
{
    hipError_t e = hipMalloc((void **)p, n * sizeof(T) + 16);
    if (e != hipSuccess)
        return mfail(COSENU_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    m->owned.push_back(*p);
    MHIP(hipMemset(*p, 0, n * sizeof(T) + 16));
    return 0;
}

int gemm(hipStream_t s, const bf16_t *A, long am, long ak, const bf16_t *B, long bk, long bn, int M,
         int N, int K, const float *bias, float *z, long ldz, bf16_t *o, long ldo, int epi)
{
    dim3 grid((N + 31) / 32, (M + 31) / 32);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, s, A, am, ak, B, bk, bn, M, N, K, bias, z, ldz,
                       o, ldo, epi);
    MHIP(hipGetLastError());
    return 0;
}

// Forward pass on X16 (M rows): Z[l], A16[l]; the last layer has no activation.
int forward(cosenu_mlp *m, int M)
{
    const bf16_t *in = m->X16;
    for (int l = 0; l < m->nl; l++) {
        const int K = m->dims[l], N = m->dims[l + 1];
        const bool last = l == m->nl - 1;
        int rc = gemm(m->s, in, K, 1, m->W16[l], N, 1, M, N, K, m->b[l], m->Z[l], N,
                      last ? nullptr : m->A16[l], N, EPI_GELU);
        if (rc)
            return rc;
        in = m->A16[l];
    }
    return 0;
}

// Loss and gradients for targets T (M rows); leaves gW, gb filled.
int backward(cosenu_mlp *m, int M)
{
    const int L = m->nl;
    const int No = m->dims[L];
    const long n = (long)M * No;
    const unsigned nb = blocks(n, 256);
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, m->s, m->Z[L - 1], m->T, n,
                       1.0f / (float)n, m->dY32, m->dZ16a, m->part);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(64), 0, m->s, m->part, (int)nb, 1.0f / (float)n, m->loss);
    MHIP(hipGetLastError());
    float *dZ32 = m->dY32;
    bf16_t *dZ16 = m->dZ16a, *other = m->dZ16b;
    for (int l = L - 1; l >= 0; l--) {
        const int K = m->dims[l], N = m->dims[l + 1];
        const bf16_t *in = l == 0 ? m->X16 : m->A16[l - 1];
        // gW = in^T dZ  (K x N):  A(k, i) = in[i*K + k]
        int rc = gemm(m->s, in, 1, K, dZ16, N, 1, K, N, M, nullptr, m->gW[l], N, nullptr, 0, EPI_NONE);
        if (rc)
            return rc;
        hipLaunchKernelGGL(k_colsum, dim3(blocks(N, 256)), dim3(256), 0, m->s, dZ32, M, N, m->gb[l]);
        MHIP(hipGetLastError());
        if (l == 0)
            break;
        // dA = dZ W^T (M x K):  B(n, k) = W[k*N + n]
        rc = gemm(m->s, dZ16, N, 1, m->W16[l], 1, N, M, K, N, nullptr, m->dA, K, nullptr, 0, EPI_NONE);
        if (rc)
            return rc;
        const long nk = (long)M * K;
        hipLaunchKernelGGL(k_gelu_bwd, dim3(blocks(nk, 256)), dim3(256), 0, m->s, m->dA, m->Z[l - 1],
                           nk, other, m->dY32);
        MHIP(hipGetLastError());
        dZ32 = m->dY32;
        std::swap(dZ16, other);
    }
    return 0;
}

// The device address of `scratch` on `device`, which must be current.
int resolve_scratch(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device)
        g_scratch_ptr.resize(device + 1, nullptr);
    if (g_scratch_ptr[device])
        return 0;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return mfail(COSENU_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
    return 0;
}

void *scratch_on(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

int upload_batch(cosenu_mlp *m, const float *X, const float *T, int M)
{
    if (M < 1 || M > m->maxB)
        return mfail(COSENU_EINVAL, "mlp: batch size out of range");
    const long nx = (long)M * m->dims[0];
    MHIP(hipMemcpyAsync(m->Xf, X, nx * sizeof(float), hipMemcpyHostToDevice, m->s));
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks(nx, 256)), dim3(256), 0, m->s, m->Xf, nx, m->X16);
    MHIP(hipGetLastError());
    if (T)
        MHIP(hipMemcpyAsync(m->T, T, (size_t)M * m->dims[m->nl] * sizeof(float),
                            hipMemcpyHostToDevice, m->s));
    return 0;
}

} // namespace

extern "C" {

const char *cosenu_mlp_last_error(void) { return g_mlp_err.c_str(); }

int cosenu_mlp_create(const int *dims, int nlayers, int max_batch, int device, cosenu_mlp **out)
{
    if (!dims || nlayers < 1 || nlayers > 16 || max_batch < 1 || !out)
        return mfail(COSENU_EINVAL, "cosenu_mlp_create: bad argument");
    for (int l = 0; l <= nlayers; l++)
        if (dims[l] < 1 || dims[l] > 4096)
            return mfail(COSENU_EINVAL, "cosenu_mlp_create: layer width out of range");
    MHIP(hipSetDevice(device));
    {
        const int rs = resolve_scratch(device); // for cosenu_mlp_forward_noalloc
        if (rs)
            return rs;
    }
    cosenu_mlp *m = new cosenu_mlp();
    m->nl = nlayers;
    m->maxB = max_batch;
    m->device = device;
    m->dims.assign(dims, dims + nlayers + 1);
    int maxw = 0;
    for (int d : m->dims)
        maxw = std::max(maxw, d);
    int rc = 0;
    auto chk = [&](int r) {
        if (!rc)
            rc = r;
    };
    m->W.resize(nlayers); m->b.resize(nlayers); m->gW.resize(nlayers); m->gb.resize(nlayers);
    m->mW.resize(nlayers); m->vW.resize(nlayers); m->mb.resize(nlayers); m->vb.resize(nlayers);
    m->W16.resize(nlayers); m->Z.resize(nlayers); m->A16.resize(nlayers);
    for (int l = 0; l < nlayers && !rc; l++) {
        const size_t nw = (size_t)m->dims[l] * m->dims[l + 1], nb = m->dims[l + 1];
        m->nparams += nw + nb;
        chk(malloc_own(m, &m->W[l], nw)); chk(malloc_own(m, &m->b[l], nb));
        chk(malloc_own(m, &m->gW[l], nw)); chk(malloc_own(m, &m->gb[l], nb));
        chk(malloc_own(m, &m->mW[l], nw)); chk(malloc_own(m, &m->vW[l], nw));
        chk(malloc_own(m, &m->mb[l], nb)); chk(malloc_own(m, &m->vb[l], nb));
        chk(malloc_own(m, &m->W16[l], nw));
        chk(malloc_own(m, &m->Z[l], (size_t)max_batch * nb));
        chk(malloc_own(m, &m->A16[l], (size_t)max_batch * nb));
    }
    chk(malloc_own(m, &m->X16, (size_t)max_batch * m->dims[0]));
    chk(malloc_own(m, &m->Xf, (size_t)max_batch * m->dims[0]));
    chk(malloc_own(m, &m->T, (size_t)max_batch * m->dims[nlayers]));
    chk(malloc_own(m, &m->dY32, (size_t)max_batch * maxw));
    chk(malloc_own(m, &m->dA, (size_t)max_batch * maxw));
    chk(malloc_own(m, &m->dZ16a, (size_t)max_batch * maxw));
    chk(malloc_own(m, &m->dZ16b, (size_t)max_batch * maxw));
    chk(malloc_own(m, &m->part, (size_t)max_batch * maxw / 256 + 8));
    chk(malloc_own(m, &m->loss, 1));
    if (!rc && hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking) != hipSuccess)
        rc = mfail(COSENU_EHIP, "cosenu_mlp_create: stream");
    if (rc) {
        for (void *p : m->owned)
            (void)hipFree(p);
        delete m;
        return rc;
    }
    *out = m;
    return 0;
}

int cosenu_mlp_destroy(cosenu_mlp *m)
{
    if (!m)
        return 0;
    (void)hipSetDevice(m->device);
    (void)hipStreamSynchronize(m->s);
    for (void *p : m->owned)
        (void)hipFree(p);
    (void)hipStreamDestroy(m->s);
    delete m;
    return 0;
}

long long cosenu_mlp_num_params(cosenu_mlp *m) { return m ? (long long)m->nparams : -1; }

// Flat FP32 parameters, layer by layer: W_l (dims[l] x dims[l+1], row-major,
// y = x W + b), then b_l.  Setting resets the AdamW state.
int cosenu_mlp_set_params(cosenu_mlp *m, const float *flat)
{
    if (!m || !flat)
        return mfail(COSENU_EINVAL, "cosenu_mlp_set_params: null argument");
    MHIP(hipSetDevice(m->device));
    size_t o = 0;
    for (int l = 0; l < m->nl; l++) {
        const size_t nw = (size_t)m->dims[l] * m->dims[l + 1], nb = m->dims[l + 1];
        MHIP(hipMemcpyAsync(m->W[l], flat + o, nw * 4, hipMemcpyHostToDevice, m->s));
        o += nw;
        MHIP(hipMemcpyAsync(m->b[l], flat + o, nb * 4, hipMemcpyHostToDevice, m->s));
        o += nb;
        hipLaunchKernelGGL(k_to_bf16, dim3(blocks(nw, 256)), dim3(256), 0, m->s, m->W[l], (long)nw,
                           m->W16[l]);
        MHIP(hipMemsetAsync(m->mW[l], 0, nw * 4, m->s));
        MHIP(hipMemsetAsync(m->vW[l], 0, nw * 4, m->s));
        MHIP(hipMemsetAsync(m->mb[l], 0, nb * 4, m->s));
        MHIP(hipMemsetAsync(m->vb[l], 0, nb * 4, m->s));
    }
    m->t = 0;
    MHIP(hipStreamSynchronize(m->s));
    return 0;
}

static int copy_out(cosenu_mlp *m, const std::vector<float *> &Wv, const std::vector<float *> &bv,
                    float *flat)
{
    size_t o = 0;
    for (int l = 0; l < m->nl; l++) {
        const size_t nw = (size_t)m->dims[l] * m->dims[l + 1], nb = m->dims[l + 1];
        MHIP(hipMemcpyAsync(flat + o, Wv[l], nw * 4, hipMemcpyDeviceToHost, m->s));
        o += nw;
        MHIP(hipMemcpyAsync(flat + o, bv[l], nb * 4, hipMemcpyDeviceToHost, m->s));
        o += nb;
    }
    MHIP(hipStreamSynchronize(m->s));
    return 0;
}

int cosenu_mlp_get_params(cosenu_mlp *m, float *flat)
{
    if (!m || !flat)
        return mfail(COSENU_EINVAL, "cosenu_mlp_get_params: null argument");
    MHIP(hipSetDevice(m->device));
    return copy_out(m, m->W, m->b, flat);
}

// Y (M x dims[L], FP32) = network(X) for X (M x dims[0], FP32, rounded to BF16).
int cosenu_mlp_forward(cosenu_mlp *m, const float *X, int M, float *Y)
{
    if (!m || !X || !Y)
        return mfail(COSENU_EINVAL, "cosenu_mlp_forward: null argument");
    MHIP(hipSetDevice(m->device));
    int rc = upload_batch(m, X, nullptr, M);
    if (!rc)
        rc = forward(m, M);
    if (rc)
        return rc;
    MHIP(hipMemcpyAsync(Y, m->Z[m->nl - 1], (size_t)M * m->dims[m->nl] * 4, hipMemcpyDeviceToHost,
                        m->s));
    MHIP(hipStreamSynchronize(m->s));
    return 0;
}

// cosenu_mlp_forward on device arrays without allocating: the same kernels
// with the BF16 input and the hidden activations in `scratch` (laid out as
// X16 [M x dims[0]], then two activation buffers [M x widest hidden layer]
// used in turn, each rounded up to 256 bytes); the last layer writes Y.
// Enqueues only.
int cosenu_mlp_forward_noalloc(cosenu_mlp *m, const float *X, int M, float *Y, void *stream)
{
    if (!m || !X || !Y)
        return mfail(COSENU_EINVAL, "cosenu_mlp_forward_noalloc: null argument");
    if (M < 1)
        return mfail(COSENU_EINVAL, "cosenu_mlp_forward_noalloc: M must be at least 1");
    int hid = 0;
    for (int l = 1; l < m->nl; l++)
        hid = std::max(hid, m->dims[l]);
    const size_t row = align256((size_t)m->dims[0] * 2) + 2 * align256((size_t)hid * 2);
    const size_t need = align256((size_t)M * m->dims[0] * 2) + 2 * align256((size_t)M * hid * 2);
    if (need > kNoallocScratchBytes)
        return mfail(COSENU_EINVAL, "cosenu_mlp_forward_noalloc: " + std::to_string(M) +
                                        " rows need " + std::to_string(need) +
                                        " bytes of scratch; it holds " +
                                        std::to_string(kNoallocScratchBytes) + " (about " +
                                        std::to_string(kNoallocScratchBytes / row) +
                                        " rows of this network)");
    MHIP(hipSetDevice(m->device));
    unsigned char *base = static_cast<unsigned char *>(scratch_on(m->device));
    if (!base)
        return mfail(COSENU_ESTATE, "cosenu_mlp_forward_noalloc: scratch not resolved on this device");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : m->s;
    bf16_t *x16 = reinterpret_cast<bf16_t *>(base);
    bf16_t *act[2];
    act[0] = reinterpret_cast<bf16_t *>(base + align256((size_t)M * m->dims[0] * 2));
    act[1] = reinterpret_cast<bf16_t *>(reinterpret_cast<unsigned char *>(act[0]) +
                                        align256((size_t)M * hid * 2));
    const long nx = (long)M * m->dims[0];
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks(nx, 256)), dim3(256), 0, s, X, nx, x16);
    MHIP(hipGetLastError());
    const bf16_t *in = x16;
    for (int l = 0; l < m->nl; l++) {
        const int K = m->dims[l], N = m->dims[l + 1];
        const bool last = l == m->nl - 1;
        int rc = gemm(s, in, K, 1, m->W16[l], N, 1, M, N, K, m->b[l], last ? Y : nullptr, N,
                      last ? nullptr : act[l & 1], N, EPI_GELU);
        if (rc)
            return rc;
        in = act[l & 1];
    }
    return 0;
}

// Mean-squared-error loss and its gradient (flat layout of set_params), no
// update.  For gradient checks.
int cosenu_mlp_gradients(cosenu_mlp *m, const float *X, const float *T, int M, float *grad,
                         float *loss)
{
    if (!m || !X || !T || !grad)
        return mfail(COSENU_EINVAL, "cosenu_mlp_gradients: null argument");
    MHIP(hipSetDevice(m->device));
    int rc = upload_batch(m, X, T, M);
    if (!rc) rc = forward(m, M);
    if (!rc) rc = backward(m, M);
    if (rc)
        return rc;
    if (loss)
        MHIP(hipMemcpyAsync(loss, m->loss, 4, hipMemcpyDeviceToHost, m->s));
    return copy_out(m, m->gW, m->gb, grad);
}

// One AdamW step on the mean-squared error of (X, T).
int cosenu_mlp_train_step(cosenu_mlp *m, const float *X, const float *T, int M, float lr,
                          float beta1, float beta2, float eps, float weight_decay, float *loss)
{
    if (!m || !X || !T)
        return mfail(COSENU_EINVAL, "cosenu_mlp_train_step: null argument");
    MHIP(hipSetDevice(m->device));
    int rc = upload_batch(m, X, T, M);
    if (!rc) rc = forward(m, M);
    if (!rc) rc = backward(m, M);
    if (rc)
        return rc;
    m->t += 1;
    const float bc1 = 1.0f - powf(beta1, (float)m->t), bc2 = 1.0f - powf(beta2, (float)m->t);
    for (int l = 0; l < m->nl; l++) {
        const long nw = (long)m->dims[l] * m->dims[l + 1], nb = m->dims[l + 1];
        hipLaunchKernelGGL(k_adamw, dim3(blocks(nw, 256)), dim3(256), 0, m->s, m->W[l], m->gW[l],
                           m->mW[l], m->vW[l], nw, lr, beta1, beta2, eps, weight_decay, bc1, bc2,
                           m->W16[l]);
        // no weight decay on biases
        hipLaunchKernelGGL(k_adamw, dim3(blocks(nb, 256)), dim3(256), 0, m->s, m->b[l], m->gb[l],
                           m->mb[l], m->vb[l], nb, lr, beta1, beta2, eps, 0.0f, bc1, bc2,
                           (bf16_t *)nullptr);
    }
    MHIP(hipGetLastError());
    if (loss) {
        MHIP(hipMemcpyAsync(loss, m->loss, 4, hipMemcpyDeviceToHost, m->s));
        MHIP(hipStreamSynchronize(m->s));
    }
    return 0;
}

int cosenu_mlp_sync(cosenu_mlp *m)
{
    if (!m)
        return mfail(COSENU_EINVAL, "cosenu_mlp_sync: null argument");
    MHIP(hipSetDevice(m->device));
    MHIP(hipStreamSynchronize(m->s));
    return 0;
}

} // extern "C"
