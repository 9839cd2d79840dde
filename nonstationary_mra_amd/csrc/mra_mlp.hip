// Warm-start network for the nonstationary parameters: an MLP
// 64 -> 256 -> 256 -> 6 with GELU, trained in FP32 on the GPU with its own
// forward, backward and Adam kernels; inference optionally with weights and
// activations rounded to bfloat16 on the bfloat16 matrix instructions (FP32
// accumulation).  No reference
// counterpart: the reference starts its optimizer from a fixed guess
// (src/optimization.cpp:160-169).
//
// Layer l: Z = X W^T + b (X: batch x in, W: out x in, row-major), A = gelu(Z)
// for the two hidden layers, Y = A2 W3^T + b3.  Loss: mean squared error over
// batch and outputs.  GEMMs run on v_mfma_f32_16x16x4_f32, one 64 x 64 output
// tile per work group.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mra.h"
#include "mra_device.hpp"

namespace mlp {

constexpr int NIN = 64, NH = 256, NOUT = 6;
constexpr int MAXB = 1024;
// flat parameter layout: W1, b1, W2, b2, W3, b3
constexpr int OW1 = 0, OB1 = OW1 + NH * NIN, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NOUT * NH, NPARAM = OB3 + NOUT;

// C (M x N, row-major ldc) = op(A) op(B) with
//   TA = false: A(k,i) = A[k*lda+i];  TA = true: A(k,i) = A[i*lda+k]
//   TB = false: B(k,j) = B[k*ldb+j];  TB = true: B(k,j) = B[j*ldb+k]
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const float* A, int lda, const float* B, int ldb,
                                              float* C, int ldc) {
    const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
    const int mv = min(64, M - i0), nv = min(64, N - j0);
    const float* Ab = TA ? A + (int64_t)i0 * lda : A + i0;
    const float* Bb = TB ? B + (int64_t)j0 * ldb : B + j0;
    mra::wg_gemm_tn_impl<float, TA, false, float, float, mra::GEMM_ANY, float, TB>(
        C + (int64_t)i0 * ldc + j0, ldc, (mv + 15) & ~15, (nv + 15) & ~15, K, 1.0f, Ab, lda, mv, Bb, ldb, nv);
}

// C[i][j] = sum_k A[i][k] B[j][k] (M x N, both operands k-contiguous) on the
// bfloat16 matrix instruction v_mfma_f32_16x16x16_bf16 (FP32 accumulation),
// for inference on operands already rounded to bfloat16 (the upper 16 bits of
// each float are the bfloat16 value).  K is a multiple of 16.  One 64 x 64
// output tile per work group, 32 x 32 per wave.  Fragments: lane t holds
// A[t%16][4(t/16) .. 4(t/16)+3] and B[4(t/16) .. +3][t%16]; D element v of
// lane t is row 4(t/16) + v, column t%16.
typedef short mra_s4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ mra_s4 bf16x4(const float* p) {
    const float4 v = *reinterpret_cast<const float4*>(p);
    return mra_s4{(short)(__float_as_uint(v.x) >> 16), (short)(__float_as_uint(v.y) >> 16),
                  (short)(__float_as_uint(v.z) >> 16), (short)(__float_as_uint(v.w) >> 16)};
}
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const float* A, int lda, const float* B,
                                                   int ldb, float* C, int ldc) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r16 = lane & 15, k4 = lane >> 4;
    const int i0 = blockIdx.y * 64 + (w >> 1) * 32, j0 = blockIdx.x * 64 + (w & 1) * 32;
    int ri[2], cj[2];
    bool vi[2], vj[2];
    for (int t = 0; t < 2; t++) {
        ri[t] = i0 + 16 * t + r16;
        cj[t] = j0 + 16 * t + r16;
        vi[t] = ri[t] < M;
        vj[t] = cj[t] < N;
        ri[t] = vi[t] ? ri[t] : 0;  // clamped rows are read and their values dropped
        cj[t] = vj[t] ? cj[t] : 0;
    }
    mra_f4 acc[2][2] = {};
    const mra_s4 z = {0, 0, 0, 0};
    for (int k0 = 0; k0 < K; k0 += 16) {
        mra_s4 a[2], b[2];
        for (int t = 0; t < 2; t++) {
            a[t] = vi[t] ? bf16x4(A + (int64_t)ri[t] * lda + k0 + 4 * k4) : z;
            b[t] = vj[t] ? bf16x4(B + (int64_t)cj[t] * ldb + k0 + 4 * k4) : z;
        }
        for (int ti = 0; ti < 2; ti++)
            for (int tj = 0; tj < 2; tj++)
                acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
    }
    for (int ti = 0; ti < 2; ti++)
        for (int tj = 0; tj < 2; tj++)
            for (int v = 0; v < 4; v++) {
                const int i = i0 + 16 * ti + 4 * k4 + v, j = j0 + 16 * tj + r16;
                if (i < M && j < N) C[(int64_t)i * ldc + j] = acc[ti][tj][v];
            }
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * expf(-0.5f * z * z) * 0.39894228040143268f;
}
// round to the nearest bfloat16 (ties to even), kept in a float
__device__ __forceinline__ float to_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;  // inf / nan unchanged
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

// Z += b (per column); A = gelu(Z) (act) or A = Z; optional bf16 rounding of A
__global__ void k_bias_act(int B, int N, float* Z, const float* b, float* A, int act, int bf16) {
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= (int64_t)B * N) return;
    const float z = Z[idx] + b[idx % N];
    Z[idx] = z;
    float a = act ? gelu(z) : z;
    if (bf16) a = to_bf16(a);
    A[idx] = a;
}

__global__ void k_round_bf16(int64_t n, const float* x, float* y) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) y[i] = to_bf16(x[i]);
}

// dY = 2 (Y - T) / (B * NOUT); per-block partial sums of (Y - T)^2
__global__ void k_loss(int B, const float* Y, const float* T, float* dY, float* part) {
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < B * NOUT; i += gridDim.x * 256) {
        const float e = Y[i] - T[i];
        dY[i] = 2.0f * e / (float)(B * NOUT);
        s += e * e;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// dZ = dA * gelu'(Z)
__global__ void k_act_back(int64_t n, const float* Z, const float* dA, float* dZ) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) dZ[i] = dA[i] * gelu_d(Z[i]);
}

// db[j] = sum_b dZ[b][j] (fixed order)
__global__ void k_colsum(int B, int N, const float* dZ, float* db) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    float s = 0.0f;
    for (int b = 0; b < B; b++) s += dZ[(int64_t)b * N + j];
    db[j] = s;
}

__global__ void k_adam(int n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace mlp

using namespace mlp;

struct mra_mlp {
    hipStream_t st = nullptr;
    float *P = nullptr, *Pb = nullptr, *G = nullptr, *Mo = nullptr, *Vo = nullptr;
    float *X = nullptr, *T = nullptr, *Z1 = nullptr, *A1 = nullptr, *Z2 = nullptr, *A2 = nullptr, *Y = nullptr,
          *dY = nullptr, *dA = nullptr, *dZ = nullptr, *part = nullptr;
    int64_t step = 0;
    std::string err;
    std::vector<void*> allocs;
};

namespace {
std::string g_mlp_err;

int fail(mra_mlp* m, const std::string& s, int code) {
    if (m) m->err = s;
    g_mlp_err = s;
    return code;
}

#define MLP_HIP(m, x)                                                              \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) return fail(m, std::string(#x) + ": " + hipGetErrorString(e_), MRA_E_HIP); \
    } while (0)

inline unsigned nb(int64_t n) { return (unsigned)((n + 255) / 256); }

template <bool TA, bool TB>
void gemm(hipStream_t st, int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc) {
    dim3 grid((N + 63) / 64, (M + 63) / 64);
    k_gemm<TA, TB><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc);
}

// C = A B^T for row-major A (M x K) and B (N x K): FP32 matrix instructions,
// or (bf16, operands already rounded to bfloat16) the bfloat16 ones
void gemm_nt(hipStream_t st, bool bf16, int M, int N, int K, const float* A, int lda, const float* B, int ldb,
             float* C, int ldc) {
    if (bf16) {
        dim3 grid((N + 63) / 64, (M + 63) / 64);
        k_gemm_bf16<<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc);
    } else {
        gemm<true, true>(st, M, N, K, A, lda, B, ldb, C, ldc);
    }
}

// forward pass on the first B rows of m->X; bf16 selects rounded weights (Pb)
// and activations on the bfloat16 matrix instructions
void forward(mra_mlp* m, int B, bool bf16) {
    const float* P = bf16 ? m->Pb : m->P;
    hipStream_t st = m->st;
    if (bf16) k_round_bf16<<<nb((int64_t)B * NIN), 256, 0, st>>>((int64_t)B * NIN, m->X, m->X);
    // Z1[b][o] = sum_k X[b][k] W1[o][k]
    gemm_nt(st, bf16, B, NH, NIN, m->X, NIN, P + OW1, NIN, m->Z1, NH);
    k_bias_act<<<nb((int64_t)B * NH), 256, 0, st>>>(B, NH, m->Z1, P + OB1, m->A1, 1, bf16);
    gemm_nt(st, bf16, B, NH, NH, m->A1, NH, P + OW2, NH, m->Z2, NH);
    k_bias_act<<<nb((int64_t)B * NH), 256, 0, st>>>(B, NH, m->Z2, P + OB2, m->A2, 1, bf16);
    gemm_nt(st, bf16, B, NOUT, NH, m->A2, NH, P + OW3, NH, m->Y, NOUT);
    k_bias_act<<<nb((int64_t)B * NOUT), 256, 0, st>>>(B, NOUT, m->Y, P + OB3, m->Y, 0, 0);
}

// loss and gradient (into m->G) for the first B rows of X, T
void backward(mra_mlp* m, int B) {
    hipStream_t st = m->st;
    float* G = m->G;
    const float* P = m->P;
    k_loss<<<8, 256, 0, st>>>(B, m->Y, m->T, m->dY, m->part);
    // layer 3: dW3[o][k] = sum_b dY[b][o] A2[b][k]; db3; dA2 = dY W3
    gemm<false, false>(st, NOUT, NH, B, m->dY, NOUT, m->A2, NH, G + OW3, NH);
    k_colsum<<<1, 256, 0, st>>>(B, NOUT, m->dY, G + OB3);
    gemm<true, false>(st, B, NH, NOUT, m->dY, NOUT, P + OW3, NH, m->dA, NH);
    k_act_back<<<nb((int64_t)B * NH), 256, 0, st>>>((int64_t)B * NH, m->Z2, m->dA, m->dZ);
    // layer 2
    gemm<false, false>(st, NH, NH, B, m->dZ, NH, m->A1, NH, G + OW2, NH);
    k_colsum<<<1, 256, 0, st>>>(B, NH, m->dZ, G + OB2);
    gemm<true, false>(st, B, NH, NH, m->dZ, NH, P + OW2, NH, m->dA, NH);
    k_act_back<<<nb((int64_t)B * NH), 256, 0, st>>>((int64_t)B * NH, m->Z1, m->dA, m->dZ);
    // layer 1
    gemm<false, false>(st, NH, NIN, B, m->dZ, NH, m->X, NIN, G + OW1, NIN);
    k_colsum<<<1, 256, 0, st>>>(B, NH, m->dZ, G + OB1);
}

int upload_batch(mra_mlp* m, const float* X, const float* T, int B) {
    if (B < 1 || B > MAXB) return fail(m, "batch size must be in [1, 1024]", MRA_E_ARG);
    MLP_HIP(m, hipMemcpyAsync(m->X, X, sizeof(float) * B * NIN, hipMemcpyHostToDevice, m->st));
    if (T) MLP_HIP(m, hipMemcpyAsync(m->T, T, sizeof(float) * B * NOUT, hipMemcpyHostToDevice, m->st));
    return MRA_OK;
}

int read_loss(mra_mlp* m, float* loss, int B) {
    float part[8];
    MLP_HIP(m, hipMemcpyAsync(part, m->part, sizeof(part), hipMemcpyDeviceToHost, m->st));
    MLP_HIP(m, hipStreamSynchronize(m->st));
    double s = 0.0;
    for (int i = 0; i < 8; i++) s += part[i];
    if (loss) *loss = (float)(s / (double)(B * NOUT));
    return MRA_OK;
}

// splitmix64 -> uniform in (-1, 1)
double unif(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((z >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
}
}  // namespace

extern "C" int mra_mlp_param_count(void) { return NPARAM; }

extern "C" const char* mra_mlp_last_error(const mra_mlp* m) { return m ? m->err.c_str() : g_mlp_err.c_str(); }

extern "C" void mra_mlp_destroy(mra_mlp* m) {
    if (!m) return;
    if (m->st) (void)hipStreamSynchronize(m->st);
    for (void* p : m->allocs) (void)hipFree(p);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
}

extern "C" int mra_mlp_create(uint64_t seed, mra_mlp** out) {
    if (!out) return MRA_E_ARG;
    *out = nullptr;
    mra_mlp* m = new mra_mlp();
    auto al = [&](float** p, size_t n) -> int {
        hipError_t e = hipMalloc((void**)p, sizeof(float) * n);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return MRA_E_NOMEM;
        }
        m->allocs.push_back(*p);
        return MRA_OK;
    };
    int rc = MRA_OK;
    if (hipStreamCreate(&m->st) != hipSuccess) rc = MRA_E_HIP;
    float** bufs[] = {&m->P, &m->Pb, &m->G, &m->Mo, &m->Vo};
    for (float** b : bufs)
        if (!rc) rc = al(b, NPARAM);
    float** acts[] = {&m->Z1, &m->A1, &m->Z2, &m->A2, &m->dA, &m->dZ};
    for (float** b : acts)
        if (!rc) rc = al(b, (size_t)MAXB * NH);
    if (!rc) rc = al(&m->X, (size_t)MAXB * NIN);
    if (!rc) rc = al(&m->T, (size_t)MAXB * NOUT);
    if (!rc) rc = al(&m->Y, (size_t)MAXB * NOUT);
    if (!rc) rc = al(&m->dY, (size_t)MAXB * NOUT);
    if (!rc) rc = al(&m->part, 8);
    if (rc) {
        g_mlp_err = "mra_mlp_create: device allocation failed";
        mra_mlp_destroy(m);
        return rc;
    }
    // initial weights: uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)), zero biases
    std::vector<float> h(NPARAM, 0.0f);
    uint64_t s = seed;
    for (int i = 0; i < NH * NIN; i++) h[OW1 + i] = (float)(unif(s) / std::sqrt((double)NIN));
    for (int i = 0; i < NH * NH; i++) h[OW2 + i] = (float)(unif(s) / std::sqrt((double)NH));
    for (int i = 0; i < NOUT * NH; i++) h[OW3 + i] = (float)(unif(s) / std::sqrt((double)NH));
    std::vector<float> z(NPARAM, 0.0f);
    if (hipMemcpy(m->P, h.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->Mo, z.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->Vo, z.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice) != hipSuccess) {
        g_mlp_err = "mra_mlp_create: upload failed";
        mra_mlp_destroy(m);
        return MRA_E_HIP;
    }
    *out = m;
    return MRA_OK;
}

extern "C" int mra_mlp_set_params(mra_mlp* m, const float* p) {
    if (!m || !p) return MRA_E_ARG;
    MLP_HIP(m, hipMemcpy(m->P, p, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    m->step = 0;
    std::vector<float> z(NPARAM, 0.0f);
    MLP_HIP(m, hipMemcpy(m->Mo, z.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    MLP_HIP(m, hipMemcpy(m->Vo, z.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    return MRA_OK;
}

extern "C" int mra_mlp_get_params(mra_mlp* m, float* p) {
    if (!m || !p) return MRA_E_ARG;
    MLP_HIP(m, hipMemcpy(p, m->P, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
    return MRA_OK;
}

extern "C" int mra_mlp_forward(mra_mlp* m, const float* X, int batch, float* Y, int bf16) {
    if (!m || !X || !Y) return MRA_E_ARG;
    int rc = upload_batch(m, X, nullptr, batch);
    if (rc) return rc;
    if (bf16) k_round_bf16<<<nb(NPARAM), 256, 0, m->st>>>(NPARAM, m->P, m->Pb);
    forward(m, batch, bf16 != 0);
    MLP_HIP(m, hipGetLastError());
    MLP_HIP(m, hipMemcpyAsync(Y, m->Y, sizeof(float) * batch * NOUT, hipMemcpyDeviceToHost, m->st));
    MLP_HIP(m, hipStreamSynchronize(m->st));
    return MRA_OK;
}

extern "C" int mra_mlp_loss_grad(mra_mlp* m, const float* X, const float* T, int batch, float* loss, float* grad) {
    if (!m || !X || !T) return MRA_E_ARG;
    int rc = upload_batch(m, X, T, batch);
    if (rc) return rc;
    forward(m, batch, false);
    backward(m, batch);
    MLP_HIP(m, hipGetLastError());
    if (grad) MLP_HIP(m, hipMemcpyAsync(grad, m->G, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, m->st));
    return read_loss(m, loss, batch);
}

extern "C" int mra_mlp_train_step(mra_mlp* m, const float* X, const float* T, int batch, float lr, float* loss) {
    if (!m || !X || !T || !(lr > 0)) return MRA_E_ARG;
    int rc = upload_batch(m, X, T, batch);
    if (rc) return rc;
    forward(m, batch, false);
    backward(m, batch);
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
    k_adam<<<nb(NPARAM), 256, 0, m->st>>>(NPARAM, m->P, m->G, m->Mo, m->Vo, lr, b1, b2, 1e-8f, c1, c2);
    MLP_HIP(m, hipGetLastError());
    return read_loss(m, loss, batch);
}
