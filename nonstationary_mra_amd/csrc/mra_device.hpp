// Device building blocks for the MRA kernels (gfx950 / CDNA4).
//
// Work-group conventions: 256 threads = 4 waves of 64.  Small dense blocks
// (r x r knot blocks, r <= 64, padded to RP = 16/32/64) live in LDS with a
// row stride of RP+1 doubles (odd stride: no LDS bank conflicts on column
// walks).  Dense products use the FP64 matrix core instruction
// v_mfma_f64_16x16x4_f64 through __builtin_amdgcn_mfma_f64_16x16x4f64:
//   A fragment (16x4): lane t holds A[t%16][t/16]
//   B fragment (4x16): lane t holds B[t/16][t%16]
//   D (16x16):         lane t, element v holds D[t/16 + 4*v][t%16]
//   (the D layout was confirmed on gfx950 by tests/test_gpu_blocks.py)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MRA_WG 256

typedef double mra_d4 __attribute__((ext_vector_type(4)));
typedef float mra_f4 __attribute__((ext_vector_type(4)));

namespace mra {

// ---------------------------------------------------------------- geometry
// Point sets are SoA on the device.  Coordinates per metric:
//   chordal: unit-sphere x, y, z;  planar: lon, lat, 0;
//   great circle: sin(lat), cos(lat), lon (radians).
// s = sqrt(scaled sill), b = scaled range, v = scaled sill + scaled nugget
// (the diagonal of a variance-covariance matrix).
struct Pts {
    const double *x, *y, *z, *s, *b, *v;
};

struct CovCtx {
    int metric;      // 0 chordal, 1 great circle, 2 planar
    int stationary;  // 1: sill * exp(-d / range)
    double sill, range;
};

__device__ __forceinline__ double pdist(const CovCtx& c, double x1, double y1, double z1, double x2, double y2,
                                        double z2) {
    if (c.metric == 1) {
        // src/evaluate_covariance.cpp:9-25
        double ang = x1 * x2 + y1 * y2 * cos(z1 - z2);
        ang = fmin(1.0, fmax(-1.0, ang));
        return 6371.0 * acos(ang);
    }
    // src/evaluate_covariance.cpp:28-37 (chordal x 6371) / planar
    double dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
    double d = sqrt(dx * dx + dy * dy + dz * dz);
    return c.metric == 0 ? 6371.0 * d : d;
}

// Cross-covariance of points i (set P) and j (set Q), no nugget
// (src/evaluate_covariance.cpp:207-223).
__device__ __forceinline__ double pcov(const CovCtx& c, const Pts& P, uint64_t i, const Pts& Q, uint64_t j) {
    double d = pdist(c, P.x[i], P.y[i], P.z[i], Q.x[j], Q.y[j], Q.z[j]);
    if (c.stationary) return c.sill * exp(-d / c.range);
    double b1 = P.b[i], b2 = Q.b[j];
    double bms = (b1 * b1 + b2 * b2) * 0.5;
    double ratio = (b1 * b2) / bms;
    return (P.s[i] * Q.s[j]) * (ratio * sqrt(ratio)) * exp(-d / sqrt(bms));
}

// ---------------------------------------------------- covariance tiles
// 1/sqrt(x) for x > 0: hardware estimate refined by two Newton steps
// y <- y + y/2 (1 - x y^2) (each step squares the relative error).
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double e = __builtin_fma(-x * y, y, 1.0);
    y = __builtin_fma(0.5 * y, e, y);
    e = __builtin_fma(-x * y, y, 1.0);
    return __builtin_fma(0.5 * y, e, y);
}

// Point data of one tile side: coordinates, q = s b^1.5 and b.  With
// q_i = s_i b_i^1.5 and t = 1/sqrt(bms), bms = (b_i^2 + b_j^2)/2, the
// nonstationary covariance s_i s_j (b_i b_j / bms)^1.5 exp(-d / sqrt(bms))
// (src/evaluate_covariance.cpp:207-223) is q_i q_j t^3 exp(-d t).
struct PtV {
    double x, y, z, q, b;
};

__device__ __forceinline__ PtV load_pt(const CovCtx& c, const Pts& P, uint64_t i) {
    PtV p;
    p.x = P.x[i];
    p.y = P.y[i];
    p.z = P.z[i];
    p.b = P.b[i];
    p.q = c.stationary ? 0.0 : P.s[i] * (p.b * sqrt(p.b));
    return p;
}

__device__ __forceinline__ double pcov_pt(const CovCtx& c, const PtV& a, const PtV& b) {
    double d;
    if (c.metric == 1) {
        d = pdist(c, a.x, a.y, a.z, b.x, b.y, b.z);
    } else {
        const double dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
        const double d2 = dx * dx + dy * dy + dz * dz;
        d = d2 > 0.0 ? d2 * rsqrt_nr(d2) : 0.0;
        if (c.metric == 0) d *= 6371.0;
    }
    if (c.stationary) return c.sill * exp(-d / c.range);
    const double t = rsqrt_nr((a.b * a.b + b.b * b.b) * 0.5);
    return (a.q * b.q) * (t * t * t) * exp(-d * t);
}

// FP32 evaluation of the same form (mixed precision): inputs rounded to
// float, hardware rsqrt (about 1 ulp) and expf.
__device__ __forceinline__ float pcov_ptf(const CovCtx& c, const PtV& a, const PtV& b) {
    float d;
    if (c.metric == 1) {
        d = (float)pdist(c, a.x, a.y, a.z, b.x, b.y, b.z);
    } else {
        const float dx = (float)b.x - (float)a.x, dy = (float)b.y - (float)a.y, dz = (float)b.z - (float)a.z;
        const float d2 = dx * dx + dy * dy + dz * dz;
        d = d2 > 0.0f ? d2 * __builtin_amdgcn_rsqf(d2) : 0.0f;
        if (c.metric == 0) d *= 6371.0f;
    }
    if (c.stationary) return (float)c.sill * expf(-d / (float)c.range);
    const float ab = (float)a.b, bb = (float)b.b;
    const float t = __builtin_amdgcn_rsqf((ab * ab + bb * bb) * 0.5f);
    return ((float)a.q * (float)b.q) * (t * t * t) * expf(-d * t);
}

// Fill the RP x NC tile C[i*LD + c] (i < RP, c < NC) with the covariance of
// row points P[p0 + i] (valid i < nr) and column points Q[q0 + c] (valid
// c < nc), zero outside.  diagv: rows and columns are the same point set and
// the diagonal takes the variance v = sigma^2 + tau^2 instead.  Row data are
// staged in `stage` (LDS, 5 * RP doubles); the caller synchronises after.
template <int RP, int NC, typename T>
__device__ __forceinline__ void cov_tile(T* C, int LD, const CovCtx& cc, const Pts& P, uint64_t p0, int nr,
                                         const Pts& Q, uint64_t q0, int nc, bool diagv, double* stage) {
    static_assert(MRA_WG % NC == 0, "tile width must divide the work-group size");
    for (int i = threadIdx.x; i < RP; i += MRA_WG) {
        if (i < nr) {
            const PtV p = load_pt(cc, P, p0 + i);
            stage[i] = p.x;
            stage[RP + i] = p.y;
            stage[2 * RP + i] = p.z;
            stage[3 * RP + i] = p.q;
            stage[4 * RP + i] = p.b;
        }
    }
    const int c = threadIdx.x % NC;
    PtV qc = {0, 0, 0, 0, 1};
    if (c < nc) qc = load_pt(cc, Q, q0 + c);
    __syncthreads();
    for (int i = threadIdx.x / NC; i < RP; i += MRA_WG / NC) {
        T v = T(0);
        if (diagv && i == c) {
            v = i < nr ? (T)P.v[p0 + i] : T(1);  // padding rows: identity
        } else if (i < nr && c < nc) {
            const PtV pr = {stage[i], stage[RP + i], stage[2 * RP + i], stage[3 * RP + i], stage[4 * RP + i]};
            if constexpr (sizeof(T) == 4)
                v = pcov_ptf(cc, pr, qc);
            else
                v = pcov_pt(cc, pr, qc);
        }
        C[i * LD + c] = v;
    }
}

__device__ __forceinline__ float pcovf(const CovCtx& c, const Pts& P, uint64_t i, const Pts& Q, uint64_t j) {
    float x1 = (float)P.x[i], y1 = (float)P.y[i], z1 = (float)P.z[i];
    float x2 = (float)Q.x[j], y2 = (float)Q.y[j], z2 = (float)Q.z[j];
    float d;
    if (c.metric == 1) {
        float ang = x1 * x2 + y1 * y2 * cosf(z1 - z2);
        ang = fminf(1.0f, fmaxf(-1.0f, ang));
        d = 6371.0f * acosf(ang);
    } else {
        float dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
        d = sqrtf(dx * dx + dy * dy + dz * dz);
        if (c.metric == 0) d *= 6371.0f;
    }
    if (c.stationary) return (float)c.sill * expf(-d / (float)c.range);
    float b1 = (float)P.b[i], b2 = (float)Q.b[j];
    float bms = (b1 * b1 + b2 * b2) * 0.5f;
    float ratio = (b1 * b2) / bms;
    return ((float)P.s[i] * (float)Q.s[j]) * (ratio * sqrtf(ratio)) * expf(-d / sqrtf(bms));
}

// ----------------------------------------------------------- MFMA GEMM TN
// C[i*ldc+j] += alpha * sum_k A(k,i) B(k,j) over the M x N tile (multiples
// of 16).  A(k,i) = A[k*lda+i] valid for k < Kv, i < Mv (else 0); B likewise
// with Nv.  K is rounded up to a multiple of 4.  C is LDS (or global) and
// must be made visible by the caller (__syncthreads before and after).
// TA: A is stored transposed, A(k,i) = A[i*lda+k] (an LDS matrix read by
// columns).  ACC = false stores C = alpha*A^T B instead of accumulating.
// TILES: GEMM_ANY checks per wave block whether every operand is in range
// and then takes the predication-free path; GEMM_FULL asserts that
// Mv >= M and Nv >= N (only that path is compiled); GEMM_PART compiles only
// the predicated path (smaller register footprint).
enum { GEMM_ANY = 0, GEMM_FULL = 1, GEMM_PART = 2 };
template <bool TA = false, bool ACC = true, typename TIA = double, typename TIB = double, int TILES = GEMM_PART>
__device__ __forceinline__ void wg_gemm_tn(double* C, int ldc, int M, int N, int Kv, double alpha, const TIA* A,
                                           int64_t lda, int Mv, const TIB* B, int64_t ldb, int Nv) {
    static_assert(!(TA && TILES == GEMM_FULL), "the full-tile path reads A untransposed");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, k4 = lane >> 4;
    const int tM = M >> 4, tN = N >> 4;
    // per-wave blocks of bh x bw tiles: 2x2 when that still gives every wave a
    // block, else 2x1 or 1x1 (all wave-uniform)
    const int nt = tM * tN;
    const int bh = (nt >= 16 || (nt >= 8 && tM >= 2)) ? 2 : 1;
    const int bw = (nt >= 16) ? 2 : 1;
    const int bM = (tM + bh - 1) / bh, bN = (tN + bw - 1) / bw;
    for (int blk = wave; blk < bM * bN; blk += 4) {
        const int ti0 = (blk / bN) * bh, tj0 = (blk % bN) * bw;
        mra_d4 acc00 = {0, 0, 0, 0}, acc01 = acc00, acc10 = acc00, acc11 = acc00;
        const int ia = ti0 * 16 + r16, ib = ia + 16;
        const int ja = tj0 * 16 + r16, jb = ja + 16;
        const bool tM2 = bh == 2 && ti0 + 1 < tM, tN2 = bw == 2 && tj0 + 1 < tN;  // second tile row / column
        const bool ha = ia < Mv, hb = tM2 && ib < Mv;
        const bool hja = ja < Nv, hjb = tN2 && jb < Nv;
        // Operand loads of U k-steps are issued together before their MFMAs so
        // that one memory latency covers U steps (all control flow is
        // uniform across the wave).
        constexpr int U = 8;
        int k0 = 0;
        const bool full = TILES == GEMM_FULL ||
                          (TILES == GEMM_ANY && !TA && ti0 * 16 + (tM2 ? 32 : 16) <= Mv && tj0 * 16 + (tN2 ? 32 : 16) <= Nv);
        if (TILES != GEMM_PART && full) {
            // every operand in range: wave-uniform row pointers plus 32-bit
            // per-lane byte offsets, no predication
            const uint32_t oa0 = (uint32_t)((k4 * lda + ia) * sizeof(TIA)), oa1 = oa0 + 16 * sizeof(TIA);
            const uint32_t ob0 = (uint32_t)((k4 * ldb + ja) * sizeof(TIB)), ob1 = ob0 + 16 * sizeof(TIB);
            for (; k0 + 4 * U <= Kv; k0 += 4 * U) {
                TIA a0[U], a1[U];
                TIB b0[U], b1[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const char* Ak = (const char*)(A + (int64_t)(k0 + 4 * u) * lda);
                    const char* Bk = (const char*)(B + (int64_t)(k0 + 4 * u) * ldb);
                    a0[u] = *(const TIA*)(Ak + oa0);
                    if (tM2) a1[u] = *(const TIA*)(Ak + oa1);
                    b0[u] = *(const TIB*)(Bk + ob0);
                    if (tN2) b1[u] = *(const TIB*)(Bk + ob1);
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const double xa0 = (double)a0[u], xb0 = (double)b0[u];
                    acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(xa0, xb0, acc00, 0, 0, 0);
                    if (tN2) acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(xa0, (double)b1[u], acc01, 0, 0, 0);
                    if (tM2) acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a1[u], xb0, acc10, 0, 0, 0);
                    if (tM2 && tN2)
                        acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64((double)a1[u], (double)b1[u], acc11, 0, 0, 0);
                }
            }
        }
        constexpr int UP = 4;
        for (; TILES != GEMM_FULL && k0 + 4 * UP <= Kv; k0 += 4 * UP) {
            // raw operands (input type) first, widened at use, so that a
            // conversion does not wait on its load before the next load issues
            TIA a0[UP], a1[UP];
            TIB b0[UP], b1[UP];
#pragma unroll
            for (int u = 0; u < UP; u++) {
                const int64_t k = k0 + 4 * u + k4;
                a0[u] = ha ? (TA ? A[ia * lda + k] : A[k * lda + ia]) : TIA(0);
                a1[u] = hb ? (TA ? A[ib * lda + k] : A[k * lda + ib]) : TIA(0);
                b0[u] = hja ? B[k * ldb + ja] : TIB(0);
                b1[u] = hjb ? B[k * ldb + jb] : TIB(0);
            }
#pragma unroll
            for (int u = 0; u < UP; u++) {
                const double xa0 = (double)a0[u], xa1 = (double)a1[u], xb0 = (double)b0[u], xb1 = (double)b1[u];
                acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(xa0, xb0, acc00, 0, 0, 0);
                if (tN2) acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(xa0, xb1, acc01, 0, 0, 0);
                if (tM2) acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(xa1, xb0, acc10, 0, 0, 0);
                if (tM2 && tN2) acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(xa1, xb1, acc11, 0, 0, 0);
            }
        }
        for (; k0 < Kv; k0 += 4) {
            const int k = k0 + k4;
            const bool kv = k < Kv;
            double a0 = (kv && ha) ? (double)(TA ? A[ia * lda + k] : A[(int64_t)k * lda + ia]) : 0.0;
            double a1 = (kv && hb) ? (double)(TA ? A[ib * lda + k] : A[(int64_t)k * lda + ib]) : 0.0;
            double b0 = (kv && hja) ? (double)B[(int64_t)k * ldb + ja] : 0.0;
            double b1 = (kv && hjb) ? (double)B[(int64_t)k * ldb + jb] : 0.0;
            acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc00, 0, 0, 0);
            if (tN2) acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc01, 0, 0, 0);
            if (tM2) acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc10, 0, 0, 0);
            if (tM2 && tN2) acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc11, 0, 0, 0);
        }
        // only the valid Mv x Nv part of C is updated; element v of lane t
        // holds row t/16 + 4v of the 16 x 16 tile
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = ti0 * 16 + k4 + 4 * v;
            const bool r0 = row < Mv, r1 = tM2 && (row + 16 < Mv);
            double* c00 = C + (int64_t)row * ldc + ja;
            double* c10 = C + (int64_t)(row + 16) * ldc + ja;
            if (ACC) {
                if (r0 && hja) c00[0] += alpha * acc00[v];
                if (r0 && hjb) c00[16] += alpha * acc01[v];
                if (r1 && hja) c10[0] += alpha * acc10[v];
                if (r1 && hjb) c10[16] += alpha * acc11[v];
            } else {
                if (r0 && hja) c00[0] = alpha * acc00[v];
                if (r0 && hjb) c00[16] = alpha * acc01[v];
                if (r1 && hja) c10[0] = alpha * acc10[v];
                if (r1 && hjb) c10[16] = alpha * acc11[v];
            }
        }
    }
}

// FP32 variant: C (float) += alpha * A^T B with v_mfma_f32_16x16x4_f32 and
// FP32 accumulation; operands are read as TIA / TIB (double or float) and
// rounded to float.  Same tiling and validity rules as wg_gemm_tn; the f32 D
// fragment has element v of lane t at row 4*(t/16) + v (unlike the f64
// instruction; confirmed on gfx950 by tests/test_gpu_blocks.py).
template <bool ACC = true, typename TIA = float, typename TIB = float, typename TC = float>
__device__ __forceinline__ void wg_gemm_tn_f32(TC* C, int ldc, int M, int N, int Kv, float alpha, const TIA* A,
                                               int64_t lda, int Mv, const TIB* B, int64_t ldb, int Nv) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, k4 = lane >> 4;
    const int tM = M >> 4, tN = N >> 4;
    const int nt = tM * tN;
    const int bh = (nt >= 16 || (nt >= 8 && tM >= 2)) ? 2 : 1;
    const int bw = (nt >= 16) ? 2 : 1;
    const int bM = (tM + bh - 1) / bh, bN = (tN + bw - 1) / bw;
    for (int blk = wave; blk < bM * bN; blk += 4) {
        const int ti0 = (blk / bN) * bh, tj0 = (blk % bN) * bw;
        mra_f4 acc00 = {0, 0, 0, 0}, acc01 = acc00, acc10 = acc00, acc11 = acc00;
        const int ia = ti0 * 16 + r16, ib = ia + 16;
        const int ja = tj0 * 16 + r16, jb = ja + 16;
        const bool tM2 = bh == 2 && ti0 + 1 < tM, tN2 = bw == 2 && tj0 + 1 < tN;
        const bool ha = ia < Mv, hb = tM2 && ib < Mv;
        const bool hja = ja < Nv, hjb = tN2 && jb < Nv;
        constexpr int U = 8;
        int k0 = 0;
        for (; k0 + 4 * U <= Kv; k0 += 4 * U) {
            TIA a0[U], a1[U];
            TIB b0[U], b1[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int64_t k = k0 + 4 * u + k4;
                a0[u] = ha ? A[k * lda + ia] : TIA(0);
                a1[u] = hb ? A[k * lda + ib] : TIA(0);
                b0[u] = hja ? B[k * ldb + ja] : TIB(0);
                b1[u] = hjb ? B[k * ldb + jb] : TIB(0);
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const float xa0 = (float)a0[u], xa1 = (float)a1[u], xb0 = (float)b0[u], xb1 = (float)b1[u];
                acc00 = __builtin_amdgcn_mfma_f32_16x16x4f32(xa0, xb0, acc00, 0, 0, 0);
                if (tN2) acc01 = __builtin_amdgcn_mfma_f32_16x16x4f32(xa0, xb1, acc01, 0, 0, 0);
                if (tM2) acc10 = __builtin_amdgcn_mfma_f32_16x16x4f32(xa1, xb0, acc10, 0, 0, 0);
                if (tM2 && tN2) acc11 = __builtin_amdgcn_mfma_f32_16x16x4f32(xa1, xb1, acc11, 0, 0, 0);
            }
        }
        for (; k0 < Kv; k0 += 4) {
            const int k = k0 + k4;
            const bool kv = k < Kv;
            const float a0 = (kv && ha) ? (float)A[(int64_t)k * lda + ia] : 0.0f;
            const float a1 = (kv && hb) ? (float)A[(int64_t)k * lda + ib] : 0.0f;
            const float b0 = (kv && hja) ? (float)B[(int64_t)k * ldb + ja] : 0.0f;
            const float b1 = (kv && hjb) ? (float)B[(int64_t)k * ldb + jb] : 0.0f;
            acc00 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc00, 0, 0, 0);
            if (tN2) acc01 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc01, 0, 0, 0);
            if (tM2) acc10 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc10, 0, 0, 0);
            if (tM2 && tN2) acc11 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc11, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = ti0 * 16 + 4 * k4 + v;
            const bool r0 = row < Mv, r1 = tM2 && (row + 16 < Mv);
            TC* c00 = C + (int64_t)row * ldc + ja;
            TC* c10 = C + (int64_t)(row + 16) * ldc + ja;
            if (ACC) {
                if (r0 && hja) c00[0] += alpha * acc00[v];
                if (r0 && hjb) c00[16] += alpha * acc01[v];
                if (r1 && hja) c10[0] += alpha * acc10[v];
                if (r1 && hjb) c10[16] += alpha * acc11[v];
            } else {
                if (r0 && hja) c00[0] = alpha * acc00[v];
                if (r0 && hjb) c00[16] = alpha * acc01[v];
                if (r1 && hja) c10[0] = alpha * acc10[v];
                if (r1 && hjb) c10[16] = alpha * acc11[v];
            }
        }
    }
}

// Same product for an output of at most 32 x 32 (one 2x2 tile block), with
// K split over the 4 waves; the partial tiles are summed in wave order
// through red (LDS: 3*256 doubles for a 16 x 16 output, else 3*4*256), so
// the result is reproducible.
template <typename TI = double>
__device__ __forceinline__ void wg_gemm_tn_ksplit(double* C, int ldc, int M, int N, int Kv, double alpha,
                                                  const TI* A, int64_t lda, int Mv, const TI* B, int64_t ldb,
                                                  int Nv, double* red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, k4 = lane >> 4;
    const int tM = M >> 4, tN = N >> 4;
    const bool tM2 = tM > 1, tN2 = tN > 1;
    mra_d4 acc[4];
#pragma unroll
    for (int q = 0; q < 4; q++) acc[q] = mra_d4{0, 0, 0, 0};
    const int ia = r16, ib = 16 + r16, ja = r16, jb = 16 + r16;
    const bool ha = ia < Mv, hb = tM2 && ib < Mv, hja = ja < Nv, hjb = tN2 && jb < Nv;
    // this wave's contiguous K range, a multiple of 4
    const int kq = ((Kv + 15) / 16) * 4;
    const int kb = wave * kq, ke = min(Kv, kb + kq);
    constexpr int U = 6;
    int k0 = kb;
    for (; k0 + 4 * U <= ke; k0 += 4 * U) {  // batched loads, one latency per U steps
        TI a0[U], a1[U], b0[U], b1[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t k = k0 + 4 * u + k4;
            a0[u] = ha ? A[k * lda + ia] : TI(0);
            a1[u] = hb ? A[k * lda + ib] : TI(0);
            b0[u] = hja ? B[k * ldb + ja] : TI(0);
            b1[u] = hjb ? B[k * ldb + jb] : TI(0);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const double xa0 = (double)a0[u], xa1 = (double)a1[u], xb0 = (double)b0[u], xb1 = (double)b1[u];
            acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa0, xb0, acc[0], 0, 0, 0);
            if (tN2) acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa0, xb1, acc[1], 0, 0, 0);
            if (tM2) acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa1, xb0, acc[2], 0, 0, 0);
            if (tM2 && tN2) acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa1, xb1, acc[3], 0, 0, 0);
        }
    }
    for (; k0 < ke; k0 += 4) {
        const int k = k0 + k4;
        const bool kv = k < ke;
        const double a0 = (kv && ha) ? (double)A[(int64_t)k * lda + ia] : 0.0;
        const double a1 = (kv && hb) ? (double)A[(int64_t)k * lda + ib] : 0.0;
        const double b0 = (kv && hja) ? (double)B[(int64_t)k * ldb + ja] : 0.0;
        const double b1 = (kv && hjb) ? (double)B[(int64_t)k * ldb + jb] : 0.0;
        acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0], 0, 0, 0);
        if (tN2) acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[1], 0, 0, 0);
        if (tM2) acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[2], 0, 0, 0);
        if (tM2 && tN2) acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[3], 0, 0, 0);
    }
    const int nq = (tM2 || tN2) ? 4 : 1;  // red holds 3*nq*256 doubles
    if (wave > 0) {
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (q < nq)
#pragma unroll
                for (int v = 0; v < 4; v++) red[((wave - 1) * nq + q) * 256 + v * 64 + lane] = acc[q][v];
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (q < nq)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                double t = acc[q][v];
                for (int w = 0; w < 3; w++) t += red[(w * nq + q) * 256 + v * 64 + lane];
                const int row = (q >> 1) * 16 + k4 + 4 * v, col = (q & 1) * 16 + r16;
                if (row < Mv && col < Nv) C[row * ldc + col] += alpha * t;
            }
    }
    __syncthreads();
}

// ------------------------------------------------------------ LDS solvers
// Lower Cholesky in place of the n x n block (row stride lda); entries with
// i < j are left alone.  Right-looking elimination on the unscaled lower
// triangle, A[i][j] -= A[i][k] A[j][k] / A[k][k] (j > k), which needs one
// barrier per column, then one pass L[i][k] = A[i][k] / sqrt(A[k][k]).  A
// non-positive pivot sets *bad and is replaced by 1 so that the kernel
// still completes.
__device__ __forceinline__ void wg_chol(double* A, int lda, int n, int* bad) {
    __syncthreads();
    for (int k = 0; k < n - 1; k++) {
        double akk = A[k * lda + k];
        if (!(akk > 0.0)) akk = 1.0;
        const double inv = 1.0 / akk;
        const int m = n - k - 1;
        for (int idx = threadIdx.x; idx < m * m; idx += MRA_WG) {
            const int i = k + 1 + idx / m, j = k + 1 + idx % m;
            if (j <= i) A[i * lda + j] -= A[i * lda + k] * A[j * lda + k] * inv;
        }
        __syncthreads();
    }
    // scale: pivots are final now
    for (int idx = threadIdx.x; idx < n * n; idx += MRA_WG) {
        const int i = idx / n, k = idx % n;
        if (k > i) continue;
        double akk = A[k * lda + k];
        if (!(akk > 0.0)) {
            if (i == k) *bad = 1;
            akk = 1.0;
        }
        if (i != k) A[i * lda + k] /= sqrt(akk);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += MRA_WG) {
        const double akk = A[k * lda + k];
        A[k * lda + k] = (akk > 0.0) ? sqrt(akk) : 1.0;
    }
    __syncthreads();
}

// ------------------------------------------------ blocked LDS factorisations
// N x N (N a multiple of 16, at most 128) lower Cholesky and triangular
// inverse on 16 x 16 blocks.  One wave factors each diagonal block and
// inverts it (wave-synchronous, no work-group barrier); panel solves, trailing
// updates and the off-diagonal blocks of the inverse are 16 x 16 x 16 MFMA
// products spread over the 4 waves.  The 16 x 16 inverse diagonal blocks
// D_b = L_bb^{-1} are kept in the strictly upper blocks (b, b+1) and
// (0, N/16-1) of the matrix itself when N >= 48 (those entries are free),
// else in the caller's buffer dbuf (N/16 blocks of 256 doubles).
// Barriers: 3 per block column for the factorisation, 2 per block row for
// the inverse.

// compiler-only ordering of LDS accesses made by one wave (LDS executes a
// wave's instructions in order)
__device__ __forceinline__ void wave_lds_fence() { __atomic_signal_fence(__ATOMIC_SEQ_CST); }

template <int N>
struct BlkLayout {
    static constexpr int NB = N / 16;
    static constexpr bool UPPER = NB >= 3;
    __device__ static double* dptr(double* A, int lda, double* dbuf, int b) {
        if (UPPER) return b < NB - 1 ? A + b * 16 * lda + (b + 1) * 16 : A + (NB - 1) * 16;
        return dbuf + b * 256;
    }
    __device__ static int ldd(int lda) { return UPPER ? lda : 16; }
};

// 16 x 16 x 16 product on one wave: acc[v] (row k4 + 4v, column r16) +=
// sum_k X(row, k) Y(k, column) with X(m, k) = X[m*ldx + k] and
// Y(k, n) = TY ? Y[n*ldy + k] : Y[k*ldy + n].
template <bool TY>
__device__ __forceinline__ mra_d4 mfma16(mra_d4 acc, const double* X, int ldx, const double* Y, int ldy) {
    const int lane = threadIdx.x & 63, r16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const int k = 4 * s + k4;
        const double a = X[r16 * ldx + k];
        const double b = TY ? Y[r16 * ldy + k] : Y[k * ldy + r16];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    return acc;
}

// One wave: factor the 16 x 16 diagonal block P in place (lower triangle)
// and write its inverse (lower, zeros above) to Dp (row stride ldd).
__device__ __forceinline__ void wave_chol16(double* P, int lda, double* Dp, int ldd, int* bad) {
    const int lane = threadIdx.x & 63, i = lane & 15, cg = lane >> 4;
    // unscaled right-looking elimination of the lower triangle:
    // A[i][j] -= A[i][k] A[j][k] / A[k][k]  (k < j <= i)
    for (int k = 0; k < 15; k++) {
        wave_lds_fence();
        double akk = P[k * lda + k];
        if (!(akk > 0.0)) akk = 1.0;
        const double inv = 1.0 / akk;
        const double aik = P[i * lda + k];
        double upd[4];
        bool act[4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const int j = cg + 4 * m;
            act[m] = j > k && j <= i;
            upd[m] = act[m] ? P[i * lda + j] - aik * P[j * lda + k] * inv : 0.0;
        }
        wave_lds_fence();
#pragma unroll
        for (int m = 0; m < 4; m++)
            if (act[m]) P[i * lda + cg + 4 * m] = upd[m];
    }
    wave_lds_fence();
    // pivots are final: L[i][j] = A[i][j] / sqrt(A[j][j]), L[i][i] = sqrt(A[i][i])
    double val[4], piv[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const int j = cg + 4 * m;
        val[m] = j <= i ? P[i * lda + j] : 0.0;
        piv[m] = P[j * lda + j];
    }
    wave_lds_fence();
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const int j = cg + 4 * m;
        double pj = piv[m];
        if (!(pj > 0.0)) {
            if (j == i) *bad = 1;
            pj = 1.0;
        }
        if (j < i) P[i * lda + j] = val[m] / sqrt(pj);
        if (j == i) P[i * lda + i] = sqrt(pj);
    }
    wave_lds_fence();
    // inverse by forward substitution, one column per lane (lanes 0..15);
    // the partial column lives in Dp (rolled loops keep register use low)
    if (lane < 16) {
        const int c = lane;
#pragma unroll 1
        for (int ii = 0; ii < 16; ii++) {
            double sacc = ii == c ? 1.0 : 0.0;
#pragma unroll 1
            for (int kk = 0; kk < ii; kk++) sacc -= P[ii * lda + kk] * Dp[kk * ldd + c];
            Dp[ii * ldd + c] = sacc / P[ii * lda + ii];
        }
    }
    wave_lds_fence();
}

template <int N>
__device__ __forceinline__ void wg_chol_blk(double* A, int lda, double* dbuf, int* bad) {
    using BL = BlkLayout<N>;
    constexpr int NB = BL::NB;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r16 = lane & 15, k4 = lane >> 4;
    const int ldd = BL::ldd(lda);
    __syncthreads();
    for (int kb = 0; kb < NB; kb++) {
        double* Pkk = A + kb * 16 * (lda + 1);
        double* Dk = BL::dptr(A, lda, dbuf, kb);
        if (wave == 0) wave_chol16(Pkk, lda, Dk, ldd, bad);
        __syncthreads();
        if (kb == NB - 1) break;
        // panel: L_{ib,kb} = A_{ib,kb} D_kb^T
        for (int ib = kb + 1 + wave; ib < NB; ib += 4) {
            double* Pib = A + ib * 16 * lda + kb * 16;
            mra_d4 acc = {0, 0, 0, 0};
            acc = mfma16<true>(acc, Pib, lda, Dk, ldd);
            wave_lds_fence();
#pragma unroll
            for (int v = 0; v < 4; v++) Pib[(k4 + 4 * v) * lda + r16] = acc[v];
        }
        __syncthreads();
        // trailing update A_{ib,jb} -= L_{ib,kb} L_{jb,kb}^T, kb < jb <= ib
        const int m = NB - 1 - kb;
        for (int q = wave; q < m * (m + 1) / 2; q += 4) {
            int ib = 0, rem = q;
            while (rem > ib) {
                rem -= ib + 1;
                ib++;
            }
            const int jb = rem;  // 0 <= jb <= ib (relative)
            const double* Li = A + (kb + 1 + ib) * 16 * lda + kb * 16;
            const double* Lj = A + (kb + 1 + jb) * 16 * lda + kb * 16;
            mra_d4 acc = {0, 0, 0, 0};
            acc = mfma16<true>(acc, Li, lda, Lj, lda);
            double* Cb = A + (kb + 1 + ib) * 16 * lda + (kb + 1 + jb) * 16;
#pragma unroll
            for (int v = 0; v < 4; v++) Cb[(k4 + 4 * v) * lda + r16] -= acc[v];
        }
        __syncthreads();
    }
}

// A <- L^{-1} in place after wg_chol_blk<N> (same dbuf); the strictly upper
// part is set to zero.
template <int N>
__device__ __forceinline__ void wg_trinv_blk(double* A, int lda, double* dbuf) {
    using BL = BlkLayout<N>;
    constexpr int NB = BL::NB;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r16 = lane & 15, k4 = lane >> 4;
    const int ldd = BL::ldd(lda);
    __syncthreads();
    for (int i = 1; i < NB; i++) {
        // T_ij = sum_{k=j}^{i-1} L_ik X_kj (X_jj = D_j; X_kj, k > j, final),
        // at most two blocks j per wave
        mra_d4 acc[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int j = wave + 4 * t;
            if (j < i)
                for (int k = j; k < i; k++) {
                    const double* Lik = A + i * 16 * lda + k * 16;
                    const double* Xkj = k == j ? BL::dptr(A, lda, dbuf, j) : A + k * 16 * lda + j * 16;
                    acc[t] = mfma16<false>(acc[t], Lik, lda, Xkj, k == j ? ldd : lda);
                }
        }
        __syncthreads();  // every L_ij of block row i has been read
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int j = wave + 4 * t;
            if (j < i) {
                double* Cb = A + i * 16 * lda + j * 16;
#pragma unroll
                for (int v = 0; v < 4; v++) Cb[(k4 + 4 * v) * lda + r16] = acc[t][v];
                wave_lds_fence();
                // X_ij = -D_i T_ij
                mra_d4 x = {0, 0, 0, 0};
                x = mfma16<false>(x, BL::dptr(A, lda, dbuf, i), ldd, Cb, lda);
                wave_lds_fence();
#pragma unroll
                for (int v = 0; v < 4; v++) Cb[(k4 + 4 * v) * lda + r16] = -x[v];
            }
        }
        __syncthreads();
    }
    // diagonal blocks <- D_b (read every D_b before any upper block is zeroed)
    constexpr int PER = (NB * 256 + MRA_WG - 1) / MRA_WG;
    double dv[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int idx = threadIdx.x + q * MRA_WG;
        dv[q] = 0.0;
        if (idx < NB * 256) {
            const int b = idx >> 8, e = idx & 255;
            dv[q] = BL::dptr(A, lda, dbuf, b)[(e >> 4) * ldd + (e & 15)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int idx = threadIdx.x + q * MRA_WG;
        if (idx < NB * 256) {
            const int b = idx >> 8, e = idx & 255, ii = e >> 4, jj = e & 15;
            A[(b * 16 + ii) * lda + b * 16 + jj] = jj <= ii ? dv[q] : 0.0;
        }
    }
    for (int idx = threadIdx.x; idx < N * N; idx += MRA_WG) {
        const int ii = idx / N, jj = idx % N;
        if ((jj >> 4) > (ii >> 4)) A[ii * lda + jj] = 0.0;
    }
    __syncthreads();
}

// In-place inverse of the lower-triangular n x n block (row stride lda),
// column by column from the last: two barriers per column.  tmp: n doubles
// of scratch.
__device__ __forceinline__ void wg_trinv(double* A, int lda, int n, double* tmp) {
    __syncthreads();
    for (int j = n - 1; j >= 0; j--) {
        const double dinv = 1.0 / A[j * lda + j];
        // x_i = sum_{k=j+1..i} Ainv[i][k] A[k][j]  for i > j (old column j)
        for (int i = j + 1 + threadIdx.x; i < n; i += MRA_WG) {
            double s = 0.0;
            for (int k = j + 1; k <= i; k++) s += A[i * lda + k] * A[k * lda + j];
            tmp[i] = s;
        }
        __syncthreads();
        for (int i = j + 1 + threadIdx.x; i < n; i += MRA_WG) A[i * lda + j] = -tmp[i] * dinv;
        if (threadIdx.x == 0) A[j * lda + j] = dinv;
        __syncthreads();
    }
}

// B <- L^{-1} B for lower L (n x n, row stride ldl), B n x ncol (ldb).
__device__ __forceinline__ void wg_trsm(const double* L, int ldl, double* B, int ldb, int n, int ncol) {
    for (int k = 0; k < n; k++) {
        const double lkk = L[k * ldl + k];
        for (int c = threadIdx.x; c < ncol; c += MRA_WG) B[k * ldb + c] /= lkk;
        __syncthreads();
        const int m = n - k - 1;
        for (int idx = threadIdx.x; idx < m * ncol; idx += MRA_WG) {
            const int i = k + 1 + idx / ncol, c = idx % ncol;
            B[i * ldb + c] -= L[i * ldl + k] * B[k * ldb + c];
        }
        __syncthreads();
    }
}

// Deterministic work-group sum of one value per thread (tree in LDS).
__device__ __forceinline__ double wg_sum(double x, double* red) {
    red[threadIdx.x] = x;
    __syncthreads();
    for (int s = MRA_WG / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

}  // namespace mra
