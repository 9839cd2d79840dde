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

#include <type_traits>

#include "mra_exp.hpp"
#include <stdint.h>

#define MRA_WG 256

typedef double mra_d4 __attribute__((ext_vector_type(4)));

// wave index within the work group, as a wave-uniform (scalar) value
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }
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
// covariance of two points given by their coordinates, variance field s and
// range field b (the arithmetic of pcov, which reads them from Pts)
__device__ __forceinline__ double pcov_v(const CovCtx& c, double x1, double y1, double z1, double s1, double b1,
                                         double x2, double y2, double z2, double s2, double b2) {
    double d = pdist(c, x1, y1, z1, x2, y2, z2);
    if (c.stationary) return c.sill * exp_neg(-d / c.range);
    double bms = (b1 * b1 + b2 * b2) * 0.5;
    double ratio = (b1 * b2) / bms;
    return (s1 * s2) * (ratio * sqrt(ratio)) * exp_neg(-d / sqrt(bms));
}

__device__ __forceinline__ double pcov(const CovCtx& c, const Pts& P, uint64_t i, const Pts& Q, uint64_t j) {
    if (c.stationary) return pcov_v(c, P.x[i], P.y[i], P.z[i], 0.0, 0.0, Q.x[j], Q.y[j], Q.z[j], 0.0, 0.0);
    return pcov_v(c, P.x[i], P.y[i], P.z[i], P.s[i], P.b[i], Q.x[j], Q.y[j], Q.z[j], Q.s[j], Q.b[j]);
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
    if (c.stationary) return c.sill * exp_neg(-d / c.range);
    const double t = rsqrt_nr((a.b * a.b + b.b * b.b) * 0.5);
    return (a.q * b.q) * (t * t * t) * exp_neg(-d * t);
}

// FP32 evaluation of the same form (mixed precision) on point data rounded
// to float: raw v_sqrt_f32, v_rsq_f32 and v_exp_f32 (about 1 ulp each; the
// exponential as 2^(-d t log2 e)).  PLANAR: the z coordinates are zero.
struct PtF {
    float x, y, z, q, b;
};

__device__ __forceinline__ PtF to_ptf(const PtV& p) {
    return PtF{(float)p.x, (float)p.y, (float)p.z, (float)p.q, (float)p.b};
}

// Explicit fused multiply-adds and no contraction, so every inlined copy
// rounds identically (vector and scalar store paths give the same bits).
template <bool PLANAR>
__device__ __forceinline__ float pcov_ff(const PtF& a, const PtF& b, bool chordal) {
#pragma clang fp contract(off)
    const float dx = b.x - a.x, dy = b.y - a.y;
    float d2 = __builtin_fmaf(dx, dx, dy * dy);
    if (!PLANAR) {
        const float dz = b.z - a.z;
        d2 = __builtin_fmaf(dz, dz, d2);
    }
    float d = __builtin_amdgcn_sqrtf(d2);
    if (chordal) d *= 6371.0f;
    const float t = __builtin_amdgcn_rsqf(__builtin_fmaf(a.b, a.b, b.b * b.b) * 0.5f);
    return (a.q * b.q) * (t * t * t) * __builtin_amdgcn_exp2f(-(d * t) * 1.44269504088896341f);
}

__device__ __forceinline__ float pcov_ptf(const CovCtx& c, const PtV& a, const PtV& b) {
    if (c.metric == 1 || c.stationary) {
        const float d = (float)pdist(c, a.x, a.y, a.z, b.x, b.y, b.z);
        if (c.stationary) return (float)c.sill * expf(-d / (float)c.range);
        const float ab = (float)a.b, bb = (float)b.b;
        const float t = __builtin_amdgcn_rsqf((ab * ab + bb * bb) * 0.5f);
        return ((float)a.q * (float)b.q) * (t * t * t) * expf(-d * t);
    }
    return pcov_ff<false>(to_ptf(a), to_ptf(b), c.metric == 0);
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

// The same tile (no diagonal variant) from row points already staged in LDS
// (st[i], i < nr) and this thread's column point qc (column threadIdx.x % NC,
// valid when < nc); the caller synchronises before reuse of C.
// vdiag (optional): rows and columns are the same point set, the diagonal
// takes vdiag[i] (identity on padding rows).
// F32 (nonstationary chordal / planar kernel): the coordinate differences and
// d^2 in FP64, the rest (root, range factor, exponential) in FP32 on the
// hardware's reciprocal square root and exponential, widened to T.
template <int RP, int NC, typename T, bool F32 = false>
__device__ __forceinline__ void cov_tile_pre(T* C, int LD, const CovCtx& cc, const PtV* st, int nr, const PtV& qc,
                                             int nc, const double* vdiag = nullptr) {
    static_assert(MRA_WG % NC == 0, "tile width must divide the work-group size");
    const int c = threadIdx.x % NC;
    if (F32 && !cc.stationary && cc.metric != 1) {
        const float dscale = cc.metric == 0 ? 6371.0f : 1.0f;
        const float qb = (float)qc.b, qb2 = qb * qb, qq = (float)qc.q;
        const bool cok = c < nc;
#pragma unroll 4
        for (int i = threadIdx.x / NC; i < RP; i += MRA_WG / NC) {
            const PtV& a = st[i];
            const double dx = qc.x - a.x, dy = qc.y - a.y, dz = qc.z - a.z;
            const float d2 = (float)(dx * dx + dy * dy + dz * dz);
            const float dd = d2 * __builtin_amdgcn_rsqf(fmaxf(d2, 1e-30f)) * dscale;
            const float ab = (float)a.b;
            const float t = __builtin_amdgcn_rsqf((ab * ab + qb2) * 0.5f);
            const float v = ((float)a.q * qq) * (t * t * t) * __expf(-dd * t);
            C[i * LD + c] = (T)((i < nr && cok) ? v : 0.0f);
        }
        if (vdiag)
            for (int i = threadIdx.x / NC; i < RP; i += MRA_WG / NC)
                if (i == c) C[i * LD + c] = (T)(i < nr ? vdiag[i] : 1.0);
        return;
    }
    if (sizeof(T) == 8 && !cc.stationary && cc.metric != 1) {
        // FP64 nonstationary chordal / planar kernel (the values of pcov_pt):
        // a straight-line body, validity by selects, four rows at a time so
        // that their exponential-table loads overlap
        // (d = d2 / sqrt(d2) with the root taken of max(d2, 1e-300): exactly
        // 0 for coincident points, the value of pcov_pt otherwise)
        const double dscale = cc.metric == 0 ? 6371.0 : 1.0;
        constexpr int STEP = MRA_WG / NC, G = RP / STEP >= 4 ? 4 : RP / STEP;
        static_assert(G >= 1 && RP % (STEP * G) == 0, "rows in groups of up to four per thread");
        for (int i0 = threadIdx.x / NC; i0 < RP; i0 += G * STEP) {
            double cv[G], xg[G], fg[G], tk[G];
            ExpParts eg[G];
#pragma unroll
            for (int g = 0; g < G; g++) {
                const PtV& a = st[i0 + g * STEP];
                const double dx = qc.x - a.x, dy = qc.y - a.y, dz = qc.z - a.z;
                const double d2 = dx * dx + dy * dy + dz * dz;
                const double d = d2 * rsqrt_nr(fmax(d2, 1e-300)) * dscale;
                const double t = rsqrt_nr((a.b * a.b + qc.b * qc.b) * 0.5);
                fg[g] = (a.q * qc.q) * (t * t * t);
                xg[g] = -d * t;
                eg[g] = exp_neg_parts(xg[g]);
            }
            // the group's table loads in flight together (one wait)
#pragma unroll
            for (int g = 0; g < G; g++) tk[g] = Exp2Frac::T[eg[g].k & 63];
            if constexpr (G == 4)
                asm volatile("" ::"v"(tk[0]), "v"(tk[1]), "v"(tk[2]), "v"(tk[3]));
#pragma unroll
            for (int g = 0; g < G; g++) cv[g] = fg[g] * exp_neg_finish(tk[g], eg[g], xg[g]);
#pragma unroll
            for (int g = 0; g < G; g++) {
                const int i = i0 + g * STEP;
                C[i * LD + c] = (T)((i < nr && c < nc) ? cv[g] : 0.0);
            }
        }
        if (vdiag)
            for (int i = threadIdx.x / NC; i < RP; i += MRA_WG / NC)
                if (i == c) C[i * LD + c] = (T)(i < nr ? vdiag[i] : 1.0);
        return;
    }
    for (int i = threadIdx.x / NC; i < RP; i += MRA_WG / NC) {
        T v = T(0);
        if (vdiag && i == c) {
            v = i < nr ? (T)vdiag[i] : T(1);
        } else if (i < nr && c < nc) {
            if constexpr (sizeof(T) == 4)
                v = pcov_ptf(cc, st[i], qc);
            else
                v = pcov_pt(cc, st[i], qc);
        }
        C[i * LD + c] = v;
    }
}

// A 64 x 64 covariance tile in the accumulator layout of the one-wave
// products (wave_mm64_acc): element v of acc[ti * 4 + tj] of lane t is
// C(rows[i], cols[c]) with i = 16 ti + t/16 + 4 v and c = 16 tj + t%16, zero
// where i or c >= n.  vdiag (rows and columns are one point set): the
// diagonal takes vdiag[i], identity on padding.  rows, cols: LDS.  Per
// column tile, each lane evaluates its 16 entries in a loop (one copy of the
// covariance code per column tile) into buf (16 * 64 doubles of LDS, this
// wave's own; entry q of lane t at buf[64 q + t]) and reads them back into
// the accumulators, whose register indices must be compile-time constants.
__device__ __forceinline__ void wave_cov64(mra_d4 (&acc)[16], const CovCtx& cc, const PtV* rows, const PtV* cols,
                                           int n, const double* vdiag, double* buf) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int tj = 0; tj < 4; tj++) {
        const int c = 16 * tj + r16;
        const PtV cq = cols[c];
#pragma unroll 1
        for (int q = 0; q < 16; q++) {
            const int i = k4 + 4 * q;  // row 16 ti + k4 + 4 v with q = 4 ti + v
            double x = (i < n && c < n) ? pcov_pt(cc, rows[i], cq) : 0.0;
            if (vdiag && i == c) x = i < n ? vdiag[i] : 1.0;
            buf[64 * q + lane] = x;
        }
#pragma unroll
        for (int q = 0; q < 16; q++) acc[(q >> 2) * 4 + tj][q & 3] = buf[64 * q + lane];
    }
}

// FP32 counterpart of pcov_v (inputs rounded to float)
__device__ __forceinline__ float pcovf_v(const CovCtx& c, double x1d, double y1d, double z1d, double s1, double b1d,
                                         double x2d, double y2d, double z2d, double s2, double b2d) {
    float x1 = (float)x1d, y1 = (float)y1d, z1 = (float)z1d;
    float x2 = (float)x2d, y2 = (float)y2d, z2 = (float)z2d;
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
    float b1 = (float)b1d, b2 = (float)b2d;
    float bms = (b1 * b1 + b2 * b2) * 0.5f;
    float ratio = (b1 * b2) / bms;
    return ((float)s1 * (float)s2) * (ratio * sqrtf(ratio)) * expf(-d / sqrtf(bms));
}

__device__ __forceinline__ float pcovf(const CovCtx& c, const Pts& P, uint64_t i, const Pts& Q, uint64_t j) {
    if (c.stationary) return pcovf_v(c, P.x[i], P.y[i], P.z[i], 0.0, 0.0, Q.x[j], Q.y[j], Q.z[j], 0.0, 0.0);
    return pcovf_v(c, P.x[i], P.y[i], P.z[i], P.s[i], P.b[i], Q.x[j], Q.y[j], Q.z[j], Q.s[j], Q.b[j]);
}

// ----------------------------------------------------------- MFMA GEMM TN
// C[i*ldc+j] += alpha * sum_k A(k,i) B(k,j) over the M x N tile (multiples
// of 16).  A(k,i) = A[k*lda+i] valid for k < Kv, i < Mv (else 0); B likewise
// with Nv (Mv, Nv >= 1).  K is rounded up to a multiple of 4.  C is LDS (or
// global) and must be made visible by the caller (__syncthreads before and
// after).  TA: A is stored transposed, A(k,i) = A[i*lda+k] (an LDS matrix
// read by columns).  ACC = false stores C = alpha*A^T B instead of
// accumulating.
// TILES: GEMM_ANY checks per wave block whether every operand is in range
// and then takes the unpredicated path; GEMM_FULL asserts that Mv >= M and
// Nv >= N (only that path is compiled); GEMM_PART compiles only the
// predicated path.
// Each wave owns blocks of up to 2 x 2 tiles of 16 x 16.  The k loop is
// instantiated per block shape (no branches inside): operand loads of U
// k-steps are issued together, at wave-uniform row addresses plus 32-bit
// per-lane offsets; out-of-range rows are read at a clamped valid row and
// replaced by zero with a select.
enum { GEMM_ANY = 0, GEMM_FULL = 1, GEMM_PART = 2 };
// TRI (bit set, caller's promise about the operands): GEMM_TRI_A: A(k,i) = 0
// for k > i (A^T lower triangular: a wave block stops K at its last row);
// GEMM_TRI_B: B(k,j) = 0 for k > j (stops K at its last column); GEMM_TRI_C:
// only C on and below the diagonal is wanted (wave blocks strictly above the
// diagonal are skipped and that part of C is left as it was).
enum { GEMM_TRI_A = 1, GEMM_TRI_B = 2, GEMM_TRI_C = 4 };

template <typename TC>
struct Mfma;
template <>
struct Mfma<double> {
    typedef mra_d4 V;
    __device__ static V op(double a, double b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    // D element v of lane t is row t/16 + 4v (column t%16)
    __device__ static int row(int k4, int v) { return k4 + 4 * v; }
};
template <>
struct Mfma<float> {
    typedef mra_f4 V;
    __device__ static V op(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    // D element v of lane t is row 4*(t/16) + v (column t%16)
    __device__ static int row(int k4, int v) { return 4 * k4 + v; }
};

// Operands of one wave block: lane offsets (bytes) of the two row tiles of
// A and the two column tiles of B at k = k4, and the per-k strides.
// TB: B stored transposed, B(k,j) = B[j*ldb+k].
template <bool TA, typename TIA, typename TIB, bool TB = false>
struct BlockOps {
    const char* A;
    const char* B;
    int64_t sa, sb;         // bytes per k step
    uint32_t oa0, oa1, ob0, ob1;
    bool ha, hb, hja, hjb;  // rows / columns in range
    __device__ BlockOps(const TIA* A_, int64_t lda, const TIB* B_, int64_t ldb, int ia, int ib, int ja, int jb,
                        bool ha_, bool hb_, bool hja_, bool hjb_, int k4)
        : A((const char*)A_), B((const char*)B_), ha(ha_), hb(hb_), hja(hja_), hjb(hjb_) {
        const int iac = ha ? ia : 0, ibc = hb ? ib : 0, jac = hja ? ja : 0, jbc = hjb ? jb : 0;
        sa = (TA ? 1 : lda) * (int64_t)sizeof(TIA);
        sb = (TB ? 1 : ldb) * (int64_t)sizeof(TIB);
        oa0 = (uint32_t)((TA ? iac * lda + k4 : k4 * lda + iac) * sizeof(TIA));
        oa1 = (uint32_t)((TA ? ibc * lda + k4 : k4 * lda + ibc) * sizeof(TIA));
        ob0 = (uint32_t)((TB ? jac * ldb + k4 : k4 * ldb + jac) * sizeof(TIB));
        ob1 = (uint32_t)((TB ? jbc * ldb + k4 : k4 * ldb + jbc) * sizeof(TIB));
    }
};

// PIPE: the operands of the next U k steps are loaded before the products of
// the current ones are issued (more registers; for kernels with spare ones).
template <typename TC, bool M2, bool N2, bool PRED, int U, bool PIPE = false, bool TA, typename TIA, typename TIB,
          bool TB>
__device__ __forceinline__ void block_kloop(typename Mfma<TC>::V (&acc)[4], const BlockOps<TA, TIA, TIB, TB>& o,
                                            int kb, int ke, int k4) {
    int k0 = kb;
    // operands of U k steps
    struct Ops {
        TIA a0[U], a1[U];
        TIB b0[U], b1[U];
    };
    auto load = [&](int kk, Ops& p) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const char* Ak = o.A + (int64_t)(kk + 4 * u) * o.sa;
            const char* Bk = o.B + (int64_t)(kk + 4 * u) * o.sb;
            p.a0[u] = *(const TIA*)(Ak + o.oa0);
            if (M2) p.a1[u] = *(const TIA*)(Ak + o.oa1);
            p.b0[u] = *(const TIB*)(Bk + o.ob0);
            if (N2) p.b1[u] = *(const TIB*)(Bk + o.ob1);
        }
    };
    auto mm = [&](const Ops& p) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const TC xa0 = (!PRED || o.ha) ? (TC)p.a0[u] : TC(0);
            const TC xb0 = (!PRED || o.hja) ? (TC)p.b0[u] : TC(0);
            acc[0] = Mfma<TC>::op(xa0, xb0, acc[0]);
            if (N2) {
                const TC xb1 = (!PRED || o.hjb) ? (TC)p.b1[u] : TC(0);
                acc[1] = Mfma<TC>::op(xa0, xb1, acc[1]);
            }
            if (M2) {
                const TC xa1 = (!PRED || o.hb) ? (TC)p.a1[u] : TC(0);
                acc[2] = Mfma<TC>::op(xa1, xb0, acc[2]);
                if (N2) {
                    const TC xb1 = (!PRED || o.hjb) ? (TC)p.b1[u] : TC(0);
                    acc[3] = Mfma<TC>::op(xa1, xb1, acc[3]);
                }
            }
        }
    };
    if constexpr (PIPE) {
        if (k0 + 4 * U <= ke) {
            Ops cur;
            load(k0, cur);
            for (; k0 + 8 * U <= ke; k0 += 4 * U) {
                Ops nxt;
                load(k0 + 4 * U, nxt);
                mm(cur);
                cur = nxt;
            }
            mm(cur);
            k0 += 4 * U;
        }
    } else {
        for (; k0 + 4 * U <= ke; k0 += 4 * U) {
            Ops cur;
            load(k0, cur);
            mm(cur);
        }
    }
    // remaining k steps: k clamped into range, zero beyond ke
    for (; k0 < ke; k0 += 4) {
        const int k = k0 + k4;
        const bool kv = k < ke;
        const int64_t dk = (int64_t)((kv ? k : ke - 1) - k4);  // offset from the lane's k4 row
        const char* Ak = o.A + dk * o.sa;
        const char* Bk = o.B + dk * o.sb;
        const TC xa0 = (kv && o.ha) ? (TC) * (const TIA*)(Ak + o.oa0) : TC(0);
        const TC xb0 = (kv && o.hja) ? (TC) * (const TIB*)(Bk + o.ob0) : TC(0);
        acc[0] = Mfma<TC>::op(xa0, xb0, acc[0]);
        if (N2) {
            const TC xb1 = (kv && o.hjb) ? (TC) * (const TIB*)(Bk + o.ob1) : TC(0);
            acc[1] = Mfma<TC>::op(xa0, xb1, acc[1]);
        }
        if (M2) {
            const TC xa1 = (kv && o.hb) ? (TC) * (const TIA*)(Ak + o.oa1) : TC(0);
            acc[2] = Mfma<TC>::op(xa1, xb0, acc[2]);
            if (N2) {
                const TC xb1 = (kv && o.hjb) ? (TC) * (const TIB*)(Bk + o.ob1) : TC(0);
                acc[3] = Mfma<TC>::op(xa1, xb1, acc[3]);
            }
        }
    }
}

// dispatch on the block shape (wave-uniform)
template <typename TC, bool PRED, int U, bool PIPE = false, bool TA, typename TIA, typename TIB, bool TB>
__device__ __forceinline__ void block_kloop_any(typename Mfma<TC>::V (&acc)[4], const BlockOps<TA, TIA, TIB, TB>& o,
                                                bool m2, bool n2, int kb, int ke, int k4) {
    if (m2 && n2)
        block_kloop<TC, true, true, PRED, U, PIPE>(acc, o, kb, ke, k4);
    else if (m2)
        block_kloop<TC, true, false, PRED, U, PIPE>(acc, o, kb, ke, k4);
    else if (n2)
        block_kloop<TC, false, true, PRED, U, PIPE>(acc, o, kb, ke, k4);
    else
        block_kloop<TC, false, false, PRED, U, PIPE>(acc, o, kb, ke, k4);
}

template <typename TC, bool TA, bool ACC, typename TIA, typename TIB, int TILES, typename TO, bool TB = false,
          bool DUAL = false, int TRI = 0, bool PIPE = false>
__device__ __forceinline__ void wg_gemm_tn_impl(TO* C, int ldc, int M, int N, int Kv, TC alpha, const TIA* A,
                                                int64_t lda, int Mv, const TIB* B, int64_t ldb, int Nv,
                                                float* C32 = nullptr) {
    static_assert(!(TA && TILES == GEMM_FULL), "the full-tile path reads A untransposed");
    typedef typename Mfma<TC>::V V;
    const int lane = threadIdx.x & 63, wave = wave_id();
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
        const int ia = ti0 * 16 + r16, ib = ia + 16;
        const int ja = tj0 * 16 + r16, jb = ja + 16;
        const bool tM2 = bh == 2 && ti0 + 1 < tM, tN2 = bw == 2 && tj0 + 1 < tN;
        if ((TRI & GEMM_TRI_C) && tj0 >= ti0 + (tM2 ? 2 : 1)) continue;  // strictly above the diagonal
        int ke = Kv;
        if (TRI & GEMM_TRI_A) ke = min(ke, ti0 * 16 + (tM2 ? 32 : 16));
        if (TRI & GEMM_TRI_B) ke = min(ke, tj0 * 16 + (tN2 ? 32 : 16));
        const bool ha = ia < Mv, hb = tM2 && ib < Mv;
        const bool hja = ja < Nv, hjb = tN2 && jb < Nv;
        const BlockOps<TA, TIA, TIB, TB> o(A, lda, B, ldb, ia, ib, ja, jb, ha, hb, hja, hjb, k4);
        V acc[4];
#pragma unroll
        for (int q = 0; q < 4; q++) acc[q] = V{0, 0, 0, 0};
        const bool full = TILES == GEMM_FULL ||
                          (TILES == GEMM_ANY && ti0 * 16 + (tM2 ? 32 : 16) <= Mv && tj0 * 16 + (tN2 ? 32 : 16) <= Nv);
        if (TILES != GEMM_PART && full)
            block_kloop_any<TC, false, 4, PIPE>(acc, o, tM2, tN2, 0, ke, k4);
        else if (TILES != GEMM_FULL)
            block_kloop_any<TC, true, 4, PIPE>(acc, o, tM2, tN2, 0, ke, k4);
        // only the valid Mv x Nv part of C is updated
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = ti0 * 16 + Mfma<TC>::row(k4, v);
            const bool r0 = row < Mv, r1 = tM2 && (row + 16 < Mv);
            TO* c00 = C + (int64_t)row * ldc + ja;
            TO* c10 = C + (int64_t)(row + 16) * ldc + ja;
            if (ACC) {
                if (r0 && hja) c00[0] += alpha * acc[0][v];
                if (r0 && hjb) c00[16] += alpha * acc[1][v];
                if (r1 && hja) c10[0] += alpha * acc[2][v];
                if (r1 && hjb) c10[16] += alpha * acc[3][v];
            } else {
                if (r0 && hja) c00[0] = alpha * acc[0][v];
                if (r0 && hjb) c00[16] = alpha * acc[1][v];
                if (r1 && hja) c10[0] = alpha * acc[2][v];
                if (r1 && hjb) c10[16] = alpha * acc[3][v];
                if (DUAL) {
                    // the same values rounded to FP32 (second output, same layout)
                    float* d00 = C32 + (int64_t)row * ldc + ja;
                    float* d10 = C32 + (int64_t)(row + 16) * ldc + ja;
                    if (r0 && hja) d00[0] = (float)(alpha * acc[0][v]);
                    if (r0 && hjb) d00[16] = (float)(alpha * acc[1][v]);
                    if (r1 && hja) d10[0] = (float)(alpha * acc[2][v]);
                    if (r1 && hjb) d10[16] = (float)(alpha * acc[3][v]);
                }
            }
        }
    }
}

// One wave accumulates a whole 64 x 64 product acc += al A^T B (A, B: K x 64
// row-major with leading dimensions lda, ldb, all in range): 4 x 4 MFMA tiles
// in registers, so every operand value loaded feeds four matrix
// instructions.  Tile (ti, tj), element v of lane t holds row
// 16 ti + t/16 + 4 v, column 16 tj + t%16.  FP64.  Waves work on different
// outputs independently (no barrier).  LOWER: only the tiles on and below the
// diagonal (tj <= ti) are accumulated.
template <bool LOWER = false>
__device__ __forceinline__ void wave_mm64_acc(mra_d4 (&acc)[16], int K, double al, const double* A, int64_t lda,
                                              const double* B, int64_t ldb) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, k4 = lane >> 4;
    const uint32_t oa = (uint32_t)((k4 * lda + r16) * sizeof(double));
    const uint32_t ob = (uint32_t)((k4 * ldb + r16) * sizeof(double));
    // operands of 8 rows; the next 8 rows are loaded before the products of
    // the current ones are issued
    auto load8 = [&](int k0, double (&av)[2][4], double (&bv)[2][4]) {
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const char* Ak = (const char*)(A + (int64_t)(k0 + 4 * u) * lda);
            const char* Bk = (const char*)(B + (int64_t)(k0 + 4 * u) * ldb);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                av[u][t] = *(const double*)(Ak + oa + 16 * t * sizeof(double));
                bv[u][t] = *(const double*)(Bk + ob + 16 * t * sizeof(double));
            }
        }
    };
    auto mm8 = [&](const double (&av)[2][4], const double (&bv)[2][4]) {
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int ti = 0; ti < 4; ti++)
#pragma unroll
                for (int tj = 0; tj < 4; tj++)
                    if (!LOWER || tj <= ti)
                        acc[ti * 4 + tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(al * av[u][ti], bv[u][tj],
                                                                                acc[ti * 4 + tj], 0, 0, 0);
    };
    int k0 = 0;
    if (K >= 8) {
        double ac[2][4], bc[2][4];
        load8(0, ac, bc);
        for (; k0 + 16 <= K; k0 += 8) {
            double an[2][4], bn[2][4];
            load8(k0 + 8, an, bn);
            mm8(ac, bc);
#pragma unroll
            for (int u = 0; u < 2; u++)
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    ac[u][t] = an[u][t];
                    bc[u][t] = bn[u][t];
                }
        }
        mm8(ac, bc);
        k0 += 8;
    }
    // remaining k steps: rows clamped into range, zero beyond K
    for (; k0 < K; k0 += 4) {
        const int k = k0 + k4;
        const bool kv = k < K;
        const int64_t kc = kv ? k : K - 1;
        double av[4], bv[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            av[t] = kv ? al * A[kc * lda + r16 + 16 * t] : 0.0;
            bv[t] = kv ? B[kc * ldb + r16 + 16 * t] : 0.0;
        }
#pragma unroll
        for (int ti = 0; ti < 4; ti++)
#pragma unroll
            for (int tj = 0; tj < 4; tj++)
                if (!LOWER || tj <= ti)
                    acc[ti * 4 + tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ti], bv[tj], acc[ti * 4 + tj], 0, 0, 0);
    }
}

// acc = A^T acc for a lower triangular A^T: A(k, i) = L[k * 64 + i], zero for
// k > i (an inverse factor stored transposed).  The accumulators are the B
// operand: k step kk = 4 tb + v (rows 4 kk .. 4 kk + 3) of column tile tj is
// element v of acc[tb * 4 + tj], exactly the operand layout of the matrix
// instruction.  Output row tile ti reads row tiles 0..ti, so the row tiles
// are produced from the last to the first, each overwriting its own rows
// once no later output reads them; every A value feeds four instructions.
__device__ __forceinline__ void wave_trmm64_lt(mra_d4 (&acc)[16], const double* L) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, k4 = lane >> 4;
    const double* Ll = L + k4 * 64 + r16;
#pragma unroll
    for (int ti = 3; ti >= 0; ti--) {
        double a[16];
#pragma unroll
        for (int kk = 0; kk < 4 * ti + 4; kk++) a[kk] = Ll[kk * 256 + 16 * ti];
        mra_d4 o[4];
#pragma unroll
        for (int tj = 0; tj < 4; tj++) o[tj] = mra_d4{0, 0, 0, 0};
#pragma unroll
        for (int kk = 0; kk < 4 * ti + 4; kk++)
#pragma unroll
            for (int tj = 0; tj < 4; tj++)
                o[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kk], acc[(kk >> 2) * 4 + tj][kk & 3], o[tj], 0, 0, 0);
#pragma unroll
        for (int tj = 0; tj < 4; tj++) acc[ti * 4 + tj] = o[tj];
    }
}

// wave_mm64_acc with al = 1 that also accumulates w = B^T z (z: K values):
// z is the A operand of four more matrix instructions per k step, the same
// in all 16 rows of A, so every row of wacc[tj] holds w[16 tj + column];
// the B values are the ones already loaded for the product.
__device__ __forceinline__ void wave_mm64_acc_w(mra_d4 (&acc)[16], mra_d4 (&wacc)[4], int K, const double* A,
                                                int64_t lda, const double* B, int64_t ldb, const double* z) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, k4 = lane >> 4;
    const uint32_t oa = (uint32_t)((k4 * lda + r16) * sizeof(double));
    const uint32_t ob = (uint32_t)((k4 * ldb + r16) * sizeof(double));
    int k0 = 0;
    for (; k0 + 8 <= K; k0 += 8) {
        double av[2][4], bv[2][4], zv[2];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const char* Ak = (const char*)(A + (int64_t)(k0 + 4 * u) * lda);
            const char* Bk = (const char*)(B + (int64_t)(k0 + 4 * u) * ldb);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                av[u][t] = *(const double*)(Ak + oa + 16 * t * sizeof(double));
                bv[u][t] = *(const double*)(Bk + ob + 16 * t * sizeof(double));
            }
            zv[u] = z[k0 + 4 * u + k4];
        }
#pragma unroll
        for (int u = 0; u < 2; u++) {
#pragma unroll
            for (int ti = 0; ti < 4; ti++)
#pragma unroll
                for (int tj = 0; tj < 4; tj++)
                    acc[ti * 4 + tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u][ti], bv[u][tj], acc[ti * 4 + tj],
                                                                            0, 0, 0);
#pragma unroll
            for (int tj = 0; tj < 4; tj++)
                wacc[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(zv[u], bv[u][tj], wacc[tj], 0, 0, 0);
        }
    }
    for (; k0 < K; k0 += 4) {
        const int k = k0 + k4;
        const bool kv = k < K;
        const int64_t kc = kv ? k : K - 1;
        double av[4], bv[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            av[t] = kv ? A[kc * lda + r16 + 16 * t] : 0.0;
            bv[t] = kv ? B[kc * ldb + r16 + 16 * t] : 0.0;
        }
        const double zv = kv ? z[kc] : 0.0;
#pragma unroll
        for (int ti = 0; ti < 4; ti++)
#pragma unroll
            for (int tj = 0; tj < 4; tj++)
                acc[ti * 4 + tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ti], bv[tj], acc[ti * 4 + tj], 0, 0, 0);
#pragma unroll
        for (int tj = 0; tj < 4; tj++) wacc[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(zv, bv[tj], wacc[tj], 0, 0, 0);
    }
}

// acc += S (a 64 x 64 row-major block with leading dimension lds), in the
// register layout of wave_mm64_acc.
__device__ __forceinline__ void wave_add64(mra_d4 (&acc)[16], const double* S, int lds) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < 4; ti++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const double* srow = S + (int64_t)(ti * 16 + k4 + 4 * v) * lds + r16;
#pragma unroll
            for (int tj = 0; tj < 4; tj++) acc[ti * 4 + tj][v] += srow[16 * tj];
        }
}

// C = acc (64 x 64 row-major, leading dimension ldc), TO = double or float.
template <typename TO>
__device__ __forceinline__ void wave_store64(TO* C, int ldc, const mra_d4 (&acc)[16]) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < 4; ti++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            TO* c = C + (int64_t)(ti * 16 + k4 + 4 * v) * ldc + r16;
#pragma unroll
            for (int tj = 0; tj < 4; tj++) c[16 * tj] = (TO)acc[ti * 4 + tj][v];
        }
}

// ---- the same 64 x 64 accumulators in the layout of the FP32 matrix
// instruction ("L32"): tile (ti, tj), element v of lane t holds row
// 16 ti + 4 (t/16) + v, column 16 tj + t%16.

// f += A^T B over K rows with FP32 matrix instructions (mixed precision:
// operands rounded to float, FP32 sums), f in the L32 layout.  TI: the
// element type of A and B in memory (double, or an FP32 copy).
template <typename TI>
__device__ __forceinline__ void wave_mm64_f32(mra_f4 (&f)[16], int K, const TI* A, int64_t lda, const TI* B,
                                              int64_t ldb) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, k4 = lane >> 4;
    const uint32_t oa = (uint32_t)((k4 * lda + r16) * sizeof(TI));
    const uint32_t ob = (uint32_t)((k4 * ldb + r16) * sizeof(TI));
    // operands of 8 rows (two k steps); the next 8 rows are loaded before the
    // products of the current ones are issued
    auto load8 = [&](int k0, TI (&av)[2][4], TI (&bv)[2][4]) {
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const char* Ak = (const char*)(A + (int64_t)(k0 + 4 * u) * lda);
            const char* Bk = (const char*)(B + (int64_t)(k0 + 4 * u) * ldb);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                av[u][t] = *(const TI*)(Ak + oa + 16 * t * sizeof(TI));
                bv[u][t] = *(const TI*)(Bk + ob + 16 * t * sizeof(TI));
            }
        }
    };
    auto mm8 = [&](const TI (&av)[2][4], const TI (&bv)[2][4]) {
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int ti = 0; ti < 4; ti++)
#pragma unroll
                for (int tj = 0; tj < 4; tj++)
                    f[ti * 4 + tj] = __builtin_amdgcn_mfma_f32_16x16x4f32((float)av[u][ti], (float)bv[u][tj],
                                                                          f[ti * 4 + tj], 0, 0, 0);
    };
    int k0 = 0;
    if (K >= 8) {
        TI ac[2][4], bc[2][4];
        load8(0, ac, bc);
        for (; k0 + 16 <= K; k0 += 8) {
            TI an[2][4], bn[2][4];
            load8(k0 + 8, an, bn);
            mm8(ac, bc);
#pragma unroll
            for (int u = 0; u < 2; u++)
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    ac[u][t] = an[u][t];
                    bc[u][t] = bn[u][t];
                }
        }
        mm8(ac, bc);
        k0 += 8;
    }
    // remaining rows: clamped into range, zero beyond K
    for (; k0 < K; k0 += 4) {
        const int k = k0 + k4;
        const bool kv = k < K;
        const int64_t kk = kv ? k : K - 1;
        float av[4], bv[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            av[t] = kv ? (float)A[kk * lda + r16 + 16 * t] : 0.0f;
            bv[t] = kv ? (float)B[kk * ldb + r16 + 16 * t] : 0.0f;
        }
#pragma unroll
        for (int ti = 0; ti < 4; ti++)
#pragma unroll
            for (int tj = 0; tj < 4; tj++)
                f[ti * 4 + tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[ti], bv[tj], f[ti * 4 + tj], 0, 0, 0);
    }
}

// ---- one wave per finest conditioning tile (FP32, L32 layout): RP = 16 KT
// knot rows (KT = 2 or 4 row tiles) x NTT column tiles of 16 observations,
// acc[ti * NTT + tj].

// acc -= A^T B over K rows (K a multiple of 4 U) on FP32 matrix
// instructions: A(k, i) = A[k * 16 KT + i], B(k, c) = B[k * ldb + c] for the
// columns c = 16 tj + t%16; a column with cv[tj] false is read at column 0
// and contributes zero.  Operands come in stages of U k steps (4 U rows);
// the next stage is loaded before the products of the current one are
// issued.
template <int NTT, int U = 4, int KT = 4>
__device__ __forceinline__ void wave_cond_f32(mra_f4 (&acc)[KT * NTT], int K, const float* A, const float* B,
                                              int64_t ldb, const bool (&cv)[NTT]) {
    constexpr int RP = 16 * KT;
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, k4 = lane >> 4;
    const uint32_t oa = (uint32_t)((k4 * RP + r16) * sizeof(float));
    uint32_t ob[NTT];
#pragma unroll
    for (int tj = 0; tj < NTT; tj++) ob[tj] = (uint32_t)((k4 * ldb + (cv[tj] ? 16 * tj + r16 : 0)) * sizeof(float));
    struct Stage {
        float a[U][KT], b[U][NTT];
    };
    auto load = [&](int k0, Stage& p) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const char* Ak = (const char*)(A + (int64_t)(k0 + 4 * u) * RP);
            const char* Bk = (const char*)(B + (int64_t)(k0 + 4 * u) * ldb);
#pragma unroll
            for (int t = 0; t < KT; t++) p.a[u][t] = *(const float*)(Ak + oa + 16 * t * sizeof(float));
#pragma unroll
            for (int tj = 0; tj < NTT; tj++) p.b[u][tj] = *(const float*)(Bk + ob[tj]);
        }
    };
    auto mm = [&](const Stage& p) {
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int tj = 0; tj < NTT; tj++) {
                const float b = cv[tj] ? -p.b[u][tj] : 0.0f;
#pragma unroll
                for (int ti = 0; ti < KT; ti++)
                    acc[ti * NTT + tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(p.a[u][ti], b, acc[ti * NTT + tj], 0, 0, 0);
            }
    };
    Stage cur;
    load(0, cur);
    int k0 = 0;
    for (; k0 + 8 * U <= K; k0 += 4 * U) {
        Stage nxt;
        load(k0 + 4 * U, nxt);
        mm(cur);
        cur = nxt;
    }
    mm(cur);
}

// acc = L acc with L (16 KT x 16 KT) lower triangular, given transposed in
// FP64 (L[i][k] = Lt[k * 16 KT + i], rounded to float): the accumulators are
// the B operand, k step (ti, v) of column tile tj being element v of
// acc[ti * NTT + tj] (rows 16 ti + 4 (t/16) + v, so A is read at those k);
// output row tiles from the last to the first, each overwriting its rows
// once no later output reads them.
template <int NTT, int KT = 4>
__device__ __forceinline__ void wave_trmm_l32(mra_f4 (&acc)[KT * NTT], const double* Lt) {
    constexpr int RP = 16 * KT;
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, k4 = lane >> 4;
    const double* Ll = Lt + 4 * k4 * RP + r16;
#pragma unroll
    for (int to = KT - 1; to >= 0; to--) {
        float a[4 * KT];
#pragma unroll
        for (int ti = 0; ti <= to; ti++)
#pragma unroll
            for (int v = 0; v < 4; v++) a[4 * ti + v] = (float)Ll[(16 * ti + v) * RP + 16 * to];
        mra_f4 o[NTT];
#pragma unroll
        for (int tj = 0; tj < NTT; tj++) o[tj] = mra_f4{0, 0, 0, 0};
#pragma unroll
        for (int ti = 0; ti <= to; ti++)
#pragma unroll
            for (int v = 0; v < 4; v++)
#pragma unroll
                for (int tj = 0; tj < NTT; tj++)
                    o[tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[4 * ti + v], acc[ti * NTT + tj][v], o[tj], 0, 0, 0);
#pragma unroll
        for (int tj = 0; tj < NTT; tj++) acc[to * NTT + tj] = o[tj];
    }
}

// wave_mm64_acc_w with the product in FP32 (as wave_mm64_f32, into f; A of
// element type TIA) and w = B^T z in FP64 (every row of wacc[tj] holds it).
template <typename TIA>
__device__ __forceinline__ void wave_mm64_w_f32(mra_f4 (&f)[16], mra_d4 (&wacc)[4], int K, const TIA* A,
                                                int64_t lda, const double* B, int64_t ldb, const double* z) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, k4 = lane >> 4;
    const uint32_t oa = (uint32_t)((k4 * lda + r16) * sizeof(TIA));
    const uint32_t ob = (uint32_t)((k4 * ldb + r16) * sizeof(double));
    auto load8 = [&](int k0, TIA (&av)[2][4], double (&bv)[2][4], double (&zv)[2]) {
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const char* Ak = (const char*)(A + (int64_t)(k0 + 4 * u) * lda);
            const char* Bk = (const char*)(B + (int64_t)(k0 + 4 * u) * ldb);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                av[u][t] = *(const TIA*)(Ak + oa + 16 * t * sizeof(TIA));
                bv[u][t] = *(const double*)(Bk + ob + 16 * t * sizeof(double));
            }
            zv[u] = z[k0 + 4 * u + k4];
        }
    };
    auto mm8 = [&](const TIA (&av)[2][4], const double (&bv)[2][4], const double (&zv)[2]) {
#pragma unroll
        for (int u = 0; u < 2; u++) {
#pragma unroll
            for (int ti = 0; ti < 4; ti++)
#pragma unroll
                for (int tj = 0; tj < 4; tj++)
                    f[ti * 4 + tj] = __builtin_amdgcn_mfma_f32_16x16x4f32((float)av[u][ti], (float)bv[u][tj],
                                                                          f[ti * 4 + tj], 0, 0, 0);
#pragma unroll
            for (int tj = 0; tj < 4; tj++)
                wacc[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(zv[u], bv[u][tj], wacc[tj], 0, 0, 0);
        }
    };
    int k0 = 0;
    if (K >= 8) {
        TIA ac[2][4];
        double bc[2][4], zc[2];
        load8(0, ac, bc, zc);
        for (; k0 + 16 <= K; k0 += 8) {
            TIA an[2][4];
            double bn[2][4], zn[2];
            load8(k0 + 8, an, bn, zn);
            mm8(ac, bc, zc);
#pragma unroll
            for (int u = 0; u < 2; u++) {
                zc[u] = zn[u];
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    ac[u][t] = an[u][t];
                    bc[u][t] = bn[u][t];
                }
            }
        }
        mm8(ac, bc, zc);
        k0 += 8;
    }
    for (; k0 < K; k0 += 4) {
        const int k = k0 + k4;
        const bool kv = k < K;
        const int64_t kk = kv ? k : K - 1;
        float av[4];
        double bv[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            av[t] = kv ? (float)A[kk * lda + r16 + 16 * t] : 0.0f;
            bv[t] = kv ? B[kk * ldb + r16 + 16 * t] : 0.0;
        }
        const double zv = kv ? z[kk] : 0.0;
#pragma unroll
        for (int ti = 0; ti < 4; ti++)
#pragma unroll
            for (int tj = 0; tj < 4; tj++)
                f[ti * 4 + tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[ti], (float)bv[tj], f[ti * 4 + tj], 0, 0, 0);
#pragma unroll
        for (int tj = 0; tj < 4; tj++) wacc[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(zv, bv[tj], wacc[tj], 0, 0, 0);
    }
}

// acc = f widened to FP64 (L32 layout).
__device__ __forceinline__ void wave_widen64(mra_d4 (&acc)[16], const mra_f4 (&f)[16]) {
#pragma unroll
    for (int q = 0; q < 16; q++)
#pragma unroll
        for (int v = 0; v < 4; v++) acc[q][v] = (double)f[q][v];
}

// acc += al A^T B in FP64 with the result in the L32 layout: the rows of A
// are permuted so that the FP64 instruction's row t/16 + 4v of lane t is
// row 4 (t/16) + v (row r of a tile read as row 4 (r%4) + r/4).
__device__ __forceinline__ void wave_mm64_acc_l32(mra_d4 (&acc)[16], int K, double al, const double* A,
                                                  int64_t lda, const double* B, int64_t ldb) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15;
    const int rp = 4 * (r16 & 3) + (r16 >> 2);
    // A is read at columns permuted by rp: shift A so that wave_mm64_acc's
    // column offset r16 lands on rp
    const double* Ap = A + (rp - r16);
    wave_mm64_acc(acc, K, al, Ap, lda, B, ldb);
}

// acc += S in the L32 layout (S 64 x 64 row-major, leading dimension lds).
__device__ __forceinline__ void wave_add64_l32(mra_d4 (&acc)[16], const double* S, int lds) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < 4; ti++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const double* srow = S + (int64_t)(ti * 16 + 4 * k4 + v) * lds + r16;
#pragma unroll
            for (int tj = 0; tj < 4; tj++) acc[ti * 4 + tj][v] += srow[16 * tj];
        }
}

// acc += S^T in the L32 layout.
__device__ __forceinline__ void wave_add64_t_l32(mra_d4 (&acc)[16], const double* S, int lds) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < 4; ti++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const double* scol = S + (int64_t)r16 * lds + ti * 16 + 4 * k4 + v;
#pragma unroll
            for (int tj = 0; tj < 4; tj++) acc[ti * 4 + tj][v] += scol[16 * tj * lds];
        }
}

// C = acc (L32 layout; 64 x 64 row-major, leading dimension ldc).
__device__ __forceinline__ void wave_store64_l32(double* C, int ldc, const mra_d4 (&acc)[16]) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < 4; ti++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            double* c = C + (int64_t)(ti * 16 + 4 * k4 + v) * ldc + r16;
#pragma unroll
            for (int tj = 0; tj < 4; tj++) c[16 * tj] = acc[ti * 4 + tj][v];
        }
}

// C = al1 * A1^T B1 + al2 * A2^T B2 stored to C (M x N, all tiles in range:
// Mv >= M, Nv >= N), both products accumulated in registers and written
// once: no LDS, no barrier.  FP64.
__device__ __forceinline__ void wg_gemm2_store(double* C, int ldc, int M, int N, int K1, double al1, const double* A1,
                                               int64_t lda1, const double* B1, int64_t ldb1, int K2, double al2,
                                               const double* A2, int64_t lda2, const double* B2, int64_t ldb2) {
    typedef Mfma<double>::V V;
    const int lane = threadIdx.x & 63, wave = wave_id();
    const int r16 = lane & 15, k4 = lane >> 4;
    const int tM = M >> 4, tN = N >> 4;
    const int nt = tM * tN;
    const int bh = (nt >= 16 || (nt >= 8 && tM >= 2)) ? 2 : 1;
    const int bw = (nt >= 16) ? 2 : 1;
    const int bM = (tM + bh - 1) / bh, bN = (tN + bw - 1) / bw;
    for (int blk = wave; blk < bM * bN; blk += 4) {
        const int ti0 = (blk / bN) * bh, tj0 = (blk % bN) * bw;
        const int ia = ti0 * 16 + r16, ib = ia + 16;
        const int ja = tj0 * 16 + r16, jb = ja + 16;
        const bool tM2 = bh == 2 && ti0 + 1 < tM, tN2 = bw == 2 && tj0 + 1 < tN;
        V acc1[4], acc2[4];
#pragma unroll
        for (int q = 0; q < 4; q++) acc1[q] = acc2[q] = V{0, 0, 0, 0};
        const BlockOps<false, double, double> o1(A1, lda1, B1, ldb1, ia, ib, ja, jb, true, tM2, true, tN2, k4);
        block_kloop_any<double, false, 4>(acc1, o1, tM2, tN2, 0, K1, k4);
        const BlockOps<false, double, double> o2(A2, lda2, B2, ldb2, ia, ib, ja, jb, true, tM2, true, tN2, k4);
        block_kloop_any<double, false, 4>(acc2, o2, tM2, tN2, 0, K2, k4);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = ti0 * 16 + k4 + 4 * v;
            double* c0 = C + (int64_t)row * ldc + ja;
            c0[0] = al1 * acc1[0][v] + al2 * acc2[0][v];
            if (tN2) c0[16] = al1 * acc1[1][v] + al2 * acc2[1][v];
            if (tM2) {
                double* c1 = c0 + (int64_t)16 * ldc;
                c1[0] = al1 * acc1[2][v] + al2 * acc2[2][v];
                if (tN2) c1[16] = al1 * acc1[3][v] + al2 * acc2[3][v];
            }
        }
    }
}

// FP64 MFMA, FP64 accumulation; operands read as TIA / TIB and widened
template <bool TA = false, bool ACC = true, typename TIA = double, typename TIB = double, int TILES = GEMM_PART,
          bool DUAL = false, int TRI = 0, bool PIPE = false>
__device__ __forceinline__ void wg_gemm_tn(double* C, int ldc, int M, int N, int Kv, double alpha, const TIA* A,
                                           int64_t lda, int Mv, const TIB* B, int64_t ldb, int Nv,
                                           float* C32 = nullptr) {
    wg_gemm_tn_impl<double, TA, ACC, TIA, TIB, TILES, double, false, DUAL, TRI, PIPE>(C, ldc, M, N, Kv, alpha, A, lda,
                                                                                     Mv, B, ldb, Nv, C32);
}

// FP32 MFMA with FP32 accumulation (operands rounded to float); C of type TC
template <bool ACC = true, typename TIA = float, typename TIB = float, typename TC = float, int TRI = 0,
          bool PIPE = false>
__device__ __forceinline__ void wg_gemm_tn_f32(TC* C, int ldc, int M, int N, int Kv, float alpha, const TIA* A,
                                               int64_t lda, int Mv, const TIB* B, int64_t ldb, int Nv) {
    wg_gemm_tn_impl<float, false, ACC, TIA, TIB, GEMM_ANY, TC, false, false, TRI, PIPE>(C, ldc, M, N, Kv, alpha, A,
                                                                                       lda, Mv, B, ldb, Nv);
}

// G = L^{-1} B for one 64 x 64 block, work balanced over the triangle: Li
// holds L^{-1} column-major (Li[k*64 + i] = L^{-1}(i, k), zero for k > i), B
// is 64 x 64 in LDS (leading dimension ldb).  Wave w computes output columns
// 16w .. 16w+15 for all four row tiles; row tile ti stops at k < 16(ti+1), so
// every wave issues the same 40 matrix instructions (a 2 x 2 tile block per
// wave would give two waves K = 64 and two K = 32).  G (ld 64) is written
// from the accumulators, and with G32 also its FP32 rounding.
__device__ __forceinline__ void wg_trmm64_cols(double* G, float* G32, const double* Li, const double* B, int ldb) {
    const int lane = threadIdx.x & 63, wave = wave_id();
    const int r16 = lane & 15, k4 = lane >> 4;
    const int col = 16 * wave + r16;
    // wave-uniform bases plus 32-bit lane offsets (row k4, column r16 / col)
    const uint32_t oa = (uint32_t)((k4 * 64 + r16) * sizeof(double));
    const uint32_t og = (uint32_t)((k4 * 64 + col) * sizeof(double));
    const double* Bl = B + k4 * ldb + col;
    mra_d4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = mra_d4{0, 0, 0, 0};
    // 8 k rows at a time (row tiles kh/2 .. 3), the next 8 rows' operands
    // loaded before the current products are issued
    auto load = [&](const int kh, double (&a)[2][4], double (&b)[2]) {
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const int k0 = 8 * kh + 4 * s;
            const char* Lk = (const char*)(Li + k0 * 64);
            b[s] = Bl[k0 * ldb];
#pragma unroll
            for (int t = kh >> 1; t < 4; t++) a[s][t] = *(const double*)(Lk + oa + 16 * t * sizeof(double));
        }
    };
    auto mm = [&](const int kh, const double (&a)[2][4], const double (&b)[2]) {
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int t = kh >> 1; t < 4; t++) acc[t] = Mfma<double>::op(a[s][t], b[s], acc[t]);
    };
    double a0[2][4], b0[2], a1[2][4], b1[2];
    load(0, a0, b0);
#pragma unroll
    for (int kh = 0; kh < 8; kh += 2) {
        load(kh + 1, a1, b1);
        mm(kh, a0, b0);
        if (kh + 2 < 8) load(kh + 2, a0, b0);
        mm(kh + 1, a1, b1);
    }
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            // row 16 t + k4 + 4 v
            const int rb = 16 * t + 4 * v;
            *(double*)((char*)(G + rb * 64) + og) = acc[t][v];
            if (G32) *(float*)((char*)(G32 + rb * 64) + og / 2) = (float)acc[t][v];
        }
}

// C -= G^T G on the ten 16 x 16 tiles on and below the diagonal of a 64 x 64
// C in LDS (ldc); G is K x 64 row-major (ld 64), K a multiple of 32.  Column
// tiles {0, 3} and {1, 2} hold five lower tiles each; waves 0 / 1 take them
// over k < K/2, waves 2 / 3 over k >= K/2 (320 matrix instructions per wave
// at K = 512 where a 2 x 2 block per wave takes 512).  A and B operands of
// the symmetric product are the same loads.  The first half is subtracted
// before the second (a work-group barrier between), so the result does not
// depend on timing.  Called by all 256 threads.
__device__ __forceinline__ void wg_syrk64_lower(double* C, int ldc, const double* G, int K) {
    const int lane = threadIdx.x & 63, wave = wave_id();
    const int r16 = lane & 15, k4 = lane >> 4;
    const bool outer = (wave & 1) == 0;  // column tiles {0, 3}, else {1, 2}
    const int kh = K >> 1, kb = (wave >> 1) * kh;
    // tiles (row tile, column tile): outer (0,0) (1,0) (2,0) (3,0) (3,3);
    // inner (1,1) (2,1) (3,1) (2,2) (3,2)
    mra_d4 acc[5];
#pragma unroll
    for (int t = 0; t < 5; t++) acc[t] = mra_d4{0, 0, 0, 0};
    const double* Gl = G + (int64_t)k4 * 64 + r16;
    // 8 k rows at a time, the next 8 rows loaded before the current products
    // are issued (kh is a multiple of 16)
    if (outer) {
        auto ld = [&](int k0, double (&g)[2][4]) {
#pragma unroll
            for (int s = 0; s < 2; s++)
#pragma unroll
                for (int t = 0; t < 4; t++) g[s][t] = Gl[(int64_t)(k0 + 4 * s) * 64 + 16 * t];
        };
        auto mm = [&](const double (&g)[2][4]) {
#pragma unroll
            for (int s = 0; s < 2; s++) {
                acc[0] = Mfma<double>::op(g[s][0], g[s][0], acc[0]);
                acc[1] = Mfma<double>::op(g[s][1], g[s][0], acc[1]);
                acc[2] = Mfma<double>::op(g[s][2], g[s][0], acc[2]);
                acc[3] = Mfma<double>::op(g[s][3], g[s][0], acc[3]);
                acc[4] = Mfma<double>::op(g[s][3], g[s][3], acc[4]);
            }
        };
        double g0[2][4], g1[2][4];
        ld(kb, g0);
        for (int k0 = kb; k0 < kb + kh; k0 += 16) {
            ld(k0 + 8, g1);
            mm(g0);
            if (k0 + 16 < kb + kh) ld(k0 + 16, g0);
            mm(g1);
        }
    } else {
        auto ld = [&](int k0, double (&g)[2][3]) {
#pragma unroll
            for (int s = 0; s < 2; s++)
#pragma unroll
                for (int t = 0; t < 3; t++) g[s][t] = Gl[(int64_t)(k0 + 4 * s) * 64 + 16 * (t + 1)];
        };
        auto mm = [&](const double (&g)[2][3]) {
#pragma unroll
            for (int s = 0; s < 2; s++) {
                acc[0] = Mfma<double>::op(g[s][0], g[s][0], acc[0]);
                acc[1] = Mfma<double>::op(g[s][1], g[s][0], acc[1]);
                acc[2] = Mfma<double>::op(g[s][2], g[s][0], acc[2]);
                acc[3] = Mfma<double>::op(g[s][1], g[s][1], acc[3]);
                acc[4] = Mfma<double>::op(g[s][2], g[s][1], acc[4]);
            }
        };
        double g0[2][3], g1[2][3];
        ld(kb, g0);
        for (int k0 = kb; k0 < kb + kh; k0 += 16) {
            ld(k0 + 8, g1);
            mm(g0);
            if (k0 + 16 < kb + kh) ld(k0 + 16, g0);
            mm(g1);
        }
    }
    auto sub = [&]() {
#pragma unroll
        for (int t = 0; t < 5; t++) {
            const int ti = outer ? (t < 4 ? t : 3) : (t < 3 ? t + 1 : t - 1);
            const int tj = outer ? (t < 4 ? 0 : 3) : (t < 3 ? 1 : 2);
#pragma unroll
            for (int v = 0; v < 4; v++)
                C[(16 * ti + Mfma<double>::row(k4, v)) * ldc + 16 * tj + r16] -= acc[t][v];
        }
    };
    if (wave < 2) sub();
    __syncthreads();
    if (wave >= 2) sub();
}

// Same product for an output of at most 32 x 32 (one 2x2 tile block), with
// K split over the 4 waves; the partial tiles are summed in wave order
// through red (LDS: 3*256 doubles for a 16 x 16 output, else 3*4*256), so
// the result is reproducible.
template <typename TIA = double, typename TIB = TIA>
__device__ __forceinline__ void wg_gemm_tn_ksplit(double* C, int ldc, int M, int N, int Kv, double alpha,
                                                  const TIA* A, int64_t lda, int Mv, const TIB* B, int64_t ldb,
                                                  int Nv, double* red) {
    const int lane = threadIdx.x & 63, wave = wave_id();
    const int r16 = lane & 15, k4 = lane >> 4;
    const int tM = M >> 4, tN = N >> 4;
    const bool tM2 = tM > 1, tN2 = tN > 1;
    mra_d4 acc[4];
#pragma unroll
    for (int q = 0; q < 4; q++) acc[q] = mra_d4{0, 0, 0, 0};
    const int ia = r16, ib = 16 + r16, ja = r16, jb = 16 + r16;
    const bool ha = ia < Mv, hb = tM2 && ib < Mv, hja = ja < Nv, hjb = tN2 && jb < Nv;
    const BlockOps<false, TIA, TIB> o(A, lda, B, ldb, ia, ib, ja, jb, ha, hb, hja, hjb, k4);
    // this wave's contiguous K range, a multiple of 4
    const int kq = ((Kv + 15) / 16) * 4;
    const int kb = min(Kv, wave * kq), ke = min(Kv, kb + kq);
    block_kloop_any<double, true, 6>(acc, o, tM2, tN2, kb, ke, k4);
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

// layout of the inverse diagonal blocks for nb = N/16 blocks
struct BlkLayout {
    int nb;
    __device__ bool upper() const { return nb >= 3; }
    __device__ double* dptr(double* A, int lda, double* dbuf, int b) const {
        if (upper()) return b < nb - 1 ? A + b * 16 * lda + (b + 1) * 16 : A + (nb - 1) * 16;
        return dbuf + b * 256;
    }
    __device__ int ldd(int lda) const { return upper() ? lda : 16; }
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

// 1/x for a positive normal x: the hardware reciprocal refined by two Newton
// steps (r += r (1 - x r)), five dependent instructions where the IEEE
// division takes about ten; within an ulp or so of the division.
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}

// One wave: factor the 16 x 16 diagonal block P in place (lower triangle)
// and write its inverse (lower, zeros above) to Dp (row stride ldd).  The
// column chain carries no IEEE division: the pivot reciprocals by rcp_nr,
// and the inverse multiplies by 1/L_ii (staged on Dp's diagonal).
__device__ __forceinline__ void wave_chol16(double* P, int lda, double* Dp, int ldd, int* bad) {
    const int lane = threadIdx.x & 63, i = lane & 15, cg = lane >> 4;
    // unscaled right-looking elimination of the lower triangle:
    // A[i][j] -= A[i][k] A[j][k] / A[k][k]  (k < j <= i)
    for (int k = 0; k < 15; k++) {
        wave_lds_fence();
        double akk = P[k * lda + k];
        if (!(akk > 0.0)) akk = 1.0;
        const double inv = rcp_nr(akk);
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
        if (j == i) {
            const double lii = sqrt(pj);
            P[i * lda + i] = lii;
            Dp[i * ldd + i] = 1.0 / lii;
        }
    }
    wave_lds_fence();
    // inverse by forward substitution, one column per lane (lanes 0..15):
    // the column is kept in registers and the factor entries are read as
    // LDS broadcasts that do not wait on earlier results
    if (lane < 16) {
        const int c = lane;
        double x[16];
#pragma unroll
        for (int ii = 0; ii < 16; ii++) {
            double sacc = ii == c ? 1.0 : 0.0;
#pragma unroll
            for (int kk = 0; kk < ii; kk++) sacc -= P[ii * lda + kk] * x[kk];
            x[ii] = sacc * Dp[ii * ldd + ii];
        }
#pragma unroll
        for (int ii = 0; ii < 16; ii++) Dp[ii * ldd + c] = x[ii];
    }
    wave_lds_fence();
}

// One wave: 32 x 32 lower Cholesky of P in place (row stride ld) and its
// inverse D = L^{-1} (lower, zeros above, row stride ld): the two 16 x 16
// diagonal blocks by wave_chol16, L_10 = P_10 D_00^T, P_11 -= L_10 L_10^T and
// D_10 = -D_11 (L_10 D_00).  Wave-synchronous (no work-group barrier).
__device__ __forceinline__ void wave_chol32(double* P, double* D, int ld, int* bad) {
    const int lane = threadIdx.x & 63, r16 = lane & 15, k4 = lane >> 4;
    double* P10 = P + 16 * ld;
    double* P11 = P10 + 16;
    double* D10 = D + 16 * ld;
    double* D11 = D10 + 16;
    wave_chol16(P, ld, D, ld, bad);
    mra_d4 acc = {0, 0, 0, 0};
    acc = mfma16<true>(acc, P10, ld, D, ld);
    wave_lds_fence();
#pragma unroll
    for (int v = 0; v < 4; v++) P10[(k4 + 4 * v) * ld + r16] = acc[v];
    wave_lds_fence();
    acc = mra_d4{0, 0, 0, 0};
    acc = mfma16<true>(acc, P10, ld, P10, ld);
#pragma unroll
    for (int v = 0; v < 4; v++) P11[(k4 + 4 * v) * ld + r16] -= acc[v];
    wave_lds_fence();
    wave_chol16(P11, ld, D11, ld, bad);
    // T = L_10 D_00 in the D_10 block, then D_10 = -D_11 T
    acc = mra_d4{0, 0, 0, 0};
    acc = mfma16<false>(acc, P10, ld, D, ld);
    wave_lds_fence();
#pragma unroll
    for (int v = 0; v < 4; v++) D10[(k4 + 4 * v) * ld + r16] = acc[v];
    wave_lds_fence();
    acc = mra_d4{0, 0, 0, 0};
    acc = mfma16<false>(acc, D11, ld, D10, ld);
    wave_lds_fence();
#pragma unroll
    for (int v = 0; v < 4; v++) {
        D10[(k4 + 4 * v) * ld + r16] = -acc[v];
        D[(k4 + 4 * v) * ld + 16 + r16] = 0.0;  // D_01
    }
    wave_lds_fence();
}

// Cholesky of the leading 16*nb x 16*nb block (nb <= 8).
__device__ __forceinline__ void wg_chol_blk(double* A, int lda, int nb, double* dbuf, int* bad) {
    const BlkLayout BL{nb};
    const int wave = wave_id(), lane = threadIdx.x & 63, r16 = lane & 15, k4 = lane >> 4;
    const int ldd = BL.ldd(lda);
    __syncthreads();
    for (int kb = 0; kb < nb; kb++) {
        double* Pkk = A + kb * 16 * (lda + 1);
        double* Dk = BL.dptr(A, lda, dbuf, kb);
        if (wave == 0) wave_chol16(Pkk, lda, Dk, ldd, bad);
        __syncthreads();
        if (kb == nb - 1) break;
        // panel: L_{ib,kb} = A_{ib,kb} D_kb^T
        for (int ib = kb + 1 + wave; ib < nb; ib += 4) {
            double* Pib = A + ib * 16 * lda + kb * 16;
            mra_d4 acc = {0, 0, 0, 0};
            acc = mfma16<true>(acc, Pib, lda, Dk, ldd);
            wave_lds_fence();
#pragma unroll
            for (int v = 0; v < 4; v++) Pib[(k4 + 4 * v) * lda + r16] = acc[v];
        }
        __syncthreads();
        // trailing update A_{ib,jb} -= L_{ib,kb} L_{jb,kb}^T, kb < jb <= ib
        const int m = nb - 1 - kb;
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

// A <- L^{-1} in place after wg_chol_blk (same nb <= NBMAX, dbuf); the
// strictly upper part of the leading 16*nb block is set to zero.
template <int NBMAX = 8>
__device__ __forceinline__ void wg_trinv_blk(double* A, int lda, int nb, double* dbuf) {
    constexpr int NBW = (NBMAX + 3) / 4;  // block columns per wave
    const BlkLayout BL{nb};
    const int wave = wave_id(), lane = threadIdx.x & 63, r16 = lane & 15, k4 = lane >> 4;
    const int ldd = BL.ldd(lda);
    const int N = 16 * nb;
    __syncthreads();
    for (int i = 1; i < nb; i++) {
        // T_ij = sum_{k=j}^{i-1} L_ik X_kj (X_jj = D_j; X_kj, k > j, final),
        // at most NBW blocks j per wave
        mra_d4 acc[NBW];
#pragma unroll
        for (int t = 0; t < NBW; t++) acc[t] = mra_d4{0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < NBW; t++) {
            const int j = wave + 4 * t;
            if (j < i)
                for (int k = j; k < i; k++) {
                    const double* Lik = A + i * 16 * lda + k * 16;
                    const double* Xkj = k == j ? BL.dptr(A, lda, dbuf, j) : A + k * 16 * lda + j * 16;
                    acc[t] = mfma16<false>(acc[t], Lik, lda, Xkj, k == j ? ldd : lda);
                }
        }
        __syncthreads();  // every L_ij of block row i has been read
#pragma unroll
        for (int t = 0; t < NBW; t++) {
            const int j = wave + 4 * t;
            if (j < i) {
                double* Cb = A + i * 16 * lda + j * 16;
#pragma unroll
                for (int v = 0; v < 4; v++) Cb[(k4 + 4 * v) * lda + r16] = acc[t][v];
                wave_lds_fence();
                // X_ij = -D_i T_ij
                mra_d4 x = {0, 0, 0, 0};
                x = mfma16<false>(x, BL.dptr(A, lda, dbuf, i), ldd, Cb, lda);
                wave_lds_fence();
#pragma unroll
                for (int v = 0; v < 4; v++) Cb[(k4 + 4 * v) * lda + r16] = -x[v];
            }
        }
        __syncthreads();
    }
    // diagonal blocks <- D_b (every D_b is read before any upper block is zeroed)
    double dv[NBMAX];
#pragma unroll
    for (int q = 0; q < NBMAX; q++) {
        const int idx = threadIdx.x + q * MRA_WG;
        dv[q] = 0.0;
        if (idx < nb * 256) {
            const int b = idx >> 8, e = idx & 255;
            dv[q] = BL.dptr(A, lda, dbuf, b)[(e >> 4) * ldd + (e & 15)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NBMAX; q++) {
        const int idx = threadIdx.x + q * MRA_WG;
        if (idx < nb * 256) {
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
