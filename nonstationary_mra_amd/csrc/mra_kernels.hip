// HIP kernels of the MRA likelihood / prediction path (gfx950).
//
// One launch per level and phase, batched over every region of the level
// (the reference runs an OpenMP loop per level over the same regions:
// src/class_approximation.cpp:108-159,
// src/class_approximation-loop-regions-at-finest-level-in-prior.cpp:52-400,
// src/class_approximation-posterior-inference.cpp:51-219,
// src/class_approximation-predict.cpp:105-189).
//
// Layouts (RP = knots per coarse region padded to a multiple of 16; pad
// knots have zero covariance and identity on the knot-matrix diagonal):
//   Gc   coarse G_{R,j}: level-l region block of l*RP*RP, row m (whitened
//        ancestor knot, stacked over j) x column c (region knot)
//   Lt   prior Cholesky factor L_R stored column-major (Lt[k*RP+i] = L[i][k])
//   LiT  L_R^{-1} column-major;   Lpt / LpiT: the posterior factor and inverse
//   Gf   finest G_{F,j}: stored per segment (finest regions sharing a
//        parent), element (m, o) at L*RP*seg_os + m*seg_n + o-seg_os
//   Bf   finest B_{F,j} observation-major: (m, o) at o*L*RP + m
//   H    L_F^{-1} B_F^T, observation-major rows of L*RP
//   Sig  L_F (row-major n_F x n_F) at sig_off[F]
//   msgs w~ (l*RP) and A~ (l(l+1)/2 slices of RP*RP, slice order
//        (0,0),(0,1)..(0,l-1),(1,1).. as src/class_approximation.cpp:62-73)
#include "mra_device.hpp"

namespace mra {

struct Tree {
    int M, J, L, r, RP, LRP;
    int xcd;                  // 1: work groups take items in XCD-contiguous ranges (xcd_block)
    uint64_t nC, nF;
    const uint64_t* parent;   // [nF] global coarse id
    const uint64_t* obs_off;  // [nF+1]
    const uint64_t* pred_off; // [nF+1] or null
    const uint64_t* ch_s;     // [n level L-1] first finest child (local), ch_e < ch_s: none
    const uint64_t* ch_e;
    const uint8_t* working;   // [nC]
    const uint64_t* sig_off;  // [nF]
    // segments: maximal runs of finest regions with the same parent; G_F / B_F
    // of a segment are stored together, element (m, o) at
    // LRP*seg_os + m*seg_n + (o - seg_os)
    const uint64_t* f_seg_os;  // [nF] first observation of F's segment
    const uint32_t* f_seg_n;   // [nF] observations in F's segment
    const uint64_t* s_os;      // [nSeg] first observation of the segment
    const uint32_t* s_n;       // [nSeg]
    const uint64_t* s_parent;  // [nSeg] global coarse id of the shared parent
    // prediction points of F's segment (the same runs of finest regions)
    const uint64_t* p_seg_os;  // [nF] first prediction point of F's segment
    const uint32_t* p_seg_n;   // [nF] prediction points in F's segment
    // groups: maximal runs of finest regions with the same grandparent; B~
    // and G^ of a group are stored together (see k_pred_chain)
    const uint64_t* p_grp_os;  // [nF] first prediction point of F's group
    const uint32_t* p_grp_n;   // [nF] prediction points in F's group
};

// Item of work group b among n when xcd = 1: the device dispatches work
// groups to its 8 XCDs round robin (group b on XCD b mod 8), so group b takes
// item b/8 of the b mod 8-th of 8 contiguous item ranges, and the items of one
// range (neighbouring regions, which read the same ancestor blocks) share one
// XCD's L2.  A bijection of [0, n).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t n, int xcd) {
    if (!xcd) return b;
    const uint32_t x = b & 7u, i = b >> 3, q = n >> 3, r = n & 7u;
    return x * q + (x < r ? x : r) + i;
}

__device__ __forceinline__ uint64_t lvl_start(int l, int J) {
    uint64_t s = 0, c = 1;
    for (int i = 0; i < l; i++) {
        s += c;
        c *= (uint64_t)J;
    }
    return s;
}

// Ancestor at level j of a region at level l (j <= l).
__device__ __forceinline__ uint64_t ancestor_up(uint64_t R, int l, int j, int J) {
    for (int i = j; i < l; i++) R = (R - 1) / (uint64_t)J;
    return R;
}

// ------------------------------------------------------------- point setup
__global__ void k_coords(uint64_t n, const double* lon, const double* lat, int metric, double* x, double* y,
                         double* z) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double lo = lon[i], la = lat[i];
    if (metric == 0) {
        // src/class_partition.cpp:507-517
        double cosLat = cos(la * M_PI / 180), cosLon = cos(lo * M_PI / 180);
        double sinLat = sin(la * M_PI / 180), sinLon = sin(lo * M_PI / 180);
        x[i] = cosLat * cosLon;
        y[i] = cosLat * sinLon;
        z[i] = sinLat;
    } else if (metric == 1) {
        double lar = la * (M_PI / 180);
        x[i] = sin(lar);
        y[i] = cos(lar);
        z[i] = lo * (M_PI / 180);
    } else {
        x[i] = lo;
        y[i] = la;
        z[i] = 0.0;
    }
}

struct FieldDev {
    int kind;
    int nb;
    double intercept, sx, sy;
    const double *bx, *by, *bz, *blon, *blat, *w;
};

// exp(log field) at one point (src/class_nonstationary_covariance.cpp:479-573).
__device__ double field_eval(const FieldDev& f, int metric, double wlen, double lon, double lat, double x, double y,
                             double z) {
    if (f.kind == 2) return exp(f.intercept + f.sx * lon + f.sy * lat);
    double v = f.intercept;
    if (f.kind == 1) {
        for (int i = 0; i < f.nb; i++) {
            double d;
            if (metric == 0) {
                double dx = f.bx[i] - x, dy = f.by[i] - y, dz = f.bz[i] - z;
                d = sqrt(dx * dx + dy * dy + dz * dz);
            } else if (metric == 1) {
                double l1 = lon * (M_PI / 180), l2 = f.blon[i] * (M_PI / 180);
                double a1 = lat * (M_PI / 180), a2 = f.blat[i] * (M_PI / 180);
                double ang = sin(a1) * sin(a2) + cos(a1) * cos(a2) * cos(l1 - l2);
                ang = fmin(1.0, fmax(-1.0, ang));
                d = acos(ang);
            } else {
                double dx = f.blon[i] - lon, dy = f.blat[i] - lat;
                d = sqrt(dx * dx + dy * dy);
            }
            d /= wlen;
            if (d >= 1) continue;
            double om = 1 - d;
            double om6 = om * om * om;
            om6 *= om6;
            v += om6 * (35 * d * d + 18 * d + 3) * f.w[i] / 3;
        }
    }
    return exp(v);
}

__global__ void k_fields(uint64_t n, const double* lon, const double* lat, int metric, double wlen, FieldDev f0,
                         FieldDev f1, FieldDev f2, double* o0, double* o1, double* o2) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double lo = lon[i], la = lat[i];
    double cosLat = cos(la * M_PI / 180), cosLon = cos(lo * M_PI / 180);
    double sinLat = sin(la * M_PI / 180), sinLon = sin(lo * M_PI / 180);
    double x = cosLat * cosLon, y = cosLat * sinLon, z = sinLat;
    o0[i] = field_eval(f0, metric, wlen, lo, la, x, y, z);
    o1[i] = field_eval(f1, metric, wlen, lo, la, x, y, z);
    o2[i] = field_eval(f2, metric, wlen, lo, la, x, y, z);
}

// Scaled per-point parameters (src/evaluate_covariance.cpp:170-189,258-269).
// prm = {sill, range, nugget, sill_scale, range_scale, nugget_scale}, read
// from device memory so that a captured evaluation graph follows parameter
// updates without re-capture
__global__ void k_scale(uint64_t n, const double* f0, const double* f1, const double* f2, int stationary,
                        const double* __restrict__ prm, double* s, double* b, double* v) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (stationary) {
        s[i] = sqrt(prm[0]);
        b[i] = prm[1];
        v[i] = prm[0] + prm[2];
    } else {
        double sg = prm[3] * f0[i];
        s[i] = sqrt(sg);
        b[i] = prm[4] * f1[i];
        v[i] = sg + prm[5] * f2[i];
    }
}

__global__ void k_fill_u64(unsigned long long* p, uint64_t n) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void k_zero_int(int* p, uint64_t n) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_fill(double* p, uint64_t n, double val) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = val;
}

// ------------------------------------------------------ covariance tiles
// Standalone covariance matrix out[i*n2 + j] (row-major n1 x n2, the layout
// evaluate_nonstationary_cross_covariance writes, src/evaluate_covariance.cpp:
// 155-234), the pairwise evaluation on its own as a throughput kernel.
// One work group covers COV_ROWS rows x 64*VW columns, VW = 16 bytes /
// sizeof(T): each lane keeps VW consecutive column points in registers and
// writes one 16-byte nontemporal store per row (the output is streamed, not
// re-read), so each wave writes 1 KiB row segments; the row points are
// staged in LDS (broadcast reads).  The pair formula is the one of the
// engine's tiles, q_i q_j t^3 exp(-d t).  vec = 0 when the output is not
// 16-byte aligned at every row; a work group whose columns reach past n2
// also takes the scalar-store path.
constexpr int COV_ROWS = 64;
typedef double mra_d2 __attribute__((ext_vector_type(2)));

// FP32, nonstationary, planar (PLANAR) or chordal metric: float point data.
template <bool PLANAR>
__global__ __launch_bounds__(256) void k_cov_tile_f32(Pts P, uint64_t n1, Pts Q, uint64_t n2, float* out, int vec) {
    __shared__ PtF rows[COV_ROWS];
    const uint64_t i0 = (uint64_t)blockIdx.y * COV_ROWS;
    const uint64_t jb = (uint64_t)blockIdx.x * 256;
    const uint64_t j0 = jb + (threadIdx.x & 63) * 4;
    const int w = threadIdx.x >> 6;
    const int nr = (int)(n1 - i0 < (uint64_t)COV_ROWS ? n1 - i0 : (uint64_t)COV_ROWS);
    const CovCtx cc{PLANAR ? 2 : 0, 0, 1.0, 1.0};
    for (int i = threadIdx.x; i < nr; i += MRA_WG) rows[i] = to_ptf(load_pt(cc, P, i0 + i));
    PtF qf[4];
    for (int v = 0; v < 4; v++) qf[v] = j0 + v < n2 ? to_ptf(load_pt(cc, Q, j0 + v)) : PtF{0, 0, 0, 0, 1};
    __syncthreads();
    if (vec && jb + 256 <= n2) {
        for (int i = w; i < nr; i += MRA_WG / 64) {
            const PtF pr = rows[i];
            mra_f4 c;
            c.x = pcov_ff<PLANAR>(pr, qf[0], !PLANAR);
            c.y = pcov_ff<PLANAR>(pr, qf[1], !PLANAR);
            c.z = pcov_ff<PLANAR>(pr, qf[2], !PLANAR);
            c.w = pcov_ff<PLANAR>(pr, qf[3], !PLANAR);
            __builtin_nontemporal_store(c, reinterpret_cast<mra_f4*>(out + (i0 + i) * n2 + j0));
        }
    } else {
        for (int i = w; i < nr; i += MRA_WG / 64) {
            const PtF pr = rows[i];
            for (int v = 0; v < 4; v++)
                if (j0 + v < n2) out[(i0 + i) * n2 + j0 + v] = pcov_ff<PLANAR>(pr, qf[v], !PLANAR);
        }
    }
}

// Any metric and precision (FP64 output; FP32 for the great-circle metric).
template <typename T>
__global__ __launch_bounds__(256) void k_cov_tile(CovCtx cc, Pts P, uint64_t n1, Pts Q, uint64_t n2, T* out,
                                                  int vec) {
    constexpr int VW = 16 / sizeof(T);
    __shared__ double st[5 * COV_ROWS];
    const uint64_t i0 = (uint64_t)blockIdx.y * COV_ROWS;
    const uint64_t jb = (uint64_t)blockIdx.x * (64 * VW);
    const uint64_t j0 = jb + (threadIdx.x & 63) * VW;
    const int w = threadIdx.x >> 6;
    const int nr = (int)(n1 - i0 < (uint64_t)COV_ROWS ? n1 - i0 : (uint64_t)COV_ROWS);
    const bool full = vec && jb + 64 * VW <= n2;
    PtV q[VW];
    for (int v = 0; v < VW; v++) q[v] = j0 + v < n2 ? load_pt(cc, Q, j0 + v) : PtV{0, 0, 0, 0, 1};
    for (int i = threadIdx.x; i < nr; i += MRA_WG) {
        const PtV p = load_pt(cc, P, i0 + i);
        st[i] = p.x;
        st[COV_ROWS + i] = p.y;
        st[2 * COV_ROWS + i] = p.z;
        st[3 * COV_ROWS + i] = p.q;
        st[4 * COV_ROWS + i] = p.b;
    }
    __syncthreads();
    for (int i = w; i < nr; i += MRA_WG / 64) {
        const PtV pr = {st[i], st[COV_ROWS + i], st[2 * COV_ROWS + i], st[3 * COV_ROWS + i], st[4 * COV_ROWS + i]};
        T c[VW];
        for (int v = 0; v < VW; v++) {
            if constexpr (sizeof(T) == 4)
                c[v] = pcov_ptf(cc, pr, q[v]);
            else
                c[v] = pcov_pt(cc, pr, q[v]);
        }
        T* o = out + (i0 + i) * n2 + j0;
        if (full) {
            if constexpr (sizeof(T) == 4) {
                const mra_f4 cv = {c[0], c[1], c[2], c[3]};
                __builtin_nontemporal_store(cv, reinterpret_cast<mra_f4*>(o));
            } else {
                const mra_d2 cv = {c[0], c[1]};
                __builtin_nontemporal_store(cv, reinterpret_cast<mra_d2*>(o));
            }
        } else {
            for (int v = 0; v < VW; v++)
                if (j0 + v < n2) o[v] = c[v];
        }
    }
}

// Variance-covariance of a point set (nugget on the diagonal).
__global__ void k_var_cov(CovCtx cc, Pts P, uint64_t n, double* out) {
    uint64_t idx = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (idx >= n * n) return;
    uint64_t i = idx / n, j = idx % n;
    out[idx] = (i == j) ? P.v[i] : pcov(cc, P, i, P, j);
}

__global__ void k_cross_cov(CovCtx cc, Pts P, uint64_t n1, Pts Q, uint64_t n2, double* out) {
    uint64_t idx = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (idx >= n1 * n2) return;
    uint64_t i = idx / n2, j = idx % n2;
    out[idx] = pcov(cc, P, i, Q, j);
}

// --------------------------------------------------------- coarse prior
// One work group per region R of level l.  For j < l:
//   B = C(Q_{a_j}, Q_R) - Gcat_{a_j}^T Gcat_R[0:j]      (MFMA, K = j*RP)
//   G_{R,j} = L_{a_j}^{-1} B                            (MFMA with L^{-1})
// then K_R = V(Q_R) - Gcat_R^T Gcat_R, L_R = chol(K_R), L_R^{-1}.
// COPY32 (mixed precision): the knot covariances in FP32 arithmetic (widened;
// the products, K_R and its factorisation stay FP64) and an FP32 copy of G.
// (src/class_approximation.cpp:108-159)
template <int RP, bool COPY32>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_prior_coarse(Tree T, CovCtx cc, Pts K, int l, uint64_t start_l,
                                                      const uint64_t* goff, double* Gc, double* Lt, double* LiT,
                                                      int* bad_flag, float* Gc32) {
    constexpr int LD = RP + 1;
    __shared__ double Cm[RP * LD];
    __shared__ PtV kst[2][RP];  // knots of the current and the next row set
    __shared__ double dbuf[RP >= 48 ? 1 : RP * 16];
    __shared__ int bad;
    const uint32_t bid = xcd_block(blockIdx.x, gridDim.x, T.xcd);
    const uint64_t R = start_l + bid;
    const int r = T.r;
    double* GR = Gc + goff[l] + (uint64_t)bid * l * RP * RP;
    if (threadIdx.x == 0) bad = 0;
    // this thread's column: a knot of R (the same in every step)
    const int cq = threadIdx.x % RP;
    const PtV qc = cq < r ? load_pt(cc, K, R * RP + cq) : PtV{0, 0, 0, 0, 1};
    // row sets: the ancestors at levels 0..l-1, then R itself (for K_R); the
    // next set's raw fields are loaded while the current step computes
    auto rows_of = [&](int j) { return j < l ? ancestor_up(R, l, j, T.J) : R; };
    if (threadIdx.x < r) kst[0][threadIdx.x] = load_pt(cc, K, rows_of(0) * RP + threadIdx.x);
    __syncthreads();
    for (int j = 0; j < l; j++) {
        const uint64_t a = ancestor_up(R, l, j, T.J);
        double nx = 0.0, ny = 0.0, nz = 0.0, ns_ = 0.0, nb = 1.0;
        if (threadIdx.x < r) {
            const uint64_t q = rows_of(j + 1) * RP + threadIdx.x;
            nx = K.x[q];
            ny = K.y[q];
            nz = K.z[q];
            ns_ = K.s[q];
            nb = K.b[q];
        }
        cov_tile_pre<RP, RP, double, COPY32>(Cm, LD, cc, kst[j & 1], r, qc, r);
        __syncthreads();
        if (j > 0) {
            const double* Ga = Gc + goff[j] + (a - lvl_start(j, T.J)) * (uint64_t)j * RP * RP;
            wg_gemm_tn<false, true, double, double, GEMM_FULL>(Cm, LD, RP, RP, j * RP, -1.0, Ga, RP, RP, GR, RP, RP);
            __syncthreads();
        }
        // G_{R,j} = L_a^{-1} B, stored straight to global memory
        // (mixed precision: also an FP32 copy for the finest-level products)
        const uint64_t o = goff[l] + (uint64_t)bid * l * RP * RP + (uint64_t)j * RP * RP;
        if constexpr (RP == 64)
            wg_trmm64_cols(GR + (uint64_t)j * RP * RP, COPY32 ? Gc32 + o : nullptr, LiT + a * RP * RP, Cm, LD);
        else
            wg_gemm_tn<false, false, double, double, GEMM_FULL, COPY32, GEMM_TRI_A>(
                GR + (uint64_t)j * RP * RP, RP, RP, RP, RP, 1.0, LiT + a * RP * RP, RP, RP, Cm, LD, RP,
                COPY32 ? Gc32 + o : nullptr);
        if (threadIdx.x < r)
            kst[(j + 1) & 1][threadIdx.x] = PtV{nx, ny, nz, cc.stationary ? 0.0 : ns_ * (nb * sqrt(nb)), nb};
        __syncthreads();
    }
    // K_R = V(Q_R) - Gcat_R^T Gcat_R: the stage now holds R's own knots
    cov_tile_pre<RP, RP, double, COPY32>(Cm, LD, cc, kst[l & 1], r, qc, r, K.v + R * RP);
    __syncthreads();
    if (l > 0) {
        // only the lower part is read by the Cholesky factorisation
        if constexpr (RP == 64)
            wg_syrk64_lower(Cm, LD, GR, l * RP);
        else
            wg_gemm_tn<false, true, double, double, GEMM_FULL, false, GEMM_TRI_C>(Cm, LD, RP, RP, l * RP, -1.0, GR, RP,
                                                                                  RP, GR, RP, RP);
        __syncthreads();
    }
    wg_chol_blk(Cm, LD, RP / 16, dbuf, &bad);
    for (int idx = threadIdx.x; idx < RP * RP; idx += MRA_WG) {
        const int k = idx / RP, i = idx % RP;  // column-major: [k*RP + i] = L[i][k]
        Lt[R * RP * RP + idx] = (i >= k) ? Cm[i * LD + k] : 0.0;
    }
    wg_trinv_blk(Cm, LD, RP / 16, dbuf);
    for (int idx = threadIdx.x; idx < RP * RP; idx += MRA_WG) {
        const int k = idx / RP, i = idx % RP;
        LiT[R * RP * RP + idx] = (i >= k) ? Cm[i * LD + k] : 0.0;
    }
    if (threadIdx.x == 0 && bad) atomicOr(bad_flag, 1);
}

// The same recursion for RP = 64, one wave per region and four regions per
// work group (the four siblings of one parent when J = 4, which read the same
// ancestor blocks), with no work-group barrier.  Per ancestor level j < l:
//   acc = C(Q_{a_j}, Q_R)                  covariance tile in registers
//   acc -= Gcat_{a_j}^T Gcat_R[0:j]         wave_mm64_acc, K = 64 j
//   G_{R,j} = L_{a_j}^{-1} acc              wave_trmm64_lt (acc is the B operand)
// then K_R = V(Q_R) - Gcat_R^T Gcat_R (tiles on and below the diagonal) goes
// to the L_R slot, where k_prior_chol factors it.
// (src/class_approximation.cpp:108-159)
template <bool COPY32>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_prior_wave(
    Tree T, CovCtx cc, Pts K, int l, uint64_t start_l, uint32_t nreg, const uint64_t* goff, double* Gc,
    const double* LiT, double* Kr, float* Gc32) {
    constexpr int RP = 64;
    __shared__ PtV kw[4][2][RP];  // per wave: R's knots, the current ancestor's knots
    __shared__ double cbuf[4][16 * 64];  // per wave: covariance entries on their way to the registers
    const int w = wave_id(), lane = threadIdx.x & 63;
    const uint32_t bid = xcd_block(blockIdx.x, gridDim.x, T.xcd) * 4u + (uint32_t)w;
    if (bid >= nreg) return;  // whole waves; no work-group barrier below
    const uint64_t R = start_l + bid;
    const int r = T.r;
    const uint64_t og = goff[l] + (uint64_t)bid * l * RP * RP;
    double* GR = Gc + og;
    PtV* own = kw[w][0];
    PtV* rows = kw[w][1];
    own[lane] = lane < r ? load_pt(cc, K, R * RP + lane) : PtV{0, 0, 0, 0, 1};
    mra_d4 acc[16];
    for (int j = 0; j < l; j++) {
        const uint64_t a = ancestor_up(R, l, j, T.J);
        rows[lane] = lane < r ? load_pt(cc, K, a * RP + lane) : PtV{0, 0, 0, 0, 1};
        wave_lds_fence();
        wave_cov64(acc, cc, rows, own, r, nullptr, cbuf[w]);
        if (j > 0)
            wave_mm64_acc(acc, j * RP, -1.0, Gc + goff[j] + (a - lvl_start(j, T.J)) * (uint64_t)j * RP * RP, RP, GR,
                          RP);
        wave_trmm64_lt(acc, LiT + a * RP * RP);
        wave_store64(GR + (uint64_t)j * RP * RP, RP, acc);
        if (COPY32) wave_store64(Gc32 + og + (uint64_t)j * RP * RP, RP, acc);
        // G_{R,j} is read back by this wave's later products
        __threadfence_block();
    }
    wave_lds_fence();
    wave_cov64(acc, cc, own, own, r, K.v + R * RP, cbuf[w]);
    if (l > 0) wave_mm64_acc<true>(acc, l * RP, -1.0, GR, RP, GR, RP);
    wave_store64(Kr + R * RP * RP, RP, acc);
}

// L_R = chol(K_R) and L_R^{-1} from the K_R that k_prior_wave left in the L_R
// slot (row-major, lower part read), stored as k_prior_coarse stores them.
// One work group per region.
template <int RP>
__global__ __launch_bounds__(256) void k_prior_chol(uint64_t start_l, double* Lt, double* LiT, int* bad_flag) {
    constexpr int LD = RP + 1;
    __shared__ double Cm[RP * LD];
    __shared__ double dbuf[RP >= 48 ? 1 : RP * 16];
    __shared__ int bad;
    const uint64_t R = start_l + blockIdx.x;
    if (threadIdx.x == 0) bad = 0;
    for (int idx = threadIdx.x; idx < RP * RP; idx += MRA_WG) Cm[(idx / RP) * LD + idx % RP] = Lt[R * RP * RP + idx];
    __syncthreads();
    wg_chol_blk(Cm, LD, RP / 16, dbuf, &bad);
    for (int idx = threadIdx.x; idx < RP * RP; idx += MRA_WG) {
        const int k = idx / RP, i = idx % RP;  // column-major: [k*RP + i] = L[i][k]
        Lt[R * RP * RP + idx] = (i >= k) ? Cm[i * LD + k] : 0.0;
    }
    wg_trinv_blk(Cm, LD, RP / 16, dbuf);
    for (int idx = threadIdx.x; idx < RP * RP; idx += MRA_WG) {
        const int k = idx / RP, i = idx % RP;
        LiT[R * RP * RP + idx] = (i >= k) ? Cm[i * LD + k] : 0.0;
    }
    if (threadIdx.x == 0 && bad) atomicOr(bad_flag, 1);
}

// ---------------------------------------------------------- finest level
// Observation columns of one segment (all observations below one parent,
// which share every ancestor), tiles of 64 columns:
//   B_{F,j} = C(Q_{a_j}, S) - Gcat_{a_j}^T Gcat_F[0:j],  G_{F,j} = L_{a_j}^{-1} B_{F,j}
// (…finest-level-in-prior.cpp:116-163)
// T = double: FP64 throughout.  T = float (mixed precision): the covariance
// tile, both products and the stored G, B are FP32 with FP32 accumulation.
template <int RP, int NT, typename T, bool PERSIST = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 3))) void k_finest_cross(Tree T_, CovCtx cc, Pts K, Pts O, const uint32_t* items,
                                                      uint32_t nitems, const uint64_t* goff, const double* Gc,
                                                      const double* LiT, T* Gf, T* Bf, const float* Gc32) {
    const Tree& T0 = T_;
    constexpr int LD = NT + 1;
    __shared__ T Cm[RP * LD];
    __shared__ PtV kst[2][RP];  // ancestor knots, double-buffered
    // one item per work group; PERSIST: a resident grid walks the items
    // k = blockIdx.x + i * gridDim.x (the j loop ends with a barrier, so
    // LDS is free for the next item).  The walk is a separate instance: as
    // a loop around the one-item body it raised the VGPR count of the
    // one-item launches.
    auto item = [&](const uint32_t k) {
        const uint32_t bid = xcd_block(k, nitems, T0.xcd);
        const uint32_t sg = items[2 * bid], c0 = items[2 * bid + 1];
        const uint64_t ob = T0.s_os[sg];
        const int nF = (int)T0.s_n[sg];
        const int n = min(NT, nF - (int)c0);
        const int r = T0.r, L = T0.L;
        T* GF = Gf + (uint64_t)T0.LRP * ob;
        T* BO = Bf + (uint64_t)T0.LRP * ob;
        // this thread's observation column (the same for every ancestor step)
        const int cc_ = threadIdx.x % NT;
        const PtV qc = cc_ < n ? load_pt(cc, O, ob + c0 + cc_) : PtV{0, 0, 0, 0, 1};
        // knots of the level-0 ancestor; each later step's knots are loaded into
        // registers while the current step computes, and staged at its end
        if (threadIdx.x < r) kst[0][threadIdx.x] = load_pt(cc, K, ancestor_up(T0.s_parent[sg], L - 1, 0, T0.J) * RP + threadIdx.x);
        __syncthreads();
        for (int j = 0; j < L; j++) {
            const uint64_t a = ancestor_up(T0.s_parent[sg], L - 1, j, T0.J);
            // raw fields only (q is formed when staged, so nothing waits on these loads here)
            double nx = 0.0, ny = 0.0, nz = 0.0, ns_ = 0.0, nb = 1.0;
            if (j + 1 < L && threadIdx.x < r) {
                const uint64_t q = ancestor_up(T0.s_parent[sg], L - 1, j + 1, T0.J) * RP + threadIdx.x;
                nx = K.x[q];
                ny = K.y[q];
                nz = K.z[q];
                ns_ = K.s[q];
                nb = K.b[q];
            }
            cov_tile_pre<RP, NT, T>(Cm, LD, cc, kst[j & 1], r, qc, n);
            __syncthreads();
            if (j > 0) {
                const uint64_t go = goff[j] + (a - lvl_start(j, T0.J)) * (uint64_t)j * RP * RP;
                const double* Ga = Gc + go;
                if constexpr (sizeof(T) == 4)
                    wg_gemm_tn_f32<true, float, T, float, 0, true>(Cm, LD, RP, NT, j * RP, -1.0f, Gc32 + go, RP, RP,
                                                                   GF + c0, nF, n);
                else
                    wg_gemm_tn<false, true, double, double, GEMM_ANY>(Cm, LD, RP, NT, j * RP, -1.0, Ga, RP, RP, GF + c0,
                                                                      nF, n);
                __syncthreads();
            }
            // B is stored observation-major: row o holds B_{F,0..L-1}(:, o)
            for (int idx = threadIdx.x; idx < RP * NT; idx += MRA_WG) {
                const int i = idx % RP, c = idx / RP;
                if (c < n) BO[(uint64_t)(c0 + c) * T0.LRP + j * RP + i] = Cm[i * LD + c];
            }
            // G_{F,j} = L_a^{-1} B, stored straight to global memory
            if constexpr (sizeof(T) == 4)
                wg_gemm_tn_f32<false, double, float, float, GEMM_TRI_A, true>(GF + (uint64_t)(j * RP) * nF + c0, nF, RP, NT,
                                                                        RP, 1.0f, LiT + a * RP * RP, RP, RP,
                                                                        (const float*)Cm, LD, n);
            else
                wg_gemm_tn<false, false, double, double, GEMM_ANY, false, GEMM_TRI_A>(GF + (uint64_t)(j * RP) * nF + c0, nF, RP, NT, RP, 1.0,
                                                                   LiT + a * RP * RP, RP, RP, (const double*)Cm, LD, n);
            if (j + 1 < L && threadIdx.x < r)
                kst[(j + 1) & 1][threadIdx.x] = PtV{nx, ny, nz, cc.stationary ? 0.0 : ns_ * (nb * sqrt(nb)), nb};
            __syncthreads();
        }
    };
    if constexpr (PERSIST) {
        for (uint32_t k = blockIdx.x; k < nitems; k += gridDim.x) item(k);
    } else if (blockIdx.x < nitems) {
        item(blockIdx.x);
    }
}

// The mixed-precision finest conditioning (RP = 32 or 64 knots per region)
// one wave per tile of NT observation columns, four tiles per work group, no
// work-group barrier:
// per ancestor level j the FP32 tile lives in the registers of the FP32
// matrix instruction's output layout (row 16 ti + 4 (t/16) + v, column
// 16 tj + t%16):
//   acc = C(Q_{a_j}, S)                      covariance (pcov_ptf, as k_finest_cross)
//   acc -= Gcat32_{a_j}^T Gcat_F[0:j]         wave_cond_f32, K = RP j
//   B_{F,j} = acc  -> observation-major rows (through the wave's LDS buffer)
//   G_{F,j} = L_{a_j}^{-1} acc               wave_trmm_l32
// (…finest-level-in-prior.cpp:116-163)
// PF: the nonstationary chordal or planar kernel, evaluated by pcov_ff on
// point data rounded to float once (the values pcov_ptf gives); otherwise
// pcov_ptf on the FP64 point data (great circle, stationary).
template <int NT, int CU, bool PF, int RP = 64>
__global__ __launch_bounds__(256) void k_finest_cross_w(Tree T0, CovCtx cc, Pts K, Pts O, const uint32_t* items,
                                                        uint32_t nitems, const uint64_t* goff, const double* LiT,
                                                        float* Gf, float* Bf, const float* Gc32) {
    constexpr int NTT = NT / 16, LDT = RP + 1, KT = RP / 16, NQ = 4 * KT;
    static_assert(RP == 32 || RP == 64, "32 or 64 knots per region");
    static_assert(NQ * 64 <= NT * LDT, "the covariance strip fits the transpose buffer");
    typedef typename std::conditional<PF, PtF, PtV>::type Pt;
    __shared__ Pt knw[4][RP];            // per wave: the current ancestor's knots
    __shared__ float bufw[4][NT * LDT];  // per wave: covariance strip, then the B transpose
    auto to_pt = [](const PtV& p) {
        if constexpr (PF)
            return to_ptf(p);
        else
            return p;
    };
    auto cov = [&](const Pt& a, const Pt& b) {
        if constexpr (PF)
            return pcov_ff<false>(a, b, cc.metric == 0);
        else
            return pcov_ptf(cc, a, b);
    };
    const int w = wave_id(), lane = threadIdx.x & 63, r16 = lane & 15, k4 = lane >> 4;
    const uint32_t it = xcd_block(blockIdx.x, gridDim.x, T0.xcd) * 4u + (uint32_t)w;
    if (it >= nitems) return;  // whole waves; no work-group barrier below
    const uint32_t sg = items[2 * it], c0 = items[2 * it + 1];
    const uint64_t ob = T0.s_os[sg];
    const int nF = (int)T0.s_n[sg];
    const int n = min(NT, nF - (int)c0);
    const int r = T0.r, L = T0.L, LRP = T0.LRP;
    float* GF = Gf + (uint64_t)LRP * ob;  // G_F(k, c) at GF[k * nF + c]
    float* BO = Bf + (uint64_t)LRP * ob;  // B of observation c: row BO + c * LRP
    Pt* kw = knw[w];
    float* bw = bufw[w];
    Pt oc[NTT];  // this lane's observations, columns 16 tj + t%16
    bool cv[NTT];
#pragma unroll
    for (int tj = 0; tj < NTT; tj++) {
        const int c = 16 * tj + r16;
        cv[tj] = c < n;
        oc[tj] = to_pt(cv[tj] ? load_pt(cc, O, ob + c0 + c) : PtV{0, 0, 0, 0, 1});
    }
    mra_f4 acc[KT * NTT];
    for (int j = 0; j < L; j++) {
        const uint64_t a = ancestor_up(T0.s_parent[sg], L - 1, j, T0.J);
        if (lane < RP) kw[lane] = to_pt(lane < r ? load_pt(cc, K, a * RP + lane) : PtV{0, 0, 0, 0, 1});
        wave_lds_fence();
        // covariance tile: entry q = 4 ti + v of column tile tj is knot
        // 16 ti + 4 k4 + v; evaluated in a loop through the lane's strip
#pragma unroll
        for (int tj = 0; tj < NTT; tj++) {
#pragma unroll 1
            for (int q = 0; q < NQ; q++) {
                const int i = 16 * (q >> 2) + 4 * k4 + (q & 3);
                bw[64 * q + lane] = (i < r && cv[tj]) ? cov(kw[i], oc[tj]) : 0.0f;
            }
#pragma unroll
            for (int q = 0; q < NQ; q++) acc[(q >> 2) * NTT + tj][q & 3] = bw[64 * q + lane];
        }
        if (j > 0)
            wave_cond_f32<NTT, CU, KT>(acc, j * RP,
                                       Gc32 + goff[j] + (a - lvl_start(j, T0.J)) * (uint64_t)j * RP * RP, GF + c0,
                                       nF, cv);
        // B_{F,j}: observation-major, one row of RP knots per observation
        wave_lds_fence();
#pragma unroll
        for (int ti = 0; ti < KT; ti++)
#pragma unroll
            for (int tj = 0; tj < NTT; tj++)
#pragma unroll
                for (int v = 0; v < 4; v++) bw[(16 * tj + r16) * LDT + 16 * ti + 4 * k4 + v] = acc[ti * NTT + tj][v];
        wave_lds_fence();
        if (lane < RP)
            for (int c = 0; c < n; c++) BO[(uint64_t)(c0 + c) * LRP + j * RP + lane] = bw[c * LDT + lane];
        wave_lds_fence();
        // G_{F,j} = L_a^{-1} B, stored for the local step and read back by
        // this wave's later products
        wave_trmm_l32<NTT, KT>(acc, LiT + a * RP * RP);
#pragma unroll
        for (int ti = 0; ti < KT; ti++)
#pragma unroll
            for (int tj = 0; tj < NTT; tj++)
                if (cv[tj])
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        GF[(uint64_t)(j * RP + 16 * ti + 4 * k4 + v) * nF + c0 + 16 * tj + r16] = acc[ti * NTT + tj][v];
        __threadfence_block();
    }
}

// Per finest region: Sigma_F = V(S) - G^T G, L_F = chol, u = L_F^{-1} y,
// H = L_F^{-1} B^T, d_F = |u|^2 + 2 sum log diag L_F.
// (…finest-level-in-prior.cpp:165-256)
// Length-bucketed: NB > 0 handles regions with n_F <= NB with Sigma_F in LDS
// (NB = 16/32: the GEMM is split over the 4 waves and the triangular solve
// for H keeps one row of B per thread in registers); NB = 0 handles larger
// regions with Sigma_F in the global workspace.
template <int NB>
struct FinestLds {
    static constexpr int LD = NB + 1;
    double S[NB * LD];
    double kred[NB <= 16 ? 3 * 256 : (NB <= 32 ? 3 * 4 * 256 : 1)];
    double D[NB <= 32 ? NB * 16 : 1];  // inverse diagonal blocks (fewer than 3 blocks)
};
template <>
struct FinestLds<0> {
    static constexpr int LD = 1;
    double S[1];
    double kred[1];
    double D[1];
};

template <int RP, int NB, typename TG, bool PERSIST = false>
__global__ __launch_bounds__(256) void k_finest_local(Tree T, CovCtx cc, Pts O, const double* yv,
                                                      const uint32_t* items, uint32_t nitems, TG* Gf, TG* Bf,
                                                      double* Sig, double* H, double* u, double* dterm,
                                                      int* bad_flag) {
    __shared__ FinestLds<NB> sh;
    __shared__ double red[MRA_WG];
    __shared__ int bad;
    // one region per work group; PERSIST: a resident grid walks the regions
    // k = blockIdx.x + i * gridDim.x, with a barrier before the next one
    auto item = [&](const uint32_t k) {
        const uint32_t F = items[k];
        const uint64_t ob = T.obs_off[F];
        const int n = (int)(T.obs_off[F + 1] - ob);
        const int LRP = T.LRP;
        double* LF = Sig + T.sig_off[F];  // L_F (row-major) kept for prediction and sampling
        double* Sg = NB > 0 ? sh.S : LF;
        const int ld = NB > 0 ? FinestLds<NB>::LD : n;
        const uint64_t so = T.f_seg_os[F];
        const int sn = (int)T.f_seg_n[F];
        const TG* GF = Gf + (uint64_t)LRP * so + (ob - so);
        if (threadIdx.x == 0) bad = 0;
        for (int idx = threadIdx.x; idx < n * n; idx += MRA_WG) {
            const int i = idx / n, c = idx % n;
            double v;
            if constexpr (sizeof(TG) == 4)
                v = (i == c) ? O.v[ob + i] : (double)pcovf(cc, O, ob + i, O, ob + c);
            else
                v = (i == c) ? O.v[ob + i] : pcov(cc, O, ob + i, O, ob + c);
            Sg[i * ld + c] = v;
        }
        __syncthreads();
        if (NB > 0 && NB <= 32) {
            wg_gemm_tn_ksplit(Sg, ld, (n + 15) & ~15, (n + 15) & ~15, LRP, -1.0, GF, sn, n, GF, sn, n, sh.kred);
        } else {
            for (int ti = 0; ti < n; ti += 64)
                for (int tj = 0; tj < n; tj += 64) {
                    const int mv = min(64, n - ti), nv = min(64, n - tj);
                    wg_gemm_tn(Sg + ti * ld + tj, ld, (mv + 15) & ~15, (nv + 15) & ~15, LRP, -1.0, GF + ti, sn, mv,
                               GF + tj, sn, nv);
                }
        }
        __syncthreads();
        const int nbk = (n + 15) >> 4;
        if (NB > 0) {
            // blocked factorisation of the 16-padded block: identity padding
            const int N16 = nbk * 16;
            for (int idx = threadIdx.x; idx < N16 * N16; idx += MRA_WG) {
                const int i = idx / N16, c = idx % N16;
                if (i >= n || c >= n) Sg[i * ld + c] = (i == c) ? 1.0 : 0.0;
            }
            wg_chol_blk(Sg, ld, nbk, sh.D, &bad);
        } else {
            wg_chol(Sg, ld, n, &bad);
        }
        // log det part, then keep L_F (lower triangle; zeros above)
        double part = 0.0;
        for (int i = threadIdx.x; i < n; i += MRA_WG) part += 2.0 * log(Sg[i * ld + i]);
        if (NB > 0)
            for (int idx = threadIdx.x; idx < n * n; idx += MRA_WG) {
                const int i = idx / n, c = idx % n;
                LF[idx] = c <= i ? Sg[i * ld + c] : 0.0;
            }
        __syncthreads();
        if (NB > 0) {
            // L_F^{-1} in place (zero strict upper part, read by the H product)
            wg_trinv_blk(Sg, ld, nbk, sh.D);
            // u = L_F^{-1} y
            double* uF = u + ob;
            for (int i = threadIdx.x; i < n; i += MRA_WG) {
                double acc = 0.0;
                for (int k = 0; k <= i; k++) acc += Sg[i * ld + k] * yv[ob + k];
                uF[i] = acc;
                part += acc * acc;
            }
            const double dF = wg_sum(part, red);
            // H = L_F^{-1} B (observation-major rows), 64 columns at a time
            const TG* BO = Bf + ob * (uint64_t)LRP;
            for (int c0 = 0; c0 < LRP; c0 += 64) {
                const int cv = min(64, LRP - c0);
                wg_gemm_tn<true, false>(H + ob * (uint64_t)LRP + c0, LRP, (n + 15) & ~15, 64, n, 1.0, Sg, ld, n,
                                        BO + c0, LRP, cv);
            }
            if (threadIdx.x == 0) dterm[T.nC + F] = dF;
        } else {
            // large regions: forward substitution in global memory
            double* uF = u + ob;
            for (int i = threadIdx.x; i < n; i += MRA_WG) uF[i] = yv[ob + i];
            __syncthreads();
            wg_trsm(Sg, ld, uF, 1, n, 1);
            for (int i = threadIdx.x; i < n; i += MRA_WG) part += uF[i] * uF[i];
            const double dF = wg_sum(part, red);
            const TG* BO = Bf + ob * (uint64_t)LRP;
            double* HF = H + ob * (uint64_t)LRP;
            for (int m = threadIdx.x; m < LRP; m += MRA_WG) {
                for (int i = 0; i < n; i++) {
                    double acc = (double)BO[(uint64_t)i * LRP + m];
                    for (int k = 0; k < i; k++) acc -= Sg[i * ld + k] * HF[(uint64_t)k * LRP + m];
                    acc /= Sg[i * ld + i];
                    HF[(uint64_t)i * LRP + m] = acc;
                }
            }
            if (threadIdx.x == 0) dterm[T.nC + F] = dF;
        }
        if (threadIdx.x == 0 && bad) atomicOr(bad_flag, 1);
        __syncthreads();
    };
    if constexpr (PERSIST) {
        for (uint32_t k = blockIdx.x; k < nitems; k += gridDim.x) item(k);
    } else if (blockIdx.x < nitems) {
        item(blockIdx.x);
    }
}

// The same local step for finest regions with 64 < n_F <= 256 observations,
// on a padded N16 x N16 workspace in global memory (ws + ws_off[item], row
// stride N16 = 16 ceil(n_F / 16), identity padding): Sigma_F = V(S) - G^T G
// on the lower 64 x 64 tiles (MFMA), blocked Cholesky and inverse on 16 x 16
// blocks (the inverse diagonal blocks kept in the unused upper triangle),
// u = L_F^{-1} y, d_F and H = L_F^{-1} B (MFMA).  L_F is copied to its
// n_F x n_F place for prediction and sampling.
template <typename TG, bool PERSIST = false>
__global__ __launch_bounds__(256) void k_finest_local_big(Tree T, CovCtx cc, Pts O, const double* yv,
                                                          const uint32_t* items, uint32_t nitems,
                                                          const uint64_t* ws_off, double* ws, const TG* Gf,
                                                          const TG* Bf, double* Sig, double* H, double* u,
                                                          double* dterm, int* bad_flag) {
    __shared__ double red[MRA_WG];
    __shared__ double dbuf[1];  // unused: with 3 or more blocks the inverse diagonal blocks live in A
    __shared__ int bad;
    // one region per work group; PERSIST: a resident grid walks the regions
    // k = blockIdx.x + i * gridDim.x, with a barrier before the next one
    auto item = [&](const uint32_t k) {
        const uint32_t F = items[k];
        const uint64_t ob = T.obs_off[F];
        const int n = (int)(T.obs_off[F + 1] - ob);
        const int LRP = T.LRP;
        const int nbk = (n + 15) >> 4, ld = nbk * 16;
        double* Sg = ws + ws_off[k];
        double* LF = Sig + T.sig_off[F];
        const uint64_t so = T.f_seg_os[F];
        const int sn = (int)T.f_seg_n[F];
        const TG* GF = Gf + (uint64_t)LRP * so + (ob - so);
        if (threadIdx.x == 0) bad = 0;
        // V(S) on the lower triangle, identity padding (the upper triangle is not read)
        for (int idx = threadIdx.x; idx < ld * ld; idx += MRA_WG) {
            const int i = idx / ld, c = idx % ld;
            if (c > i) continue;
            double v;
            if (i < n && c < n) {
                if (i == c)
                    v = O.v[ob + i];
                else if constexpr (sizeof(TG) == 4)
                    v = (double)pcovf(cc, O, ob + i, O, ob + c);
                else
                    v = pcov(cc, O, ob + i, O, ob + c);
            } else {
                v = (i == c) ? 1.0 : 0.0;
            }
            Sg[(uint64_t)i * ld + c] = v;
        }
        __syncthreads();
        // - G^T G on the lower 64 x 64 tiles (only the valid n x n part changes)
        for (int ti = 0; ti < n; ti += 64)
            for (int tj = 0; tj <= ti; tj += 64) {
                const int mv = min(64, n - ti), nv = min(64, n - tj);
                wg_gemm_tn(Sg + (uint64_t)ti * ld + tj, ld, (mv + 15) & ~15, (nv + 15) & ~15, LRP, -1.0, GF + ti, sn, mv,
                           GF + tj, sn, nv);
            }
        __syncthreads();
        wg_chol_blk(Sg, ld, nbk, dbuf, &bad);
        double part = 0.0;
        for (int i = threadIdx.x; i < n; i += MRA_WG) part += 2.0 * log(Sg[(uint64_t)i * ld + i]);
        for (int idx = threadIdx.x; idx < n * n; idx += MRA_WG) {
            const int i = idx / n, c = idx % n;
            LF[idx] = c <= i ? Sg[(uint64_t)i * ld + c] : 0.0;
        }
        __syncthreads();
        wg_trinv_blk<16>(Sg, ld, nbk, dbuf);
        // u = L_F^{-1} y
        for (int i = threadIdx.x; i < n; i += MRA_WG) {
            double acc = 0.0;
            for (int q = 0; q <= i; q++) acc += Sg[(uint64_t)i * ld + q] * yv[ob + q];
            u[ob + i] = acc;
            part += acc * acc;
        }
        const double dF = wg_sum(part, red);
        // H = L_F^{-1} B (observation-major rows), 64 columns at a time
        const TG* BO = Bf + ob * (uint64_t)LRP;
        for (int c0 = 0; c0 < LRP; c0 += 64) {
            const int cv = min(64, LRP - c0);
            wg_gemm_tn<true, false>(H + ob * (uint64_t)LRP + c0, LRP, ld, 64, n, 1.0, Sg, ld, n, BO + c0, LRP, cv);
        }
        if (threadIdx.x == 0) {
            dterm[T.nC + F] = dF;
            if (bad) atomicOr(bad_flag, 1);
        }
        __syncthreads();
    };
    if constexpr (PERSIST) {
        for (uint32_t k = blockIdx.x; k < nitems; k += gridDim.x) item(k);
    } else if (blockIdx.x < nitems) {
        item(blockIdx.x);
    }
}

// The same local step for finest regions with at most 16 observations, one
// wave per region and four regions per work group, with no work-group
// barrier: Sigma_F = V(S) - G_F^T G_F is one 16 x 16 MFMA tile (K = L RP in
// four interleaved chains, summed in a fixed order; A and B fragments are the
// same G values, so one load feeds both), factored with its inverse by
// wave_chol16 in the wave's LDS tile; u = L_F^{-1} y, d_F and
// H = L_F^{-1} B (16 x 16 output tiles across the L RP columns).
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

template <typename TG, bool PERSIST = false>
__global__ __launch_bounds__(256) void k_finest_local_w16(Tree T, CovCtx cc, Pts O, const double* yv,
                                                          const uint32_t* items, uint32_t nitems, const TG* Gf,
                                                          const TG* Bf, double* Sig, double* H, double* u,
                                                          double* dterm, int* bad_flag) {
    constexpr int LDP = 17;
    __shared__ double Ps[4][16 * LDP];
    __shared__ double Ds[4][16 * 16];
    __shared__ int badw[4];
    const int w = wave_id();
    const int lane = threadIdx.x & 63, r16 = lane & 15, k4 = lane >> 4;
    // one region per wave: item blockIdx.x * 4 + w; PERSIST: every
    // 4 * gridDim.x-th item from there (waves share no LDS)
    auto item = [&](const uint32_t it) {
        const uint32_t F = items[it];
        const uint64_t ob = T.obs_off[F];
        const int n = (int)(T.obs_off[F + 1] - ob);
        const int LRP = T.LRP;
        const uint64_t so = T.f_seg_os[F];
        const int sn = (int)T.f_seg_n[F];
        const TG* GF = Gf + (uint64_t)LRP * so + (ob - so);  // G_F(k, i) at GF[k * sn + i]
        double* P = Ps[w];
        double* Dinv = Ds[w];
        if (lane == 0) badw[w] = 0;
        // -G^T G in four chains
        const bool iv = r16 < n;
        const int ic = iv ? r16 : 0;
        mra_d4 c4[4];
#pragma unroll
        for (int q = 0; q < 4; q++) c4[q] = mra_d4{0, 0, 0, 0};
#pragma unroll 4
        for (int k0 = 0; k0 < LRP; k0 += 16) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const double g = iv ? (double)GF[(uint64_t)(k0 + 4 * q + k4) * sn + ic] : 0.0;
                c4[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(g, g, c4[q], 0, 0, 0);
            }
        }
        // Sigma_F = V(S) - G^T G, identity padding (element v: row k4 + 4v, column r16)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = k4 + 4 * v, c = r16;
            const double gg = ((c4[0][v] + c4[1][v]) + c4[2][v]) + c4[3][v];
            double val;
            if (i < n && c < n) {
                double cv;
                if (i == c)
                    cv = O.v[ob + i];
                else if constexpr (sizeof(TG) == 4)
                    cv = (double)pcovf(cc, O, ob + i, O, ob + c);
                else
                    cv = pcov(cc, O, ob + i, O, ob + c);
                val = cv - gg;
            } else {
                val = (i == c) ? 1.0 : 0.0;
            }
            P[i * LDP + c] = val;
        }
        wave_chol16(P, LDP, Dinv, 16, &badw[w]);
        // log det part, L_F kept (row-major n x n, zeros above)
        double* LF = Sig + T.sig_off[F];
        double part = lane < n ? 2.0 * log(P[lane * LDP + lane]) : 0.0;
        for (int idx = lane; idx < n * n; idx += 64) {
            const int i = idx / n, c = idx % n;
            LF[idx] = c <= i ? P[i * LDP + c] : 0.0;
        }
        // u = L_F^{-1} y
        if (lane < n) {
            double acc = 0.0;
            for (int k = 0; k <= lane; k++) acc += Dinv[lane * 16 + k] * yv[ob + k];
            u[ob + lane] = acc;
            part += acc * acc;
        }
        const double dF = wave_sum(part);
        if (lane == 0) dterm[T.nC + F] = dF;
        // H = L_F^{-1} B, observation-major rows of L RP
        const TG* BO = Bf + ob * (uint64_t)LRP;
        double* HF = H + ob * (uint64_t)LRP;
        double al[4];
#pragma unroll
        for (int q = 0; q < 4; q++) al[q] = Dinv[r16 * 16 + 4 * q + k4];  // A(i = r16, k = 4q + k4)
#pragma unroll 4
        for (int m0 = 0; m0 < LRP; m0 += 16) {
            mra_d4 hacc = {0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int k = 4 * q + k4;
                const double b = k < n ? (double)BO[(uint64_t)k * LRP + m0 + r16] : 0.0;
                hacc = __builtin_amdgcn_mfma_f64_16x16x4f64(al[q], b, hacc, 0, 0, 0);
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = k4 + 4 * v;
                if (i < n) HF[(uint64_t)i * LRP + m0 + r16] = hacc[v];
            }
        }
        if (lane == 0 && badw[w]) atomicOr(bad_flag, 1);
        wave_lds_fence();
    };
    if constexpr (PERSIST) {
        for (uint32_t it = blockIdx.x * 4u + (uint32_t)w; it < nitems; it += 4u * gridDim.x) item(it);
    } else if (blockIdx.x * 4u + (uint32_t)w < nitems) {
        item(blockIdx.x * 4u + (uint32_t)w);
    }
}

// The local step for finest regions of more than 64 observations (any size)
// as a left-looking Cholesky on 32-wide block columns, with no Sigma_F
// workspace.  Block column c (rows and columns c0 = 32c ..):
//  (1) panel tiles P_rc = V_rc - G_r^T G_c - L_{r,<c} L_{c,<c}^T (r >= c),
//      32 x 32 MFMA products whose operands are G_F (segment layout) and the
//      rows of L_F already final in Sig; P_cc goes to LDS, the others to
//      their places in Sig; t = y_c - L_{c,<c} u_{<c};
//  (2) one wave factors P_cc and forms D_c = L_cc^{-1} and u_c = D_c t while
//      the other three form R_c = B_c - L_{c,<c} H_{<c} for their 32-column
//      strips of H (stored in H's rows of the block);
//  (3) L_rc = P_rc D_c^T in place in Sig, H_c = D_c R_c in place in H, and
//      L_cc with zeros above the diagonal into Sig.
// Three work-group barriers per block column; d_F = |u|^2 + 2 sum log diag
// L_F summed by one wave in a fixed order.
// (…finest-level-in-prior.cpp:165-256)
template <typename TG, bool PERSIST = false>
__global__ __launch_bounds__(256, 2) void k_finest_local_ll(Tree T, CovCtx cc, Pts O, const double* yv,
                                                         const uint32_t* items, uint32_t nitems, const TG* Gf,
                                                         const TG* Bf, double* Sig, double* H, double* u,
                                                         double* dterm, int* bad_flag) {
    constexpr int LDP = 33;
    __shared__ double Pd[32 * LDP];  // P_cc, factored in place to L_cc
    __shared__ double Dd[32 * LDP];  // D_c
    __shared__ double tv[32];        // y_c - L_{c,<c} u_{<c}
    __shared__ double pw[4][5][64];  // per wave: x, y, z, s, b of the tile's rows (0-31) and columns (32-63)
    __shared__ int bad;
    const int wave = wave_id(), lane = threadIdx.x & 63, r16 = lane & 15, k4 = lane >> 4;
    auto item = [&](const uint32_t k) {
        const uint32_t F = items[k];
        const uint64_t ob = T.obs_off[F];
        const int n = (int)(T.obs_off[F + 1] - ob);
        const int LRP = T.LRP;
        const int nb = (n + 31) >> 5, nstrip = (LRP + 31) >> 5;
        double* LF = Sig + T.sig_off[F];
        const uint64_t so = T.f_seg_os[F];
        const int sn = (int)T.f_seg_n[F];
        const TG* GF = Gf + (uint64_t)LRP * so + (ob - so);  // G(m, i) = GF[m*sn + i]
        const TG* BO = Bf + ob * (uint64_t)LRP;              // B(i, m) = BO[i*LRP + m]
        double* HF = H + ob * (uint64_t)LRP;
        double* uF = u + ob;
        if (threadIdx.x == 0) bad = 0;
        double part = 0.0;  // wave 0: 2 log L_ii + u_i^2 of its lanes
        for (int c = 0; c < nb; c++) {
            const int c0 = 32 * c, cn = min(32, n - c0);
            // (1) panel tiles r = c + wave + 4q
            for (int r = c + wave; r < nb; r += 4) {
                const int r0 = 32 * r, rn = min(32, n - r0);
                // the points of the tile's rows and columns, staged in the
                // wave's LDS (one round of loads for the 16 covariances below)
                wave_lds_fence();
                {
                    const int pi = lane < 32 ? r0 + lane : c0 + lane - 32;
                    if (pi < n) {
                        const uint64_t g = ob + pi;
                        pw[wave][0][lane] = O.x[g];
                        pw[wave][1][lane] = O.y[g];
                        pw[wave][2][lane] = O.z[g];
                        pw[wave][3][lane] = O.s[g];
                        pw[wave][4][lane] = O.b[g];
                    }
                }
                wave_lds_fence();
                // V_rc (lower part of V_cc; identity padding) into the tile's
                // place, in the lane layout of the product below (each lane
                // reads back only what it wrote)
#pragma unroll 1
                for (int e = 0; e < 16; e++) {
                    const int i = 16 * (e >> 3) + k4 + 4 * (e & 3), j = 16 * ((e >> 2) & 1) + r16;
                    const bool in = i < rn && j < cn;
                    double val = (r0 + i == c0 + j) ? 1.0 : 0.0;
                    if (in && (r > c || j <= i)) {
                        const double(*q)[64] = pw[wave];
                        if (r0 + i == c0 + j)
                            val = O.v[ob + r0 + i];
                        else if constexpr (sizeof(TG) == 4)
                            val = (double)pcovf_v(cc, q[0][i], q[1][i], q[2][i], q[3][i], q[4][i], q[0][32 + j],
                                                  q[1][32 + j], q[2][32 + j], q[3][32 + j], q[4][32 + j]);
                        else
                            val = pcov_v(cc, q[0][i], q[1][i], q[2][i], q[3][i], q[4][i], q[0][32 + j], q[1][32 + j],
                                         q[2][32 + j], q[3][32 + j], q[4][32 + j]);
                    }
                    if (r == c)
                        Pd[i * LDP + j] = val;
                    else if (in)
                        LF[(uint64_t)(r0 + i) * n + c0 + j] = val;
                }
                mra_d4 acc[4];
#pragma unroll
                for (int q = 0; q < 4; q++) acc[q] = mra_d4{0, 0, 0, 0};
                {
                    // A(k, i) = G(k, r0 + i), B(k, j) = G(k, c0 + j), K = LRP
                    const BlockOps<false, TG, TG> o(GF + r0, sn, GF + c0, sn, r16, r16 + 16, r16, r16 + 16, r16 < rn,
                                                    r16 + 16 < rn, r16 < cn, r16 + 16 < cn, k4);
                    block_kloop<double, true, true, true, 4>(acc, o, 0, LRP, k4);
                }
                if (c0 > 0) {
                    // A(k, i) = L(r0 + i, k), B(k, j) = L(c0 + j, k), K = c0
                    const BlockOps<true, double, double, true> o(LF + (uint64_t)r0 * n, n, LF + (uint64_t)c0 * n, n, r16,
                                                                 r16 + 16, r16, r16 + 16, r16 < rn, r16 + 16 < rn,
                                                                 r16 < cn, r16 + 16 < cn, k4);
                    block_kloop<double, true, true, true, 4>(acc, o, 0, c0, k4);
                }
                // P = V - acc; element v of tile (ti, tj) is row 16 ti + k4 + 4 v, column 16 tj + r16
#pragma unroll
                for (int ti = 0; ti < 2; ti++)
#pragma unroll
                    for (int tj = 0; tj < 2; tj++)
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            const int i = 16 * ti + k4 + 4 * v, j = 16 * tj + r16;
                            if (r == c) {
                                if (i < rn && j <= i) Pd[i * LDP + j] -= acc[2 * ti + tj][v];
                            } else if (i < rn) {
                                LF[(uint64_t)(r0 + i) * n + c0 + j] -= acc[2 * ti + tj][v];
                            }
                        }
            }
            if (wave == 3) {
                // t = y_c - L_{c,<c} u_{<c}: row lane & 31, halves of the sum in the two half-waves
                const int i = lane & 31;
                double s = 0.0;
                if (i < cn) {
                    const double* Lr = LF + (uint64_t)(c0 + i) * n;
#pragma unroll 8
                    for (int q = lane >> 5; q < c0; q += 2) s += Lr[q] * uF[q];
                }
                s += __shfl_xor(s, 32);
                if (lane < 32) tv[lane] = i < cn ? yv[ob + c0 + i] - s : 0.0;
            }
            __syncthreads();
            if (wave == 0) {
                // (2) L_cc, D_c, u_c = D_c t
                wave_chol32(Pd, Dd, LDP, &bad);
                if (lane < cn) part += 2.0 * log(Pd[lane * LDP + lane]);
                if (lane < 32) {
                    double s = 0.0;
#pragma unroll
                    for (int q = 0; q < 32; q++) s += Dd[lane * LDP + q] * tv[q];
                    if (lane < cn) {
                        uF[c0 + lane] = s;
                        part += s * s;
                    }
                }
            } else {
                // (2) R_c = B_c - L_{c,<c} H_{<c} for strips wave - 1 + 3q, into H's rows c0 ..
                for (int s = wave - 1; s < nstrip; s += 3) {
                    const int m0 = 32 * s, mv = min(32, LRP - m0);
                    mra_d4 acc[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) acc[q] = mra_d4{0, 0, 0, 0};
                    if (c0 > 0) {
                        // A(k, i) = L(c0 + i, k), B(k, m) = H(k, m0 + m), K = c0
                        const BlockOps<true, double, double> o(LF + (uint64_t)c0 * n, n, HF + m0, LRP, r16, r16 + 16,
                                                               r16, r16 + 16, r16 < cn, r16 + 16 < cn, r16 < mv,
                                                               r16 + 16 < mv, k4);
                        block_kloop<double, true, true, true, 4>(acc, o, 0, c0, k4);
                    }
#pragma unroll
                    for (int ti = 0; ti < 2; ti++)
#pragma unroll
                        for (int tj = 0; tj < 2; tj++)
#pragma unroll
                            for (int v = 0; v < 4; v++) {
                                const int i = 16 * ti + k4 + 4 * v, m = 16 * tj + r16;
                                if (i < cn && m < mv) {
                                    const uint64_t e = (uint64_t)(c0 + i) * LRP + m0 + m;
                                    HF[e] = (double)BO[e] - acc[2 * ti + tj][v];
                                }
                            }
                }
            }
            __syncthreads();
            // (3) L_rc = P_rc D_c^T (r > c, tiles r = c + 1 + wave + 4q), in place
            for (int r = c + 1 + wave; r < nb; r += 4) {
                const int r0 = 32 * r, rn = min(32, n - r0);
                mra_d4 acc[4];
#pragma unroll
                for (int q = 0; q < 4; q++) acc[q] = mra_d4{0, 0, 0, 0};
                // A(k, i) = P(r0 + i, c0 + k), B(k, j) = D(j, k); cn = 32 here (c < nb - 1)
                const BlockOps<true, double, double, true> o(LF + (uint64_t)r0 * n + c0, n, Dd, LDP, r16, r16 + 16, r16,
                                                             r16 + 16, r16 < rn, r16 + 16 < rn, true, true, k4);
                block_kloop<double, true, true, true, 4>(acc, o, 0, 32, k4);
#pragma unroll
                for (int ti = 0; ti < 2; ti++)
#pragma unroll
                    for (int tj = 0; tj < 2; tj++)
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            const int i = 16 * ti + k4 + 4 * v, j = 16 * tj + r16;
                            if (i < rn) LF[(uint64_t)(r0 + i) * n + c0 + j] = acc[2 * ti + tj][v];
                        }
            }
            if (wave > 0) {
                // (3) H_c = D_c R_c for strips wave - 1 + 3q, in place
                for (int s = wave - 1; s < nstrip; s += 3) {
                    const int m0 = 32 * s, mv = min(32, LRP - m0);
                    mra_d4 acc[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) acc[q] = mra_d4{0, 0, 0, 0};
                    // A(k, j) = D(j, k), B(k, m) = R(k, m) = H(c0 + k, m0 + m), K = cn
                    const BlockOps<true, double, double> o(Dd, LDP, HF + (uint64_t)c0 * LRP + m0, LRP, r16, r16 + 16,
                                                           r16, r16 + 16, true, true, r16 < mv, r16 + 16 < mv, k4);
                    block_kloop<double, true, true, true, 4>(acc, o, 0, cn, k4);
#pragma unroll
                    for (int ti = 0; ti < 2; ti++)
#pragma unroll
                        for (int tj = 0; tj < 2; tj++)
#pragma unroll
                            for (int v = 0; v < 4; v++) {
                                const int j = 16 * ti + k4 + 4 * v, m = 16 * tj + r16;
                                if (j < cn && m < mv) HF[(uint64_t)(c0 + j) * LRP + m0 + m] = acc[2 * ti + tj][v];
                            }
                }
            }
            // L_cc and the zeros above it (rows 0 .. c0 + cn of block column c)
            for (int idx = threadIdx.x; idx < (c0 + cn) * cn; idx += MRA_WG) {
                const int i = idx / cn, j = idx - i * cn;
                double val = 0.0;
                if (i >= c0 && j <= i - c0) val = Pd[(i - c0) * LDP + j];
                LF[(uint64_t)i * n + c0 + j] = val;
            }
            __syncthreads();
        }
        if (wave == 0) {
            part = wave_sum(part);
            if (lane == 0) {
                dterm[T.nC + F] = part;
                if (bad) atomicOr(bad_flag, 1);
            }
        }
        __syncthreads();
    };
    if constexpr (PERSIST) {
        for (uint32_t k = blockIdx.x; k < nitems; k += gridDim.x) item(k);
    } else if (blockIdx.x < nitems) {
        item(blockIdx.x);
    }
}

// -------------------------------------------------------------- posterior
// One work group per working region R of level l:
//   A_{jk} = sum over children (from the stacked finest rows H when l = L-1,
//            else from the children's A~ messages)
//   L~ = chol(L L^T + A_{ll}), v = L~^{-1} w_l, P_j = L~^{-1} A_{jl}^T,
//   w~_j = w_j - P_j^T v, A~_{jk} = A_{jk} - P_j^T P_k   (j <= k < l)
//   d_R = 2 sum log diag L~ - 2 sum log diag L - v^T v
// (src/class_approximation-posterior-inference.cpp:51-131,150-214;
//  src/class_approximation.cpp:186-306)
struct PostArgs {
    int l;
    uint64_t start_l;
    const uint32_t* items;
    const uint64_t* goff;
    const double* Lt;
    double* Lpt;
    double* LpiT;
    double* P;
    double* v;
    double* dterm;
    // children messages (level l+1) ...
    const double* wc;
    const double* Ac;
    uint64_t start_c;
    // ... or stacked rows: A_{jk} = X_j^T X_k, w_j = X_j^T z over the rows
    // [row_s, row_e) of the region (X = L_F^{-1} B_F^T rows of the finest
    // regions below, or factor rows X' of the children)
    const double* X;
    const double* z;
    int ldx;
    const uint64_t* row_s;
    const uint64_t* row_e;
    // outputs (level l)
    double* wo;
    double* Ao;
    int write_msgs;  // 0: only L~, L~^{-1}, v, P (prediction after a factored level)
    int write_d;
    const int32_t* slot;   // message slot of each region of the level (null: its local index)
    const int32_t* cslot;  // rows mode: dense-message slot of each child (-1: the child sent rows)
    double* Xz;            // rows mode of a dense region in a factored level: zero its rows
    int ldxz;
    double* zz;
    float* Xz32;           // (mixed precision) the FP32 copy of those rows, zeroed too
    // mixed precision: FP32 copy of X (same layout) read by the F32 row products
    const float* X32;
    int* bad_flag;
    int gram_sep;  // the message blocks A~_jk come from k_post_gram (RP = 64)
    int pj_sep;    // P_j and w~_j come from k_post_pj (RP = 64, stacked rows)
    // dense children: Ac holds the sum of the children's message blocks, one
    // block set per region of this level (written by k_post_gram<true>)
    int csum;
    // k_post_gram<true>: items are the parents (level l-1, from start_p)
    uint64_t start_p;
};

// kred: LDS for the K-split products of RP <= 32 (3*4*256 doubles)
template <int RP, bool FROM_OBS>
__device__ __forceinline__ void post_add_A(const Tree& T, const PostArgs& a, uint64_t R, int j, int k, double* C,
                                           int ld, bool transpose, double* kred) {
    // C += A_{jk} (or its transpose)
    if (FROM_OBS) {
        const uint64_t loc = R - a.start_l;
        const uint64_t os = a.row_s[loc], oe = a.row_e[loc];
        if (oe > os) {
            // (a region may read no rows: all its children sent dense messages)
            const int ns = (int)(oe - os);
            const double* X = a.X + os * a.ldx;
            if constexpr (RP <= 32) {
                // an RP x RP output is one 2 x 2 tile block: the rows are split
                // over the 4 waves (each wave's products are then four per
                // operand load, and its load latency is paid on a quarter of the rows)
                const int p = transpose ? k : j, q = transpose ? j : k;
                wg_gemm_tn_ksplit(C, ld, RP, RP, ns, 1.0, X + p * RP, a.ldx, RP, X + q * RP, a.ldx, RP, kred);
            } else if (!transpose)
                wg_gemm_tn<false, true, double, double, GEMM_FULL>(C, ld, RP, RP, ns, 1.0, X + j * RP, a.ldx, RP,
                                                                  X + k * RP, a.ldx, RP);
            else
                wg_gemm_tn<false, true, double, double, GEMM_FULL>(C, ld, RP, RP, ns, 1.0, X + k * RP, a.ldx, RP,
                                                                  X + j * RP, a.ldx, RP);
        }
        if (a.cslot) {
            // children that sent a dense message (their rows are zero)
            __syncthreads();
            const int lc = a.l + 1;
            const int nslc = lc * (lc + 1) / 2;
            const int sl = j * lc - j * (j - 1) / 2 + (k - j);
            for (int c = 0; c < T.J; c++) {
                const uint64_t ch = R * T.J + 1 + c;
                const int32_t cs = a.cslot[ch - a.start_c];
                if (cs < 0) continue;
                const double* S = a.Ac + ((uint64_t)cs * nslc + sl) * (uint64_t)RP * RP;
                for (int idx = threadIdx.x; idx < RP * RP; idx += MRA_WG) {
                    const int i = idx / RP, q = idx % RP;
                    C[i * ld + q] += transpose ? S[q * RP + i] : S[idx];
                }
                __syncthreads();
            }
        }
    } else {
        const int lc = a.l + 1;
        const int nslc = lc * (lc + 1) / 2;
        const int sl = j * lc - j * (j - 1) / 2 + (k - j);
        if (a.csum) {
            const double* S = a.Ac + ((R - a.start_l) * nslc + sl) * (uint64_t)RP * RP;
            for (int idx = threadIdx.x; idx < RP * RP; idx += MRA_WG) {
                const int i = idx / RP, q = idx % RP;
                C[i * ld + q] += transpose ? S[q * RP + i] : S[idx];
            }
            __syncthreads();
            return;
        }
        for (int c = 0; c < T.J; c++) {
            const uint64_t ch = R * T.J + 1 + c;
            if (!T.working[ch]) continue;
            const double* S = a.Ac + ((ch - a.start_c) * nslc + sl) * (uint64_t)RP * RP;
            for (int idx = threadIdx.x; idx < RP * RP; idx += MRA_WG) {
                const int i = idx / RP, q = idx % RP;
                C[i * ld + q] += transpose ? S[q * RP + i] : S[idx];
            }
            __syncthreads();
        }
    }
}

// Rows part of w_j at a level that reads stacked rows: rw[q] = sum over the
// region's rows o of X[o][j RP + q] z[o] (q < RP), by RP columns x MRA_WG/RP
// row phases whose partial sums are added in phase order (red: MRA_WG
// doubles).  All threads of the work group call it.
template <int RP>
__device__ __forceinline__ void wg_rows_w(const PostArgs& a, uint64_t R, int j, double* rw, double* red) {
    constexpr int PH = MRA_WG / RP;
    const uint64_t loc = R - a.start_l;
    const uint64_t os = a.row_s[loc], oe = a.row_e[loc];
    const int q = (int)threadIdx.x % RP, p = (int)threadIdx.x / RP;
    double s = 0.0;
    for (uint64_t o = os + p; o < oe; o += PH) s += a.X[o * a.ldx + j * RP + q] * a.z[o];
    red[threadIdx.x] = s;
    __syncthreads();
    if ((int)threadIdx.x < RP) {
        double t = 0.0;
        for (int ph = 0; ph < PH; ph++) t += red[ph * RP + threadIdx.x];
        rw[threadIdx.x] = t;
    }
    __syncthreads();
}

// w_j[q]; at a level that reads stacked rows, rows = the rows part (wg_rows_w)
template <int RP, bool FROM_OBS>
__device__ __forceinline__ double post_w(const Tree& T, const PostArgs& a, uint64_t R, int j, int q, double rows) {
    double s = 0.0;
    if (FROM_OBS) {
        s = rows;
        if (a.cslot) {
            const int lc = a.l + 1;
            for (int c = 0; c < T.J; c++) {
                const uint64_t ch = R * T.J + 1 + c;
                const int32_t cs = a.cslot[ch - a.start_c];
                if (cs >= 0) s += a.wc[(uint64_t)cs * lc * RP + j * RP + q];
            }
        }
    } else {
        const int lc = a.l + 1;
        bool first = true;
        for (int c = 0; c < T.J; c++) {
            const uint64_t ch = R * T.J + 1 + c;
            if (!T.working[ch]) continue;
            const double x = a.wc[(ch - a.start_c) * (uint64_t)lc * RP + j * RP + q];
            s = first ? x : s + x;
            first = false;
        }
    }
    return s;
}

template <int RP, bool FROM_OBS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_posterior(Tree T, PostArgs a) {
    constexpr int LD = RP + 1;
    __shared__ double Km[RP * LD];
    __shared__ double vv[RP];
    __shared__ double wl[RP];
    __shared__ double red[MRA_WG];
    __shared__ double dbuf[RP >= 48 ? 1 : RP * 16];
    __shared__ double kred[FROM_OBS && RP <= 32 ? 3 * 4 * 256 : 1];
    __shared__ int bad;
    const int l = a.l;
    const uint64_t R = a.start_l + a.items[blockIdx.x];
    const uint64_t msg = a.slot ? (uint64_t)a.slot[R - a.start_l] : R - a.start_l;
    const int r = T.r;
    if (threadIdx.x == 0) bad = 0;
    const double* LR = a.Lt + R * RP * RP;  // column-major L
    if (a.Xz) {
        // this region sends a dense message: its factor rows contribute nothing
        const uint64_t os = a.row_s[R - a.start_l], oe = a.row_e[R - a.start_l];
        for (uint64_t idx = threadIdx.x; idx < (oe - os) * a.ldxz; idx += MRA_WG) a.Xz[os * a.ldxz + idx] = 0.0;
        if (a.Xz32)
            for (uint64_t idx = threadIdx.x; idx < (oe - os) * a.ldxz; idx += MRA_WG) a.Xz32[os * a.ldxz + idx] = 0.0f;
        for (uint64_t o = os + threadIdx.x; o < oe; o += MRA_WG) a.zz[o] = 0.0;
    }
    for (int idx = threadIdx.x; idx < RP * RP; idx += MRA_WG) Km[(idx / RP) * LD + idx % RP] = 0.0;
    __syncthreads();
    // L L^T
    wg_gemm_tn<false, true, double, double, GEMM_FULL, false, GEMM_TRI_A | GEMM_TRI_B | GEMM_TRI_C>(
        Km, LD, RP, RP, RP, 1.0, LR, RP, RP, LR, RP, RP);
    __syncthreads();
    post_add_A<RP, FROM_OBS>(T, a, R, l, l, Km, LD, false, kred);
    __syncthreads();
    double lp = 0.0;
    for (int i = threadIdx.x; i < r; i += MRA_WG) lp += log(LR[i * RP + i]);
    const double logdet_prior = wg_sum(lp, red);
    wg_chol_blk(Km, LD, RP / 16, dbuf, &bad);
    double lq = 0.0;
    for (int i = threadIdx.x; i < r; i += MRA_WG) lq += 2.0 * log(Km[i * LD + i]);
    for (int idx = threadIdx.x; idx < RP * RP; idx += MRA_WG) {
        const int k = idx / RP, i = idx % RP;
        a.Lpt[R * RP * RP + idx] = (i >= k) ? Km[i * LD + k] : 0.0;
    }
    wg_trinv_blk(Km, LD, RP / 16, dbuf);  // Km = L~^{-1}
    for (int idx = threadIdx.x; idx < RP * RP; idx += MRA_WG) {
        const int k = idx / RP, i = idx % RP;
        a.LpiT[R * RP * RP + idx] = (i >= k) ? Km[i * LD + k] : 0.0;
    }
    if constexpr (FROM_OBS) wg_rows_w<RP>(a, R, l, wl, red);
    for (int q = threadIdx.x; q < RP; q += MRA_WG)
        wl[q] = (q < r) ? post_w<RP, FROM_OBS>(T, a, R, l, q, FROM_OBS ? wl[q] : 0.0) : 0.0;
    __syncthreads();
    for (int i = threadIdx.x; i < RP; i += MRA_WG) {
        double s = 0.0;
        for (int k = 0; k <= i; k++) s += Km[i * LD + k] * wl[k];
        vv[i] = s;
        a.v[R * RP + i] = s;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < r; i += MRA_WG) lq -= vv[i] * vv[i];
    const double post_part = wg_sum(lq, red);
    if (threadIdx.x == 0 && a.write_d) a.dterm[R] = -2.0 * logdet_prior + post_part;
    if (l > 0 && a.pj_sep) {
        if (threadIdx.x == 0 && bad) atomicOr(a.bad_flag, 1);
        return;
    }
    if (l > 0) {
        double* PR = a.P + a.goff[l] + (R - a.start_l) * (uint64_t)l * RP * RP;
        const double* LpiTR = a.LpiT + R * RP * RP;
        // P_j = L~^{-1} A_{jl}^T  (Km reused as scratch for A_{jl}^T)
        for (int j = 0; j < l; j++) {
            for (int idx = threadIdx.x; idx < RP * RP; idx += MRA_WG) Km[(idx / RP) * LD + idx % RP] = 0.0;
            __syncthreads();
            post_add_A<RP, FROM_OBS>(T, a, R, j, l, Km, LD, true, kred);
            __syncthreads();
            // P_j stored straight to global memory (row-major RP x RP)
            double* Pj = PR + (uint64_t)j * RP * RP;
            wg_gemm_tn<false, false, double, double, GEMM_FULL, false, GEMM_TRI_A>(Pj, RP, RP, RP, RP, 1.0, LpiTR, RP,
                                                                                  RP, Km, LD, RP);
            __syncthreads();
            // w~_j = w_j - P_j^T v (wl, free after v, holds the rows part of w_j)
            if (FROM_OBS && a.write_msgs) wg_rows_w<RP>(a, R, j, wl, red);
            if (a.write_msgs)
                for (int q = threadIdx.x; q < RP; q += MRA_WG) {
                    double wj = (q < r) ? post_w<RP, FROM_OBS>(T, a, R, j, q, FROM_OBS ? wl[q] : 0.0) : 0.0;
                    double s = 0.0;
                    for (int m = 0; m < RP; m++) s += Pj[m * RP + q] * vv[m];
                    a.wo[msg * (uint64_t)l * RP + j * RP + q] = wj - s;
                }
            __syncthreads();
        }
        if (!a.write_msgs) {
            if (threadIdx.x == 0 && bad) atomicOr(a.bad_flag, 1);
            return;
        }
        // A~_{jk} = A_{jk} - P_j^T P_k
        if (a.gram_sep) {
            if (threadIdx.x == 0 && bad) atomicOr(a.bad_flag, 1);
            return;
        }
        const int nsl = l * (l + 1) / 2;
        if (FROM_OBS) {
            // A_{jk} = X_j^T X_k over the stacked rows: both products in
            // registers, stored straight to the message (no LDS, no barrier
            // between blocks); children that sent dense messages are added after
            const uint64_t loc = R - a.start_l;
            const uint64_t os = a.row_s[loc], oe = a.row_e[loc];
            const int ns = oe > os ? (int)(oe - os) : 0;
            const double* X = a.X + os * a.ldx;
            for (int j = 0; j < l; j++)
                for (int k = j; k < l; k++) {
                    const int sl = j * l - j * (j - 1) / 2 + (k - j);
                    double* out = a.Ao + (msg * nsl + sl) * (uint64_t)RP * RP;
                    if constexpr (RP <= 32) {
                        // X_j^T X_k with the rows split over the waves, then - P_j^T P_k
                        for (int idx = threadIdx.x; idx < RP * RP; idx += MRA_WG) Km[(idx / RP) * LD + idx % RP] = 0.0;
                        __syncthreads();
                        if (ns > 0)
                            wg_gemm_tn_ksplit(Km, LD, RP, RP, ns, 1.0, X + j * RP, a.ldx, RP, X + k * RP, a.ldx, RP,
                                              kred);
                        wg_gemm_tn(Km, LD, RP, RP, RP, -1.0, PR + (uint64_t)j * RP * RP, RP, RP,
                                   PR + (uint64_t)k * RP * RP, RP, RP);
                        __syncthreads();
                        for (int idx = threadIdx.x; idx < RP * RP; idx += MRA_WG) out[idx] = Km[(idx / RP) * LD + idx % RP];
                        __syncthreads();
                    } else {
                        wg_gemm2_store(out, RP, RP, RP, ns, 1.0, X + j * RP, a.ldx, X + k * RP, a.ldx, RP, -1.0,
                                       PR + (uint64_t)j * RP * RP, RP, PR + (uint64_t)k * RP * RP, RP);
                    }
                }
            if (a.cslot) {
                __syncthreads();
                const int lc = l + 1;
                const int nslc = lc * (lc + 1) / 2;
                for (int c = 0; c < T.J; c++) {
                    const uint64_t ch = R * T.J + 1 + c;
                    const int32_t cs = a.cslot[ch - a.start_c];
                    if (cs < 0) continue;
                    for (int j = 0; j < l; j++)
                        for (int k = j; k < l; k++) {
                            const int sl = j * l - j * (j - 1) / 2 + (k - j);
                            const int slc = j * lc - j * (j - 1) / 2 + (k - j);
                            double* out = a.Ao + (msg * nsl + sl) * (uint64_t)RP * RP;
                            const double* S = a.Ac + ((uint64_t)cs * nslc + slc) * (uint64_t)RP * RP;
                            for (int idx = threadIdx.x; idx < RP * RP; idx += MRA_WG) out[idx] += S[idx];
                        }
                }
            }
        }
        for (int j = 0; j < (FROM_OBS ? 0 : l); j++)
            for (int k = j; k < l; k++) {
                for (int idx = threadIdx.x; idx < RP * RP; idx += MRA_WG) Km[(idx / RP) * LD + idx % RP] = 0.0;
                __syncthreads();
                post_add_A<RP, FROM_OBS>(T, a, R, j, k, Km, LD, false, kred);
                __syncthreads();
                wg_gemm_tn(Km, LD, RP, RP, RP, -1.0, PR + (uint64_t)j * RP * RP, RP, RP, PR + (uint64_t)k * RP * RP,
                           RP, RP);
                __syncthreads();
                const int sl = j * l - j * (j - 1) / 2 + (k - j);
                double* out = a.Ao + (msg * nsl + sl) * (uint64_t)RP * RP;
                for (int idx = threadIdx.x; idx < RP * RP; idx += MRA_WG) out[idx] = Km[(idx / RP) * LD + idx % RP];
                __syncthreads();
            }
    }
    if (threadIdx.x == 0 && bad) atomicOr(a.bad_flag, 1);
}

// Message blocks A~_jk = A_jk - P_j^T P_k (j <= k < l) of the regions of a
// posterior level (src/class_approximation.cpp:289-303), RP = 64: at a level
// that reads stacked rows A_jk = X_j^T X_k over the region's rows
// (observation rows H, or the factor rows of the level below) plus the dense
// messages of children that sent one (cslot); at a dense level (a.X null)
// A_jk is the sum of the children's messages (one block per region when the
// level below stored them summed, csum).  One wave per 64 x 64 block with its 4 x 4
// MFMA tiles in registers (each operand value loaded once per block, not
// once per wave of a shared block), four blocks of the same region per work
// group, launched after k_posterior has stored P.  Work group g runs on XCD
// g mod 8 (round-robin dispatch); the mapping below gives every work group
// of one region the same g mod 8, so the region's rows are re-read from that
// XCD's L2.
// SUM: the items are parents (level l-1); one wave sums block (j, k) of the
// messages of the parent's working children (in child order) and stores one
// block set per parent, which is all the parent reads (PostArgs::csum).
// F32 (mixed precision): X_j^T X_k over all rows of the item (the rows of
// every summed child) with FP32 matrix instructions and FP32 sums, widened
// to FP64 once; the dense child messages and -P_j^T P_k are added in FP64.
// Only the FP32 accumulators are live during the row loop.
// DENSE: a dense level (no rows; compiled separately to keep the row path's
// registers).
template <bool SUM, bool F32 = false, bool DENSE = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_post_gram(Tree T, PostArgs a,
                                                                                           uint32_t nitems) {
    constexpr int RP = 64;
    const int l = a.l;
    const int nsl = l * (l + 1) / 2;
    const uint32_t nq = (uint32_t)(nsl + 3) / 4;
    const uint32_t g = blockIdx.x, xcd = g & 7u, t = g >> 3;
    const uint32_t it = (t / nq) * 8u + xcd;
    const int b = (int)(t % nq) * 4 + wave_id();
    if (it >= nitems || b >= nsl) return;
    // (j, k) of message slot b: (0,0), (0,1), ..., (0,l-1), (1,1), ...
    int j = 0, rem = b;
    while (rem >= l - j) {
        rem -= l - j;
        j++;
    }
    const int k = j + rem;
    mra_d4 acc[16];
#pragma unroll
    for (int q = 0; q < 16; q++) acc[q] = mra_d4{0, 0, 0, 0};
    const int lc = l + 1;
    const int nslc = lc * (lc + 1) / 2;
    const int slc = j * lc - j * (j - 1) / 2 + (k - j);
    // FP32 row products of region loc (into f)
    auto rows_f32 = [&](mra_f4 (&f)[16], uint64_t loc) {
        const uint64_t os = a.row_s[loc], oe = a.row_e[loc];
        if (oe <= os) return;
        if (a.X32) {
            const float* X = a.X32 + os * (uint64_t)a.ldx;
            wave_mm64_f32(f, (int)(oe - os), X + j * RP, a.ldx, X + k * RP, a.ldx);
        } else {
            const double* X = a.X + os * (uint64_t)a.ldx;
            wave_mm64_f32(f, (int)(oe - os), X + j * RP, a.ldx, X + k * RP, a.ldx);
        }
    };
    if constexpr (F32) {
        mra_f4 f[16];
#pragma unroll
        for (int q = 0; q < 16; q++) f[q] = mra_f4{0, 0, 0, 0};
        if constexpr (SUM) {
            const uint64_t Q = a.start_p + a.items[it];
            for (int c = 0; c < T.J; c++) {
                const uint64_t R = Q * T.J + 1 + c;
                if (T.working[R]) rows_f32(f, R - a.start_l);
            }
        } else {
            rows_f32(f, a.items[it]);
        }
        wave_widen64(acc, f);
    }
    // A~_jk of region R (local index loc) added to acc (in F32 mode without
    // the row products, already in acc)
    auto add_block = [&](const double* S) {
        if constexpr (F32)
            wave_add64_l32(acc, S, RP);
        else
            wave_add64(acc, S, RP);
    };
    auto add_region = [&](uint64_t loc) {
        const uint64_t R = a.start_l + loc;
        const double* PR = a.P + a.goff[l] + loc * (uint64_t)l * RP * RP;
        if constexpr (!DENSE) {
            // rows, then the dense messages of children that sent one
            const uint64_t os = a.row_s[loc], oe = a.row_e[loc];
            if (!F32 && oe > os) {
                const double* X = a.X + os * (uint64_t)a.ldx;
                wave_mm64_acc(acc, (int)(oe - os), 1.0, X + j * RP, a.ldx, X + k * RP, a.ldx);
            }
            if (a.cslot)
                for (int c = 0; c < T.J; c++) {
                    const int32_t cs = a.cslot[R * T.J + 1 + c - a.start_c];
                    if (cs >= 0) add_block(a.Ac + ((uint64_t)cs * nslc + slc) * (uint64_t)RP * RP);
                }
        } else if (a.csum) {
            // dense level: the children's messages, summed per region below
            add_block(a.Ac + (loc * nslc + slc) * (uint64_t)RP * RP);
        } else {
            for (int c = 0; c < T.J; c++) {
                const uint64_t ch = R * T.J + 1 + c;
                if (T.working[ch]) add_block(a.Ac + ((ch - a.start_c) * nslc + slc) * (uint64_t)RP * RP);
            }
        }
        (void)R;
        if constexpr (F32)
            wave_mm64_acc_l32(acc, RP, -1.0, PR + (uint64_t)j * RP * RP, RP, PR + (uint64_t)k * RP * RP, RP);
        else
            wave_mm64_acc(acc, RP, -1.0, PR + (uint64_t)j * RP * RP, RP, PR + (uint64_t)k * RP * RP, RP);
    };
    auto store = [&](double* out) {
        if constexpr (F32)
            wave_store64_l32(out, RP, acc);
        else
            wave_store64(out, RP, acc);
    };
    if constexpr (SUM) {
        const uint64_t ploc = a.items[it];
        const uint64_t Q = a.start_p + ploc;
        for (int c = 0; c < T.J; c++) {
            const uint64_t R = Q * T.J + 1 + c;
            if (T.working[R]) add_region(R - a.start_l);
        }
        store(a.Ao + (ploc * nsl + b) * (uint64_t)RP * RP);
    } else {
        const uint64_t loc = a.items[it];
        add_region(loc);
        const uint64_t msg = a.slot ? (uint64_t)a.slot[loc] : loc;
        store(a.Ao + (msg * nsl + b) * (uint64_t)RP * RP);
    }
}

// P_j = L~^{-1} A_jl^T and w~_j = w_j - P_j^T v (j < l) of the regions of a
// posterior level that reads stacked rows (src/class_approximation.cpp:
// 266-306), RP = 64, after k_posterior has stored L~^{-1} and v: one wave per
// j.  A_jl^T = X_l^T X_j over the region's rows (plus the transposed dense
// messages of children that sent one) accumulates in registers; the FP64 MFMA
// result layout is the B-operand layout, so it feeds L~^{-1} A_jl^T directly
// (zero upper blocks of L~^{-1} skipped).  Work groups of one region share
// g mod 8 (one XCD), as in k_post_gram.
__device__ __forceinline__ void wave_add64_t(mra_d4 (&acc)[16], const double* S, int lds) {
    // acc += S^T (S 64 x 64 row-major), in the register layout of wave_mm64_acc
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < 4; ti++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const double* scol = S + (int64_t)r16 * lds + ti * 16 + k4 + 4 * v;
#pragma unroll
            for (int tj = 0; tj < 4; tj++) acc[ti * 4 + tj][v] += scol[16 * tj * lds];
        }
}

// F32 (mixed precision): X_l^T X_j with FP32 matrix instructions (as in
// k_post_gram), held in the L32 layout; its rows feed L~^{-1} as the k index
// of the product, so L~^{-1} is read at the same permuted k.  w_j = X_j^T z
// stays FP64.
// DENSE: a dense level (P_j from the children's messages only).
template <bool F32 = false, bool DENSE = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_post_pj(Tree T, PostArgs a,
                                                                                         uint32_t nitems) {
    constexpr int RP = 64;
    const int l = a.l;
    const uint32_t nq = (uint32_t)(l + 3) / 4;
    const uint32_t g = blockIdx.x, xcd = g & 7u, t = g >> 3;
    const uint32_t it = (t / nq) * 8u + xcd;
    const int j = (int)(t % nq) * 4 + wave_id();
    if (it >= nitems || j >= l) return;
    const int lane = threadIdx.x & 63, r16 = lane & 15, k4 = lane >> 4;
    const uint64_t loc = a.items[it];
    const uint64_t R = a.start_l + loc;
    const uint64_t msg = a.slot ? (uint64_t)a.slot[loc] : loc;
    const int r = T.r;
    // rows mode or dense level (children's messages only)
    const uint64_t os = DENSE ? 0 : a.row_s[loc], oe = DENSE ? 0 : a.row_e[loc];
    const double* X = DENSE ? nullptr : a.X + os * (uint64_t)a.ldx;
    const int ns = oe > os ? (int)(oe - os) : 0;
    const int lc = l + 1;
    const int nslc = lc * (lc + 1) / 2;
    const int slc = j * lc - j * (j - 1) / 2 + (l - j);  // child slot of (j, l)
    mra_d4 acc[16], wacc[4];
#pragma unroll
    for (int q = 0; q < 16; q++) acc[q] = mra_d4{0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; q++) wacc[q] = mra_d4{0, 0, 0, 0};
    // acc = X_l^T X_j, wacc = X_j^T z (every row)
    if constexpr (DENSE) {
    } else if constexpr (F32) {
        // FP32 row products (only f live in the loop), widened once
        mra_f4 f[16];
#pragma unroll
        for (int q = 0; q < 16; q++) f[q] = mra_f4{0, 0, 0, 0};
        if (ns > 0 && a.X32)
            wave_mm64_w_f32(f, wacc, ns, a.X32 + os * (uint64_t)a.ldx + l * RP, a.ldx, X + j * RP, a.ldx, a.z + os);
        else if (ns > 0)
            wave_mm64_w_f32(f, wacc, ns, X + l * RP, a.ldx, X + j * RP, a.ldx, a.z + os);
        wave_widen64(acc, f);
    } else if (ns > 0) {
        wave_mm64_acc_w(acc, wacc, ns, X + l * RP, a.ldx, X + j * RP, a.ldx, a.z + os);
    }
    double wq[4];  // w_j[16 tj + r16]: every row of wacc holds it
#pragma unroll
    for (int tj = 0; tj < 4; tj++) wq[tj] = wacc[tj][0];
    auto add_block_t = [&](const double* S) {
        if constexpr (F32)
            wave_add64_t_l32(acc, S, RP);
        else
            wave_add64_t(acc, S, RP);
    };
    if (!DENSE && a.cslot) {
        for (int c = 0; c < T.J; c++) {
            const int32_t cs = a.cslot[R * T.J + 1 + c - a.start_c];
            if (cs >= 0) add_block_t(a.Ac + ((uint64_t)cs * nslc + slc) * (uint64_t)RP * RP);
        }
    } else if (DENSE && a.csum) {
        add_block_t(a.Ac + (loc * nslc + slc) * (uint64_t)RP * RP);
    } else if (DENSE) {
        for (int c = 0; c < T.J; c++) {
            const uint64_t ch = R * T.J + 1 + c;
            if (T.working[ch]) add_block_t(a.Ac + ((ch - a.start_c) * nslc + slc) * (uint64_t)RP * RP);
        }
    }
    // P_j = L~^{-1} acc by row tiles tq (k tiles kt <= tq), stored row-major;
    // pv[tj] collects this lane's part of (P_j^T v)[16 tj + r16]
    const char* Li = (const char*)(a.LpiT + R * RP * RP);  // [k*RP + i] = L~^{-1}[i][k]
    double* Pj = a.P + a.goff[l] + loc * (uint64_t)l * RP * RP + (uint64_t)j * RP * RP;
    char* Pc = (char*)Pj;
    const double* vR = a.v + R * RP;
    double pv[4] = {0.0, 0.0, 0.0, 0.0};
    // element v of lane t of acc tile kt is row 16 kt + k4 + 4v (FP64
    // layout) or 16 kt + 4 k4 + v (L32): the k of L~^{-1} read with it
    // (the product gq is in the FP64 layout either way: lo addresses its rows)
    constexpr uint32_t KV = F32 ? 1u : 4u;
    uint32_t lo = (uint32_t)(k4 * RP + r16);
    const uint32_t lk = F32 ? (uint32_t)(3 * k4 * RP) : 0u;
#pragma unroll
    for (int tq = 0; tq < 4; tq++) {
        asm volatile("" : "+v"(lo));  // keep the displaced offsets in the loop
        mra_d4 gq[4];
#pragma unroll
        for (int tj = 0; tj < 4; tj++) gq[tj] = mra_d4{0, 0, 0, 0};
#pragma unroll
        for (int kt = 0; kt <= tq; kt++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const double av = *(const double*)(Li + 8u * (lo + lk + (16 * kt + KV * v) * RP + 16 * tq));
#pragma unroll
                for (int tj = 0; tj < 4; tj++)
                    gq[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, acc[kt * 4 + tj][v], gq[tj], 0, 0, 0);
            }
        uint32_t vo = (uint32_t)k4;
        asm volatile("" : "+v"(vo));  // v loaded here, not hoisted above the products
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const double vm = vR[16 * tq + vo + 4 * v];
#pragma unroll
            for (int tj = 0; tj < 4; tj++) {
                *(double*)(Pc + 8u * (lo + (16 * tq + 4 * v) * RP + 16 * tj)) = gq[tj][v];
                pv[tj] += gq[tj][v] * vm;
            }
        }
    }
    if (!a.write_msgs) return;
    // w~_j = w_j - P_j^T v: the four lane groups' parts summed across lanes
#pragma unroll
    for (int tj = 0; tj < 4; tj++) {
        pv[tj] += __shfl_xor(pv[tj], 16);
        pv[tj] += __shfl_xor(pv[tj], 32);
    }
    if (k4 == 0)
#pragma unroll
        for (int tj = 0; tj < 4; tj++) {
            const int q = 16 * tj + r16;
            double wj = 0.0;
            if (q < r && !DENSE) {
                wj = wq[tj];
                if (a.cslot)
                    for (int c = 0; c < T.J; c++) {
                        const int32_t cs = a.cslot[R * T.J + 1 + c - a.start_c];
                        if (cs >= 0) wj += a.wc[(uint64_t)cs * lc * RP + j * RP + q];
                    }
            } else if (q < r) {
                // dense level: w_j summed over the working children in order
                bool first = true;
                for (int c = 0; c < T.J; c++) {
                    const uint64_t ch = R * T.J + 1 + c;
                    if (!T.working[ch]) continue;
                    const double x = a.wc[(ch - a.start_c) * (uint64_t)lc * RP + j * RP + q];
                    wj = first ? x : wj + x;
                    first = false;
                }
            }
            a.wo[msg * (uint64_t)l * RP + j * RP + q] = wj - pv[tj];
        }
}

static inline uint32_t post_pj_grid(uint32_t nitems, int l) {
    return ((nitems + 7u) / 8u) * ((uint32_t)(l + 3) / 4) * 8u;
}

// Grid of k_post_gram for nitems regions of level l.
static inline uint32_t post_gram_grid(uint32_t nitems, int l) {
    const uint32_t nq = (uint32_t)(l * (l + 1) / 2 + 3) / 4;
    return ((nitems + 7u) / 8u) * nq * 8u;
}

// Factored posterior of a finest-but-one region R (level l = L-1).  The
// message A~ = X^T S X, w~ = X^T S z of the dense update (rank <= n_s, the
// observations below R) is emitted as factor rows instead:
//   W = L_R^{-1} X_l^T,  I + W^T W = Q Q^T (S = Q^{-T} Q^{-1} by Woodbury),
//   X' = Q^{-1} X[:, 0:l*RP],  z' = Q^{-1} z   =>  A~ = X'^T X', w~ = X'^T z'
//   d_R = 2 sum log diag Q - z^T z + z'^T z'   (determinant lemma)
// The parent level reads the rows X' like stacked observation rows.
// (Replaces the dense update of src/class_approximation.cpp:266-306 and
// src/class_approximation-posterior-inference.cpp:51-131 for this level.)
struct LowRankArgs {
    int l;
    uint64_t start_l;
    const uint32_t* items;
    const uint64_t* row_s;
    const uint64_t* row_e;
    const double* X;  // input rows, width (l+1)*RP
    int ldx;
    const double* z;
    const double* LiT;  // prior L^{-1}, column-major
    double* Xo;         // output rows, width l*RP
    float* Xo32;        // mixed precision: their FP32 copy (same layout), or null
    int ldo;
    double* zo;
    double* dterm;
    double* ws;  // W = L^{-1} X_l^T of each region: RP x n_s at ws + row_s * RP
    double* qws; // split form: (Q^{-1})^T of item k, NB x NB at qws + k NB^2
    int* bad_flag;
};

// Regions with at most NB (16, 32, 64 or 128) rows: Q in LDS, W through the
// global workspace, X' stored straight from the product to the output rows.
// SPLIT: (Q^{-1})^T is stored to qws instead and k_post_lowrank_x forms X'
// by column blocks in a launch of its own (the product is most of the work
// and one work group per region leaves most of the chip idle for NB = 128,
// whose LDS allows one work group per CU).
template <int RP, int NB, bool SPLIT = false>
__global__ __launch_bounds__(256) void k_post_lowrank(Tree T, LowRankArgs a) {
    constexpr int LDA = NB + 1;
    __shared__ double Aq[NB * LDA];
    __shared__ double dbuf[NB >= 48 ? 1 : NB * 16];
    __shared__ double red[MRA_WG];
    __shared__ int bad;
    const uint32_t loc = a.items[blockIdx.x];
    const uint64_t R = a.start_l + loc;
    const uint64_t os = a.row_s[loc];
    const int ns = (int)(a.row_e[loc] - os);
    const int l = a.l;
    if (threadIdx.x == 0) bad = 0;
    if (ns <= 0) {
        if (threadIdx.x == 0) a.dterm[R] = 0.0;
        return;
    }
    const int nsp = (ns + 15) & ~15, nbk = nsp >> 4;
    const double* Xr = a.X + os * (uint64_t)a.ldx;
    double* W = a.ws + os * (uint64_t)RP;  // RP x ns, row-major
    // 1. W = L^{-1} X_l^T:  W[i][o] = sum_k LiT[k*RP + i] X_l[o][k]
    wg_gemm_tn_impl<double, false, false, double, double, GEMM_ANY, double, true>(
        W, ns, RP, nsp, RP, 1.0, a.LiT + R * RP * RP, RP, RP, Xr + l * RP, a.ldx, ns);
    // identity (and padding) of Q's block while W is written
    for (int idx = threadIdx.x; idx < nsp * nsp; idx += MRA_WG) {
        const int o = idx / nsp, q = idx % nsp;
        Aq[o * LDA + q] = (o == q) ? 1.0 : 0.0;
    }
    __syncthreads();
    // 2. A = I + W^T W  (ns x ns in LDS)
    wg_gemm_tn(Aq, LDA, nsp, nsp, RP, 1.0, W, ns, ns, W, ns, ns);
    // 3. Q = chol(A), 4. Q^{-1} in place
    wg_chol_blk(Aq, LDA, nbk, dbuf, &bad);
    double part = 0.0;
    for (int o = threadIdx.x; o < ns; o += MRA_WG) part += 2.0 * log(Aq[o * LDA + o]);
    wg_trinv_blk(Aq, LDA, nbk, dbuf);
    // 5. z' = Q^{-1} z, d_R = 2 sum log diag Q - z^T z + z'^T z'
    for (int o = threadIdx.x; o < ns; o += MRA_WG) {
        double s = 0.0;
        for (int k = 0; k <= o; k++) s += Aq[o * LDA + k] * a.z[os + k];
        const double z0 = a.z[os + o];
        a.zo[os + o] = s;
        part += s * s - z0 * z0;
    }
    const double d = wg_sum(part, red);
    if (threadIdx.x == 0) a.dterm[R] = d;
    // 6. X' = Q^{-1} X[:, 0:l*RP] straight to the output rows
    const int Wd = l * RP;
    if constexpr (SPLIT) {
        // Qt[k*nsp + i] = Q^{-1}[i][k] (zero for k > i)
        double* Qt = a.qws + (uint64_t)blockIdx.x * NB * NB;
        for (int idx = threadIdx.x; idx < nsp * nsp; idx += MRA_WG) {
            const int k = idx / nsp, i = idx % nsp;
            Qt[idx] = (i >= k) ? Aq[i * LDA + k] : 0.0;
        }
    } else if (Wd > 0 && a.Xo32) {
        wg_gemm_tn<true, false, double, double, GEMM_PART, true>(a.Xo + os * (uint64_t)a.ldo, a.ldo, nsp, Wd, ns, 1.0,
                                                                  Aq, LDA, ns, Xr, a.ldx, Wd,
                                                                  a.Xo32 + os * (uint64_t)a.ldo);
    } else if (Wd > 0) {
        wg_gemm_tn<true, false>(a.Xo + os * (uint64_t)a.ldo, a.ldo, nsp, Wd, ns, 1.0, Aq, LDA, ns, Xr, a.ldx, Wd);
    }
    if (threadIdx.x == 0 && bad) atomicOr(a.bad_flag, 1);
}

// X' = Q^{-1} X[:, 0:l*RP] for the regions of one k_post_lowrank<.., NB, true>
// launch (same items, same order): work group g forms the 64 columns
// 64 (g mod nblk) .. of item g / nblk from the stored (Q^{-1})^T; the upper
// zeros of Q^{-1} are skipped (K of each wave block stops at its last row).
// wd = l*RP columns, nblk = ceil(wd / 64).
template <int NB>
__global__ __launch_bounds__(256) void k_post_lowrank_x(LowRankArgs a, int wd, int nblk) {
    const uint32_t it = blockIdx.x / (uint32_t)nblk;
    const int cb = (int)(blockIdx.x % (uint32_t)nblk);
    const uint32_t loc = a.items[it];
    const uint64_t os = a.row_s[loc];
    const int ns = (int)(a.row_e[loc] - os);
    if (ns <= 0) return;
    const int nsp = (ns + 15) & ~15;
    const double* Qt = a.qws + (uint64_t)it * NB * NB;
    const uint64_t oo = os * (uint64_t)a.ldo + cb * 64;
    if (a.Xo32)
        wg_gemm_tn<false, false, double, double, GEMM_ANY, true, GEMM_TRI_A>(
            a.Xo + oo, a.ldo, nsp, 64, ns, 1.0, Qt, nsp, ns, a.X + os * (uint64_t)a.ldx + cb * 64, a.ldx,
            min(64, wd - cb * 64), a.Xo32 + oo);
    else
        wg_gemm_tn<false, false, double, double, GEMM_ANY, false, GEMM_TRI_A>(
            a.Xo + oo, a.ldo, nsp, 64, ns, 1.0, Qt, nsp, ns, a.X + os * (uint64_t)a.ldx + cb * 64, a.ldx,
            min(64, wd - cb * 64));
}

// ------------------------------------------------------------- prediction
// Pre-pass (…finest-level-in-prior.cpp:271-386) in two launches: the
// ancestor chain per tile of a segment's points (G^_j, E_j into Bt, var0) and
// each finest region's own part per tile of its points (G^_L, mean0, var0,
// B~_j = E_j - H_j^T G^_L); then the top-down post-pass per segment tile.
struct PredArgs {
    const uint32_t* items;  // (F, p0) pairs or, for k_pred_local, (F, p0, scratch slot) triples
    const uint64_t* goff;
    const double* Gc;
    const double* LiT;
    const void* Gf;  // double or float (mixed precision)
    const double* Sig;
    const double* H;
    const double* u;
    double* Bt;    // B~ per group: (m, p) at L*RP*p_grp_os + m*p_grp_n + (p - p_grp_os)
    double* Gh;    // G^ of the ancestor chain, the same layout
    double* GL;    // scratch: gl_off[slot] (n_F x 64)
    const uint64_t* gl_off;
    double* mean;
    double* var;
    // posterior
    const double* LpiT;
    const double* P;
    const double* v;
};

// Prediction pre-pass, part 1: the conditioning chain on the ancestors, in
// tiles of NT prediction points of a group (the finest regions below one
// grandparent: every ancestor but the parent is shared), items (F, p0), F
// the group's first finest region.  Per level j: E_j = C(Q_{a_j}, p) -
// G_{a_j}^T G^ (stored as the rows of B~), G^_j = L_{a_j}^{-1} E_j (stored),
// var = V(p) - sum_m G^(m, p)^2, mean = 0.  Levels below the parent's run
// once on the whole tile; the parent's level once per segment (the points
// below one parent) of the tile, on its column range.  B~ and G^ of a group
// are stored together: (m, c) at LRP*p_grp_os + m*p_grp_n + (c - p_grp_os).
// (…finest-level-in-prior.cpp:271-317)
template <int RP, int NT>
__global__ __launch_bounds__(256) void k_pred_chain(Tree T, CovCtx cc, Pts K, Pts Pp, PredArgs a) {
    constexpr int LD = 2 * NT + 1;  // room for a segment's column offset (< NT) plus a tile
    __shared__ double Cm[RP * LD];
    __shared__ double Wm[RP * LD];
    __shared__ double stage[5 * RP];
    __shared__ double red[MRA_WG];
    const uint32_t F0 = a.items[2 * blockIdx.x], p0 = a.items[2 * blockIdx.x + 1];
    const uint64_t pg = T.p_grp_os[F0];
    const int pG = (int)T.p_grp_n[F0];
    const int np = min(NT, pG - (int)p0);
    const int r = T.r, L = T.L, LRP = T.LRP;
    double* BtG = a.Bt + (uint64_t)LRP * pg + p0;  // (m, c) at BtG[m * pG + c], c tile column
    double* GhG = a.Gh + (uint64_t)LRP * pg + p0;  // the same layout
    const int tc = threadIdx.x % NT;               // this thread's tile column in every loop below
    double sq = 0.0;                               // its part of sum_m G^(m, tc)^2
    // one level j on the tile columns [off, off + cnt) with ancestor an
    auto level = [&](int j, uint64_t an, int off, int cnt) {
        cov_tile<RP, NT>(Cm + off, LD, cc, K, an * RP, r, Pp, pg + p0 + off, cnt, false, stage);
        __syncthreads();
        if (j > 0) {
            const double* Ga = a.Gc + a.goff[j] + (an - lvl_start(j, T.J)) * (uint64_t)j * RP * RP;
            wg_gemm_tn(Cm + off, LD, RP, NT, j * RP, -1.0, Ga, RP, RP, GhG + off, pG, cnt);
            __syncthreads();
        }
        for (int idx = threadIdx.x; idx < RP * NT; idx += MRA_WG) {
            const int i = idx / NT;
            if (tc >= off && tc < off + cnt) BtG[(uint64_t)(j * RP + i) * pG + tc] = Cm[i * LD + tc];
        }
        // Wm = L^{-1} Cm on the range (every element of the NT columns stored)
        wg_gemm_tn<false, false, double, double, GEMM_FULL, false, GEMM_TRI_A>(Wm + off, LD, RP, NT, RP, 1.0,
                                                                              a.LiT + an * RP * RP, RP, RP,
                                                                              Cm + off, LD, NT);
        __syncthreads();
        for (int idx = threadIdx.x; idx < RP * NT; idx += MRA_WG) {
            const int i = idx / NT;
            if (tc >= off && tc < off + cnt) {
                const double w = Wm[i * LD + tc];
                GhG[(uint64_t)(j * RP + i) * pG + tc] = w;
                sq += w * w;
            }
        }
        __syncthreads();
    };
    for (int j = 0; j + 1 < L; j++) level(j, ancestor_up(T.parent[F0], L - 1, j, T.J), 0, np);
    // the parent's level, segment by segment over the tile's columns
    uint32_t f = F0;
    for (int c = 0; c < np;) {
        const uint64_t pc = pg + p0 + c;  // global index of the point at tile column c
        while (T.pred_off[f + 1] <= pc) f++;
        const uint64_t se = T.p_seg_os[f] + T.p_seg_n[f];
        const int cnt = (int)min((uint64_t)(np - c), se - pc);
        level(L - 1, T.parent[f], c, cnt);
        c += cnt;
    }
    red[threadIdx.x] = sq;
    __syncthreads();
    if (threadIdx.x < np) {
        double s = 0.0;
        for (int q = threadIdx.x; q < MRA_WG; q += NT) s += red[q];
        a.var[pg + p0 + threadIdx.x] = Pp.v[pg + p0 + threadIdx.x] - s;
        a.mean[pg + p0 + threadIdx.x] = 0.0;
    }
}

// Prediction pre-pass, part 2, per finest region F with observations, in
// tiles of NT of its points (items (F, p0, scratch slot)): the conditional
// cross block with F's observations G^_L = L_F^{-1} (C(S_F, p) - G_F^T G^),
// mean = u^T G^_L, var -= |G^_L|^2, and B~ = E - H^T G^_L in place.
// (…finest-level-in-prior.cpp:318-386)
template <int RP, typename TG, int NT>
__global__ __launch_bounds__(256) void k_pred_local(Tree T, CovCtx cc, Pts O, Pts Pp, PredArgs a) {
    constexpr int LD = NT + 1;
    __shared__ double Cm[RP * LD];
    __shared__ double stage[5 * RP];
    __shared__ double kred[NT == 16 ? 3 * 4 * 256 : 1];  // K-split partial tiles
    const uint32_t F = a.items[3 * blockIdx.x], p0 = a.items[3 * blockIdx.x + 1], slot = a.items[3 * blockIdx.x + 2];
    const uint64_t ob = T.obs_off[F];
    const int nF = (int)(T.obs_off[F + 1] - ob);
    if (nF == 0) return;
    const uint64_t pb = T.pred_off[F];
    const int pF = (int)(T.pred_off[F + 1] - pb);
    const int np = min(NT, pF - (int)p0);
    const int LRP = T.LRP;
    const uint64_t ps = T.p_grp_os[F];
    const int pS = (int)T.p_grp_n[F];
    const uint64_t cb = (pb - ps) + p0;                 // the tile's first column in the group
    double* BtS = a.Bt + (uint64_t)LRP * ps + cb;       // (m, c) at BtS[m * pS + c]
    const double* GhS = a.Gh + (uint64_t)LRP * ps + cb;
    const uint64_t so = T.f_seg_os[F];
    const int sn = (int)T.f_seg_n[F];
    const TG* GF = (const TG*)a.Gf + (uint64_t)LRP * so + (ob - so);
    const double* LF = a.Sig + T.sig_off[F];
    const int c = threadIdx.x;
    // G^_L in Cm (one row tile) or in the scratch GL (ld NT)
    double* GL = nF <= RP ? Cm : a.GL + a.gl_off[slot];
    const int ldg = nF <= RP ? LD : NT;
    if (nF <= RP) {
        cov_tile<RP, NT>(Cm, LD, cc, O, ob, nF, Pp, pb + p0, np, false, stage);
        __syncthreads();
        if (NT == 16 && nF <= 32)  // at most 2 x 1 output tiles: K over the four waves
            wg_gemm_tn_ksplit<TG, double>(Cm, LD, (nF + 15) & ~15, NT, LRP, -1.0, GF, sn, nF, GhS, pS, np, kred);
        else
            wg_gemm_tn(Cm, LD, (nF + 15) & ~15, NT, LRP, -1.0, GF, sn, nF, GhS, pS, np);
        __syncthreads();
    } else {
        for (int t0 = 0; t0 < nF; t0 += RP) {  // row tiles of RP (the tile height of Cm)
            const int mv = min(RP, nF - t0);
            cov_tile<RP, NT>(Cm, LD, cc, O, ob + t0, mv, Pp, pb + p0, np, false, stage);
            __syncthreads();
            wg_gemm_tn(Cm, LD, (mv + 15) & ~15, NT, LRP, -1.0, GF + t0, sn, mv, GhS, pS, np);
            __syncthreads();
            for (int idx = threadIdx.x; idx < mv * NT; idx += MRA_WG)
                GL[(uint64_t)(t0 + idx / NT) * NT + idx % NT] = Cm[(idx / NT) * LD + idx % NT];
            __syncthreads();
        }
    }
    // G^_L <- L_F^{-1} G^_L, one column per thread
    if (c < NT) {
        for (int i = 0; i < nF; i++) {
            double s = GL[(uint64_t)i * ldg + c];
            for (int k = 0; k < i; k++) s -= LF[(uint64_t)i * nF + k] * GL[(uint64_t)k * ldg + c];
            s /= LF[(uint64_t)i * nF + i];
            GL[(uint64_t)i * ldg + c] = s;
        }
        if (c < np) {
            double sv = 0.0, sm = 0.0;
            for (int i = 0; i < nF; i++) {
                const double g = GL[(uint64_t)i * ldg + c];
                sv += g * g;
                sm += g * a.u[ob + i];
            }
            a.var[pb + p0 + c] -= sv;
            a.mean[pb + p0 + c] = sm;
        }
    }
    __syncthreads();
    // B~ = E - H^T G^_L for all L*RP rows at once, accumulated in place on
    // the stored E (rows m = j*RP + i, the tile's np valid columns)
    wg_gemm_tn(BtS, pS, LRP, NT, nF, -1.0, a.H + ob * LRP, LRP, LRP, GL, ldg, np);
}

// Top-down posterior prediction (src/class_approximation-predict.cpp:123-188):
//   Z_j = L~_{a_j}^{-1} (B~_j - sum_{j'>j} P_{a_j',j}^T Z_j'),
//   mean += Z_j^T v_{a_j}, var += colsumsq(Z_j).
// In 16-point tiles of a group (items as k_pred_chain): the parent's level
// first, once per segment of the tile on its column range, then the levels
// above on the whole tile (every column is independent of the others).
template <int RP, int NT>
__global__ __launch_bounds__(256) void k_pred_post(Tree T, PredArgs a) {
    constexpr int LD = 2 * NT + 1;  // room for a segment's column offset (< NT) plus a tile
    __shared__ double Zm[RP * LD];
    const uint32_t F0 = a.items[2 * blockIdx.x], p0 = a.items[2 * blockIdx.x + 1];
    const uint64_t pg = T.p_grp_os[F0];
    const int pG = (int)T.p_grp_n[F0];
    const int np = min(NT, pG - (int)p0);
    const int L = T.L, LRP = T.LRP;
    double* BtG = a.Bt + (uint64_t)LRP * pg + p0;  // B~(m, c) at BtG[m * pG + c], c tile column
    double macc = 0.0, vacc = 0.0;
    const int c = threadIdx.x;
    // level jj with ancestor an on the tile columns [off, off + cnt)
    auto step = [&](int jj, uint64_t an, int off, int cnt) {
        // Z = L~^{-1} B~_jj, B~ read from global memory
        wg_gemm_tn<false, false, double, double, GEMM_PART, false, GEMM_TRI_A>(
            Zm + off, LD, RP, NT, RP, 1.0, a.LpiT + an * RP * RP, RP, RP, BtG + (uint64_t)jj * RP * pG + off, pG,
            cnt);
        __syncthreads();
        if (c >= off && c < off + cnt) {
            const double* vA = a.v + an * RP;
            double sm = 0.0, sv = 0.0;
            for (int m = 0; m < RP; m++) {
                const double z = Zm[m * LD + c];
                sm += z * vA[m];
                sv += z * z;
            }
            macc += sm;
            vacc += sv;
        }
        // B~_jp -= P_{an,jp}^T Z in place (disjoint rows for each jp)
        if (jj > 0) {
            const double* PA = a.P + a.goff[jj] + (an - lvl_start(jj, T.J)) * (uint64_t)jj * RP * RP;
            for (int jp = 0; jp < jj; jp++)
                wg_gemm_tn(BtG + (uint64_t)jp * RP * pG + off, pG, RP, NT, RP, -1.0, PA + (uint64_t)jp * RP * RP, RP,
                           RP, Zm + off, LD, cnt);
        }
        __syncthreads();
    };
    uint32_t f = F0;
    for (int c0 = 0; c0 < np;) {
        const uint64_t pc = pg + p0 + c0;  // global index of the point at tile column c0
        while (T.pred_off[f + 1] <= pc) f++;
        const uint64_t se = T.p_seg_os[f] + T.p_seg_n[f];
        const int cnt = (int)min((uint64_t)(np - c0), se - pc);
        step(L - 1, T.parent[f], c0, cnt);
        c0 += cnt;
    }
    for (int jj = L - 2; jj >= 0; jj--) step(jj, ancestor_up(T.parent[F0], L - 1, jj, T.J), 0, np);
    if (c < np) {
        a.mean[pg + p0 + c] += macc;
        a.var[pg + p0 + c] += vacc;
    }
}

// ---------------------------------------------------------- prior sampler
// Counter-based standard normals: splitmix64 of (seed, index), Box-Muller.
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ double std_normal(uint64_t seed, uint64_t idx) {
    const uint64_t a = splitmix64(seed ^ splitmix64(2 * idx + 1));
    const uint64_t b = splitmix64(a ^ 0xD1B54A32D192ED03ull);
    const double u1 = ((a >> 11) + 1.0) * (1.0 / 9007199254740993.0);  // (0, 1]
    const double u2 = (b >> 11) * (1.0 / 9007199254740992.0);           // [0, 1)
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

__global__ void k_normals(uint64_t seed, uint64_t base, uint64_t n, double* out) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = std_normal(seed, base + i);
}

// y_F = sum_k G_{F,k}^T xi_{a_k} + L_F xi_F: a draw from N(0, Sigma_MRA), since
// G_{F,k}^T G_{F',k} = C_k(S_F, Q) K^{-1} C_k(Q, S_F') and L_F L_F^T = Sigma_F.
// One work group per non-empty finest region, one thread per observation.
template <typename TG>
__global__ __launch_bounds__(256) void k_sample_finest(Tree T, const uint32_t* items, const TG* Gf,
                                                       const double* Sig, const double* xi_c, const double* xi_o,
                                                       double* y) {
    const uint32_t F = items[blockIdx.x];
    const uint64_t ob = T.obs_off[F];
    const int n = (int)(T.obs_off[F + 1] - ob);
    const int RP = T.RP, L = T.L;
    const uint64_t so = T.f_seg_os[F];
    const int sn = (int)T.f_seg_n[F];
    const TG* GF = Gf + (uint64_t)T.LRP * so + (ob - so);
    const double* LF = Sig + T.sig_off[F];
    for (int o = threadIdx.x; o < n; o += MRA_WG) {
        double s = 0.0;
        for (int j = 0; j < L; j++) {
            const uint64_t a = ancestor_up(T.parent[F], L - 1, j, T.J);
            for (int m = 0; m < T.r; m++) s += GF[(uint64_t)(j * RP + m) * sn + o] * xi_c[a * RP + m];
        }
        for (int k = 0; k <= o; k++) s += LF[(uint64_t)o * n + k] * xi_o[ob + k];
        y[ob + o] = s;
    }
}

// -------------------------------------------------------------- self test
// Exercises the work-group building blocks on fixed inputs (one work group):
//   Cout = C0 - A^T B (MFMA, K rows, Mv x Nv valid part), Lout = chol(S),
//   Xout = L^{-1} Bt.   All 64 x 64 row-major except A, B (K x 64).
__global__ __launch_bounds__(256) void k_selftest(const double* A, const double* B, int K, int Mv, int Nv,
                                                  const double* C0, const double* S, const double* Bt, double* Cout,
                                                  double* Lout, double* Xout, int* bad_out) {
    constexpr int LD = 65;
    __shared__ double Cm[64 * LD];
    __shared__ double Wm[64 * LD];
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    for (int idx = threadIdx.x; idx < 64 * 64; idx += MRA_WG) Cm[(idx / 64) * LD + idx % 64] = C0[idx];
    __syncthreads();
    wg_gemm_tn(Cm, LD, 64, 64, K, -1.0, A, 64, Mv, B, 64, Nv);
    __syncthreads();
    for (int idx = threadIdx.x; idx < 64 * 64; idx += MRA_WG) Cout[idx] = Cm[(idx / 64) * LD + idx % 64];
    for (int idx = threadIdx.x; idx < 64 * 64; idx += MRA_WG) {
        Cm[(idx / 64) * LD + idx % 64] = S[idx];
        Wm[(idx / 64) * LD + idx % 64] = Bt[idx];
    }
    __syncthreads();
    wg_chol(Cm, LD, 64, &bad);
    wg_trsm(Cm, LD, Wm, LD, 64, 64);
    for (int idx = threadIdx.x; idx < 64 * 64; idx += MRA_WG) {
        Lout[idx] = Cm[(idx / 64) * LD + idx % 64];
        Xout[idx] = Wm[(idx / 64) * LD + idx % 64];
    }
    if (threadIdx.x == 0) *bad_out = bad;
}

// -------------------------------------------------- variogram features
// Empirical semivariogram of the observations in 4 quadrants x 16 lag bins
// (features of the warm-start network).  All pairs of the (at most VG_S)
// points of a quadrant; lag d falls in bin floor(16 sqrt(d / hmax)).  Sums of
// (y_i - y_j)^2 are accumulated in fixed point (2^-24 units): each thread
// keeps its 16 bins in registers, then one integer atomic per bin and thread,
// so the result does not depend on the order of the additions.
constexpr int VG_S = 16384, VG_BINS = 16;
constexpr double VG_SCALE = 16777216.0;

__global__ __launch_bounds__(256) void k_variogram(const uint32_t* idx, const uint32_t* cnt, const double* hmax,
                                                   const double* x, const double* y, const double* val,
                                                   unsigned long long* acc) {
    __shared__ unsigned long long s_sum[VG_BINS];
    __shared__ unsigned long long s_cnt[VG_BINS];
    const int q = blockIdx.y;
    const int S = (int)cnt[q];
    if (threadIdx.x < VG_BINS) {
        s_sum[threadIdx.x] = 0ull;
        s_cnt[threadIdx.x] = 0ull;
    }
    __syncthreads();
    const uint32_t* id = idx + (int64_t)q * VG_S;
    const double hm = hmax[q];
    unsigned long long ls[VG_BINS];
    unsigned int lc[VG_BINS];
#pragma unroll
    for (int k = 0; k < VG_BINS; k++) {
        ls[k] = 0ull;
        lc[k] = 0u;
    }
    const int64_t np = (int64_t)S * S;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(p / S), j = (int)(p % S);
        if (j <= i) continue;
        const uint32_t a = id[i], b = id[j];
        const double dx = x[a] - x[b], dy = y[a] - y[b];
        const double d = sqrt(dx * dx + dy * dy);
        if (d >= hm) continue;
        const int bin = min(VG_BINS - 1, (int)(VG_BINS * sqrt(d / hm)));
        const double e = val[a] - val[b];
        const unsigned long long v = (unsigned long long)(e * e * VG_SCALE + 0.5);
#pragma unroll
        for (int k = 0; k < VG_BINS; k++) {
            const bool m = k == bin;
            ls[k] += m ? v : 0ull;
            lc[k] += m ? 1u : 0u;
        }
    }
#pragma unroll
    for (int k = 0; k < VG_BINS; k++)
        if (lc[k]) {
            atomicAdd(&s_sum[k], ls[k]);
            atomicAdd(&s_cnt[k], (unsigned long long)lc[k]);
        }
    __syncthreads();
    if (threadIdx.x < VG_BINS) {
        atomicAdd(&acc[(q * VG_BINS + threadIdx.x) * 2], s_sum[threadIdx.x]);
        atomicAdd(&acc[(q * VG_BINS + threadIdx.x) * 2 + 1], s_cnt[threadIdx.x]);
    }
}

// Blocked Cholesky / triangular inverse on one N x N matrix S (row-major):
// Lout = lower factor (strict upper reported as zero), Xout = L^{-1} as left
// in place (its strict upper part must be zero).
template <int N>
__global__ __launch_bounds__(256) void k_selftest_blk(const double* S, double* Lout, double* Xout, int* bad_out) {
    constexpr int LD = N + 1;
    __shared__ double Cm[N * LD];
    __shared__ double dbuf[N >= 48 ? 1 : N * 16];
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    for (int idx = threadIdx.x; idx < N * N; idx += MRA_WG) Cm[(idx / N) * LD + idx % N] = S[idx];
    wg_chol_blk(Cm, LD, N / 16, dbuf, &bad);
    for (int idx = threadIdx.x; idx < N * N; idx += MRA_WG) {
        const int i = idx / N, k = idx % N;
        Lout[idx] = k <= i ? Cm[i * LD + k] : 0.0;
    }
    wg_trinv_blk(Cm, LD, N / 16, dbuf);
    for (int idx = threadIdx.x; idx < N * N; idx += MRA_WG) Xout[idx] = Cm[(idx / N) * LD + idx % N];
    if (threadIdx.x == 0) *bad_out = bad;
}

__global__ __launch_bounds__(256) void k_selftest_f32(const float* A, const float* B, int K, int Mv, int Nv,
                                                      const float* C0, float* Cout) {
    constexpr int LD = 65;
    __shared__ float Cm[64 * LD];
    for (int idx = threadIdx.x; idx < 64 * 64; idx += MRA_WG) Cm[(idx / 64) * LD + idx % 64] = C0[idx];
    __syncthreads();
    wg_gemm_tn_f32(Cm, LD, 64, 64, K, -1.0f, A, 64, Mv, B, 64, Nv);
    __syncthreads();
    for (int idx = threadIdx.x; idx < 64 * 64; idx += MRA_WG) Cout[idx] = Cm[(idx / 64) * LD + idx % 64];
}

// ------------------------------------------------------------------ batch
// Scatter of a batch's parameters: field i's six parameters prm[6i..6i+5]
// to that field's own parameter buffer dst[i].
__global__ void k_batch_params(int n, const double* __restrict__ prm, double* const* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 6 * n) dst[i / 6][i % 6] = prm[i];
}

// Gather of a batch's results: out[2i] = D of field i, out[2i+1] = its
// non-positive-pivot flag.
__global__ void k_batch_results(int n, const double* const* __restrict__ D, const int* const* __restrict__ bad,
                                double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        out[2 * i] = *D[i];
        out[2 * i + 1] = *bad[i] ? 1.0 : 0.0;
    }
}

// ------------------------------------------------------------- reductions
// Deterministic sum of x[0..n) into out[0] (one work group, fixed order).
__global__ __launch_bounds__(256) void k_reduce(const double* x, uint64_t n, double* out) {
    __shared__ double red[MRA_WG];
    double s = 0.0;
    for (uint64_t i = threadIdx.x; i < n; i += MRA_WG) s += x[i];
    double t = wg_sum(s, red);
    if (threadIdx.x == 0) *out = t;
}

// D = finest term + level terms (M-2 .. 0), in that order.
__global__ void k_total(const double* lev, int M, double* D) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double d = 0.0;
        for (int i = 0; i < M; i++) d += lev[i];
        *D = d;
    }
}

}  // namespace mra
