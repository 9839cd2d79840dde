// GPU form of the program partition (Partition::build_partition_by_program,
// src/class_partition.cpp:459-654): one thread per location descends the
// region boxes level by level (half-open boxes, :544) and, for an
// observation, tests the knots of its ancestors (:547-565); a stable radix
// sort by finest region then groups the locations with their input order
// kept inside each region, as the host builder (mra_partition.cpp) does.
// The region boxes and knot axes are built on the host by the same code as
// the host builder, so knots, assignment and order are identical.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mra.h"
#include "mra_partition_grid.hpp"

namespace {

using namespace mra_part;

thread_local std::string g_part_err;

struct DevGrid {
    const double *lo0, *lo1, *la0, *la1;  // region boxes (all levels)
    const double *klon, *klat;            // knot axes per coarse region
    int M, J, nLon, nLat;
    uint64_t nC;
    uint32_t nF;
};

__device__ __forceinline__ bool inside_box(const DevGrid& g, uint64_t R, double lon, double lat) {
    return lon >= g.lo0[R] && lon < g.lo1[R] && lat >= g.la0[R] && lat < g.la1[R];
}

// key[i] = finest region (local index) of location i, or nF when it lies in
// no finest box or (observations) on a knot of one of its ancestors.
__global__ void k_locate(DevGrid g, uint64_t n, const double* __restrict__ lon, const double* __restrict__ lat,
                         int obs, uint32_t* __restrict__ key, uint32_t* __restrict__ idx) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = lon[i], y = lat[i];
    uint32_t f = g.nF;
    if (inside_box(g, 0, x, y)) {
        uint64_t R = 0;
        bool found = true;
        for (int l = 0; l < g.M - 1 && found; l++) {
            const uint64_t c0 = R * g.J + 1;
            found = false;
            for (int c = 0; c < g.J; c++)
                if (inside_box(g, c0 + c, x, y)) {
                    R = c0 + c;
                    found = true;
                    break;
                }
        }
        if (found) {
            f = (uint32_t)(R - g.nC);
            if (obs) {
                uint64_t a = R;
                for (int k = 0; k < g.M - 1; k++) {
                    a = (a - 1) / g.J;
                    bool lonHit = false, latHit = false;
                    for (int q = 0; q < g.nLon && !lonHit; q++) lonHit = g.klon[a * g.nLon + q] == x;
                    if (!lonHit) continue;
                    for (int q = 0; q < g.nLat && !latHit; q++) latHit = g.klat[a * g.nLat + q] == y;
                    if (latHit) {
                        f = g.nF;
                        break;
                    }
                }
            }
        }
    }
    key[i] = f;
    idx[i] = (uint32_t)i;
}

// offsets[f] = first position of region f among the sorted keys (lower bound).
__global__ void k_offsets(const uint32_t* __restrict__ keys, uint64_t n, uint32_t nF, uint64_t* __restrict__ off) {
    const uint64_t f = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (f > nF) return;
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if (keys[mid] < f)
            lo = mid + 1;
        else
            hi = mid;
    }
    off[f] = lo;
}

__global__ void k_gather(uint64_t m, const uint32_t* __restrict__ order, const double* __restrict__ a,
                         const double* __restrict__ b, const double* __restrict__ c, double* __restrict__ oa,
                         double* __restrict__ ob, double* __restrict__ oc, uint64_t* __restrict__ oidx) {
    const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (k >= m) return;
    const uint32_t i = order[k];
    oa[k] = a[i];
    ob[k] = b[i];
    if (c) oc[k] = c[i];
    oidx[k] = i;
}

struct DevBuf {
    std::vector<void*> list;
    ~DevBuf() {
        for (void* p : list) (void)hipFree(p);
    }
    template <typename T>
    T* get(uint64_t count) {
        void* p = nullptr;
        if (hipMalloc(&p, sizeof(T) * (count ? count : 1)) != hipSuccess) return nullptr;
        list.push_back(p);
        return (T*)p;
    }
};

#define GP_TRY(x)                                                                    \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            g_part_err = std::string(#x) + ": " + hipGetErrorString(e_);             \
            return MRA_E_HIP;                                                        \
        }                                                                            \
    } while (0)

template <typename T>
int put(DevBuf& b, T** d, const T* h, uint64_t n) {
    *d = b.get<T>(n);
    if (!*d) {
        g_part_err = "hipMalloc failed";
        return MRA_E_NOMEM;
    }
    if (n) GP_TRY(hipMemcpy(*d, h, sizeof(T) * n, hipMemcpyHostToDevice));
    return MRA_OK;
}

// Locations of one kind grouped by finest region on the device: offsets (nF+1)
// and the kept locations in region order, copied back to the host arrays.
int group_on_device(DevBuf& B, const DevGrid& g, uint64_t n, const double* lon, const double* lat,
                    const double* val, int obs, uint64_t* offsets, double* out_lon, double* out_lat,
                    double* out_val, uint64_t* out_index, uint64_t* kept) {
    double *dl = nullptr, *da = nullptr, *dv = nullptr;
    int rc = put(B, &dl, lon, n);
    if (!rc) rc = put(B, &da, lat, n);
    if (!rc && val) rc = put(B, &dv, val, n);
    if (rc) return rc;
    uint32_t *key = B.get<uint32_t>(n), *idx = B.get<uint32_t>(n), *key2 = B.get<uint32_t>(n),
             *idx2 = B.get<uint32_t>(n);
    uint64_t* off = B.get<uint64_t>((uint64_t)g.nF + 1);
    if (!key || !idx || !key2 || !idx2 || !off) {
        g_part_err = "hipMalloc failed";
        return MRA_E_NOMEM;
    }
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (n) k_locate<<<blocks, 256>>>(g, n, dl, da, obs, key, idx);
    GP_TRY(hipGetLastError());
    int end_bit = 1;
    while (end_bit < 32 && (1ull << end_bit) <= (uint64_t)g.nF) end_bit++;
    size_t tmp = 0;
    if (n) {
        GP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, key, key2, idx, idx2, (int)n, 0, end_bit));
        void* dtmp = B.get<char>(tmp);
        if (!dtmp) {
            g_part_err = "hipMalloc failed";
            return MRA_E_NOMEM;
        }
        GP_TRY(hipcub::DeviceRadixSort::SortPairs(dtmp, tmp, key, key2, idx, idx2, (int)n, 0, end_bit));
    }
    k_offsets<<<(unsigned)((g.nF + 1 + 255) / 256), 256>>>(key2, n, g.nF, off);
    GP_TRY(hipGetLastError());
    GP_TRY(hipMemcpy(offsets, off, sizeof(uint64_t) * ((uint64_t)g.nF + 1), hipMemcpyDeviceToHost));
    const uint64_t m = offsets[g.nF];
    *kept = m;
    if (m) {
        double *ol = B.get<double>(m), *oa = B.get<double>(m), *ov = val ? B.get<double>(m) : nullptr;
        uint64_t* oi = B.get<uint64_t>(m);
        if (!ol || !oa || !oi || (val && !ov)) {
            g_part_err = "hipMalloc failed";
            return MRA_E_NOMEM;
        }
        k_gather<<<(unsigned)((m + 255) / 256), 256>>>(m, idx2, dl, da, dv, ol, oa, ov, oi);
        GP_TRY(hipGetLastError());
        GP_TRY(hipMemcpy(out_lon, ol, sizeof(double) * m, hipMemcpyDeviceToHost));
        GP_TRY(hipMemcpy(out_lat, oa, sizeof(double) * m, hipMemcpyDeviceToHost));
        if (val) GP_TRY(hipMemcpy(out_val, ov, sizeof(double) * m, hipMemcpyDeviceToHost));
        if (out_index) GP_TRY(hipMemcpy(out_index, oi, sizeof(uint64_t) * m, hipMemcpyDeviceToHost));
    }
    return MRA_OK;
}

}  // namespace

extern "C" const char* mra_partition_gpu_last_error(void) { return g_part_err.c_str(); }

extern "C" int mra_partition_program_gpu(const double* lon, const double* lat, const double* val, uint64_t n,
                                         const double* plon, const double* plat, uint64_t np, int M, int J, int r,
                                         double offset, const double* bounds, double* knot_lon, double* knot_lat,
                                         uint64_t* finest_parent, uint64_t* obs_offsets, double* obs_lon,
                                         double* obs_lat, double* obs_val, uint64_t* obs_index,
                                         uint64_t* pred_offsets, double* pred_lon, double* pred_lat,
                                         uint64_t* pred_index, uint64_t* sizes) {
    g_part_err.clear();
    if (!valid_args(M, J, r, bounds) || !knot_lon || !knot_lat || !obs_offsets || !sizes ||
        (n && (!lon || !lat || !val || !obs_lon || !obs_lat || !obs_val)) ||
        (pred_offsets && np && (!plon || !plat || !pred_lon || !pred_lat))) {
        g_part_err = "invalid argument";
        return MRA_E_ARG;
    }
    const Grid g = make_grid(M, J, r);
    if (g.nFinest >= 0xffffffffull || n >= 0x7fffffffull || np >= 0x7fffffffull) {
        g_part_err = "more than 2^31 locations or 2^32 finest regions";
        return MRA_E_UNSUPPORTED;
    }
    const Built b = build_boxes(g, bounds, offset);
    expand_knots(b, knot_lon, knot_lat);
    if (finest_parent)
        for (uint64_t f = 0; f < g.nFinest; f++) finest_parent[f] = (g.nCoarse + f - 1) / (uint64_t)J;
    std::vector<double> bx[4];
    for (auto& v : bx) v.resize(g.nTotal);
    for (uint64_t R = 0; R < g.nTotal; R++) {
        bx[0][R] = b.box[R].lonMin;
        bx[1][R] = b.box[R].lonMax;
        bx[2][R] = b.box[R].latMin;
        bx[3][R] = b.box[R].latMax;
    }
    DevBuf B;
    double* d[6] = {};
    for (int k = 0; k < 4; k++)
        if (int rc = put(B, &d[k], bx[k].data(), g.nTotal)) return rc;
    if (int rc = put(B, &d[4], b.klon.data(), b.klon.size())) return rc;
    if (int rc = put(B, &d[5], b.klat.data(), b.klat.size())) return rc;
    const DevGrid dg{d[0], d[1], d[2], d[3], d[4], d[5], M, J, g.nLon, g.nLat, g.nCoarse, (uint32_t)g.nFinest};
    uint64_t kept = 0, keptP = 0;
    if (int rc = group_on_device(B, dg, n, lon, lat, val, 1, obs_offsets, obs_lon, obs_lat, obs_val, obs_index, &kept))
        return rc;
    if (pred_offsets) {
        if (int rc = group_on_device(B, dg, np, plon, plat, nullptr, 0, pred_offsets, pred_lon, pred_lat, nullptr,
                                     pred_index, &keptP))
            return rc;
    }
    GP_TRY(hipDeviceSynchronize());
    sizes[0] = g.nCoarse * (uint64_t)g.nKnots;
    sizes[1] = g.nFinest;
    sizes[2] = kept;
    sizes[3] = keptP;
    return MRA_OK;
}
