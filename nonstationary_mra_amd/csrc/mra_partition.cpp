// Partition builder: the program-placed knot grid and the assignment of
// observation / prediction locations to finest regions.
//
// Same semantics as Partition::build_partition_by_program
// (src/class_partition.cpp:459-654): region boxes split at midpoints
// (create_partitions :49-89), knot grid with OFFSET margins (create_knots
// :24-45, grid ceil(sqrt r) x floor(r/ceil(sqrt r)) :468-470), half-open boxes
// (:544), observations that coincide with a knot of any ancestor dropped
// (:547-565), locations kept in input order inside a region.  The reference
// tests every location against every finest box (O(n*F)); here each location
// descends the tree (O(n*M)) and a stable counting sort groups them.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/mra.h"

#include "mra_partition_grid.hpp"

namespace {

using namespace mra_part;

inline bool inside(const Box& b, double lon, double lat) {
    return lon >= b.lonMin && lon < b.lonMax && lat >= b.latMin && lat < b.latMax;
}

// Finest region (local index) containing the point, or -1.
int64_t locate(const Built& b, double lon, double lat) {
    if (!inside(b.box[0], lon, lat)) return -1;
    uint64_t R = 0;
    for (int l = 0; l < b.g.M - 1; l++) {
        uint64_t c0 = R * b.g.J + 1;
        uint64_t next = UINT64_MAX;
        for (int c = 0; c < b.g.J; c++)
            if (inside(b.box[c0 + c], lon, lat)) {
                next = c0 + c;
                break;
            }
        if (next == UINT64_MAX) return -1;
        R = next;
    }
    return (int64_t)(R - b.g.nCoarse);
}

bool on_ancestor_knot(const Built& b, int64_t f, double lon, double lat) {
    uint64_t a = b.g.nCoarse + (uint64_t)f;
    for (int k = 0; k < b.g.M - 1; k++) {
        a = (a - 1) / b.g.J;
        bool lonHit = false, latHit = false;
        for (int i = 0; i < b.g.nLon && !lonHit; i++) lonHit = (b.klon[a * b.g.nLon + i] == lon);
        if (!lonHit) continue;
        for (int j = 0; j < b.g.nLat && !latHit; j++) latHit = (b.klat[a * b.g.nLat + j] == lat);
        if (latHit) return true;
    }
    return false;
}

// Stable counting sort of kept items by finest region.
void group(const std::vector<int64_t>& reg, uint64_t nF, std::vector<uint64_t>& offs, std::vector<uint64_t>& order) {
    offs.assign(nF + 1, 0);
    for (int64_t f : reg)
        if (f >= 0) offs[(uint64_t)f + 1]++;
    for (uint64_t f = 0; f < nF; f++) offs[f + 1] += offs[f];
    order.assign(offs[nF], 0);
    std::vector<uint64_t> pos(offs.begin(), offs.end() - 1);
    for (uint64_t i = 0; i < reg.size(); i++)
        if (reg[i] >= 0) order[pos[(uint64_t)reg[i]]++] = i;
}

void assign_obs(const Built& b, const double* lon, const double* lat, uint64_t n, std::vector<int64_t>& reg) {
    reg.assign(n, -1);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        int64_t f = locate(b, lon[i], lat[i]);
        if (f >= 0 && on_ancestor_knot(b, f, lon[i], lat[i])) f = -1;
        reg[i] = f;
    }
}

void assign_pred(const Built& b, const double* lon, const double* lat, uint64_t n, std::vector<int64_t>& reg) {
    reg.assign(n, -1);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) reg[i] = locate(b, lon[i], lat[i]);
}

}  // namespace

extern "C" int mra_partition_program_sizes(const double* lon, const double* lat, uint64_t n, const double* plon,
                                           const double* plat, uint64_t np, int M, int J, int r, double offset,
                                           const double* bounds, uint64_t* sizes) {
    if (!valid_args(M, J, r, bounds) || !sizes || (n && (!lon || !lat)) || (np && (!plon || !plat))) return MRA_E_ARG;
    Grid g = make_grid(M, J, r);
    Built b = build_boxes(g, bounds, offset);
    std::vector<int64_t> reg;
    assign_obs(b, lon, lat, n, reg);
    uint64_t no = 0, nq = 0;
    for (int64_t f : reg) no += (f >= 0);
    if (np) {
        assign_pred(b, plon, plat, np, reg);
        for (int64_t f : reg) nq += (f >= 0);
    }
    sizes[0] = g.nCoarse * (uint64_t)g.nKnots;
    sizes[1] = g.nFinest;
    sizes[2] = no;
    sizes[3] = nq;
    return MRA_OK;
}

extern "C" int mra_partition_program(const double* lon, const double* lat, const double* val, uint64_t n,
                                     const double* plon, const double* plat, uint64_t np, int M, int J, int r,
                                     double offset, const double* bounds, double* knot_lon, double* knot_lat,
                                     uint64_t* finest_parent, uint64_t* obs_offsets, double* obs_lon, double* obs_lat,
                                     double* obs_val, uint64_t* obs_index, uint64_t* pred_offsets, double* pred_lon,
                                     double* pred_lat, uint64_t* pred_index) {
    if (!valid_args(M, J, r, bounds) || !knot_lon || !knot_lat || !obs_offsets) return MRA_E_ARG;
    if (n && (!lon || !lat || !val || !obs_lon || !obs_lat || !obs_val)) return MRA_E_ARG;
    Grid g = make_grid(M, J, r);
    Built b = build_boxes(g, bounds, offset);
    expand_knots(b, knot_lon, knot_lat);
    if (finest_parent)
        for (uint64_t f = 0; f < g.nFinest; f++) finest_parent[f] = (g.nCoarse + f - 1) / (uint64_t)J;
    std::vector<int64_t> reg;
    std::vector<uint64_t> offs, order;
    assign_obs(b, lon, lat, n, reg);
    group(reg, g.nFinest, offs, order);
    std::memcpy(obs_offsets, offs.data(), sizeof(uint64_t) * (g.nFinest + 1));
    for (uint64_t k = 0; k < order.size(); k++) {
        uint64_t i = order[k];
        obs_lon[k] = lon[i];
        obs_lat[k] = lat[i];
        obs_val[k] = val[i];
        if (obs_index) obs_index[k] = i;
    }
    if (pred_offsets) {
        if (np && (!plon || !plat || !pred_lon || !pred_lat)) return MRA_E_ARG;
        assign_pred(b, plon, plat, np, reg);
        group(reg, g.nFinest, offs, order);
        std::memcpy(pred_offsets, offs.data(), sizeof(uint64_t) * (g.nFinest + 1));
        for (uint64_t k = 0; k < order.size(); k++) {
            uint64_t i = order[k];
            pred_lon[k] = plon[i];
            pred_lat[k] = plat[i];
            if (pred_index) pred_index[k] = i;
        }
    }
    return MRA_OK;
}
