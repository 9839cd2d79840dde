// Region boxes and knot grids of the program partition, shared by the host
// builder (mra_partition.cpp) and the GPU builder (mra_partition_gpu.hip).
//
// Same semantics as Partition::build_partition_by_program
// (src/class_partition.cpp:459-654): region boxes split at midpoints
// (create_partitions :49-89) and the knot grid with OFFSET margins
// (create_knots :24-45, grid ceil(sqrt r) x floor(r/ceil(sqrt r)) :468-470).
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

namespace mra_part {

struct Grid {
    int M, J, nLon, nLat, nKnots;
    uint64_t nCoarse, nFinest, nTotal;
};

inline Grid make_grid(int M, int J, int r) {
    Grid g;
    g.M = M;
    g.J = J;
    g.nLon = (int)std::ceil(std::sqrt((double)r));
    g.nLat = r / g.nLon;
    g.nKnots = g.nLon * g.nLat;
    uint64_t tot = 1, cnt = 1;
    for (int l = 1; l < M - 1; l++) {
        cnt *= (uint64_t)J;
        tot += cnt;
    }
    g.nCoarse = tot;
    g.nFinest = cnt * (uint64_t)J;
    g.nTotal = tot + g.nFinest;
    return g;
}

struct Box {
    double lonMin, lonMax, latMin, latMax;
};

// Children boxes of a box, in the reference's child order.
inline void split(const Box& b, int J, Box* c) {
    if (J == 2) {
        if ((b.lonMax - b.lonMin) >= (b.latMax - b.latMin)) {
            double mid = (b.lonMax + b.lonMin) / 2;
            c[0] = {b.lonMin, mid, b.latMin, b.latMax};
            c[1] = {mid, b.lonMax, b.latMin, b.latMax};
        } else {
            double mid = (b.latMax + b.latMin) / 2;
            c[0] = {b.lonMin, b.lonMax, b.latMin, mid};
            c[1] = {b.lonMin, b.lonMax, mid, b.latMax};
        }
    } else {
        double lm = (b.lonMax + b.lonMin) / 2, am = (b.latMax + b.latMin) / 2;
        c[0] = {b.lonMin, lm, b.latMin, am};
        c[1] = {lm, b.lonMax, b.latMin, am};
        c[2] = {b.lonMin, lm, am, b.latMax};
        c[3] = {lm, b.lonMax, am, b.latMax};
    }
}

// Knot grid of one box: lon values (nLon) and lat values (nLat); knot
// (i, j) sits at index i*nLat + j.
inline void knot_axes(const Box& b, const Grid& g, double offset, double* lonv, double* latv) {
    double offLon = (b.lonMax - b.lonMin) * offset, offLat = (b.latMax - b.latMin) * offset;
    double lonStart = b.lonMin + offLon, lonEnd = b.lonMax - offLon;
    double latStart = b.latMin + offLat, latEnd = b.latMax - offLat;
    double incLon = (g.nLon == 1 ? 0 : (lonEnd - lonStart) / (g.nLon - 1));
    double incLat = (g.nLat == 1 ? 0 : (latEnd - latStart) / (g.nLat - 1));
    for (int i = 0; i < g.nLon; i++) lonv[i] = lonEnd - (g.nLon - i - 1) * incLon;
    for (int j = 0; j < g.nLat; j++) latv[j] = latEnd - (g.nLat - j - 1) * incLat;
}

struct Built {
    Grid g;
    std::vector<Box> box;           // all regions
    std::vector<double> klon, klat; // knot axes per coarse region: nLon / nLat values
};

inline Built build_boxes(const Grid& g, const double* bounds, double offset) {
    Built b;
    b.g = g;
    b.box.resize(g.nTotal);
    b.box[0] = {bounds[0], bounds[1], bounds[2], bounds[3]};
    for (uint64_t R = 0; R < g.nCoarse; R++) split(b.box[R], g.J, &b.box[R * g.J + 1]);
    b.klon.resize(g.nCoarse * g.nLon);
    b.klat.resize(g.nCoarse * g.nLat);
    for (uint64_t R = 0; R < g.nCoarse; R++)
        knot_axes(b.box[R], g, offset, &b.klon[R * g.nLon], &b.klat[R * g.nLat]);
    return b;
}

// Knots of every coarse region, knot q of region R at R*nKnots+q.
inline void expand_knots(const Built& b, double* knot_lon, double* knot_lat) {
    const Grid& g = b.g;
    for (uint64_t R = 0; R < g.nCoarse; R++)
        for (int i = 0; i < g.nLon; i++)
            for (int j = 0; j < g.nLat; j++) {
                knot_lon[R * g.nKnots + i * g.nLat + j] = b.klon[R * g.nLon + i];
                knot_lat[R * g.nKnots + i * g.nLat + j] = b.klat[R * g.nLat + j];
            }
}

inline bool valid_args(int M, int J, int r, const double* bounds) {
    return M >= 2 && M <= 24 && (J == 2 || J == 4) && r >= 1 && bounds != nullptr;
}

}  // namespace mra_part
