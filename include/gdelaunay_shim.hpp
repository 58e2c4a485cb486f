// gStar4D for MI355X — header-only C++ shim with the reference's four-call interface
// (GDelaunay/GDelaunay/GDelaunay.h:39-42) on top of the C ABI in gstar4d.h.
//
//   gdelInit(config, points, scaled)   remembers the caller's points (and options)
//   gdelCompute(tPba, tStar, tSplay, tOut)
//                                       triangulates; like PointData::init
//                                       (GD/GDelaunay/GDelData.cu:125-130) it rewrites the
//                                       caller's points (and a non-empty `scaled`) in Morton
//                                       order, and tetra ids refer to that order
//   getTetraFromGpu(mesh)              copies tetrahedra with adjacency into mesh
//   gdelDeInit()                       releases the device context
//
// Differences from the reference, by design: errors throw std::runtime_error instead of
// exit(-1); the grid scaling is internal (the caller passes raw FP64 coordinates and
// `scaled` is only reordered, never read); duplicates are reported, not tolerated.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "gstar4d.h"

namespace gdel {

struct Point3 {
    double _p[3];
};
typedef std::vector<Point3> Point3HVec;

struct Tetrahedron {
    int _v[4];   // orient3d(_v[0], _v[1], _v[2], _v[3]) < 0
    int _opp[4]; // tetrahedron across the face opposite _v[i], or -1
};
typedef std::vector<Tetrahedron> TetraHVec;

// The Config fields the reference's triangulation reads (GD/Common/Config.h:59-79).
struct Config {
    int _gridSize = 256;
    int _facetMax = -1;       // not needed: HBM capacity bounds nothing here
    bool _doSorting = true;   // always true in the reference
    bool _doCheck = false;
    int _minWorkset = 8;
    int _device = 0;
};

struct TetraMesh {
    Point3HVec _pointVec;
    TetraHVec _tetraVec;
    void setPoints(const Point3HVec &p) { _pointVec = p; }
    void setTetra(const TetraHVec &t) { _tetraVec = t; }
};

namespace detail {
struct State {
    g4d_ctx *ctx = nullptr;
    Point3HVec *points = nullptr;
    Point3HVec *scaled = nullptr;
    TetraHVec tets;
};
inline State &state()
{
    static State s;
    return s;
}
inline void check(int rc, g4d_ctx *ctx)
{
    if (rc != G4D_OK) throw std::runtime_error(std::string("gstar4d: ") + (ctx ? g4d_last_error(ctx) : "error"));
}
} // namespace detail

inline void gdelDeInit()
{
    detail::State &s = detail::state();
    if (s.ctx) g4d_destroy(s.ctx);
    s = detail::State();
}

inline void gdelInit(const Config &config, Point3HVec &points, Point3HVec &scaled)
{
    gdelDeInit();
    detail::State &s = detail::state();
    g4d_options o;
    g4d_default_options(&o);
    o.grid_size = config._gridSize;
    o.min_workset = config._minWorkset;
    o.device = config._device;
    o.index_space = G4D_INDEX_MORTON;
    o.compute_adjacency = 1;
    detail::check(g4d_create(&s.ctx, &o), s.ctx);
    s.points = &points;
    s.scaled = &scaled;
}

inline void gdelCompute(double &tPba, double &tStar, double &tSplay, double &tOut)
{
    detail::State &s = detail::state();
    if (!s.ctx || !s.points) throw std::runtime_error("gstar4d: gdelInit not called");
    Point3HVec &pts = *s.points;
    g4d_result r;
    detail::check(g4d_triangulate(s.ctx, &pts[0]._p[0], (int32_t)pts.size(), &r), s.ctx);
    Point3HVec sorted(pts.size());
    for (size_t k = 0; k < pts.size(); k++) sorted[k] = pts[(size_t)r.perm[k]];
    if (s.scaled && !s.scaled->empty()) {
        Point3HVec sc(s.scaled->size());
        for (size_t k = 0; k < sc.size(); k++) sc[k] = (*s.scaled)[(size_t)r.perm[k]];
        s.scaled->swap(sc);
    }
    pts.swap(sorted);
    s.tets.resize((size_t)r.num_tets);
    for (int64_t t = 0; t < r.num_tets; t++)
        for (int i = 0; i < 4; i++) {
            s.tets[(size_t)t]._v[i] = r.tets[4 * t + i];
            s.tets[(size_t)t]._opp[i] = r.opp ? r.opp[4 * t + i] : -1;
        }
    g4d_stats st;
    detail::check(g4d_get_stats(s.ctx, &st), s.ctx);
    tPba = st.t_pba_ms;
    tStar = st.t_pairs_ms + st.t_star_ms;
    tSplay = st.t_splay_ms;
    tOut = st.t_out_ms + st.t_adj_ms;
}

inline void getTetraFromGpu(TetraMesh &mesh)
{
    mesh.setTetra(detail::state().tets);
}

} // namespace gdel
