// gStar4D for MI355X — point and mesh files (host only, no device work).
//
// Restates the reference's file formats:
//   .asc point reader   GD/Main/Application.cpp:475-566 (readPoints): whitespace-separated
//                       tokens, each read as a float, three per point; exact duplicates
//                       (lexicographic Point3 order, GD/Common/Geometry.cu:49-68) are dropped,
//                       keeping the first occurrence; min/max over every coordinate.
//   grid scaling        GD/Main/Application.cpp:568-585 (_scalePoint): translate by the float
//                       minimum, multiply by (G - 3), divide by the float range, add 1.
//   point dump          GD/Main/Application.cpp:231-243 ("x y z" per line).
//   PLY writer          GD/Common/Geometry.cu:395-466 (TetraMesh::writeToFile): ASCII PLY with
//                       float vertex properties and three faces per tetrahedron.
// Numbers are written as C++ streams write a double by default (%g, 6 significant digits).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "gstar4d.h"

namespace {

struct P3 {
    double p[3];
    bool operator<(const P3 &o) const
    {
        if (p[0] < o.p[0]) return true;
        if (p[0] > o.p[0]) return false;
        if (p[1] < o.p[1]) return true;
        if (p[1] > o.p[1]) return false;
        return p[2] < o.p[2];
    }
};

// Application::_scalePoint: double arithmetic except the range, which is a float difference.
double scale_point(double grid, float min_val, float max_val, double v)
{
    v = v - min_val;
    const float range = max_val - min_val;
    v = (grid - 3.0f) * v / range;
    v += 1.0f;
    return v;
}

} // namespace

extern "C" int g4d_read_asc(const char *path, int32_t grid_size, g4d_point_file *out)
{
    if (!path || !out) return G4D_ERR_INVALID_INPUT;
    memset(out, 0, sizeof(*out));
    if (grid_size < 4) return G4D_ERR_INVALID_INPUT;
    std::ifstream in(path);
    if (!in) return G4D_ERR_IO;
    std::vector<P3> pts;
    std::set<P3> seen;
    std::string tok;
    P3 cur{{0, 0, 0}};
    int idx = 0;
    int64_t read = 0;
    float val = 0.0f, vmin = 999.0f, vmax = -999.0f; // the reference's initial bounds
    while (in >> tok) {
        std::istringstream iss(tok);
        iss >> val; // a token that does not start with a number reads as 0
        cur.p[idx++] = val;
        if (val < vmin) vmin = val;
        if (val > vmax) vmax = val;
        if (idx == 3) {
            idx = 0;
            read++;
            if (seen.insert(cur).second) pts.push_back(cur);
        }
    }
    const int64_t n = (int64_t)pts.size();
    if (n > INT32_MAX) return G4D_ERR_INVALID_INPUT;
    out->num_read = read;
    out->num_duplicates = read - n;
    out->min_val = vmin;
    out->max_val = vmax;
    out->num_points = (int32_t)n;
    out->points = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    out->scaled = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    if (!out->points || !out->scaled) {
        g4d_free_point_file(out);
        return G4D_ERR_OOM;
    }
    std::set<P3> scaled_seen;
    for (int64_t i = 0; i < n; i++) {
        P3 s;
        for (int c = 0; c < 3; c++) {
            out->points[3 * i + c] = pts[i].p[c];
            s.p[c] = scale_point((double)grid_size, vmin, vmax, pts[i].p[c]);
            out->scaled[3 * i + c] = s.p[c];
        }
        scaled_seen.insert(s);
    }
    // The reference asserts that scaling creates no duplicates; here it is an error.
    if ((int64_t)scaled_seen.size() != n) return G4D_ERR_DUPLICATE_POINTS;
    return G4D_OK;
}

extern "C" void g4d_free_point_file(g4d_point_file *f)
{
    if (!f) return;
    free(f->points);
    free(f->scaled);
    f->points = nullptr;
    f->scaled = nullptr;
    f->num_points = 0;
}

extern "C" int g4d_write_asc(const char *path, const double *xyz, int32_t n)
{
    if (!path || (n > 0 && !xyz) || n < 0) return G4D_ERR_INVALID_INPUT;
    FILE *f = fopen(path, "w");
    if (!f) return G4D_ERR_IO;
    for (int32_t i = 0; i < n; i++)
        fprintf(f, "%g %g %g\n", xyz[3 * (int64_t)i], xyz[3 * (int64_t)i + 1], xyz[3 * (int64_t)i + 2]);
    return fclose(f) == 0 ? G4D_OK : G4D_ERR_IO;
}

extern "C" int g4d_write_ply(const char *path, const double *xyz, int32_t n, const int32_t *tets, int64_t num_tets)
{
    if (!path || n < 0 || num_tets < 0 || (n > 0 && !xyz) || (num_tets > 0 && !tets)) return G4D_ERR_INVALID_INPUT;
    FILE *f = fopen(path, "w");
    if (!f) return G4D_ERR_IO;
    fprintf(f, "ply\nformat ascii 1.0\nelement vertex %d\nproperty float x\nproperty float y\nproperty float z\n"
               "element face %lld\nproperty list uchar int vertex_index\nend_header\n",
            (int)n, (long long)(num_tets * 3));
    for (int32_t i = 0; i < n; i++)
        fprintf(f, "%g %g %g \n", xyz[3 * (int64_t)i], xyz[3 * (int64_t)i + 1], xyz[3 * (int64_t)i + 2]);
    static const int faces[3][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}};
    for (int64_t t = 0; t < num_tets; t++) {
        const int32_t *q = tets + 4 * t;
        for (int k = 0; k < 3; k++) fprintf(f, "3 %d %d %d \n", q[faces[k][0]], q[faces[k][1]], q[faces[k][2]]);
    }
    return fclose(f) == 0 ? G4D_OK : G4D_ERR_IO;
}

extern "C" int g4d_struct_sizes(int64_t *out, int32_t max_out)
{
    const int64_t sz[6] = {(int64_t)sizeof(g4d_options), (int64_t)sizeof(g4d_result), (int64_t)sizeof(g4d_stats),
                           (int64_t)sizeof(g4d_check_report), (int64_t)sizeof(g4d_point_file),
                           (int64_t)sizeof(g4d_batch_result)};
    for (int i = 0; i < 6 && out && i < max_out; i++) out[i] = sz[i];
    return 6;
}
