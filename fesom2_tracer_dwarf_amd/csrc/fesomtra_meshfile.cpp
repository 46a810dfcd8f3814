// FESOM2 mesh files and the spherical mesh set-up of the reference (host code), the C++ form of
// fesom2_tracer_dwarf_amd/mesh_io.py behind tracer_load_mesh_from_files:
//   files            lib/oce_mesh.F90 read_mesh (nod2d.out, aux3d.out), load_edges (edgenum.out,
//                    edges.out, edge_tri.out), find_levels (nlvls.out, elvls.out)
//   rotation         g2r with the default Euler angles (lib/gen_modules_rotate_grid.F90:33-119,
//                    lib/gen_modules_config.F90:86-88)
//   orientation      test_tri (lib/oce_mesh.F90:1565-1607): clockwise in (lon, lat)
//   areas            mesh_areas (:2082-2195): longitude scaled by cos of the mean latitude
//   auxiliary        mesh_auxiliary_arrays (:2385-2534): element centres, elem_cos, edge_dxdy,
//                    edge_cross_dxdy, gradient_sca
//   MUSCL support    find_up_downwind_triangles (lib/oce_muscl_adv.F90:245-338), nboundary_lay
// Longitude differences are wrapped into [-pi, pi] (the reference's trim_cyclic compares radian
// differences with cyclic_length = 360 and never wraps).  Only single-partition meshes without
// cavities are read.
#include "fesomtra_meshfile.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace fesomtra {
namespace {

constexpr double kR = 6367500.0;  // lib/oce_modules.F90:24
constexpr double kPi = 3.14159265358979323846;

bool read_tokens(const std::string &path, std::vector<std::string> &tok, std::string &err) {
    std::ifstream f(path);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::string t;
    while (f >> t) tok.push_back(t);
    return true;
}
bool to_int(const std::string &s, long &v) {
    char *end = nullptr;
    v = std::strtol(s.c_str(), &end, 10);
    return end && *end == '\0';
}
bool to_dbl(const std::string &s, double &v) {
    char *end = nullptr;
    v = std::strtod(s.c_str(), &end);
    return end && *end == '\0';
}
// n integers from tok starting at `at`
bool ints(const std::vector<std::string> &tok, size_t at, size_t n, std::vector<int> &out, const std::string &what,
          std::string &err) {
    if (tok.size() < at + n) {
        err = what + ": file too short";
        return false;
    }
    out.resize(n);
    for (size_t i = 0; i < n; ++i) {
        long v;
        if (!to_int(tok[at + i], v)) {
            err = what + ": not an integer: " + tok[at + i];
            return false;
        }
        out[i] = (int)v;
    }
    return true;
}
double trim_lon(double d) {
    if (d > kPi) d -= 2.0 * kPi;
    if (d < -kPi) d += 2.0 * kPi;
    return d;
}

}  // namespace

int read_mesh_dir(const std::string &dir, FileMesh &m, std::string &err) {
    const std::string base = dir.empty() || dir.back() == '/' ? dir : dir + "/";
    std::vector<std::string> tok;
    // nodes: N, then index lon lat coast (degrees)
    if (!read_tokens(base + "nod2d.out", tok, err)) return -1;
    long N;
    if (tok.empty() || !to_int(tok[0], N) || N < 3 || tok.size() < 1 + 4 * (size_t)N) {
        err = "nod2d.out: bad header or too short";
        return -1;
    }
    std::vector<double> lon(N), lat(N);
    for (long n = 0; n < N; ++n) {
        if (!to_dbl(tok[1 + 4 * n + 1], lon[n]) || !to_dbl(tok[1 + 4 * n + 2], lat[n])) {
            err = "nod2d.out: bad coordinate at node " + std::to_string(n + 1);
            return -1;
        }
    }
    // elements
    tok.clear();
    if (!read_tokens(base + "elem2d.out", tok, err)) return -1;
    long T;
    if (tok.empty() || !to_int(tok[0], T) || T < 1) {
        err = "elem2d.out: bad header";
        return -1;
    }
    std::vector<int> en;
    if (!ints(tok, 1, 3 * (size_t)T, en, "elem2d.out", err)) return -1;
    // vertical: nl, then zbar
    tok.clear();
    if (!read_tokens(base + "aux3d.out", tok, err)) return -1;
    long nl;
    if (tok.empty() || !to_int(tok[0], nl) || nl < 3 || tok.size() < 1 + (size_t)nl) {
        err = "aux3d.out: bad header or too short";
        return -1;
    }
    m.zbar.resize(nl);
    for (long k = 0; k < nl; ++k)
        if (!to_dbl(tok[1 + k], m.zbar[k])) {
            err = "aux3d.out: bad depth";
            return -1;
        }
    if (m.zbar[1] > 0)
        for (double &z : m.zbar) z = -z;
    // levels, edges
    tok.clear();
    if (!read_tokens(base + "nlvls.out", tok, err) || !ints(tok, 0, N, m.nlev_n, "nlvls.out", err)) return -1;
    tok.clear();
    if (!read_tokens(base + "elvls.out", tok, err) || !ints(tok, 0, T, m.nlev_e, "elvls.out", err)) return -1;
    tok.clear();
    std::vector<int> enum2;
    if (!read_tokens(base + "edgenum.out", tok, err) || !ints(tok, 0, 1, enum2, "edgenum.out", err)) return -1;
    const long E = enum2[0];
    tok.clear();
    if (!read_tokens(base + "edges.out", tok, err) || !ints(tok, 0, 2 * (size_t)E, m.edges, "edges.out", err)) return -1;
    tok.clear();
    if (!read_tokens(base + "edge_tri.out", tok, err) || !ints(tok, 0, 2 * (size_t)E, m.edge_tri, "edge_tri.out", err))
        return -1;
    for (int v : en)
        if (v < 1 || v > N) {
            err = "elem2d.out: node index out of range";
            return -1;
        }
    for (int v : m.edges)
        if (v < 1 || v > N) {
            err = "edges.out: node index out of range";
            return -1;
        }
    for (int &v : m.edge_tri) {
        if (v < 0) v = 0;
        if (v > T) {
            err = "edge_tri.out: element index out of range";
            return -1;
        }
    }
    for (long t = 0; t < T; ++t)
        if (m.nlev_e[t] < 2 || m.nlev_e[t] > nl) {
            err = "elvls.out: level count outside [2, nl]";
            return -1;
        }
    m.nl = (int)nl;
    m.N = (int)N;
    m.T = (int)T;
    m.E = (int)E;

    // rotation to model coordinates (g2r)
    const double al = 50.0 * kPi / 180.0, be = 15.0 * kPi / 180.0, ga = -90.0 * kPi / 180.0;
    const double r[3][3] = {
        {std::cos(ga) * std::cos(al) - std::sin(ga) * std::cos(be) * std::sin(al),
         std::cos(ga) * std::sin(al) + std::sin(ga) * std::cos(be) * std::cos(al), std::sin(ga) * std::sin(be)},
        {-std::sin(ga) * std::cos(al) - std::cos(ga) * std::cos(be) * std::sin(al),
         -std::sin(ga) * std::sin(al) + std::cos(ga) * std::cos(be) * std::cos(al), std::cos(ga) * std::sin(be)},
        {std::sin(be) * std::sin(al), -std::sin(be) * std::cos(al), std::cos(be)}};
    m.coord.resize(2 * (size_t)N);
    for (long n = 0; n < N; ++n) {
        const double gl = lon[n] * kPi / 180.0, gp = lat[n] * kPi / 180.0;
        const double g[3] = {std::cos(gp) * std::cos(gl), std::cos(gp) * std::sin(gl), std::sin(gp)};
        double x[3];
        for (int i = 0; i < 3; ++i) x[i] = r[i][0] * g[0] + r[i][1] * g[1] + r[i][2] * g[2];
        m.coord[2 * n + 1] = std::asin(std::min(1.0, std::max(-1.0, x[2])));
        m.coord[2 * n] = (x[0] == 0.0 && x[1] == 0.0) ? 0.0 : std::atan2(x[1], x[0]);
    }
    auto X = [&](int n0) { return m.coord[2 * (size_t)n0]; };      // 0-based node
    auto Y = [&](int n0) { return m.coord[2 * (size_t)n0 + 1]; };

    // test_tri: clockwise
    m.enodes = en;
    for (long t = 0; t < T; ++t) {
        int *e = &m.enodes[3 * t];
        const double bx = trim_lon(X(e[1] - 1) - X(e[0] - 1)), by = Y(e[1] - 1) - Y(e[0] - 1);
        const double cx = trim_lon(X(e[2] - 1) - X(e[0] - 1)), cy = Y(e[2] - 1) - Y(e[0] - 1);
        if (bx * cy - by * cx > 0) std::swap(e[1], e[2]);
    }
    // nod_in_elem2D, ascending element order; node level minimum over elements
    m.nie_num.assign(N, 0);
    for (long t = 0; t < T; ++t)
        for (int c = 0; c < 3; ++c) m.nie_num[m.enodes[3 * t + c] - 1]++;
    m.max_adj = *std::max_element(m.nie_num.begin(), m.nie_num.end());
    if (m.max_adj < 1) {
        err = "a node belongs to no element";
        return -1;
    }
    m.nie.assign((size_t)m.max_adj * N, 0);
    std::vector<int> fill(N, 0);
    for (long t = 0; t < T; ++t)
        for (int c = 0; c < 3; ++c) {
            const int n0 = m.enodes[3 * t + c] - 1;
            m.nie[(size_t)n0 * m.max_adj + fill[n0]++] = (int)t + 1;
        }
    for (long n = 0; n < N; ++n)
        if (m.nie_num[n] < 1) {
            err = "node " + std::to_string(n + 1) + " belongs to no element";
            return -1;
        }
    m.ulev_n.assign(N, 1);
    m.ulev_e.assign(T, 1);

    // areas (mesh_areas)
    m.elem_area.resize(T);
    for (long t = 0; t < T; ++t) {
        const int *e = &m.enodes[3 * t];
        const double ay = std::cos((Y(e[0] - 1) + Y(e[1] - 1) + Y(e[2] - 1)) / 3.0);
        const double a0 = trim_lon(X(e[1] - 1) - X(e[0] - 1)) * ay, a1 = Y(e[1] - 1) - Y(e[0] - 1);
        const double b0 = trim_lon(X(e[2] - 1) - X(e[0] - 1)) * ay, b1 = Y(e[2] - 1) - Y(e[0] - 1);
        m.elem_area[t] = 0.5 * std::fabs(a0 * b1 - b0 * a1);
    }
    m.area.assign((size_t)nl * N, 0.0);
    for (long n = 0; n < N; ++n)
        for (int k = 0; k < m.nie_num[n]; ++k) {
            const int t = m.nie[(size_t)n * m.max_adj + k] - 1;
            for (int nz = 1; nz <= m.nlev_e[t] - 1; ++nz) m.area[(size_t)n * nl + nz - 1] += m.elem_area[t] / 3.0;
        }
    m.areasvol.assign((size_t)nl * N, 0.0);
    for (long n = 0; n < N; ++n)
        for (int nz = 1; nz <= m.nlev_n[n] - 1; ++nz) m.areasvol[(size_t)n * nl + nz - 1] = m.area[(size_t)n * nl + nz - 1];
    for (double &v : m.elem_area) v = v * kR * kR;
    for (double &v : m.area) v = v * kR * kR;
    for (double &v : m.areasvol) v = v * kR * kR;

    // element centres and cos (elem_center: longitudes unwrapped around the smallest)
    std::vector<double> cx(T), cy(T);
    m.elem_cos.resize(T);
    for (long t = 0; t < T; ++t) {
        const int *e = &m.enodes[3 * t];
        double ax[3] = {X(e[0] - 1), X(e[1] - 1), X(e[2] - 1)};
        const double amin = std::min(ax[0], std::min(ax[1], ax[2]));
        for (double &a : ax) {
            if (a - amin >= kPi) a -= 2.0 * kPi;
            if (a - amin < -kPi) a += 2.0 * kPi;
        }
        cx[t] = (ax[0] + ax[1] + ax[2]) / 3.0;
        cy[t] = (Y(e[0] - 1) + Y(e[1] - 1) + Y(e[2] - 1)) / 3.0;
        m.elem_cos[t] = std::cos(cy[t]);
    }
    // edge vectors, cross distances
    m.edxdy.resize(2 * (size_t)E);
    m.ecd.assign(4 * (size_t)E, 0.0);
    for (long e = 0; e < E; ++e) {
        const int a = m.edges[2 * e] - 1, b = m.edges[2 * e + 1] - 1;
        m.edxdy[2 * e] = trim_lon(X(b) - X(a));
        m.edxdy[2 * e + 1] = Y(b) - Y(a);
        double ax = X(a), bx = X(b);
        const double d = ax - bx;
        if (d > kPi) ax -= 2.0 * kPi;
        if (d < -kPi) bx -= 2.0 * kPi;
        const double mx = 0.5 * (ax + bx), my = 0.5 * (Y(a) + Y(b));
        for (int s = 0; s < 2; ++s) {
            const int t = m.edge_tri[2 * e + s] - 1;
            if (t < 0) continue;
            m.ecd[4 * e + 2 * s] = trim_lon(cx[t] - mx) * m.elem_cos[t] * kR;
            m.ecd[4 * e + 2 * s + 1] = (cy[t] - my) * kR;
        }
    }
    // P1 gradients
    m.gsca.resize(6 * (size_t)T);
    for (long t = 0; t < T; ++t) {
        const int *e = &m.enodes[3 * t];
        const double x31 = trim_lon(X(e[2] - 1) - X(e[0] - 1)) * m.elem_cos[t];
        const double x21 = trim_lon(X(e[1] - 1) - X(e[0] - 1)) * m.elem_cos[t];
        const double y31 = Y(e[2] - 1) - Y(e[0] - 1), y21 = Y(e[1] - 1) - Y(e[0] - 1);
        const double df = -0.5 * kR / m.elem_area[t];
        double *g = &m.gsca[6 * t];
        g[0] = (-y31 + y21) * df;
        g[1] = y31 * df;
        g[2] = -y21 * df;
        g[3] = (x31 - x21) * df;
        g[4] = -x31 * df;
        g[5] = x21 * df;
    }
    // vertical grid (full cells)
    const int L = (int)nl - 1;
    m.hnode.resize((size_t)L * N);
    m.Z.resize((size_t)L * N);
    m.zbar3.resize((size_t)nl * N);
    m.helem.resize((size_t)L * T);
    for (long n = 0; n < N; ++n) {
        for (int k = 0; k < L; ++k) {
            m.hnode[(size_t)n * L + k] = m.zbar[k] - m.zbar[k + 1];
            m.Z[(size_t)n * L + k] = 0.5 * (m.zbar[k] + m.zbar[k + 1]);
        }
        for (int k = 0; k < nl; ++k) m.zbar3[(size_t)n * nl + k] = m.zbar[k];
    }
    for (long t = 0; t < T; ++t)
        for (int k = 0; k < L; ++k) m.helem[(size_t)t * L + k] = m.zbar[k] - m.zbar[k + 1];

    // up- and down-wind triangles: for edge n1 -> n2 with x = X(n2) - X(n1), the element around n1
    // whose corner contains -x, and around n2 the one containing +x; the last match in
    // nod_in_elem2D order wins, 0 if none (edges at a coast)
    m.updn.assign(2 * (size_t)E, 0);
    for (long e = 0; e < E; ++e) {
        const int n1 = m.edges[2 * e] - 1, n2 = m.edges[2 * e + 1] - 1;
        const double xv0 = trim_lon(X(n2) - X(n1)), xv1 = Y(n2) - Y(n1);
        for (int side = 0; side < 2; ++side) {
            const int node = side == 0 ? n1 : n2;
            const double sg = side == 0 ? -1.0 : 1.0;
            const double x0 = sg * xv0, x1 = sg * xv1;
            for (int k = 0; k < m.nie_num[node]; ++k) {
                const int el = m.nie[(size_t)node * m.max_adj + k] - 1;
                const int *c3 = &m.enodes[3 * (size_t)el];
                const int pos = c3[0] - 1 == node ? 0 : (c3[1] - 1 == node ? 1 : 2);
                const int ib = pos == 0 ? 1 : 0, ic = pos == 2 ? 1 : 2;
                const int nb = c3[ib] - 1, nc = c3[ic] - 1;
                const double b0 = trim_lon(X(nb) - X(node)), b1 = Y(nb) - Y(node);
                const double c0 = trim_lon(X(nc) - X(node)), c1 = Y(nc) - Y(node);
                const double cr = c0 * c0 + c1 * c1;
                const double bx = (b0 * c0 + b1 * c1) / cr, by = (-b0 * c1 + b1 * c0) / cr;
                const double xx = (x0 * c0 + x1 * c1) / cr, xy = (-x0 * c1 + x1 * c0) / cr;
                const double ab = std::atan2(by, bx), ax = std::atan2(xy, xx);
                if ((ab > 0 && ax > 0 && ax < ab) || (ab < 0 && ax < 0 && ax > ab) || ab == ax || ax == 0)
                    m.updn[2 * e + side] = el + 1;
            }
        }
    }
    // nboundary_lay: boundary edges force 0, else the shallower of the edge's two elements
    m.nbl.assign(N, (int)nl - 1);
    for (long e = 0; e < E; ++e) {
        const int a = m.edges[2 * e] - 1, b = m.edges[2 * e + 1] - 1;
        const int t1 = m.edge_tri[2 * e] - 1, t2 = m.edge_tri[2 * e + 1] - 1;
        const int lev = (t1 < 0 || t2 < 0) ? 0 : std::min(m.nlev_e[t1], m.nlev_e[t2]) - 1;
        m.nbl[a] = std::min(m.nbl[a], lev);
        m.nbl[b] = std::min(m.nbl[b], lev);
    }
    return 0;
}

}  // namespace fesomtra
