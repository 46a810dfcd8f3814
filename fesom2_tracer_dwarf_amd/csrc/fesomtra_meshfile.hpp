// FESOM2 mesh directory reader and the geometry of a mesh on the sphere (host code).
#pragma once

#include <string>
#include <vector>

namespace fesomtra {

// Every array the library's mesh setters take, in their layouts (1-based indices, Fortran order).
struct FileMesh {
    int nl = 0, N = 0, T = 0, E = 0, max_adj = 0;
    std::vector<double> coord;                      // (2, N) radians, model (rotated) coordinates
    std::vector<int> enodes, edges, edge_tri;       // (3, T), (2, E), (2, E); 0: no element
    std::vector<int> nie, nie_num;                  // (max_adj, N), (N)
    std::vector<int> nlev_n, ulev_n, nlev_e, ulev_e;
    std::vector<double> elem_area, area, areasvol;  // (T), (nl, N), (nl, N)
    std::vector<double> edxdy, ecd, elem_cos, gsca; // (2, E), (4, E), (T), (6, T)
    std::vector<double> zbar, hnode, helem, Z, zbar3;
    std::vector<int> updn, nbl;                     // (2, E) 0: none, (N)
};

// Reads nod2d.out, elem2d.out, aux3d.out, nlvls.out, elvls.out, edgenum.out, edges.out and
// edge_tri.out from `dir` and derives the geometry; returns 0, or -1 with a message in err.
int read_mesh_dir(const std::string &dir, FileMesh &m, std::string &err);

}  // namespace fesomtra
