// Host-side scene data: mesh, surface topology, elastic constraints, mass.
// Reference counterparts: MeshData (src/mesh/mesh_container.h:26-596),
// Mesh2Constraint and the constraint constructors (src/constraint/*.h).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace pa {

struct SceneMesh {
    int n_all = 0, n_dyn = 0;
    std::vector<double> x, density;
    std::vector<int32_t> tris; // 3 per face
    std::vector<double> tri_k;
    std::vector<double> bend_k; // per vertex
    std::vector<int32_t> tets; // 4 per tet
    std::vector<double> tet_k;
    std::vector<int32_t> tet_boundary; // 3 per facet
    std::vector<int32_t> pins;
};

struct SceneTopology {
    std::vector<int32_t> surf_verts;     // sorted unique vertices of surface tris
    std::vector<int32_t> surf_tris;      // tri faces then tet boundary facets
    std::vector<int32_t> surf_edges;     // 2 per edge (min, max)
    std::vector<int32_t> adj_ptr, adj;   // sorted unique tri-face neighbours per vertex
};

struct SceneConstraints {
    int tri_row0 = 0, tet_row0 = 0, bend_row0 = 0, pin_row0 = 0, m_rows = 0;
    std::vector<int32_t> tri_inds;
    std::vector<double> tri_binv, tri_w, tri_area;
    std::vector<int32_t> tet_inds;
    std::vector<double> tet_binv, tet_w, tet_vol;
    std::vector<int32_t> bend_ptr{0}, bend_inds;
    std::vector<double> bend_coef, bend_w;
    std::vector<int32_t> pin_inds;
    std::vector<double> pin_pos, pin_w;
};

struct Scene {
    SceneMesh mesh;
    SceneTopology topo;
    SceneConstraints cons;
    std::vector<double> mass;
};

// Throws std::runtime_error on degenerate input.
void scene_finalize(Scene& s);
Scene scene_generate(int kind, int side, uint64_t seed, double radius, double drop);

// Several scenes as one block-diagonal scene: the dynamic vertices of all
// parts first (part order), then their static vertices; elements keep their
// part order.  The result is finalized like any scene.  seg[v] = part of
// merged vertex v; part p owns dynamic vertices [dyn_off[p], dyn_off[p+1])
// and static vertices [stat_off[p], stat_off[p+1]).
Scene scene_merge(const std::vector<const Scene*>& parts, std::vector<int32_t>& seg, std::vector<int32_t>& dyn_off,
                  std::vector<int32_t>& stat_off);

} // namespace pa
