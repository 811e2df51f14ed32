// Scene setup on the host: benchmark scene generators, surface topology,
// elastic constraints and lumped mass.
//
// Topology semantics follow MeshData::precompute_geometry_info
// (src/mesh/mesh_container.h:232-335): surface triangles are the tri faces
// followed by the tet boundary facets; surface edges are the unique edges of
// the tri faces then those of the boundary facets, each ordered by
// (max, min) vertex as igl::edges emits them; the bending ring of a vertex
// is its sorted unique neighbour list (igl::adjacency_list, sorted = false).
// Constraint rows: tri (2), tet (3), bend (1), pin (1)
// (src/constraint/mesh_to_constraint.h:29-153, app_initializer.cpp:145-172).
#include <climits>
#include "pa_scene.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <random>
#include <stdexcept>

namespace pa {
namespace {

struct D3 {
    double x, y, z;
};
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator*(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(D3 a, D3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double len(D3 a) { return std::sqrt(dot(a, a)); }
inline D3 unit(D3 a) {
    double q = dot(a, a);
    if (q > 0) {
        double l = std::sqrt(q);
        return {a.x / l, a.y / l, a.z / l};
    }
    return a;
}
inline D3 vtx(const std::vector<double>& x, int i) { return {x[3 * i], x[3 * i + 1], x[3 * i + 2]}; }

// Unique undirected edges of a face list, ordered by (max, min).
std::vector<std::array<int32_t, 2>> face_edges(const int32_t* f, size_t nf) {
    std::vector<uint64_t> keys;
    keys.reserve(nf * 3);
    for (size_t i = 0; i < nf; i++)
        for (int k = 0; k < 3; k++) {
            uint32_t a = (uint32_t)f[3 * i + k], b = (uint32_t)f[3 * i + (k + 1) % 3];
            uint32_t lo = std::min(a, b), hi = std::max(a, b);
            keys.push_back(((uint64_t)hi << 32) | lo);
        }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    std::vector<std::array<int32_t, 2>> e(keys.size());
    for (size_t i = 0; i < keys.size(); i++) e[i] = {(int32_t)(keys[i] & 0xffffffffu), (int32_t)(keys[i] >> 32)};
    return e;
}

void build_topology(Scene& s) {
    auto& M = s.mesh;
    auto& T = s.topo;
    T.surf_tris = M.tris;
    T.surf_tris.insert(T.surf_tris.end(), M.tet_boundary.begin(), M.tet_boundary.end());
    std::vector<uint8_t> on(M.n_all, 0);
    for (int32_t v : T.surf_tris) on[v] = 1;
    T.surf_verts.clear();
    for (int i = 0; i < M.n_all; i++)
        if (on[i]) T.surf_verts.push_back(i);
    T.surf_edges.clear();
    for (auto& e : face_edges(M.tris.data(), M.tris.size() / 3)) T.surf_edges.insert(T.surf_edges.end(), e.begin(), e.end());
    for (auto& e : face_edges(M.tet_boundary.data(), M.tet_boundary.size() / 3))
        T.surf_edges.insert(T.surf_edges.end(), e.begin(), e.end());
    // neighbour lists over tri faces
    std::vector<std::vector<int32_t>> nb(M.n_all);
    for (size_t f = 0; f < M.tris.size() / 3; f++)
        for (int k = 0; k < 3; k++) {
            int32_t a = M.tris[3 * f + k], b = M.tris[3 * f + (k + 1) % 3];
            nb[a].push_back(b);
            nb[b].push_back(a);
        }
    T.adj_ptr.assign(M.n_all + 1, 0);
    T.adj.clear();
    for (int i = 0; i < M.n_all; i++) {
        auto& l = nb[i];
        std::sort(l.begin(), l.end());
        l.erase(std::unique(l.begin(), l.end()), l.end());
        T.adj.insert(T.adj.end(), l.begin(), l.end());
        T.adj_ptr[i + 1] = (int32_t)T.adj.size();
    }
}

// Triangle rest data: Binv = (basis^T [e12 e13])^{-1} with an orthonormal
// basis of the triangle plane (triangle_constraint.h:86-102).
void tri_rest(const std::vector<double>& x, const int32_t* f, double binv[4], double* area) {
    D3 x0 = vtx(x, f[0]), e12 = vtx(x, f[1]) - x0, e13 = vtx(x, f[2]) - x0;
    D3 n1 = unit(e12);
    D3 n2 = unit(e13 - dot(e13, n1) * n1);
    double a = dot(n1, e12), b = dot(n1, e13), c = dot(n2, e12), d = dot(n2, e13);
    double det = a * d - b * c;
    if (!(det > 0)) throw std::runtime_error("degenerate or inverted triangle in rest pose");
    binv[0] = d / det;
    binv[1] = -b / det;
    binv[2] = -c / det;
    binv[3] = a / det;
    *area = det / 2;
}

// Tet rest data: Binv = [e12 e13 e14]^{-1}, volume |det|/6 (tetrahedral_constraint.h:43-54).
void tet_rest(const std::vector<double>& x, const int32_t* t, double binv[9], double* vol) {
    D3 x0 = vtx(x, t[0]);
    D3 c[3] = {vtx(x, t[1]) - x0, vtx(x, t[2]) - x0, vtx(x, t[3]) - x0};
    double B[9] = {c[0].x, c[1].x, c[2].x, c[0].y, c[1].y, c[2].y, c[0].z, c[1].z, c[2].z};
    double det = B[0] * (B[4] * B[8] - B[5] * B[7]) - B[1] * (B[3] * B[8] - B[5] * B[6]) + B[2] * (B[3] * B[7] - B[4] * B[6]);
    if (det == 0) throw std::runtime_error("degenerate tet in rest pose");
    binv[0] = (B[4] * B[8] - B[5] * B[7]) / det;
    binv[1] = (B[2] * B[7] - B[1] * B[8]) / det;
    binv[2] = (B[1] * B[5] - B[2] * B[4]) / det;
    binv[3] = (B[5] * B[6] - B[3] * B[8]) / det;
    binv[4] = (B[0] * B[8] - B[2] * B[6]) / det;
    binv[5] = (B[2] * B[3] - B[0] * B[5]) / det;
    binv[6] = (B[3] * B[7] - B[4] * B[6]) / det;
    binv[7] = (B[1] * B[6] - B[0] * B[7]) / det;
    binv[8] = (B[0] * B[4] - B[1] * B[3]) / det;
    *vol = std::fabs(det / 6);
}

// Bending row over {middle, ring...}: cotangent weights of the angles at the
// middle vertex towards the previous and next ring entries (skipping collinear
// triples), scaled by 1/(2 varea)/100 (bending_constraint.h:24-240).
void bend_row(const std::vector<double>& x, int mid, const int32_t* ring, int nr, std::vector<double>& coef, double* varea) {
    D3 c = vtx(x, mid);
    double r = 0;
    for (int i = 0; i < nr; i++) r += len(vtx(x, ring[i]) - c);
    r /= nr;
    *varea = 3.1415926 * r * r / 3.0;
    auto cot_at_mid = [&](D3 a, D3 b) {
        D3 u = a - c, w = b - c;
        D3 cr = cross(u, w);
        if (dot(cr, cr) < 1e-7) return 0.0; // collinear: term skipped
        return dot(u, w) / len(cr);
    };
    coef.assign(nr + 1, 0.0);
    double sum = 0;
    for (int i = 0; i < nr; i++) {
        D3 v = vtx(x, ring[i]);
        double cw = cot_at_mid(v, vtx(x, ring[(i + 1) % nr]));
        double ccw = cot_at_mid(v, vtx(x, ring[(i + nr - 1) % nr]));
        double wi = (cw + ccw) / (2 * *varea) / 100.0;
        sum += wi;
        coef[i + 1] = -wi;
    }
    coef[0] = sum;
}

} // namespace

void scene_finalize(Scene& s) {
    auto& M = s.mesh;
    if (M.n_dyn < 0 || M.n_dyn > M.n_all) throw std::runtime_error("n_dyn out of range");
    build_topology(s);
    auto& C = s.cons;
    C = SceneConstraints{};
    int row = 0;
    // triangles of dynamic meshes
    C.tri_row0 = row;
    for (size_t f = 0; f < M.tris.size() / 3; f++) {
        const int32_t* fa = &M.tris[3 * f];
        if (fa[0] >= M.n_dyn) continue;
        double bi[4], area;
        tri_rest(M.x, fa, bi, &area);
        C.tri_inds.insert(C.tri_inds.end(), fa, fa + 3);
        C.tri_binv.insert(C.tri_binv.end(), bi, bi + 4);
        C.tri_area.push_back(area);
        C.tri_w.push_back(std::sqrt(M.tri_k[f] * area));
        row += 2;
    }
    C.tet_row0 = row;
    for (size_t t = 0; t < M.tets.size() / 4; t++) {
        double bi[9], vol;
        tet_rest(M.x, &M.tets[4 * t], bi, &vol);
        C.tet_inds.insert(C.tet_inds.end(), &M.tets[4 * t], &M.tets[4 * t] + 4);
        C.tet_binv.insert(C.tet_binv.end(), bi, bi + 9);
        C.tet_vol.push_back(vol);
        C.tet_w.push_back(std::sqrt(M.tet_k[t] * vol));
        row += 3;
    }
    C.bend_row0 = row;
    const auto& T = s.topo;
    for (int v = 0; v < M.n_dyn; v++) {
        int nr = T.adj_ptr[v + 1] - T.adj_ptr[v];
        if (nr < 3) continue;
        std::vector<double> coef;
        double varea;
        bend_row(M.x, v, &T.adj[T.adj_ptr[v]], nr, coef, &varea);
        C.bend_inds.push_back(v);
        C.bend_inds.insert(C.bend_inds.end(), &T.adj[T.adj_ptr[v]], &T.adj[T.adj_ptr[v]] + nr);
        C.bend_coef.insert(C.bend_coef.end(), coef.begin(), coef.end());
        C.bend_ptr.push_back((int32_t)C.bend_inds.size());
        C.bend_w.push_back(std::sqrt(M.bend_k[v] * varea));
        row += 1;
    }
    C.pin_row0 = row;
    for (int32_t p : M.pins) {
        if (p < 0 || p >= M.n_all) throw std::runtime_error("pin index out of range");
        if (p >= M.n_dyn) continue;
        C.pin_inds.push_back(p);
        C.pin_pos.insert(C.pin_pos.end(), &M.x[3 * p], &M.x[3 * p] + 3);
        C.pin_w.push_back(std::sqrt(1e5)); // PinConstraint k = 1e5 (mesh_to_constraint.h:47)
        row += 1;
    }
    C.m_rows = row;
    // lumped mass: area/3 * rho per tri vertex, vol/4 * rho per tet vertex
    s.mass.assign(M.n_all, 0.0);
    for (size_t f = 0; f < M.tris.size() / 3; f++) {
        const int32_t* fa = &M.tris[3 * f];
        D3 a = vtx(M.x, fa[0]);
        double area = 0.5 * len(cross(vtx(M.x, fa[1]) - a, vtx(M.x, fa[2]) - a));
        for (int k = 0; k < 3; k++) s.mass[fa[k]] += area / 3.0 * M.density[fa[k]];
    }
    for (size_t t = 0; t < M.tets.size() / 4; t++) {
        const int32_t* ta = &M.tets[4 * t];
        D3 a = vtx(M.x, ta[0]);
        double vol = std::fabs(dot(vtx(M.x, ta[1]) - a, cross(vtx(M.x, ta[2]) - a, vtx(M.x, ta[3]) - a))) / 6.0;
        for (int k = 0; k < 4; k++) s.mass[ta[k]] += vol / 4.0 * M.density[ta[k]];
    }
    for (int i = 0; i < M.n_dyn; i++)
        if (!(s.mass[i] > 0)) throw std::runtime_error("dynamic vertex with zero mass");
}

// ---------------------------------------------------------------------------
// Generators (y-up; the benchmark text is z-up, so its z=h plane is y=h here).
// ---------------------------------------------------------------------------
namespace {

void add_cloth(SceneMesh& M, int nx, int nz, double spacing, double y, double cx, double cz, double density, double k_stretch,
               double k_bend) {
    // Vertex (i, j) -> index i * nz + j with x = x0 + i h, z = z0 + j h, and
    // quads split along alternating diagonals, as in the reference's grid
    // meshes (resource/flat/grid_*.obj).
    int base = M.n_all;
    double x0 = cx - 0.5 * spacing * (nx - 1), z0 = cz - 0.5 * spacing * (nz - 1);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < nz; j++) {
            M.x.push_back(x0 + spacing * i);
            M.x.push_back(y);
            M.x.push_back(z0 + spacing * j);
            M.density.push_back(density);
            M.bend_k.push_back(k_bend);
        }
    for (int i = 0; i + 1 < nx; i++)
        for (int j = 0; j + 1 < nz; j++) {
            int a = base + i * nz + j, b = a + 1, c = a + nz, d = c + 1;
            if (((i + j) & 1) == 0) M.tris.insert(M.tris.end(), {a, b, c, b, d, c});
            else M.tris.insert(M.tris.end(), {a, b, d, a, d, c});
            M.tri_k.push_back(k_stretch);
            M.tri_k.push_back(k_stretch);
        }
    M.n_all += nx * nz;
}

void add_static_tris(SceneMesh& M, const std::vector<double>& v, const std::vector<int32_t>& f) {
    int base = M.n_all;
    M.x.insert(M.x.end(), v.begin(), v.end());
    for (size_t i = 0; i < v.size() / 3; i++) {
        M.density.push_back(1e12);
        M.bend_k.push_back(0.0);
    }
    for (int32_t q : f) M.tris.push_back(q + base);
    for (size_t i = 0; i < f.size() / 3; i++) M.tri_k.push_back(0.0);
    M.n_all += (int)(v.size() / 3);
}

// Icosphere: icosahedron, `level` rounds of midpoint subdivision projected to the sphere.
void icosphere(int level, double r, D3 c, std::vector<double>& v, std::vector<int32_t>& f) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<D3> p = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                         {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& q : p) q = unit(q);
    std::vector<std::array<int, 3>> F = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                                         {11, 10, 2}, {10, 7, 6}, {7, 1, 8},   {3, 9, 4},  {3, 4, 2},   {3, 2, 6}, {3, 6, 8},
                                         {3, 8, 9},  {4, 9, 5},  {2, 4, 11},  {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
    for (int l = 0; l < level; l++) {
        std::map<std::pair<int, int>, int> mid;
        auto midpoint = [&](int a, int b) {
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            p.push_back(unit(0.5 * (p[a] + p[b])));
            int id = (int)p.size() - 1;
            mid[key] = id;
            return id;
        };
        std::vector<std::array<int, 3>> G;
        for (auto& tri : F) {
            int a = midpoint(tri[0], tri[1]), b = midpoint(tri[1], tri[2]), cc = midpoint(tri[2], tri[0]);
            G.push_back({tri[0], a, cc});
            G.push_back({tri[1], b, a});
            G.push_back({tri[2], cc, b});
            G.push_back({a, b, cc});
        }
        F.swap(G);
    }
    v.clear();
    f.clear();
    for (auto& q : p) {
        D3 w = c + r * q;
        v.insert(v.end(), {w.x, w.y, w.z});
    }
    for (auto& tri : F) f.insert(f.end(), tri.begin(), tri.end());
}

// Closed cylinder side along z: `around` segments, `along` rings.
void cylinder(double r, double length, int around, int along, D3 c, std::vector<double>& v, std::vector<int32_t>& f) {
    v.clear();
    f.clear();
    for (int j = 0; j < along; j++) {
        double z = c.z - 0.5 * length + length * j / (along - 1);
        for (int i = 0; i < around; i++) {
            double a = 2 * M_PI * i / around;
            v.insert(v.end(), {c.x + r * std::cos(a), c.y + r * std::sin(a), z});
        }
    }
    for (int j = 0; j + 1 < along; j++)
        for (int i = 0; i < around; i++) {
            int a = j * around + i, b = j * around + (i + 1) % around, cc = b + around, d = a + around;
            f.insert(f.end(), {a, b, cc, a, cc, d});
        }
}

// Box of n^3 cells split into 5 tets per cell; the split alternates with
// the cell parity so that neighbouring faces match.
void tet_box(int n, double h, std::vector<double>& v, std::vector<int32_t>& tets, std::vector<int32_t>& boundary) {
    int m = n + 1;
    auto id = [&](int i, int j, int k) { return (k * m + j) * m + i; };
    v.clear();
    for (int k = 0; k < m; k++)
        for (int j = 0; j < m; j++)
            for (int i = 0; i < m; i++) v.insert(v.end(), {h * i - 0.5 * h * n, h * j - 0.5 * h * n, h * k - 0.5 * h * n});
    tets.clear();
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                int c[8] = {id(i, j, k),         id(i + 1, j, k),         id(i + 1, j + 1, k),         id(i, j + 1, k),
                            id(i, j, k + 1),     id(i + 1, j, k + 1),     id(i + 1, j + 1, k + 1),     id(i, j + 1, k + 1)};
                static const int A[5][4] = {{0, 1, 3, 4}, {1, 2, 3, 6}, {1, 4, 5, 6}, {3, 4, 6, 7}, {1, 3, 4, 6}};
                static const int B[5][4] = {{1, 2, 0, 5}, {2, 3, 0, 7}, {2, 5, 6, 7}, {0, 5, 7, 4}, {0, 2, 5, 7}};
                const int(*S)[4] = ((i + j + k) & 1) ? B : A;
                for (int t = 0; t < 5; t++) {
                    int q[4] = {c[S[t][0]], c[S[t][1]], c[S[t][2]], c[S[t][3]]};
                    // orient positively
                    D3 p0 = vtx(v, q[0]);
                    double d = dot(vtx(v, q[1]) - p0, cross(vtx(v, q[2]) - p0, vtx(v, q[3]) - p0));
                    if (d < 0) std::swap(q[2], q[3]);
                    tets.insert(tets.end(), q, q + 4);
                }
            }
    // boundary facets: faces used by exactly one tet
    std::map<std::array<int, 3>, std::pair<int, std::array<int, 3>>> faces;
    static const int FT[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
    for (size_t t = 0; t < tets.size() / 4; t++)
        for (int fi = 0; fi < 4; fi++) {
            std::array<int, 3> fa = {tets[4 * t + FT[fi][0]], tets[4 * t + FT[fi][1]], tets[4 * t + FT[fi][2]]};
            std::array<int, 3> key = fa;
            std::sort(key.begin(), key.end());
            auto it = faces.find(key);
            if (it == faces.end()) faces[key] = {1, fa};
            else it->second.first++;
        }
    boundary.clear();
    for (auto& kv : faces)
        if (kv.second.first == 1) boundary.insert(boundary.end(), kv.second.second.begin(), kv.second.second.end());
}

void rotate_about(std::vector<double>& v, size_t begin, size_t end, const double R[9], D3 c) {
    for (size_t i = begin; i < end; i++) {
        D3 p = vtx(v, (int)i) - c;
        v[3 * i] = c.x + R[0] * p.x + R[1] * p.y + R[2] * p.z;
        v[3 * i + 1] = c.y + R[3] * p.x + R[4] * p.y + R[5] * p.z;
        v[3 * i + 2] = c.z + R[6] * p.x + R[7] * p.y + R[8] * p.z;
    }
}

} // namespace

Scene scene_generate(int kind, int side, uint64_t seed, double radius, double drop) {
    Scene s;
    SceneMesh& M = s.mesh;
    if (kind == 1) {
        // 32x32 cloth, 1 cm spacing, 0.1 kg, k 1e3 / 1e-2, at y = 0.5 over an
        // icosphere (level 3, r 0.2 m) at the origin; corners (0,0),(31,0) pinned.
        const int n = 32;
        const double sp = 0.01, area = (sp * (n - 1)) * (sp * (n - 1));
        add_cloth(M, n, n, sp, 0.5, 0.0, 0.0, 0.1 / area, 1e3, 1e-2);
        M.n_dyn = M.n_all;
        M.pins = {0, (n - 1) * n}; // grid corners (0,0) and (31,0)
        std::vector<double> v;
        std::vector<int32_t> f;
        icosphere(3, 0.2, {0, 0, 0}, v, f);
        add_static_tris(M, v, f);
    } else if (kind == 2) {
        // 20^3-cell 5-tet box, 0.2 m edge, rho 1000, k = E/(1+nu) = 1e5/1.4,
        // random rotation (seed) with its centre at y = 0.3 m, over a ground quad at y = 0.
        const int n = 20;
        std::vector<double> v;
        std::vector<int32_t> t, b;
        tet_box(n, 0.2 / n, v, t, b);
        M.x = v;
        M.n_all = (int)(v.size() / 3);
        M.density.assign(M.n_all, 1000.0);
        M.bend_k.assign(M.n_all, 0.0);
        M.tets = t;
        M.tet_k.assign(t.size() / 4, 1e5 / (1.0 + 0.4));
        M.tet_boundary = b;
        std::mt19937_64 rng(seed);
        std::normal_distribution<double> nd(0.0, 1.0);
        double q[4] = {nd(rng), nd(rng), nd(rng), nd(rng)};
        double qn = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        for (double& e : q) e /= qn;
        double w = q[0], x = q[1], y = q[2], z = q[3];
        double R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - w * z),     2 * (x * z + w * y),
                       2 * (x * y + w * z),     1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                       2 * (x * z - w * y),     2 * (y * z + w * x),     1 - 2 * (x * x + y * y)};
        rotate_about(M.x, 0, M.n_all, R, {0, 0, 0});
        double h0 = drop > 0 ? drop : 0.3;
        for (int i = 0; i < M.n_all; i++) M.x[3 * i + 1] += h0;
        M.n_dyn = M.n_all;
        std::vector<double> gv;
        std::vector<int32_t> gf;
        const int g = 2;
        for (int j = 0; j <= g; j++)
            for (int i = 0; i <= g; i++) gv.insert(gv.end(), {-1.0 + 2.0 * i / g, 0.0, -1.0 + 2.0 * j / g});
        for (int j = 0; j < g; j++)
            for (int i = 0; i < g; i++) {
                int a = j * (g + 1) + i, bb = a + 1, c = a + g + 2, d = a + g + 1;
                gf.insert(gf.end(), {a, bb, c, a, c, d});
            }
        add_static_tris(M, gv, gf);
    } else if (kind == 3) {
        // 256x256 cloth at 2 mm (same areal density and stiffness as the 32x32
        // cloth), 0.2 mm above a sphere (r 0.15 m) at the origin, beside a
        // cylinder (r 0.05 m) along z at x = 0.22, y = 0.05.
        const int n = 256;
        const double sp = 0.002;
        add_cloth(M, n, n, sp, 0.15 + 0.0002, 0.08, 0.0, 0.1 / (0.31 * 0.31), 1e3, 1e-2);
        M.n_dyn = M.n_all;
        std::vector<double> v;
        std::vector<int32_t> f;
        icosphere(4, 0.15, {0, 0, 0}, v, f);
        add_static_tris(M, v, f);
        cylinder(0.05, 0.6, 48, 41, {0.22, 0.05, 0.0}, v, f);
        add_static_tris(M, v, f);
    } else if (kind == 4) {
        // One batch member: side x side cloth (1 cm), 0.1 kg per 0.31^2 m^2,
        // `drop` m above a sphere of radius `radius` at the origin.
        int n = side > 0 ? side : 32;
        double r = radius > 0 ? radius : 0.2;
        double h0 = drop > 0 ? drop : 0.4;
        add_cloth(M, n, n, 0.01, r + h0, 0.0, 0.0, 0.1 / (0.31 * 0.31), 1e3, 1e-2);
        M.n_dyn = M.n_all;
        std::vector<double> v;
        std::vector<int32_t> f;
        icosphere(3, r, {0, 0, 0}, v, f);
        add_static_tris(M, v, f);
    } else {
        throw std::runtime_error("unknown scene kind");
    }
    scene_finalize(s);
    return s;
}

Scene scene_merge(const std::vector<const Scene*>& parts, std::vector<int32_t>& seg, std::vector<int32_t>& dyn_off,
                  std::vector<int32_t>& stat_off) {
    if (parts.empty()) throw std::runtime_error("no scenes to merge");
    const size_t P = parts.size();
    dyn_off.assign(P + 1, 0);
    stat_off.assign(P + 1, 0);
    int64_t nd = 0, ns = 0;
    for (size_t p = 0; p < P; p++) {
        nd += parts[p]->mesh.n_dyn;
        ns += parts[p]->mesh.n_all - parts[p]->mesh.n_dyn;
    }
    if (nd + ns > INT32_MAX / 4) throw std::runtime_error("merged scene too large");
    Scene s;
    SceneMesh& M = s.mesh;
    M.n_dyn = (int)nd;
    M.n_all = (int)(nd + ns);
    M.x.resize(3 * (size_t)M.n_all);
    M.density.resize(M.n_all);
    M.bend_k.resize(M.n_all);
    seg.resize(M.n_all);
    int d0 = 0, s0 = (int)nd;
    for (size_t p = 0; p < P; p++) {
        const SceneMesh& Q = parts[p]->mesh;
        dyn_off[p] = d0;
        stat_off[p] = s0;
        auto map = [&](int v) { return v < Q.n_dyn ? d0 + v : s0 + (v - Q.n_dyn); };
        for (int v = 0; v < Q.n_all; v++) {
            const int g = map(v);
            for (int k = 0; k < 3; k++) M.x[3 * (size_t)g + k] = Q.x[3 * (size_t)v + k];
            M.density[g] = Q.density[v];
            M.bend_k[g] = Q.bend_k[v];
            seg[g] = (int32_t)p;
        }
        for (int32_t v : Q.tris) M.tris.push_back(map(v));
        M.tri_k.insert(M.tri_k.end(), Q.tri_k.begin(), Q.tri_k.end());
        for (int32_t v : Q.tets) M.tets.push_back(map(v));
        M.tet_k.insert(M.tet_k.end(), Q.tet_k.begin(), Q.tet_k.end());
        for (int32_t v : Q.tet_boundary) M.tet_boundary.push_back(map(v));
        for (int32_t v : Q.pins) M.pins.push_back(map(v));
        d0 += Q.n_dyn;
        s0 += Q.n_all - Q.n_dyn;
    }
    dyn_off[P] = d0;
    stat_off[P] = s0;
    scene_finalize(s);
    return s;
}

} // namespace pa
