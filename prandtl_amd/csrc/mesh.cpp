// Quadrilateral and hexahedral mesh files (MFEM mesh v1.0, Gmsh 2.2 ASCII),
// uniform refinement of quadrilateral meshes, and the per-node / per-face
// geometry the general-geometry kernels read (with the axis tables of
// axisymmetric meshes).  Host code only.
#include "mesh.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <sstream>
#include <unordered_map>
#include <utility>

struct prandtl_mesh_file {
    int dim = 2;                   // 2: quadrilaterals, 3: hexahedra
    std::vector<double> vert;      // [nv][dim]
    std::vector<int> elem;         // [ne][4] (2-D) or [ne][8] (3-D)
    std::vector<double> ecoords;   // empty, or [ne][4][2] (2-D) / [ne][8][3] (3-D): periodic meshes
    std::vector<int> bdr;          // [nb][2] segments (2-D) or [nb][4] quadrilaterals (3-D)
    std::vector<int> battr;        // [nb]
    std::vector<std::pair<std::string, int>> names;   // boundary physical groups
};

namespace {

thread_local std::string g_mesh_err;

int mesh_fail(const std::string &msg)
{
    g_mesh_err = msg;
    return PRANDTL_ERR_IO;
}

// a + b and a * b rounded as written (no fused multiply-add), so that
// coordinates match a restatement compiled without contraction
double add_nc(double a, double b)
{
    volatile double r = a + b;
    return r;
}
double mul_nc(double a, double b)
{
    volatile double r = a * b;
    return r;
}

long long edge_key(int a, int b)
{
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    return ((long long)lo << 32) | (unsigned int)hi;
}

// corner k (vertex order) of element e
inline void corner(const prandtl_mesh_file &m, int e, int k, double &x, double &y)
{
    if (!m.ecoords.empty()) {
        x = m.ecoords[(size_t)e * 8 + 2 * k];
        y = m.ecoords[(size_t)e * 8 + 2 * k + 1];
    } else {
        const int v = m.elem[(size_t)e * 4 + k];
        x = m.vert[2 * (size_t)v];
        y = m.vert[2 * (size_t)v + 1];
    }
}

// Elements with clockwise corners get the order (v0, v3, v2, v1), as
// Mesh::Finalize's orientation fix does for quadrilaterals.
void orient(prandtl_mesh_file &m)
{
    const int ne = (int)(m.elem.size() / 4);
    for (int e = 0; e < ne; e++) {
        double c[4][2];
        for (int k = 0; k < 4; k++) corner(m, e, k, c[k][0], c[k][1]);
        double a = 0.0;
        for (int k = 0; k < 4; k++) {
            const int k1 = (k + 1) & 3;
            a += c[k][0] * c[k1][1] - c[k1][0] * c[k][1];
        }
        if (a < 0.0) {
            std::swap(m.elem[(size_t)e * 4 + 1], m.elem[(size_t)e * 4 + 3]);
            if (!m.ecoords.empty())
                for (int i = 0; i < 2; i++)
                    std::swap(m.ecoords[(size_t)e * 8 + 2 + i], m.ecoords[(size_t)e * 8 + 6 + i]);
        }
    }
}

std::string next_line(std::istream &in)
{
    std::string s;
    while (std::getline(in, s)) {
        const size_t a = s.find_first_not_of(" \t\r");
        if (a == std::string::npos) continue;
        if (s[a] == '#') continue;
        const size_t b = s.find_last_not_of(" \t\r");
        return s.substr(a, b - a + 1);
    }
    return std::string();
}

int read_gmsh(std::istream &in, prandtl_mesh_file &m)
{
    std::string line = next_line(in);   // the version line after $MeshFormat
    {
        std::istringstream ls(line);
        double ver = 0.0;
        int ftype = -1;
        ls >> ver >> ftype;
        if (ver < 2.0 || ver >= 3.0 || ftype != 0) return mesh_fail("only Gmsh 2.x ASCII files are read");
    }
    std::unordered_map<long long, int> node_of;
    bool have_nodes = false, have_elems = false;
    std::vector<double> v3;                        // node coordinates, 3 per node
    std::vector<int> quads, qattr, lines, lattr, hexes;
    std::vector<std::pair<std::string, int>> names[4];   // physical groups by dimension
    std::vector<std::pair<long long, long long>> periodic;   // (slave, master) node ids
    while (in) {
        line = next_line(in);
        if (line.empty()) break;
        if (line == "$PhysicalNames") {
            const int n = std::atoi(next_line(in).c_str());
            for (int i = 0; i < n; i++) {
                std::istringstream ls(next_line(in));
                int dim = 0, tag = 0;
                std::string name;
                ls >> dim >> tag;
                std::getline(ls, name);
                const size_t a = name.find('"'), b = name.rfind('"');
                if (a != std::string::npos && b > a) name = name.substr(a + 1, b - a - 1);
                if (dim >= 0 && dim <= 3) names[dim].emplace_back(name, tag);
            }
        } else if (line == "$Nodes") {
            const int n = std::atoi(next_line(in).c_str());
            if (n <= 0) return mesh_fail("empty $Nodes");
            v3.resize(3 * (size_t)n);
            for (int i = 0; i < n; i++) {
                std::istringstream ls(next_line(in));
                long long id = 0;
                double x = 0, y = 0, z = 0;
                if (!(ls >> id >> x >> y >> z)) return mesh_fail("bad $Nodes record");
                node_of[id] = i;
                v3[3 * (size_t)i] = x;
                v3[3 * (size_t)i + 1] = y;
                v3[3 * (size_t)i + 2] = z;
            }
            have_nodes = true;
        } else if (line == "$Elements") {
            if (!have_nodes) return mesh_fail("$Elements before $Nodes");
            const int n = std::atoi(next_line(in).c_str());
            for (int i = 0; i < n; i++) {
                std::istringstream ls(next_line(in));
                long long id = 0;
                int type = 0, ntags = 0;
                if (!(ls >> id >> type >> ntags)) return mesh_fail("bad $Elements record");
                std::vector<long long> tags(ntags);
                for (int t = 0; t < ntags; t++) ls >> tags[t];
                const int nn = type == 1 ? 2 : type == 3 ? 4 : type == 5 ? 8 : type == 15 ? 1 : -1;
                if (nn < 0)
                    return mesh_fail("element type " + std::to_string(type) +
                                     ": only quadrilaterals, hexahedra and their boundary lines / quadrilaterals are read");
                int v[8];
                for (int k = 0; k < nn; k++) {
                    long long nid = 0;
                    if (!(ls >> nid)) return mesh_fail("short $Elements record");
                    auto it = node_of.find(nid);
                    if (it == node_of.end()) return mesh_fail("element references an unknown node");
                    v[k] = it->second;
                }
                const int attr = ntags > 0 ? (int)tags[0] : 1;
                if (type == 5)   // Gmsh's 8-node hexahedron has MFEM's vertex order
                    for (int k = 0; k < 8; k++) hexes.push_back(v[k]);
                else if (type == 3) {
                    for (int k = 0; k < 4; k++) quads.push_back(v[k]);
                    qattr.push_back(attr);
                } else if (type == 1) {
                    lines.push_back(v[0]);
                    lines.push_back(v[1]);
                    lattr.push_back(attr);
                }
            }
            have_elems = true;
        } else if (line == "$Periodic") {
            // entity pairs, each with its node correspondences (an "Affine"
            // line may precede the node count)
            const int ne = std::atoi(next_line(in).c_str());
            for (int i = 0; i < ne; i++) {
                next_line(in);   // dimension slave-entity master-entity
                std::string cnt = next_line(in);
                if (cnt.compare(0, 6, "Affine") == 0) cnt = next_line(in);
                const int nn = std::atoi(cnt.c_str());
                for (int k = 0; k < nn; k++) {
                    std::istringstream ls(next_line(in));
                    long long a = 0, b = 0;
                    if (!(ls >> a >> b)) return mesh_fail("bad $Periodic record");
                    periodic.emplace_back(a, b);
                }
            }
        } else if (line[0] == '$' && line.compare(0, 4, "$End") != 0) {
            // skip an unknown section
            const std::string end = "$End" + line.substr(1);
            while (in && next_line(in) != end) {}
        }
    }
    if (!have_elems) return mesh_fail("no $Elements section");
    // periodic meshes: each element keeps its corner coordinates, and slave
    // nodes are identified with their masters (chains followed to the end),
    // as mfem::Mesh does for Gmsh's $Periodic section
    std::vector<double> ecoords;
    if (!periodic.empty()) {
        const int nn = (int)(v3.size() / 3);
        std::vector<int> rep(nn);
        for (int i = 0; i < nn; i++) rep[i] = i;
        for (auto &pr : periodic) {
            auto a = node_of.find(pr.first), b = node_of.find(pr.second);
            if (a == node_of.end() || b == node_of.end()) return mesh_fail("$Periodic references an unknown node");
            if (a->second != b->second) rep[a->second] = b->second;
        }
        auto root = [&](int i) {
            int steps = 0;
            while (rep[i] != i) {
                i = rep[i];
                if (++steps > nn) return -1;   // a cycle
            }
            return i;
        };
        std::vector<int> &el = hexes.empty() ? quads : hexes;
        const int dim = hexes.empty() ? 2 : 3;
        ecoords.resize(el.size() * dim);
        for (size_t i = 0; i < el.size(); i++) {
            for (int d = 0; d < dim; d++) ecoords[i * dim + d] = v3[3 * (size_t)el[i] + d];
            const int r = root(el[i]);
            if (r < 0) return mesh_fail("$Periodic correspondences form a cycle");
            el[i] = r;
        }
        std::vector<int> &bd = hexes.empty() ? lines : quads;
        for (auto &v : bd) {
            const int r = root(v);
            if (r < 0) return mesh_fail("$Periodic correspondences form a cycle");
            v = r;
        }
    }
    if (!hexes.empty()) {
        // 3-D: hexahedra, boundary quadrilaterals, 2-D physical names
        m.dim = 3;
        m.vert.swap(v3);
        m.elem.swap(hexes);
        m.bdr.swap(quads);
        m.battr.swap(qattr);
        m.ecoords.swap(ecoords);
        m.names = names[2];
        return PRANDTL_OK;
    }
    if (quads.empty()) return mesh_fail("no quadrilateral or hexahedral elements");
    m.dim = 2;
    const size_t nv = v3.size() / 3;
    m.vert.resize(2 * nv);
    for (size_t i = 0; i < nv; i++) {
        m.vert[2 * i] = v3[3 * i];
        m.vert[2 * i + 1] = v3[3 * i + 1];
    }
    m.elem.swap(quads);
    m.bdr.swap(lines);
    m.battr.swap(lattr);
    m.ecoords.swap(ecoords);
    m.names = names[1];
    return PRANDTL_OK;
}

int read_mfem(std::istream &in, prandtl_mesh_file &m)
{
    int dim = 0, nv = -1;
    std::string line;
    while (in) {
        line = next_line(in);
        if (line.empty()) break;
        if (line == "dimension") {
            dim = std::atoi(next_line(in).c_str());
            if (dim != 2 && dim != 3) return mesh_fail("only 2-D and 3-D MFEM meshes are read");
            m.dim = dim;
        } else if (line == "elements") {
            // geometry 3: quadrilateral (2-D), 5: hexahedron (3-D, MFEM's vertex order)
            const int n = std::atoi(next_line(in).c_str());
            const int nvx = dim == 3 ? 8 : 4, want = dim == 3 ? 5 : 3;
            for (int i = 0; i < n; i++) {
                std::istringstream ls(next_line(in));
                int attr = 0, geom = 0, v[8];
                ls >> attr >> geom;
                if (geom != want)
                    return mesh_fail(dim == 3 ? "only hexahedral elements (geometry 5) are read in 3-D"
                                              : "only quadrilateral elements (geometry 3) are read");
                for (int k = 0; k < nvx; k++)
                    if (!(ls >> v[k])) return mesh_fail("short element record");
                for (int k = 0; k < nvx; k++) m.elem.push_back(v[k]);
            }
        } else if (line == "boundary") {
            const int n = std::atoi(next_line(in).c_str());
            const int nvx = dim == 3 ? 4 : 2, want = dim == 3 ? 3 : 1;
            for (int i = 0; i < n; i++) {
                std::istringstream ls(next_line(in));
                int attr = 0, geom = 0, v[4];
                if (!(ls >> attr >> geom) || geom != want) return mesh_fail("bad boundary record");
                for (int k = 0; k < nvx; k++)
                    if (!(ls >> v[k])) return mesh_fail("bad boundary record");
                for (int k = 0; k < nvx; k++) m.bdr.push_back(v[k]);
                m.battr.push_back(attr);
            }
        } else if (line == "vertices") {
            nv = std::atoi(next_line(in).c_str());
            const std::string nxt = next_line(in);
            if (nxt == "nodes") {
                // first-order L2 nodes with Gauss-Lobatto points: per element
                // the corners in lexicographic order ((0,0), (1,0), (0,1),
                // (1,1), then the same at z = 1 in 3-D)
                std::string fes = next_line(in), fec = next_line(in), vd = next_line(in), ord = next_line(in);
                const char *want = dim == 3 ? "L2_T1_3D_P1" : "L2_T1_2D_P1";
                if (fec.find(want) == std::string::npos)
                    return mesh_fail(std::string("mesh nodes: only the ") + want + " collection is read");
                const int ordering = std::atoi(ord.substr(ord.find(':') + 1).c_str());
                const int nc = dim == 3 ? 8 : 4;   // corners per element
                const int ne = (int)(m.elem.size() / nc);
                const size_t nval = (size_t)ne * nc * dim;
                std::vector<double> vals;
                vals.reserve(nval);
                double val;
                while (vals.size() < nval && in >> val) vals.push_back(val);
                if (vals.size() < nval) return mesh_fail("short nodes section");
                m.ecoords.assign(nval, 0.0);
                static const int lex_to_vertex[8] = {0, 1, 3, 2, 4, 5, 7, 6};
                for (int e = 0; e < ne; e++)
                    for (int q = 0; q < nc; q++)
                        for (int c = 0; c < dim; c++) {
                            const double x = ordering == 1 ? vals[((size_t)e * nc + q) * dim + c]
                                                           : vals[(size_t)c * ne * nc + (size_t)e * nc + q];
                            m.ecoords[((size_t)e * nc + lex_to_vertex[q]) * dim + c] = x;
                        }
                m.vert.assign((size_t)dim * std::max(nv, 0), 0.0);
                for (int e = 0; e < ne; e++)
                    for (int k = 0; k < nc; k++) {
                        const int v = m.elem[(size_t)e * nc + k];
                        if (v >= 0 && v < nv)
                            for (int c = 0; c < dim; c++)
                                m.vert[(size_t)dim * v + c] = m.ecoords[((size_t)e * nc + k) * dim + c];
                    }
            } else {
                const int vdim = std::atoi(nxt.c_str());
                if (vdim != dim) return mesh_fail("vertex dimension must equal the mesh dimension");
                m.vert.resize((size_t)vdim * nv);
                for (int i = 0; i < nv; i++) {
                    std::istringstream ls(next_line(in));
                    for (int d = 0; d < vdim; d++)
                        if (!(ls >> m.vert[(size_t)vdim * i + d])) return mesh_fail("bad vertex record");
                }
            }
        } else if (line == "mfem_mesh_end") {
            break;
        }
    }
    if (m.elem.empty() || nv <= 0) return mesh_fail("no elements or vertices");
    for (int v : m.elem)
        if (v < 0 || v >= nv) return mesh_fail("element vertex out of range");
    for (int v : m.bdr)
        if (v < 0 || v >= nv) return mesh_fail("boundary vertex out of range");
    return PRANDTL_OK;
}

} // namespace

namespace prandtl {
namespace quad {

namespace {

// Jacobian of the bilinear map at (xi, eta): J(i,j) = sum_k c_k[i] dN_k/dxi_j,
// summed over the corners in order (IsoparametricTransformation::EvalJacobian
// with BiLinear2DFiniteElement's derivatives).
void jacobian(const double c[4][2], double xi, double eta, double J[2][2])
{
    const double ds[4][2] = {{eta - 1.0, xi - 1.0}, {1.0 - eta, -xi}, {eta, xi}, {-eta, 1.0 - xi}};
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            double a = 0.0;
            for (int k = 0; k < 4; k++) a += ds[k][j] * c[k][i];
            J[i][j] = a;
        }
}

// face f's end points in l-increasing order (vertex positions 0..3)
const int FACE_VERT[4][2] = {{0, 3}, {1, 2}, {0, 1}, {3, 2}};
// face f's end points in counterclockwise traversal (the outward normal's side)
const int FACE_CCW[4][2] = {{3, 0}, {1, 2}, {0, 1}, {2, 3}};

inline int line_node2(int NP, int r, int l, int m) { return r == 0 ? l * NP + m : m * NP + l; }

} // namespace

// The axis tables of an axisymmetric mesh.  The state recovery
// (DGSEMOperator.cpp:161-424) picks, for a node on the axis, the two
// nearest off-axis nodes of its element at its z (within 1e-12), else the
// two nearest overall, and for the density-floor fallback the nearest
// off-axis node; these depend on the geometry only and are found here once.
// The axis index holds the nodes with |r| <= 1e-14 of elements with an axis
// boundary face (DGSEMOperator.cpp:446-485); nodes with r <= 0 outside it
// are recovered the same way.
bool axis_tables(const prandtl_quad_mesh *m, int NP, const std::vector<double> &x,
                 const std::vector<char> &axis_el, Geometry &g, std::string &err)
{
    const int NE = g.NE, K = NP * NP;
    const size_t NK = (size_t)NE * K;
    bool axis_bc = false;
    for (int e = 0; e < NE; e++) axis_bc = axis_bc || axis_el[e];
    g.rnode.assign(NK, 0.0);
    std::vector<double> z(NK, 0.0);
    for (int e = 0; e < NE; e++) {
        double c[4][2];
        for (int k = 0; k < 4; k++)
            for (int i = 0; i < 2; i++)
                c[k][i] = m->elem_coords ? m->elem_coords[(size_t)e * 8 + 2 * k + i]
                                         : m->vert[2 * (size_t)m->elem[(size_t)e * 4 + k] + i];
        for (int j = 0; j < NP; j++)
            for (int i = 0; i < NP; i++) {
                // the bilinear map, corners in vertex order (Tr.Transform)
                const double xi = x[i], eta = x[j];
                const double Nk[4] = {(1 - xi) * (1 - eta), xi * (1 - eta), xi * eta, (1 - xi) * eta};
                double zz = 0.0, rr = 0.0;
                for (int k = 0; k < 4; k++) {
                    zz += Nk[k] * c[k][0];
                    rr += Nk[k] * c[k][1];
                }
                const size_t q = (size_t)e * K + j * NP + i;
                z[q] = zz;
                g.rnode[q] = rr;
                if (rr < -1e-14) {
                    err = "element " + std::to_string(e) + " has a node below the axis (r < 0)";
                    return false;
                }
            }
    }
    g.axm.assign(NK, -1);
    g.axj.clear();
    g.axr.clear();
    for (int e = 0; e < NE; e++) {
        const double *re = &g.rnode[(size_t)e * K], *ze = &z[(size_t)e * K];
        for (int ld = 0; ld < K; ld++) {
            const bool in_idx = axis_el[e] && std::fabs(re[ld]) <= 1e-14;
            if (!in_idx && re[ld] > 0.0) continue;
            if (!axis_bc) {
                err = "nodes on the axis (r = 0) need an \"axis\" boundary";
                return false;
            }
            int j1 = -1, j2 = -1, jn = -1;
            double r1 = INFINITY, r2 = INFINITY, rn = -1.0;
            auto take = [&](int j) {
                if (re[j] < r1) { r2 = r1; j2 = j1; r1 = re[j]; j1 = j; }
                else if (re[j] < r2) { r2 = re[j]; j2 = j; }
            };
            for (int j = 0; j < K; j++)
                if (re[j] > 0.0 && std::fabs(ze[j] - ze[ld]) <= 1e-12) take(j);
            if (j1 < 0)
                for (int j = 0; j < K; j++)
                    if (re[j] > 0.0) take(j);
            for (int j = 0; j < K; j++)
                if (re[j] > 0.0 && (jn < 0 || re[j] < rn)) { rn = re[j]; jn = j; }
            if (jn < 0) {
                err = "element " + std::to_string(e) + " has no node off the axis";
                return false;
            }
            const int d = (int)(g.axj.size() / 3);
            g.axj.insert(g.axj.end(), {j1, j2, jn});
            g.axr.insert(g.axr.end(), {j1 >= 0 ? r1 : 0.0, j2 >= 0 ? r2 : 0.0, rn});
            g.axm[(size_t)e * K + ld] = 2 * d + (in_idx ? 1 : 0);
        }
    }
    return true;
}

bool build(const prandtl_quad_mesh *m, int order, const std::vector<double> &x,
           const std::vector<double> &w, const std::vector<double> &D, bool axisymmetric, Geometry &g,
           std::string &err)
{
    const int NE = m->nelem, NP = order + 1, K = NP * NP, NFP = NP, N = order;
    if (NE < 1 || !m->elem || (!m->vert && !m->elem_coords)) { err = "empty mesh"; return false; }
    g.NE = NE; g.NP = NP; g.K = K; g.NFP = NFP;
    auto corners = [&](int e, double c[4][2]) {
        for (int k = 0; k < 4; k++)
            for (int i = 0; i < 2; i++)
                c[k][i] = m->elem_coords ? m->elem_coords[(size_t)e * 8 + 2 * k + i]
                                         : m->vert[2 * (size_t)m->elem[(size_t)e * 4 + k] + i];
    };
    for (size_t i = 0; i < (size_t)NE * 4; i++)
        if (m->elem[i] < 0 || (m->vert && !m->elem_coords && m->elem[i] >= m->nvert)) {
            err = "element vertex out of range";
            return false;
        }
    // per-node metrics
    const size_t NK = (size_t)NE * K;
    g.met.assign(4 * NK, 0.0);
    g.jac.assign(NK, 0.0);
    g.hmin = INFINITY;
    for (int e = 0; e < NE; e++) {
        double c[4][2];
        corners(e, c);
        for (int j = 0; j < NP; j++)
            for (int i = 0; i < NP; i++) {
                double J[2][2];
                jacobian(c, x[i], x[j], J);
                const size_t q = (size_t)e * K + j * NP + i;
                // CalcAdjugate: rows (J11, -J01) and (-J10, J00)
                g.met[0 * NK + q] = J[1][1];
                g.met[1 * NK + q] = -J[0][1];
                g.met[2 * NK + q] = -J[1][0];
                g.met[3 * NK + q] = J[0][0];
                const double det = J[0][0] * J[1][1] - J[1][0] * J[0][1];
                if (!(det > 0.0)) {
                    err = "element " + std::to_string(e) + " has a non-positive Jacobian";
                    return false;
                }
                g.jac[q] = det;
            }
        // GetElementSize(e, 1): the smallest singular value of J at the centre
        double J[2][2];
        jacobian(c, 0.5, 0.5, J);
        const double S = J[0][0] * J[0][0] + J[0][1] * J[0][1] + J[1][0] * J[1][0] + J[1][1] * J[1][1];
        const double det = std::fabs(J[0][0] * J[1][1] - J[1][0] * J[0][1]);
        const double smax = std::sqrt(0.5 * (S + std::sqrt(std::max(S * S - 4.0 * det * det, 0.0))));
        g.hmin = std::min(g.hmin, det / smax);
    }
    // faces: element edges keyed by their vertex pair
    std::unordered_map<long long, std::pair<int, int>> first;   // key -> (e, f)
    first.reserve((size_t)NE * 4);
    std::vector<int> nbr_e((size_t)NE * 4, -1), nbr_f((size_t)NE * 4, -1);
    for (int e = 0; e < NE; e++)
        for (int f = 0; f < 4; f++) {
            const int a = m->elem[(size_t)e * 4 + FACE_VERT[f][0]], b = m->elem[(size_t)e * 4 + FACE_VERT[f][1]];
            if (a == b) { err = "degenerate element edge"; return false; }
            const long long key = edge_key(a, b);
            auto it = first.find(key);
            if (it == first.end()) {
                first.emplace(key, std::make_pair(e, f));
            } else {
                const int e2 = it->second.first, f2 = it->second.second;
                if (nbr_e[(size_t)e2 * 4 + f2] >= 0) { err = "an edge is shared by more than two elements"; return false; }
                nbr_e[(size_t)e2 * 4 + f2] = e; nbr_f[(size_t)e2 * 4 + f2] = f;
                nbr_e[(size_t)e * 4 + f] = e2; nbr_f[(size_t)e * 4 + f] = f2;
            }
        }
    std::unordered_map<long long, int> battr;
    for (int i = 0; i < m->nbdr; i++) battr[edge_key(m->bdr[2 * i], m->bdr[2 * i + 1])] = m->bdr_attr[i];
    g.fface.assign((size_t)NE * 4, 0);
    g.fnb.assign((size_t)NE * 4 * NFP, -1);
    g.fnf.assign((size_t)NE * 4, -1);
    g.fnor.assign((size_t)NE * 4 * NFP * 2, 0.0);
    g.fxy.assign((size_t)NE * 4 * NFP * 2, 0.0);
    std::vector<char> axis_el(NE, 0);   // elements with an axis boundary face
    for (int e = 0; e < NE; e++)
        for (int f = 0; f < 4; f++) {
            const size_t ef = (size_t)e * 4 + f;
            const int e2 = nbr_e[ef];
            int e1 = e, f1 = f;   // the face's first element and its local face
            int code = 0;
            if (e2 >= 0) {
                if (e2 < e) { e1 = e2; f1 = nbr_f[ef]; }
                else code = FACE_OWN;
                const int f2 = nbr_f[ef];
                const int s_me = m->elem[(size_t)e * 4 + FACE_VERT[f][0]];
                const int s_nb = m->elem[(size_t)e2 * 4 + FACE_VERT[f2][0]];
                const bool same = s_me == s_nb;
                const int r2 = f2 >> 1, side2 = f2 & 1;
                g.fnf[ef] = 2 * (e2 * 4 + f2) + (same ? 0 : 1);
                for (int l = 0; l < NFP; l++) {
                    const int l2 = same ? l : N - l;
                    g.fnb[ef * NFP + l] = e2 * K + line_node2(NP, r2, l2, side2 ? N : 0);
                }
            } else {
                const int a = m->elem[(size_t)e * 4 + FACE_VERT[f][0]], b = m->elem[(size_t)e * 4 + FACE_VERT[f][1]];
                auto it = battr.find(edge_key(a, b));
                if (it == battr.end()) {
                    err = "edge of element " + std::to_string(e) + " has no neighbour and no boundary segment";
                    return false;
                }
                const int attr = it->second;
                if (attr < 1 || attr > m->nattr || !m->attr_bc) {
                    err = "boundary attribute " + std::to_string(attr) + " has no boundary condition";
                    return false;
                }
                int t = m->attr_bc[attr - 1];
                if (t < PRANDTL_BC_SLIP || t > PRANDTL_BC_AXIS) {
                    err = "boundary attribute " + std::to_string(attr) + " needs a non-periodic boundary type";
                    return false;
                }
                if (t == PRANDTL_BC_AXIS) {
                    // the symmetry flux (Simulation.cpp:520-524); the axis
                    // nodes are this element's
                    axis_el[e] = 1;
                    t = PRANDTL_BC_SYMMETRY;
                }
                code = FACE_OWN | (t << 1) | ((attr - 1) << 4);
                // the boundary points' coordinates, for outer states that
                // depend on position (the bilinear map of the corners, summed
                // in corner order without contraction)
                double c[4][2];
                corners(e, c);
                const int r = f >> 1, side = f & 1;
                for (int l = 0; l < NFP; l++) {
                    const int nd = line_node2(NP, r, l, side ? N : 0);
                    const double xi = x[nd % NP], eta = x[nd / NP];
                    const double Nk[4] = {(1 - xi) * (1 - eta), xi * (1 - eta), xi * eta, (1 - xi) * eta};
                    for (int a = 0; a < 2; a++) {
                        double xa = 0.0;
                        for (int k = 0; k < 4; k++) xa = add_nc(xa, mul_nc(Nk[k], c[k][a]));
                        g.fxy[(ef * NFP + l) * 2 + a] = xa;
                    }
                }
            }
            g.fface[ef] = code;
            // CalcOrtho of the first element's edge, counterclockwise: (dy, -dx)
            double c[4][2];
            corners(e1, c);
            const int P = FACE_CCW[f1][0], Q = FACE_CCW[f1][1];
            const double dx = c[Q][0] - c[P][0], dy = c[Q][1] - c[P][1];
            for (int l = 0; l < NFP; l++) {
                g.fnor[(ef * NFP + l) * 2 + 0] = dy;
                g.fnor[(ef * NFP + l) * 2 + 1] = -dx;
            }
        }
    // subcell interface normals n_{i+1/2} = Ja^r(line start) + sum_{l<=i} w_l sum_q D_lq Ja^r_q
    g.subn.assign((size_t)NE * 2 * NFP * N * 2, 0.0);
    for (int e = 0; e < NE; e++)
        for (int r = 0; r < 2; r++)
            for (int l = 0; l < NFP; l++) {
                const int n0 = line_node2(NP, r, l, 0);
                for (int i = 1; i < NP; i++) {
                    double nor[2];
                    for (int d = 0; d < 2; d++) nor[d] = g.met[(2 * r + d) * NK + (size_t)e * K + n0];
                    for (int ll = 0; ll < i; ll++) {
                        double tmp[2] = {0.0, 0.0};
                        for (int q = 0; q < NP; q++) {
                            const int nq = line_node2(NP, r, l, q);
                            for (int d = 0; d < 2; d++) tmp[d] += g.met[(2 * r + d) * NK + (size_t)e * K + nq] * D[ll * NP + q];
                        }
                        for (int d = 0; d < 2; d++) { tmp[d] *= w[ll]; nor[d] += tmp[d]; }
                    }
                    for (int d = 0; d < 2; d++)
                        g.subn[((((size_t)e * 2 + r) * NFP + l) * N + (i - 1)) * 2 + d] = nor[d];
                }
            }
    if (axisymmetric && !axis_tables(m, NP, x, axis_el, g, err)) return false;
    return true;
}

} // namespace quad
} // namespace prandtl

extern "C" {

int prandtl_mesh_read(const char *path, prandtl_mesh_file **out)
{
    if (!path || !out) {
        g_mesh_err = "null argument";
        return PRANDTL_ERR_ARG;
    }
    *out = nullptr;
    std::ifstream in(path);
    if (!in) return mesh_fail(std::string("cannot open ") + path);
    std::string head;
    std::getline(in, head);
    auto *m = new (std::nothrow) prandtl_mesh_file();
    if (!m) {
        g_mesh_err = "host allocation";
        return PRANDTL_ERR_NOMEM;
    }
    int rc;
    if (head.compare(0, 11, "$MeshFormat") == 0) rc = read_gmsh(in, *m);
    else if (head.compare(0, 9, "MFEM mesh") == 0) rc = read_mfem(in, *m);
    else rc = mesh_fail("unknown mesh format (MFEM mesh v1.0 or Gmsh 2.2 ASCII)");
    if (rc != PRANDTL_OK) {
        delete m;
        return rc;
    }
    if (m->dim == 2) orient(*m);
    *out = m;
    return PRANDTL_OK;
}

int prandtl_mesh_dim(const prandtl_mesh_file *m) { return m ? m->dim : 0; }

int prandtl_mesh_view_hex(const prandtl_mesh_file *m, prandtl_hex_mesh *v)
{
    if (!m || !v) return PRANDTL_ERR_ARG;
    if (m->dim != 3) {
        g_mesh_err = "a quadrilateral mesh: use prandtl_mesh_view";
        return PRANDTL_ERR_ARG;
    }
    v->nvert = (int)(m->vert.size() / 3);
    v->vert = m->vert.data();
    v->nelem = (int)(m->elem.size() / 8);
    v->elem = m->elem.data();
    v->elem_coords = m->ecoords.empty() ? nullptr : m->ecoords.data();
    v->nbdr = (int)m->battr.size();
    v->bdr = m->bdr.data();
    v->bdr_attr = m->battr.data();
    return PRANDTL_OK;
}

// One uniform refinement of a hexahedral mesh: every element splits into
// eight (children of element i at 8i..8i+7, child (a, b, c) at 8i + a + 2b +
// 4c occupying [a/2, (a+1)/2] x [b/2, (b+1)/2] x [c/2, (c+1)/2] of the
// parent's reference cube, corners in MFEM's vertex order), with new
// vertices at edge midpoints, face centres and the element centre (the
// trilinear map there: the mean of the edge's, face's or element's
// corners); boundary quadrilaterals split into four.
static int refine_hex(prandtl_mesh_file *m)
{
    const int ne = (int)(m->elem.size() / 8);
    int nv = (int)(m->vert.size() / 3);
    std::vector<double> vert = m->vert;
    std::map<std::vector<int>, int> made;   // sorted parent vertex set -> new vertex
    auto vertex_of = [&](std::vector<int> key, const double *x) {
        std::sort(key.begin(), key.end());
        auto it = made.find(key);
        if (it != made.end()) return it->second;
        for (int d = 0; d < 3; d++) vert.push_back(x[d]);
        made.emplace(key, nv);
        return nv++;
    };
    static const int C[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    auto trilinear = [&](const double c[8][3], const double xi[3], double *x) {
        for (int d = 0; d < 3; d++) x[d] = 0.0;
        for (int k = 0; k < 8; k++) {
            const double w = (C[k][0] ? xi[0] : 1 - xi[0]) * (C[k][1] ? xi[1] : 1 - xi[1]) * (C[k][2] ? xi[2] : 1 - xi[2]);
            for (int d = 0; d < 3; d++) x[d] += w * c[k][d];
        }
    };
    const bool per = !m->ecoords.empty();
    std::vector<int> elem((size_t)ne * 64);
    std::vector<double> ecoords(per ? (size_t)ne * 192 : 0);
    for (int e = 0; e < ne; e++) {
        const int *v = &m->elem[(size_t)e * 8];
        double c[8][3];
        for (int k = 0; k < 8; k++)
            for (int d = 0; d < 3; d++)
                c[k][d] = per ? m->ecoords[((size_t)e * 8 + k) * 3 + d] : m->vert[3 * (size_t)v[k] + d];
        // the vertex at reference point h/2 (h in {0,1,2}^3): the parent corners
        // whose coordinate matches h wherever h is 0 or 2
        int centre = -1;
        auto at = [&](int hx, int hy, int hz) {
            const int h[3] = {hx, hy, hz};
            std::vector<int> key;
            for (int k = 0; k < 8; k++) {
                bool in = true;
                for (int d = 0; d < 3; d++)
                    if (h[d] != 1 && C[k][d] * 2 != h[d]) in = false;
                if (in) key.push_back(v[k]);
            }
            if (key.size() == 1) return key[0];
            const double xi[3] = {0.5 * hx, 0.5 * hy, 0.5 * hz};
            double x[3];
            trilinear(c, xi, x);
            if (key.size() == 8) {   // the element centre is this element's own
                if (centre < 0) {
                    for (int d = 0; d < 3; d++) vert.push_back(x[d]);
                    centre = nv++;
                }
                return centre;
            }
            return vertex_of(key, x);
        };
        for (int cz = 0; cz < 2; cz++)
            for (int cy = 0; cy < 2; cy++)
                for (int cx = 0; cx < 2; cx++) {
                    const size_t ch = (size_t)e * 8 + cx + 2 * cy + 4 * cz;
                    for (int k = 0; k < 8; k++) {
                        const int hx = cx + C[k][0], hy = cy + C[k][1], hz = cz + C[k][2];
                        elem[ch * 8 + k] = at(hx, hy, hz);
                        if (per) {
                            const double xi[3] = {0.5 * hx, 0.5 * hy, 0.5 * hz};
                            trilinear(c, xi, &ecoords[(ch * 8 + k) * 3]);
                        }
                    }
                }
    }
    // boundary quadrilaterals (a, b, c, d in cyclic order): four children
    std::vector<int> bdr, battr;
    const int nb = (int)m->battr.size();
    for (int i = 0; i < nb; i++) {
        const int *q = &m->bdr[(size_t)i * 4];
        auto find = [&](std::vector<int> key) {
            std::sort(key.begin(), key.end());
            auto it = made.find(key);
            return it == made.end() ? -1 : it->second;
        };
        const int e01 = find({q[0], q[1]}), e12 = find({q[1], q[2]}), e23 = find({q[2], q[3]});
        const int e30 = find({q[3], q[0]}), fc = find({q[0], q[1], q[2], q[3]});
        if (e01 < 0 || e12 < 0 || e23 < 0 || e30 < 0 || fc < 0) {
            // a quadrilateral that bounds no element stays whole
            bdr.insert(bdr.end(), q, q + 4);
            battr.push_back(m->battr[i]);
            continue;
        }
        const int kids[4][4] = {{q[0], e01, fc, e30}, {e01, q[1], e12, fc}, {fc, e12, q[2], e23}, {e30, fc, e23, q[3]}};
        for (auto &k : kids) {
            bdr.insert(bdr.end(), k, k + 4);
            battr.push_back(m->battr[i]);
        }
    }
    m->vert.swap(vert);
    m->elem.swap(elem);
    m->ecoords.swap(ecoords);
    m->bdr.swap(bdr);
    m->battr.swap(battr);
    return PRANDTL_OK;
}

int prandtl_mesh_refine(prandtl_mesh_file *m)
{
    if (!m) return PRANDTL_ERR_ARG;
    if (m->dim == 3) return refine_hex(m);
    const int ne = (int)(m->elem.size() / 4);
    int nv = (int)(m->vert.size() / 2);
    std::unordered_map<long long, int> mid;
    std::vector<double> vert = m->vert;
    auto midpoint = [&](int a, int b) {
        const long long key = edge_key(a, b);
        auto it = mid.find(key);
        if (it != mid.end()) return it->second;
        vert.push_back(0.5 * (m->vert[2 * (size_t)a] + m->vert[2 * (size_t)b]));
        vert.push_back(0.5 * (m->vert[2 * (size_t)a + 1] + m->vert[2 * (size_t)b + 1]));
        mid.emplace(key, nv);
        return nv++;
    };
    std::vector<int> elem((size_t)ne * 16);
    std::vector<double> ecoords(m->ecoords.empty() ? 0 : (size_t)ne * 32);
    for (int e = 0; e < ne; e++) {
        const int *v = &m->elem[(size_t)e * 4];
        const int m01 = midpoint(v[0], v[1]), m12 = midpoint(v[1], v[2]);
        const int m23 = midpoint(v[2], v[3]), m30 = midpoint(v[3], v[0]);
        const int ctr = nv++;
        double c[4][2];
        corner(*m, e, 0, c[0][0], c[0][1]);
        corner(*m, e, 1, c[1][0], c[1][1]);
        corner(*m, e, 2, c[2][0], c[2][1]);
        corner(*m, e, 3, c[3][0], c[3][1]);
        vert.push_back(0.25 * (((c[0][0] + c[1][0]) + c[2][0]) + c[3][0]));
        vert.push_back(0.25 * (((c[0][1] + c[1][1]) + c[2][1]) + c[3][1]));
        const int child[4][4] = {{v[0], m01, ctr, m30}, {m01, v[1], m12, ctr},
                                 {ctr, m12, v[2], m23}, {m30, ctr, m23, v[3]}};
        // child corners in the parent's (xi, eta)
        const double pc[4][4][2] = {{{0, 0}, {0.5, 0}, {0.5, 0.5}, {0, 0.5}},
                                    {{0.5, 0}, {1, 0}, {1, 0.5}, {0.5, 0.5}},
                                    {{0.5, 0.5}, {1, 0.5}, {1, 1}, {0.5, 1}},
                                    {{0, 0.5}, {0.5, 0.5}, {0.5, 1}, {0, 1}}};
        for (int k = 0; k < 4; k++)
            for (int q = 0; q < 4; q++) {
                elem[((size_t)e * 4 + k) * 4 + q] = child[k][q];
                if (!ecoords.empty()) {
                    const double xi = pc[k][q][0], eta = pc[k][q][1];
                    for (int i = 0; i < 2; i++)
                        ecoords[(((size_t)e * 4 + k) * 4 + q) * 2 + i] =
                            (1 - xi) * (1 - eta) * c[0][i] + xi * (1 - eta) * c[1][i] + xi * eta * c[2][i] +
                            (1 - xi) * eta * c[3][i];
                }
            }
    }
    std::vector<int> bdr, battr;
    const int nb = (int)m->battr.size();
    for (int i = 0; i < nb; i++) {
        const int a = m->bdr[2 * i], b = m->bdr[2 * i + 1];
        auto it = mid.find(edge_key(a, b));
        if (it == mid.end()) {
            // a segment that bounds no element stays whole
            bdr.push_back(a); bdr.push_back(b); battr.push_back(m->battr[i]);
            continue;
        }
        bdr.push_back(a); bdr.push_back(it->second); battr.push_back(m->battr[i]);
        bdr.push_back(it->second); bdr.push_back(b); battr.push_back(m->battr[i]);
    }
    m->vert.swap(vert);
    m->elem.swap(elem);
    m->ecoords.swap(ecoords);
    m->bdr.swap(bdr);
    m->battr.swap(battr);
    return PRANDTL_OK;
}

int prandtl_mesh_view(const prandtl_mesh_file *m, prandtl_quad_mesh *v)
{
    if (!m || !v) return PRANDTL_ERR_ARG;
    if (m->dim != 2) {
        g_mesh_err = "a hexahedral mesh: use prandtl_mesh_view_hex";
        return PRANDTL_ERR_ARG;
    }
    v->nvert = (int)(m->vert.size() / 2);
    v->vert = m->vert.data();
    v->nelem = (int)(m->elem.size() / 4);
    v->elem = m->elem.data();
    v->elem_coords = m->ecoords.empty() ? nullptr : m->ecoords.data();
    v->nbdr = (int)m->battr.size();
    v->bdr = m->bdr.data();
    v->bdr_attr = m->battr.data();
    return PRANDTL_OK;
}

const char *prandtl_mesh_name(const prandtl_mesh_file *m, int i, int *attr)
{
    if (!m || i < 0 || i >= (int)m->names.size()) return nullptr;
    if (attr) *attr = m->names[i].second;
    return m->names[i].first.c_str();
}

void prandtl_mesh_free(prandtl_mesh_file *m) { delete m; }

const char *prandtl_mesh_last_error(void) { return g_mesh_err.c_str(); }

} // extern "C"

namespace prandtl {
namespace quad {

namespace {

// Eigenvalues of a symmetric 3x3 matrix by cyclic Jacobi rotations (the
// off-diagonal sum falls quadratically); returns the smallest.
double smallest_eig_sym3(double A[3][3])
{
    for (int sweep = 0; sweep < 50; sweep++) {
        double off = 0.0, dia = 0.0;
        for (int i = 0; i < 3; i++) {
            dia += A[i][i] * A[i][i];
            for (int j = i + 1; j < 3; j++) off += A[i][j] * A[i][j];
        }
        if (!(off > 1e-32 * dia)) break;
        for (int p = 0; p < 2; p++)
            for (int q = p + 1; q < 3; q++) {
                if (A[p][q] == 0.0) continue;
                const double tau = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
                const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                for (int k = 0; k < 3; k++) {
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; k++) {
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
            }
    }
    return std::min(A[0][0], std::min(A[1][1], A[2][2]));
}

// reference-cube corner k of MFEM's hexahedron vertex order
const int HEX_C[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
// corners of local face f = 2r + side at face coordinates (a, b) = (0,0),
// (1,0), (0,1), (1,1), a and b the other two directions in increasing order
const int HEX_FACE[6][4] = {{0, 3, 4, 7}, {1, 2, 5, 6}, {0, 1, 4, 5}, {3, 2, 7, 6}, {0, 1, 3, 2}, {4, 5, 7, 6}};

inline int line_node3(int NP, int r, int l, int m)
{
    const int a = l % NP, b = l / NP;
    if (r == 0) return (b * NP + a) * NP + m;
    if (r == 1) return (b * NP + m) * NP + a;
    return (m * NP + b) * NP + a;
}

// Jacobian of the trilinear map at xi (TriLinear3DFiniteElement derivatives,
// summed over the corners in order)
void jacobian3(const double c[8][3], const double xi[3], double J[3][3])
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) J[i][j] = 0.0;
    for (int k = 0; k < 8; k++) {
        double f[3], df[3];
        for (int d = 0; d < 3; d++) {
            f[d] = HEX_C[k][d] ? xi[d] : 1.0 - xi[d];
            df[d] = HEX_C[k][d] ? 1.0 : -1.0;
        }
        const double dN[3] = {df[0] * f[1] * f[2], f[0] * df[1] * f[2], f[0] * f[1] * df[2]};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) J[i][j] += dN[j] * c[k][i];
    }
}

struct Key4 {
    int v[4];
    bool operator<(const Key4 &o) const { return std::lexicographical_compare(v, v + 4, o.v, o.v + 4); }
};

Key4 key4(int a, int b, int c, int d)
{
    Key4 k{{a, b, c, d}};
    std::sort(k.v, k.v + 4);
    return k;
}

} // namespace

bool build_hex(const prandtl_hex_mesh *m, int order, const std::vector<double> &x,
               const std::vector<double> &w, const std::vector<double> &D, Geometry &g, std::string &err)
{
    const int NE = m->nelem, NP = order + 1, K = NP * NP * NP, NFP = NP * NP, N = order;
    if (NE < 1 || !m->elem || (!m->vert && !m->elem_coords)) { err = "empty mesh"; return false; }
    g.NE = NE; g.NP = NP; g.K = K; g.NFP = NFP;
    for (size_t i = 0; i < (size_t)NE * 8; i++)
        if (m->elem[i] < 0 || (!m->elem_coords && m->elem[i] >= m->nvert)) {
            err = "element vertex out of range";
            return false;
        }
    auto corners = [&](int e, double c[8][3]) {
        for (int k = 0; k < 8; k++)
            for (int i = 0; i < 3; i++)
                c[k][i] = m->elem_coords ? m->elem_coords[(size_t)e * 24 + 3 * k + i]
                                         : m->vert[3 * (size_t)m->elem[(size_t)e * 8 + k] + i];
    };
    const size_t NK = (size_t)NE * K;
    g.met.assign(9 * NK, 0.0);
    g.jac.assign(NK, 0.0);
    g.corner.assign((size_t)NE * 24, 0.0);
    g.hmin = INFINITY;
    for (int e = 0; e < NE; e++) {
        double c[8][3];
        corners(e, c);
        for (int k = 0; k < 8; k++)
            for (int d = 0; d < 3; d++) g.corner[(size_t)e * 24 + 3 * k + d] = c[k][d];
        for (int k = 0; k < NP; k++)
            for (int j = 0; j < NP; j++)
                for (int i = 0; i < NP; i++) {
                    const double xi[3] = {x[i], x[j], x[k]};
                    double J[3][3];
                    jacobian3(c, xi, J);
                    // adj(J) (CalcAdjugate): row r is the contravariant metric Ja^r
                    double a[3][3];
                    a[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
                    a[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
                    a[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
                    a[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
                    a[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
                    a[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
                    a[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
                    a[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
                    a[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
                    const double det = J[0][0] * a[0][0] + J[0][1] * a[1][0] + J[0][2] * a[2][0];
                    if (!(det > 0.0)) {
                        err = "element " + std::to_string(e) + " has a non-positive Jacobian";
                        return false;
                    }
                    const size_t q = (size_t)e * K + (k * NP + j) * NP + i;
                    for (int r = 0; r < 3; r++)
                        for (int d = 0; d < 3; d++) g.met[(3 * r + d) * NK + q] = a[r][d];
                    g.jac[q] = det;
                }
        // GetElementSize(e, 1): the smallest singular value of J at the centre
        const double xc[3] = {0.5, 0.5, 0.5};
        double J[3][3], A[3][3];
        jacobian3(c, xc, J);
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                A[a][b] = 0.0;
                for (int k = 0; k < 3; k++) A[a][b] += J[k][a] * J[k][b];
            }
        g.hmin = std::min(g.hmin, std::sqrt(std::max(smallest_eig_sym3(A), 0.0)));
    }
    // faces: element faces keyed by their four vertices
    std::map<Key4, std::pair<int, int>> first;
    std::vector<int> nbr_e((size_t)NE * 6, -1), nbr_f((size_t)NE * 6, -1);
    auto vtx = [&](int e, int k) { return m->elem[(size_t)e * 8 + k]; };
    for (int e = 0; e < NE; e++)
        for (int f = 0; f < 6; f++) {
            const int *fc = HEX_FACE[f];
            const Key4 kk = key4(vtx(e, fc[0]), vtx(e, fc[1]), vtx(e, fc[2]), vtx(e, fc[3]));
            if (kk.v[0] == kk.v[1] || kk.v[1] == kk.v[2] || kk.v[2] == kk.v[3]) {
                err = "degenerate element face";
                return false;
            }
            auto it = first.find(kk);
            if (it == first.end()) {
                first.emplace(kk, std::make_pair(e, f));
            } else {
                const int e2 = it->second.first, f2 = it->second.second;
                if (nbr_e[(size_t)e2 * 6 + f2] >= 0) { err = "a face is shared by more than two elements"; return false; }
                nbr_e[(size_t)e2 * 6 + f2] = e; nbr_f[(size_t)e2 * 6 + f2] = f;
                nbr_e[(size_t)e * 6 + f] = e2; nbr_f[(size_t)e * 6 + f] = f2;
            }
        }
    std::map<Key4, int> battr;
    for (int i = 0; i < m->nbdr; i++) {
        const int *b = &m->bdr[4 * (size_t)i];
        battr[key4(b[0], b[1], b[2], b[3])] = m->bdr_attr[i];
    }
    g.fface.assign((size_t)NE * 6, 0);
    g.fnb.assign((size_t)NE * 6 * NFP, -1);
    g.fnf.assign((size_t)NE * 6, -1);
    g.fnor.assign((size_t)NE * 6 * NFP * 3, 0.0);
    for (int e = 0; e < NE; e++)
        for (int f = 0; f < 6; f++) {
            const size_t ef = (size_t)e * 6 + f;
            const int e2 = nbr_e[ef];
            int e1 = e, f1 = f, code = 0;
            if (e2 >= 0) {
                const int f2 = nbr_f[ef];
                if (e2 < e) { e1 = e2; f1 = f2; }
                else code = FACE_OWN;
                // where this element's face corners (0,0), (1,0), (0,1) sit
                // on the neighbour's face
                int ca[3], cb[3];
                for (int t = 0; t < 3; t++) {
                    const int gv = vtx(e, HEX_FACE[f][t]);
                    int hit = -1;
                    for (int u = 0; u < 4; u++)
                        if (vtx(e2, HEX_FACE[f2][u]) == gv) hit = u;
                    if (hit < 0) { err = "face vertices do not match"; return false; }
                    ca[t] = hit & 1;
                    cb[t] = hit >> 1;
                }
                g.fnf[ef] = 2 * (e2 * 6 + f2);
                const int r2 = f2 >> 1, s2 = f2 & 1;
                for (int l = 0; l < NFP; l++) {
                    const int a = l % NP, b = l / NP;
                    const int a2 = N * ca[0] + (ca[1] - ca[0]) * a + (ca[2] - ca[0]) * b;
                    const int b2 = N * cb[0] + (cb[1] - cb[0]) * a + (cb[2] - cb[0]) * b;
                    g.fnb[ef * NFP + l] = e2 * K + line_node3(NP, r2, a2 + NP * b2, s2 ? N : 0);
                }
            } else {
                const int *fc = HEX_FACE[f];
                auto it = battr.find(key4(vtx(e, fc[0]), vtx(e, fc[1]), vtx(e, fc[2]), vtx(e, fc[3])));
                if (it == battr.end()) {
                    err = "face of element " + std::to_string(e) + " has no neighbour and no boundary quadrilateral";
                    return false;
                }
                const int attr = it->second;
                if (attr < 1 || attr > m->nattr || !m->attr_bc) {
                    err = "boundary attribute " + std::to_string(attr) + " has no boundary condition";
                    return false;
                }
                const int t = m->attr_bc[attr - 1];
                if (t < PRANDTL_BC_SLIP || t > PRANDTL_BC_NO_SLIP_ADIABATIC) {
                    err = "boundary attribute " + std::to_string(attr) + " needs a non-periodic boundary type";
                    return false;
                }
                code = FACE_OWN | (t << 1) | ((attr - 1) << 4);
            }
            g.fface[ef] = code;
        }
    // normals: CalcOrtho of the face Jacobian, the first element's adjugate
    // row at the face point with the sign of its side; the second element's
    // points take the first element's value at the matching point
    for (int e = 0; e < NE; e++)
        for (int f = 0; f < 6; f++) {
            const size_t ef = (size_t)e * 6 + f;
            const bool own = g.fface[ef] & FACE_OWN;
            for (int l = 0; l < NFP; l++) {
                int e1 = e, f1 = f, n1 = line_node3(NP, f >> 1, l, (f & 1) ? N : 0);
                if (!own) {
                    const int gn = g.fnb[ef * NFP + l];
                    e1 = gn / K;
                    n1 = gn - e1 * K;
                    f1 = (g.fnf[ef] >> 1) % 6;
                }
                const double sg = (f1 & 1) ? 1.0 : -1.0;
                const int r1 = f1 >> 1;
                for (int d = 0; d < 3; d++)
                    g.fnor[(ef * NFP + l) * 3 + d] = sg * g.met[(3 * r1 + d) * NK + (size_t)e1 * K + n1];
            }
        }
    // subcell interface normals n_{i+1/2} = Ja^r(line start) + sum_{l<=i} w_l sum_q D_lq Ja^r_q
    g.subn.assign((size_t)NE * 3 * NFP * N * 3, 0.0);
    for (int e = 0; e < NE; e++)
        for (int r = 0; r < 3; r++)
            for (int l = 0; l < NFP; l++) {
                const int n0 = line_node3(NP, r, l, 0);
                for (int i = 1; i < NP; i++) {
                    double nor[3];
                    for (int d = 0; d < 3; d++) nor[d] = g.met[(3 * r + d) * NK + (size_t)e * K + n0];
                    for (int ll = 0; ll < i; ll++) {
                        double tmp[3] = {0.0, 0.0, 0.0};
                        for (int q = 0; q < NP; q++) {
                            const int nq = line_node3(NP, r, l, q);
                            for (int d = 0; d < 3; d++) tmp[d] += g.met[(3 * r + d) * NK + (size_t)e * K + nq] * D[ll * NP + q];
                        }
                        for (int d = 0; d < 3; d++) { tmp[d] *= w[ll]; nor[d] += tmp[d]; }
                    }
                    for (int d = 0; d < 3; d++)
                        g.subn[((((size_t)e * 3 + r) * NFP + l) * N + (i - 1)) * 3 + d] = nor[d];
                }
            }
    return true;
}

} // namespace quad
} // namespace prandtl
