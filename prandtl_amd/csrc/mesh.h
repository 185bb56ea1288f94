// Host-side geometry of unstructured quadrilateral and hexahedral meshes for
// the general-geometry kernels.
#ifndef PRANDTL_MESH_H
#define PRANDTL_MESH_H

#include <string>
#include <vector>

#include "../../include/prandtl_hip.h"

namespace prandtl {
namespace quad {

// Per-face code (fface): bit 0 set when the element is the face's first
// element (lower index: the owner of the normal's orientation), bits 1-3 the
// boundary type (0 for an interior face), bits 4 and up the attribute - 1.
constexpr int FACE_OWN = 1;
inline int face_type(int code) { return (code >> 1) & 7; }
inline int face_attr(int code) { return code >> 4; }

// Local faces f = 2 r + side: r = 0 the xi = 0 / xi = 1 faces, r = 1 the
// eta = 0 / eta = 1 faces (and r = 2 the zeta faces in 3-D); face point l
// runs along the other coordinate (3-D: l = a + NP b over the other two in
// increasing order).
struct Geometry {
    int NE = 0, NP = 0, K = 0, NFP = 0;
    std::vector<double> met;    // [4][NE*K]: adjugate row r, component d at 2r+d
    std::vector<double> jac;    // [NE*K]: det of the Jacobian
    std::vector<int> fface;     // [NE*4]
    std::vector<int> fnb;       // [NE*4*NFP]: neighbour node e2*K + n2 (interior faces)
    std::vector<int> fnf;       // [NE*4]: 2 (e2*4 + f2) + reversed, the neighbour's face (interior faces)
    std::vector<double> fnor;   // [NE*4*NFP*2]: normal outward of the face's first element
    std::vector<double> fxy;    // [NE*4*NFP*2]: coordinates of boundary face points (0 on interior faces)
    std::vector<double> subn;   // [NE][2][NFP][NP-1][2]: subcell interface normals
    double hmin = 0.0;
    std::vector<double> corner; // hexahedral meshes: [NE][8][3] element corners
    // axisymmetric meshes (r = y): r per node; per node -1 off the axis, else
    // 2 d + z with d the node's axis descriptor and z = 1 when the node is in
    // the axis index (its radial-momentum RHS is zeroed); per descriptor the
    // element-local nodes j1, j2 of the ray fit and jn of the density-floor
    // fallback (-1: none) and their r
    std::vector<double> rnode;  // [NE*K]
    std::vector<int> axm;       // [NE*K]
    std::vector<int> axj;       // [nax][3]
    std::vector<double> axr;    // [nax][3]
};

// x, w: LGL nodes and weights on [0,1]; D: derivative matrix, row-major
// NP x NP; axisymmetric: also the axis tables.  Returns false with err set
// for a mesh it cannot use.
bool build(const prandtl_quad_mesh *m, int order, const std::vector<double> &x,
           const std::vector<double> &w, const std::vector<double> &D, bool axisymmetric, Geometry &g,
           std::string &err);

// The same tables for a hexahedral mesh (3-D layouts: met [9][NE*K], six
// faces per element, normals with 3 components).  fnf holds 2 (e2*6 + f2);
// the neighbour's face point follows from fnb's node.
bool build_hex(const prandtl_hex_mesh *m, int order, const std::vector<double> &x,
               const std::vector<double> &w, const std::vector<double> &D, Geometry &g, std::string &err);

} // namespace quad
} // namespace prandtl

#endif
