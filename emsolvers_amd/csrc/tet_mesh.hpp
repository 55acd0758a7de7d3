// Structured tetrahedral mesh (Kuhn split of an nx x ny x nz hex grid):
// edge and face numbering, barycentric geometry and point location.  Host
// code of the tetrahedral edge-element solver; no device code.
//
// Each cube (i,j,k) is split into 6 tetrahedra, one per permutation p of the
// axes, with vertices v0 = 0, v1 = e_p0, v2 = e_p0 + e_p1, v3 = (1,1,1)
// (offsets from the cube's corner).  Along this path the global vertex index
// increases, so a local edge (a < b) points the same way as its global edge
// (lower to higher vertex): every orientation sign is +1.  The edges of the
// mesh fall into 7 families by direction: x, y, z, x+y, x+z, y+z, x+y+z.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

namespace ems {
namespace tet {

static const int kFam[7][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
static const int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
static const int kLocEdge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
// Faces: a face with vertices p < q < r (global order) is identified by its
// base vertex p and the edge families of q - p and r - p; the 12 families
// that occur in the split, in this order (pairs of kFam indices):
static const int kFFam[12][2] = {{0, 3}, {0, 4}, {0, 6}, {1, 3}, {1, 5}, {1, 6},
                                 {2, 4}, {2, 5}, {2, 6}, {3, 6}, {4, 6}, {5, 6}};
// local faces of a tetrahedron (local vertex triples)
static const int kLocTri[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

struct Mesh {
    int nx = 0, ny = 0, nz = 0;
    double h = 0;
    long off[8] = {};    // first edge of each family
    long foff[13] = {};  // first face of each face family
    long ne = 0, nf = 0, nt = 0;

    void init(int x, int y, int z, double hh)
    {
        nx = x;
        ny = y;
        nz = z;
        h = hh;
        long o = 0;
        for (int f = 0; f < 7; ++f) {
            off[f] = o;
            o += (long)(nx + 1 - kFam[f][0]) * (ny + 1 - kFam[f][1]) * (nz + 1 - kFam[f][2]);
        }
        off[7] = o;
        ne = o;
        o = 0;
        for (int q = 0; q < 12; ++q) {
            foff[q] = o;
            const int *d = kFam[kFFam[q][1]];  // the farther vertex spans the face's extent
            o += (long)(nx + 1 - d[0]) * (ny + 1 - d[1]) * (nz + 1 - d[2]);
        }
        foff[12] = o;
        nf = o;
        nt = 6L * nx * ny * nz;
    }
    static int face_family(int f1, int f2)
    {
        for (int q = 0; q < 12; ++q)
            if (kFFam[q][0] == f1 && kFFam[q][1] == f2) return q;
        return -1;
    }
    long face(int q, int I, int J, int K) const
    {
        const int *d = kFam[kFFam[q][1]];
        const long cx = nx + 1 - d[0], cy = ny + 1 - d[1];
        return foff[q] + I + cx * (J + cy * (long)K);
    }
    // global faces of tet (cube i,j,k; type t) in local face order
    void tet_faces(int i, int j, int k, int t, long fc[4]) const
    {
        int v[4][3];
        verts(t, v);
        for (int q = 0; q < 4; ++q) {
            const int a = kLocTri[q][0], b = kLocTri[q][1], c = kLocTri[q][2];
            const int f1 = family(v[b][0] - v[a][0], v[b][1] - v[a][1], v[b][2] - v[a][2]);
            const int f2 = family(v[c][0] - v[a][0], v[c][1] - v[a][1], v[c][2] - v[a][2]);
            fc[q] = face(face_family(f1, f2), i + v[a][0], j + v[a][1], k + v[a][2]);
        }
    }
    void face_coords(long fid, int &q, int &I, int &J, int &K) const
    {
        q = 0;
        while (fid >= foff[q + 1]) ++q;
        const long r = fid - foff[q];
        const int *d = kFam[kFFam[q][1]];
        const long cx = nx + 1 - d[0], cy = ny + 1 - d[1];
        I = (int)(r % cx);
        J = (int)((r / cx) % cy);
        K = (int)(r / (cx * cy));
    }
    // a face lying in an outer face of the box
    bool face_boundary(long fid) const
    {
        int q, I, J, K;
        face_coords(fid, q, I, J, K);
        const int *d1 = kFam[kFFam[q][0]], *d2 = kFam[kFFam[q][1]];
        const int p[3] = {I, J, K}, n[3] = {nx, ny, nz};
        for (int a = 0; a < 3; ++a)
            if (d1[a] == 0 && d2[a] == 0 && (p[a] == 0 || p[a] == n[a])) return true;
        return false;
    }
    static int family_of(int dx, int dy, int dz)
    {
        for (int f = 0; f < 7; ++f)
            if (kFam[f][0] == dx && kFam[f][1] == dy && kFam[f][2] == dz) return f;
        return -1;
    }
    int family(int dx, int dy, int dz) const { return family_of(dx, dy, dz); }
    long edge(int f, int I, int J, int K) const
    {
        const long cx = nx + 1 - kFam[f][0], cy = ny + 1 - kFam[f][1];
        return off[f] + I + cx * (J + cy * (long)K);
    }
    // vertex offsets of tet type t
    static void verts(int t, int v[4][3])
    {
        for (int a = 0; a < 4; ++a)
            for (int c = 0; c < 3; ++c) v[a][c] = 0;
        v[1][kPerm[t][0]] = 1;
        v[2][kPerm[t][0]] = 1;
        v[2][kPerm[t][1]] = 1;
        v[3][0] = v[3][1] = v[3][2] = 1;
    }
    // global edges of tet (cube i,j,k; type t) in local edge order
    void tet_edges(int i, int j, int k, int t, long e[6]) const
    {
        int v[4][3];
        verts(t, v);
        for (int l = 0; l < 6; ++l) {
            const int a = kLocEdge[l][0], b = kLocEdge[l][1];
            const int f = family(v[b][0] - v[a][0], v[b][1] - v[a][1], v[b][2] - v[a][2]);
            e[l] = edge(f, i + v[a][0], j + v[a][1], k + v[a][2]);
        }
    }
    // edge family / base vertex of a global edge
    void edge_coords(long e, int &f, int &I, int &J, int &K) const
    {
        f = 0;
        while (e >= off[f + 1]) ++f;
        long r = e - off[f];
        const long cx = nx + 1 - kFam[f][0], cy = ny + 1 - kFam[f][1];
        I = (int)(r % cx);
        J = (int)((r / cx) % cy);
        K = (int)(r / (cx * cy));
    }
    // an edge with both end points on one outer face (PEC: tangential E = 0)
    bool boundary(long e) const
    {
        int f, I, J, K;
        edge_coords(e, f, I, J, K);
        const int d[3] = {kFam[f][0], kFam[f][1], kFam[f][2]};
        const int p[3] = {I, J, K}, n[3] = {nx, ny, nz};
        for (int a = 0; a < 3; ++a)
            if (d[a] == 0 && (p[a] == 0 || p[a] == n[a])) return true;
        return false;
    }
};

// Barycentric gradients and volume of tet type t scaled by h.
inline void bary(int t, double h, double g[4][3], double &vol)
{
    int v[4][3];
    Mesh::verts(t, v);
    double B[3][3];  // columns x1-x0, x2-x0, x3-x0
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) B[r][c] = h * (v[c + 1][r] - v[0][r]);
    const double det = B[0][0] * (B[1][1] * B[2][2] - B[1][2] * B[2][1]) -
                       B[0][1] * (B[1][0] * B[2][2] - B[1][2] * B[2][0]) +
                       B[0][2] * (B[1][0] * B[2][1] - B[1][1] * B[2][0]);
    double inv[3][3];
    inv[0][0] = (B[1][1] * B[2][2] - B[1][2] * B[2][1]) / det;
    inv[0][1] = (B[0][2] * B[2][1] - B[0][1] * B[2][2]) / det;
    inv[0][2] = (B[0][1] * B[1][2] - B[0][2] * B[1][1]) / det;
    inv[1][0] = (B[1][2] * B[2][0] - B[1][0] * B[2][2]) / det;
    inv[1][1] = (B[0][0] * B[2][2] - B[0][2] * B[2][0]) / det;
    inv[1][2] = (B[0][2] * B[1][0] - B[0][0] * B[1][2]) / det;
    inv[2][0] = (B[1][0] * B[2][1] - B[1][1] * B[2][0]) / det;
    inv[2][1] = (B[0][1] * B[2][0] - B[0][0] * B[2][1]) / det;
    inv[2][2] = (B[0][0] * B[1][1] - B[0][1] * B[1][0]) / det;
    // lambda_i (i = 1..3) = row i-1 of B^-1 applied to x - x0
    for (int i = 1; i < 4; ++i)
        for (int c = 0; c < 3; ++c) g[i][c] = inv[i - 1][c];
    for (int c = 0; c < 3; ++c) g[0][c] = -(g[1][c] + g[2][c] + g[3][c]);
    vol = std::fabs(det) / 6.0;
}

// Tet (type) containing point x (physical) of cube (i,j,k) and the
// barycentric coordinates there; returns -1 outside the mesh.
inline int locate(const Mesh &m, const double x[3], int &i, int &j, int &k, double lam[4])
{
    const double u[3] = {x[0] / m.h, x[1] / m.h, x[2] / m.h};
    const int n[3] = {m.nx, m.ny, m.nz};
    int c[3];
    double r[3];
    for (int a = 0; a < 3; ++a) {
        if (!(u[a] >= 0 && u[a] <= n[a])) return -1;
        c[a] = std::min((int)std::floor(u[a]), n[a] - 1);
        r[a] = u[a] - c[a];
    }
    i = c[0];
    j = c[1];
    k = c[2];
    // the Kuhn tet of the permutation that sorts the local coordinates descending
    int p[3] = {0, 1, 2};
    for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b)
            if (r[p[b]] > r[p[a]]) std::swap(p[a], p[b]);
    int t = 0;
    for (; t < 6; ++t)
        if (kPerm[t][0] == p[0] && kPerm[t][1] == p[1] && kPerm[t][2] == p[2]) break;
    lam[0] = 1.0 - r[p[0]];
    lam[1] = r[p[0]] - r[p[1]];
    lam[2] = r[p[1]] - r[p[2]];
    lam[3] = r[p[2]];
    return t;
}

}  // namespace tet
}  // namespace ems
