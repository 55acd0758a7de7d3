// Structured tetrahedral mesh (Kuhn split of an nx x ny x nz hex grid) and
// the first-order Nedelec (Whitney) element matrices.  Host-side geometry
// and assembly of the tetrahedral edge-element solver; no device code.
//
// Each cube (i,j,k) is split into 6 tetrahedra, one per permutation p of the
// axes, with vertices v0 = 0, v1 = e_p0, v2 = e_p0 + e_p1, v3 = (1,1,1)
// (offsets from the cube's corner).  Along this path the global vertex index
// increases, so a local edge (a < b) points the same way as its global edge
// (lower to higher vertex): every orientation sign is +1.  The edges of the
// mesh fall into 7 families by direction: x, y, z, x+y, x+z, y+z, x+y+z.
#pragma once

#include <array>
#include <cmath>
#include <vector>

namespace ems {
namespace tet {

static const int kFam[7][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
static const int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
static const int kLocEdge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

struct Mesh {
    int nx = 0, ny = 0, nz = 0;
    double h = 0;
    long off[8] = {};  // first edge of each family
    long ne = 0, nt = 0;

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
        nt = 6L * nx * ny * nz;
    }
    int family(int dx, int dy, int dz) const
    {
        for (int f = 0; f < 7; ++f)
            if (kFam[f][0] == dx && kFam[f][1] == dy && kFam[f][2] == dz) return f;
        return -1;
    }
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

// Whitney element matrices of tet type t (local edge order kLocEdge):
//   M_lm = int w_l . w_m,   w_ab = l_a grad l_b - l_b grad l_a,
//   int l_i l_j = V (1 + d_ij) / 20;
//   S_lm = int curl w_l . curl w_m = V (2 g_a x g_b) . (2 g_c x g_d).
inline void whitney(int t, double h, double M[6][6], double S[6][6])
{
    double g[4][3], V;
    bary(t, h, g, V);
    auto dot = [](const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    auto I2 = [&](int i, int j) { return V * (i == j ? 2.0 : 1.0) / 20.0; };
    double C[6][3];
    for (int l = 0; l < 6; ++l) {
        const int a = kLocEdge[l][0], b = kLocEdge[l][1];
        C[l][0] = 2.0 * (g[a][1] * g[b][2] - g[a][2] * g[b][1]);
        C[l][1] = 2.0 * (g[a][2] * g[b][0] - g[a][0] * g[b][2]);
        C[l][2] = 2.0 * (g[a][0] * g[b][1] - g[a][1] * g[b][0]);
    }
    for (int l = 0; l < 6; ++l)
        for (int m = 0; m < 6; ++m) {
            const int a = kLocEdge[l][0], b = kLocEdge[l][1], c = kLocEdge[m][0], d = kLocEdge[m][1];
            M[l][m] = dot(g[b], g[d]) * I2(a, c) - dot(g[b], g[c]) * I2(a, d) - dot(g[a], g[d]) * I2(b, c) +
                      dot(g[a], g[c]) * I2(b, d);
            S[l][m] = V * dot(C[l], C[m]);
        }
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

// Values of the 6 local Whitney functions at barycentric point lam.
inline void whitney_at(int t, double h, const double lam[4], double w[6][3])
{
    double g[4][3], V;
    bary(t, h, g, V);
    for (int l = 0; l < 6; ++l) {
        const int a = kLocEdge[l][0], b = kLocEdge[l][1];
        for (int c = 0; c < 3; ++c) w[l][c] = lam[a] * g[b][c] - lam[b] * g[a][c];
    }
}

}  // namespace tet
}  // namespace ems
