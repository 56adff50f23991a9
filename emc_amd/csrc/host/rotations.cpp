// Rotation sampling: refined 600-cell quaternion grid with volume weights.
//
// Behaviour of generate_rotation_list (rotations.c:77-458) is a drop-in contract: the order of
// the list (it indexes the rows of the responsibility matrix), the FP32 rounding of every point,
// the hemisphere pruning and the weights all match it.  The reference keeps `real` = float and
// evaluates pow()/sqrt()/acos() in double before storing to float; the same places do so here.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../../include/emc.h"
#include "host_internal.h"

namespace {

struct Quat {
    float v[4];
};

// pow(a,2)+pow(b,2)+pow(c,2)+pow(d,2) of floats, evaluated in double left to right
// (rotations.c:145-149, :181-189, :224-232, :281-284).  Squares of floats are exact in double.
inline double norm2d(float a, float b, float c, float d)
{
    return (double)a * a + (double)b * b + (double)c * c + (double)d * d;
}

class Cell600 {
public:
    Quat vert[120];
    int edge[720][2];
    int face[1200][3];
    int cell[600][4];
    float centre[600][4];
    int n_edge = 0, n_face = 0, n_cell = 0;

    bool build()
    {
        make_vertices();
        return find_edges() && find_faces() && find_cells();
    }

private:
    void make_vertices()
    {
        std::memset(vert, 0, sizeof vert);
        const float tau = (float)((1.0 + std::sqrt(5.0)) / 2.0);
        // (+-1/2)^4, bit 3 of the index drives q0 (rotations.c:87-98)
        for (int a = 0; a < 16; ++a)
            for (int k = 0; k < 4; ++k) vert[a].v[k] = (float)(-0.5 + (float)((a >> (3 - k)) & 1));
        // (+-1, 0, 0, 0) and permutations (rotations.c:101-103)
        for (int a = 0; a < 8; ++a) vert[16 + a].v[a / 2] = (float)(-1.0 + 2.0 * (float)(a % 2));
        // even permutations of (+-1/2, +-tau/2, +-1/(2 tau), 0) (rotations.c:106-132); the
        // permutation order fixes the vertex order and is kept as the reference lists it.
        static const unsigned char even_perm[12][4] = {
            {0, 1, 2, 3}, {0, 3, 1, 2}, {0, 2, 3, 1}, {1, 2, 0, 3}, {1, 3, 2, 0}, {1, 0, 3, 2},
            {2, 0, 1, 3}, {2, 3, 0, 1}, {2, 1, 3, 0}, {3, 1, 0, 2}, {3, 2, 1, 0}, {3, 0, 2, 1}};
        for (int g = 0; g < 12; ++g)
            for (int bits = 0; bits < 8; ++bits) {
                const int s1 = (bits >> 2) & 1, s2 = (bits >> 1) & 1, s3 = bits & 1;
                float *q = vert[24 + 8 * g + bits].v;
                q[even_perm[g][0]] = (float)(-0.5 + 1.0 * (float)s1);
                q[even_perm[g][1]] = (float)(tau * (-0.5 + 1.0 * (float)s2));
                q[even_perm[g][2]] = (float)(1.0 / tau * (-0.5 + 1.0 * (float)s3));
                q[even_perm[g][3]] = 0.0f;
            }
    }

    float sum_norm2(std::initializer_list<int> ids) const
    {
        float s[4] = {0, 0, 0, 0};
        bool first = true;
        for (int id : ids)
            for (int k = 0; k < 4; ++k) {
                if (first) s[k] = vert[id].v[k];
                else s[k] = s[k] + vert[id].v[k];
                if (k == 3) first = false;
            }
        return (float)norm2d(s[0], s[1], s[2], s[3]);
    }

    // neighbouring vertices: |vi + vj|^2 > 3 (rotations.c:134-158)
    bool find_edges()
    {
        for (int i = 0; i < 120; ++i)
            for (int j = 0; j < i; ++j)
                if (sum_norm2({i, j}) > 3.0f) {
                    if (n_edge == 720) return false;
                    edge[n_edge][0] = i;
                    edge[n_edge][1] = j;
                    ++n_edge;
                }
        return n_edge == 720;
    }

    // triangles: an edge plus a vertex untouched by this or any earlier edge (rotations.c:161-198)
    bool find_faces()
    {
        bool used[120] = {};
        for (int e = 0; e < n_edge; ++e) {
            used[edge[e][0]] = used[edge[e][1]] = true;
            for (int j = 0; j < 120; ++j) {
                if (used[j]) continue;
                if (sum_norm2({j, edge[e][0], edge[e][1]}) > 7.5f) {
                    if (n_face == 1200) return false;
                    face[n_face][0] = edge[e][0];
                    face[n_face][1] = edge[e][1];
                    face[n_face][2] = j;
                    ++n_face;
                }
            }
        }
        return n_face == 1200;
    }

    // tetrahedra: a face plus a vertex, each cell once (rotations.c:203-255)
    bool find_cells()
    {
        bool used[120] = {};
        for (int j = 0; j < 120; ++j) {
            used[j] = true;
            for (int f = 0; f < n_face; ++f) {
                const int a = face[f][0], b = face[f][1], c = face[f][2];
                if (used[a] || used[b] || used[c]) continue;
                if (sum_norm2({a, b, c, j}) > 13.5f) {
                    if (n_cell == 600) return false;
                    int *t = cell[n_cell];
                    t[0] = a; t[1] = b; t[2] = c; t[3] = j;
                    float *ce = centre[n_cell];
                    for (int k = 0; k < 4; ++k)
                        ce[k] = vert[a].v[k] + vert[b].v[k] + vert[c].v[k] + vert[j].v[k];
                    const float nrm = (float)std::sqrt(norm2d(ce[0], ce[1], ce[2], ce[3]));
                    for (int k = 0; k < 4; ++k) ce[k] /= nrm;
                    ++n_cell;
                }
            }
        }
        return n_cell == 600;
    }
};

// q.c / |q| with c the cell centre of largest such value (first one wins ties)
// (rotations.c:49-75)
float best_centre_cos(const float *q, const Cell600 &g)
{
    const double nrm = std::sqrt(norm2d(q[0], q[1], q[2], q[3]));
    float best = 0.0f, best_val = 0.0f;
    int best_i = 0;
    for (int i = 0; i < 600; ++i) {
        const float *c = g.centre[i];
        const float dot = q[0] * c[0] + q[1] * c[1] + q[2] * c[2] + q[3] * c[3];
        const float d = (float)((double)dot / nrm);
        if (d > best) {
            best = d;
            best_i = i;
        }
    }
    const float *c = g.centre[best_i];
    const float dot = q[0] * c[0] + q[1] * c[1] + q[2] * c[2] + q[3] * c[3];
    best_val = (float)((double)dot / nrm);
    return best_val;
}

inline float norm_cubed(const float *q)
{
    return (float)std::pow(norm2d(q[0], q[1], q[2], q[3]), (double)((float)3 / (float)2));
}

// hemisphere rule of rotations.c:402-419: first non-zero component positive
inline bool upper_hemisphere(const float *q)
{
    for (int k = 0; k < 4; ++k) {
        if (q[k] > 0) return true;
        if (!(q[k] == 0.0f)) return false;
    }
    return false;
}

}  // namespace

extern "C" int32_t emc_rotation_count(int32_t n)
{
    if (n < 1 || n > 200) return -1;
    return (int32_t)(20.0 * (n + 5.0 * std::pow((double)n, 3.0))) / 2;
}

extern "C" emc_status emc_generate_rotations(int32_t n, float *q_out, float *w_out)
{
    if (n < 1 || n > 200 || !q_out || !w_out) return emc_host_fail(EMC_E_INVALID, "generate_rotations: bad argument");
    Cell600 g;
    if (!g.build()) return emc_host_fail(EMC_E_INVALID, "generate_rotations: 600-cell construction failed");

    const float alpha = (float)std::acos(1.0 / 3.0);
    // weight factor of each simplex type (rotations.c:258-262)
    const float f_vert = (float)(20.0 * (3.0 * alpha - M_PI) / 4.0 / M_PI);
    const float f_edge = (float)(5.0 * alpha / 2.0 / M_PI);
    const float f_face = 1.0f, f_cell = 1.0f;

    const int N = (int)(20.0 * (n + 5.0 * std::pow((double)n, 3.0)));
    const int half = N / 2;
    const float Nf = (float)N;
    int kept = 0;
    float wsum_dummy = 0.0f;
    (void)wsum_dummy;

    // Points are produced in the reference's order (vertices, edges, faces, cells) and pruned on
    // the fly; kept points fill q_out/w_out in order.
    auto emit = [&](const float *q, float f) -> bool {
        if (!upper_hemisphere(q)) return true;
        if (kept >= half) return false;
        const float d3 = norm_cubed(q);
        const float c = best_centre_cos(q, g);
        for (int k = 0; k < 4; ++k) q_out[4 * kept + k] = q[k];
        w_out[kept] = f * c / Nf / d3;
        ++kept;
        return true;
    };

    float q[4];
    for (int i = 0; i < 120; ++i) {
        for (int k = 0; k < 4; ++k) q[k] = g.vert[i].v[k];
        if (!emit(q, f_vert)) return emc_host_fail(EMC_E_INVALID, "generate_rotations: pruning overflow");
    }
    const int m = n;  // points per edge + 1
    for (int e = 0; e < 720; ++e) {
        const float *a = g.vert[g.edge[e][0]].v, *b = g.vert[g.edge[e][1]].v;
        for (int j = 0; j < m - 1; ++j) {
            const float ta = (float)(j + 1) / (float)m, tb = (float)(m - 1 - j) / (float)m;
            for (int k = 0; k < 4; ++k) q[k] = ta * a[k] + tb * b[k];
            if (!emit(q, f_edge)) return emc_host_fail(EMC_E_INVALID, "generate_rotations: pruning overflow");
        }
    }
    if (n >= 3)
        for (int f = 0; f < 1200; ++f) {
            const float *a = g.vert[g.face[f][0]].v, *b = g.vert[g.face[f][1]].v,
                        *c = g.vert[g.face[f][2]].v;
            for (int ka = 2; ka < m; ++ka)
                for (int kb = 2; kb < m; ++kb) {
                    if (ka + kb <= m) continue;
                    const int kc = 2 * m - ka - kb;
                    const float den = (float)(3 * m - ka - kb - kc);
                    const float wa = (float)(m - ka) / den, wb = (float)(m - kb) / den,
                                wc = (float)(m - kc) / den;
                    for (int k = 0; k < 4; ++k) q[k] = wa * a[k] + wb * b[k] + wc * c[k];
                    if (!emit(q, f_face)) return emc_host_fail(EMC_E_INVALID, "generate_rotations: pruning overflow");
                }
        }
    if (n >= 4)
        for (int t = 0; t < 600; ++t) {
            const float *a = g.vert[g.cell[t][0]].v, *b = g.vert[g.cell[t][1]].v,
                        *c = g.vert[g.cell[t][2]].v, *d = g.vert[g.cell[t][3]].v;
            for (int ka = 3; ka < m; ++ka)
                for (int kb = 3; kb < m; ++kb)
                    for (int kc = 3; kc < m; ++kc) {
                        const int kd = 3 * m - ka - kb - kc;
                        if (kd < 3 || kd >= m) continue;
                        const float den = (float)(4 * m - ka - kb - kc - kd);
                        const float wa = (float)(m - ka) / den, wb = (float)(m - kb) / den,
                                    wc = (float)(m - kc) / den, wd = (float)(m - kd) / den;
                        for (int k = 0; k < 4; ++k) q[k] = wa * a[k] + wb * b[k] + wc * c[k] + wd * d[k];
                        if (!emit(q, f_cell)) return emc_host_fail(EMC_E_INVALID, "generate_rotations: pruning overflow");
                    }
        }
    if (kept != half) return emc_host_fail(EMC_E_INVALID, "generate_rotations: hemisphere did not keep half the points");

    // weights normalised with a float running sum (rotations.c:429-439)
    float wsum = 0.0f;
    for (int i = 0; i < half; ++i) wsum += w_out[i];
    for (int i = 0; i < half; ++i) w_out[i] /= wsum;
    // unit quaternions (rotations.c:446-453 -> quaternion_normalize, rotations.c:21-28)
    for (int i = 0; i < half; ++i) {
        float *p = q_out + 4 * i;
        const float a = (float)std::sqrt(norm2d(p[0], p[1], p[2], p[3]));
        for (int k = 0; k < 4; ++k) p[k] /= a;
    }
    return EMC_OK;
}
