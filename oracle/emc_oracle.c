/*
 * emc_oracle.c -- CPU restatement of one EMC (Expand-Maximize-Compress) iteration of JinLi971/EMC.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only to compare against (or to time beside)
 * the HIP library.  The product path (emc_amd/libemc.so) never links or calls it.
 *
 * Arithmetic:
 *   - Everything that the reference computes as numbers (slices, log-likelihoods, softmax,
 *     M-step sums, divide, blur, normalisation) is done here in FP64.
 *   - Everything that is an index or a discrete decision (rotation list order and pruning,
 *     detector coordinates, rotation matrices, voxel indices, trilinear fractions, validity)
 *     is computed in FP32 with an explicitly specified operation sequence (fmaf / single
 *     products, no contraction: build with -ffp-contract=off).  The HIP kernels use the same
 *     sequence, so both sides take identical discrete decisions and differ only by rounding
 *     of the interpolated values.
 *
 * Every function names the reference lines it restates.  Parity status of each function is
 * listed in the comment above it ("pinned by": the tests in tests/test_oracle_*.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------------------------ */
/* a1: rotation sampling (rotations.c:43, 49-75, 77-458)                                       */
/* FP32-faithful: `real` is float in the reference build, libm calls are made in double and     */
/* stored to float exactly where the reference does.                                             */
/* Pinned by: count closed form, 600-cell combinatorics, textbook vertex set, weight symmetry    */
/* under the binary icosahedral group, sum of weights, n=1 equal weights (test_oracle_rot.py).   */
/* ------------------------------------------------------------------------------------------ */

/* rotations.c:43  n_to_samples = 20*(n+5n^3); the list is pruned to half of it. */
ORC_API int orc_rotation_count(int n) { return (int)(20.0 * (n + 5.0 * pow((double)n, 3.0))) / 2; }

/* sum of squares the way rotations.c:145-149 etc. evaluate pow(float,2) in double */
static double sq4d(float a, float b, float c, float d)
{
    return (double)a * a + (double)b * b + (double)c * c + (double)d * d;
}

/* rotations.c:49-75 */
static float best_center_dot(const float *q, const float *centers)
{
    double nrm = sqrt(sq4d(q[0], q[1], q[2], q[3]));
    float best = 0.0f;
    int bi = 0;
    for (int i = 0; i < 600; i++) {
        const float *c = centers + 4 * i;
        float num = q[0] * c[0] + q[1] * c[1] + q[2] * c[2] + q[3] * c[3];
        float d = (float)((double)num / nrm);
        if (d > best) { best = d; bi = i; }
    }
    const float *c = centers + 4 * bi;
    float num = q[0] * c[0] + q[1] * c[1] + q[2] * c[2] + q[3] * c[3];
    return (float)((double)num / nrm);
}

static float dist3_of(const float *q)
{
    return (float)pow(sq4d(q[0], q[1], q[2], q[3]), (double)((float)3 / (float)2));
}

/* stats (optional, may be NULL): [0]=edges [1]=faces [2]=cells [3]=kept [4]=total */
ORC_API int orc_rotations(int n, float *q_out, float *w_out, int *stats)
{
    if (n < 1) return -1;
    const float tau = (float)((1.0 + sqrt(5.0)) / 2.0);
    float V[120][4];
    memset(V, 0, sizeof V);
    /* rotations.c:87-98 : 16 vertices (+-1/2, +-1/2, +-1/2, +-1/2) in binary order */
    for (int a = 0; a < 16; a++)
        for (int k = 0; k < 4; k++) V[a][k] = (float)(-0.5 + (float)((a >> (3 - k)) & 1));
    /* rotations.c:101-103 : 8 vertices (+-1,0,0,0) and permutations */
    for (int i = 0; i < 8; i++) V[16 + i][i / 2] = (float)(-1.0 + 2.0 * (float)(i % 2));
    /* rotations.c:106-132 : 96 vertices, even permutations of (+-1/2, +-tau/2, +-1/(2tau), 0) */
    static const int perm[12][4] = {{1, 2, 3, 4}, {1, 4, 2, 3}, {1, 3, 4, 2}, {2, 3, 1, 4},
                                    {2, 4, 3, 1}, {2, 1, 4, 3}, {3, 1, 2, 4}, {3, 4, 1, 2},
                                    {3, 2, 4, 1}, {4, 2, 1, 3}, {4, 3, 2, 1}, {4, 1, 3, 2}};
    for (int i = 0; i < 12; i++)
        for (int j1 = 0; j1 < 2; j1++)
            for (int j2 = 0; j2 < 2; j2++)
                for (int j3 = 0; j3 < 2; j3++) {
                    float *v = V[24 + 8 * i + 4 * j1 + 2 * j2 + j3];
                    v[perm[i][0] - 1] = (float)(-0.5 + 1.0 * (float)j1);
                    v[perm[i][1] - 1] = (float)(tau * (-0.5 + 1.0 * (float)j2));
                    v[perm[i][2] - 1] = (float)(1.0 / tau * (-0.5 + 1.0 * (float)j3));
                    v[perm[i][3] - 1] = 0.0f;
                }
    /* rotations.c:134-158 : edges, |vi+vj|^2 > 3 */
    int E[720][2], ne = 0;
    for (int i = 0; i < 120; i++)
        for (int j = 0; j < i; j++) {
            float d2 = (float)sq4d(V[i][0] + V[j][0], V[i][1] + V[j][1], V[i][2] + V[j][2],
                                   V[i][3] + V[j][3]);
            if (d2 > 3.0f) {
                if (ne >= 720) return -2;
                E[ne][0] = i; E[ne][1] = j; ne++;
            }
        }
    /* rotations.c:161-198 : faces, vertices not yet touched by an earlier edge */
    int F[1200][3], nf = 0, done[120];
    memset(done, 0, sizeof done);
    for (int i = 0; i < ne; i++) {
        done[E[i][0]] = 1; done[E[i][1]] = 1;
        for (int j = 0; j < 120; j++) {
            if (done[j]) continue;
            const float *a = V[j], *b = V[E[i][0]], *c = V[E[i][1]];
            float d2 = (float)sq4d(a[0] + b[0] + c[0], a[1] + b[1] + c[1], a[2] + b[2] + c[2],
                                   a[3] + b[3] + c[3]);
            if (d2 > 7.5f) {
                if (nf >= 1200) return -3;
                F[nf][0] = E[i][0]; F[nf][1] = E[i][1]; F[nf][2] = j; nf++;
            }
        }
    }
    /* rotations.c:203-255 : cells and normalised cell centres */
    int C[600][4], nc = 0;
    float centers[4 * 600];
    memset(done, 0, sizeof done);
    for (int j = 0; j < 120; j++) {
        done[j] = 1;
        for (int i = 0; i < nf; i++) {
            if (done[F[i][0]] || done[F[i][1]] || done[F[i][2]]) continue;
            const float *a = V[F[i][0]], *b = V[F[i][1]], *c = V[F[i][2]], *d = V[j];
            float d2 = (float)sq4d(a[0] + b[0] + c[0] + d[0], a[1] + b[1] + c[1] + d[1],
                                   a[2] + b[2] + c[2] + d[2], a[3] + b[3] + c[3] + d[3]);
            if (d2 > 13.5f) {
                if (nc >= 600) return -4;
                C[nc][0] = F[i][0]; C[nc][1] = F[i][1]; C[nc][2] = F[i][2]; C[nc][3] = j;
                float *ce = centers + 4 * nc;
                for (int k = 0; k < 4; k++) ce[k] = a[k] + b[k] + c[k] + d[k];
                float nrm = (float)sqrt(sq4d(ce[0], ce[1], ce[2], ce[3]));
                for (int k = 0; k < 4; k++) ce[k] /= nrm;
                nc++;
            }
        }
    }
    if (ne != 720 || nf != 1200 || nc != 600) return -5;
    /* rotations.c:258-262 */
    const float alpha = (float)acos(1.0 / 3.0);
    const float f1 = (float)(5.0 * alpha / 2.0 / M_PI);
    const float f0 = (float)(20.0 * (3.0 * alpha - M_PI) / 4.0 / M_PI);
    const float f2 = 1.0f, f3 = 1.0f;
    const int N = (int)(20.0 * (n + 5.0 * pow((double)n, 3.0)));
    float *Q = (float *)calloc((size_t)N * 4, sizeof(float));
    float *W = (float *)calloc((size_t)N, sizeof(float));
    if (!Q || !W) { free(Q); free(W); return -6; }
    /* rotations.c:276-287 : vertices */
    for (int i = 0; i < 120; i++) {
        float *q = Q + 4 * i;
        for (int k = 0; k < 4; k++) q[k] = V[i][k];
        float d3 = dist3_of(q);
        float sp = best_center_dot(q, centers);
        W[i] = f0 * sp / (float)N / d3;
    }
    /* rotations.c:290-307 : n-1 points per edge */
    const int ev = n - 1;
    for (int i = 0; i < 720; i++)
        for (int j = 0; j < ev; j++) {
            int idx = 120 + ev * i + j;
            float *q = Q + 4 * idx;
            for (int k = 0; k < 4; k++)
                q[k] = (float)(j + 1) / (float)(ev + 1) * V[E[i][0]][k] +
                       (float)(ev - j) / (float)(ev + 1) * V[E[i][1]][k];
            float d3 = dist3_of(q);
            float sp = best_center_dot(q, centers);
            W[idx] = f1 * sp / (float)N / d3;
        }
    /* rotations.c:310-343 : (n-1)(n-2)/2 points per face */
    const int fv = ((n - 1) * (n - 2)) / 2;
    const int fbase = 120 + 720 * ev;
    if (fv > 0)
        for (int i = 0; i < 1200; i++) {
            int cnt = 0;
            for (int ka = 2; ka < ev + 1; ka++)
                for (int kb = 2; kb < ev + 1; kb++) {
                    if (ka + kb <= ev + 1) continue;
                    int kc = 2 * (ev + 1) - ka - kb;
                    float den = (float)(3 * (ev + 1) - ka - kb - kc);
                    float a = (float)(ev + 1 - ka) / den;
                    float b = (float)(ev + 1 - kb) / den;
                    float c = (float)(ev + 1 - kc) / den;
                    int idx = fbase + fv * i + cnt;
                    float *q = Q + 4 * idx;
                    for (int k = 0; k < 4; k++)
                        q[k] = a * V[F[i][0]][k] + b * V[F[i][1]][k] + c * V[F[i][2]][k];
                    float d3 = dist3_of(q);
                    float sp = best_center_dot(q, centers);
                    W[idx] = f2 * sp / (float)N / d3;
                    cnt++;
                }
        }
    /* rotations.c:346-384 : (n-1)(n-2)(n-3)/6 points per cell */
    const int cv = ((n - 1) * (n - 2) * (n - 3)) / 6;
    const int cbase = fbase + 1200 * fv;
    if (cv > 0)
        for (int i = 0; i < 600; i++) {
            int cnt = 0;
            for (int ka = 3; ka < ev + 1; ka++)
                for (int kb = 3; kb < ev + 1; kb++)
                    for (int kc = 3; kc < ev + 1; kc++) {
                        int kd = 3 * (ev + 1) - ka - kb - kc;
                        if (!(kd >= 3 && kd < ev + 1)) continue;
                        float den = (float)(4 * (ev + 1) - ka - kb - kc - kd);
                        float a = (float)(ev + 1 - ka) / den;
                        float b = (float)(ev + 1 - kb) / den;
                        float c = (float)(ev + 1 - kc) / den;
                        float d = (float)(ev + 1 - kd) / den;
                        int idx = cbase + cv * i + cnt;
                        float *q = Q + 4 * idx;
                        for (int k = 0; k < 4; k++)
                            q[k] = a * V[C[i][0]][k] + b * V[C[i][1]][k] + c * V[C[i][2]][k] +
                                   d * V[C[i][3]][k];
                        float d3 = dist3_of(q);
                        float sp = best_center_dot(q, centers);
                        W[idx] = f3 * sp / (float)N / d3;
                        cnt++;
                    }
        }
    /* rotations.c:394-425 : keep the q0>0 hemisphere (ties broken on q1, q2, q3) */
    const int half = N / 2;
    int kept = 0;
    for (int i = 0; i < N; i++) {
        const float *q = Q + 4 * i;
        int keep = 0;
        if (q[0] > 0) keep = 1;
        else if (q[0] == 0.0f) {
            if (q[1] > 0) keep = 1;
            else if (q[1] == 0.0f) {
                if (q[2] > 0) keep = 1;
                else if (q[2] == 0.0f && q[3] > 0) keep = 1;
            }
        }
        if (keep) {
            if (kept < half) {
                for (int k = 0; k < 4; k++) q_out[4 * kept + k] = q[k];
                w_out[kept] = W[i];
            }
            kept++;
        }
    }
    if (stats) { stats[0] = ne; stats[1] = nf; stats[2] = nc; stats[3] = kept; stats[4] = N; }
    free(Q); free(W);
    if (kept != half) return -7;
    /* rotations.c:429-439 : normalise weights (float running sum) */
    float wsum = 0.0f;
    for (int i = 0; i < half; i++) wsum += w_out[i];
    for (int i = 0; i < half; i++) w_out[i] /= wsum;
    /* rotations.c:446-453 -> quaternion_normalize (rotations.c:21-28) */
    for (int i = 0; i < half; i++) {
        float *q = q_out + 4 * i;
        float a = (float)sqrt(sq4d(q[0], q[1], q[2], q[3]));
        for (int k = 0; k < 4; k++) q[k] /= a;
    }
    return half;
}

/* ------------------------------------------------------------------------------------------ */
/* a2: detector coordinates                                                                    */
/* ------------------------------------------------------------------------------------------ */

/* emc_helper.cpp:55-79, FP32-faithful.  Output index p = y*L + x.
 * Pinned by: closed form X = x-L/2+0.5, Y likewise, Z = r(1-cos t)/sin t (test_oracle_geom.py). */
ORC_API void orc_coords_ewald(int L, float pixel_size, float detector_distance, float wavelength,
                              float *X, float *Y, float *Z)
{
    for (int x = 0; x < L; x++)
        for (int y = 0; y < L; y++) {
            double dx = (double)(float)(x - L / 2) + 0.5, dy = (double)(float)(y - L / 2) + 0.5;
            float pr = (float)sqrt(dx * dx + dy * dy);
            float rr = pr * pixel_size;
            float ang = (float)atan2((double)rr, (double)detector_distance);
            float fr = (float)(sin((double)ang) / (double)wavelength);
            float fz = (float)((1. - cos((double)ang)) / (double)wavelength);
            X[y * L + x] = (float)dx;
            Y[y * L + x] = (float)dy;
            Z[y * L + x] = fz / fr * pr;
        }
}

/* Flat detector plane used by the synthetic workloads: Z = 0. */
ORC_API void orc_coords_flat(int L, float *X, float *Y, float *Z)
{
    for (int y = 0; y < L; y++)
        for (int x = 0; x < L; x++) {
            X[y * L + x] = (float)((double)(x - L / 2) + 0.5);
            Y[y * L + x] = (float)((double)(y - L / 2) + 0.5);
            Z[y * L + x] = 0.0f;
        }
}

/* Beamstop/outer-radius mask on pixel-centre radius: 1 iff rb <= r <= ro. */
ORC_API int orc_mask_annulus(int L, double rb, double ro, int32_t *mask)
{
    int n = 0;
    for (int y = 0; y < L; y++)
        for (int x = 0; x < L; x++) {
            double dx = x - L / 2 + 0.5, dy = y - L / 2 + 0.5;
            double r = sqrt(dx * dx + dy * dy);
            int m = (r >= rb && r <= ro) ? 1 : 0;
            mask[y * L + x] = m;
            n += m;
        }
    return n;
}

/* ------------------------------------------------------------------------------------------ */
/* Geometry spec shared with the HIP kernels (emc_cuda.cu:113-138, emc_atomic.cu:50-69)         */
/* ------------------------------------------------------------------------------------------ */

/* Rotation matrix of a unit quaternion (emc_cuda.cu:113-121) with an explicit FMA sequence. */
ORC_API void orc_rotmat(const float *q, float *m)
{
    const float q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    m[0] = fmaf(q0, q0, fmaf(q1, q1, fmaf(-q2, q2, -(q3 * q3))));
    m[1] = 2.0f * fmaf(q1, q2, -(q0 * q3));
    m[2] = 2.0f * fmaf(q1, q3, q0 * q2);
    m[3] = 2.0f * fmaf(q1, q2, q0 * q3);
    m[4] = fmaf(q0, q0, fmaf(-q1, q1, fmaf(q2, q2, -(q3 * q3))));
    m[5] = 2.0f * fmaf(q2, q3, -(q0 * q1));
    m[6] = 2.0f * fmaf(q1, q3, -(q0 * q2));
    m[7] = 2.0f * fmaf(q2, q3, q0 * q1);
    m[8] = fmaf(q0, q0, fmaf(-q1, q1, fmaf(-q2, q2, q3 * q3)));
}

/* Model-grid coordinate of detector point (X,Y,Z) under rotation m: u = (L/2-0.5) + m.x */
static inline void rot_point(const float *m, float X, float Y, float Z, float c, float *u)
{
    u[0] = c + fmaf(m[0], X, fmaf(m[1], Y, m[2] * Z));
    u[1] = c + fmaf(m[3], X, fmaf(m[4], Y, m[5] * Z));
    u[2] = c + fmaf(m[6], X, fmaf(m[7], Y, m[8] * Z));
}

/* Trilinear cell: valid iff 0 <= u <= L-1 on every axis; base index clamped to L-2. */
static inline int tri_cell(const float *u, int L, int *b, float *f)
{
    for (int a = 0; a < 3; a++) {
        if (!(u[a] >= 0.0f && u[a] <= (float)(L - 1))) return 0;
        int i0 = (int)floorf(u[a]);
        if (i0 > L - 2) i0 = L - 2;
        b[a] = i0;
        f[a] = u[a] - (float)i0;
    }
    return 1;
}

/* ------------------------------------------------------------------------------------------ */
/* a3: expand -- slicing (emc_cuda.cu:102-157)                                                 */
/* interp 0 = nearest (reference parity: lroundf, bounds 0 < v < L), 1 = trilinear.             */
/* Invalid samples are -1 (reference value).  Trilinear: invalid if any of the 8 corner voxels  */
/* is negative (a missing voxel) or the point is outside [0, L-1]^3.                             */
/* Pinned by: brute force (pure Python), identity rotation -> z-plane, 90-degree permutations,  */
/* adjoint identity with orc_mstep's insertion (test_oracle_stages.py).                          */
/* ------------------------------------------------------------------------------------------ */
ORC_API void orc_slices(const double *model, int L, const float *quat, int S, const float *X,
                        const float *Y, const float *Z, int P, int interp, double *Wout)
{
    const float c = 0.5f * (float)L - 0.5f;
#pragma omp parallel for schedule(static)
    for (int s = 0; s < S; s++) {
        float m[9];
        orc_rotmat(quat + 4 * (size_t)s, m);
        double *W = Wout + (size_t)s * P;
        for (int p = 0; p < P; p++) {
            float u[3];
            rot_point(m, X[p], Y[p], Z[p], c, u);
            if (interp == 0) {
                int vx = (int)roundf(u[0]), vy = (int)roundf(u[1]), vz = (int)roundf(u[2]);
                if (vx > 0 && vx < L && vy > 0 && vy < L && vz > 0 && vz < L)
                    W[p] = model[((size_t)vz * L + vy) * L + vx];
                else
                    W[p] = -1.0;
            } else {
                int b[3];
                float f[3];
                if (!tri_cell(u, L, b, f)) { W[p] = -1.0; continue; }
                double acc = 0.0;
                int bad = 0;
                for (int dz = 0; dz < 2; dz++)
                    for (int dy = 0; dy < 2; dy++)
                        for (int dx = 0; dx < 2; dx++) {
                            double v = model[((size_t)(b[2] + dz) * L + (b[1] + dy)) * L + (b[0] + dx)];
                            if (v < 0.0) bad = 1;
                            double wgt = (dx ? (double)f[0] : 1.0 - (double)f[0]) *
                                         (dy ? (double)f[1] : 1.0 - (double)f[1]) *
                                         (dz ? (double)f[2] : 1.0 - (double)f[2]);
                            acc += wgt * v;
                        }
                W[p] = bad ? -1.0 : acc;
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* a5: likelihood logits R[s*I + i]  (emc_cuda.cu:160-216, 242-276)                            */
/* diff: 0 absolute, 1 "poisson" (the reference's chi^2 form), 2 relative, 3 true Poisson.      */
/* V_s = {p : mask[p] != 0 and W[s,p] > 0}, n_s = |V_s|.                                         */
/*  chi^2 kinds: R = log w_s - D/(2 n_s sigma^2); D = sum over V_s of                            */
/*      abs: (W-K/phi)^2   poisson: (W-K/phi)^2/(W+0.02)   relative: ((W-K/phi)/W)^2             */
/*      (n_s = 0 gives D/(2 n_s sigma^2) := 0; the reference produces 0/0 there.)                */
/*  true Poisson: R = log w_s + sum_{V_s} K log(phi0 W) - phi0 sum_{V_s} W                        */
/* phi (per s,i scaling) may be NULL (= 1).                                                       */
/* Pinned by: pure-Python brute force, constant-W closed form, GEMM identity (tests).            */
/* ------------------------------------------------------------------------------------------ */
ORC_API void orc_logl(const double *W, const double *K, const int32_t *mask, int S, int I, int P,
                      int diff, double sigma, const double *wrot, const double *phi, double phi0,
                      double *R, int32_t *n_s_out)
{
#pragma omp parallel
    {
        double *lw = (double *)malloc(sizeof(double) * (size_t)P);
        int *vp = (int *)malloc(sizeof(int) * (size_t)P);
#pragma omp for schedule(dynamic, 4)
        for (int s = 0; s < S; s++) {
            const double *Ws = W + (size_t)s * P;
            int ns = 0;
            double sumW = 0.0;
            for (int p = 0; p < P; p++)
                if (mask[p] != 0 && Ws[p] > 0.0) {
                    vp[ns++] = p;
                    sumW += Ws[p];
                }
            if (n_s_out) n_s_out[s] = ns;
            const double lws = log(wrot[s]);
            if (diff == 3)
                for (int k = 0; k < ns; k++) lw[k] = log(phi0 * Ws[vp[k]]);
            for (int i = 0; i < I; i++) {
                const double *Ki = K + (size_t)i * P;
                const double ph = phi ? phi[(size_t)s * I + i] : 1.0;
                double acc = 0.0;
                if (diff == 3) {
                    for (int k = 0; k < ns; k++) acc += Ki[vp[k]] * lw[k];
                    R[(size_t)s * I + i] = lws + acc - phi0 * sumW;
                } else {
                    for (int k = 0; k < ns; k++) {
                        const int p = vp[k];
                        double d = Ws[p] - Ki[p] / ph;
                        if (diff == 0) acc += d * d;
                        else if (diff == 1) acc += d * d / (Ws[p] + 0.02);
                        else acc += (d / Ws[p]) * (d / Ws[p]);
                    }
                    double t = ns > 0 ? acc / 2.0 / (double)ns / (sigma * sigma) : 0.0;
                    R[(size_t)s * I + i] = lws - t;
                }
            }
        }
        free(lw);
        free(vp);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* a5b: per (rotation, pattern) intensity scaling for known_intensity = 0 (main.cpp:283-295,      */
/* emc_cuda.cu:438-455 calculate_scaling_absolute, 516-531 update_scaling_full_kernel).           */
/*   phi[s, i] = sum K^2 / sum K W over {p : mask[p] != 0, W[s,p] > 1e-10, K[i,p] > 1e-10}       */
/* Exact double sums.  The reference accumulates each thread's sum K W in an int (the float      */
/* product is truncated on every add); this restatement keeps the real-valued sum.  phi = 1 when */
/* either sum is zero (the reference produces 0/0 or x/0 there).                                 */
/* Pinned by: closed form for one-pixel and proportional patterns (K = c W gives phi = c) and     */
/* pure-Python brute force (tests).                                                              */
/* ------------------------------------------------------------------------------------------ */
ORC_API void orc_scaling(const double *W, const double *K, const int32_t *mask, int S, int I, int P,
                         double *phi)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (int s = 0; s < S; s++) {
        const double *Ws = W + (size_t)s * P;
        for (int i = 0; i < I; i++) {
            const double *Ki = K + (size_t)i * P;
            double kk = 0.0, kw = 0.0;
            for (int p = 0; p < P; p++)
                if (mask[p] != 0 && Ws[p] > 1e-10 && Ki[p] > 1e-10) {
                    kk += Ki[p] * Ki[p];
                    kw += Ki[p] * Ws[p];
                }
            phi[(size_t)s * I + i] = (kk > 0.0 && kw > 0.0) ? kk / kw : 1.0;
        }
    }
}

/* Reference-emulation mode of the same scaling (parity mode only; the library computes the exact */
/* sums above).  calculate_scaling_absolute with 512 threads (TNUM, emc_cuda_help.h:8): thread t   */
/* takes pixels t, t+512, ...; its sum of K^2 is a float, its sum of K W an int that truncates the */
/* float (int + K W) on every add; inblock_reduce (emc_cuda.cu:52-60) then adds the 512 partials in */
/* a halving tree (float and int); phi = float / int (0 gives inf or nan, as in the reference).    */
/* Pinned by: equal to orc_scaling when every K W and every partial sum is an integer; below it  */
/* (truncation) otherwise (tests).                                                                */
ORC_API void orc_scaling_reference(const float *W, const float *K, const int32_t *mask, int S, int I, int P,
                                   double *phi)
{
    enum { T = 512 };
#pragma omp parallel for schedule(dynamic, 4)
    for (int s = 0; s < S; s++) {
        float fs[T];
        int iw[T];
        const float *Ws = W + (size_t)s * P;
        for (int i = 0; i < I; i++) {
            const float *Ki = K + (size_t)i * P;
            for (int t = 0; t < T; t++) {
                fs[t] = 0.0f;
                iw[t] = 0;
                for (int p = t; p < P; p += T)
                    if (mask[p] > 0 && (double)Ws[p] > 1e-10 && (double)Ki[p] > 1e-10) {
                        fs[t] = fs[t] + Ki[p] * Ki[p];
                        iw[t] = (int)((float)iw[t] + Ki[p] * Ws[p]);
                    }
            }
            for (int h = T / 2; h > 0; h >>= 1)
                for (int t = 0; t < h; t++) {
                    fs[t] += fs[t + h];
                    iw[t] += iw[t + h];
                }
            phi[(size_t)s * I + i] = (double)(fs[0] / (float)iw[0]);
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Model initialisation (emc_helper.cpp:440-534) and pattern normalisation (:389-419).          */
/* Pinned by: constant patterns give the constant inside L/2 and -1 outside; a brute-force      */
/* numpy shell average; one pattern at the identity rotation lands shifted by (+1, +1) on plane */
/* z = L/2 + 1 (the +0.5 of the host insert); normalised patterns sum to P over the mask (tests). */
/* ------------------------------------------------------------------------------------------ */
/* model_input = 3.  radavg[L/2] (read past the reference's array) is 0 here. */
ORC_API void orc_model_radial(int L, const float *K, int I, double *model)
{
    const int H = L / 2;
    double *sum = (double *)calloc((size_t)H + 1, sizeof(double));
    double *cnt = (double *)calloc((size_t)H + 1, sizeof(double));
    for (int i = 0; i < I; i++)
        for (int x = 0; x < L; x++)
            for (int y = 0; y < L; y++) {
                const double dx = x - L / 2.0 + 0.5, dy = y - L / 2.0 + 0.5;
                const double r = sqrt(dx * dx + dy * dy);
                if (r < L / 2.0) {
                    sum[(int)r] += K[(size_t)i * L * L + (size_t)y * L + x];
                    cnt[(int)r] += 1.0;
                }
            }
    for (int r = 0; r < H; r++) sum[r] = cnt[r] > 0.0 ? sum[r] / cnt[r] : 0.0;
    sum[H] = 0.0;
    for (int x = 0; x < L; x++)
        for (int y = 0; y < L; y++)
            for (int z = 0; z < L; z++) {
                const double dx = x - L / 2.0 + 0.5, dy = y - L / 2.0 + 0.5, dz = z - L / 2.0 + 0.5;
                const double rad = sqrt(dx * dx + dy * dy + dz * dz);
                const int r = (int)rad;
                model[((size_t)z * L + y) * L + x] = rad < L / 2.0 ? sum[r] + (sum[r + 1] - sum[r]) * (rad - r) : -1.0;
            }
    free(sum);
    free(cnt);
}

/* model_input = 1 with the orientations q [I][4] given: insert_slice (emc_helper.cpp:80-123) of every
 * pattern with weight 1 at round(L/2 + 0.5 + R x) (R x rounded to float, as the reference stores it),
 * then num / weight (0 where no weight). */
ORC_API void orc_model_orientations(int L, const float *K, int I, const int32_t *mask, const float *X,
                                    const float *Y, const float *Z, const float *q, double *model)
{
    const size_t N3 = (size_t)L * L * L, P = (size_t)L * L;
    double *num = (double *)calloc(N3, sizeof(double));
    double *wt = (double *)calloc(N3, sizeof(double));
    for (int i = 0; i < I; i++) {
        const double q0 = q[4 * i], q1 = q[4 * i + 1], q2 = q[4 * i + 2], q3 = q[4 * i + 3];
        for (size_t p = 0; p < P; p++) {
            if (mask[p] != 1) continue;
            const float nx = (float)((q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3) * X[p] + 2.0 * (q1 * q2 - q0 * q3) * Y[p] +
                                     2.0 * (q1 * q3 + q0 * q2) * Z[p]);
            const float ny = (float)(2.0 * (q1 * q2 + q0 * q3) * X[p] + (q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3) * Y[p] +
                                     2.0 * (q2 * q3 - q0 * q1) * Z[p]);
            const float nz = (float)(2.0 * (q1 * q3 - q0 * q2) * X[p] + 2.0 * (q2 * q3 + q0 * q1) * Y[p] +
                                     (q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3) * Z[p]);
            const long vx = lround(L / 2.0 + 0.5 + nx), vy = lround(L / 2.0 + 0.5 + ny), vz = lround(L / 2.0 + 0.5 + nz);
            if (vx < 0 || vx >= L || vy < 0 || vy >= L || vz < 0 || vz >= L) continue;
            const size_t o = ((size_t)vz * L + vy) * L + vx;
            num[o] += K[(size_t)i * P + p];
            wt[o] += 1.0;
        }
    }
    for (size_t o = 0; o < N3; o++) model[o] = wt[o] > 0.0 ? num[o] / wt[o] : 0.0;
    free(num);
    free(wt);
}

/* normalize_images: K * P / sum_{mask == 1} K (unchanged when that sum is 0). */
ORC_API void orc_normalize_patterns(const float *K, int I, int P, const int32_t *mask, double *out)
{
    for (int i = 0; i < I; i++) {
        double s = 0.0;
        for (int p = 0; p < P; p++)
            if (mask[p] == 1) s += K[(size_t)i * P + p];
        const double f = s != 0.0 ? (double)P / s : 1.0;
        for (int p = 0; p < P; p++) out[(size_t)i * P + p] = K[(size_t)i * P + p] * f;
    }
}

/* Global fluence for true Poisson (the G8 decision; the reference's model carries no photon
 * units): mean photons per pattern over mean slice sum, both over unmasked pixels, the slice sum
 * over V_s (W > 0).  1 when either mean is 0 (no photon or no valid sample: undefined).  No
 * per-pattern scaling.
 * Pinned by: closed forms on constructed slices and patterns, the fallbacks and the scaling laws
 * in K and W (test_oracle_fluence.py); the device's value is compared with it on every
 * true-Poisson GPU iteration test. */
ORC_API double orc_fluence(const double *W, const double *K, const int32_t *mask, int S, int I,
                           int P)
{
    double sk = 0.0, sw = 0.0;
    for (int i = 0; i < I; i++)
        for (int p = 0; p < P; p++)
            if (mask[p] != 0) sk += K[(size_t)i * P + p];
    for (int s = 0; s < S; s++)
        for (int p = 0; p < P; p++)
            if (mask[p] != 0 && W[(size_t)s * P + p] > 0.0) sw += W[(size_t)s * P + p];
    if (sw <= 0.0 || sk <= 0.0 || S == 0 || I == 0) return 1.0;
    return (sk / I) / (sw / S);
}

/* ------------------------------------------------------------------------------------------ */
/* a7-a10: responsibilities (emc_cuda.cu:1041-1056, 1175-1202, 1253-1260; main.cpp:367-392;      */
/* emc_helper.cpp:579-587).  In place: R (logits, [S][I]) becomes P.  lse[i] written; returns   */
/* the total log-likelihood sum_i lse_i.                                                         */
/* Pinned by: scipy.special.logsumexp, sum_s P = 1, shift invariance (tests).                    */
/* ------------------------------------------------------------------------------------------ */
ORC_API double orc_softmax(double *R, int S, int I, double *lse)
{
    double total = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : total)
    for (int i = 0; i < I; i++) {
        double mx = -1.0e10;
        for (int s = 0; s < S; s++) {
            double v = R[(size_t)s * I + i];
            if (v > mx) mx = v;
        }
        double sum = 0.0;
        for (int s = 0; s < S; s++) {
            double d = R[(size_t)s * I + i] - mx;
            if (d > -1.0e10) sum += exp(d);
        }
        double l = mx + log(sum);
        lse[i] = l;
        total += l;
        for (int s = 0; s < S; s++) R[(size_t)s * I + i] = exp(R[(size_t)s * I + i] - l);
    }
    return total;
}

/* ------------------------------------------------------------------------------------------ */
/* a11-a12: maximize -- weighted slice mean and insertion (emc_atomic.cu:33-132, 183-200)        */
/*   T_s = sum_i P[s,i];  U[s,p] = sum_i K[i,p] P[s,i] / phi[s,i] (mask != 0) else -1;            */
/*   Ubar = U / T_s if T_s > 1e-10.  Insert only if T_s > 1e-10, for p with mask == 1, Ubar >= 0: */
/*     nearest:   v = round(u), 0 <= v < L  ->  Mnum[v] += T_s Ubar, Wt[v] += T_s                 */
/*     trilinear: 8 corners with weights g ->  Mnum[c] += g T_s Ubar, Wt[c] += g T_s              */
/* Mnum and Wt are accumulated (caller zeroes them).  T (may be NULL) receives T_s.              */
/* Pinned by: one-hot re-insertion, conservation of sum Wt, adjoint identity with orc_slices.    */
/* ------------------------------------------------------------------------------------------ */
ORC_API void orc_mstep(const double *Presp, const double *K, const int32_t *mask, int S, int I,
                       int P, const double *phi, const float *quat, const float *X, const float *Y,
                       const float *Z, int L, int interp, double *Mnum, double *Wt, double *Tout)
{
    const float c = 0.5f * (float)L - 0.5f;
    const size_t N3 = (size_t)L * L * L;
    int nth = 1;
#ifdef _OPENMP
    nth = omp_get_max_threads();
#endif
    double *accM = (double *)calloc(N3 * (size_t)nth, sizeof(double));
    double *accW = (double *)calloc(N3 * (size_t)nth, sizeof(double));
#pragma omp parallel
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        double *aM = accM + N3 * t, *aW = accW + N3 * t;
        double *U = (double *)malloc(sizeof(double) * (size_t)P);
#pragma omp for schedule(dynamic, 4)
        for (int s = 0; s < S; s++) {
            double T = 0.0;
            for (int i = 0; i < I; i++) T += Presp[(size_t)s * I + i];
            if (Tout) Tout[s] = T;
            for (int p = 0; p < P; p++) U[p] = 0.0;
            for (int i = 0; i < I; i++) {
                const double pr = Presp[(size_t)s * I + i];
                if (pr == 0.0) continue;
                const double a = pr / (phi ? phi[(size_t)s * I + i] : 1.0);
                const double *Ki = K + (size_t)i * P;
                for (int p = 0; p < P; p++) U[p] += Ki[p] * a;
            }
            for (int p = 0; p < P; p++)
                if (mask[p] == 0) U[p] = -1.0;
            if (!(T > 1e-10)) continue;
            for (int p = 0; p < P; p++)
                if (mask[p] != 0) U[p] /= T;
            float m[9];
            orc_rotmat(quat + 4 * (size_t)s, m);
            for (int p = 0; p < P; p++) {
                if (mask[p] != 1 || !(U[p] >= 0.0)) continue;
                float u[3];
                rot_point(m, X[p], Y[p], Z[p], c, u);
                if (interp == 0) {
                    int vx = (int)roundf(u[0]), vy = (int)roundf(u[1]), vz = (int)roundf(u[2]);
                    if (vx >= 0 && vx < L && vy >= 0 && vy < L && vz >= 0 && vz < L) {
                        size_t v = ((size_t)vz * L + vy) * L + vx;
                        aM[v] += T * U[p];
                        aW[v] += T;
                    }
                } else {
                    int b[3];
                    float f[3];
                    if (!tri_cell(u, L, b, f)) continue;
                    for (int dz = 0; dz < 2; dz++)
                        for (int dy = 0; dy < 2; dy++)
                            for (int dx = 0; dx < 2; dx++) {
                                double g = (dx ? (double)f[0] : 1.0 - (double)f[0]) *
                                           (dy ? (double)f[1] : 1.0 - (double)f[1]) *
                                           (dz ? (double)f[2] : 1.0 - (double)f[2]);
                                size_t v = ((size_t)(b[2] + dz) * L + (b[1] + dy)) * L + (b[0] + dx);
                                aM[v] += g * T * U[p];
                                aW[v] += g * T;
                            }
                }
            }
        }
        free(U);
    }
    for (int t = 0; t < nth; t++)
        for (size_t v = 0; v < N3; v++) {
            Mnum[v] += accM[N3 * t + v];
            Wt[v] += accW[N3 * t + v];
        }
    free(accM);
    free(accW);
}

/* ------------------------------------------------------------------------------------------ */
/* a13: divide by weight (emc_cuda.cu:758-786): M = Wt > 0 ? Mnum/Wt : -1                        */
/* ------------------------------------------------------------------------------------------ */
ORC_API void orc_compress(double *M, const double *Mnum, const double *Wt, long n)
{
    for (long v = 0; v < n; v++) M[v] = Wt[v] > 0.0 ? Mnum[v] / Wt[v] : -1.0;
}

/* ------------------------------------------------------------------------------------------ */
/* a15: blur -- the reference's intended masked Fourier-Gaussian (emc_cuda.cu:1404-1470):      */
/*   mask = M >= 0; M<0 -> 0; multiply the DFT by exp(-2 pi^2 sigma^2 |d|^2 / L^2),             */
/*   d = min(k, L-k) per axis; inverse DFT; masked voxels keep their (negative) value: -1 after  */
/*   the divide, -1/avg once normalised (the reference's missing-voxel marker either way).        */
/* Restated as the equivalent separable periodic convolution with taps                           */
/*   g[d] = (1/L) sum_k exp(-2 pi^2 sigma^2 min(k,L-k)^2 / L^2) cos(2 pi k d / L).                */
/* Pinned by: numpy FFT of the same multiplier, delta response, constants preserved (tests).    */
/* ------------------------------------------------------------------------------------------ */
ORC_API void orc_blur_taps(int L, double sigma, double *g)
{
    for (int d = 0; d < L; d++) {
        double acc = 0.0;
        for (int k = 0; k < L; k++) {
            int kk = k < L - k ? k : L - k;
            acc += exp(-2.0 * M_PI * M_PI * sigma * sigma * (double)kk * kk / ((double)L * L)) *
                   cos(2.0 * M_PI * (double)k * d / L);
        }
        g[d] = acc / L;
    }
}

ORC_API void orc_blur(double *M, int L, double sigma)
{
    const size_t N3 = (size_t)L * L * L;
    double *g = (double *)malloc(sizeof(double) * L);
    double *A = (double *)malloc(sizeof(double) * N3);
    double *B = (double *)malloc(sizeof(double) * N3);
    unsigned char *msk = (unsigned char *)malloc(N3);
    orc_blur_taps(L, sigma, g);
    for (size_t v = 0; v < N3; v++) {
        msk[v] = M[v] >= 0.0;
        A[v] = M[v] < 0.0 ? 0.0 : M[v];
    }
    /* axis x, then y, then z; out[x] = sum_t g[(x - t) mod L] in[t] */
    for (int axis = 0; axis < 3; axis++) {
        const size_t stride = axis == 0 ? 1 : (axis == 1 ? (size_t)L : (size_t)L * L);
#pragma omp parallel for schedule(static)
        for (long line = 0; line < (long)L * L; line++) {
            size_t base;
            int a = (int)(line % L), b = (int)(line / L);
            if (axis == 0) base = ((size_t)b * L + a) * L;
            else if (axis == 1) base = (size_t)b * L * L + a;
            else base = (size_t)b * L + a;
            for (int x = 0; x < L; x++) {
                double acc = 0.0;
                for (int t = 0; t < L; t++) acc += g[((x - t) % L + L) % L] * A[base + t * stride];
                B[base + x * stride] = acc;
            }
        }
        double *tmp = A; A = B; B = tmp;
    }
    for (size_t v = 0; v < N3; v++) M[v] = msk[v] ? A[v] : M[v];
    free(g); free(A); free(B); free(msk);
}

/* ------------------------------------------------------------------------------------------ */
/* a14: normalise (emc_cuda.cu:639-661, 788-798): avg = mean{M > 0}; M *= 1/avg on every voxel. */
/* Returns the count of positive voxels; avg written to *avg (0 and no scaling if none).        */
/* ------------------------------------------------------------------------------------------ */
ORC_API long orc_normalize(double *M, long n, double *avg)
{
    double sum = 0.0;
    long cnt = 0;
    for (long v = 0; v < n; v++)
        if (M[v] > 0.0) { sum += M[v]; cnt++; }
    if (cnt == 0) { *avg = 0.0; return 0; }
    double a = sum / (double)cnt;
    *avg = a;
    double inv = 1.0 / a;
    for (long v = 0; v < n; v++) M[v] *= inv;
    return cnt;
}

/* a14 as the iteration applies it after the divide: avg = the mean of M over the voxels that
 * received weight (Wt > 0, zeros included) -- the set the reference's positive voxels would be if
 * no responsibility underflowed; M *= 1/avg on every voxel.  Returns the voxel count (0: none, no
 * scaling).  Pinned by closed forms and by its equality with orc_normalize when no supported voxel
 * is zero (test_oracle_stages.py). */
ORC_API long orc_normalize_support(double *M, const double *Wt, long n, double *avg)
{
    double sum = 0.0;
    long cnt = 0;
    for (long v = 0; v < n; v++)
        if (Wt[v] > 0.0) { sum += M[v]; cnt++; }
    *avg = cnt ? sum / (double)cnt : 0.0;
    if (cnt == 0 || !(*avg > 0.0)) return 0;
    const double inv = 1.0 / *avg;
    for (long v = 0; v < n; v++) M[v] *= inv;
    return cnt;
}

/* a6: annealed width (main.cpp:251) */
ORC_API double orc_sigma(int iteration, double s_start, double s_final, int half_life)
{
    return (double)(float)(s_final + (s_start - s_final) *
                                         exp(-iteration / (float)half_life * log(2.)));
}

/* ------------------------------------------------------------------------------------------ */
/* a16: one full iteration (main.cpp:239-480 order of operations), FP64.  scaling != 0 applies  */
/* the per (s, i) intensity scaling of known_intensity = 0 to the chi^2 logits and the M-step.  */
/* model (in/out, L^3), Wt_out (L^3, may be NULL), P_out ([S][I], may be NULL), lse (I, may be  */
/* NULL).  Returns total logL (sum_i lse_i).  stats (may be NULL): [0]=phi0 [1]=model average    */
/* before normalisation.                                                                         */
/* ------------------------------------------------------------------------------------------ */
ORC_API double orc_iterate(double *model, int L, const float *quat, const double *wrot, int S,
                           const double *K, const int32_t *mask, int I, const float *X,
                           const float *Y, const float *Z, int interp, int diff, double sigma,
                           double fluence, double blur_sigma, int scaling, double *Wt_out,
                           double *P_out, double *lse_out, double *stats)
{
    const int P = L * L;
    const size_t N3 = (size_t)L * L * L;
    double *W = (double *)malloc(sizeof(double) * (size_t)S * P);
    double *R = P_out ? P_out : (double *)malloc(sizeof(double) * (size_t)S * I);
    double *lse = lse_out ? lse_out : (double *)malloc(sizeof(double) * I);
    double *Mnum = (double *)calloc(N3, sizeof(double));
    double *Wt = (double *)calloc(N3, sizeof(double));
    orc_slices(model, L, quat, S, X, Y, Z, P, interp, W);
    double phi0 = 1.0;
    if (diff == 3) phi0 = fluence > 0.0 ? fluence : orc_fluence(W, K, mask, S, I, P);
    double *phi = NULL;
    if (scaling) {
        phi = (double *)malloc(sizeof(double) * (size_t)S * I);
        orc_scaling(W, K, mask, S, I, P, phi);
    }
    orc_logl(W, K, mask, S, I, P, diff, sigma, wrot, phi, phi0, R, NULL);
    double total = orc_softmax(R, S, I, lse);
    orc_mstep(R, K, mask, S, I, P, phi, quat, X, Y, Z, L, interp, Mnum, Wt, NULL);
    free(phi);
    /* divide, normalise over the voxels with weight, then the blur */
    orc_compress(model, Mnum, Wt, (long)N3);
    double avg = 0.0;
    orc_normalize_support(model, Wt, (long)N3, &avg);
    if (blur_sigma > 0.0) orc_blur(model, L, blur_sigma);
    if (stats) { stats[0] = phi0; stats[1] = avg; }
    if (Wt_out) memcpy(Wt_out, Wt, sizeof(double) * N3);
    free(W);
    if (!P_out) free(R);
    if (!lse_out) free(lse);
    free(Mnum);
    free(Wt);
    return total;
}
