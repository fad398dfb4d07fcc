/*
 * oracle.c -- FP64 CPU restatement of ChaNGa's gravity force-evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY: linked by nothing in the product. Used by tests/,
 * by __graft_entry__.smoke() as the checker and by bench.py's cpu_baseline.
 *
 * Formulation. Radius-scaled reduced moments (the FMOMR of moments.h:64-68)
 * are the order-2..4 symmetric trace-free (STF) tensors of the mass
 * distribution about its centre of mass. This file manipulates them through
 * their "complete" components C(a,b,c) (a x's, b y's, c z's), obtained from
 * the stored ones by tracelessness, and through full 3^n tensors where a
 * projection onto the STF subspace is needed (P2M, M2M). The potential of an
 * STF multipole is phi = -sum_n (2n-1)!!/n! (Q_n . r^n) / r^(2n+1); the Ewald
 * real-space kernels use the g-recursion g_(n+1) = -(1/r) d g_n / dr. The
 * reference's hand-scheduled arithmetic (moments.c, Ewald.cpp) is followed in
 * meaning, not in operation order; the pinning tests compare against
 * moments.c compiled from the reference.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* -------------------------------------------------------------------------
 * Index bookkeeping
 * ------------------------------------------------------------------------- */

/* stored component k -> index counts (a,b,c); order of moments.h:64-68 */
static const int RED[ORC_NMOM][3] = {
    {0, 0, 0},
    {2, 0, 0}, {0, 2, 0}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1},
    {3, 0, 0}, {1, 2, 0}, {2, 1, 0}, {0, 3, 0}, {2, 0, 1}, {0, 2, 1}, {1, 1, 1},
    {4, 0, 0}, {1, 3, 0}, {3, 1, 0}, {0, 4, 0}, {3, 0, 1}, {0, 3, 1}, {2, 2, 0},
    {2, 1, 1}, {1, 2, 1}};

/* complete components are kept in C[a][b][c] (a+b+c <= 4) */
typedef struct {
    double v[5][5][5];
} cmom;

static int red_index(int a, int b, int c) {
    for (int k = 0; k < ORC_NMOM; ++k)
        if (RED[k][0] == a && RED[k][1] == b && RED[k][2] == c) return k;
    return -1;
}

/* complete STF component from stored ones: tracelessness in z,
   T(a,b,c) = -T(a+2,b,c-2) - T(a,b+2,c-2) */
static double stf_comp(const double *q, int a, int b, int c) {
    if (c >= 2) return -(stf_comp(q, a + 2, b, c - 2) + stf_comp(q, a, b + 2, c - 2));
    int k = red_index(a, b, c);
    return k < 0 ? 0.0 : q[k];
}

/* complete components, order n multiplied by scale^n */
static void to_complete(const double *q, double scale, cmom *C) {
    memset(C, 0, sizeof(*C));
    for (int n = 0; n <= 4; ++n) {
        if (n == 1) continue;
        double s = pow(scale, n);
        for (int a = n; a >= 0; --a)
            for (int b = n - a; b >= 0; --b) {
                int c = n - a - b;
                C->v[a][b][c] = s * stf_comp(q, a, b, c);
            }
    }
}

static double fact(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; ++i) f *= i;
    return f;
}
static double multinom(int a, int b, int c) { return fact(a + b + c) / (fact(a) * fact(b) * fact(c)); }
static double ipow(double x, int n) {
    double r = 1.0;
    for (int i = 0; i < n; ++i) r *= x;
    return r;
}

/* P_n = Q_n . r^n and V_n = Q_n . r^(n-1) (a vector) */
static void contract(const cmom *C, int n, const double r[3], double *P, double V[3]) {
    double p = 0.0;
    for (int a = n; a >= 0; --a)
        for (int b = n - a; b >= 0; --b) {
            int c = n - a - b;
            p += multinom(a, b, c) * C->v[a][b][c] * ipow(r[0], a) * ipow(r[1], b) * ipow(r[2], c);
        }
    *P = p;
    V[0] = V[1] = V[2] = 0.0;
    if (n == 0) return;
    int m = n - 1;
    for (int a = m; a >= 0; --a)
        for (int b = m - a; b >= 0; --b) {
            int c = m - a - b;
            double w = multinom(a, b, c) * ipow(r[0], a) * ipow(r[1], b) * ipow(r[2], c);
            V[0] += w * C->v[a + 1][b][c];
            V[1] += w * C->v[a][b + 1][c];
            V[2] += w * C->v[a][b][c + 1];
        }
}

/* -------------------------------------------------------------------------
 * Full rank-n tensors (n <= 4) and the STF projection
 * ------------------------------------------------------------------------- */
static int p3(int n) { return n == 0 ? 1 : 3 * p3(n - 1); }

static void digits(int idx, int n, int *d) {
    for (int p = n - 1; p >= 0; --p) {
        d[p] = idx % 3;
        idx /= 3;
    }
}
static int flat(const int *d, int n) {
    int idx = 0;
    for (int p = 0; p < n; ++p) idx = idx * 3 + d[p];
    return idx;
}
static double kd(int i, int j) { return i == j ? 1.0 : 0.0; }

/* symmetric trace-free part of a symmetric rank-n tensor, n in {2,3,4} */
static void stf_project(const double *T, int n, double *S) {
    int d[4], e[4];
    if (n == 2) {
        double tr = T[0] + T[4] + T[8];
        for (int i = 0; i < 9; ++i) {
            digits(i, 2, d);
            S[i] = T[i] - kd(d[0], d[1]) * tr / 3.0;
        }
    } else if (n == 3) {
        double V[3];
        for (int i = 0; i < 3; ++i) {
            V[i] = 0;
            for (int k = 0; k < 3; ++k) {
                e[0] = i; e[1] = k; e[2] = k;
                V[i] += T[flat(e, 3)];
            }
        }
        for (int i = 0; i < 27; ++i) {
            digits(i, 3, d);
            S[i] = T[i] - (kd(d[0], d[1]) * V[d[2]] + kd(d[0], d[2]) * V[d[1]] +
                           kd(d[1], d[2]) * V[d[0]]) / 5.0;
        }
    } else {
        double W[9], w = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = 0;
                for (int k = 0; k < 3; ++k) {
                    e[0] = i; e[1] = j; e[2] = k; e[3] = k;
                    s += T[flat(e, 4)];
                }
                W[i * 3 + j] = s;
            }
        w = W[0] + W[4] + W[8];
        for (int i = 0; i < 81; ++i) {
            digits(i, 4, d);
            int a = d[0], b = d[1], c = d[2], f = d[3];
            double t2 = kd(a, b) * W[c * 3 + f] + kd(a, c) * W[b * 3 + f] + kd(a, f) * W[b * 3 + c] +
                        kd(b, c) * W[a * 3 + f] + kd(b, f) * W[a * 3 + c] + kd(c, f) * W[a * 3 + b];
            double t4 = kd(a, b) * kd(c, f) + kd(a, c) * kd(b, f) + kd(a, f) * kd(b, c);
            S[i] = T[i] - t2 / 7.0 + w * t4 / 35.0;
        }
    }
}

/* read stored components of order n out of a full STF tensor */
static void store_order(const double *S, int n, double *q) {
    for (int k = 0; k < ORC_NMOM; ++k) {
        int a = RED[k][0], b = RED[k][1], c = RED[k][2];
        if (a + b + c != n) continue;
        int d[4], p = 0;
        for (int i = 0; i < a; ++i) d[p++] = 0;
        for (int i = 0; i < b; ++i) d[p++] = 1;
        for (int i = 0; i < c; ++i) d[p++] = 2;
        q[k] = S[flat(d, n)];
    }
}

/* full tensor of stored order n (scaled units) */
static void full_from_stored(const double *q, int n, double *T) {
    int d[4];
    for (int i = 0; i < p3(n); ++i) {
        digits(i, n, d);
        int cnt[3] = {0, 0, 0};
        for (int p = 0; p < n; ++p) cnt[d[p]]++;
        T[i] = stf_comp(q, cnt[0], cnt[1], cnt[2]);
    }
}

/* -------------------------------------------------------------------------
 * Spline softening (gravity.h:147-182; Hernquist & Katz 1989, ApJS 70, 419,
 * eqs. A1-A2). a is the potential kernel, b = -(1/r) da/dr.
 * ------------------------------------------------------------------------- */
void orc_spline(double r2, double twoh, double *a, double *b) {
    double r = sqrt(r2);
    if (r < twoh) {
        double dih = 2.0 / twoh; /* 1/h */
        double u = r * dih;
        double u2 = u * u;
        if (u < 1.0) {
            *a = dih * (7.0 / 5.0 + u2 * (-2.0 / 3.0 + u2 * (3.0 / 10.0 - u / 10.0)));
            *b = dih * dih * dih * (4.0 / 3.0 + u2 * (-6.0 / 5.0 + u / 2.0));
        } else {
            double ir = 1.0 / r;
            *a = -ir / 15.0 +
                 dih * (8.0 / 5.0 + u2 * (-4.0 / 3.0 + u * (1.0 + u * (-3.0 / 10.0 + u / 30.0))));
            *b = -ir * ir * ir / 15.0 +
                 dih * dih * dih * (8.0 / 3.0 + u * (-3.0 + u * (6.0 / 5.0 - u / 6.0)));
        }
    } else {
        *a = 1.0 / r;
        *b = (*a) * (*a) * (*a);
    }
}

/* -------------------------------------------------------------------------
 * Multipole algebra
 * ------------------------------------------------------------------------- */
double orc_mom_make(double q[ORC_NMOM], double m, double u, double x, double y, double z) {
    double r[3] = {x / u, y / u, z / u};
    double T[81], S[81];
    int d[4];
    memset(q, 0, sizeof(double) * ORC_NMOM);
    q[0] = m;
    for (int n = 2; n <= 4; ++n) {
        for (int i = 0; i < p3(n); ++i) {
            digits(i, n, d);
            double t = m;
            for (int p = 0; p < n; ++p) t *= r[d[p]];
            T[i] = t;
        }
        stf_project(T, n, S);
        store_order(S, n, q);
    }
    return r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
}

void orc_mom_shift(double q[ORC_NMOM], double u, double dx, double dy, double dz) {
    double d[3] = {dx / u, dy / u, dz / u};
    double Q2[9], Q3[27], T[81], S[81];
    double m = q[0];
    double out[ORC_NMOM];
    int id[4], sub[4];
    full_from_stored(q, 2, Q2);
    full_from_stored(q, 3, Q3);
    memcpy(out, q, sizeof(out));
    /* Q'_n = STF( sum over subsets P of the index positions, |P| = k, of
       Q_k(indices in P) * prod_{p not in P} d ), k in {0, 2, ..., n}:
       the binomial shift of raw moments; STF parts of lower orders suffice
       because traces vanish under the outer projection. Q_1 = 0 about the
       old centre of mass. */
    for (int n = 2; n <= 4; ++n) {
        for (int i = 0; i < p3(n); ++i) {
            digits(i, n, id);
            double acc = 0.0;
            for (int mask = 0; mask < (1 << n); ++mask) {
                int k = __builtin_popcount(mask);
                if (k == 1 || k == n) continue; /* dipole zero; k==n is Q_n itself */
                double t;
                if (k == 0) {
                    t = m;
                } else {
                    int np = 0;
                    for (int p = 0; p < n; ++p)
                        if (mask & (1 << p)) sub[np++] = id[p];
                    t = (k == 2) ? Q2[flat(sub, 2)] : Q3[flat(sub, 3)];
                }
                for (int p = 0; p < n; ++p)
                    if (!(mask & (1 << p))) t *= d[id[p]];
                acc += t;
            }
            T[i] = acc;
        }
        stf_project(T, n, S);
        double add[ORC_NMOM];
        memset(add, 0, sizeof(add));
        store_order(S, n, add);
        for (int k = 0; k < ORC_NMOM; ++k) {
            int o = RED[k][0] + RED[k][1] + RED[k][2];
            if (o == n) out[k] = q[k] + add[k];
        }
    }
    memcpy(q, out, sizeof(out));
}

void orc_mom_rescale(double q[ORC_NMOM], double unew, double uold) {
    double f = uold / unew;
    for (int k = 1; k < ORC_NMOM; ++k) {
        int o = RED[k][0] + RED[k][1] + RED[k][2];
        q[k] *= ipow(f, o);
    }
}

void orc_mom_scaled_add(double qr[ORC_NMOM], double ur, const double qa[ORC_NMOM], double ua) {
    double f = ua / ur;
    qr[0] += qa[0];
    for (int k = 1; k < ORC_NMOM; ++k) {
        int o = RED[k][0] + RED[k][1] + RED[k][2];
        qr[k] += ipow(f, o) * qa[k];
    }
}

/* MOMC order (moments.h:42-50): the 22 stored ones, then
   zz, xzz, yzz, zzz, xxzz, xyzz, xzzz, yyzz, yzzz, zzzz */
void orc_mom_to_momc(const double q[ORC_NMOM], double mc[32]) {
    static const int EXTRA[10][3] = {{0, 0, 2}, {1, 0, 2}, {0, 1, 2}, {0, 0, 3}, {2, 0, 2},
                                     {1, 1, 2}, {1, 0, 3}, {0, 2, 2}, {0, 1, 3}, {0, 0, 4}};
    for (int k = 0; k < ORC_NMOM; ++k) mc[k] = q[k];
    for (int k = 0; k < 10; ++k) mc[22 + k] = stf_comp(q, EXTRA[k][0], EXTRA[k][1], EXTRA[k][2]);
}

static void eval_complete(const cmom *C, const double r[3], double *pot, double a[3]) {
    /* (2n-1)!!/n! */
    static const double cn[5] = {1.0, 0.0, 1.5, 2.5, 4.375};
    double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    double ir2 = 1.0 / r2;
    double ir = sqrt(ir2);
    double P, V[3];
    double irn = ir; /* 1/r^(2n+1) */
    for (int n = 0; n <= 4; ++n) {
        if (n != 1) {
            contract(C, n, r, &P, V);
            *pot -= cn[n] * P * irn;
            for (int i = 0; i < 3; ++i)
                a[i] += cn[n] * (n * V[i] * irn - (2 * n + 1) * P * r[i] * irn * ir2);
        }
        irn *= ir2;
    }
}

void orc_mom_eval(const double q[ORC_NMOM], double u, double x, double y, double z,
                  double *pot, double a[3]) {
    cmom C;
    double r[3] = {x, y, z};
    to_complete(q, u, &C);
    eval_complete(&C, r, pot, a);
}

/* -------------------------------------------------------------------------
 * Tree moments build
 * ------------------------------------------------------------------------- */
static void box_reset(double *lo, double *hi) {
    for (int i = 0; i < 3; ++i) {
        lo[i] = HUGE_VAL;
        hi[i] = -HUGE_VAL;
    }
}

void orc_build_moments(int n_nodes, const int32_t *node_first, const int32_t *node_count,
                       const int32_t *node_child, const double *node_box_radius,
                       const double *pos, const double *mass, const double *soft,
                       double *geo, double *mom) {
    /* children have larger indices than parents: a reverse sweep is a
       post-order traversal */
    for (int n = n_nodes - 1; n >= 0; --n) {
        double *g = geo + (size_t)n * 12;
        double *q = mom + (size_t)n * ORC_NMOM;
        int c0 = node_child[2 * n], c1 = node_child[2 * n + 1];
        memset(q, 0, sizeof(double) * ORC_NMOM);
        if (c0 < 0) {
            /* bucket: makeBucket (GenericTreeNode.h:219-257) */
            int f = node_first[n], cnt = node_count[n];
            double M = 0, sx = 0, sy = 0, sz = 0, ss = 0;
            box_reset(g + 6, g + 9);
            for (int i = f; i < f + cnt; ++i) {
                M += mass[i];
                sx += mass[i] * pos[3 * i];
                sy += mass[i] * pos[3 * i + 1];
                sz += mass[i] * pos[3 * i + 2];
                ss += mass[i] * soft[i];
                for (int k = 0; k < 3; ++k) {
                    if (pos[3 * i + k] < g[6 + k]) g[6 + k] = pos[3 * i + k];
                    if (pos[3 * i + k] > g[9 + k]) g[9 + k] = pos[3 * i + k];
                }
            }
            g[0] = sx / M; g[1] = sy / M; g[2] = sz / M;
            g[4] = ss / M;
            g[5] = M;
            /* calculateRadiusFarthestParticle, falling back to the initial
               box radius when every particle sits on the centre of mass */
            double r2max = 0.0;
            for (int i = f; i < f + cnt; ++i) {
                double dx = g[0] - pos[3 * i], dy = g[1] - pos[3 * i + 1], dz = g[2] - pos[3 * i + 2];
                double d2 = dx * dx + dy * dy + dz * dz;
                if (d2 > r2max) r2max = d2;
            }
            double u = (cnt > 1 && r2max > 0.0) ? sqrt(r2max) : node_box_radius[n];
            g[3] = u;
            double qp[ORC_NMOM];
            for (int i = f; i < f + cnt; ++i) {
                orc_mom_make(qp, mass[i], u, pos[3 * i] - g[0], pos[3 * i + 1] - g[1],
                             pos[3 * i + 2] - g[2]);
                for (int k = 0; k < ORC_NMOM; ++k) q[k] += qp[k];
            }
        } else {
            /* doneChildren (Compute.cpp:2534-2580): operator+= of both
               children, then calculateRadiusFarthestCorner */
            const double *g0 = geo + (size_t)c0 * 12, *g1 = geo + (size_t)c1 * 12;
            const double *q0 = mom + (size_t)c0 * ORC_NMOM, *q1 = mom + (size_t)c1 * ORC_NMOM;
            double M0 = g0[5], M1 = g1[5], M = M0 + M1;
            g[5] = M;
            g[4] = (M0 * g0[4] + M1 * g1[4]) / M;
            for (int k = 0; k < 3; ++k) g[k] = (M0 * g0[k] + M1 * g1[k]) / M;
            for (int k = 0; k < 3; ++k) {
                g[6 + k] = g0[6 + k] < g1[6 + k] ? g0[6 + k] : g1[6 + k];
                g[9 + k] = g0[9 + k] > g1[9 + k] ? g0[9 + k] : g1[9 + k];
            }
            double a[ORC_NMOM], b[ORC_NMOM];
            memcpy(a, q0, sizeof(a));
            memcpy(b, q1, sizeof(b));
            orc_mom_shift(a, g0[3], g0[0] - g[0], g0[1] - g[1], g0[2] - g[2]);
            orc_mom_shift(b, g1[3], g1[0] - g[0], g1[1] - g[1], g1[2] - g[2]);
            orc_mom_scaled_add(a, g0[3], b, g1[3]);
            double d2 = 0.0;
            for (int k = 0; k < 3; ++k) {
                double e1 = g[k] - g[6 + k], e2 = g[9 + k] - g[k];
                double e = e1 > e2 ? e1 : e2;
                d2 += e * e;
            }
            double u = sqrt(d2);
            orc_mom_rescale(a, u, g0[3]);
            g[3] = u;
            memcpy(q, a, sizeof(a));
        }
    }
}

/* -------------------------------------------------------------------------
 * Forces on interaction lists
 * ------------------------------------------------------------------------- */
/* Space::intersect(box, sphere): squared distance from the centre to the
   box (zero inside) against r^2 (CUDAMoments.cu:137-159 states it). */
static int box_sphere(const double *lo, const double *hi, const double c[3], double rad) {
    double dsq = 0.0;
    for (int k = 0; k < 3; ++k) {
        double e = 0.0;
        if (c[k] < lo[k]) e = lo[k] - c[k];
        else if (c[k] > hi[k]) e = c[k] - hi[k];
        dsq += e * e;
    }
    return dsq <= rad * rad;
}

int orc_open_softening(const double *cg, const double *bg, double ox, double oy, double oz) {
    double c[3] = {cg[0] + ox, cg[1] + oy, cg[2] + oz};
    double rs = 2.0 * cg[4], rb = 2.0 * bg[4];
    double dx = c[0] - bg[0], dy = c[1] - bg[1], dz = c[2] - bg[2];
    if (dx * dx + dy * dy + dz * dz <= (rs + rb) * (rs + rb)) return 1;
    return box_sphere(bg + 6, bg + 9, c, rs);
}

static void decode_offset(int32_t oid, double period, double off[3]) {
    uint32_t code = ((uint32_t)oid >> 22) & 0x1ff;
    off[0] = ((int)(code & 7) - 3) * period;
    off[1] = ((int)((code >> 3) & 7) - 3) * period;
    off[2] = ((int)((code >> 6) & 7) - 3) * period;
}

/* the replica offset with a period per axis (a box that is not a cube, as
   gpuLocalTreeWalk's fperiod / fperiodY / fperiodZ, HostCUDA.cu:789-791) */
static void decode_offset3(int32_t oid, const double period[3], double off[3]) {
    uint32_t code = ((uint32_t)oid >> 22) & 0x1ff;
    off[0] = ((int)(code & 7) - 3) * period[0];
    off[1] = ((int)((code >> 3) & 7) - 3) * period[1];
    off[2] = ((int)((code >> 6) & 7) - 3) * period[2];
}

/* partBucketForce for one source (position already offset) */
static int64_t pp_source(const double sp[3], double sm, double ss, int first, int count,
                         const double *pos, const double *mass, const double *soft,
                         const int32_t *rung, int active_rung, double *acc, double *pot,
                         double *dtgrav) {
    int64_t n = 0;
    for (int j = first; j < first + count; ++j) {
        if (rung && rung[j] < active_rung) continue;
        ++n;
        double r[3] = {sp[0] - pos[3 * j], sp[1] - pos[3 * j + 1], sp[2] - pos[3 * j + 2]};
        double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        if (r2 == 0.0) continue;
        double a, b;
        orc_spline(r2, ss + soft[j], &a, &b);
        for (int k = 0; k < 3; ++k) acc[3 * j + k] += r[k] * b * sm;
        pot[j] -= sm * a;
        double idt2 = (mass[j] + sm) * b;
        if (idt2 > dtgrav[j]) dtgrav[j] = idt2;
    }
    return n;
}

void orc_forces_lists(int n_buckets, const int32_t *bucket_node,
                      const int32_t *node_first, const int32_t *node_count,
                      const double *geo, const double *mom,
                      const double *pos, const double *mass, const double *soft,
                      const int32_t *rung, int active_rung,
                      const int32_t *cells, const int64_t *cell_marks,
                      const int32_t *parts, const int64_t *part_marks,
                      double period, int soft_flags, int n_threads,
                      double *acc, double *pot, double *dtgrav, int64_t counts[3]) {
    const double p3[3] = {period, period, period};
    orc_forces_lists3(n_buckets, bucket_node, node_first, node_count, geo, mom, pos, mass, soft,
                      rung, active_rung, cells, cell_marks, parts, part_marks, p3, soft_flags,
                      n_threads, acc, pot, dtgrav, counts);
}

void orc_forces_lists3(int n_buckets, const int32_t *bucket_node,
                       const int32_t *node_first, const int32_t *node_count,
                       const double *geo, const double *mom,
                       const double *pos, const double *mass, const double *soft,
                       const int32_t *rung, int active_rung,
                       const int32_t *cells, const int64_t *cell_marks,
                       const int32_t *parts, const int64_t *part_marks,
                       const double period[3], int soft_flags, int n_threads,
                       double *acc, double *pot, double *dtgrav, int64_t counts[3]) {
    int64_t n_pc = 0, n_soft = 0, n_pp = 0;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : n_pc, n_soft, n_pp)
#endif
    for (int b = 0; b < n_buckets; ++b) {
        int bn = bucket_node[b];
        int first = node_first[bn], count = node_count[bn];
        const double *bg = geo + (size_t)bn * 12;
        for (int64_t e = cell_marks[b]; e < cell_marks[b + 1]; ++e) {
            int cn = cells[2 * e];
            double off[3];
            decode_offset3(cells[2 * e + 1], period, off);
            const double *cg = geo + (size_t)cn * 12;
            int softened = soft_flags ? cells[2 * e + 1] < 0
                                      : orc_open_softening(cg, bg, off[0], off[1], off[2]);
            if (softened) {
                /* monopole pseudo-particle (gravity.h:414-421) */
                double sp[3] = {cg[0] + off[0], cg[1] + off[1], cg[2] + off[2]};
                n_soft += pp_source(sp, cg[5], cg[4], first, count, pos, mass, soft, rung,
                                    active_rung, acc, pot, dtgrav);
                continue;
            }
            cmom C;
            to_complete(mom + (size_t)cn * ORC_NMOM, cg[3], &C);
            for (int j = first; j < first + count; ++j) {
                if (rung && rung[j] < active_rung) continue;
                ++n_pc;
                double r[3] = {pos[3 * j] - (cg[0] + off[0]), pos[3 * j + 1] - (cg[1] + off[1]),
                               pos[3 * j + 2] - (cg[2] + off[2])};
                double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                double dir = 1.0 / sqrt(r2);
                eval_complete(&C, r, &pot[j], &acc[3 * j]);
                double idt2 = (mass[j] + cg[5]) * dir * dir * dir;
                if (idt2 > dtgrav[j]) dtgrav[j] = idt2;
            }
        }
        for (int64_t e = part_marks[b]; e < part_marks[b + 1]; ++e) {
            int idx = parts[3 * e], num = parts[3 * e + 2];
            double off[3];
            decode_offset3(parts[3 * e + 1], period, off);
            for (int s = idx; s < idx + num; ++s) {
                double sp[3] = {pos[3 * s] + off[0], pos[3 * s + 1] + off[1], pos[3 * s + 2] + off[2]};
                n_pp += pp_source(sp, mass[s], soft[s], first, count, pos, mass, soft, rung,
                                  active_rung, acc, pot, dtgrav);
            }
        }
    }
    counts[0] = n_pc;
    counts[1] = n_soft;
    counts[2] = n_pp;
}

/* -------------------------------------------------------------------------
 * Ewald
 * ------------------------------------------------------------------------- */
int orc_ewald_table(const double q_root[ORC_NMOM], double u_root, double L, double hcut,
                    double *ewt, int max_n) {
    cmom C;
    to_complete(q_root, u_root, &C);
    double alpha = 2.0 / L;
    int hreps = (int)ceil(hcut);
    int n = 0;
    for (int hx = -hreps; hx <= hreps; ++hx)
        for (int hy = -hreps; hy <= hreps; ++hy)
            for (int hz = -hreps; hz <= hreps; ++hz) {
                int h2 = hx * hx + hy * hy + hz * hz;
                if (h2 == 0 || h2 > hcut * hcut) continue;
                if (n >= max_n) return -1;
                double k[3] = {2 * M_PI / L * hx, 2 * M_PI / L * hy, 2 * M_PI / L * hz};
                double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
                /* (4 pi / V) exp(-k^2 / 4 alpha^2) / k^2 */
                double w = 4 * M_PI / (L * L * L) * exp(-k2 / (4 * alpha * alpha)) / k2;
                /* rho(k) = sum_n (-i)^n / n! Q_n . k^n */
                double P0, P2, P3, P4, V[3];
                contract(&C, 0, k, &P0, V);
                contract(&C, 2, k, &P2, V);
                contract(&C, 3, k, &P3, V);
                contract(&C, 4, k, &P4, V);
                double re = P0 - P2 / 2.0 + P4 / 24.0;
                double im = P3 / 6.0;
                ewt[5 * n + 0] = k[0];
                ewt[5 * n + 1] = k[1];
                ewt[5 * n + 2] = k[2];
                ewt[5 * n + 3] = -w * re;
                ewt[5 * n + 4] = w * im;
                ++n;
            }
    return n;
}

void orc_bucket_ewald(int n_targets, const int32_t *targets, const double *pos,
                      const double root_cm[3], double root_mass,
                      const double q_root[ORC_NMOM], double u_root,
                      const double *ewt, int n_ewt, double L, int n_reps, double ew_cut,
                      int n_threads, double *acc, double *pot, int64_t counts[2]) {
    cmom C;
    to_complete(q_root, u_root, &C);
    C.v[0][0][0] = q_root[0];
    int nEwReps = (int)ceil(ew_cut);
    if (nEwReps < n_reps) nEwReps = n_reps;
    double cut2 = ew_cut * ew_cut * L * L;
    double inner2 = 1.2e-3 * L * L; /* Ewald.cpp:119 */
    double alpha = 2.0 / L, alpha2 = alpha * alpha;
    double k1 = M_PI / (alpha2 * L * L * L);
    double ka = 2.0 * alpha / sqrt(M_PI);
    int64_t n_real = 0;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : n_real)
#endif
    for (int t = 0; t < n_targets; ++t) {
        int j = targets[t];
        double d[3] = {pos[3 * j] - root_cm[0], pos[3 * j + 1] - root_cm[1], pos[3 * j + 2] - root_cm[2]};
        double fPot = root_mass * k1, a[3] = {0, 0, 0};
        for (int ix = -nEwReps; ix <= nEwReps; ++ix)
            for (int iy = -nEwReps; iy <= nEwReps; ++iy)
                for (int iz = -nEwReps; iz <= nEwReps; ++iz) {
                    int hole = abs(ix) <= n_reps && abs(iy) <= n_reps && abs(iz) <= n_reps;
                    double r[3] = {d[0] + ix * L, d[1] + iy * L, d[2] + iz * L};
                    double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                    if (r2 > cut2 && !hole) continue;
                    double g[6];
                    if (r2 < inner2) {
                        /* two-term series about r = 0 (Ewald.cpp:139-156):
                           g_n = ka (2 alpha^2)^n (alpha^2 r^2/(2n+3) - 1/(2n+1)) */
                        double s = ka, x = alpha2 * r2;
                        for (int n = 0; n < 6; ++n) {
                            g[n] = s * (x / (2 * n + 3) - 1.0 / (2 * n + 1));
                            s *= 2 * alpha2;
                        }
                    } else {
                        double r1 = sqrt(r2), ir2 = 1.0 / r2;
                        double e = ka * exp(-alpha2 * r2) * ir2;
                        g[0] = (hole ? -erf(alpha * r1) : erfc(alpha * r1)) / r1;
                        double s = 1.0;
                        for (int n = 0; n < 5; ++n) {
                            g[n + 1] = (2 * n + 1) * g[n] * ir2 + s * e;
                            s *= 2 * alpha2;
                        }
                    }
                    /* phi -= sum_n (1/n!) (Q_n . r^n) g_n;
                       a += sum_n (1/n!) [n (Q_n . r^(n-1)) g_n - (Q_n . r^n) g_(n+1) r] */
                    for (int n = 0; n <= 4; ++n) {
                        if (n == 1) continue;
                        double P, V[3], inv = 1.0 / fact(n);
                        contract(&C, n, r, &P, V);
                        fPot -= inv * P * g[n];
                        for (int k = 0; k < 3; ++k) a[k] += inv * (n * V[k] * g[n] - P * g[n + 1] * r[k]);
                    }
                    ++n_real;
                }
        for (int i = 0; i < n_ewt; ++i) {
            const double *e = ewt + 5 * i;
            double hd = e[0] * d[0] + e[1] * d[1] + e[2] * d[2];
            double c = cos(hd), s = sin(hd);
            fPot += e[3] * c + e[4] * s;
            double t = e[3] * s - e[4] * c;
            a[0] += e[0] * t;
            a[1] += e[1] * t;
            a[2] += e[2] * t;
        }
        pot[j] += fPot;
        acc[3 * j] += a[0];
        acc[3 * j + 1] += a[1];
        acc[3 * j + 2] += a[2];
    }
    counts[0] = n_real;
    counts[1] = (int64_t)n_targets * n_ewt;
}

/* -------------------------------------------------------------------------
 * Accuracy yardsticks
 * ------------------------------------------------------------------------- */
void orc_direct_sum(int n, const double *pos, const double *mass, const double *soft,
                    int n_targets, const int32_t *targets, int n_threads,
                    double *acc, double *pot) {
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (int t = 0; t < n_targets; ++t) {
        int j = targets[t];
        double a[3] = {0, 0, 0}, p = 0.0;
        for (int s = 0; s < n; ++s) {
            double r[3] = {pos[3 * s] - pos[3 * j], pos[3 * s + 1] - pos[3 * j + 1],
                           pos[3 * s + 2] - pos[3 * j + 2]};
            double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            if (r2 == 0.0) continue;
            double sa, sb;
            orc_spline(r2, soft[s] + soft[j], &sa, &sb);
            for (int k = 0; k < 3; ++k) a[k] += r[k] * sb * mass[s];
            p -= mass[s] * sa;
        }
        acc[3 * t] = a[0];
        acc[3 * t + 1] = a[1];
        acc[3 * t + 2] = a[2];
        pot[t] = p;
    }
}

void orc_ewald_exact(int n, const double *pos, const double *mass, const double *soft,
                     int n_targets, const int32_t *targets, double L, double alpha,
                     double rcut, int hmax, int n_threads, double *acc, double *pot) {
    const double V = L * L * L;
    const double ka = 2.0 * alpha / sqrt(M_PI);
    double Mtot = 0.0;
    for (int i = 0; i < n; ++i) Mtot += mass[i];
    /* half space of h: hx > 0, or hx == 0 && hy > 0, or hx == hy == 0 && hz > 0 */
    int nh = 0;
    int H = 2 * hmax + 1;
    int *hl = (int *)malloc(sizeof(int) * 3 * H * H * H);
    for (int hx = 0; hx <= hmax; ++hx)
        for (int hy = -hmax; hy <= hmax; ++hy)
            for (int hz = -hmax; hz <= hmax; ++hz) {
                if (hx == 0 && (hy < 0 || (hy == 0 && hz <= 0))) continue;
                if (hx * hx + hy * hy + hz * hz > hmax * hmax) continue;
                hl[3 * nh] = hx;
                hl[3 * nh + 1] = hy;
                hl[3 * nh + 2] = hz;
                ++nh;
            }
    double *Sre = (double *)calloc(nh, sizeof(double));
    double *Sim = (double *)calloc(nh, sizeof(double));
    int nt = 1;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel
    {
#pragma omp single
        nt = omp_get_num_threads();
    }
#endif
    double *part = (double *)calloc((size_t)nt * 2 * nh, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *pre = part + (size_t)tid * 2 * nh, *pim = pre + nh;
        double *cx = (double *)malloc(sizeof(double) * 2 * (hmax + 1));
        double *cy = (double *)malloc(sizeof(double) * 2 * (2 * hmax + 1));
        double *cz = (double *)malloc(sizeof(double) * 2 * (2 * hmax + 1));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int i = 0; i < n; ++i) {
            /* e^{-i k x} for k = 2 pi h / L, h = 0..hmax (x) and -hmax..hmax (y, z) */
            double th[3];
            for (int k = 0; k < 3; ++k) th[k] = -2 * M_PI * pos[3 * i + k] / L;
            for (int h = 0; h <= hmax; ++h) {
                cx[2 * h] = cos(h * th[0]);
                cx[2 * h + 1] = sin(h * th[0]);
            }
            for (int h = -hmax; h <= hmax; ++h) {
                cy[2 * (h + hmax)] = cos(h * th[1]);
                cy[2 * (h + hmax) + 1] = sin(h * th[1]);
                cz[2 * (h + hmax)] = cos(h * th[2]);
                cz[2 * (h + hmax) + 1] = sin(h * th[2]);
            }
            for (int q = 0; q < nh; ++q) {
                const double *a = cx + 2 * hl[3 * q];
                const double *b = cy + 2 * (hl[3 * q + 1] + hmax);
                const double *c = cz + 2 * (hl[3 * q + 2] + hmax);
                double re = a[0] * b[0] - a[1] * b[1], im = a[0] * b[1] + a[1] * b[0];
                double re2 = re * c[0] - im * c[1], im2 = re * c[1] + im * c[0];
                pre[q] += mass[i] * re2;
                pim[q] += mass[i] * im2;
            }
        }
        free(cx);
        free(cy);
        free(cz);
    }
    for (int t = 0; t < nt; ++t)
        for (int q = 0; q < nh; ++q) {
            Sre[q] += part[(size_t)t * 2 * nh + q];
            Sim[q] += part[(size_t)t * 2 * nh + nh + q];
        }
    free(part);

    /* cell grid for real space */
    int nc = (int)floor(L / rcut);
    if (nc < 1) nc = 1;
    if (nc > 64) nc = 64;
    double cs = L / nc;
    int ncell = nc * nc * nc;
    int *cstart = (int *)calloc(ncell + 1, sizeof(int));
    int *cidx = (int *)malloc(sizeof(int) * n);
    int *pc = (int *)malloc(sizeof(int) * n);
    for (int i = 0; i < n; ++i) {
        int c[3];
        for (int k = 0; k < 3; ++k) {
            double x = pos[3 * i + k] - floor(pos[3 * i + k] / L) * L;
            c[k] = (int)(x / cs);
            if (c[k] >= nc) c[k] = nc - 1;
        }
        pc[i] = (c[0] * nc + c[1]) * nc + c[2];
        cstart[pc[i] + 1]++;
    }
    for (int c = 0; c < ncell; ++c) cstart[c + 1] += cstart[c];
    int *fill = (int *)calloc(ncell, sizeof(int));
    for (int i = 0; i < n; ++i) cidx[cstart[pc[i]] + fill[pc[i]]++] = i;
    free(fill);
    int rng = (int)ceil(rcut / cs);
    if (2 * rng + 1 > nc) rng = -1; /* scan every cell */

#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (int t = 0; t < n_targets; ++t) {
        int i = targets[t];
        const double *xi = pos + 3 * i;
        double p = Mtot * M_PI / (alpha * alpha * V);
        double a[3] = {0, 0, 0};
        /* self image energy, real-space part: lim (erfc(ar)/r - 1/r) */
        p -= mass[i] * (-ka);
        int ci[3];
        for (int k = 0; k < 3; ++k) {
            double x = xi[k] - floor(xi[k] / L) * L;
            ci[k] = (int)(x / cs);
            if (ci[k] >= nc) ci[k] = nc - 1;
        }
        int lo = rng < 0 ? 0 : -rng, hi = rng < 0 ? nc - 1 : rng;
        for (int ox = lo; ox <= hi; ++ox)
            for (int oy = lo; oy <= hi; ++oy)
                for (int oz = lo; oz <= hi; ++oz) {
                    int cx = rng < 0 ? ox : ((ci[0] + ox) % nc + nc) % nc;
                    int cy = rng < 0 ? oy : ((ci[1] + oy) % nc + nc) % nc;
                    int cz = rng < 0 ? oz : ((ci[2] + oz) % nc + nc) % nc;
                    int c = (cx * nc + cy) * nc + cz;
                    for (int q = cstart[c]; q < cstart[c + 1]; ++q) {
                        int j = cidx[q];
                        if (j == i) continue;
                        double d[3];
                        for (int k = 0; k < 3; ++k) {
                            d[k] = xi[k] - pos[3 * j + k];
                            d[k] -= L * floor(d[k] / L + 0.5);
                        }
                        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                        if (r2 >= rcut * rcut || r2 == 0.0) continue;
                        double r = sqrt(r2);
                        double g0 = erfc(alpha * r) / r;
                        double g1 = g0 / r2 + ka * exp(-alpha * alpha * r2) / r2;
                        p -= mass[j] * g0;
                        for (int k = 0; k < 3; ++k) a[k] -= mass[j] * g1 * d[k];
                        double twoh = soft[i] + soft[j];
                        if (r < twoh) {
                            double sa, sb;
                            orc_spline(r2, twoh, &sa, &sb);
                            p -= mass[j] * (sa - 1.0 / r);
                            for (int k = 0; k < 3; ++k) a[k] -= mass[j] * (sb - 1.0 / (r2 * r)) * d[k];
                        }
                    }
                }
        /* k space */
        double th[3];
        for (int k = 0; k < 3; ++k) th[k] = 2 * M_PI * xi[k] / L;
        for (int q = 0; q < nh; ++q) {
            double kv[3] = {2 * M_PI / L * hl[3 * q], 2 * M_PI / L * hl[3 * q + 1], 2 * M_PI / L * hl[3 * q + 2]};
            double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
            double w = 2.0 * 4 * M_PI / V * exp(-k2 / (4 * alpha * alpha)) / k2; /* x2: h and -h */
            double ph = hl[3 * q] * th[0] + hl[3 * q + 1] * th[1] + hl[3 * q + 2] * th[2];
            double c = cos(ph), s = sin(ph);
            /* e^{ik.x_i} S(h) */
            double re = c * Sre[q] - s * Sim[q], im = c * Sim[q] + s * Sre[q];
            p -= w * re;
            for (int k = 0; k < 3; ++k) a[k] -= w * kv[k] * im;
        }
        pot[t] = p;
        acc[3 * t] = a[0];
        acc[3 * t + 1] = a[1];
        acc[3 * t + 2] = a[2];
    }
    free(cstart);
    free(cidx);
    free(pc);
    free(hl);
    free(Sre);
    free(Sim);
}

/* -------------------------------------------------------------------------
 * Adjoint of the list evaluation (no reference counterpart).
 *
 * L = sum_i w_i . a_i, with a_i the tree acceleration of orc_forces_lists on
 * fixed lists (softened cells as flagged in the lists). The exact gradient
 * with respect to every particle position and mass:
 *   - target side: d(w_i . a_i)/dx_i = (Hessian of the potential of each list
 *     entry at x_i) . w_i;
 *   - particle sources: pairwise terms;
 *   - cell sources: a cell's term depends on its particles through M, cm and
 *     the moments Q_n about cm. With s = x_k - cm, sum_i w_i . a_ic =
 *     sum_k m_k H_c(s_k), H_c(s) = sum_alpha h_alpha s^alpha, h_alpha =
 *     (-1)^n / alpha! D_alpha, D_alpha = sum_i w_i . grad d^alpha (1/r)(r_i)
 *     (n = |alpha| in {0,2,3,4}). Through cm (dcm/dx_k = m_k/M, dcm/dm_k =
 *     s_k/M) this gives dL/dm_k = Hhat_c(s_k), dL/dx_k = m_k grad Hhat_c(s_k)
 *     with Hhat_c(s) = H_c(s) - U_c . s, U_c = (sum_k m_k grad H_c(s_k) +
 *     sum_i t_ic) / M, t_ic the target-side term of pair (i, c).
 * The polynomials Hhat are summed down the tree (re-expanded exactly about
 * each child's centre of mass) and evaluated at the particles of the
 * buckets. Softening lengths are constants (the mass-weighted softening of
 * a cell is not differentiated; it is constant when its particles share one
 * softening).
 * ------------------------------------------------------------------------- */
#define ORC_TD 7
typedef double orc_tder[ORC_TD][ORC_TD][ORC_TD];

/* T[a][b][c] = d^(a+b+c)(1/|r|) / dx^a dy^b dz^c for a+b+c <= nmax (<= 6):
   differentiating r^2 d_x f + x f = 0 (f = 1/r) by Leibniz' rule gives
   r^2 T(al+e_x) = -(2a+1) x T(al) - 2b y T(al+e_x-e_y) - 2c z T(al+e_x-e_z)
                   - a^2 T(al-e_x) - b(b-1) T(al+e_x-2e_y) - c(c-1) T(al+e_x-2e_z)
   for al = (a,b,c), and the same with the axes permuted. */
static void rinv_derivs(const double r[3], int nmax, orc_tder T) {
    const double x = r[0], y = r[1], z = r[2];
    const double ir2 = 1.0 / (x * x + y * y + z * z);
    T[0][0][0] = sqrt(ir2);
    for (int n = 0; n < nmax; ++n)
        for (int a = n + 1; a >= 0; --a)
            for (int b = n + 1 - a; b >= 0; --b) {
                const int c = n + 1 - a - b;
                double v;
                if (a > 0) { /* raise x from (a-1, b, c) */
                    const int A = a - 1;
                    v = -(2 * A + 1) * x * T[A][b][c];
                    if (b >= 1) v -= 2.0 * b * y * T[a][b - 1][c];
                    if (c >= 1) v -= 2.0 * c * z * T[a][b][c - 1];
                    if (A >= 1) v -= (double)A * A * T[A - 1][b][c];
                    if (b >= 2) v -= (double)b * (b - 1) * T[a][b - 2][c];
                    if (c >= 2) v -= (double)c * (c - 1) * T[a][b][c - 2];
                } else if (b > 0) { /* raise y from (0, b-1, c) */
                    const int B = b - 1;
                    v = -(2 * B + 1) * y * T[0][B][c];
                    if (c >= 1) v -= 2.0 * c * z * T[0][b][c - 1];
                    if (B >= 1) v -= (double)B * B * T[0][B - 1][c];
                    if (c >= 2) v -= (double)c * (c - 1) * T[0][b][c - 2];
                } else { /* raise z from (0, 0, c-1) */
                    const int C = c - 1;
                    v = -(2 * C + 1) * z * T[0][0][C];
                    if (C >= 1) v -= (double)C * C * T[0][0][C - 1];
                }
                T[a][b][c] = v * ir2;
            }
}

void orc_rinv_derivs(const double r[3], int nmax, double *out) {
    orc_tder T;
    memset(T, 0, sizeof(T));
    rinv_derivs(r, nmax, T);
    memcpy(out, T, sizeof(T));
}

/* -(1/r) db/dr of the spline force kernel b (orc_spline): the pairwise
   Jacobian is d(d b)/dd = b I - c d d^T. */
static double spline_c(double r2, double twoh) {
    double r = sqrt(r2);
    if (r < twoh) {
        double dih = 2.0 / twoh, u = r * dih;
        double d5 = dih * dih * dih * dih * dih;
        if (u < 1.0) return d5 * (12.0 / 5.0 - 1.5 * u);
        return -0.2 / (r2 * r2 * r) + d5 * (3.0 / u - 12.0 / 5.0 + 0.5 * u);
    }
    return 3.0 / (r2 * r2 * r);
}

/* monomials of degree <= 4: index and inverse factorials */
static int mono_idx[5][5][5];
static int mono_abc[35][3];
static int mono_n = 0;
static void mono_init(void) {
    if (mono_n) return;
    int k = 0;
    for (int n = 0; n <= 4; ++n)
        for (int a = n; a >= 0; --a)
            for (int b = n - a; b >= 0; --b) {
                int c = n - a - b;
                mono_idx[a][b][c] = k;
                mono_abc[k][0] = a;
                mono_abc[k][1] = b;
                mono_abc[k][2] = c;
                ++k;
            }
    mono_n = k;
}
static double inv_fact3(int a, int b, int c) { return 1.0 / (fact(a) * fact(b) * fact(c)); }

/* per-cell adjoint accumulator: D_alpha (35 monomial slots, orders 0,2,3,4
   used) and V = sum_i t_ic */
#define ORC_ADJ 38

static void atomic_add(double *p, double v) {
#ifdef _OPENMP
#pragma omp atomic
#endif
    *p += v;
}

/* pairwise Jacobian term (b I - c d d^T) w */
static void pair_jw(const double d[3], double b, double c, const double w[3], double out[3]) {
    double dw = d[0] * w[0] + d[1] * w[1] + d[2] * w[2];
    for (int k = 0; k < 3; ++k) out[k] = b * w[k] - c * d[k] * dw;
}

void orc_adjoint_lists(int n_nodes, const int32_t *node_child, int n_buckets,
                       const int32_t *bucket_node, const int32_t *node_first,
                       const int32_t *node_count, const double *geo, const double *mom,
                       const double *pos, const double *mass, const double *soft,
                       const int32_t *cells, const int64_t *cell_marks,
                       const int32_t *parts, const int64_t *part_marks, double period,
                       const double *w, const double *root_extra, int n_threads, double *grad) {
    mono_init();
    double *acc = (double *)calloc((size_t)n_nodes * ORC_ADJ, sizeof(double));
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 8)
#endif
    for (int bk = 0; bk < n_buckets; ++bk) {
        const int bn = bucket_node[bk];
        const int first = node_first[bn], count = node_count[bn];
        /* every term is linear in w: targets with w = 0 contribute nothing */
        int live = 0;
        for (int i = first; i < first + count && !live; ++i)
            live = w[3 * i] != 0.0 || w[3 * i + 1] != 0.0 || w[3 * i + 2] != 0.0;
        if (!live) continue;
        for (int64_t e = cell_marks[bk]; e < cell_marks[bk + 1]; ++e) {
            const int cn = cells[2 * e];
            double off[3];
            decode_offset(cells[2 * e + 1], period, off);
            const double *cg = geo + (size_t)cn * 12;
            double *A = acc + (size_t)cn * ORC_ADJ;
            const double cx = cg[0] + off[0], cy = cg[1] + off[1], cz = cg[2] + off[2];
            if (cells[2 * e + 1] < 0) {
                /* softened: spline monopole pseudo-particle at the cm */
                const double M = cg[5];
                double d0 = 0.0, V[3] = {0, 0, 0};
                for (int i = first; i < first + count; ++i) {
                    const double *wi = w + 3 * i;
                    double d[3] = {cx - pos[3 * i], cy - pos[3 * i + 1], cz - pos[3 * i + 2]};
                    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                    if (r2 == 0.0) continue;
                    double a, b;
                    orc_spline(r2, soft[i] + cg[4], &a, &b);
                    double c = spline_c(r2, soft[i] + cg[4]), jw[3];
                    pair_jw(d, b, c, wi, jw);
                    for (int k = 0; k < 3; ++k) {
                        atomic_add(&grad[4 * i + k], -M * jw[k]);
                        V[k] -= M * jw[k];
                    }
                    d0 += b * (wi[0] * d[0] + wi[1] * d[1] + wi[2] * d[2]);
                }
                atomic_add(&A[0], d0);
                for (int k = 0; k < 3; ++k) atomic_add(&A[35 + k], V[k]);
                continue;
            }
            cmom C;
            to_complete(mom + (size_t)cn * ORC_NMOM, cg[3], &C);
            double Dl[35], V[3] = {0, 0, 0};
            memset(Dl, 0, sizeof(Dl));
            for (int i = first; i < first + count; ++i) {
                const double *wi = w + 3 * i;
                if (wi[0] == 0.0 && wi[1] == 0.0 && wi[2] == 0.0) continue;
                double r[3] = {pos[3 * i] - cx, pos[3 * i + 1] - cy, pos[3 * i + 2] - cz};
                orc_tder T;
                rinv_derivs(r, 6, T);
                /* target side: t_j = sum_n sum_al (-1)^n/al! C_al sum_k w_k T[al+e_j+e_k] */
                double t[3] = {0, 0, 0};
                for (int q = 0; q < 35; ++q) {
                    const int a = mono_abc[q][0], b = mono_abc[q][1], c = mono_abc[q][2];
                    const int n = a + b + c;
                    if (n == 1) continue;
                    const double f = ((n & 1) ? -1.0 : 1.0) * inv_fact3(a, b, c) * C.v[a][b][c];
                    if (f == 0.0) continue;
                    t[0] += f * (wi[0] * T[a + 2][b][c] + wi[1] * T[a + 1][b + 1][c] +
                                 wi[2] * T[a + 1][b][c + 1]);
                    t[1] += f * (wi[0] * T[a + 1][b + 1][c] + wi[1] * T[a][b + 2][c] +
                                 wi[2] * T[a][b + 1][c + 1]);
                    t[2] += f * (wi[0] * T[a + 1][b][c + 1] + wi[1] * T[a][b + 1][c + 1] +
                                 wi[2] * T[a][b][c + 2]);
                }
                for (int k = 0; k < 3; ++k) {
                    atomic_add(&grad[4 * i + k], t[k]);
                    V[k] += t[k];
                }
                /* source side: D_al += sum_k w_k T[al+e_k], |al| in {0,2,3,4} */
                for (int q = 0; q < 35; ++q) {
                    const int a = mono_abc[q][0], b = mono_abc[q][1], c = mono_abc[q][2];
                    if (a + b + c == 1) continue;
                    Dl[q] += wi[0] * T[a + 1][b][c] + wi[1] * T[a][b + 1][c] + wi[2] * T[a][b][c + 1];
                }
            }
            for (int q = 0; q < 35; ++q) atomic_add(&A[q], Dl[q]);
            for (int k = 0; k < 3; ++k) atomic_add(&A[35 + k], V[k]);
        }
        for (int64_t e = part_marks[bk]; e < part_marks[bk + 1]; ++e) {
            const int idx = parts[3 * e], num = parts[3 * e + 2];
            double off[3];
            decode_offset(parts[3 * e + 1], period, off);
            for (int s = idx; s < idx + num; ++s) {
                double gs[4] = {0, 0, 0, 0};
                for (int i = first; i < first + count; ++i) {
                    const double *wi = w + 3 * i;
                    double d[3] = {pos[3 * s] + off[0] - pos[3 * i],
                                   pos[3 * s + 1] + off[1] - pos[3 * i + 1],
                                   pos[3 * s + 2] + off[2] - pos[3 * i + 2]};
                    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                    if (r2 == 0.0) continue;
                    double a, b;
                    orc_spline(r2, soft[i] + soft[s], &a, &b);
                    double c = spline_c(r2, soft[i] + soft[s]), jw[3];
                    pair_jw(d, b, c, wi, jw);
                    for (int k = 0; k < 3; ++k) {
                        atomic_add(&grad[4 * i + k], -mass[s] * jw[k]);
                        gs[k] += mass[s] * jw[k];
                    }
                    gs[3] += b * (wi[0] * d[0] + wi[1] * d[1] + wi[2] * d[2]);
                }
                for (int k = 0; k < 4; ++k) atomic_add(&grad[4 * s + k], gs[k]);
            }
        }
    }

    /* top-down: Hhat per node, summed with the parent's polynomial
       re-expanded about this node's centre of mass */
    double *P = (double *)calloc((size_t)n_nodes * 35, sizeof(double));
    int *parent = (int *)malloc(sizeof(int) * (size_t)n_nodes);
    for (int n = 0; n < n_nodes; ++n) parent[n] = -1;
    for (int n = 0; n < n_nodes; ++n)
        for (int k = 0; k < 2; ++k)
            if (node_child[2 * n + k] >= 0) parent[node_child[2 * n + k]] = n;
    for (int n = 0; n < n_nodes; ++n) { /* parents precede children */
        const double *g = geo + (size_t)n * 12;
        const double *A = acc + (size_t)n * ORC_ADJ;
        double *Pn = P + (size_t)n * 35;
        const double M = g[5];
        cmom C;
        to_complete(mom + (size_t)n * ORC_NMOM, g[3], &C);
        /* H coefficients */
        for (int q = 0; q < 35; ++q) {
            const int a = mono_abc[q][0], b = mono_abc[q][1], c = mono_abc[q][2];
            const int o = a + b + c;
            if (o == 1) continue;
            Pn[q] = ((o & 1) ? -1.0 : 1.0) * inv_fact3(a, b, c) * A[q];
        }
        /* U = (sum_k m_k grad H(s_k) + V) / M; sum_k m_k d_j H(s_k) =
           sum_{|be| in {2,3}} (-1)^(|be|+1) / be! D[be+e_j] C[be] */
        double U[3];
        for (int j = 0; j < 3; ++j) {
            double S = 0.0;
            for (int q = 0; q < 35; ++q) {
                const int a = mono_abc[q][0], b = mono_abc[q][1], c = mono_abc[q][2];
                const int o = a + b + c;
                if (o != 2 && o != 3) continue;
                const int e[3] = {a + (j == 0), b + (j == 1), c + (j == 2)};
                S += (((o + 1) & 1) ? -1.0 : 1.0) * inv_fact3(a, b, c) *
                     A[mono_idx[e[0]][e[1]][e[2]]] * C.v[a][b][c];
            }
            U[j] = M > 0.0 ? (S + A[35 + j]) / M : 0.0;
        }
        Pn[mono_idx[1][0][0]] = -U[0];
        Pn[mono_idx[0][1][0]] = -U[1];
        Pn[mono_idx[0][0][1]] = -U[2];
        if (n == 0 && root_extra)
            for (int q = 0; q < 35; ++q) Pn[q] += root_extra[q];
        const int p = parent[n];
        if (p >= 0) {
            const double *gp = geo + (size_t)p * 12;
            const double dd[3] = {g[0] - gp[0], g[1] - gp[1], g[2] - gp[2]};
            const double *Pp = P + (size_t)p * 35;
            /* Pp(s + dd) = sum_al p_al prod_axis sum_k C(al_axis, k) s^k dd^(al_axis - k) */
            for (int q = 0; q < 35; ++q) {
                if (Pp[q] == 0.0) continue;
                const int a = mono_abc[q][0], b = mono_abc[q][1], c = mono_abc[q][2];
                for (int i = 0; i <= a; ++i)
                    for (int j = 0; j <= b; ++j)
                        for (int k = 0; k <= c; ++k) {
                            double f = Pp[q] * multinom(i, a - i, 0) * multinom(j, b - j, 0) *
                                       multinom(k, c - k, 0) * ipow(dd[0], a - i) *
                                       ipow(dd[1], b - j) * ipow(dd[2], c - k);
                            Pn[mono_idx[i][j][k]] += f;
                        }
            }
        }
    }
    /* buckets: dL/dm_k += P(s_k), dL/dx_k += m_k grad P(s_k) */
    for (int n = 0; n < n_nodes; ++n) {
        if (node_child[2 * n] >= 0) continue;
        const double *g = geo + (size_t)n * 12;
        const double *Pn = P + (size_t)n * 35;
        for (int k = node_first[n]; k < node_first[n] + node_count[n]; ++k) {
            const double s[3] = {pos[3 * k] - g[0], pos[3 * k + 1] - g[1], pos[3 * k + 2] - g[2]};
            double val = 0.0, gr[3] = {0, 0, 0};
            for (int q = 0; q < 35; ++q) {
                const int a = mono_abc[q][0], b = mono_abc[q][1], c = mono_abc[q][2];
                const double pa = ipow(s[0], a), pb = ipow(s[1], b), pc = ipow(s[2], c);
                val += Pn[q] * pa * pb * pc;
                if (a) gr[0] += Pn[q] * a * ipow(s[0], a - 1) * pb * pc;
                if (b) gr[1] += Pn[q] * b * pa * ipow(s[1], b - 1) * pc;
                if (c) gr[2] += Pn[q] * c * pa * pb * ipow(s[2], c - 1);
            }
            for (int j = 0; j < 3; ++j) grad[4 * k + j] += mass[k] * gr[j];
            grad[4 * k + 3] += val;
        }
    }
    free(parent);
    free(P);
    free(acc);
}

/* Exact gradient of L = sum_i w_i . a_i for the softened direct sum:
   dL/dx_k = sum_j J(x_k - x_j) (m_k w_j - m_j w_k), J = b I - c d d^T;
   dL/dm_k = sum_i w_i . g(x_k - x_i), g(d) = d b(|d|). */
void orc_direct_grad(int n, const double *pos, const double *mass, const double *soft,
                     const double *w, int n_threads, double *grad) {
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(static)
#endif
    for (int k = 0; k < n; ++k) {
        double gx[3] = {0, 0, 0}, gm = 0.0;
        for (int j = 0; j < n; ++j) {
            if (j == k) continue;
            double d[3] = {pos[3 * k] - pos[3 * j], pos[3 * k + 1] - pos[3 * j + 1],
                           pos[3 * k + 2] - pos[3 * j + 2]};
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 == 0.0) continue;
            double a, b;
            orc_spline(r2, soft[k] + soft[j], &a, &b);
            double c = spline_c(r2, soft[k] + soft[j]);
            double v[3], jv[3];
            for (int q = 0; q < 3; ++q) v[q] = mass[k] * w[3 * j + q] - mass[j] * w[3 * k + q];
            pair_jw(d, b, c, v, jv);
            for (int q = 0; q < 3; ++q) gx[q] += jv[q];
            gm += b * (w[3 * j] * d[0] + w[3 * j + 1] * d[1] + w[3 * j + 2] * d[2]);
        }
        for (int q = 0; q < 3; ++q) grad[4 * k + q] = gx[q];
        grad[4 * k + 3] = gm;
    }
}

/* -------------------------------------------------------------------------
 * Adjoint of the Ewald correction (no reference counterpart).
 *
 * L_E = sum_t w_t . a_E(x_t - cm_root) with a_E the BucketEwald acceleration
 * of orc_bucket_ewald and the h-table of orc_ewald_table built from the root
 * moments. Per target t (d = x_t - cm):
 *   - target side: the gradient of w . a for every real-space replica term
 *     a = sum_n (1/n!) [n V_n g_n - P_n g_(n+1) r] (V_n = Q_n.r^(n-1),
 *     P_n = Q_n.r^n), with grad g_m = gam_m r, gam_m = -g_(m+1) for the
 *     closed form and 2 B_m for the two-term series g_m = A_m + B_m r^2 (the
 *     series' a is not an exact gradient field, so this is the transposed
 *     Jacobian applied to w, not its derivative along w); and of the k-terms,
 *     sum_h k (k.w)(hC cos + hS sin);
 *   - source side: G_n = dL/dQ_n (symmetric tensors, dL = sum_n G_n . dQ_n):
 *     real space (1/n!) [g_n sum_j alpha_j w_j r^(alpha-e_j) - (w.r) g_(n+1)
 *     r^alpha]; k space through hC, hS (linear in M, Q_2, Q_3, Q_4).
 * Then, as for a cell, H(s) = G_0 + sum_n STF(G_n) . s^n, U = (sum_k m_k
 * grad H(s_k) + sum_t (target-side terms)) / M, Hhat = H - U . s about the
 * root's centre of mass; hroot receives Hhat's 35 monomial coefficients.
 * ------------------------------------------------------------------------- */

/* W = Q_n . (r^(n-2) w): W_i = sum_{j, |be| = n-2} mult(be) C(e_i+e_j+be) w_j r^be */
static void contract_w(const cmom *C, int n, const double r[3], const double w[3], double W[3]) {
    W[0] = W[1] = W[2] = 0.0;
    if (n < 2) return;
    int m = n - 2;
    for (int a = m; a >= 0; --a)
        for (int b = m - a; b >= 0; --b) {
            int c = m - a - b;
            double f = multinom(a, b, c) * ipow(r[0], a) * ipow(r[1], b) * ipow(r[2], c);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    int e[3] = {a, b, c};
                    e[i]++;
                    e[j]++;
                    W[i] += f * w[j] * C->v[e[0]][e[1]][e[2]];
                }
        }
}

void orc_adjoint_ewald(int n, const double *pos, const double *w, const double root_cm[3],
                       double root_mass, const double q_root[ORC_NMOM], double u_root,
                       const double *ewt, int n_ewt, double L, int n_reps, double ew_cut,
                       int n_threads, double *grad, double hroot[35]) {
    mono_init();
    cmom C;
    to_complete(q_root, u_root, &C);
    C.v[0][0][0] = q_root[0];
    int nEwReps = (int)ceil(ew_cut);
    if (nEwReps < n_reps) nEwReps = n_reps;
    const double cut2 = ew_cut * ew_cut * L * L, inner2 = 1.2e-3 * L * L;
    const double alpha = 2.0 / L, alpha2 = alpha * alpha, ka = 2.0 * alpha / sqrt(M_PI);
    double G[35], V[3] = {0, 0, 0};
    double *AB = (double *)calloc(2 * (size_t)(n_ewt > 0 ? n_ewt : 1), sizeof(double));
    memset(G, 0, sizeof(G));
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel
#endif
    {
        double Gl[35], Vl[3] = {0, 0, 0};
        double *ABl = (double *)calloc(2 * (size_t)(n_ewt > 0 ? n_ewt : 1), sizeof(double));
        memset(Gl, 0, sizeof(Gl));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int t = 0; t < n; ++t) {
            const double *wt = w + 3 * t;
            if (wt[0] == 0.0 && wt[1] == 0.0 && wt[2] == 0.0) continue;
            double d[3] = {pos[3 * t] - root_cm[0], pos[3 * t + 1] - root_cm[1],
                           pos[3 * t + 2] - root_cm[2]};
            double tt[3] = {0, 0, 0};
            for (int ix = -nEwReps; ix <= nEwReps; ++ix)
                for (int iy = -nEwReps; iy <= nEwReps; ++iy)
                    for (int iz = -nEwReps; iz <= nEwReps; ++iz) {
                        int hole = abs(ix) <= n_reps && abs(iy) <= n_reps && abs(iz) <= n_reps;
                        double r[3] = {d[0] + ix * L, d[1] + iy * L, d[2] + iz * L};
                        double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                        if (r2 > cut2 && !hole) continue;
                        double rw = r[0] * wt[0] + r[1] * wt[1] + r[2] * wt[2];
                        /* g_m and gam_m = (1/r) dg_m/dr, so grad g_m = gam_m r */
                        double g[7], gam[6], dg[6];
                        if (r2 < inner2) {
                            double s = ka, x = alpha2 * r2;
                            for (int k = 0; k < 7; ++k) {
                                g[k] = s * (x / (2 * k + 3) - 1.0 / (2 * k + 1));
                                if (k < 6) gam[k] = 2.0 * s * alpha2 / (2 * k + 3);
                                s *= 2 * alpha2;
                            }
                        } else {
                            double r1 = sqrt(r2), ir2 = 1.0 / r2;
                            double e = ka * exp(-alpha2 * r2) * ir2;
                            g[0] = (hole ? -erf(alpha * r1) : erfc(alpha * r1)) / r1;
                            double s = 1.0;
                            for (int k = 0; k < 6; ++k) {
                                g[k + 1] = (2 * k + 1) * g[k] * ir2 + s * e;
                                s *= 2 * alpha2;
                            }
                            for (int k = 0; k < 6; ++k) gam[k] = -g[k + 1];
                        }
                        for (int k = 0; k < 6; ++k) dg[k] = gam[k] * rw;
                        for (int nn = 0; nn <= 4; ++nn) {
                            if (nn == 1) continue;
                            double P, Vn[3], W[3], inv = 1.0 / fact(nn);
                            contract(&C, nn, r, &P, Vn);
                            contract_w(&C, nn, r, wt, W);
                            double Vw = Vn[0] * wt[0] + Vn[1] * wt[1] + Vn[2] * wt[2];
                            for (int k = 0; k < 3; ++k)
                                tt[k] += inv * (nn * (nn - 1) * W[k] * g[nn] + nn * gam[nn] * Vw * r[k] -
                                                nn * Vn[k] * g[nn + 1] * rw - P * dg[nn + 1] * r[k] -
                                                P * g[nn + 1] * wt[k]);
                            for (int q = 0; q < 35; ++q) {
                                int a = mono_abc[q][0], b = mono_abc[q][1], c = mono_abc[q][2];
                                if (a + b + c != nn) continue;
                                double s = 0.0;
                                if (a) s += a * wt[0] * ipow(r[0], a - 1) * ipow(r[1], b) * ipow(r[2], c);
                                if (b) s += b * wt[1] * ipow(r[0], a) * ipow(r[1], b - 1) * ipow(r[2], c);
                                if (c) s += c * wt[2] * ipow(r[0], a) * ipow(r[1], b) * ipow(r[2], c - 1);
                                Gl[q] += inv * (g[nn] * s - rw * g[nn + 1] * ipow(r[0], a) *
                                                                  ipow(r[1], b) * ipow(r[2], c));
                            }
                        }
                    }
            for (int i = 0; i < n_ewt; ++i) {
                const double *e = ewt + 5 * i;
                double th = e[0] * d[0] + e[1] * d[1] + e[2] * d[2];
                double c = cos(th), s = sin(th);
                double kw = e[0] * wt[0] + e[1] * wt[1] + e[2] * wt[2];
                double f = kw * (e[3] * c + e[4] * s);
                for (int k = 0; k < 3; ++k) tt[k] += e[k] * f;
                ABl[2 * i] += kw * s;
                ABl[2 * i + 1] -= kw * c;
            }
            for (int k = 0; k < 3; ++k) {
                grad[4 * t + k] += tt[k];
                Vl[k] += tt[k];
            }
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            for (int q = 0; q < 35; ++q) G[q] += Gl[q];
            for (int k = 0; k < 3; ++k) V[k] += Vl[k];
            for (int i = 0; i < 2 * n_ewt; ++i) AB[i] += ABl[i];
        }
        free(ABl);
    }
    /* k space: hC = -w_h (M - Q2.kk/2 + Q4.k^4/24), hS = w_h Q3.k^3/6 */
    for (int i = 0; i < n_ewt; ++i) {
        const double *e = ewt + 5 * i;
        double k2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
        double wh = 4 * M_PI / (L * L * L) * exp(-k2 / (4 * alpha2)) / k2;
        double A = AB[2 * i], B = AB[2 * i + 1];
        for (int q = 0; q < 35; ++q) {
            int a = mono_abc[q][0], b = mono_abc[q][1], c = mono_abc[q][2];
            int o = a + b + c;
            double kp = ipow(e[0], a) * ipow(e[1], b) * ipow(e[2], c);
            if (o == 0) G[q] += -A * wh;
            else if (o == 2) G[q] += A * wh * 0.5 * kp;
            else if (o == 3) G[q] += B * wh / 6.0 * kp;
            else if (o == 4) G[q] += -A * wh / 24.0 * kp;
        }
    }
    free(AB);
    /* H coefficients: h_0 = G_0, h_al = mult(al) STF(G_n)_al */
    memset(hroot, 0, sizeof(double) * 35);
    hroot[0] = G[0];
    for (int nn = 2; nn <= 4; ++nn) {
        double T[81], S[81];
        int dgt[4];
        for (int i = 0; i < p3(nn); ++i) {
            digits(i, nn, dgt);
            int cnt[3] = {0, 0, 0};
            for (int p = 0; p < nn; ++p) cnt[dgt[p]]++;
            T[i] = G[mono_idx[cnt[0]][cnt[1]][cnt[2]]];
        }
        stf_project(T, nn, S);
        for (int q = 0; q < 35; ++q) {
            int a = mono_abc[q][0], b = mono_abc[q][1], c = mono_abc[q][2];
            if (a + b + c != nn) continue;
            int dd[4], p = 0;
            for (int i = 0; i < a; ++i) dd[p++] = 0;
            for (int i = 0; i < b; ++i) dd[p++] = 1;
            for (int i = 0; i < c; ++i) dd[p++] = 2;
            hroot[q] = multinom(a, b, c) * S[flat(dd, nn)];
        }
    }
    /* U = (sum_k m_k grad H(s_k) + V) / M; sum_k m_k d_j H = sum_al h_al al_j C(al - e_j) */
    for (int j = 0; j < 3; ++j) {
        double S = 0.0;
        for (int q = 0; q < 35; ++q) {
            int a = mono_abc[q][0], b = mono_abc[q][1], c = mono_abc[q][2];
            int e[3] = {a, b, c};
            if (e[j] == 0 || a + b + c < 3) continue;
            e[j]--;
            S += hroot[q] * (j == 0 ? a : j == 1 ? b : c) * C.v[e[0]][e[1]][e[2]];
        }
        hroot[mono_idx[j == 0][j == 1][j == 2]] = -(S + V[j]) / root_mass;
    }
}
