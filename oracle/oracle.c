/*
 * CPU oracle for the motion-planning hot path -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off; fmaf() is glibc's correctly rounded
 * fused multiply-add, sqrtf() is correctly rounded).  Every FP32 expression below is written
 * with explicit fmaf / mul / add so that its rounding sequence is fixed by the source.
 *
 * FP32 evaluation rules (the kernels in the HIP library evaluate the same expressions; they
 * drop terms with a constant zero coefficient, which leaves every result bit-identical):
 *   dot3(a,b)        = fmaf(a0,b0, fmaf(a1,b1, a2*b2))
 *   affine row       = fmaf(a0,b0, fmaf(a1,b1, fmaf(a2,b2, t)))
 *   compose(A,B)     : R = A.R*B.R by dot3, t = affine row of A.R with B.t, plus A.t
 *   joint transforms : fixed  T = T_parent o O
 *                      revol. T = (T_parent o O) o Rot(axis, q)   (Rot exact for +-x/y/z axes)
 *                      prism. T = T_parent o (O.R, O.R*(axis*q) + O.t)
 *   constants        : origin / axis entries within 1e-12 of 0 or +-1 are snapped, then
 *                      rounded to float
 *   sin/cos          : orc_sincosf (Cody-Waite pi/2 reduction + Cephes sinf/cosf minimax
 *                      polynomials), not libm
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define MAXL 64
#define MAXDOF 16

enum { JT_ROOT = 0, JT_REV = 1, JT_PRISM = 2, JT_FIXED = 5 };
enum { AX_NONE = -1, AX_X = 0, AX_Y = 1, AX_Z = 2, AX_GENERIC = 3 };

struct orc_robot {
    int L, S, dof, P;
    int jtype[MAXL], parent[MAXL], active[MAXL], active_index[MAXL];
    double origin64[MAXL][16];
    double axis64[MAXL][3];
    float Of[MAXL][12];   /* snapped float origin: R row-major 3x3, then t */
    float axisf[MAXL][3]; /* snapped float axis (unit for generic revolute) */
    int axmode[MAXL], axsign[MAXL];
    float lower[MAXL], upper[MAXL], defaults[MAXL];
    float act_lo[MAXDOF], act_hi[MAXDOF];
    int* sphere_link;
    float* sphere_pos;
    float* sphere_r;
    int* pi_;
    int* pj_;
    float* pthr; /* fl((ri+rj)^2) */
    float* psr;  /* sqrtf(thr) */
    double fk_flops;
};

struct orc_scene {
    int ns, nb, nc, ncap;
    float* sph;  /* [ns][4] */
    float* box;  /* [nb][18] inverse pose m0..m11, min3, max3 */
    float* cyl;  /* [nc][14] inverse pose m0..m11, radius, half height */
    float* cap;  /* [ncap][8] p1(3), v(3), rdv, radius */
    int has_bnd; /* 0: the robot's joint limits; 1: bnd below */
    float bnd[32]; /* lower[16], upper[16] over the active joints (+-inf: unbounded) */
};

/* ----------------------------------------------------------------------------------------- */
/* FP32 helpers                                                                              */
/* ----------------------------------------------------------------------------------------- */
static inline float dot3f(float a0, float b0, float a1, float b1, float a2, float b2) {
    return fmaf(a0, b0, fmaf(a1, b1, a2 * b2));
}
static inline float rowf(float a0, float b0, float a1, float b1, float a2, float b2, float t) {
    return fmaf(a0, b0, fmaf(a1, b1, fmaf(a2, b2, t)));
}

void orc_sincosf(float x, float* s_out, float* c_out) {
    /* Cody-Waite reduction by pi/2 (three-part constant) and the Cephes single precision
     * minimax polynomials for sin/cos on [-pi/4, pi/4] (S. L. Moshier, Cephes Math Library). */
    const float two_over_pi = 0.636619772367581343f;
    const float p1 = 1.5703125f;
    const float p2 = 4.837512969970703125e-4f;
    const float p3 = 7.54978995489188216e-8f;
    float kf = rintf(x * two_over_pi);
    int k = (int)kf;
    float r = fmaf(-kf, p1, x);
    r = fmaf(-kf, p2, r);
    r = fmaf(-kf, p3, r);
    float z = r * r;
    float sp = fmaf(fmaf(fmaf(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f), z * r, r);
    float cp = fmaf(fmaf(fmaf(2.443315711809948e-5f, z, -1.388731625493765e-3f), z, 4.166664568298827e-2f),
                    z * z, fmaf(-0.5f, z, 1.0f));
    switch (k & 3) {
        case 0: *s_out = sp; *c_out = cp; break;
        case 1: *s_out = cp; *c_out = -sp; break;
        case 2: *s_out = -sp; *c_out = -cp; break;
        default: *s_out = -cp; *c_out = sp; break;
    }
}

static double snap(double v) {
    if (fabs(v) < 1e-12) return 0.0;
    if (fabs(v - 1.0) < 1e-12) return 1.0;
    if (fabs(v + 1.0) < 1e-12) return -1.0;
    return v;
}

/* ----------------------------------------------------------------------------------------- */
/* Robot                                                                                      */
/* ----------------------------------------------------------------------------------------- */
orc_robot* orc_robot_create(int L, const int* joint_types, const double* joint_origins, const double* joint_axes,
                            const int* parents, const int* active, const float* lower, const float* upper,
                            const float* defaults, int S, const int* sphere_link, const float* sphere_pos,
                            const float* sphere_radius, const unsigned char* enabled) {
    if (L <= 0 || L > MAXL) return NULL;
    orc_robot* r = (orc_robot*)calloc(1, sizeof(orc_robot));
    r->L = L;
    r->S = S;
    int dof = 0;
    double fk = 0.0;
    for (int i = 0; i < L; i++) {
        r->jtype[i] = joint_types[i];
        r->parent[i] = parents[i];
        r->active[i] = active[i] ? 1 : 0;
        r->active_index[i] = r->active[i] ? dof : -1;
        if (r->active[i]) {
            if (dof >= MAXDOF) { free(r); return NULL; }
            r->act_lo[dof] = lower[i];
            r->act_hi[dof] = upper[i];
            dof++;
        }
        memcpy(r->origin64[i], joint_origins + 16 * i, 16 * sizeof(double));
        memcpy(r->axis64[i], joint_axes + 3 * i, 3 * sizeof(double));
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) r->Of[i][3 * a + b] = (float)snap(joint_origins[16 * i + 4 * a + b]);
        for (int a = 0; a < 3; a++) r->Of[i][9 + a] = (float)snap(joint_origins[16 * i + 4 * a + 3]);
        double ax[3] = {snap(joint_axes[3 * i]), snap(joint_axes[3 * i + 1]), snap(joint_axes[3 * i + 2])};
        r->axmode[i] = AX_NONE;
        r->axsign[i] = 1;
        int nz = 0, which = -1;
        for (int a = 0; a < 3; a++)
            if (ax[a] != 0.0) { nz++; which = a; }
        if (nz == 0) {
            r->axmode[i] = AX_NONE;
        } else if (nz == 1 && fabs(ax[which]) == 1.0) {
            r->axmode[i] = which;
            r->axsign[i] = ax[which] > 0 ? 1 : -1;
        } else {
            r->axmode[i] = AX_GENERIC;
        }
        if (joint_types[i] == JT_REV && r->axmode[i] == AX_GENERIC) {
            double n = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
            for (int a = 0; a < 3; a++) r->axisf[i][a] = (float)(ax[a] / n);
        } else {
            for (int a = 0; a < 3; a++) r->axisf[i][a] = (float)ax[a];
        }
        r->lower[i] = lower[i];
        r->upper[i] = upper[i];
        r->defaults[i] = defaults[i];
        if (i > 0) fk += joint_types[i] == JT_REV ? 130.0 : joint_types[i] == JT_PRISM ? 87.0 : 63.0;
    }
    r->dof = dof;
    r->fk_flops = fk;
    r->sphere_link = (int*)malloc(sizeof(int) * (S > 0 ? S : 1));
    r->sphere_pos = (float*)malloc(sizeof(float) * 3 * (S > 0 ? S : 1));
    r->sphere_r = (float*)malloc(sizeof(float) * (S > 0 ? S : 1));
    memcpy(r->sphere_link, sphere_link, sizeof(int) * S);
    memcpy(r->sphere_pos, sphere_pos, sizeof(float) * 3 * S);
    memcpy(r->sphere_r, sphere_radius, sizeof(float) * S);
    /* pair list: SelfCollisionConstraint.cu:29-39 / RobotSolver.cu:86-96 */
    int P = 0;
    for (int i = 0; i < S; i++)
        for (int j = i + 1; j < S; j++)
            if (sphere_link[i] != sphere_link[j] && enabled[sphere_link[i] * L + sphere_link[j]]) P++;
    r->P = P;
    r->pi_ = (int*)malloc(sizeof(int) * (P > 0 ? P : 1));
    r->pj_ = (int*)malloc(sizeof(int) * (P > 0 ? P : 1));
    r->pthr = (float*)malloc(sizeof(float) * (P > 0 ? P : 1));
    r->psr = (float*)malloc(sizeof(float) * (P > 0 ? P : 1));
    /* Pairs are stored grouped by link pair: link pairs ordered by the first sphere index of
     * each link (the link of the lower sphere index first), then sphere pairs in index order.
     * This is the summation order of the self-collision cost; the set of pairs is that of the
     * reference. */
    int* first = (int*)malloc(sizeof(int) * L);
    for (int l = 0; l < L; l++) first[l] = 1 << 30;
    for (int s = S - 1; s >= 0; s--) first[sphere_link[s]] = s;
    P = 0;
    for (int a = 0; a < S; a++) {
        if (first[sphere_link[a]] != a) continue; /* a = first sphere of link la */
        int la = sphere_link[a];
        for (int b = a + 1; b < S; b++) {
            if (first[sphere_link[b]] != b) continue;
            int lb = sphere_link[b];
            if (la == lb || !enabled[la * L + lb]) continue;
            for (int i = 0; i < S; i++) {
                if (sphere_link[i] != la) continue;
                for (int j = 0; j < S; j++) {
                    if (sphere_link[j] != lb) continue;
                    float sr = sphere_radius[i] + sphere_radius[j];
                    float thr = sr * sr;
                    r->pi_[P] = i < j ? i : j;
                    r->pj_[P] = i < j ? j : i;
                    r->pthr[P] = thr;
                    r->psr[P] = sqrtf(thr);
                    P++;
                }
            }
        }
    }
    free(first);
    return r;
}

void orc_robot_destroy(orc_robot* r) {
    if (!r) return;
    free(r->sphere_link);
    free(r->sphere_pos);
    free(r->sphere_r);
    free(r->pi_);
    free(r->pj_);
    free(r->pthr);
    free(r->psr);
    free(r);
}
int orc_robot_dof(const orc_robot* r) { return r->dof; }
int orc_robot_num_pairs(const orc_robot* r) { return r->P; }
void orc_robot_pairs(const orc_robot* r, int* out_i, int* out_j, float* out_thr) {
    for (int p = 0; p < r->P; p++) {
        if (out_i) out_i[p] = r->pi_[p];
        if (out_j) out_j[p] = r->pj_[p];
        if (out_thr) out_thr[p] = r->pthr[p];
    }
}

double orc_full_flops(const orc_robot* r, const orc_scene* s) {
    double S = r->S;
    return r->fk_flops + 18.0 * S + 9.0 * r->P +
           S * (11.0 * s->ns + 34.0 * s->nb + 35.0 * s->nc + 28.0 * s->ncap);
}

/* ----------------------------------------------------------------------------------------- */
/* Scene                                                                                      */
/* ----------------------------------------------------------------------------------------- */
/* computeInversePoseMatrix (EnvConstraintCuboid.cu:13-34): [R^T | -R^T p] in float */
static void inverse_pose(const float* p, const float* R, float* m) {
    float t00 = R[0], t01 = R[3], t02 = R[6];
    float t10 = R[1], t11 = R[4], t12 = R[7];
    float t20 = R[2], t21 = R[5], t22 = R[8];
    m[0] = t00; m[1] = t01; m[2] = t02;
    m[4] = t10; m[5] = t11; m[6] = t12;
    m[8] = t20; m[9] = t21; m[10] = t22;
    m[3] = -dot3f(t00, p[0], t01, p[1], t02, p[2]);
    m[7] = -dot3f(t10, p[0], t11, p[1], t12, p[2]);
    m[11] = -dot3f(t20, p[0], t21, p[1], t22, p[2]);
}

orc_scene* orc_scene_create(const float* spheres, int ns, const float* boxes, int nb, const float* cylinders,
                            int nc, const float* capsules, int ncap) {
    orc_scene* s = (orc_scene*)calloc(1, sizeof(orc_scene));
    s->ns = ns; s->nb = nb; s->nc = nc; s->ncap = ncap;
    s->sph = (float*)malloc(sizeof(float) * 4 * (ns + 1));
    s->box = (float*)malloc(sizeof(float) * 18 * (nb + 1));
    s->cyl = (float*)malloc(sizeof(float) * 14 * (nc + 1));
    s->cap = (float*)malloc(sizeof(float) * 8 * (ncap + 1));
    if (ns) memcpy(s->sph, spheres, sizeof(float) * 4 * ns);
    for (int i = 0; i < nb; i++) {
        const float* b = boxes + 18 * i;
        inverse_pose(b, b + 3, s->box + 18 * i);
        memcpy(s->box + 18 * i + 12, b + 12, 6 * sizeof(float)); /* min3, max3 */
    }
    for (int i = 0; i < nc; i++) {
        const float* c = cylinders + 14 * i;
        inverse_pose(c, c + 3, s->cyl + 14 * i);
        s->cyl[14 * i + 12] = c[12];
        s->cyl[14 * i + 13] = c[13] * 0.5f;
    }
    for (int i = 0; i < ncap; i++) {
        const float* c = capsules + 7 * i;
        float* o = s->cap + 8 * i;
        o[0] = c[0]; o[1] = c[1]; o[2] = c[2];
        o[3] = c[3] - c[0]; o[4] = c[4] - c[1]; o[5] = c[5] - c[2];
        o[6] = 1.0f / dot3f(o[3], o[3], o[4], o[4], o[5], o[5]);
        o[7] = c[6];
    }
    return s;
}

void orc_scene_set_joint_bounds(orc_scene* s, int mode, const float* lower, const float* upper, int n) {
    s->has_bnd = mode != 0;
    for (int j = 0; j < 16; j++) {
        int given = mode == 2 && j < n;
        s->bnd[j] = given ? lower[j] : -INFINITY;
        s->bnd[16 + j] = given ? upper[j] : INFINITY;
    }
}

void orc_scene_destroy(orc_scene* s) {
    if (!s) return;
    free(s->sph); free(s->box); free(s->cyl); free(s->cap);
    free(s);
}

/* ----------------------------------------------------------------------------------------- */
/* FK                                                                                         */
/* ----------------------------------------------------------------------------------------- */
static void full_joints_d(const orc_robot* r, const double* q, double* full) {
    for (int i = 0; i < r->L; i++) full[i] = r->active[i] ? q[r->active_index[i]] : (double)r->defaults[i];
}

static void mul4(const double* A, const double* B, double* C) {
    /* multiply4x4 (KernelGenerator.h / RobotSolver.cu:850-870), last row fixed [0 0 0 1] */
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 4; j++)
            C[4 * i + j] = A[4 * i] * B[j] + A[4 * i + 1] * B[4 + j] + A[4 * i + 2] * B[8 + j] + A[4 * i + 3] * B[12 + j];
    C[12] = 0; C[13] = 0; C[14] = 0; C[15] = 1;
}

void orc_fk64(const orc_robot* r, const double* q, double* centres, double* link_poses) {
    double T[MAXL][16];
    double full[MAXL];
    full_joints_d(r, q, full);
    memset(T[0], 0, sizeof(T[0]));
    T[0][0] = T[0][5] = T[0][10] = T[0][15] = 1.0;
    for (int i = 1; i < r->L; i++) {
        const double* P = T[r->parent[i]];
        const double* O = r->origin64[i];
        if (r->jtype[i] == JT_REV) {
            /* make_rotation_axis_angle (KernelGenerator.h:289-313), double precision */
            double x = r->axis64[i][0], y = r->axis64[i][1], z = r->axis64[i][2];
            double len = sqrt(x * x + y * y + z * z);
            double valid = len >= 1e-12 ? 1.0 : 0.0;
            double inv = 1.0 / (len > 1e-12 ? len : 1e-12);
            double nx = x * inv * valid, ny = y * inv * valid, nz = z * inv * valid;
            double c = cos(full[i]), s = sin(full[i]), oc = 1.0 - c;
            double R[16] = {c + nx * nx * oc, nx * ny * oc - nz * s, nx * nz * oc + ny * s, 0,
                            ny * nx * oc + nz * s, c + ny * ny * oc, ny * nz * oc - nx * s, 0,
                            nz * nx * oc - ny * s, nz * ny * oc + nx * s, c + nz * nz * oc, 0,
                            0, 0, 0, 1};
            if (valid == 0.0) {
                memset(R, 0, sizeof(R));
                R[0] = R[5] = R[10] = R[15] = 1.0;
            }
            double tmp[16];
            mul4(P, O, tmp);
            mul4(tmp, R, T[i]);
        } else if (r->jtype[i] == JT_PRISM) {
            double Tq[16] = {1, 0, 0, r->axis64[i][0] * full[i], 0, 1, 0, r->axis64[i][1] * full[i],
                             0, 0, 1, r->axis64[i][2] * full[i], 0, 0, 0, 1};
            double OT[16];
            mul4(O, Tq, OT);
            mul4(P, OT, T[i]);
        } else if (r->jtype[i] == JT_FIXED) {
            mul4(P, O, T[i]);
        } else {
            memcpy(T[i], P, sizeof(T[i]));
        }
    }
    for (int s = 0; s < r->S; s++) {
        const double* M = T[r->sphere_link[s]];
        double px = r->sphere_pos[3 * s], py = r->sphere_pos[3 * s + 1], pz = r->sphere_pos[3 * s + 2];
        centres[3 * s + 0] = M[0] * px + M[1] * py + M[2] * pz + M[3];
        centres[3 * s + 1] = M[4] * px + M[5] * py + M[6] * pz + M[7];
        centres[3 * s + 2] = M[8] * px + M[9] * py + M[10] * pz + M[11];
    }
    if (link_poses) memcpy(link_poses, T, sizeof(double) * 16 * r->L);
}

typedef struct { float R[9]; float t[3]; } aff;

static void compose(const aff* A, const aff* B, aff* C) {
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            C->R[3 * i + j] = dot3f(A->R[3 * i], B->R[j], A->R[3 * i + 1], B->R[3 + j], A->R[3 * i + 2], B->R[6 + j]);
        C->t[i] = rowf(A->R[3 * i], B->t[0], A->R[3 * i + 1], B->t[1], A->R[3 * i + 2], B->t[2], A->t[i]);
    }
}

static void rot_of(const orc_robot* r, int i, float s, float c, aff* out) {
    memset(out, 0, sizeof(*out));
    int m = r->axmode[i];
    float ss = r->axsign[i] > 0 ? s : -s;
    if (m == AX_Z) {
        out->R[0] = c; out->R[1] = -ss; out->R[3] = ss; out->R[4] = c; out->R[8] = 1.0f;
    } else if (m == AX_Y) {
        out->R[0] = c; out->R[2] = ss; out->R[4] = 1.0f; out->R[6] = -ss; out->R[8] = c;
    } else if (m == AX_X) {
        out->R[0] = 1.0f; out->R[4] = c; out->R[5] = -ss; out->R[7] = ss; out->R[8] = c;
    } else if (m == AX_GENERIC) {
        float nx = r->axisf[i][0], ny = r->axisf[i][1], nz = r->axisf[i][2];
        float oc = 1.0f - c;
        out->R[0] = fmaf(nx * nx, oc, c);
        out->R[1] = fmaf(nx * ny, oc, -(nz * s));
        out->R[2] = fmaf(nx * nz, oc, ny * s);
        out->R[3] = fmaf(ny * nx, oc, nz * s);
        out->R[4] = fmaf(ny * ny, oc, c);
        out->R[5] = fmaf(ny * nz, oc, -(nx * s));
        out->R[6] = fmaf(nz * nx, oc, -(ny * s));
        out->R[7] = fmaf(nz * ny, oc, nx * s);
        out->R[8] = fmaf(nz * nz, oc, c);
    } else {
        out->R[0] = out->R[4] = out->R[8] = 1.0f;
    }
}

/* FP32 link frames T[L] of one configuration (the GPU's composition and operation order). */
static void fk32_frames(const orc_robot* r, const float* q, aff* T) {
    memset(&T[0], 0, sizeof(aff));
    T[0].R[0] = T[0].R[4] = T[0].R[8] = 1.0f;
    for (int i = 1; i < r->L; i++) {
        const aff* P = &T[r->parent[i]];
        aff O;
        memcpy(O.R, r->Of[i], 9 * sizeof(float));
        memcpy(O.t, r->Of[i] + 9, 3 * sizeof(float));
        float v = r->active[i] ? q[r->active_index[i]] : r->defaults[i];
        if (r->jtype[i] == JT_REV) {
            aff M, Rq;
            compose(P, &O, &M);
            float s, c;
            orc_sincosf(v, &s, &c);
            rot_of(r, i, s, c, &Rq);
            compose(&M, &Rq, &T[i]);
        } else if (r->jtype[i] == JT_PRISM) {
            aff B;
            float tq[3] = {r->axisf[i][0] * v, r->axisf[i][1] * v, r->axisf[i][2] * v};
            memcpy(B.R, O.R, sizeof(B.R));
            for (int a = 0; a < 3; a++) B.t[a] = rowf(O.R[3 * a], tq[0], O.R[3 * a + 1], tq[1], O.R[3 * a + 2], tq[2], O.t[a]);
            compose(P, &B, &T[i]);
        } else if (r->jtype[i] == JT_FIXED) {
            compose(P, &O, &T[i]);
        } else {
            T[i] = *P;
        }
    }
}

/* Link poses [L][16] row-major 4x4 (last row 0 0 0 1): the FK output of
 * kin_forward_nvrtc_kernel (KernelGenerator.h:222-406) in FP32. */
void orc_link_poses32(const orc_robot* r, const float* q, float* poses) {
    aff T[MAXL];
    fk32_frames(r, q, T);
    for (int i = 0; i < r->L; i++) {
        float* P = poses + 16 * i;
        for (int a = 0; a < 3; a++) {
            for (int b = 0; b < 3; b++) P[4 * a + b] = T[i].R[3 * a + b];
            P[4 * a + 3] = T[i].t[a];
        }
        P[12] = 0.0f; P[13] = 0.0f; P[14] = 0.0f; P[15] = 1.0f;
    }
}

void orc_fk32(const orc_robot* r, const float* q, float* centres) {
    aff T[MAXL];
    fk32_frames(r, q, T);
    for (int s = 0; s < r->S; s++) {
        const aff* M = &T[r->sphere_link[s]];
        float px = r->sphere_pos[3 * s], py = r->sphere_pos[3 * s + 1], pz = r->sphere_pos[3 * s + 2];
        for (int a = 0; a < 3; a++)
            centres[3 * s + a] = rowf(M->R[3 * a], px, M->R[3 * a + 1], py, M->R[3 * a + 2], pz, M->t[a]);
    }
}

/* ----------------------------------------------------------------------------------------- */
/* Collision                                                                                  */
/* ----------------------------------------------------------------------------------------- */
#define F_SELF 9.0
#define F_SPH 11.0
#define F_BOX 34.0
#define F_CYL 35.0
#define F_CAP 28.0

/* Per-pair contributions; each returns the reference cost term (>= 0), > 0 <=> contact. */
static inline float self_term(float d2, float thr, float sr) {
    /* SelfCollisionConstraint.cu:97-115 */
    return d2 < thr ? sr - sqrtf(d2) : 0.0f;
}
static inline float sphere_term(const float* c, float r, const float* o) {
    /* EnvConstraintSphere.cu:80-95 */
    float dx = c[0] - o[0], dy = c[1] - o[1], dz = c[2] - o[2];
    float d2 = dot3f(dx, dx, dy, dy, dz, dz);
    float sr = r + o[3];
    float thr = sr * sr;
    return d2 < thr ? fmaxf(0.0f, sr - sqrtf(d2)) : 0.0f;
}
static inline float box_term(const float* c, float r, const float* m) {
    /* EnvConstraintCuboid.cu:133-160 */
    float lx = rowf(m[0], c[0], m[1], c[1], m[2], c[2], m[3]);
    float ly = rowf(m[4], c[0], m[5], c[1], m[6], c[2], m[7]);
    float lz = rowf(m[8], c[0], m[9], c[1], m[10], c[2], m[11]);
    float cx = fminf(fmaxf(lx, m[12]), m[15]);
    float cy = fminf(fmaxf(ly, m[13]), m[16]);
    float cz = fminf(fmaxf(lz, m[14]), m[17]);
    float dx = lx - cx, dy = ly - cy, dz = lz - cz;
    float d2 = dot3f(dx, dx, dy, dy, dz, dz);
    return d2 < r * r ? r - sqrtf(d2) : 0.0f;
}
static inline float cyl_term(const float* c, float r, const float* m) {
    /* EnvConstraintCylinder.cu:150-180 */
    float lx = rowf(m[0], c[0], m[1], c[1], m[2], c[2], m[3]);
    float ly = rowf(m[4], c[0], m[5], c[1], m[6], c[2], m[7]);
    float lz = rowf(m[8], c[0], m[9], c[1], m[10], c[2], m[11]);
    float dxy = sqrtf(fmaf(lx, lx, ly * ly));
    float dx = dxy - m[12];
    float dz = fabsf(lz) - m[13];
    float ox = dx > 0.0f ? dx : 0.0f;
    float oz = dz > 0.0f ? dz : 0.0f;
    float outside = sqrtf(fmaf(ox, ox, oz * oz));
    float inside = fminf(fmaxf(dx, dz), 0.0f);
    float sdf = outside + inside;
    float pen = r - sdf;
    return pen > 0.0f ? pen : 0.0f;
}
static inline float cap_term(const float* c, float r, const float* k) {
    /* sphere_capsule (pRRTC/src/planning/utils.cuh:113-127): the closest point of the segment
     * p1 + t v, t = clamp((c - p1).v / |v|^2, 0, 1); contact <=> |c - closest|^2 < (r+R)^2.  The
     * cost is a penetration depth, as for env spheres (EnvConstraintSphere.cu:89-93):
     * max(0, (r+R) - |c - closest|) */
    float dot = dot3f(c[0] - k[0], k[3], c[1] - k[1], k[4], c[2] - k[2], k[5]);
    float cdf = fminf(fmaxf(dot * k[6], 0.0f), 1.0f);
    float px = fmaf(k[3], cdf, k[0]), py = fmaf(k[4], cdf, k[1]), pz = fmaf(k[5], cdf, k[2]);
    float dx = c[0] - px, dy = c[1] - py, dz = c[2] - pz;
    float d2 = dot3f(dx, dx, dy, dy, dz, dz);
    float rs = r + k[7];
    float rr = rs * rs;
    return d2 < rr ? fmaxf(0.0f, rs - sqrtf(d2)) : 0.0f;
}

static int bounds_ok(const orc_robot* r, const orc_scene* sc, const float* q, float* cost) {
    /* BoundaryConstraint.cu:56-92 with tolerance 1e-4 (BoundaryConstraint.h:13); the bounds are
     * the scene's when it carries its own (a space's BoundaryConstraint, or none) */
    float acc = 0.0f;
    for (int j = 0; j < r->dof; j++) {
        float v = q[j];
        float lo = sc && sc->has_bnd ? sc->bnd[j] : r->act_lo[j];
        float hi = sc && sc->has_bnd ? sc->bnd[16 + j] : r->act_hi[j];
        if (v < lo) { float d = lo - v; acc = fmaf(d, d, acc); }
        else if (v > hi) { float d = v - hi; acc = fmaf(d, d, acc); }
    }
    if (acc < 1e-4f) { if (cost) *cost = 0.0f; return 1; }
    if (cost) *cost = sqrtf(acc);
    return 0;
}

/* Per-constraint costs in the fixed order env spheres, boxes, cylinders, capsules (each
 * obstacle-major, then robot sphere), self pairs, bounds (ORC_COST_KINDS values), and their total
 * summed in that order.  Returns 1 if valid (total == 0). */
#define ORC_COST_KINDS 6
static int check_one_cost(const orc_robot* r, const orc_scene* sc, const float* q, const float* cen, float* total,
                          float* parts) {
    float pk[ORC_COST_KINDS];
    int S = r->S;
    float part = 0.0f;
    for (int o = 0; o < sc->ns; o++)
        for (int s = 0; s < S; s++) part += sphere_term(cen + 3 * s, r->sphere_r[s], sc->sph + 4 * o);
    pk[0] = part;
    part = 0.0f;
    for (int o = 0; o < sc->nb; o++)
        for (int s = 0; s < S; s++) part += box_term(cen + 3 * s, r->sphere_r[s], sc->box + 18 * o);
    pk[1] = part;
    part = 0.0f;
    for (int o = 0; o < sc->nc; o++)
        for (int s = 0; s < S; s++) part += cyl_term(cen + 3 * s, r->sphere_r[s], sc->cyl + 14 * o);
    pk[2] = part;
    part = 0.0f;
    for (int o = 0; o < sc->ncap; o++)
        for (int s = 0; s < S; s++) part += cap_term(cen + 3 * s, r->sphere_r[s], sc->cap + 8 * o);
    pk[3] = part;
    part = 0.0f;
    for (int p = 0; p < r->P; p++) {
        const float* a = cen + 3 * r->pi_[p];
        const float* b = cen + 3 * r->pj_[p];
        float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
        part += self_term(dot3f(dx, dx, dy, dy, dz, dz), r->pthr[p], r->psr[p]);
    }
    pk[4] = part;
    bounds_ok(r, sc, q, &pk[5]);
    float tot = 0.0f;
    for (int k = 0; k < ORC_COST_KINDS; k++) tot += pk[k];
    if (parts) memcpy(parts, pk, sizeof(pk));
    *total = tot;
    return tot == 0.0f;
}

void orc_constraint_costs(const orc_robot* r, const orc_scene* s, const float* q, int64_t n, float* parts) {
    float* cen = (float*)malloc(sizeof(float) * 3 * (r->S + 1));
    float pk[ORC_COST_KINDS], tot;
    for (int64_t i = 0; i < n; i++) {
        orc_fk32(r, q + i * r->dof, cen);
        check_one_cost(r, s, q + i * r->dof, cen, &tot, pk);
        for (int k = 0; k < ORC_COST_KINDS; k++) parts[k * n + i] = pk[k];
    }
    free(cen);
}

/* Early-out verdict in canonical order; returns 1 if valid; *flops = canonical work. */
static int check_one_canon(const orc_robot* r, const orc_scene* sc, const float* q, const float* cen, double* flops) {
    double f = r->fk_flops + 18.0 * r->S;
    int valid = bounds_ok(r, sc, q, NULL);
    if (valid) {
        for (int p = 0; p < r->P && valid; p++) {
            const float* a = cen + 3 * r->pi_[p];
            const float* b = cen + 3 * r->pj_[p];
            float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
            f += F_SELF;
            if (self_term(dot3f(dx, dx, dy, dy, dz, dz), r->pthr[p], r->psr[p]) > 0.0f) valid = 0;
        }
        for (int s = 0; s < r->S && valid; s++) {
            const float* c = cen + 3 * s;
            float rr = r->sphere_r[s];
            for (int o = 0; o < sc->ns && valid; o++) { f += F_SPH; if (sphere_term(c, rr, sc->sph + 4 * o) > 0.0f) valid = 0; }
            for (int o = 0; o < sc->nb && valid; o++) { f += F_BOX; if (box_term(c, rr, sc->box + 18 * o) > 0.0f) valid = 0; }
            for (int o = 0; o < sc->nc && valid; o++) { f += F_CYL; if (cyl_term(c, rr, sc->cyl + 14 * o) > 0.0f) valid = 0; }
            for (int o = 0; o < sc->ncap && valid; o++) { f += F_CAP; if (cap_term(c, rr, sc->cap + 8 * o) > 0.0f) valid = 0; }
        }
    }
    if (flops) *flops = f;
    return valid;
}

static int check_q(const orc_robot* r, const orc_scene* sc, const float* q) {
    float cen[3 * 512];
    orc_fk32(r, q, cen);
    return check_one_canon(r, sc, q, cen, NULL);
}

typedef struct {
    const orc_robot* r;
    const orc_scene* s;
    const float* q;
    int64_t lo, hi;
    uint8_t* verdict;
    float* cost;
    float* centres;
    double* canon;
} check_job;

static void* check_worker(void* arg) {
    check_job* j = (check_job*)arg;
    const orc_robot* r = j->r;
    int S = r->S, dof = r->dof;
    float* tmp = (float*)malloc(sizeof(float) * 3 * (S + 1));
    for (int64_t i = j->lo; i < j->hi; i++) {
        const float* q = j->q + i * dof;
        float* cen = j->centres ? j->centres + i * 3 * S : tmp;
        orc_fk32(r, q, cen);
        if (j->cost) {
            float tot;
            int v = check_one_cost(r, j->s, q, cen, &tot, NULL);
            j->cost[i] = tot;
            if (j->verdict) j->verdict[i] = (uint8_t)v;
            if (j->canon) check_one_canon(r, j->s, q, cen, &j->canon[i]);
        } else {
            double f;
            int v = check_one_canon(r, j->s, q, cen, &f);
            if (j->verdict) j->verdict[i] = (uint8_t)v;
            if (j->canon) j->canon[i] = f;
        }
    }
    free(tmp);
    return NULL;
}

void orc_check32(const orc_robot* r, const orc_scene* s, const float* q, int64_t n, uint8_t* verdict, float* cost,
                 float* centres, double* canon_flops, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if (n < nthreads) nthreads = n > 0 ? (int)n : 1;
    pthread_t th[256];
    check_job jobs[256];
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = (check_job){r, s, q, n * t / nthreads, n * (t + 1) / nthreads, verdict, cost, centres, canon_flops};
        if (nthreads > 1) pthread_create(&th[t], NULL, check_worker, &jobs[t]);
    }
    if (nthreads == 1) check_worker(&jobs[0]);
    else
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

void orc_check_centres32(const orc_robot* r, const orc_scene* s, const float* q, const float* centres, int64_t n,
                         uint8_t* verdict) {
    for (int64_t i = 0; i < n; i++)
        verdict[i] = (uint8_t)check_one_canon(r, s, q + i * r->dof, centres + i * 3 * r->S, NULL);
}

/* ----------------------------------------------------------------------------------------- */
/* Edges, NN, validation                                                                      */
/* ----------------------------------------------------------------------------------------- */
void orc_edges32(const orc_robot* r, const orc_scene* s, const float* q0, const float* q1, int m, float resolution,
                 uint8_t* ok, float* length) {
    /* checkUnconstrainedMotions (SingleArmSpace.cu:646-848): steps = floor(d/res)+1; states
     * q0 + dir*i (i < steps-1) and q1; feasible iff every state is valid; cost = d */
    int dof = r->dof;
    float st[MAXDOF], dir[MAXDOF];
    for (int e = 0; e < m; e++) {
        const float* a = q0 + e * dof;
        const float* b = q1 + e * dof;
        float d2 = 0.0f;
        for (int j = 0; j < dof; j++) { float df = a[j] - b[j]; d2 = fmaf(df, df, d2); }
        float d = sqrtf(d2);
        int steps = (int)(d / resolution) + 1;
        for (int j = 0; j < dof; j++) dir[j] = (b[j] - a[j]) / (float)steps;
        int good = 1;
        for (int i = 0; i < steps && good; i++) {
            if (i == steps - 1) memcpy(st, b, sizeof(float) * dof);
            else for (int j = 0; j < dof; j++) st[j] = fmaf(dir[j], (float)i, a[j]);
            good = check_q(r, s, st);
        }
        ok[e] = (uint8_t)good;
        length[e] = d;
    }
}

void orc_steer(const float* from, float* to, int n, int dof, float max_distance) {
    /* interpolate_kernel (SingleArmSpace.cu:219-249): d = |to - from|; when d > max, to <- from +
     * (max / d)(to - from).  FP32 order: d2 accumulated by fmaf over the joints in order,
     * correctly rounded sqrtf and division, each coordinate fmaf(f, to - from, from). */
    for (int i = 0; i < n; i++) {
        const float* a = from + (int64_t)i * dof;
        float* b = to + (int64_t)i * dof;
        float d2 = 0.0f;
        for (int j = 0; j < dof; j++) { float df = a[j] - b[j]; d2 = fmaf(df, df, d2); }
        float d = sqrtf(d2);
        if (d > max_distance) {
            float f = max_distance / d;
            for (int j = 0; j < dof; j++) b[j] = fmaf(f, b[j] - a[j], a[j]);
        }
    }
}

int orc_nn(const float* tree, int n, int dof, const float* q) {
    int best = -1;
    float bd = INFINITY;
    for (int i = 0; i < n; i++) {
        float acc = 0.0f;
        for (int j = 0; j < dof; j++) { float df = q[j] - tree[i * dof + j]; acc = fmaf(df, df, acc); }
        if (acc < bd) { bd = acc; best = i; }
    }
    return best;
}

/* Nearest member of q among members [cnt][dof] (a group's states in order of addition): squared
 * L2 accumulated by fmaf over the joints, ties to the earlier member. */
static int group_nearest(const float* members, int cnt, int dof, const float* q) {
    int best = 0;
    float bd = INFINITY;
    for (int m = 0; m < cnt; m++) {
        float acc = 0.0f;
        for (int j = 0; j < dof; j++) { float d0 = q[j] - members[m * dof + j]; acc = fmaf(d0, d0, acc); }
        if (acc < bd) { bd = acc; best = m; }
    }
    return best;
}

int orc_rrg(const orc_robot* r, const orc_scene* s, const float* starts, int ns, const float* goals, int ng, int k,
            float max_travel, float resolution, uint64_t seed, uint64_t offset, int max_passes, float* vertices,
            int* groups, int vcap, int* edges, float* weights, int ecap, int* start_members, int* goal_members,
            orc_rrg_info* info) {
    /* CUDAMPLib::RRG::solve's extend-connect loop (RRG.cpp:241-424), pass by pass. */
    const int dof = r->dof;
    if (ns < 1 || ng < 1 || k < 1 || ns + ng > vcap) return -1;
    float* gq[2] = {(float*)malloc(sizeof(float) * dof * vcap), (float*)malloc(sizeof(float) * dof * vcap)};
    int* gv[2] = {start_members, goal_members};
    float* smp = (float*)malloc(sizeof(float) * dof * k);
    float* from = (float*)malloc(sizeof(float) * dof * k);
    float* q0 = (float*)malloc(sizeof(float) * dof * 2 * k);
    float* q1 = (float*)malloc(sizeof(float) * dof * 2 * k);
    uint8_t* freed = (uint8_t*)malloc(k);
    uint8_t* ok = (uint8_t*)malloc(2 * k);
    float* len = (float*)malloc(sizeof(float) * 2 * k);
    int* nn1 = (int*)malloc(sizeof(int) * k);
    int* nn2 = (int*)malloc(sizeof(int) * k);
    int nv = ns + ng, ne = 0, cnt[2] = {ns, ng}, t = 0, pass = 0, solved = 0, rc = 0;
    float travel[2] = {max_travel, max_travel};
    /* vertices 0..ns-1: the start states (group 1), then the goal states (group 2) */
    for (int v = 0; v < nv; v++) {
        const float* q = v < ns ? starts + v * dof : goals + (v - ns) * dof;
        memcpy(vertices + v * dof, q, sizeof(float) * dof);
        groups[v] = v < ns ? 1 : 2;
        int g = v < ns ? 0 : 1, m = v < ns ? v : v - ns;
        memcpy(gq[g] + m * dof, q, sizeof(float) * dof);
        gv[g][m] = v;
    }
    while (pass < max_passes && !solved) {
        const int side = t & 1;
        /* sample, pull to within the side's travel distance of the nearest side member, check */
        int nfree = 0;
        for (int i = 0; i < k; i++) {
            orc_sample(r, seed, offset + (uint64_t)pass * (uint64_t)k + (uint64_t)i, smp + i * dof);
            int m = group_nearest(gq[side], cnt[side], dof, smp + i * dof);
            memcpy(from + i * dof, gq[side] + m * dof, sizeof(float) * dof);
        }
        orc_steer(from, smp, k, dof, travel[side]);
        for (int i = 0; i < k; i++) {
            freed[i] = (uint8_t)check_q(r, s, smp + i * dof);
            nfree += freed[i];
        }
        if (nfree == 0) { /* RRG.cpp:300-306 */
            t += 1;
            pass++;
            continue;
        }
        /* motions from each free sample to its nearest member of each group */
        int m2 = 0;
        for (int i = 0; i < k; i++) {
            if (!freed[i]) continue;
            nn1[i] = group_nearest(gq[0], cnt[0], dof, smp + i * dof);
            nn2[i] = group_nearest(gq[1], cnt[1], dof, smp + i * dof);
        }
        for (int i = 0; i < k; i++) {
            memcpy(q0 + i * dof, smp + i * dof, sizeof(float) * dof);
            memcpy(q0 + (k + i) * dof, smp + i * dof, sizeof(float) * dof);
            memcpy(q1 + i * dof, gq[0] + (freed[i] ? nn1[i] : 0) * dof, sizeof(float) * dof);
            memcpy(q1 + (k + i) * dof, gq[1] + (freed[i] ? nn2[i] : 0) * dof, sizeof(float) * dof);
        }
        for (int e = 0; e < 2 * k; e++) {
            int i = e < k ? e : e - k;
            if (freed[i]) orc_edges32(r, s, q0 + e * dof, q1 + e * dof, 1, resolution, ok + e, len + e);
            else { ok[e] = 0; len[e] = 0.0f; }
        }
        int reach[2] = {0, 0};
        for (int i = 0; i < k; i++) {
            reach[0] += freed[i] && ok[i];
            reach[1] += freed[i] && ok[k + i];
        }
        (void)m2;
        if (reach[side] == 0) { /* halve the side's travel distance (floor 0.1), same side */
            double half = (double)travel[side] * 0.5;
            travel[side] = half > 0.1 ? (float)half : 0.1f;
            t += 2;
            pass++;
            continue;
        }
        if (nv + nfree > vcap || ne + 2 * nfree > ecap) { rc = -2; break; }
        travel[side] = max_travel;
        for (int i = 0; i < k; i++) {
            int to1 = freed[i] && ok[i], to2 = freed[i] && ok[k + i];
            if (!to1 && !to2) continue;
            int v = nv++;
            memcpy(vertices + v * dof, smp + i * dof, sizeof(float) * dof);
            groups[v] = to1 && to2 ? 0 : (to1 ? 1 : 2);
            if (to1) { edges[2 * ne] = v; edges[2 * ne + 1] = gv[0][nn1[i]]; weights[ne] = len[i]; ne++; }
            if (to2) { edges[2 * ne] = v; edges[2 * ne + 1] = gv[1][nn2[i]]; weights[ne] = len[k + i]; ne++; }
            if (to1 && to2) solved = 1;
            else {
                int g = to1 ? 0 : 1;
                memcpy(gq[g] + cnt[g] * dof, smp + i * dof, sizeof(float) * dof);
                gv[g][cnt[g]++] = v;
            }
        }
        t += 1;
        pass++;
    }
    if (info) {
        info->solved = solved;
        info->passes = pass;
        info->loop_index = t;
        info->vertices = nv;
        info->edges = ne;
        info->start_group = cnt[0];
        info->goal_group = cnt[1];
        info->travel_start = travel[0];
        info->travel_goal = travel[1];
    }
    free(gq[0]); free(gq[1]); free(smp); free(from); free(q0); free(q1); free(freed); free(ok); free(len);
    free(nn1); free(nn2);
    return rc;
}

int orc_validate_path(const orc_robot* r, const orc_scene* s, const float* path, int k, float resolution,
                      int* bad_segment) {
    int dof = r->dof;
    float st[MAXDOF];
    if (bad_segment) *bad_segment = -1;
    if (k <= 0) return 0;
    if (!check_q(r, s, path)) { if (bad_segment) *bad_segment = -2; return 0; }
    for (int seg = 0; seg + 1 < k; seg++) {
        const float* a = path + seg * dof;
        const float* b = path + (seg + 1) * dof;
        double d2 = 0.0;
        for (int j = 0; j < dof; j++) { double df = (double)b[j] - a[j]; d2 += df * df; }
        int n = (int)ceil(sqrt(d2) / resolution);
        if (n < 1) n = 1;
        for (int i = 1; i <= n; i++) {
            double t = (double)i / n;
            for (int j = 0; j < dof; j++) st[j] = (float)(a[j] + ((double)b[j] - a[j]) * t);
            if (!check_q(r, s, st)) { if (bad_segment) *bad_segment = seg; return 0; }
        }
    }
    return 1;
}

/* ----------------------------------------------------------------------------------------- */
/* RNG                                                                                        */
/* ----------------------------------------------------------------------------------------- */
void orc_philox4x32(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t out[4]) {
    for (int round = 0; round < 10; round++) {
        if (round > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_sample(const orc_robot* r, uint64_t seed, uint64_t idx, float* q) {
    uint32_t w[4];
    for (int j = 0; j < r->dof; j++) {
        if ((j & 3) == 0)
            orc_philox4x32((uint32_t)idx, (uint32_t)(idx >> 32), (uint32_t)(j >> 2), 0x6d70u, (uint32_t)seed,
                           (uint32_t)(seed >> 32), w);
        float u = (float)(w[j & 3] >> 8) * (1.0f / 16777216.0f);
        q[j] = fmaf(u, r->act_hi[j] - r->act_lo[j], r->act_lo[j]);
    }
}

/* Halton sample (pRRTC.cu:43-105 draws a stateful Halton sequence per worker over the prime
 * bases 3, 5, 7, ... of its table at :60-63).  Stateless form: sample idx is the Halton point
 * i = idx + 1, dimension j the radical inverse of i in base ORC_HALTON_BASES[j], exact in
 * integers (digits reversed into num over den = b^digits) and rounded once by a float division.
 * The point set is randomised per seed by a Cranley-Patterson rotation (shift every coordinate
 * by r_j in [0, 1), modulo 1), r_j from the Philox stream of the seed at counter (0, 0, j/4,
 * 0x6861); coordinates are scaled into the joint bounds as in orc_sample. */
static const uint32_t ORC_HALTON_BASES[16] = {3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59};

float orc_radical_inverse(uint32_t i, uint32_t b) {
    uint64_t num = 0, den = 1;
    while (i) {
        num = num * b + i % b;
        den *= b;
        i /= b;
    }
    return (float)num / (float)den;
}

void orc_sample_halton(const orc_robot* r, uint64_t seed, uint64_t idx, float* q) {
    uint32_t w[4];
    const uint32_t i = (uint32_t)(idx + 1);
    for (int j = 0; j < r->dof; j++) {
        if ((j & 3) == 0)
            orc_philox4x32(0u, 0u, (uint32_t)(j >> 2), 0x6861u, (uint32_t)seed, (uint32_t)(seed >> 32), w);
        const float rot = (float)(w[j & 3] >> 8) * (1.0f / 16777216.0f);
        float v = orc_radical_inverse(i, ORC_HALTON_BASES[j]) + rot;
        if (v >= 1.0f) v -= 1.0f;
        q[j] = fmaf(v, r->act_hi[j] - r->act_lo[j], r->act_lo[j]);
    }
}

/* ----------------------------------------------------------------------------------------- */
/* RRT-Connect (OMPL semantics)                                                               */
/* ----------------------------------------------------------------------------------------- */
typedef struct {
    float* q;
    int* parent;
    int n, cap, dof;
} tree_t;

static void tree_init(tree_t* t, int dof, const float* root) {
    t->dof = dof;
    t->cap = 1024;
    t->n = 1;
    t->q = (float*)malloc(sizeof(float) * dof * t->cap);
    t->parent = (int*)malloc(sizeof(int) * t->cap);
    memcpy(t->q, root, sizeof(float) * dof);
    t->parent[0] = -1;
}
static int tree_add(tree_t* t, const float* q, int parent) {
    if (t->n == t->cap) {
        t->cap *= 2;
        t->q = (float*)realloc(t->q, sizeof(float) * t->dof * t->cap);
        t->parent = (int*)realloc(t->parent, sizeof(int) * t->cap);
    }
    memcpy(t->q + t->n * t->dof, q, sizeof(float) * t->dof);
    t->parent[t->n] = parent;
    return t->n++;
}
static void tree_free(tree_t* t) { free(t->q); free(t->parent); }

static float qdist(const float* a, const float* b, int dof) {
    double s = 0.0;
    for (int j = 0; j < dof; j++) { double d = (double)a[j] - b[j]; s += d * d; }
    return (float)sqrt(s);
}

typedef struct {
    const orc_robot* r;
    const orc_scene* s;
    float range, res;
    int64_t checked;
} rrtc_ctx;

/* DiscreteMotionValidator::checkMotion: s2 first, then interior states j/nd, nd = ceil(d/res) */
static int check_motion(rrtc_ctx* c, const float* s1, const float* s2) {
    int dof = c->r->dof;
    c->checked++;
    if (!check_q(c->r, c->s, s2)) return 0;
    float d = qdist(s1, s2, dof);
    int nd = (int)ceil(d / c->res);
    float st[MAXDOF];
    for (int j = 1; j < nd; j++) {
        float t = (float)j / (float)nd;
        for (int k = 0; k < dof; k++) st[k] = s1[k] + (s2[k] - s1[k]) * t;
        c->checked++;
        if (!check_q(c->r, c->s, st)) return 0;
    }
    return 1;
}

enum { TRAPPED = 0, ADVANCED = 1, REACHED = 2 };

/* RRTConnect::growTree */
static int grow(rrtc_ctx* c, tree_t* t, const float* target, int is_start, int* added) {
    int dof = c->r->dof;
    int nn = orc_nn(t->q, t->n, dof, target);
    const float* nq = t->q + nn * dof;
    float d = qdist(nq, target, dof);
    float x[MAXDOF];
    const float* dst = target;
    int reach = 1;
    if (d > c->range) {
        float f = c->range / d;
        for (int k = 0; k < dof; k++) x[k] = nq[k] + (target[k] - nq[k]) * f;
        if (memcmp(x, nq, sizeof(float) * dof) == 0) return TRAPPED;
        dst = x;
        reach = 0;
    }
    int ok = is_start ? check_motion(c, nq, dst) : (check_q(c->r, c->s, dst) && check_motion(c, dst, nq));
    if (!ok) return TRAPPED;
    *added = tree_add(t, dst, nn);
    return reach ? REACHED : ADVANCED;
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int orc_rrtconnect(const orc_robot* r, const orc_scene* s, const float* start, const float* goal, float range,
                   float resolution, uint64_t seed, int max_iterations, double time_limit, float* out_path, int out_cap,
                   orc_plan_stats* stats) {
    double t0 = now_s();
    int dof = r->dof;
    rrtc_ctx c = {r, s, range, resolution, 0};
    orc_plan_stats st;
    memset(&st, 0, sizeof(st));
    int written = 0;
    if (!check_q(r, s, start) || !check_q(r, s, goal)) {
        st.solved = -1;
        st.seconds = now_s() - t0;
        if (stats) *stats = st;
        return 0;
    }
    tree_t ts, tg;
    tree_init(&ts, dof, start);
    tree_init(&tg, dof, goal);
    int start_tree = 1;
    float rq[MAXDOF];
    uint64_t sample_idx = 0;
    int it;
    for (it = 0; it < max_iterations; it++) {
        if (time_limit > 0 && (it & 15) == 0 && now_s() - t0 > time_limit) break;
        tree_t* tree = start_tree ? &ts : &tg;
        int tree_is_start = start_tree;
        start_tree = !start_tree;
        tree_t* other = start_tree ? &ts : &tg;
        orc_sample(r, seed, sample_idx++, rq);
        int added = -1;
        int gs = grow(&c, tree, rq, tree_is_start, &added);
        if (gs == TRAPPED) continue;
        float target[MAXDOF];
        memcpy(target, tree->q + added * dof, sizeof(float) * dof);
        int oadded = -1, gsc;
        do {
            gsc = grow(&c, other, target, !tree_is_start, &oadded);
        } while (gsc == ADVANCED);
        if (gsc == REACHED) {
            int smot = tree_is_start ? added : oadded;
            int gmot = tree_is_start ? oadded : added;
            /* start half root->smot, goal half gmot->root */
            int* chain = (int*)malloc(sizeof(int) * (size_t)ts.n);
            int n1 = 0;
            for (int v = smot; v >= 0 && n1 < ts.n; v = ts.parent[v]) chain[n1++] = v;
            for (int i = n1 - 1; i >= 0 && written < out_cap; i--)
                memcpy(out_path + (written++) * dof, ts.q + chain[i] * dof, sizeof(float) * dof);
            free(chain);
            int first = 1;
            for (int v = gmot; v >= 0 && written < out_cap; v = tg.parent[v]) {
                if (first && memcmp(tg.q + v * dof, out_path + (written - 1) * dof, sizeof(float) * dof) == 0) {
                    first = 0;
                    continue;
                }
                first = 0;
                memcpy(out_path + (written++) * dof, tg.q + v * dof, sizeof(float) * dof);
            }
            st.solved = 1;
            it++;
            break;
        }
    }
    st.iterations = it;
    st.start_tree = ts.n;
    st.goal_tree = tg.n;
    st.states_checked = c.checked;
    st.seconds = now_s() - t0;
    tree_free(&ts);
    tree_free(&tg);
    if (stats) *stats = st;
    return written;
}


/* ----------------------------------------------------------------------------------------- */
/* CPRRTC planner, one team per query                                                        */
/* ----------------------------------------------------------------------------------------- */
/* The parallel RRT-Connect of CPRRTC (RobotSolver.cu:507-799) run as the GPU planner runs it
 * with one team per query (exclusive mode), which makes it sequential and deterministic per seed.
 *
 * One iteration (64 lanes): draw sample i of the query's Philox stream (orc_sample); grow the
 * smaller tree, the start tree on a tie (RobotSolver.cu:570-588), from its node nearest to the
 * sample (squared distance accumulated with fmaf over the joints, ties to the earlier node).  The
 * extension covers ms = min(floor(d / res), max_step) states from that node toward the sample,
 * spaced res apart (state k = fmaf(delta, k, from), delta = ((sample - from) / d) * res); with
 * ms = 0 nothing happens.  Every state is checked and the last one becomes the new node.  Then
 * the same tree connects greedily toward the other tree's node nearest to the new node: while
 * the gap is at least 2 res, a chunk of min(floor(gap / res), max_step) states from the current
 * node, inserted at its last state when all are free, else the connection stops; a gap below
 * 2 res joins the trees unchecked (RobotSolver.cu:686-794).
 *
 * Paired iterations (32 lanes, the reference's 32-thread blocks two to a wave): samples 2p and
 * 2p + 1 (the second only within the budget) run one iteration each against the tree sizes read
 * before the pair, so both grow the same tree and search the same nodes; their extensions are
 * inserted first's first, and their connections advance in lockstep, one chunk each per step,
 * inserts in the same order; the first connection to close (the first sample's when both close
 * in one step) solves the query.
 *
 * The path is the start tree's branch from its root to the joining node, then the goal tree's
 * from its joining node to its root (RobotSolver.cu:297-346).  Statuses as MP_PLAN_*: 1 solved,
 * 2 exhausted, 3 tree full (an insert beyond cap nodes), 4 invalid start or goal.  iterations =
 * min(samples drawn, max_iters). */
static int tree_nearest(const tree_t* t, int cnt, const float* q, float* best_out) {
    float best = INFINITY;
    int bi = 0;
    for (int i = 0; i < cnt; i++) {
        float acc = 0.0f;
        for (int j = 0; j < t->dof; j++) {
            float d0 = q[j] - t->q[i * t->dof + j];
            acc = fmaf(d0, d0, acc);
        }
        if (acc < best) { best = acc; bi = i; }
    }
    *best_out = best;
    return bi;
}

static int chunk_free(const orc_robot* r, const orc_scene* s, const float* from, const float* delta, int n) {
    float q[MAXDOF];
    for (int k = 0; k < n; k++) {
        for (int j = 0; j < r->dof; j++) q[j] = fmaf(delta[j], (float)k, from[j]);
        if (!check_q(r, s, q)) return 0;
    }
    return 1;
}

typedef struct {
    int live, parent, onn;
    float cur[MAXDOF], dcur;
} cp_branch;

/* The greedy connections of the live branches, in lockstep; returns the status (0 running). */
static int cp_connect(const orc_robot* r, const orc_scene* s, tree_t* T, int tgt, cp_branch* b, int nb, float res,
                      int max_step, int cap, int* join) {
    const int dof = r->dof, oth = 1 - tgt;
    for (;;) {
        int solved = 0, live = 0;
        for (int h = 0; h < nb; h++) {
            if (!b[h].live || !(b[h].dcur < 2.0f * res)) continue;
            if (!solved) {
                join[tgt] = b[h].parent;
                join[oth] = b[h].onn;
                solved = 1;
            }
            b[h].live = 0;
        }
        if (solved) return 1;
        float n1[2][MAXDOF];
        for (int h = 0; h < nb; h++) {
            if (!b[h].live) continue;
            const float* to = T[oth].q + b[h].onn * dof;
            int ms1 = (int)(b[h].dcur / res);
            ms1 = ms1 < max_step ? ms1 : max_step;
            float d1v[MAXDOF];
            for (int j = 0; j < dof; j++) {
                d1v[j] = ((to[j] - b[h].cur[j]) / b[h].dcur) * res;
                n1[h][j] = fmaf(d1v[j], (float)(ms1 - 1), b[h].cur[j]);
            }
            if (!chunk_free(r, s, b[h].cur, d1v, ms1)) b[h].live = 0;  /* trapped */
            live += b[h].live;
        }
        if (!live) return 0;
        for (int h = 0; h < nb; h++) {
            if (!b[h].live) continue;
            if (T[tgt].n >= cap) return 3;
            b[h].parent = tree_add(&T[tgt], n1[h], b[h].parent);
        }
        for (int h = 0; h < nb; h++) {
            if (!b[h].live) continue;
            const float* to = T[oth].q + b[h].onn * dof;
            memcpy(b[h].cur, n1[h], sizeof(float) * dof);
            float acc = 0.0f;
            for (int j = 0; j < dof; j++) {
                float df = to[j] - n1[h][j];
                acc = fmaf(df, df, acc);
            }
            b[h].dcur = sqrtf(acc);
        }
    }
}

int orc_cprrtc(const orc_robot* r, const orc_scene* s, const float* start, const float* goal, float res,
               int max_step, uint64_t seed, int max_iters, int cap, int lanes, float* out_path, int out_cap,
               orc_cprrtc_info* info) {
    const int dof = r->dof;
    const int step = lanes == 32 ? 2 : 1;
    orc_cprrtc_info inf;
    memset(&inf, 0, sizeof(inf));
    if (!check_q(r, s, start) || !check_q(r, s, goal)) {
        inf.status = 4;
        inf.start_tree = inf.goal_tree = 1;
        if (info) *info = inf;
        return 0;
    }
    tree_t T[2];
    tree_init(&T[0], dof, start);
    tree_init(&T[1], dof, goal);
    int status = 0, join[2] = {-1, -1};
    long drawn = 0;
    for (int sidx = 0; status == 0; sidx += step) {
        drawn = (long)sidx + step;
        if (sidx >= max_iters) { status = 2; break; }
        const int nb = step == 2 && sidx + 1 < max_iters ? 2 : 1;
        const int tgt = T[0].n > T[1].n ? 1 : 0, oth = 1 - tgt;
        const int cnt = T[tgt].n;
        cp_branch b[2];
        float node[2][MAXDOF];
        int nn[2], live = 0;
        for (int h = 0; h < nb; h++) {
            float smp[MAXDOF], delta[MAXDOF], bd;
            orc_sample(r, seed, (uint64_t)(sidx + h), smp);
            nn[h] = tree_nearest(&T[tgt], cnt, smp, &bd);
            const float d = sqrtf(bd);
            int ms = (int)(d / res);
            ms = ms < max_step ? ms : max_step;
            b[h].live = ms > 0;
            if (!b[h].live) continue;
            const float* from = T[tgt].q + nn[h] * dof;
            for (int j = 0; j < dof; j++) delta[j] = ((smp[j] - from[j]) / d) * res;
            b[h].live = chunk_free(r, s, from, delta, ms);
            for (int j = 0; j < dof; j++) node[h][j] = fmaf(delta[j], (float)(ms - 1), from[j]);
            live += b[h].live;
        }
        if (!live) continue;
        for (int h = 0; h < nb && status == 0; h++) {
            if (!b[h].live) continue;
            if (T[tgt].n >= cap) status = 3;
            else b[h].parent = tree_add(&T[tgt], node[h], nn[h]);
        }
        if (status) break;
        for (int h = 0; h < nb; h++) {
            if (!b[h].live) continue;
            float od;
            b[h].onn = tree_nearest(&T[oth], T[oth].n, node[h], &od);
            memcpy(b[h].cur, node[h], sizeof(float) * dof);
            b[h].dcur = sqrtf(od);
        }
        status = cp_connect(r, s, T, tgt, b, nb, res, max_step, cap, join);
    }
    inf.status = status;
    inf.iterations = (int)(drawn < max_iters ? drawn : max_iters);
    inf.start_tree = T[0].n;
    inf.goal_tree = T[1].n;
    int written = 0;
    if (status == 1) {
        int n1 = 0, n2 = 0;
        for (int v = join[0]; v >= 0; v = T[0].parent[v]) n1++;
        for (int v = join[1]; v >= 0; v = T[1].parent[v]) n2++;
        inf.path_len = n1 + n2;
        if (n1 + n2 <= out_cap) {
            int k = n1 - 1;
            for (int v = join[0]; v >= 0; v = T[0].parent[v], k--)
                memcpy(out_path + k * dof, T[0].q + v * dof, sizeof(float) * dof);
            k = n1;
            for (int v = join[1]; v >= 0; v = T[1].parent[v], k++)
                memcpy(out_path + k * dof, T[1].q + v * dof, sizeof(float) * dof);
            written = n1 + n2;
        }
    }
    tree_free(&T[0]);
    tree_free(&T[1]);
    if (info) *info = inf;
    return written;
}
