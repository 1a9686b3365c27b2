/*
 * odis_oracle.c -- CPU FP64 restatement of GeodesicODIS's linear LTE step.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  It is the checker for the HIP path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  Plain C99, compiled with -ffp-contract=off so every rounding step is
 * the one written here.
 *
 * The reference (hamishHay/GeodesicODIS) cannot be compiled in this image: it
 * needs the Eigen 3 headers (include/mesh.h:13-15), which are absent.  This
 * file therefore restates the algorithm from the reference sources.  Pins:
 *
 *   part                          reference                         pinned by
 *   geometry primitives           include/mathRoutines.h:19-432     closed forms (tests/test_oracle_pins.py)
 *   grid ingest / topology        src/mesh.cpp:436-506, 508-666     Euler counts, V=2N-4, E=3N-6; reference grid files
 *   control-volume area           src/mesh.cpp:1828-1867            sum of areas vs 4 pi R^2 on reference grids
 *   face length / node distance   src/mesh.cpp:545-623              sum d_e l_e / 4 pi R^2 = 1.99873 (l4), 1.99992 (l6)
 *   TRiSK weights                 src/mesh.cpp:677-910              antisymmetry, sum rules, Coriolis analytic field
 *   G, Div, C operators           src/mesh.cpp:2808-3287            analytic fields (src/tests.cpp:38-200) with
 *                                                                     convergence across levels; sum_i A_i Div_i = 0
 *   forcing                       src/tidalPotentials.cpp:29-281    closed forms, degree-2 harmonic check
 *   AB3                           src/temporalOperators.cpp:17-68   exact on quadratic-in-time tendencies
 *   step orchestration            src/timeIntegrator.cpp:205-277    OBLIQ_WEST analytic solution (src/analyticalLTE.cpp)
 *   interpolateVelocity / energy  src/interpolation.cpp:26-62,      closed-form dissipation of analytic fields
 *                                 src/energy.cpp:13-62
 *   dt quantisation, period       src/mesh.cpp:1566-1623,           Enceladus: P=118394 s, dt=29.5985 s, N_s=4000
 *                                 src/globals.cpp:209-214
 *   membraneNuBeta                src/membraneConstants.cpp:8-144   shipped beta table l<=18 at h_s=20 km (in-container);
 *                                                                     limits: no shell and no ocean = FREE_LOADING,
 *                                                                     fluid core, rigid core (cubic in h_s); the
 *                                                                     core-shell coupling terms only loosely
 *   PLANET forcing                src/tidalPotentials.cpp:176-225   pure degree 2; G m R^2 / d^3 towards the companion
 *   FULL2 forcing                 src/tidalPotentials.cpp:133-156   first order: ECC at zero obliquity, OBLIQ a
 *                                                                     quarter period later at zero eccentricity;
 *                                                                     second order: the exact tide of an eccentric,
 *                                                                     inclined orbit (the theta^2 sectoral term as
 *                                                                     the reference writes it); third- and
 *                                                                     fourth-order terms parity unpinned
 *   OBLIQ_WEST closed form        src/analyticalLTE.cpp:133-180     exact continuity by centred differences
 *   warm-start network            (no reference counterpart)        closed forms; gradient by centred differences
 */
#include "odis_oracle.h"

#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

static const double PI = 3.1415926535897932384626433832795028841971693993751058;

/* ------------------------------------------------------------------------ */
/* geometry primitives: include/mathRoutines.h                               */
/* ------------------------------------------------------------------------ */

/* mapFactorAtPoint, include/mathRoutines.h:30-33 */
static double map_factor(double lat1, double lat2, double lon1, double lon2)
{
    return 2.0 / (1.0 + sin(lat1) * sin(lat2) + cos(lat1) * cos(lat2) * cos(lon2 - lon1));
}

/* mapAtPoint, include/mathRoutines.h:61-68 (lat1/lon1 = projection centre) */
static void map_at_point(double *m, double *x, double *y, double lat1, double lat2,
                         double lon1, double lon2, double r)
{
    *m = map_factor(lat1, lat2, lon1, lon2);
    *x = *m * r * (cos(lat2) * sin(lon2 - lon1));
    *y = *m * r * (sin(lat2) * cos(lat1) - cos(lat2) * sin(lat1) * cos(lon2 - lon1));
}

/* distanceBetweenSph, include/mathRoutines.h:133-137 */
static double dist_sph_acos(double lat1, double lat2, double lon1, double lon2, double r)
{
    return r * acos(sin(lat1) * sin(lat2) + cos(lat1) * cos(lat2) * cos(lon2 - lon1));
}

/* sph2cart with colatitude, include/mathRoutines.h:19-25 */
static void sph2cart(double *x, double *y, double *z, double r, double colat, double lon)
{
    *x = r * sin(colat) * cos(lon);
    *y = r * sin(colat) * sin(lon);
    *z = r * cos(colat);
}

/* distanceBetweenSph2 (atan2 form), include/mathRoutines.h:139-160 */
static double dist_sph_atan2(const double *s1, const double *s2)
{
    double a[3], b[3];
    sph2cart(&a[0], &a[1], &a[2], 1.0, PI * 0.5 - s1[0], s1[1]);
    sph2cart(&b[0], &b[1], &b[2], 1.0, PI * 0.5 - s2[0], s2[1]);
    double c = pow(a[1] * b[2] - a[2] * b[1], 2.0);
    c += pow(-(a[0] * b[2] - a[2] * b[0]), 2.0);
    c += pow(a[0] * b[1] - a[1] * b[0], 2.0);
    c = sqrt(c);
    double d = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    return atan2(c, d);
}

/* midpointBetweenSph, include/mathRoutines.h:167-186 */
static void midpoint_sph(const double *s1, const double *s2, double *sc)
{
    double p[3], q[3], c[3];
    p[0] = cos(s1[0]) * cos(s1[1]);
    p[1] = cos(s1[0]) * sin(s1[1]);
    p[2] = sin(s1[0]);
    q[0] = cos(s2[0]) * cos(s2[1]);
    q[1] = cos(s2[0]) * sin(s2[1]);
    q[2] = sin(s2[0]);
    for (int k = 0; k < 3; k++) c[k] = 0.5 * (p[k] + q[k]);
    sc[0] = atan2(c[2], sqrt(pow(c[0], 2.0) + pow(c[1], 2.0)));
    sc[1] = atan2(c[1], c[0]);
}

/* normalVectorBetweenMap, include/mathRoutines.h:208-243 */
static void normal_map(double *n, const double *s1, const double *s2)
{
    double mid[2], c1[2], c2[2];
    midpoint_sph(s1, s2, mid);
    double latc = mid[0], lonc = mid[1];
    double m = map_factor(latc, s1[0], lonc, s1[1]);
    c1[0] = m * 1.0 * (cos(s1[0]) * sin(s1[1] - lonc));
    c1[1] = m * 1.0 * (sin(s1[0]) * cos(latc) - cos(s1[0]) * sin(latc) * cos(s1[1] - lonc));
    m = map_factor(latc, s2[0], lonc, s2[1]);
    c2[0] = m * 1.0 * (cos(s2[0]) * sin(s2[1] - lonc));
    c2[1] = m * 1.0 * (sin(s2[0]) * cos(latc) - cos(s2[0]) * sin(latc) * cos(s2[1] - lonc));
    double xx = c2[0] - c1[0], yy = c2[1] - c1[1];
    double mag = sqrt(xx * xx + yy * yy);
    xx /= mag;
    yy /= mag;
    n[0] = -yy;
    n[1] = xx;
}

/* intersectPointSph, include/mathRoutines.h:303-370 */
static void intersect_sph(const double *s1, const double *s2, const double *s3,
                          const double *s4, double *out)
{
    double c1[3], c2[3], c3[3], c4[3], n1[3], n2[3], mp[3];
    c1[0] = cos(s1[0]) * cos(s1[1]);
    c2[0] = cos(s2[0]) * cos(s2[1]);
    c3[0] = cos(s3[0]) * cos(s3[1]);
    c4[0] = cos(s4[0]) * cos(s4[1]);
    c1[1] = cos(s1[0]) * sin(s1[1]);
    c2[1] = cos(s2[0]) * sin(s2[1]);
    c3[1] = cos(s3[0]) * sin(s3[1]);
    c4[1] = cos(s4[0]) * sin(s4[1]);
    c1[2] = sin(s1[0]);
    c2[2] = sin(s2[0]);
    c3[2] = sin(s3[0]);
    c4[2] = sin(s4[0]);
    n1[0] = c1[1] * c2[2] - c1[2] * c2[1];
    n1[1] = -(c1[0] * c2[2] - c1[2] * c2[0]);
    n1[2] = c1[0] * c2[1] - c1[1] * c2[0];
    n2[0] = c3[1] * c4[2] - c3[2] * c4[1];
    n2[1] = -(c3[0] * c4[2] - c3[2] * c4[0]);
    n2[2] = c3[0] * c4[1] - c3[1] * c4[0];
    double m1 = sqrt(n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2]);
    double m2 = sqrt(n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2]);
    for (int k = 0; k < 3; k++) { n1[k] /= m1; n2[k] /= m2; }
    mp[0] = n1[1] * n2[2] - n2[1] * n1[2];
    mp[1] = n2[0] * n1[2] - n1[0] * n2[2];
    mp[2] = n1[0] * n2[1] - n2[0] * n1[1];
    m1 = sqrt(mp[0] * mp[0] + mp[1] * mp[1] + mp[2] * mp[2]);
    for (int k = 0; k < 3; k++) mp[k] /= m1;
    out[0] = asin(mp[2] / 1.0);
    out[1] = atan2(mp[1], mp[0]);
    if (out[1] < 0.0) out[1] += 2 * PI;
}

/* angleBetweenMap, include/mathRoutines.h:372-400 (atan2 difference form) */
static double angle_map(const double *s1, const double *s2, const double *sc)
{
    double m, x1, y1, x2, y2;
    map_at_point(&m, &x1, &y1, sc[0], s1[0], sc[1], s1[1], 1.0);
    map_at_point(&m, &x2, &y2, sc[0], s2[0], sc[1], s2[1], 1.0);
    double a1 = atan2(y1 - 0.0, x1 - 0.0);
    double a2 = atan2(y2 - 0.0, x2 - 0.0);
    double ang = a1 - a2;
    if (ang > PI) ang -= 2 * PI;
    else if (ang < -PI) ang += 2 * PI;
    return ang;
}

/* triangularArea (map plane), include/mathRoutines.h:403-408 */
static double tri_area_map(double xc, double yc, double x1, double x2, double y1, double y2)
{
    return 0.5 * fabs(xc * (y1 - y2) + x1 * (y2 - yc) + x2 * (yc - y1));
}

/* triangularAreaSph (acos spherical-excess form), include/mathRoutines.h:410-430 */
static double tri_area_sph(double lat1, double lat2, double lat3, double lon1,
                           double lon2, double lon3, double r)
{
    double c = fabs(acos(sin(lat1) * sin(lat2) + cos(lat1) * cos(lat2) * cos(fabs(lon2 - lon1))));
    double a = fabs(acos(sin(lat2) * sin(lat3) + cos(lat2) * cos(lat3) * cos(fabs(lon3 - lon2))));
    double b = fabs(acos(sin(lat3) * sin(lat1) + cos(lat3) * cos(lat1) * cos(fabs(lon1 - lon3))));
    double A = fabs(acos((cos(a) - cos(b) * cos(c)) / (sin(b) * sin(c))));
    double B = fabs(acos((cos(b) - cos(a) * cos(c)) / (sin(a) * sin(c))));
    double C = fabs(acos((cos(c) - cos(b) * cos(a)) / (sin(b) * sin(a))));
    return pow(r, 2.0) * fabs((A + B + C) - PI);
}

/* ------------------------------------------------------------------------ */
/* mesh                                                                      */
/* ------------------------------------------------------------------------ */

typedef struct {
    int N, E, V;
    double R;
    double *lat, *lon;          /* [N]   */
    int *fr;                    /* [N*6] */
    double *clat, *clon;        /* [N*6] */
    int *fnum;                  /* [N] 5 or 6 */
    double *cmap_x, *cmap_y;    /* [N*6] centroid map coords around node */
    double *cv_area;            /* [N]   */
    double *trig;               /* [N*12] cosLat sinLat cosLon sinLon cos2Lat sin2Lat cos2Lon sin2Lon cos2Lat^2.. */
    int *vert;                  /* [N*6] vertex ids */
    double *vlat, *vlon;        /* [V]   */
    int *vnodes;                /* [V*3] */
    int *faces;                 /* [N*6] */
    int *nfdir;                 /* [N*6] */
    int *fnodes;                /* [E*2] */
    double *flen, *fdist;       /* [E]   */
    double *fclat, *fclon;      /* [E]   face centre */
    double *fm;                 /* [E*2] map factor at face centre w.r.t. node 0/1 */
    double *fnorm;              /* [E*2] */
    double *fint;               /* [E*2] intercept (lat, lon) */
    double *farea;              /* [E]   */
    int *icount;                /* [E]   */
    int *ifr;                   /* [E*10] */
    double *iw;                 /* [E*10] */
} mesh_t;

static void *xcalloc(size_t n, size_t s) { return calloc(n ? n : 1, s); }

/* CalcMappingCoords centroid part, src/mesh.cpp:1434-1460 */
static void mesh_mapping(mesh_t *M)
{
    for (int i = 0; i < M->N; i++) {
        for (int j = 0; j < 6; j++) {
            double m, x, y;
            if (M->fr[i * 6 + j] == -1) {
                x = -1.0; y = -1.0;
            } else {
                map_at_point(&m, &x, &y, M->lat[i], M->clat[i * 6 + j], M->lon[i],
                             M->clon[i * 6 + j], M->R);
            }
            M->cmap_x[i * 6 + j] = x;
            M->cmap_y[i * 6 + j] = y;
        }
    }
}

/* CalcControlVolumeArea, src/mesh.cpp:1828-1867 (map-plane polygon area) */
static void mesh_cv_area(mesh_t *M)
{
    for (int i = 0; i < M->N; i++) {
        double m, xc, yc;
        map_at_point(&m, &xc, &yc, M->lat[i], M->lat[i], M->lon[i], M->lon[i], M->R);
        int fn = M->fnum[i];
        double area = 0.0;
        for (int j = 0; j < fn; j++) {
            int k = (j + 1) % fn;
            area += tri_area_map(xc, yc, M->cmap_x[i * 6 + j], M->cmap_x[i * 6 + k],
                                 M->cmap_y[i * 6 + j], M->cmap_y[i * 6 + k]);
        }
        M->cv_area[i] = area;
    }
}

/* CalcTrigFunctions, src/mesh.cpp:2122-2152 */
static void mesh_trig(mesh_t *M)
{
    for (int i = 0; i < M->N; i++) {
        double la = M->lat[i], lo = M->lon[i];
        double *t = M->trig + i * 12;
        t[0] = cos(la);        t[1] = sin(la);
        t[2] = cos(lo);        t[3] = sin(lo);
        t[4] = cos(2.0 * la);  t[5] = sin(2.0 * la);
        t[6] = cos(2.0 * lo);  t[7] = sin(2.0 * lo);
        t[8] = pow(cos(la), 2.0);  t[9] = pow(sin(la), 2.0);
        t[10] = pow(cos(lo), 2.0); t[11] = pow(sin(lo), 2.0);
    }
}

/* CalcControlVolumeVertexR vertex numbering, src/mesh.cpp:436-506 */
static int mesh_vertices(mesh_t *M)
{
    int N = M->N, vc = 0, vmax = 2 * N; /* generous; checked below */
    M->vlat = xcalloc(vmax, sizeof(double));
    M->vlon = xcalloc(vmax, sizeof(double));
    M->vnodes = xcalloc((size_t)vmax * 3, sizeof(int));
    for (int i = 0; i < N * 6; i++) M->vert[i] = -1;
    for (int i = 0; i < N; i++) {
        int fn = M->fnum[i];
        for (int j = 0; j < fn; j++) {
            if (M->vert[i * 6 + j] >= 0) continue;
            if (vc >= vmax) return -2;
            M->vert[i * 6 + j] = vc;
            double vla = M->clat[i * 6 + j], vlo = M->clon[i * 6 + j];
            M->vlat[vc] = vla;
            M->vlon[vc] = vlo;
            int f1 = M->fr[i * 6 + (j % fn)];
            int f2 = M->fr[i * 6 + ((j + 1) % fn)];
            M->vnodes[vc * 3 + 0] = i;
            M->vnodes[vc * 3 + 1] = f1;
            M->vnodes[vc * 3 + 2] = f2;
            int fl[2] = {f1, f2};
            for (int q = 0; q < 2; q++) {
                int f = fl[q];
                for (int k = 0; k < M->fnum[f]; k++) {
                    double a = M->clat[f * 6 + k], b = M->clon[f * 6 + k];
                    if (a < vla + 1e-9 && a > vla - 1e-9 && b < vlo + 1e-9 && b > vlo - 1e-9)
                        M->vert[f * 6 + k] = vc;
                }
            }
            vc++;
        }
    }
    M->V = vc;
    return vc == 2 * N - 4 ? 0 : -3;
}

typedef struct { int id; double angle; } ang_item;

/* descending sort by angle (std::sort with lhs.angle > rhs.angle), <= 6 items */
static void sort_desc(ang_item *a, int n)
{
    for (int i = 1; i < n; i++) {
        ang_item t = a[i];
        int j = i - 1;
        while (j >= 0 && a[j].angle < t.angle) { a[j + 1] = a[j]; j--; }
        a[j + 1] = t;
    }
}

/* AssignFaces, src/mesh.cpp:508-1094 (hot-path parts) */
static int mesh_faces(mesh_t *M)
{
    int N = M->N, E = 3 * N - 6;
    double r = M->R;
    for (int i = 0; i < N * 6; i++) { M->faces[i] = -1; M->nfdir[i] = 0; }
    int fc = 0;
    /* face numbering and per-face geometry: src/mesh.cpp:524-666 */
    for (int i = 0; i < N; i++) {
        int fn = M->fnum[i];
        for (int j = 0; j < fn; j++) {
            if (M->faces[i * 6 + j] >= 0) continue;
            if (fc >= E) return -4;
            M->faces[i * 6 + j] = fc;
            M->fnodes[fc * 2 + 0] = i;
            M->nfdir[i * 6 + j] = 1;
            int f = M->fr[i * 6 + j];
            int jm = (j + fn - 1) % fn;
            double s1[2] = {M->clat[i * 6 + jm], M->clon[i * 6 + jm]};
            double s2[2] = {M->clat[i * 6 + j], M->clon[i * 6 + j]};
            M->flen[fc] = fabs(dist_sph_acos(s1[0], s2[0], s1[1], s2[1], r));
            double sc[2];
            midpoint_sph(s1, s2, sc);
            if (sc[1] < 0.0) sc[1] += 2 * PI;
            M->fclat[fc] = sc[0];
            M->fclon[fc] = sc[1];
            M->fm[fc * 2 + 0] = map_factor(sc[0], M->lat[i], sc[1], M->lon[i]);
            M->fm[fc * 2 + 1] = map_factor(sc[0], M->lat[f], sc[1], M->lon[f]);
            double n1[2] = {M->lat[i], M->lon[i]}, n2[2] = {M->lat[f], M->lon[f]};
            M->fdist[fc] = dist_sph_atan2(n1, n2) * r;
            double nv[2];
            normal_map(nv, s1, s2);
            M->fnorm[fc * 2 + 0] = nv[0];
            M->fnorm[fc * 2 + 1] = nv[1];
            double ip[2];
            intersect_sph(s1, s2, n1, n2, ip);
            M->fint[fc * 2 + 0] = ip[0];
            M->fint[fc * 2 + 1] = ip[1];
            for (int j2 = 0; j2 < M->fnum[f]; j2++) {
                if (M->fr[f * 6 + j2] == i) {
                    M->faces[f * 6 + j2] = fc;
                    M->fnodes[fc * 2 + 1] = f;
                    M->nfdir[f * 6 + j2] = -1;
                }
            }
            fc++;
        }
    }
    if (fc != E) return -5;
    M->E = E;

    /* TRiSK tangential weights, src/mesh.cpp:677-910 */
    for (int e = 0; e < E; e++) {
        int jadd = 0;
        for (int k = 0; k < 2; k++) {
            int nd = M->fnodes[e * 2 + k];
            int fn = M->fnum[nd];
            double sc[2] = {M->lat[nd], M->lon[nd]};
            double si[2] = {M->fint[e * 2], M->fint[e * 2 + 1]};
            ang_item fl[6], vl[6];
            for (int j = 0; j < fn; j++) {
                int id = M->faces[nd * 6 + j];
                double s2[2] = {M->fint[id * 2], M->fint[id * 2 + 1]};
                fl[j].id = id;
                fl[j].angle = angle_map(si, s2, sc) * 180. / PI;
                if (fl[j].angle > 0.0 + 1e-8) fl[j].angle -= 360.0;
            }
            for (int j = 0; j < fn; j++) {
                int vid = M->vert[nd * 6 + j];
                double s2[2] = {M->vlat[vid], M->vlon[vid]};
                vl[j].id = vid;
                vl[j].angle = angle_map(si, s2, sc) * 180. / PI;
                if (vl[j].angle > 0.0 + 1e-8) vl[j].angle -= 360.0;
            }
            sort_desc(fl, fn);
            sort_desc(vl, fn);
            double area_cv = 0.0;
            for (int j = 0; j < fn; j++) {
                int v1 = vl[j].id, v2 = vl[(j + 1) % fn].id;
                area_cv += tri_area_sph(M->lat[nd], M->vlat[v1], M->vlat[v2], M->lon[nd],
                                        M->vlon[v1], M->vlon[v2], r);
            }
            double Rw[6];
            for (int j = 0; j < fn; j++) {
                int v = vl[j].id;
                int f1 = fl[j].id, f2 = fl[(j + 1) % fn].id;
                double a1 = tri_area_sph(M->lat[nd], M->vlat[v], M->fint[f1 * 2], M->lon[nd],
                                         M->vlon[v], M->fint[f1 * 2 + 1], r);
                double a2 = tri_area_sph(M->lat[nd], M->vlat[v], M->fint[f2 * 2], M->lon[nd],
                                         M->vlon[v], M->fint[f2 * 2 + 1], r);
                Rw[j] = (a1 + a2) / area_cv;
            }
            int t_ev = (M->fnodes[e * 2 + 0] == nd) ? 1 : -1;
            for (int j = 0; j < fn; j++) {
                if (fl[j].id == e) continue;
                if (jadd >= 10) return -6;
                double w = 0.0;
                for (int j2 = 0; j2 < j; j2++) w += Rw[j2];
                w -= 0.5;
                for (int j2 = 0; j2 < fn; j2++)
                    if (fl[j].id == M->faces[nd * 6 + j2]) w *= M->nfdir[nd * 6 + j2];
                w *= t_ev;
                M->ifr[e * 10 + jadd] = fl[j].id;
                M->iw[e * 10 + jadd] = w;
                jadd++;
            }
        }
        M->icount[e] = jadd;
        for (int q = jadd; q < 10; q++) { M->ifr[e * 10 + q] = -1; M->iw[e * 10 + q] = 0.0; }
    }

    /* face_area = d_e * l_e (no 1/2), src/mesh.cpp:1093 */
    for (int e = 0; e < E; e++) M->farea[e] = M->fdist[e] * M->flen[e];
    return 0;
}

void *orc_mesh_build(int n, const double *lat, const double *lon, const int *friends,
                     const double *clat, const double *clon, double radius, int *err)
{
    mesh_t *M = xcalloc(1, sizeof(mesh_t));
    int E = 3 * n - 6;
    M->N = n;
    M->R = radius;
    M->lat = xcalloc(n, sizeof(double));
    M->lon = xcalloc(n, sizeof(double));
    M->fr = xcalloc((size_t)n * 6, sizeof(int));
    M->clat = xcalloc((size_t)n * 6, sizeof(double));
    M->clon = xcalloc((size_t)n * 6, sizeof(double));
    memcpy(M->lat, lat, sizeof(double) * n);
    memcpy(M->lon, lon, sizeof(double) * n);
    memcpy(M->fr, friends, sizeof(int) * n * 6);
    memcpy(M->clat, clat, sizeof(double) * n * 6);
    memcpy(M->clon, clon, sizeof(double) * n * 6);
    M->fnum = xcalloc(n, sizeof(int));
    for (int i = 0; i < n; i++) M->fnum[i] = (friends[i * 6 + 5] < 0) ? 5 : 6;
    M->cmap_x = xcalloc((size_t)n * 6, sizeof(double));
    M->cmap_y = xcalloc((size_t)n * 6, sizeof(double));
    M->cv_area = xcalloc(n, sizeof(double));
    M->trig = xcalloc((size_t)n * 12, sizeof(double));
    M->vert = xcalloc((size_t)n * 6, sizeof(int));
    M->faces = xcalloc((size_t)n * 6, sizeof(int));
    M->nfdir = xcalloc((size_t)n * 6, sizeof(int));
    M->fnodes = xcalloc((size_t)E * 2, sizeof(int));
    M->flen = xcalloc(E, sizeof(double));
    M->fdist = xcalloc(E, sizeof(double));
    M->fclat = xcalloc(E, sizeof(double));
    M->fclon = xcalloc(E, sizeof(double));
    M->fm = xcalloc((size_t)E * 2, sizeof(double));
    M->fnorm = xcalloc((size_t)E * 2, sizeof(double));
    M->fint = xcalloc((size_t)E * 2, sizeof(double));
    M->farea = xcalloc(E, sizeof(double));
    M->icount = xcalloc(E, sizeof(int));
    M->ifr = xcalloc((size_t)E * 10, sizeof(int));
    M->iw = xcalloc((size_t)E * 10, sizeof(double));
    int rc;
    mesh_mapping(M);
    mesh_cv_area(M);
    mesh_trig(M);
    rc = mesh_vertices(M);
    if (rc == 0) rc = mesh_faces(M);
    if (err) *err = rc;
    if (rc != 0) { orc_mesh_free(M); return NULL; }
    return M;
}

void orc_mesh_free(void *p)
{
    mesh_t *M = p;
    if (!M) return;
    void *arrs[] = {M->lat, M->lon, M->fr, M->clat, M->clon, M->fnum, M->cmap_x, M->cmap_y,
                    M->cv_area, M->trig, M->vert, M->vlat, M->vlon, M->vnodes, M->faces,
                    M->nfdir, M->fnodes, M->flen, M->fdist, M->fclat, M->fclon, M->fm,
                    M->fnorm, M->fint, M->farea, M->icount, M->ifr, M->iw};
    for (size_t i = 0; i < sizeof(arrs) / sizeof(arrs[0]); i++) free(arrs[i]);
    free(M);
}

void orc_mesh_sizes(void *p, int *n, int *e, int *v)
{
    mesh_t *M = p;
    *n = M->N; *e = M->E; *v = M->V;
}

int orc_mesh_get(void *p, const char *name, void *dst)
{
    mesh_t *M = p;
    struct { const char *nm; void *src; size_t cnt; size_t sz; } tab[] = {
        {"cv_area", M->cv_area, (size_t)M->N, sizeof(double)},
        {"trig", M->trig, (size_t)M->N * 12, sizeof(double)},
        {"faces", M->faces, (size_t)M->N * 6, sizeof(int)},
        {"node_face_dir", M->nfdir, (size_t)M->N * 6, sizeof(int)},
        {"vertexes", M->vert, (size_t)M->N * 6, sizeof(int)},
        {"vertex_lat", M->vlat, (size_t)M->V, sizeof(double)},
        {"vertex_lon", M->vlon, (size_t)M->V, sizeof(double)},
        {"vertex_nodes", M->vnodes, (size_t)M->V * 3, sizeof(int)},
        {"face_nodes", M->fnodes, (size_t)M->E * 2, sizeof(int)},
        {"face_len", M->flen, (size_t)M->E, sizeof(double)},
        {"face_node_dist", M->fdist, (size_t)M->E, sizeof(double)},
        {"face_centre_lat", M->fclat, (size_t)M->E, sizeof(double)},
        {"face_centre_lon", M->fclon, (size_t)M->E, sizeof(double)},
        {"face_centre_m", M->fm, (size_t)M->E * 2, sizeof(double)},
        {"face_normal", M->fnorm, (size_t)M->E * 2, sizeof(double)},
        {"face_intercept", M->fint, (size_t)M->E * 2, sizeof(double)},
        {"face_area", M->farea, (size_t)M->E, sizeof(double)},
        {"interp_count", M->icount, (size_t)M->E, sizeof(int)},
        {"interp_friends", M->ifr, (size_t)M->E * 10, sizeof(int)},
        {"interp_weights", M->iw, (size_t)M->E * 10, sizeof(double)},
    };
    for (size_t i = 0; i < sizeof(tab) / sizeof(tab[0]); i++) {
        if (strcmp(tab[i].nm, name) == 0) {
            if (dst) memcpy(dst, tab[i].src, tab[i].cnt * tab[i].sz);
            return (int)tab[i].cnt;
        }
    }
    return -1;
}

/* Build CSR from per-row (col, val) lists, sorting columns ascending as
 * Eigen's setFromTriplets does (duplicates never occur on these operators). */
static void csr_sort_row(int *col, double *val, int n)
{
    for (int i = 1; i < n; i++) {
        int c = col[i];
        double v = val[i];
        int j = i - 1;
        while (j >= 0 && col[j] > c) { col[j + 1] = col[j]; val[j + 1] = val[j]; j--; }
        col[j + 1] = c;
        val[j + 1] = v;
    }
}

/* CalcGradOperatorCoeffs, src/mesh.cpp:3005-3120 (without the g factor) */
int orc_mesh_grad_csr(void *p, int *rowptr, int *col, double *val)
{
    mesh_t *M = p;
    int nnz = 0;
    for (int e = 0; e < M->E; e++) {
        rowptr[e] = nnz;
        double d = M->fdist[e];
        col[nnz] = M->fnodes[e * 2 + 0];
        val[nnz] = (-M->fm[e * 2 + 0]) / d;
        col[nnz + 1] = M->fnodes[e * 2 + 1];
        val[nnz + 1] = (M->fm[e * 2 + 1]) / d;
        csr_sort_row(col + nnz, val + nnz, 2);
        nnz += 2;
    }
    rowptr[M->E] = nnz;
    return nnz;
}

/* CalcDivOperatorCoeffs, src/mesh.cpp:3177-3287 */
int orc_mesh_div_csr(void *p, int *rowptr, int *col, double *val)
{
    mesh_t *M = p;
    int nnz = 0;
    for (int i = 0; i < M->N; i++) {
        rowptr[i] = nnz;
        int fn = M->fnum[i];
        for (int j = 0; j < fn; j++) {
            int f = M->faces[i * 6 + j];
            col[nnz + j] = f;
            val[nnz + j] = -M->nfdir[i * 6 + j] * M->flen[f] / M->cv_area[i];
        }
        csr_sort_row(col + nnz, val + nnz, fn);
        nnz += fn;
    }
    rowptr[M->N] = nnz;
    return nnz;
}

int orc_mesh_coriolis_nnz(void *p)
{
    mesh_t *M = p;
    int nnz = 0;
    for (int e = 0; e < M->E; e++) nnz += M->icount[e];
    return nnz;
}

/* CalcCoriolisOperatorCoeffs, src/mesh.cpp:2808-2901 */
int orc_mesh_coriolis_csr(void *p, double omega, int *rowptr, int *col, double *val)
{
    mesh_t *M = p;
    int nnz = 0;
    for (int e = 0; e < M->E; e++) {
        rowptr[e] = nnz;
        int c = M->icount[e];
        for (int j = 0; j < c; j++) {
            int f = M->ifr[e * 10 + j];
            col[nnz + j] = f;
            val[nnz + j] = -2.0 * omega * sin(M->fclat[e]) * M->iw[e * 10 + j] * M->flen[f]
                           / M->fdist[e];
        }
        csr_sort_row(col + nnz, val + nnz, c);
        nnz += c;
    }
    rowptr[M->E] = nnz;
    return nnz;
}

/* ------------------------------------------------------------------------ */
/* time stepping: src/timeIntegrator.cpp:57-322                              */
/* ------------------------------------------------------------------------ */

typedef struct { int n; int *rp; int *col; double *val; } csr_t;

typedef struct {
    mesh_t *M;
    orc_params p;
    csr_t G, C, D;
    double *v, *dv, *eta, *de;  /* v[E], dv[E*3], eta[N], de[N*3] */
    double *tmpE, *tmpE2, *U, *uv, *tmpN;
    int iter;
    /* optional pressure-height correction eta_eff = eta + Pb (Pc eta) */
    int pk;
    double *Pb, *Pc, *pcoef, *eta_eff;
} sim_t;

static void csr_alloc(csr_t *A, int rows, int nnz)
{
    A->n = rows;
    A->rp = xcalloc(rows + 1, sizeof(int));
    A->col = xcalloc(nnz, sizeof(int));
    A->val = xcalloc(nnz, sizeof(double));
}

/* Eigen RowMajor sparse * dense: tmp = 0; tmp += (s*a_ij) * x_j in column order */
static void spmv(const csr_t *A, double s, const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < A->n; i++) {
        double t = 0.0;
        for (int k = A->rp[i]; k < A->rp[i + 1]; k++) t += (s * A->val[k]) * x[A->col[k]];
        y[i] = t;
    }
}

void *orc_sim_create(void *mesh, const orc_params *p)
{
    mesh_t *M = mesh;
    sim_t *S = xcalloc(1, sizeof(sim_t));
    S->M = M;
    S->p = *p;
    int E = M->E, N = M->N;
    csr_alloc(&S->G, E, 2 * E);
    orc_mesh_grad_csr(M, S->G.rp, S->G.col, S->G.val);
    int cn = orc_mesh_coriolis_nnz(M);
    csr_alloc(&S->C, E, cn);
    orc_mesh_coriolis_csr(M, p->omega, S->C.rp, S->C.col, S->C.val);
    csr_alloc(&S->D, N, 6 * N);
    orc_mesh_div_csr(M, S->D.rp, S->D.col, S->D.val);
    S->v = xcalloc(E, sizeof(double));
    S->dv = xcalloc((size_t)E * 3, sizeof(double));
    S->eta = xcalloc(N, sizeof(double));
    S->de = xcalloc((size_t)N * 3, sizeof(double));
    S->tmpE = xcalloc(E, sizeof(double));
    S->tmpE2 = xcalloc(E, sizeof(double));
    S->U = xcalloc(N, sizeof(double));
    S->uv = xcalloc((size_t)E * 2, sizeof(double));
    S->tmpN = xcalloc(N, sizeof(double));
    return S;
}

void orc_sim_free(void *p)
{
    sim_t *S = p;
    if (!S) return;
    csr_t *cs[] = {&S->G, &S->C, &S->D};
    for (int i = 0; i < 3; i++) { free(cs[i]->rp); free(cs[i]->col); free(cs[i]->val); }
    free(S->v); free(S->dv); free(S->eta); free(S->de);
    free(S->tmpE); free(S->tmpE2); free(S->U); free(S->uv); free(S->tmpN);
    free(S->Pb); free(S->Pc); free(S->pcoef); free(S->eta_eff);
    free(S);
}

void orc_sim_set_state(void *p, const double *v, const double *dv, const double *eta,
                       const double *de, int iter)
{
    sim_t *S = p;
    int E = S->M->E, N = S->M->N;
    if (v) memcpy(S->v, v, sizeof(double) * E);
    if (dv) memcpy(S->dv, dv, sizeof(double) * E * 3);
    if (eta) memcpy(S->eta, eta, sizeof(double) * N);
    if (de) memcpy(S->de, de, sizeof(double) * N * 3);
    S->iter = iter;
}

void orc_sim_get_state(void *p, double *v, double *dv, double *eta, double *de, int *iter)
{
    sim_t *S = p;
    int E = S->M->E, N = S->M->N;
    if (v) memcpy(v, S->v, sizeof(double) * E);
    if (dv) memcpy(dv, S->dv, sizeof(double) * E * 3);
    if (eta) memcpy(eta, S->eta, sizeof(double) * N);
    if (de) memcpy(de, S->de, sizeof(double) * N * 3);
    if (iter) *iter = S->iter;
}

/* forcing, src/tidalPotentials.cpp:29-281 */
void orc_sim_forcing(void *p, double time, double *U)
{
    sim_t *S = p;
    const orc_params *P = &S->p;
    int N = S->M->N;
    double radius = P->radius;
    double omega = P->omega;
    if (P->surface == ORC_LID_LOVE || P->surface == ORC_LID_MEMBR) radius += P->shell_thickness;
    double cosM = cos(omega * time), sinM = sin(omega * time);
    double cos2M = cos(2 * omega * time);
    double cos3M = cos(3 * omega * time), cos4M = cos(4 * omega * time);
    double ecc = P->ecc, obl = P->obliquity;
    const double *T = S->M->trig;
    switch (P->tide) {
    case ORC_ECC: {
        double factor = 0.75 * P->love_reduct * pow(omega, 2.0) * pow(radius, 2.0) * ecc;
        for (int i = 0; i < N; i++) {
            const double *t = T + i * 12;
            U[i] = factor * ((1. - 3. * t[9]) * cosM + t[8] * (3. * cosM * t[6] + 4. * sinM * t[7]));
        }
    } break;
    case ORC_OBLIQ: {
        double factor = -3. / 2. * P->love_reduct * pow(omega, 2.0) * pow(radius, 2.0) * obl;
        for (int i = 0; i < N; i++) {
            const double *t = T + i * 12;
            U[i] = factor * cosM * t[5] * t[2];
        }
    } break;
    case ORC_OBLIQ_WEST: {
        double factor = 0.5 * P->love_reduct * pow(omega, 2.0) * pow(radius, 2.0) * obl;
        for (int i = 0; i < N; i++) {
            const double *t = T + i * 12;
            U[i] = 3 * factor * t[1] * t[0] * (t[2] * cosM - t[3] * sinM);
        }
    } break;
    case ORC_FULL2: {
        double factor = 1 / 32. * P->love_reduct * pow(omega, 2.0) * pow(radius, 2.0);
        for (int i = 0; i < N; i++) {
            const double *t = T + i * 12;
            double T1, T2, T3;
            T1 = 3. * ecc * (4. - 7. * obl * obl) * cosM
                 + 6 * (obl * obl + ecc * ecc * (3 - 7 * obl * obl)) * cos2M;
            T1 += 3 * ecc * obl * obl * (7 * cos3M + 17 * ecc * cos4M);
            T1 *= -(1 - 3 * t[4]);
            T2 = (4 + 15 * ecc * ecc + 20 * ecc * cosM + 43 * ecc * ecc * cos2M) * t[2];
            T2 += 2 * ecc * (4 + 25 * ecc * cosM) * sinM * t[3];
            T2 *= 24 * obl * t[0] * t[1] * sinM;
            T3 = obl * obl * (2 + 3 * ecc * ecc + 6 * ecc * cosM + 9 * ecc * ecc * cos2M)
                 * (cosM * t[2] + sinM * t[3]);
            T3 += -(obl * obl - 2)
                  * ((6 * ecc * cosM + 17 * ecc * ecc * cos2M) * t[6]
                     + 2 * ecc * (4 + 17 * ecc * cosM) * sinM * t[7]);
            T3 *= 6 * t[8];
            U[i] = factor * (T1 + T2 + T3);
        }
    } break;
    case ORC_FULL: {
        double factor = 0.75 * P->love_reduct * pow(omega, 2.0) * pow(radius, 2.0) * ecc;
        double factor2 = -3. / 2. * P->love_reduct * pow(omega, 2.0) * pow(radius, 2.0) * obl;
        for (int i = 0; i < N; i++) {
            const double *t = T + i * 12;
            U[i] = factor * ((1 - 3 * t[9]) * cosM + t[8] * (3 * cosM * t[6] + 4 * sinM * t[7]))
                   + factor2 * cosM * t[5] * t[2];
        }
    } break;
    case ORC_PLANET: {
        double m2 = 8.931938e+22, a1 = 421800000.0, a2 = P->semimajor;
        double n2 = omega, n1 = n2 * 2.0, nij = n1 - n2;
        double cosnt = cos(nij * time), sinnt = sin(nij * time);
        double pp = pow(a1, 2.0) + pow(a2, 2.0) - 2. * a1 * a2 * cosnt;
        double cosphi = a1 - a2 * cosnt, sinphi = a2 * sinnt;
        double factor = 0.5 * 6.67408e-11 * m2 * pow(radius / pp, 2.0) / sqrt(pp);
        for (int i = 0; i < N; i++) {
            const double *t = T + i * 12;
            double cg = t[0] * (t[2] * cosphi + t[3] * sinphi);
            U[i] = factor * (3. * pow(cg, 2.0) - pp);
        }
    } break;
    default:
        /* Potential types with no live case leave the zero-initialised
         * potential untouched (Array1D zero-initialises, include/array1d.h:99). */
        break;
    }
}

/* interpolateVelocity, src/interpolation.cpp:26-62 */
void orc_sim_interp_velocity(void *p, const double *vn, double *uv)
{
    sim_t *S = p;
    mesh_t *M = S->M;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < M->E; e++) {
        double vt = 0.0;
        int c = M->icount[e];
        for (int j = 0; j < c; j++) {
            int f = M->ifr[e * 10 + j];
            vt += vn[f] * M->iw[e * 10 + j] * M->flen[f];
        }
        vt /= M->fdist[e];
        double nx = M->fnorm[e * 2], ny = M->fnorm[e * 2 + 1];
        double ty = -nx, tx = ny;
        uv[e * 2 + 0] = nx * vn[e] + tx * vt;
        uv[e * 2 + 1] = ny * vn[e] + ty * vt;
    }
}

/* updateEnergy, src/energy.cpp:13-62 */
static double energy(sim_t *S, const double *uv)
{
    mesh_t *M = S->M;
    const orc_params *P = &S->p;
    double avg = 0.0;
    if (P->friction == ORC_LINEAR) {
        for (int e = 0; e < M->E; e++) {
            double u = uv[e * 2], v = uv[e * 2 + 1];
            double ef = P->alpha * 1000.0 * P->h * (u * u + v * v);
            avg += ef * M->farea[e];
        }
    } else {
        for (int e = 0; e < M->E; e++) {
            double u = uv[e * 2], v = uv[e * 2 + 1];
            double ef = P->alpha / P->h * sqrt(u * u + v * v) * (u * u + v * v);
            avg += ef * M->farea[e];
        }
    }
    return avg / (4 * PI * pow(P->radius, 2.0));
}

double orc_sim_dissipation(void *p)
{
    sim_t *S = p;
    orc_sim_interp_velocity(S, S->v, S->uv);
    return energy(S, S->uv);
}

/* Self-gravity / shell coupling of the pressure term (the reference's
 * commented pressureGradientSH design, src/spatialOperators.cpp:387-460):
 * eta_eff = eta + Pb (Pc eta) with Pb [N x K] the beta-weighted synthesis
 * basis and Pc [K x N] the least-squares analysis operator, both supplied by
 * the caller.  Plain row-by-row dense products. */
void orc_sim_set_projector(void *p, int K, const double *Pb, const double *Pc)
{
    sim_t *S = p;
    int N = S->M->N;
    free(S->Pb); free(S->Pc); free(S->pcoef); free(S->eta_eff);
    S->pk = K;
    S->Pb = NULL; S->Pc = NULL; S->pcoef = NULL; S->eta_eff = NULL;
    if (K <= 0) return;
    S->Pb = xcalloc((size_t)N * K, sizeof(double));
    S->Pc = xcalloc((size_t)K * N, sizeof(double));
    S->pcoef = xcalloc(K, sizeof(double));
    S->eta_eff = xcalloc(N, sizeof(double));
    memcpy(S->Pb, Pb, sizeof(double) * N * K);
    memcpy(S->Pc, Pc, sizeof(double) * K * N);
}

static const double *effective_height(sim_t *S, const double *eta)
{
    if (S->pk <= 0) return eta;
    int N = S->M->N, K = S->pk;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < K; k++) {
        double t = 0.0;
        for (int i = 0; i < N; i++) t += S->Pc[(size_t)k * N + i] * eta[i];
        S->pcoef[k] = t;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        double t = 0.0;
        for (int k = 0; k < K; k++) t += S->Pb[(size_t)i * K + k] * S->pcoef[k];
        S->eta_eff[i] = eta[i] + t;
    }
    return S->eta_eff;
}

/* updateMomentum linear branch, src/updateMomentum.cpp:42 */
void orc_sim_momentum(void *p, const double *v, const double *eta, double *dvdt)
{
    sim_t *S = p;
    int E = S->M->E;
    eta = effective_height(S, eta);
    spmv(&S->G, -S->p.g, eta, S->tmpE2);
    spmv(&S->C, 1.0, v, dvdt);
    for (int e = 0; e < E; e++) dvdt[e] = S->tmpE2[e] + dvdt[e];
}

/* integrateAB3scalar, src/temporalOperators.cpp:17-68 */
static void ab3(double *s, double *ds, int iter, int n, double dt, int init_load)
{
    const double a = 23. / 12., b = -16. / 12., c = 5. / 12.;
    if (iter > 1 || init_load) {
        for (int i = 0; i < n; i++) {
            s[i] += (a * ds[i * 3] + b * ds[i * 3 + 1] + c * ds[i * 3 + 2]) * dt;
            ds[i * 3 + 2] = ds[i * 3 + 1];
            ds[i * 3 + 1] = ds[i * 3];
        }
    } else if (iter == 0) {
        for (int i = 0; i < n; i++) {
            s[i] += ds[i * 3] * dt;
            ds[i * 3 + 2] = ds[i * 3];
        }
    } else {
        for (int i = 0; i < n; i++) {
            s[i] += ds[i * 3] * dt;
            ds[i * 3 + 1] = ds[i * 3];
        }
    }
}

/* One pass of the ab3Explicit loop body, src/timeIntegrator.cpp:207-277.
 * quadraticDrag (src/drag.cpp:19-38) is applied to the momentum tendency only
 * when quad_drag_momentum is set (the reference never calls it in AB3). */
int orc_sim_step(void *p, int nsteps, double *diss)
{
    sim_t *S = p;
    mesh_t *M = S->M;
    const orc_params *P = &S->p;
    int E = M->E, N = M->N;
    double dt = P->dt;
    for (int s = 0; s < nsteps; s++) {
        double t = dt * S->iter;
        orc_sim_momentum(S, S->v, S->eta, S->tmpE);
        if (P->quad_drag_momentum) {
            orc_sim_interp_velocity(S, S->v, S->uv);
            double ih = P->alpha / P->h;
            for (int e = 0; e < E; e++) {
                double u = S->uv[e * 2], w = S->uv[e * 2 + 1];
                S->tmpE[e] -= ih * sqrt(u * u + w * w) * S->v[e];
            }
        }
        for (int e = 0; e < E; e++) S->dv[e * 3] = S->tmpE[e];
        orc_sim_forcing(S, t + dt, S->U);
        /* drag_term = D v + G U with D = -alpha I (src/mesh.cpp:2963) */
        double dcoef = P->quad_drag_momentum ? 0.0 : -P->alpha;
        spmv(&S->G, 1.0, S->U, S->tmpE2);
        for (int e = 0; e < E; e++) S->tmpE[e] = (0.0 + dcoef * S->v[e]) + S->tmpE2[e];
        ab3(S->v, S->dv, S->iter, E, dt, P->init_load);
        for (int e = 0; e < E; e++) S->v[e] += dt * S->tmpE[e];
        spmv(&S->D, P->h, S->v, S->tmpN);
        for (int i = 0; i < N; i++) S->de[i * 3] = S->tmpN[i];
        ab3(S->eta, S->de, S->iter, N, dt, P->init_load);
        if (diss) diss[s] = orc_sim_dissipation(S);
        S->iter++;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* scalar helpers                                                            */
/* ------------------------------------------------------------------------ */

/* Period rounding to an even number of seconds, src/globals.cpp:209-214 */
void orc_round_period(double ang_vel, double *period, double *omega)
{
    double per = 2. * PI / ang_vel;
    int it = (int)round(per / 2) * 2;
    *period = (double)it;
    *omega = 2 * PI / (double)it;
}

/* CalcMaxTimeStep quantisation, src/mesh.cpp:1601-1618 */
void orc_quantise_dt(double period, double target, double *dt, int *n_steps)
{
    int num = 100;
    double d = period;
    while (d > target) {
        d = period / num;
        num += 100;
    }
    num -= 100;
    *dt = d;
    *n_steps = num;
}

/* analyticalLTE OBLIQ_WEST case, src/analyticalLTE.cpp:30-180 */
void orc_analytical_obliq_west(const orc_params *p, double lat, double lon, double t, double *o)
{
    double radius = p->radius, Om = p->omega, g = p->g, h = p->h;
    double w = -Om, obl = p->obliquity, alpha = p->alpha;
    double lam = w / (2 * Om);
    double lambs = 4 * Om * Om * radius * radius / (g * h);
    double complex J = I;
    /* pn(2,1), qn(1,1) (src/analyticalLTE.cpp:30-36) */
    double pn21 = (2 + 1) * (2 + 1) / (2.0 * (2 * 2 + 1));
    double qn11 = 1 * (1 + 1 - 1) / ((1 + 1.0) * (2 * 1 + 1));
    /* Ln(1,1), Kn(2,1) (src/analyticalLTE.cpp:38-44) */
    double complex L1 = lam + 1.0 / (1 * 2) + J * alpha / (2 * Om);
    double complex K2 = lam + 1.0 / (2 * 3) - 2 * 3 / (lambs * lam) + J * alpha / (2 * Om);
    double complex U21 = 0.5 * Om * Om * radius * radius * obl;
    double complex Psi11 = U21 / (2 * Om) / (qn11 - K2 * L1 / pn21);
    double complex Phi21 = -J * L1 / pn21 * Psi11;
    double colat = PI * 0.5 - lat;
    double complex eil = cexp(J * 1.0 * lon);
    double complex Y31 = 1.5 * (5 * pow(cos(colat), 2.0) - 1) * sin(colat) * eil;
    double complex Y21 = 3 * cos(colat) * sin(colat) * eil;
    double complex Y11 = sin(colat) * eil;
    double complex ew = cexp(-J * w * t);
    double complex U = (-1 / 3. * (Psi11 * Y21) + J * Phi21 * Y21) * ew;
    double complex dU = -J * w * U;
    U += conj(U);
    U /= 2 * radius * sin(colat);
    dU += conj(dU);
    dU /= 2 * radius * sin(colat);
    double complex V = -(J * Psi11 * Y11 + Phi21 * (4.0 * Y31 - 9.0 * Y11) / 5.) * ew;
    double complex dV = -J * w * V;
    V += conj(V);
    V /= 2 * radius * sin(colat);
    dV += conj(dV);
    dV /= 2 * radius * sin(colat);
    double complex ETA = J * 6.0 / pow(radius, 2.0) * h / w * Phi21 * Y21 * ew;
    double complex dE = -J * w * ETA;
    ETA += conj(ETA);
    ETA *= 0.5;
    dE += conj(dE);
    dE *= 0.5;
    o[0] = creal(U); o[1] = creal(V); o[2] = creal(dU);
    o[3] = creal(dV); o[4] = creal(ETA); o[5] = creal(dE);
}

/* membraneNuBeta, src/membraneConstants.cpp:8-144, with the body's mass and
 * the core and shell rigidities as arguments (the reference hard-codes
 * Enceladus: 1.08e20 kg, 40 GPa, 3.5 GPa), so that the tests can drive the
 * formula to its fluid-core, rigid-core and vanishing-shell limits. */
void orc_membrane_nu_beta_gen(double radius, double hs, double ho, int l_max, double mass_total,
                              double rig_core, double rig_shell, double *nu, double *beta,
                              double *radius_ocean, double *grav_ocean)
{
    double G = 6.67384e-11, pois = 0.5;
    double rc = radius - (hs + ho), ro = radius - hs;
    double den_ocean = 1000.0, den_shell = 940.0;
    double grav_surf = G * mass_total / pow(radius, 2.0);
    double vol_total = 4. / 3. * PI * pow(radius, 3.0);
    double vol_core = 4. / 3. * PI * pow(rc, 3.0);
    double vol_ocean = 4. / 3. * PI * pow(ro, 3.0) - vol_core;
    double vol_shell = vol_total - 4. / 3. * PI * pow(ro, 3.0);
    double den_bulk = mass_total / vol_total;
    double mass_ocean = vol_ocean * den_ocean, mass_shell = vol_shell * den_shell;
    double mass_core = mass_total - (mass_ocean + mass_shell);
    double den_core = mass_core / vol_core;
    double grav_core = G * mass_core / pow(rc, 2.0);
    *grav_ocean = G * (mass_core + mass_ocean) / pow(ro, 2.0);
    *radius_ocean = ro;
    for (int l = 0; l <= l_max; l++) {
        double rig_eff = (double)(2 * l * l + 4 * l + 3) / ((double)l) * rig_core
                         / (den_core * grav_core * rc);
        double rf = 1. / (1. + rig_eff);
        double kt = rf * 3. / ((double)(2 * (l - 1)));
        double ht = rf * (double)(2 * l + 1) / ((double)(2 * (l - 1)));
        if (l == 1) { kt = 0.0; ht = 0.0; }
        double kl = -rf, hl = -rf * (double)(2 * l + 1) / 3.0;
        double gam_tide = 1.0 + kt - ht, gam_load = 1.0 + kl - hl;
        double x = (double)((l - 1) * (l + 2));
        double bend = rig_shell * pow(hs, 3.0) / (6. * (1. - pois));
        double smem = 2. * x * (1. + pois) / (x + 1. + pois) * rig_shell
                      / (den_ocean * grav_surf * radius) * hs / radius;
        double sbend = pow(x, 2.0) * (x + 2.) / (x + 1. + pois) * bend
                       / (den_ocean * grav_surf * pow(radius, 4.0));
        double spr = smem + sbend;
        double xi = 3.0 / (2. * (double)l + 1.0) * (den_ocean / den_bulk);
        double dspr = 1. - pow((1. + xi * hl), 2.0) / (1. + xi * (ht - hl) * spr);
        dspr *= -spr;
        double dgam = (1. + xi * hl) * ht / (1. + xi * (ht - hl) * spr);
        dgam *= -spr;
        beta[l] = 1. - xi * gam_load + spr + dspr;
        nu[l] = gam_tide + dgam;
    }
}

/* membraneNuBeta with the reference's Enceladus constants
 * (src/membraneConstants.cpp:8-144). */
void orc_membrane_nu_beta(double radius, double hs, double ho, int l_max, double *nu,
                          double *beta, double *radius_ocean, double *grav_ocean)
{
    orc_membrane_nu_beta_gen(radius, hs, ho, l_max, 1.08e20, 40e9, 3.5e9, nu, beta, radius_ocean,
                             grav_ocean);
}

void orc_set_threads(int n) {
  if (n >= 1) omp_set_num_threads(n);
}

int orc_get_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------------ */
/* FP64 restatement of the warm-start network (no reference counterpart):   */
/* n_in -> width x depth (tanh) -> n_out (linear); parameters per layer l:  */
/* W_l [out][in] row-major, then b_l [out]; loss = sum (y - t)^2 / (2 n     */
/* n_out).  The checker of the GPU network's forward pass and gradients.    */
/* ------------------------------------------------------------------------ */

static void mlp_dims(int n_in, int width, int depth, int n_out, int l, int *ni, int *no)
{
    *ni = l == 0 ? n_in : width;
    *no = l == depth ? n_out : width;
}

/* activations of n rows: a[0] = x, a[l + 1] = layer l output; returns malloc'd block */
static double **mlp_run(int n_in, int width, int depth, int n_out, const double *p, int n, const double *x)
{
    double **a = calloc((size_t)depth + 2, sizeof(double *));
    a[0] = malloc(sizeof(double) * (size_t)n * n_in);
    memcpy(a[0], x, sizeof(double) * (size_t)n * n_in);
    size_t off = 0;
    for (int l = 0; l <= depth; l++) {
        int ni, no;
        mlp_dims(n_in, width, depth, n_out, l, &ni, &no);
        const double *W = p + off, *b = p + off + (size_t)ni * no;
        off += (size_t)ni * no + no;
        a[l + 1] = malloc(sizeof(double) * (size_t)n * no);
        for (int r = 0; r < n; r++)
            for (int j = 0; j < no; j++) {
                double s = 0.0;
                for (int k = 0; k < ni; k++) s += a[l][(size_t)r * ni + k] * W[(size_t)j * ni + k];
                s += b[j];
                a[l + 1][(size_t)r * no + j] = l < depth ? tanh(s) : s;
            }
    }
    return a;
}

static void mlp_free(double **a, int depth)
{
    for (int l = 0; l <= depth + 1; l++) free(a[l]);
    free(a);
}

void orc_mlp_forward(int n_in, int width, int depth, int n_out, const double *p, int n, const double *x,
                     double *y)
{
    double **a = mlp_run(n_in, width, depth, n_out, p, n, x);
    memcpy(y, a[depth + 1], sizeof(double) * (size_t)n * n_out);
    mlp_free(a, depth);
}

double orc_mlp_loss_grad(int n_in, int width, int depth, int n_out, const double *p, int n, const double *x,
                         const double *t, double *grad)
{
    double **a = mlp_run(n_in, width, depth, n_out, p, n, x);
    const double cnt = (double)n * n_out;
    double loss = 0.0;
    int wmax = width > n_out ? width : n_out;
    if (n_in > wmax) wmax = n_in;
    double *dz = malloc(sizeof(double) * (size_t)n * wmax), *dx = malloc(sizeof(double) * (size_t)n * wmax);
    for (size_t i = 0; i < (size_t)n * n_out; i++) {
        double d = a[depth + 1][i] - t[i];
        loss += d * d;
        dz[i] = d / cnt;
    }
    loss /= 2.0 * cnt;
    size_t *offs = malloc(sizeof(size_t) * ((size_t)depth + 1));
    size_t off = 0;
    for (int l = 0; l <= depth; l++) {
        int ni, no;
        mlp_dims(n_in, width, depth, n_out, l, &ni, &no);
        offs[l] = off;
        off += (size_t)ni * no + no;
    }
    for (int l = depth; l >= 0; l--) {
        int ni, no;
        mlp_dims(n_in, width, depth, n_out, l, &ni, &no);
        if (l < depth)  /* through tanh: dz *= 1 - a^2 */
            for (size_t i = 0; i < (size_t)n * no; i++) dz[i] *= 1.0 - a[l + 1][i] * a[l + 1][i];
        double *gW = grad + offs[l], *gb = grad + offs[l] + (size_t)ni * no;
        const double *W = p + offs[l];
        for (int j = 0; j < no; j++) {
            double sb = 0.0;
            for (int r = 0; r < n; r++) sb += dz[(size_t)r * no + j];
            gb[j] = sb;
            for (int k = 0; k < ni; k++) {
                double s = 0.0;
                for (int r = 0; r < n; r++) s += dz[(size_t)r * no + j] * a[l][(size_t)r * ni + k];
                gW[(size_t)j * ni + k] = s;
            }
        }
        if (l > 0) {
            for (int r = 0; r < n; r++)
                for (int k = 0; k < ni; k++) {
                    double s = 0.0;
                    for (int j = 0; j < no; j++) s += dz[(size_t)r * no + j] * W[(size_t)j * ni + k];
                    dx[(size_t)r * ni + k] = s;
                }
            double *tmp = dz;
            dz = dx;
            dx = tmp;
        }
    }
    free(offs);
    free(dz);
    free(dx);
    mlp_free(a, depth);
    return loss;
}
