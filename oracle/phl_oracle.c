/*
 * phl_oracle.c -- CPU oracle for the unstructured residual + LU-SGS path.
 *
 * TEST INFRASTRUCTURE ONLY: a sequential FP64 restatement of the reference
 * solver (PHengLEI NSSolverUnstruct, one zone, perfect gas) used to check the
 * HIP implementation.  It is loaded by tests/, by the smoke check and by the
 * CPU-baseline leg of bench.py, never by the product library.
 *
 * Build with -O2 -ffp-contract=off so that every multiply and add rounds as
 * written, like the reference's own x86-64 build.  Built with -fopenmp as
 * well (liborc_omp.so), the same loops run on all host cores with results
 * identical to the sequential build (see the OpenMP notes below).
 *
 * Index conventions follow the reference: faces 0..nBFace-1 are boundary
 * faces, the ghost cell of boundary face f is nCell+f
 * (API/Geometry/src/Geo_UnstructGrid.cpp:2007), cell arrays are SoA over
 * nTotal = nCell+nBFace, primitive variables are (rho,u,v,w,p)
 * (API/Common/include/Constants.h:252-282).
 */
#include "phl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IR 0
#define IU 1
#define IV 2
#define IW 3
#define IP 4

static const double SMALL = 1.0e-40; /* API/Common/include/Precision.h:25 */
static const double TINY  = 1.0e-40; /* API/Common/include/Precision.h:26 */
static const double LARGE = 1.0e40;  /* API/Common/include/Precision.h:24 */
static const double PI_   = 3.141592653589793; /* Constants.h:110 */

/* ------------------------------------------------------------------ */
/* OpenMP build (liborc_omp.so, the CPU baseline over the host's cores): */
/* thread t owns a contiguous range of cells (a ghost belongs to its     */
/* face's owner).  A face loop runs, on every thread, the faces that     */
/* touch one of its cells in increasing order and applies a face's       */
/* contributions only to the cells the thread owns, so each cell still   */
/* receives its faces in increasing face order, exactly as the           */
/* sequential loop adds them: the results are bit for bit those of the   */
/* sequential build.  A face between two threads' cells is evaluated by  */
/* both.  The LU-SGS sweeps run level by level (the cells of a level are */
/* independent); SolutionFix stays sequential.  Without OpenMP the       */
/* macros below reduce to the plain sequential loops.                    */
/* ------------------------------------------------------------------ */
struct orc_part;
#ifdef _OPENMP
#include <omp.h>
struct orc_part {
    int nThreads;
    int *owner;   /* [nTotal] thread of each cell and ghost */
    int *fptr;    /* [nThreads+1] */
    int *flist;   /* faces of each thread, increasing */
    int *lvPtrF, *lvCellsF, nLvF;  /* forward sweep levels (cells in increasing order within a level) */
    int *lvPtrB, *lvCellsB, nLvB;  /* backward sweep levels (cells in decreasing order) */
};
static struct orc_part *orc_partition(orc_solver *s);
#define FACE_LOOP(s, f0, f1, f)                                                                  \
    {                                                                                            \
        const struct orc_part *P_ = orc_partition(s);                                           \
        _Pragma("omp parallel num_threads(P_->nThreads)")                                        \
        for (int T_ = omp_get_thread_num(); T_ < P_->nThreads; T_ += omp_get_num_threads())      \
            for (int i_ = P_->fptr[T_]; i_ < P_->fptr[T_ + 1]; ++i_) {                             \
                const int f = P_->flist[i_];                                                      \
                if (f < (f0) || f >= (f1)) continue;
#define FACE_LOOP_END }}
#define OWN(c) (P_->owner[c] == T_)
#define CELL_LOOP _Pragma("omp parallel for schedule(static)")
#else
#define FACE_LOOP(s, f0, f1, f) { for (int f = (f0); f < (f1); ++f) {
#define FACE_LOOP_END }}
#define OWN(c) 1
#define CELL_LOOP
#endif

static inline double MIN(double a, double b) { return (a < b) ? a : b; } /* Math_BasisFunction.h:31 */
static inline double MAX(double a, double b) { return (a > b) ? a : b; } /* Math_BasisFunction.h:44 */
static inline double ABS(double a) { return (a < 0) ? -a : a; }          /* Math_BasisFunction.h:58 */

struct orc_solver {
    orc_params p;
    int nCell, nBFace, nFace, nTotal;
    int *left, *right, *bcType;
    double *xfc, *yfc, *zfc, *xfn, *yfn, *zfn, *area;
    double *xcc, *ycc, *zcc, *vol, *skew;
    double avgVol;
    int dim;  /* 3, or 2 for a 2-D grid (orc_set_dim) */
    int initPhase;  /* isInIniting: first-order faces and the Van Leer flux (orc_set_init_phase) */
    double nsSweepNormals[8], saSweepNormals[8];  /* squared changes of each sweep pair (first 8) */
    int nsSweepsDone, saSweepsDone;
    int *c2fPtr, *c2f;
    /* flow fields */
    double *q[5], *res[5], *dqdx[5], *dqdy[5], *dqdz[5], *dq[5], *dflux[5];
    double *lim, *specInv, *specVis, *dt, *diag, *gama, *t, *visl, *vist;
    double *dtdx, *dtdy, *dtdz;
    /* pressure factor of the pressure-sensor entropy fixes (rtem; methods 2, 6) */
    double *rtem;
    /* least-squares gradient (gradient == 1): face weights and the inverse
     * normal-matrix entries per cell */
    double *iwt, *ixx, *iyy, *izz, *ixy, *ixz, *iyz;
    /* node-based Green-Gauss gradient (gradient == 2): node coordinates,
     * face and cell node lists, node values of the primitives and of T,
     * and whether the next node-value pass first makes t match q */
    int nNode;
    double *xn, *yn, *zn;
    int *fnPtr, *fn, *cnPtr, *cn;
    double *qNode[5], *tNode, *nodeWeight;
    int *nodeBC;
    int tFresh;
    /* Spalart-Allmaras (viscous == 3): nu~ and its residual, spectrum,
     * face matrices, increments, gradient and the wall distance */
    double *nut, *resT, *specT, *matL, *matR, *dqT, *dfluxT, *gnx, *gny, *gnz, *wdist;
    /* OpenMP build: thread partition and sweep levels (NULL otherwise) */
    struct orc_part *part;
    /* derived constants */
    double primInf[5];
    double R;           /* coefficientOfStateEquation, PerfectGas.cpp:1401 */
    double pTotal;      /* PerfectGas.cpp:1616-1641 */
    double entropyCoef; /* Param_NSSolver.cpp:225-266 */
    double tsuth;       /* 110.4 / T_ref, PerfectGas.cpp:1480-1494 */
};

/* ------------------------------------------------------------------ */
/* Reference constants                                                 */
/* ------------------------------------------------------------------ */

/* Param_NSSolver::SetEntropyFixCoefficients, CFD/NSSolver/src/Param_NSSolver.cpp:225-266 */
double orc_entropy_coef(double mach, double scale)
{
    double e;
    if (mach <= 0.6) {
        e = 0.0001;
    } else if (mach <= 2.0) {
        double x0 = 0.6, x1 = 2.0, y0 = 0.0001, y1 = 0.01;
        e = y0 * (mach - x1) / (x0 - x1) + y1 * (mach - x0) / (x1 - x0);
    } else if (mach <= 5.0) {
        double x0 = 2.0, x1 = 5.0, y0 = 0.01, y1 = 0.5;
        e = y0 * (mach - x1) / (x0 - x1) + y1 * (mach - x0) / (x1 - x0);
    } else {
        e = 1.0;
    }
    return e * scale;
}

/* PerfectGas::ComputePTotal, API/Physics/src/PerfectGas.cpp:1616-1641 */
double orc_ptotal(double mach, double gama)
{
    double R = 1.0 / (gama * mach * mach);
    double pref = R * 1.0 * 1.0 / 1.0;
    double m12 = mach * mach;
    double gamp1 = gama + 1.0;
    double gamm1 = gama - 1.0;
    double c1 = 2.0 / gamm1 + m12;
    double c2 = 2.0 * gama / gamm1 * m12 - 1.0;
    double cc = gama / (gama - 1.0);
    double t0t = 1.0 + 0.5 * gamm1 * c1 / c2;
    if (mach > 1.0) pref *= (2.0 * gama / gamp1 * m12 - gamm1 / gamp1);
    return pref * pow(t0t, cc);
}

/* ------------------------------------------------------------------ */
/* Geometry                                                            */
/* ------------------------------------------------------------------ */

/* 2-D grids (UnstructGrid::ComputeMetrics, Geo_UnstructGrid.cpp:5695-5709):
 * ComputeFaceNormal2D (:6158-6187) and ComputeFaceCenter2D (:6370-6390).  A
 * face is the edge p1 -> p2; normal (y2 - y1, x1 - x2)/|.|, area its length,
 * centre its midpoint. */
void orc_face_metrics_2d(int nFace, const int *fp, const int *fn,
                         const double *x, const double *y, const double *z,
                         double *xfc, double *yfc, double *zfc,
                         double *xfn, double *yfn, double *zfn, double *area)
{
    for (int f = 0; f < nFace; ++f) {
        int p1 = fn[fp[f]], p2 = fn[fp[f] + 1];
        xfn[f] = y[p2] - y[p1];
        yfn[f] = x[p1] - x[p2];
        zfn[f] = 0.0;
        area[f] = sqrt(xfn[f] * xfn[f] + yfn[f] * yfn[f] + zfn[f] * zfn[f]);
        double oarea = 1.0 / (area[f] + SMALL);
        xfn[f] *= oarea;
        yfn[f] *= oarea;
        zfn[f] *= oarea;
    }
    for (int f = 0; f < nFace; ++f) {
        int p1 = fn[fp[f]], p2 = fn[fp[f] + 1];
        xfc[f] = 0.5 * (x[p1] + x[p2]);
        yfc[f] = 0.5 * (y[p1] + y[p2]);
        zfc[f] = 0.5 * (z[p1] + z[p2]);
    }
}

/* ComputeCellCenterVol2D (Geo_UnstructGrid.cpp:7058-7190): the divergence
 * theorem over the edges, boundary faces then interior faces (x . n) A
 * accumulated, centroid / (1.5 sum), area sum / 2; ghosts mirrored across
 * their face (degenerate faces: point reflection), ghost area = owner's. */
void orc_cell_metrics_2d(int nCell, int nBFace, int nFace, const int *left, const int *right,
                         const double *xfc, const double *yfc, const double *zfc,
                         const double *xfn, const double *yfn, const double *zfn, const double *area,
                         double *xcc, double *ycc, double *zcc, double *vol)
{
    int nT = nCell + nBFace;
    for (int c = 0; c < nT; ++c) { xcc[c] = 0.0; ycc[c] = 0.0; zcc[c] = 0.0; }
    for (int c = 0; c < nCell; ++c) vol[c] = 0.0;
    for (int f = 0; f < nBFace; ++f) {
        int le = left[f];
        double dot = (xfc[f] * xfn[f] + yfc[f] * yfn[f] + zfc[f] * zfn[f]) * area[f];
        xcc[le] += xfc[f] * dot; ycc[le] += yfc[f] * dot; zcc[le] += zfc[f] * dot;
        vol[le] += dot;
    }
    for (int f = nBFace; f < nFace; ++f) {
        int le = left[f], re = right[f];
        double dot = (xfc[f] * xfn[f] + yfc[f] * yfn[f] + zfc[f] * zfn[f]) * area[f];
        xcc[le] += xfc[f] * dot; ycc[le] += yfc[f] * dot; zcc[le] += zfc[f] * dot;
        vol[le] += dot;
        xcc[re] -= xfc[f] * dot; ycc[re] -= yfc[f] * dot; zcc[re] -= zfc[f] * dot;
        vol[re] -= dot;
    }
    for (int c = 0; c < nCell; ++c) {
        double tmp = 1.0 / (1.5 * vol[c] + SMALL);
        xcc[c] *= tmp; ycc[c] *= tmp; zcc[c] *= tmp;
        vol[c] *= 0.5;
    }
    for (int f = 0; f < nBFace; ++f) {
        int le = left[f], re = f + nCell;
        if (area[f] > SMALL) {
            double tmp = 2.0 * ((xcc[le] - xfc[f]) * xfn[f] + (ycc[le] - yfc[f]) * yfn[f] + (zcc[le] - zfc[f]) * zfn[f]);
            xcc[re] = xcc[le] - xfn[f] * tmp;
            ycc[re] = ycc[le] - yfn[f] * tmp;
            zcc[re] = zcc[le] - zfn[f] * tmp;
        } else {
            xcc[re] = -xcc[le] + 2.0 * xfc[f];
            ycc[re] = -ycc[le] + 2.0 * yfc[f];
            zcc[re] = -zcc[le] + 2.0 * zfc[f];
        }
        vol[re] = vol[le];
    }
}

/* UnstructGrid::ComputeFaceCenterAreaNormal3DNew2,
 * API/Geometry/src/Geo_UnstructGrid.cpp:5911-6157 (static grid). */
void orc_face_metrics(int nFace, const int *fp, const int *fn,
                      const double *x, const double *y, const double *z,
                      double *xfc, double *yfc, double *zfc,
                      double *xfn, double *yfn, double *zfn, double *area)
{
    for (int f = 0; f < nFace; ++f) {
        int nn = fp[f + 1] - fp[f];
        xfc[f] = 0.0; yfc[f] = 0.0; zfc[f] = 0.0;
        for (int j = 0; j < nn; ++j) {
            int n1 = fn[fp[f] + j];
            xfc[f] += x[n1]; yfc[f] += y[n1]; zfc[f] += z[n1];
        }
        double coef = 1.0 / nn;
        xfc[f] *= coef; yfc[f] *= coef; zfc[f] *= coef;
    }
    for (int cyc = 0; cyc <= 4; ++cyc) {
        int last = (cyc == 4);
        for (int f = 0; f < nFace; ++f) {
            int nn = fp[f + 1] - fp[f];
            double xt = 0.0, yt = 0.0, zt = 0.0, at = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
            for (int j = 0; j < nn; ++j) {
                int p1 = fn[fp[f] + j];
                int p2 = fn[fp[f] + (j + 1) % nn];
                double dx1 = x[p1] - xfc[f], dy1 = y[p1] - yfc[f], dz1 = z[p1] - zfc[f];
                double dx2 = x[p2] - xfc[f], dy2 = y[p2] - yfc[f], dz2 = z[p2] - zfc[f];
                double anx = dy1 * dz2 - dy2 * dz1;
                double any = dz1 * dx2 - dz2 * dx1;
                double anz = dx1 * dy2 - dx2 * dy1;
                double a = sqrt(anx * anx + any * any + anz * anz);
                double x00 = x[p1] + x[p2] + xfc[f];
                double y00 = y[p1] + y[p2] + yfc[f];
                double z00 = z[p1] + z[p2] + zfc[f];
                at += a;
                xt += a * x00; yt += a * y00; zt += a * z00;
                sx += anx; sy += any; sz += anz;
            }
            if (at > SMALL) {
                double a = 1.0 / at / 3.0;
                xfc[f] = xt * a; yfc[f] = yt * a; zfc[f] = zt * a;
            }
            if (last) {
                double a = sqrt(sx * sx + sy * sy + sz * sz);
                area[f] = a * 0.5;
                double oa = 1.0 / (a + SMALL);
                xfn[f] = sx * oa; yfn[f] = sy * oa; zfn[f] = sz * oa;
            }
        }
    }
}

/* UnstructGrid::ComputeCellCenterVol3DNew2,
 * API/Geometry/src/Geo_UnstructGrid.cpp:6545-6820.  Output arrays are sized
 * nTotal; the ghost centroid is the mirror of the owner across the face. */
void orc_cell_metrics(int nCell, int nBFace, int nFace, const int *left, const int *right,
                      const int *fp, const int *fn, const int *cp, const int *cn,
                      const double *x, const double *y, const double *z,
                      const double *xfc, const double *yfc, const double *zfc,
                      const double *xfn, const double *yfn, const double *zfn,
                      const double *area,
                      double *xcc, double *ycc, double *zcc, double *vol)
{
    for (int c = 0; c < nCell; ++c) {
        int nn = cp[c + 1] - cp[c];
        xcc[c] = 0.0; ycc[c] = 0.0; zcc[c] = 0.0;
        for (int j = 0; j < nn; ++j) {
            int n2 = cn[cp[c] + j];
            xcc[c] += x[n2]; ycc[c] += y[n2]; zcc[c] += z[n2];
        }
        xcc[c] /= nn; ycc[c] /= nn; zcc[c] /= nn;
    }
    double *xt = (double *)calloc((size_t)nCell, sizeof(double));
    double *yt = (double *)calloc((size_t)nCell, sizeof(double));
    double *zt = (double *)calloc((size_t)nCell, sizeof(double));
    for (int c = 0; c < nCell; ++c) vol[c] = 0.0;
    for (int f = 0; f < nFace; ++f) {
        int nn = fp[f + 1] - fp[f];
        int le = left[f], re = right[f];
        int interior = f >= nBFace;
        for (int j = 0; j < nn; ++j) {
            int p2 = fn[fp[f] + j];
            int p3 = fn[fp[f] + (j + 1) % nn];
            double x21 = x[p2] - xfc[f], y21 = y[p2] - yfc[f], z21 = z[p2] - zfc[f];
            double x31 = x[p3] - xfc[f], y31 = y[p3] - yfc[f], z31 = z[p3] - zfc[f];
            double nx = y21 * z31 - y31 * z21;
            double ny = z21 * x31 - z31 * x21;
            double nz = x21 * y31 - x31 * y21;
            double x41 = xcc[le] - xfc[f], y41 = ycc[le] - yfc[f], z41 = zcc[le] - zfc[f];
            double xc = xfc[f] + x[p2] + x[p3] + xcc[le];
            double yc = yfc[f] + y[p2] + y[p3] + ycc[le];
            double zc = zfc[f] + z[p2] + z[p3] + zcc[le];
            double tmp = nx * x41 + ny * y41 + nz * z41;
            tmp = -tmp;
            xc *= tmp; yc *= tmp; zc *= tmp;
            xt[le] += xc; yt[le] += yc; zt[le] += zc; vol[le] += tmp;
            if (interior) {
                x41 = xcc[re] - xfc[f]; y41 = ycc[re] - yfc[f]; z41 = zcc[re] - zfc[f];
                xc = xfc[f] + x[p2] + x[p3] + xcc[re];
                yc = yfc[f] + y[p2] + y[p3] + ycc[re];
                zc = zfc[f] + z[p2] + z[p3] + zcc[re];
                tmp = nx * x41 + ny * y41 + nz * z41;
                xc *= tmp; yc *= tmp; zc *= tmp;
                xt[re] += xc; yt[re] += yc; zt[re] += zc; vol[re] += tmp;
            }
        }
    }
    for (int c = 0; c < nCell; ++c) {
        double tmp = 1.0 / (4.0 * vol[c] + SMALL);
        xt[c] *= tmp; yt[c] *= tmp; zt[c] *= tmp;
        vol[c] /= 6.0;
    }
    int *mark = (int *)malloc((size_t)nCell * sizeof(int));
    for (int c = 0; c < nCell; ++c) mark[c] = 1;
    for (int f = 0; f < nFace; ++f) {
        int le = left[f];
        double dx = xt[le] - xfc[f], dy = yt[le] - yfc[f], dz = zt[le] - zfc[f];
        double tmp = dx * xfn[f] + dy * yfn[f] + dz * zfn[f];
        if (tmp > 0.0) mark[le] = 0;
        if (f >= nBFace) {
            int re = right[f];
            dx = xt[re] - xfc[f]; dy = yt[re] - yfc[f]; dz = zt[re] - zfc[f];
            tmp = dx * xfn[f] + dy * yfn[f] + dz * zfn[f];
            if (tmp < 0.0) mark[re] = 0;
        }
    }
    for (int c = 0; c < nCell; ++c) {
        if (mark[c]) { xcc[c] = xt[c]; ycc[c] = yt[c]; zcc[c] = zt[c]; }
    }
    for (int f = 0; f < nBFace; ++f) {
        int le = left[f];
        int re = nCell + f;
        if (area[f] > SMALL) {
            double tmp = 2.0 * ((xcc[le] - xfc[f]) * xfn[f] + (ycc[le] - yfc[f]) * yfn[f] +
                                (zcc[le] - zfc[f]) * zfn[f]);
            xcc[re] = xcc[le] - xfn[f] * tmp;
            ycc[re] = ycc[le] - yfn[f] * tmp;
            zcc[re] = zcc[le] - zfn[f] * tmp;
        } else {
            xcc[re] = -xcc[le] + 2.0 * xfc[f];
            ycc[re] = -ycc[le] + 2.0 * yfc[f];
            zcc[re] = -zcc[le] + 2.0 * zfc[f];
        }
        vol[re] = vol[le];
    }
    free(mark); free(xt); free(yt); free(zt);
}

/* UnstructGrid::SkewnessSummary, API/Geometry/src/Geo_UnstructGrid.cpp:2206-2300:
 * per-cell minimum over its faces of the complement angle (degrees). */
void orc_skewness(int nCell, int nBFace, int nFace, const int *left, const int *right,
                  const double *xfc, const double *yfc, const double *zfc,
                  const double *xfn, const double *yfn, const double *zfn, const double *area,
                  const double *xcc, const double *ycc, const double *zcc, double *skew)
{
    const double straightAngle = 180.0;
    for (int c = 0; c < nCell; ++c) skew[c] = straightAngle;
    for (int f = nBFace; f < nFace; ++f) {
        if (area[f] <= TINY) continue;
        int le = left[f], re = right[f];
        double x1 = xfc[f] - xcc[le], y1 = yfc[f] - ycc[le], z1 = zfc[f] - zcc[le];
        double d1 = sqrt(x1 * x1 + y1 * y1 + z1 * z1);
        double p1 = (xfn[f] * x1 + yfn[f] * y1 + zfn[f] * z1) / (d1 + TINY);
        if (p1 > 1.0) p1 = 1.0;
        if (p1 < -1.0) p1 = -1.0;
        double x2 = xcc[re] - xfc[f], y2 = ycc[re] - yfc[f], z2 = zcc[re] - zfc[f];
        double d2 = sqrt(x2 * x2 + y2 * y2 + z2 * z2);
        double p2 = (xfn[f] * x2 + yfn[f] * y2 + zfn[f] * z2) / (d2 + TINY);
        if (p2 > 1.0) p2 = 1.0;
        if (p2 < -1.0) p2 = -1.0;
        double angle = MIN(asin(p1), asin(p2)) * straightAngle / PI_;
        skew[le] = MIN(skew[le], angle);
        skew[re] = MIN(skew[re], angle);
    }
    for (int f = 0; f < nBFace; ++f) {
        if (area[f] <= TINY) continue;
        int le = left[f];
        double x1 = xfc[f] - xcc[le], y1 = yfc[f] - ycc[le], z1 = zfc[f] - zcc[le];
        double d1 = sqrt(x1 * x1 + y1 * y1 + z1 * z1);
        double p1 = (xfn[f] * x1 + yfn[f] * y1 + zfn[f] * z1) / (d1 + TINY);
        if (p1 > 1.0) p1 = 1.0;
        if (p1 < -1.0) p1 = -1.0;
        double angle = asin(p1) * straightAngle / PI_;
        skew[le] = MIN(skew[le], angle);
    }
}

/* ------------------------------------------------------------------ */
/* Solver state                                                        */
/* ------------------------------------------------------------------ */

static double *dup_d(const double *a, int n)
{
    double *b = (double *)malloc((size_t)n * sizeof(double));
    memcpy(b, a, (size_t)n * sizeof(double));
    return b;
}
static int *dup_i(const int *a, int n)
{
    int *b = (int *)malloc((size_t)n * sizeof(int));
    memcpy(b, a, (size_t)n * sizeof(int));
    return b;
}
static double *zeros(int n) { return (double *)calloc((size_t)n, sizeof(double)); }

/* Freestream primitive state: PerfectGas::ComputeReferencePrimitive
 * (API/Physics/src/PerfectGas.cpp:1276-1300) with the velocity direction of
 * GlobalBoundaryCondition::SetSpeedDrictionForReferenceVar
 * (API/Geometry/src/Geo_SimpleBC.cpp:1200-1215). */
void orc_freestream(const orc_solver *s, double *prim)
{
    double attack = s->p.alpha_deg * PI_ / 180.0;
    double sideslip = s->p.beta_deg * PI_ / 180.0;
    prim[IR] = 1.0;
    prim[IU] = 1.0 * (cos(attack) * cos(sideslip));
    prim[IV] = 1.0 * (sin(attack) * cos(sideslip));
    prim[IW] = 1.0 * sin(sideslip);
    prim[IP] = s->R * 1.0 * 1.0 / 1.0;
}

/* UnstructGrid::ComputeWeight, Geo_UnstructGrid.cpp:3761-3940, 3-D branch
 * with every face marked (fMark = 1): face weight 1/(|dx|^2 + SMALL) over
 * the centroid difference right - left, the 3x3 normal matrix of each cell
 * summed in face order, then its inverse (zero if the determinant is below
 * SMALL). */
static void lsq_weights(orc_solver *s)
{
    int nC = s->nCell, nF = s->nFace;
    s->iwt = zeros(nF);
    s->ixx = zeros(nC); s->iyy = zeros(nC); s->izz = zeros(nC);
    s->ixy = zeros(nC); s->ixz = zeros(nC); s->iyz = zeros(nC);
    for (int f = 0; f < nF; ++f) {
        int le = s->left[f], re = f < s->nBFace ? f + nC : s->right[f];
        double dx = (s->xcc[re] - s->xcc[le]) * 1.0;
        double dy = (s->ycc[re] - s->ycc[le]) * 1.0;
        double dz = (s->zcc[re] - s->zcc[le]) * 1.0;
        double ods = 1.0 / (dx * dx + dy * dy + dz * dz + SMALL);
        double txx = ods * dx * dx, tyy = ods * dy * dy, tzz = ods * dz * dz;
        double txy = ods * dx * dy, txz = ods * dx * dz, tyz = ods * dy * dz;
        s->iwt[f] = ods;
        s->ixx[le] += txx; s->iyy[le] += tyy; s->izz[le] += tzz;
        s->ixy[le] += txy; s->ixz[le] += txz; s->iyz[le] += tyz;
        if (f < s->nBFace) continue;
        s->ixx[re] += txx; s->iyy[re] += tyy; s->izz[re] += tzz;
        s->ixy[re] += txy; s->ixz[re] += txz; s->iyz[re] += tyz;
    }
    for (int c = 0; c < nC; ++c) {
        double txx = s->ixx[c], tyy = s->iyy[c], tzz = s->izz[c];
        double txy = s->ixy[c], txz = s->ixz[c], tyz = s->iyz[c];
        double delt = (txx * tyy * tzz + 2.0 * txy * txz * tyz - txx * tyz * tyz - tyy * txz * txz - tzz * txy * txy);
        if (ABS(delt) < SMALL) {
            s->ixx[c] = 0.0; s->iyy[c] = 0.0; s->izz[c] = 0.0;
            s->ixy[c] = 0.0; s->ixz[c] = 0.0; s->iyz[c] = 0.0;
        } else {
            double odelt = 1.0 / delt;
            s->ixx[c] = (tyy * tzz - tyz * tyz) * odelt;
            s->iyy[c] = (txx * tzz - txz * txz) * odelt;
            s->izz[c] = (txx * tyy - txy * txy) * odelt;
            s->ixy[c] = (txz * tyz - txy * tzz) * odelt;
            s->ixz[c] = (txy * tyz - txz * tyy) * odelt;
            s->iyz[c] = (txy * txz - tyz * txx) * odelt;
        }
    }
}

orc_solver *orc_create(const orc_params *p, int nCell, int nBFace, int nFace,
                       const int *left, const int *right, const int *bcType,
                       const double *xfc, const double *yfc, const double *zfc,
                       const double *xfn, const double *yfn, const double *zfn,
                       const double *area, const double *xcc, const double *ycc,
                       const double *zcc, const double *vol, const double *skew)
{
    orc_solver *s = (orc_solver *)calloc(1, sizeof(orc_solver));
    s->p = *p;
    s->nCell = nCell; s->nBFace = nBFace; s->nFace = nFace;
    int nT = s->nTotal = nCell + nBFace;
    s->left = dup_i(left, nFace); s->right = dup_i(right, nFace);
    s->bcType = dup_i(bcType, nBFace > 0 ? nBFace : 1);
    s->xfc = dup_d(xfc, nFace); s->yfc = dup_d(yfc, nFace); s->zfc = dup_d(zfc, nFace);
    s->xfn = dup_d(xfn, nFace); s->yfn = dup_d(yfn, nFace); s->zfn = dup_d(zfn, nFace);
    s->area = dup_d(area, nFace);
    s->xcc = dup_d(xcc, nT); s->ycc = dup_d(ycc, nT); s->zcc = dup_d(zcc, nT);
    s->vol = dup_d(vol, nT); s->skew = dup_d(skew, nCell);
    /* Region.cpp:3203: mean volume over real cells. */
    double tv = 0.0;
    for (int c = 0; c < nCell; ++c) tv += vol[c];
    s->avgVol = tv / nCell;
    s->dim = 3;
    /* cell -> face in increasing face order (as built from the face loop). */
    s->c2fPtr = (int *)calloc((size_t)nCell + 1, sizeof(int));
    for (int f = 0; f < nFace; ++f) {
        s->c2fPtr[left[f] + 1]++;
        if (f >= nBFace) s->c2fPtr[right[f] + 1]++;
    }
    for (int c = 0; c < nCell; ++c) s->c2fPtr[c + 1] += s->c2fPtr[c];
    s->c2f = (int *)malloc((size_t)s->c2fPtr[nCell] * sizeof(int));
    int *fill = (int *)calloc((size_t)nCell, sizeof(int));
    for (int f = 0; f < nFace; ++f) {
        int le = left[f];
        s->c2f[s->c2fPtr[le] + fill[le]++] = f;
        if (f >= nBFace) { int re = right[f]; s->c2f[s->c2fPtr[re] + fill[re]++] = f; }
    }
    free(fill);
    for (int m = 0; m < 5; ++m) {
        s->q[m] = zeros(nT); s->res[m] = zeros(nT); s->dq[m] = zeros(nT);
        s->dqdx[m] = zeros(nT); s->dqdy[m] = zeros(nT); s->dqdz[m] = zeros(nT);
        s->dflux[m] = zeros(nT);
    }
    s->lim = zeros(nT); s->specInv = zeros(nT); s->specVis = zeros(nT);
    s->dt = zeros(nT); s->diag = zeros(nT); s->gama = zeros(nT); s->t = zeros(nT);
    s->visl = zeros(nT); s->vist = zeros(nT);
    s->nut = zeros(nT); s->resT = zeros(nT); s->specT = zeros(nT);
    s->matL = zeros(nFace); s->matR = zeros(nFace); s->dqT = zeros(nT); s->dfluxT = zeros(nT);
    s->gnx = zeros(nT); s->gny = zeros(nT); s->gnz = zeros(nT); s->wdist = zeros(nT);
    s->dtdx = zeros(nT); s->dtdy = zeros(nT); s->dtdz = zeros(nT);
    s->rtem = zeros(nT);
    s->R = 1.0 / (p->gamma * p->mach * p->mach);
    s->pTotal = orc_ptotal(p->mach, p->gamma);
    s->entropyCoef = orc_entropy_coef(p->mach, p->entropy_scale);
    s->tsuth = 110.4 / p->t_ref;
    orc_freestream(s, s->primInf);
    for (int c = 0; c < nT; ++c) s->gama[c] = p->gamma;
    lsq_weights(s);
    return s;
}

#ifdef _OPENMP
/* Levels of a sweep: a cell's level is one more than the highest level of
 * its real neighbours swept before it (lower indices forward, higher
 * backward).  Cells of a level depend only on earlier levels. */
static void sweep_levels(const orc_solver *s, int forward, int **ptrOut, int **cellsOut, int *nOut)
{
    const int nC = s->nCell;
    int *lv = (int *)calloc((size_t)nC, sizeof(int));
    int nl = 0;
    for (int i = 0; i < nC; ++i) {
        const int c = forward ? i : nC - 1 - i;
        int l = 0;
        for (int k = s->c2fPtr[c]; k < s->c2fPtr[c + 1]; ++k) {
            const int f = s->c2f[k];
            if (f < s->nBFace) continue;
            const int nb = s->left[f] == c ? s->right[f] : s->left[f];
            if ((forward && nb < c) || (!forward && nb > c)) l = lv[nb] + 1 > l ? lv[nb] + 1 : l;
        }
        lv[c] = l;
        if (l + 1 > nl) nl = l + 1;
    }
    int *ptr = (int *)calloc((size_t)nl + 1, sizeof(int)), *cells = (int *)malloc(sizeof(int) * (size_t)nC);
    for (int c = 0; c < nC; ++c) ptr[lv[c] + 1]++;
    for (int l = 0; l < nl; ++l) ptr[l + 1] += ptr[l];
    int *fill = (int *)calloc((size_t)nl, sizeof(int));
    for (int i = 0; i < nC; ++i) {
        const int c = forward ? i : nC - 1 - i;
        cells[ptr[lv[c]] + fill[lv[c]]++] = c;
    }
    free(fill);
    free(lv);
    *ptrOut = ptr; *cellsOut = cells; *nOut = nl;
}

static struct orc_part *orc_partition(orc_solver *s)
{
    if (s->part) return s->part;
    struct orc_part *P = (struct orc_part *)calloc(1, sizeof(struct orc_part));
    omp_set_dynamic(0);
    int T = omp_get_max_threads();
    if (T < 1) T = 1;
    const int nC = s->nCell, nT = s->nTotal;
    P->nThreads = T;
    P->owner = (int *)malloc(sizeof(int) * (size_t)nT);
    for (int c = 0; c < nC; ++c) P->owner[c] = (int)((long long)c * T / nC);
    for (int f = 0; f < s->nBFace; ++f) P->owner[nC + f] = P->owner[s->left[f]];
    P->fptr = (int *)calloc((size_t)T + 1, sizeof(int));
    for (int f = 0; f < s->nFace; ++f) {
        const int a = P->owner[s->left[f]], b = P->owner[s->right[f]];
        P->fptr[a + 1]++;
        if (b != a) P->fptr[b + 1]++;
    }
    for (int t = 0; t < T; ++t) P->fptr[t + 1] += P->fptr[t];
    P->flist = (int *)malloc(sizeof(int) * (size_t)(P->fptr[T] > 0 ? P->fptr[T] : 1));
    int *fill = (int *)calloc((size_t)T, sizeof(int));
    for (int f = 0; f < s->nFace; ++f) {
        const int a = P->owner[s->left[f]], b = P->owner[s->right[f]];
        P->flist[P->fptr[a] + fill[a]++] = f;
        if (b != a) P->flist[P->fptr[b] + fill[b]++] = f;
    }
    free(fill);
    sweep_levels(s, 1, &P->lvPtrF, &P->lvCellsF, &P->nLvF);
    sweep_levels(s, 0, &P->lvPtrB, &P->lvCellsB, &P->nLvB);
    s->part = P;
    return P;
}
#endif

void orc_destroy(orc_solver *s)
{
    if (!s) return;
#ifdef _OPENMP
    if (s->part) {
        free(s->part->owner); free(s->part->fptr); free(s->part->flist);
        free(s->part->lvPtrF); free(s->part->lvCellsF); free(s->part->lvPtrB); free(s->part->lvCellsB);
        free(s->part);
    }
#endif
    free(s->left); free(s->right); free(s->bcType);
    free(s->xfc); free(s->yfc); free(s->zfc); free(s->xfn); free(s->yfn); free(s->zfn);
    free(s->area); free(s->xcc); free(s->ycc); free(s->zcc); free(s->vol); free(s->skew);
    free(s->c2fPtr); free(s->c2f);
    for (int m = 0; m < 5; ++m) {
        free(s->q[m]); free(s->res[m]); free(s->dq[m]); free(s->dqdx[m]);
        free(s->dqdy[m]); free(s->dqdz[m]); free(s->dflux[m]);
    }
    free(s->lim); free(s->specInv); free(s->specVis); free(s->dt); free(s->diag);
    free(s->gama); free(s->t); free(s->visl); free(s->vist);
    free(s->nut); free(s->resT); free(s->specT); free(s->matL); free(s->matR); free(s->dqT); free(s->dfluxT);
    free(s->gnx); free(s->gny); free(s->gnz); free(s->wdist);
    free(s->dtdx); free(s->dtdy); free(s->dtdz); free(s->rtem);
    free(s->iwt); free(s->ixx); free(s->iyy); free(s->izz); free(s->ixy); free(s->ixz); free(s->iyz);
    free(s->xn); free(s->yn); free(s->zn); free(s->fnPtr); free(s->fn); free(s->cnPtr); free(s->cn);
    for (int m = 0; m < 5; ++m) free(s->qNode[m]);
    free(s->tNode); free(s->nodeWeight); free(s->nodeBC);
    free(s);
}

static double **field5(orc_solver *s, int field)
{
    switch (field) {
    case ORC_Q: return s->q;
    case ORC_RES: return s->res;
    case ORC_DQDX: return s->dqdx;
    case ORC_DQDY: return s->dqdy;
    case ORC_DQDZ: return s->dqdz;
    case ORC_DQ: return s->dq;
    case ORC_DELTA_FLUX: return s->dflux;
    default: return NULL;
    }
}
static double *field1(orc_solver *s, int field)
{
    switch (field) {
    case ORC_LIMIT: return s->lim;
    case ORC_SPEC_INV: return s->specInv;
    case ORC_SPEC_VIS: return s->specVis;
    case ORC_DT: return s->dt;
    case ORC_DIAG: return s->diag;
    case ORC_T: return s->t;
    case ORC_VISL: return s->visl;
    case ORC_DTDX: return s->dtdx;
    case ORC_DTDY: return s->dtdy;
    case ORC_DTDZ: return s->dtdz;
    case ORC_VIST: return s->vist;
    case ORC_NUT: return s->nut;
    case ORC_RES_T: return s->resT;
    case ORC_SPEC_T: return s->specT;
    case ORC_DQ_T: return s->dqT;
    case ORC_WDIST: return s->wdist;
    case ORC_DNUDX: return s->gnx;
    case ORC_DNUDY: return s->gny;
    case ORC_DNUDZ: return s->gnz;
    case ORC_PRESSURE_FACTOR: return s->rtem;
    default: return NULL;
    }
}

/* Fields are exchanged SoA over nTotal: five-component fields as [5][nTotal]. */
int orc_get(orc_solver *s, int field, double *out)
{
    double **f5 = field5(s, field);
    if (f5) {
        for (int m = 0; m < 5; ++m) memcpy(out + (size_t)m * s->nTotal, f5[m], (size_t)s->nTotal * sizeof(double));
        return 5;
    }
    double *f1 = field1(s, field);
    if (f1) { memcpy(out, f1, (size_t)s->nTotal * sizeof(double)); return 1; }
    return 0;
}

int orc_set(orc_solver *s, int field, const double *in)
{
    double **f5 = field5(s, field);
    if (field == ORC_Q) s->tFresh = 1;  /* a new state: its own temperature for the node values */
    if (f5) {
        for (int m = 0; m < 5; ++m) memcpy(f5[m], in + (size_t)m * s->nTotal, (size_t)s->nTotal * sizeof(double));
        return 5;
    }
    double *f1 = field1(s, field);
    if (f1) { memcpy(f1, in, (size_t)s->nTotal * sizeof(double)); return 1; }
    return 0;
}

/* Controller::InitFlow: uniform freestream on every cell and ghost. */
void orc_init_freestream(orc_solver *s)
{
    s->tFresh = 1;
    for (int m = 0; m < 5; ++m)
        for (int c = 0; c < s->nTotal; ++c) s->q[m][c] = s->primInf[m];
    /* TurbSolverUnstr::InitFlowAsRestart (:910-934): nu~ = freestream value;
     * the eddy viscosity starts at zero (NSSolverUnstruct.cpp:688-692) */
    if (s->p.viscous == 3)
        for (int c = 0; c < s->nTotal; ++c) { s->nut[c] = s->p.sa_nu_inf; s->vist[c] = 0.0; }
}

/* ------------------------------------------------------------------ */
/* Boundary conditions: NSSolverUnstruct::Boundary                     */
/* CFD/NSSolver/src/NSSolverUnstruct.cpp:13404-13530                  */
/* ------------------------------------------------------------------ */

static int is_wall(int bc) { return bc == ORC_BC_SOLID_SURFACE; } /* Geo_SimpleBC.h:389 */

/* SymmetryBCRegion, NSSolverUnstruct.cpp:19413-19449 (also the inviscid wall). */
static void bc_symmetry(orc_solver *s, int f)
{
    int le = s->left[f], re = s->right[f];
    for (int m = 0; m < 5; ++m) s->q[m][re] = s->q[m][le];
    double uL = s->q[IU][le], vL = s->q[IV][le], wL = s->q[IW][le];
    double vn = s->xfn[f] * uL + s->yfn[f] * vL + s->zfn[f] * wL - 0.0;
    s->q[IU][re] = uL - 2.0 * s->xfn[f] * vn;
    s->q[IV][re] = vL - 2.0 * s->yfn[f] * vn;
    s->q[IW][re] = wL - 2.0 * s->zfn[f] * vn;
}

/* ViscousAdiabaticWallBCRegion, NSSolverUnstruct.cpp:20988-21104 (static wall). */
static void bc_viscous_wall(orc_solver *s, int f)
{
    int le = s->left[f], re = s->right[f];
    for (int m = 0; m < 5; ++m) s->q[m][re] = s->q[m][le];
    s->q[IU][re] = -s->q[IU][le] + 2.0 * 0.0;
    s->q[IV][re] = -s->q[IV][le] + 2.0 * 0.0;
    s->q[IW][re] = -s->q[IW][le] + 2.0 * 0.0;
    if (s->p.t_wall > 0.0) {
        /* ViscousIsotropicWallBCRegion (:20791-20915, perfect gas): wall
         * pressure kept, temperature mirrored about wallTemperature / T_ref
         * (floored at a tenth of it), density from the equation of state */
        double tw = s->p.t_wall / s->p.t_ref;
        double tL = (1.0 / s->R) * s->q[IP][le] / s->q[IR][le];
        double pG = s->q[IP][le];
        double tG = 2.0 * tw - tL;
        if (tG <= tw * 0.1) tG = tw * 0.1;
        s->q[IR][re] = pG / (s->R * tG * 1.0);
        s->q[IP][re] = pG;
    }
}

/* FarfieldBCRegion, NSSolverUnstruct.cpp:21195-21330 (no preconditioning). */
static void bc_farfield(orc_solver *s, int f)
{
    int le = s->left[f], re = s->right[f];
    const double *pf = s->primInf;
    double gl = s->gama[le];
    double refGama = s->p.gamma;
    double rO = pf[IR], uO = pf[IU], vO = pf[IV], wO = pf[IW], pO = pf[IP];
    double cO = sqrt(ABS(refGama * pO / rO));
    for (int m = 0; m < 5; ++m) s->q[m][re] = s->q[m][le];
    double rI = s->q[IR][le], uI = s->q[IU][le], vI = s->q[IV][le], wI = s->q[IW][le], pI = s->q[IP][le];
    double nx = s->xfn[f], ny = s->yfn[f], nz = s->zfn[f];
    double vnO = nx * uO + ny * vO + nz * wO - 0.0;
    double vnI = nx * uI + ny * vI + nz * wI - 0.0;
    double cI = sqrt(ABS(gl * pI / rI));
    double gamm1 = gl - 1.0;
    double velI = sqrt(uI * uI + vI * vI + wI * wI);
    if (velI > cI) {
        if (vnI >= 0.0) {
            for (int m = 0; m < 5; ++m) s->q[m][re] = s->q[m][le];
        } else {
            for (int m = 0; m < 5; ++m) s->q[m][re] = pf[m];
        }
    } else {
        double Rp = vnI + 2.0 * cI / gamm1;
        double Rm = vnO - 2.0 * cO / gamm1;
        double vnB = 0.5 * (Rp + Rm);
        double cB = 0.25 * (Rp - Rm) * gamm1;
        double S, tx, ty, tz;
        if (vnB >= 0.0) {
            S = pI / pow(rI, gl);
            tx = uI - nx * vnI; ty = vI - ny * vnI; tz = wI - nz * vnI;
        } else {
            S = pO / pow(rO, gl);
            tx = uO - nx * vnO; ty = vO - ny * vnO; tz = wO - nz * vnO;
        }
        double rB = pow((cB * cB / (S * gl)), 1.0 / gamm1);
        s->q[IR][re] = rB;
        s->q[IU][re] = tx + nx * vnB;
        s->q[IV][re] = ty + ny * vnB;
        s->q[IW][re] = tz + nz * vnB;
        s->q[IP][re] = cB * cB * rB / gl;
    }
}

/* OutflowBCRegion, NSSolverUnstruct.cpp:19326-19366. */
static void bc_outflow(orc_solver *s, int f)
{
    int le = s->left[f], re = s->right[f];
    for (int m = 0; m < 5; ++m) s->q[m][re] = s->q[m][le];
}

void orc_boundary(orc_solver *s)
{
    CELL_LOOP
    for (int f = 0; f < s->nBFace; ++f) {
        int bc = s->bcType[f];
        if (bc == ORC_BC_EXTRAPOLATION || bc == ORC_BC_OUTFLOW) bc_outflow(s, f);
        else if (is_wall(bc)) {
            if (!s->p.viscous) bc_symmetry(s, f);
            else bc_viscous_wall(s, f); /* adiabatic, or isothermal with t_wall > 0 */
        } else if (bc == ORC_BC_SYMMETRY) bc_symmetry(s, f);
        else if (bc == ORC_BC_FARFIELD) bc_farfield(s, f);
    }
}

/* ComputeGamaAndTemperature, NSSolverUnstruct.cpp:4666-4732 (perfect gas). */
void orc_gamma_t(orc_solver *s)
{
    CELL_LOOP
    for (int c = 0; c < s->nTotal; ++c) {
        s->gama[c] = s->p.gamma;
        s->t[c] = s->q[IP][c] / (s->R * s->q[IR][c]);
    }
}

/* Controller::PressureFactor (CFD/Solver/src/Controller.cpp:1202-1234) with
 * NSSolverUnstruct::CalPressureFactor (NSSolverUnstruct.cpp:6976-7045): the
 * cell sensor rtem of the pressure-sensor entropy fixes, nothing for the
 * other methods.  Method 2: over the cell's faces (in face order), the sums
 * of the pressures on the faces' left and right sides (ghost states on
 * boundary faces), |sumR - sumL| / |sumR + sumL| / nFaces.  Method 6: from
 * the largest and smallest pressure of the cell and its interior-face
 * neighbours (UnstructGrid::ComputeCell2Cell, Geo_UnstructGrid.cpp:
 * 1104-1156), ratio = max/min, f = clip((2.6 - ratio) / (2.6 - 1.3), 0, 1),
 * rtem = (1 - cos(f pi)) / 2 with pi = 3.141592653589.  Ghosts take their
 * owner's value (SetGhostCellExceptInterface, Geo_UnstructGrid.cpp:
 * 1857-1881).  Called after the boundary conditions and gamma/T. */
void orc_pressure_factor(orc_solver *s)
{
    const int method = s->p.entropy_fix;
    if (method != 2 && method != 6) return;
    const double *p = s->q[IP];
    CELL_LOOP
    for (int c = 0; c < s->nCell; ++c) {
        const int b = s->c2fPtr[c], e = s->c2fPtr[c + 1];
        if (method == 2) {
            double sl = 0.0, sr = 0.0;
            for (int k = b; k < e; ++k) {
                const int f = s->c2f[k];
                sl += p[s->left[f]];
                sr += p[s->right[f]];
            }
            double r = 0.0;
            r += ABS(sr - sl) / ABS(sr + sl);
            r /= (double)(e - b);
            s->rtem[c] = r;
        } else {
            double lo = p[c], hi = p[c];
            for (int k = b; k < e; ++k) {
                const int f = s->c2f[k];
                if (f < s->nBFace) continue;
                const int nb = s->left[f] == c ? s->right[f] : s->left[f];
                if (p[nb] < lo) lo = p[nb];
                if (p[nb] > hi) hi = p[nb];
            }
            const double ratio = hi / lo;
            const double fn = MIN(1.0, MAX(0.0, (2.6 - ratio) / (2.6 - 1.3)));
            double cf = cos(fn * 3.141592653589);
            cf = MAX(-1.0, MIN(cf, 1.0));
            s->rtem[c] = 0.5 * (1.0 - cf);
        }
    }
    for (int f = 0; f < s->nBFace; ++f) s->rtem[s->right[f]] = s->rtem[s->left[f]];
}

/* ComputeViscousCoefficientWithoutChemical, NSSolverUnstruct.cpp:6722-6746. */
void orc_viscosity(orc_solver *s)
{
    double ts = s->tsuth;
    CELL_LOOP
    for (int c = 0; c < s->nTotal; ++c) {
        double tt = s->t[c];
        s->visl[c] = tt * sqrt(tt) * (1.0 + ts) / (tt + ts);
    }
}

/* ------------------------------------------------------------------ */
/* Gradients: UnstructGrid::CompGradientGGCell                         */
/* API/Geometry/src/Geo_UnstructGrid.cpp:3220-3284 (ghost gradient 0)  */
/* ------------------------------------------------------------------ */
static void grad_ggcell(orc_solver *s, const double *q, double *gx, double *gy, double *gz)
{
    int nT = s->nTotal;
    CELL_LOOP
    for (int c = 0; c < nT; ++c) { gx[c] = 0.0; gy[c] = 0.0; gz[c] = 0.0; }
    FACE_LOOP(s, 0, s->nBFace, f)
        int le = s->left[f], re = f + s->nCell;
        double qfc = 0.5 * (q[le] + q[re]);
        double nx = s->xfn[f] * s->area[f], ny = s->yfn[f] * s->area[f], nz = s->zfn[f] * s->area[f];
        if (OWN(le)) { gx[le] += qfc * nx; gy[le] += qfc * ny; gz[le] += qfc * nz; }
    FACE_LOOP_END
    FACE_LOOP(s, s->nBFace, s->nFace, f)
        int le = s->left[f], re = s->right[f];
        double qfc = 0.5 * (q[le] + q[re]);
        double nx = s->xfn[f] * s->area[f], ny = s->yfn[f] * s->area[f], nz = s->zfn[f] * s->area[f];
        if (OWN(le)) { gx[le] += qfc * nx; gy[le] += qfc * ny; gz[le] += qfc * nz; }
        if (OWN(re)) { gx[re] -= qfc * nx; gy[re] -= qfc * ny; gz[re] -= qfc * nz; }
    FACE_LOOP_END
    CELL_LOOP
    for (int c = 0; c < s->nCell; ++c) {
        double ovol = 1.0 / s->vol[c];
        gx[c] *= ovol; gy[c] *= ovol; gz[c] *= ovol;
    }
}

/* UnstructGrid::CompGradientLSQ, Geo_UnstructGrid.cpp:3995-4085: the
 * weighted differences right - left summed into both cells in face order,
 * times the cell's inverse normal matrix; a ghost takes its owner's gradient. */
static void grad_lsq(orc_solver *s, const double *q, double *gx, double *gy, double *gz)
{
    int nT = s->nTotal, nC = s->nCell;
    CELL_LOOP
    for (int c = 0; c < nT; ++c) { gx[c] = 0.0; gy[c] = 0.0; gz[c] = 0.0; }
    FACE_LOOP(s, 0, s->nFace, f)
        int le = s->left[f], re = f < s->nBFace ? f + nC : s->right[f];
        double dx = s->xcc[re] - s->xcc[le], dy = s->ycc[re] - s->ycc[le], dz = s->zcc[re] - s->zcc[le];
        double dq = s->iwt[f] * (q[re] - q[le]) * 1.0;
        double tmpx = dq * dx, tmpy = dq * dy, tmpz = dq * dz;
        if (OWN(le)) { gx[le] += tmpx; gy[le] += tmpy; gz[le] += tmpz; }
        if (f < s->nBFace) continue;
        if (OWN(re)) { gx[re] += tmpx; gy[re] += tmpy; gz[re] += tmpz; }
    FACE_LOOP_END
    CELL_LOOP
    for (int c = 0; c < nC; ++c) {
        double tx = gx[c], ty = gy[c], tz = gz[c];
        gx[c] = s->ixx[c] * tx + s->ixy[c] * ty + s->ixz[c] * tz;
        gy[c] = s->ixy[c] * tx + s->iyy[c] * ty + s->iyz[c] * tz;
        gz[c] = s->ixz[c] * tx + s->iyz[c] * ty + s->izz[c] * tz;
    }
    for (int f = 0; f < s->nBFace; ++f) {
        int le = s->left[f], re = f + nC;
        gx[re] = gx[le]; gy[re] = gy[le]; gz[re] = gz[le];
    }
}

/* ------------------------------------------------------------------ */
/* Node-based Green-Gauss gradient ("ggnode", the reference default)    */
/* ------------------------------------------------------------------ */

/* The node lists the node-based gradient needs (orc_create has only the
 * face and cell metrics). */
void orc_set_nodes(orc_solver *s, int nNode, const double *x, const double *y, const double *z,
                   const int *faceNodePtr, const int *faceNodes, const int *cellNodePtr, const int *cellNodes)
{
    s->nNode = nNode;
    s->xn = dup_d(x, nNode); s->yn = dup_d(y, nNode); s->zn = dup_d(z, nNode);
    s->fnPtr = dup_i(faceNodePtr, s->nFace + 1);
    s->fn = dup_i(faceNodes, faceNodePtr[s->nFace]);
    s->cnPtr = dup_i(cellNodePtr, s->nCell + 1);
    s->cn = dup_i(cellNodes, cellNodePtr[s->nCell]);
    for (int m = 0; m < 5; ++m) s->qNode[m] = zeros(nNode);
    s->tNode = zeros(nNode);
    s->nodeWeight = zeros(nNode);
    s->nodeBC = (int *)calloc((size_t)nNode, sizeof(int));
    s->tFresh = 1;
}

/* Boundary regions: the reference walks its BC regions in order and each
 * region's faces in order; a mesh here has one region per BC code, taken in
 * increasing code order, each with its faces in increasing index. */
#define FOR_BFACES_OF(s, code, f) \
    for (int f = 0; f < (s)->nBFace; ++f) if ((s)->bcType[f] == (code))

static void node_add(orc_solver *s, int point, double px, double py, double pz, const double *qv, double tv)
{
    double dx = s->xn[point] - px, dy = s->yn[point] - py, dz = s->zn[point] - pz;
    double dist = sqrt(dx * dx + dy * dy + dz * dz);
    double w = 1.0 / dist;
    for (int m = 0; m < 5; ++m) s->qNode[m][point] += qv[m] * w;
    s->tNode[point] += tv * w;
    s->nodeWeight[point] += w;
}

/* NSSolverUnstruct::ComputeNodeValue (NSSolverUnstruct.cpp:4734-5081) and
 * ModifyNodeValue (:5083-5117): inverse-distance node averages of q and t.
 * A node on a solid surface takes only the solid faces around it (their
 * face averages of the cell and ghost), a farfield node the farfield faces,
 * a node of another non-symmetry boundary those faces; every other node
 * (symmetry-plane nodes included) takes the real cells around it. */
void orc_node_values(orc_solver *s)
{
    const int nN = s->nNode, nC = s->nCell;
    if (s->tFresh) {  /* InitDependentVariables (NSSolver.cpp:761-794): t of the state itself */
        orc_gamma_t(s);
        s->tFresh = 0;
    }
    for (int i = 0; i < nN; ++i) {
        s->nodeWeight[i] = 0.0;
        for (int m = 0; m < 5; ++m) s->qNode[m][i] = 0;
        s->tNode[i] = 0;
        s->nodeBC[i] = 0;  /* NO_BOUNDARY_CONDITION */
    }
    /* node BC priority: solid surface, then farfield, then the other codes */
    FOR_BFACES_OF(s, ORC_BC_SOLID_SURFACE, f)
        for (int j = s->fnPtr[f]; j < s->fnPtr[f + 1]; ++j) s->nodeBC[s->fn[j]] = ORC_BC_SOLID_SURFACE;
    FOR_BFACES_OF(s, ORC_BC_FARFIELD, f)
        for (int j = s->fnPtr[f]; j < s->fnPtr[f + 1]; ++j)
            if (s->nodeBC[s->fn[j]] != ORC_BC_SOLID_SURFACE) s->nodeBC[s->fn[j]] = ORC_BC_FARFIELD;
    for (int code = 1; code <= 6; ++code) {
        if (code == ORC_BC_SOLID_SURFACE || code == ORC_BC_FARFIELD || code == ORC_BC_SYMMETRY) continue;
        FOR_BFACES_OF(s, code, f)
            for (int j = s->fnPtr[f]; j < s->fnPtr[f + 1]; ++j) {
                int point = s->fn[j];
                if (s->nodeBC[point] == ORC_BC_SOLID_SURFACE || s->nodeBC[point] == ORC_BC_FARFIELD) continue;
                s->nodeBC[point] = code;
            }
    }
    /* Step 1-3: boundary faces in region order, each face's average of its
     * cell and ghost */
    for (int step = 0; step < 3; ++step)
        for (int code = 1; code <= 6; ++code) {
            if (code == ORC_BC_SYMMETRY) continue;
            if (step == 0 && code != ORC_BC_SOLID_SURFACE) continue;
            if (step == 1 && code != ORC_BC_FARFIELD) continue;
            if (step == 2 && (code == ORC_BC_SOLID_SURFACE || code == ORC_BC_FARFIELD)) continue;
            FOR_BFACES_OF(s, code, f) {
                int le = s->left[f], re = s->right[f];
                double qf[5], tf;
                for (int m = 0; m < 5; ++m) qf[m] = 0.5 * (s->q[m][le] + s->q[m][re]);
                tf = 0.5 * (s->t[le] + s->t[re]);
                for (int j = s->fnPtr[f]; j < s->fnPtr[f + 1]; ++j) {
                    int point = s->fn[j];
                    if (step == 1 && s->nodeBC[point] == ORC_BC_SOLID_SURFACE) continue;
                    if (step == 2 && (s->nodeBC[point] == ORC_BC_SOLID_SURFACE ||
                                      s->nodeBC[point] == ORC_BC_FARFIELD)) continue;
                    node_add(s, point, s->xfc[f], s->yfc[f], s->zfc[f], qf, tf);
                }
            }
        }
    /* Step 4: interior (and symmetry) nodes from the real cells around them */
    for (int c = 0; c < nC; ++c) {
        double qc[5];
        for (int m = 0; m < 5; ++m) qc[m] = s->q[m][c];
        for (int j = s->cnPtr[c]; j < s->cnPtr[c + 1]; ++j) {
            int point = s->cn[j];
            if (s->nodeBC[point] != 0) continue;
            node_add(s, point, s->xcc[c], s->ycc[c], s->zcc[c], qc, s->t[c]);
        }
    }
    /* ModifyNodeValue */
    for (int i = 0; i < nN; ++i) {
        if (s->nodeWeight[i] > 1.0e-40) {  /* SMALL, Precision.h:27 */
            for (int m = 0; m < 5; ++m) s->qNode[m][i] /= s->nodeWeight[i];
            s->tNode[i] /= s->nodeWeight[i];
        } else {
            for (int m = 0; m < 5; ++m) s->qNode[m][i] = 0.0;
            s->tNode[i] = 0.0;
        }
    }
}

/* Node values [6][nNode]: the five primitives, then T. */
void orc_get_node_values(const orc_solver *s, double *out)
{
    for (int m = 0; m < 5; ++m)
        for (int i = 0; i < s->nNode; ++i) out[(size_t)m * s->nNode + i] = s->qNode[m][i];
    for (int i = 0; i < s->nNode; ++i) out[(size_t)5 * s->nNode + i] = s->tNode[i];
}

static double face_node_mean(const orc_solver *s, int f, const double *qnode)
{
    double qf = 0.0;
    for (int j = s->fnPtr[f]; j < s->fnPtr[f + 1]; ++j) qf += qnode[s->fn[j]];
    qf /= (s->fnPtr[f + 1] - s->fnPtr[f]);
    return qf;
}

/* UnstructGrid::CompGradientGGNode_NEW, Geo_UnstructGrid.cpp:4202-4315:
 * face values are node means (interior and symmetry faces) or the average
 * of the cell and its ghost (other boundary faces); the sums are divided by
 * the volume, and every ghost takes its owner's gradient. */
static void grad_ggnode(orc_solver *s, const double *q, const double *qnode, double *gx, double *gy, double *gz)
{
    int nT = s->nTotal;
    CELL_LOOP
    for (int c = 0; c < nT; ++c) { gx[c] = 0.0; gy[c] = 0.0; gz[c] = 0.0; }
    FACE_LOOP(s, 0, s->nBFace, f)
        int le = s->left[f], re = f + s->nCell;
        double qf = s->bcType[f] == ORC_BC_SYMMETRY ? face_node_mean(s, f, qnode) : 0.5 * (q[le] + q[re]);
        double nx = s->xfn[f] * s->area[f], ny = s->yfn[f] * s->area[f], nz = s->zfn[f] * s->area[f];
        if (OWN(le)) { gx[le] += qf * nx; gy[le] += qf * ny; gz[le] += qf * nz; }
    FACE_LOOP_END
    FACE_LOOP(s, s->nBFace, s->nFace, f)
        int le = s->left[f], re = s->right[f];
        double qf = face_node_mean(s, f, qnode);
        double nx = s->xfn[f] * s->area[f], ny = s->yfn[f] * s->area[f], nz = s->zfn[f] * s->area[f];
        if (OWN(le)) { gx[le] += qf * nx; gy[le] += qf * ny; gz[le] += qf * nz; }
        if (OWN(re)) { gx[re] -= qf * nx; gy[re] -= qf * ny; gz[re] -= qf * nz; }
    FACE_LOOP_END
    CELL_LOOP
    for (int c = 0; c < s->nCell; ++c) { gx[c] /= s->vol[c]; gy[c] /= s->vol[c]; gz[c] /= s->vol[c]; }
    for (int f = 0; f < s->nBFace; ++f) {
        int le = s->left[f], re = f + s->nCell;
        gx[re] = gx[le]; gy[re] = gy[le]; gz[re] = gz[le];
    }
}

/* NSSolverUnstruct::GetGradientField, NSSolverUnstruct.cpp:18244-18338:
 * gradients of the five primitives and of the temperature (gradientName
 * "ggcell" or "lsq", :18266-18272, :18308-18314). */
void orc_gradient(orc_solver *s)
{
    if (s->p.gradient == 1) {
        for (int m = 0; m < 5; ++m) grad_lsq(s, s->q[m], s->dqdx[m], s->dqdy[m], s->dqdz[m]);
        grad_lsq(s, s->t, s->dtdx, s->dtdy, s->dtdz);
        return;
    }
    if (s->p.gradient == 2) {
        for (int m = 0; m < 5; ++m) grad_ggnode(s, s->q[m], s->qNode[m], s->dqdx[m], s->dqdy[m], s->dqdz[m]);
        grad_ggnode(s, s->t, s->tNode, s->dtdx, s->dtdy, s->dtdz);
        return;
    }
    for (int m = 0; m < 5; ++m) grad_ggcell(s, s->q[m], s->dqdx[m], s->dqdy[m], s->dqdz[m]);
    grad_ggcell(s, s->t, s->dtdx, s->dtdy, s->dtdz);
}

/* ------------------------------------------------------------------ */
/* Venkatakrishnan limiter (ivencat = 7, scalar, limited by rho and p)  */
/* Limiter::Calculation, API/Physics/src/Limiter.cpp:47-145            */
/* ------------------------------------------------------------------ */

/* VenFun, API/Math/include/Math_Limiter.h:174-180 */
double orc_venfun(double x, double y, double eps)
{
    double x2 = x * x;
    double xy = x * y;
    double y2 = y * y;
    return ((x2 + eps + 2.0 * xy) / (x2 + 2.0 * y2 + xy + eps));
}

/* Limiter::MinMaxDiff, Limiter.cpp:1202-1254 (no interface faces here). */
static void min_max_diff(orc_solver *s, double *dmin, double *dmax, const double *q)
{
    CELL_LOOP
    for (int c = 0; c < s->nCell; ++c) { dmin[c] = q[c]; dmax[c] = q[c]; }
    FACE_LOOP(s, s->nBFace, s->nFace, f)
        int le = s->left[f], re = s->right[f];
        if (OWN(le)) { dmin[le] = MIN(dmin[le], q[re]); dmax[le] = MAX(dmax[le], q[re]); }
        if (OWN(re)) { dmin[re] = MIN(dmin[re], q[le]); dmax[re] = MAX(dmax[re], q[le]); }
    FACE_LOOP_END
    CELL_LOOP
    for (int c = 0; c < s->nCell; ++c) { dmin[c] -= q[c]; dmax[c] -= q[c]; }
}

/* Limiter::VencatLimiter7, Limiter.cpp:1091-1200. */
static void vencat7(orc_solver *s, double *limit, const double *q, const double *gx,
                    const double *gy, const double *gz)
{
    int nC = s->nCell;
    double *dmin = (double *)malloc((size_t)s->nTotal * sizeof(double));
    double *dmax = (double *)malloc((size_t)s->nTotal * sizeof(double));
    min_max_diff(s, dmin, dmax, q);
    double n = 3.0;
    double *eps = (double *)malloc((size_t)nC * sizeof(double));
    double eps_tmp = s->p.venkat_k * s->p.venkat_k * s->p.venkat_k;
    eps_tmp *= 1.0 / pow(pow(s->avgVol, 1.0 / s->dim), n);
    CELL_LOOP
    for (int c = 0; c < nC; ++c) {
        eps[c] = eps_tmp * q[c] * q[c];
        eps[c] *= pow(pow(s->vol[c], 1.0 / s->dim), n);
    }
    FACE_LOOP(s, 0, s->nFace, f)
        int le = s->left[f], re = s->right[f];
        if (OWN(le)) {
            double dx = s->xfc[f] - s->xcc[le], dy = s->yfc[f] - s->ycc[le], dz = s->zfc[f] - s->zcc[le];
            double dqf = gx[le] * dx + gy[le] * dy + gz[le] * dz;
            if (dqf > SMALL) limit[le] = MIN(limit[le], orc_venfun(dmax[le], dqf, eps[le]));
            else if (dqf < -SMALL) limit[le] = MIN(limit[le], orc_venfun(dmin[le], dqf, eps[le]));
        }
        if (f < s->nBFace || !OWN(re)) continue;
        double dx = s->xfc[f] - s->xcc[re], dy = s->yfc[f] - s->ycc[re], dz = s->zfc[f] - s->zcc[re];
        double dqf = gx[re] * dx + gy[re] * dy + gz[re] * dz;
        if (dqf > SMALL) limit[re] = MIN(limit[re], orc_venfun(dmax[re], dqf, eps[re]));
        else if (dqf < -SMALL) limit[re] = MIN(limit[re], orc_venfun(dmin[re], dqf, eps[re]));
    FACE_LOOP_END
    free(dmin); free(dmax); free(eps);
}

/* SetLimitBoundary, Geo_UnstructGrid.cpp:10744-10777: zero on both sides of
 * every boundary face that is not interface, symmetry, wall or outflow. */
static void set_limit_boundary(orc_solver *s, double *limit)
{
    for (int f = 0; f < s->nBFace; ++f) {
        int bc = s->bcType[f];
        if (bc != ORC_BC_SYMMETRY && !is_wall(bc) && bc != ORC_BC_OUTFLOW) {
            limit[s->left[f]] = 0.0;
            limit[s->right[f]] = 0.0;
        }
    }
}

/* NSSolverUnstruct::ComputeLimiter, NSSolverUnstruct.cpp:18189-18242. */
void orc_limiter(orc_solver *s)
{
    int nT = s->nTotal;
    if (s->p.limiter == 1) { /* ILMT_NOLIM */
        for (int c = 0; c < nT; ++c) s->lim[c] = 1.0;
        return;
    }
    for (int c = 0; c < nT; ++c) s->lim[c] = 0.0;
    /* first order (uns_limiter_name "1st", or isInIniting): the limiter
     * array stays zero and is unused (ComputeLimiter :18195-18225) */
    if (s->p.limiter == 2 || s->initPhase) return;
    for (int c = 0; c < nT; ++c) s->lim[c] = 1.0; /* InitLimitData */
    vencat7(s, s->lim, s->q[IR], s->dqdx[IR], s->dqdy[IR], s->dqdz[IR]);
    vencat7(s, s->lim, s->q[IP], s->dqdx[IP], s->dqdy[IP], s->dqdz[IP]);
    set_limit_boundary(s, s->lim);
}

/* ------------------------------------------------------------------ */
/* Roe flux, conservative form                                         */
/* ------------------------------------------------------------------ */

/* Harten, API/Physics/include/Flux_RoeEntropyCorrection.h:66-74 */
double orc_harten(double eig, double limit)
{
    double e = ABS(eig);
    if (e < limit) e = 0.5 * (eig * eig + limit * limit) / limit;
    return e;
}

/* Flux_RoeEntropyCorrection, API/Physics/src/Flux_RoeEntropyCorrection.cpp:93-166
 * (the build without USE_GMRESSOLVER).  pcL, pcR: the pressure factors of
 * the face's two cells (methods 2 and 6); absVel: |V| of the Roe average. */
static void entropy_fix(double *e1, double *e2, double *e3, double vn, double cm, double absVel,
                        double pcL, double pcR, int method, double acoustic, double convect)
{
    if (method == 1) {
        double mx = MAX(*e2, *e3);
        double t0 = mx * acoustic, t1 = mx * acoustic;
        *e1 = MAX(t0, *e1); *e2 = MAX(t1, *e2); *e3 = MAX(t1, *e3);
    } else if (method == 2) {
        /* Lin, JCP 1995 */
        const double k1 = 0.25, k2 = 5.0, k3 = 15.0;
        const double ps = MAX(pcL, pcR);
        const double l1 = (ABS(vn) + cm) * k3 * ps;
        const double l2 = (ABS(vn) + cm) * (k1 + k2 * ps);
        *e1 = orc_harten(*e1, l1);
        *e2 = orc_harten(*e2, l2);
        *e3 = orc_harten(*e3, l2);
    } else if (method == 3) {
        double me = ABS(vn) + cm;
        double al = acoustic * me;
        double cl = convect * me;
        *e1 = orc_harten(*e1, cl);
        *e2 = orc_harten(*e2, al);
        *e3 = orc_harten(*e3, al);
    } else if (method == 4) {
        const double mx = MAX(*e2, *e3);
        const double ratio = MIN(1.0, fabs(vn) / cm);
        const double t0 = MAX(mx * acoustic * ratio, mx * acoustic / 100.0);
        const double t1 = MAX(mx * acoustic * ratio, mx * acoustic / 100.0);
        *e1 = MAX(t0, *e1); *e2 = MAX(t1, *e2); *e3 = MAX(t1, *e3);
    } else if (method == 5) {
        double mx = MAX(*e2, *e3);
        double d = mx * acoustic;
        *e1 += d; *e2 += d; *e3 += d;
    } else if (method == 6) {
        const double ps = MIN(pcL, pcR);
        const double lim = (1.0 - ps) * (absVel + cm) + 0.0001 * cm;
        *e1 = orc_harten(*e1, lim);
        *e2 = orc_harten(*e2, lim);
        *e3 = orc_harten(*e3, lim);
    }
}

/* Roe_Scheme_ConservativeForm, API/Physics/src/Flux_Inviscid.cpp:136-442
 * (perfect gas, no preconditioning, static mesh), one face, flux per unit
 * area; pcL, pcR are the face's pressureCoeffL/R (GetPressureFactorLR,
 * NSSolverUnstruct.cpp:5574-5592). */
void orc_roe_flux_pc(const double *qLin, const double *qRin, double xsn, double ysn, double zsn, double gama_,
                     int fix_method, double acoustic, double convect, double pcL, double pcR, double *flux)
{
    const double LITTLE = 1.0E-10;
    double gmL = gama_, gmR = gama_;
    double rL = qLin[IR], uL = qLin[IU], vL = qLin[IV], wL = qLin[IW], pL = qLin[IP];
    double rR = qRin[IR], uR = qRin[IU], vR = qRin[IV], wR = qRin[IW], pR = qRin[IP];
    rL = MAX(rL, LITTLE); rR = MAX(rR, LITTLE);
    pL = MAX(pL, LITTLE); pR = MAX(pR, LITTLE);
    double v2L = uL * uL + vL * vL + wL * wL;
    double v2R = uR * uR + vR * vR + wR * wR;
    double hintL = (gmL / (gmL - 1.0)) * (pL / rL);
    double hintR = (gmR / (gmR - 1.0)) * (pR / rR);
    double hL = hintL + 0.5 * v2L;
    double hR = hintR + 0.5 * v2R;
    double vb = 0.0;
    double vnL = xsn * uL + ysn * vL + zsn * wL - vb;
    double vnR = xsn * uR + ysn * vR + zsn * wR - vb;
    double rvnL = rL * vnL, rvnR = rR * vnR;
    flux[0] = 0.5 * (rvnL + rvnR);
    flux[1] = 0.5 * (rvnL * uL + xsn * pL + rvnR * uR + xsn * pR);
    flux[2] = 0.5 * (rvnL * vL + ysn * pL + rvnR * vR + ysn * pR);
    flux[3] = 0.5 * (rvnL * wL + zsn * pL + rvnR * wR + zsn * pR);
    flux[4] = 0.5 * (rvnL * hL + vb * pL + rvnR * hR + vb * pR);

    double ratio = sqrt(rR / rL);
    double coef = 1.0 / (1.0 + ratio);
    double rRoe = sqrt(rL * rR);
    double uRoe = (uL + uR * ratio) * coef;
    double vRoe = (vL + vR * ratio) * coef;
    double wRoe = (wL + wR * ratio) * coef;
    double pRoe = (pL + pR * ratio) * coef;
    double gama = (gmL + gmR * ratio) * coef;
    double gamm1 = gama - 1.0;
    double v2Roe = uRoe * uRoe + vRoe * vRoe + wRoe * wRoe;
    double absVel = sqrt(v2Roe);
    double hRoe = gama / gamm1 * pRoe / rRoe + 0.5 * v2Roe;
    double vn = xsn * uRoe + ysn * vRoe + zsn * wRoe - vb;
    double theta = vn + vb;
    double c2 = gamm1 * (hRoe - 0.5 * v2Roe);
    double cm = sqrt(ABS(c2));
    double e1 = ABS(vn), e2 = ABS(vn + cm), e3 = ABS(vn - cm);
    double EL = (1.0 / (gmL - 1.0)) * (pL / rL) + 0.5 * v2L;
    double ER = (1.0 / (gmR - 1.0)) * (pR / rR) + 0.5 * v2R;
    double dq0 = rR - rL;
    double dq1 = rR * uR - rL * uL;
    double dq2 = rR * vR - rL * vL;
    double dq3 = rR * wR - rL * wL;
    double dq4 = rR * ER - rL * EL;
    entropy_fix(&e1, &e2, &e3, vn, cm, absVel, pcL, pcR, fix_method, acoustic, convect);
    double xi1 = (2.0 * e1 - e2 - e3) / (2.0 * c2);
    double xi2 = (e2 - e3) / (2.0 * cm);
    double dc = theta * dq0 - xsn * dq1 - ysn * dq2 - zsn * dq3;
    double c2dc = c2 * dc;
    double dh = -gamm1 * (uRoe * dq1 + vRoe * dq2 + wRoe * dq3 - dq4) + 0.5 * gamm1 * v2Roe * dq0;
    flux[0] -= 0.5 * (e1 * dq0 - dh * xi1 - dc * xi2);
    flux[1] -= 0.5 * (e1 * dq1 + (xsn * c2dc - uRoe * dh) * xi1 + (xsn * dh - uRoe * dc) * xi2);
    flux[2] -= 0.5 * (e1 * dq2 + (ysn * c2dc - vRoe * dh) * xi1 + (ysn * dh - vRoe * dc) * xi2);
    flux[3] -= 0.5 * (e1 * dq3 + (zsn * c2dc - wRoe * dh) * xi1 + (zsn * dh - wRoe * dc) * xi2);
    flux[4] -= 0.5 * (e1 * dq4 + (theta * c2dc - hRoe * dh) * xi1 + (theta * dh - hRoe * dc) * xi2);
}

void orc_roe_flux(const double *qLin, const double *qRin, double xsn, double ysn, double zsn,
                  double gama_, int fix_method, double acoustic, double convect, double *flux)
{
    orc_roe_flux_pc(qLin, qRin, xsn, ysn, zsn, gama_, fix_method, acoustic, convect, 0.0, 0.0, flux);
}

/* ------------------------------------------------------------------ */
/* Van Leer flux-vector splitting (Vanleer_Scheme, API/Physics/src/     */
/* Flux_Inviscid.cpp:12697-12949): perfect gas, no preconditioning,     */
/* static grid (face velocity 0), one face, flux per unit area.  It is  */
/* the flux of the flowInitStep phase (ComputeInviscidFlux :14229-14236) */
/* ------------------------------------------------------------------ */
void orc_vanleer_flux(const double *qL, const double *qR, double xsn, double ysn, double zsn, double gama,
                      double *flux)
{
    const double vb = 0.0, gmL = gama, gmR = gama;
    double rL = qL[IR], uL = qL[IU], vL = qL[IV], wL = qL[IW], pL = qL[IP];
    double rR = qR[IR], uR = qR[IU], vR = qR[IV], wR = qR[IW], pR = qR[IP];
    double v2L = uL * uL + vL * vL + wL * wL;
    double v2R = uR * uR + vR * vR + wR * wR;
    double hintL = (gmL / (gmL - 1.0)) * (pL / rL);
    double hintR = (gmR / (gmR - 1.0)) * (pR / rR);
    double HL = hintL + 0.5 * v2L;
    double HR = hintR + 0.5 * v2R;
    double vnL = xsn * uL + ysn * vL + zsn * wL - vb;
    double vnR = xsn * uR + ysn * vR + zsn * wR - vb;
    double rhoVnL = rL * vnL, rhoVnR = rR * vnR;
    double cL = sqrt(ABS(gmL * pL / rL)), cR = sqrt(ABS(gmR * pR / rR));
    double machL = vnL / cL, machR = vnR / cR;
    double fL[5], fR[5];
    if (machL > 1.0) {
        fL[IR] = rhoVnL;
        fL[IU] = rhoVnL * uL + xsn * pL;
        fL[IV] = rhoVnL * vL + ysn * pL;
        fL[IW] = rhoVnL * wL + zsn * pL;
        fL[IP] = rhoVnL * HL + vb * pL;
    } else if (machL < -1.0) {
        for (int m = 0; m < 5; ++m) fL[m] = 0.0;
    } else {
        double fmass = 0.25 * rL * cL * (machL + 1.0) * (machL + 1.0);
        double tmp = (-vnL + 2.0 * cL) / gmL;
        fL[IR] = fmass;
        fL[IU] = fmass * (uL + xsn * tmp);
        fL[IV] = fmass * (vL + ysn * tmp);
        fL[IW] = fmass * (wL + zsn * tmp);
        fL[IP] = fmass * (HL + vb * tmp);
    }
    if (machR < -1.0) {
        fR[IR] = rhoVnR;
        fR[IU] = rhoVnR * uR + xsn * pR;
        fR[IV] = rhoVnR * vR + ysn * pR;
        fR[IW] = rhoVnR * wR + zsn * pR;
        fR[IP] = rhoVnR * HR + vb * pR;
    } else if (machR > 1.0) {
        for (int m = 0; m < 5; ++m) fR[m] = 0.0;
    } else {
        double fmass = -0.25 * rR * cR * (machR - 1.0) * (machR - 1.0);
        double tmp = (-vnR - 2.0 * cR) / gmR;
        fR[IR] = fmass;
        fR[IU] = fmass * (uR + xsn * tmp);
        fR[IV] = fmass * (vR + ysn * tmp);
        fR[IW] = fmass * (wR + zsn * tmp);
        fR[IP] = fmass * (HR + vb * tmp);
    }
    for (int m = 0; m < 5; ++m) flux[m] = fL[m] + fR[m];
}

/* isInIniting of the iterations 1..flowInitStep (Simulation::PreSolve,
 * Main/src/Simulation.cpp:1190-1236): first-order faces, the Van Leer flux
 * and a zero limiter array. */
void orc_set_init_phase(orc_solver *s, int on) { s->initPhase = on ? 1 : 0; }

/* ------------------------------------------------------------------ */
/* Inviscid flux: NSSolverUnstruct::InviscidFlux (:6912-6971)          */
/* GetQlQr (:5507), ReConstructFaceValue reconmeth=1 (:14124-14189),   */
/* BoundaryQlQrFix (:13725-13940), InviscidFluxWrap area scaling        */
/* (:14476-14487), LoadFlux (:6435-6494).                               */
/* ------------------------------------------------------------------ */

static int positive_check(const double *q) /* FaceProxy.h:209-213 */
{
    return !(q[IR] <= 0.0 || q[IP] <= 0.0);
}

static void face_values(const orc_solver *s, int f, double *qL, double *qR)
{
    int le = s->left[f], re = s->right[f];
    for (int m = 0; m < 5; ++m) { qL[m] = s->q[m][le]; qR[m] = s->q[m][re]; }
    /* ReConstructFaceValue returns at once for first-order faces (:14005-14021) */
    if (s->p.limiter != 2 && !s->initPhase) {
        double qt[5];
        double lmin = MIN(s->lim[le], s->lim[re]);
        double dx = s->xfc[f] - s->xcc[le], dy = s->yfc[f] - s->ycc[le], dz = s->zfc[f] - s->zcc[le];
        for (int m = 0; m < 5; ++m) qt[m] = qL[m];
        for (int m = 0; m < 5; ++m)
            qt[m] += lmin * (s->dqdx[m][le] * dx + s->dqdy[m][le] * dy + s->dqdz[m][le] * dz);
        if (positive_check(qt)) for (int m = 0; m < 5; ++m) qL[m] = qt[m];
        dx = s->xfc[f] - s->xcc[re]; dy = s->yfc[f] - s->ycc[re]; dz = s->zfc[f] - s->zcc[re];
        for (int m = 0; m < 5; ++m) qt[m] = qR[m];
        for (int m = 0; m < 5; ++m)
            qt[m] += lmin * (s->dqdx[m][re] * dx + s->dqdy[m][re] * dy + s->dqdz[m][re] * dz);
        if (positive_check(qt)) for (int m = 0; m < 5; ++m) qR[m] = qt[m];
    }
    if (f < s->nBFace) {
        int bc = s->bcType[f];
        if (bc != ORC_BC_SOLID_SURFACE && bc != ORC_BC_SYMMETRY) return;
        for (int m = 0; m < 5; ++m) {
            double tmp = 0.5 * (s->q[m][le] + s->q[m][re]);
            qL[m] = tmp; qR[m] = tmp;
        }
        if (bc == ORC_BC_SYMMETRY) {
            double nx = s->xfn[f], ny = s->yfn[f], nz = s->zfn[f];
            double vn = 2.0 * (nx * qL[IU] + ny * qL[IV] + nz * qL[IW] - 0.0);
            qR[IR] = qL[IR];
            qR[IU] = qL[IU] - vn * nx;
            qR[IV] = qL[IV] - vn * ny;
            qR[IW] = qL[IW] - vn * nz;
            qR[IP] = qL[IP];
        }
        if (bc == ORC_BC_SOLID_SURFACE && s->p.viscous) {
            qL[IU] = 0.0; qL[IV] = 0.0; qL[IW] = 0.0;
            qR[IU] = 0.0; qR[IV] = 0.0; qR[IW] = 0.0;
        }
    }
}

void orc_face_states(const orc_solver *s, int f, double *qL, double *qR)
{
    face_values(s, f, qL, qR);
}

void orc_zero_residual(orc_solver *s)
{
    /* ZeroResiduals (:5642) then LoadResiduals res = -rhs with rhs = 0 (:4116). */
    for (int m = 0; m < 5; ++m)
        for (int c = 0; c < s->nTotal; ++c) s->res[m][c] = -0.0;
}

void orc_inviscid_flux(orc_solver *s)
{
    FACE_LOOP(s, 0, s->nFace, f)
        double qL[5], qR[5], fl[5];
        face_values(s, f, qL, qR);
        if (s->initPhase)
            orc_vanleer_flux(qL, qR, s->xfn[f], s->yfn[f], s->zfn[f], s->p.gamma, fl);
        else
            orc_roe_flux_pc(qL, qR, s->xfn[f], s->yfn[f], s->zfn[f], s->p.gamma, s->p.entropy_fix,
                            s->entropyCoef, s->entropyCoef, s->rtem[s->left[f]], s->rtem[s->right[f]], fl);
        double a = s->area[f];
        for (int m = 0; m < 5; ++m) fl[m] *= a;
        int le = s->left[f], re = s->right[f];
        if (OWN(le))
            for (int m = 0; m < 5; ++m) s->res[m][le] -= fl[m];
        if (f >= s->nBFace && OWN(re))
            for (int m = 0; m < 5; ++m) s->res[m][re] += fl[m];
    FACE_LOOP_END
}

/* ------------------------------------------------------------------ */
/* Viscous flux: NSSolverUnstruct::ViscousFlux (:7058-7114), FaceWeight */
/* (Geo_UnstructGrid.cpp:2420-2484), GetVisFaceValue (:7126-7331),      */
/* ComputeVisflux (:12977-13396).  Laminar, perfect gas, static mesh.   */
/* ------------------------------------------------------------------ */
void orc_viscous_flux(orc_solver *s)
{
    if (!s->p.viscous) return;
    const double two3rd = 2.0 / 3.0;
    const double skewAngle = 60.0;
    double oRe = 1.0 / s->p.reynolds;
    double oprl = 1.0 / s->p.prandtl, oprt = 1.0 / s->p.prandtl_turb;
    double M2 = s->p.mach * s->p.mach;
    double cp = 1.0 / ((s->p.gamma - 1.0) * M2);
    FACE_LOOP(s, 0, s->nFace, f)
        int le = s->left[f], re = s->right[f];
        int bnd = f < s->nBFace;
        int wall = bnd && is_wall(s->bcType[f]);
        /* FaceWeight */
        double dxL0 = s->xfc[f] - s->xcc[le], dyL0 = s->yfc[f] - s->ycc[le], dzL0 = s->zfc[f] - s->zcc[le];
        double dxR0 = s->xfc[f] - s->xcc[re], dyR0 = s->yfc[f] - s->ycc[re], dzR0 = s->zfc[f] - s->zcc[re];
        double d1 = sqrt(dxL0 * dxL0 + dyL0 * dyL0 + dzL0 * dzL0);
        double d2 = sqrt(dxR0 * dxR0 + dyR0 * dyR0 + dzR0 * dzR0);
        double dd = 1.0 / (d1 + d2 + SMALL);
        double wL = d2 * dd, wR = d1 * dd;
        if (bnd) { wL = 1.0; wR = 0.0; }
        /* GetVisFaceValue */
        double mul = 0.5 * (s->visl[le] + s->visl[re]);
        double mut = 0.5 * (s->vist[le] + s->vist[re]);
        double primF[5];
        for (int m = 0; m < 5; ++m) primF[m] = 0.5 * (s->q[m][le] + s->q[m][re]);
        double tmF = 0.5 * (s->t[le] + s->t[re]);
        double kcp = (mul * oprl + mut * oprt) * cp;
        if (wall) { primF[IU] = 0.0; primF[IV] = 0.0; primF[IW] = 0.0; }
        /* ComputeVisflux */
        double nxs = s->xfn[f], nys = s->yfn[f], nzs = s->zfn[f];
        double t1x, t1y, t1z;
        if (ABS(nxs) > SMALL) { t1x = nys; t1y = -nxs; t1z = 0.0; }
        else if (ABS(nys) > SMALL) { t1x = -nys; t1y = nxs; t1z = 0.0; }
        else if (ABS(nzs) > SMALL) { t1x = 0.0; t1y = -nzs; t1z = nys; }
        else continue; /* degenerate face: zero flux */
        double on = 1.0 / sqrt(t1x * t1x + t1y * t1y + t1z * t1z);
        t1x *= on; t1y *= on; t1z *= on;
        double t2x = nys * t1z - nzs * t1y;
        double t2y = nzs * t1x - nxs * t1z;
        double t2z = nxs * t1y - nys * t1x;
        double dxL = s->xcc[le] - s->xfc[f], dyL = s->ycc[le] - s->yfc[f], dzL = s->zcc[le] - s->zfc[f];
        double dxR = s->xcc[re] - s->xfc[f], dyR = s->ycc[re] - s->yfc[f], dzR = s->zcc[re] - s->zfc[f];
        double dL = nxs * dxL + nys * dyL + nzs * dzL;
        double dR = nxs * dxR + nys * dyR + nzs * dzR;
        double dt_ = -dL / (sqrt(dxL * dxL + dyL * dyL + dzL * dzL) + SMALL);
        dt_ = MIN(dt_, 1.0); dt_ = MAX(dt_, -1.0);
        double angle1 = asin(dt_) * 180.0 / PI_;
        dt_ = dR / (sqrt(dxR * dxR + dyR * dyR + dzR * dzR) + SMALL);
        dt_ = MIN(dt_, 1.0); dt_ = MAX(dt_, -1.0);
        double angle2 = asin(dt_) * 180.0 / PI_;
        double dxnL = nxs * dL - dxL, dynL = nys * dL - dyL, dznL = nzs * dL - dzL;
        double dxnR = nxs * dR - dxR, dynR = nys * dR - dyR, dznR = nzs * dR - dzR;
        double f1[5], f2[5], fM[5];
        for (int m = 0; m < 5; ++m) { f1[m] = s->q[m][le]; f2[m] = s->q[m][re]; }
        double tL = s->t[le], tR = s->t[re];
        for (int m = 0; m < 5; ++m) fM[m] = 0.5 * (f1[m] + f2[m]);
        double tM = 0.5 * (tL + tR);
        if (wall) {
            tM = tmF;
            fM[IR] = fM[IP] / (s->R * tM);
        }
        if (angle1 > skewAngle && angle2 > skewAngle) {
            for (int m = 0; m < 5; ++m) {
                f1[m] += s->dqdx[m][le] * dxnL + s->dqdy[m][le] * dynL + s->dqdz[m][le] * dznL;
                f2[m] += s->dqdx[m][re] * dxnR + s->dqdy[m][re] * dynR + s->dqdz[m][re] * dznR;
            }
            tL += s->dtdx[le] * dxnL + s->dtdy[le] * dynL + s->dtdz[le] * dznL;
            tR += s->dtdx[re] * dxnR + s->dtdy[re] * dynR + s->dtdz[re] * dznR;
            if (tL < SMALL) tL = s->t[le];
            if (tR < SMALL) tR = s->t[re];
            for (int m = 0; m < 5; ++m) fM[m] = primF[m];
            tM = tmF;
        }
        double dfdn[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        double dtdn = 0.0;
        if (angle1 > 0.0 && angle2 > 0.0 && ABS(dL) > TINY && ABS(dR) > TINY) {
            double dfd1[5], dfd2[5];
            for (int m = 0; m < 5; ++m) dfd1[m] = (f1[m] - fM[m]) / dL;
            for (int m = 0; m < 5; ++m) dfd2[m] = (f2[m] - fM[m]) / dR;
            double dtd1 = (tL - tM) / dL;
            double dtd2 = (tR - tM) / dR;
            double dtmp = dL * dL + dR * dR;
            double wl = dL * dL / dtmp;
            double wr = 1.0 - wl;
            if (bnd) { wl = 1.0; wr = 0.0; }
            for (int m = 0; m < 5; ++m) dfdn[m] = wl * dfd1[m] + wr * dfd2[m];
            dtdn = dtd1 * wl + dtd2 * wr;
        }
        double dfdx[5], dfdy[5], dfdz[5], dft1[5], dft2[5];
        for (int m = 0; m < 5; ++m) {
            dfdx[m] = wL * s->dqdx[m][le] + wR * s->dqdx[m][re];
            dfdy[m] = wL * s->dqdy[m][le] + wR * s->dqdy[m][re];
            dfdz[m] = wL * s->dqdz[m][le] + wR * s->dqdz[m][re];
        }
        for (int m = IU; m <= IW; ++m) {
            dft1[m] = t1x * dfdx[m] + t1y * dfdy[m] + t1z * dfdz[m];
            dft2[m] = t2x * dfdx[m] + t2y * dfdy[m] + t2z * dfdz[m];
        }
        for (int m = IU; m <= IW; ++m) {
            dfdx[m] = nxs * dfdn[m] + t1x * dft1[m] + t2x * dft2[m];
            dfdy[m] = nys * dfdn[m] + t1y * dft1[m] + t2y * dft2[m];
            dfdz[m] = nzs * dfdn[m] + t1z * dft1[m] + t2z * dft2[m];
        }
        double dudx = dfdx[IU], dudy = dfdy[IU], dudz = dfdz[IU];
        double dvdx = dfdx[IV], dvdy = dfdy[IV], dvdz = dfdz[IV];
        double dwdx = dfdx[IW], dwdy = dfdy[IW], dwdz = dfdz[IW];
        double qNorm = 0.0;
        qNorm += kcp * dtdn;
        double divv2p3 = two3rd * (dudx + dvdy + dwdz);
        double vis = mul + mut;
        double um = fM[1], vm = fM[2], wm = fM[3];
        double txx = vis * (2.0 * dudx - divv2p3);
        double tyy = vis * (2.0 * dvdy - divv2p3);
        double tzz = vis * (2.0 * dwdz - divv2p3);
        double txy = vis * (dudy + dvdx);
        double txz = vis * (dudz + dwdx);
        double tyz = vis * (dvdz + dwdy);
        double fv[5];
        fv[0] = 0.0;
        fv[1] = nxs * txx + nys * txy + nzs * txz;
        fv[2] = nxs * txy + nys * tyy + nzs * tyz;
        fv[3] = nxs * txz + nys * tyz + nzs * tzz;
        fv[4] = um * fv[1] + vm * fv[2] + wm * fv[3] + qNorm;
        double fl[5];
        for (int m = 0; m < 5; ++m) fl[m] = -oRe * s->area[f] * fv[m];
        if (OWN(le))
            for (int m = 0; m < 5; ++m) s->res[m][le] -= fl[m];
        if (!bnd && OWN(re))
            for (int m = 0; m < 5; ++m) s->res[m][re] += fl[m];
    FACE_LOOP_END
}

/* Controller::UpdateResiduals (Controller.cpp:1002-1061) + LoadResiduals:
 * boundary, independent fields, then the inviscid and viscous face loops. */
/* NSSolver::UpdateResiduals (CFD/NSSolver/src/NSSolver.cpp:71-82): the
 * inviscid and viscous face loops, added to res. */
void orc_update_residuals(orc_solver *s)
{
    orc_inviscid_flux(s);
    orc_viscous_flux(s);
}

/* Controller::LoadResiduals + Controller::UpdateResiduals
 * (CFD/Solver/src/Controller.cpp:1176,1002-1083): res = -rhs, Boundary
 * (SetBoundaryValueOfField :1467), UpdateIndependentField (:1715: gamma/T,
 * viscosity, gradient, limiter), then NSSolver::UpdateResiduals. */
void orc_residual(orc_solver *s)
{
    orc_zero_residual(s);
    orc_boundary(s);
    /* ggnode: the node values the reference's PostSolve of the previous
     * iteration computed (Controller.cpp:5151-5198): the state with fresh
     * ghosts, and the temperature of that iteration's UpdateIndependentField */
    if (s->p.gradient == 2) orc_node_values(s);
    orc_gamma_t(s);
    orc_pressure_factor(s);
    if (s->p.viscous) orc_viscosity(s);
    orc_gradient(s);
    orc_limiter(s);
    orc_update_residuals(s);
}

/* ------------------------------------------------------------------ */
/* Spectral radius and time step                                       */
/* ------------------------------------------------------------------ */

/* SpectrumRadiusInviscid (:5749-5850), SpectrumRadiusViscous method 2 (:5926-5963). */
void orc_spectrum(orc_solver *s)
{
    int nC = s->nCell;
    CELL_LOOP
    for (int c = 0; c < nC; ++c) s->specInv[c] = 0.0;
    FACE_LOOP(s, 0, s->nFace, f)
        int le = s->left[f], re = s->right[f];
        double nx = s->xfn[f], ny = s->yfn[f], nz = s->zfn[f];
        double rl = s->q[IR][le], ul = s->q[IU][le], vl = s->q[IV][le], wl = s->q[IW][le], pl = s->q[IP][le];
        double vnl = nx * ul + ny * vl + nz * wl - 0.0;
        double rr = s->q[IR][re], ur = s->q[IU][re], vr = s->q[IV][re], wr = s->q[IW][re], pr = s->q[IP][re];
        double vnr = nx * ur + ny * vr + nz * wr - 0.0;
        double g = 0.5 * (s->gama[le] + s->gama[re]);
        double pm = 0.5 * (pl + pr);
        double rm = 0.5 * (rl + rr);
        double vn = 0.5 * (vnl + vnr);
        double cm = sqrt(g * pm / rm);
        double sr = 0.5 * s->area[f] * (ABS(vn) + cm);
        if (OWN(le)) s->specInv[le] += sr;
        if (re < nC && OWN(re)) s->specInv[re] += sr;
    FACE_LOOP_END
    CELL_LOOP
    for (int c = 0; c < nC; ++c) s->specVis[c] = 0.0;
    if (!s->p.viscous) return;
    FACE_LOOP(s, 0, s->nFace, f)
        int le = s->left[f], re = s->right[f];
        double dx = s->xcc[re] - s->xcc[le], dy = s->ycc[re] - s->ycc[le], dz = s->zcc[re] - s->zcc[le];
        double ds = ABS(s->xfn[f] * dx + s->yfn[f] * dy + s->zfn[f] * dz);
        double vl = 0.5 * (s->visl[le] + s->visl[re]);
        double vt = 0.5 * (s->vist[le] + s->vist[re]);
        double rho = 0.5 * (s->q[IR][le] + s->q[IR][re]);
        double vis = vl + vt;
        double sr = 2.0 * vis / (rho * ds * s->p.reynolds + SMALL);
        sr *= 0.5 * s->area[f];
        if (OWN(le)) s->specVis[le] += sr;
        if (re < nC && OWN(re)) s->specVis[re] += sr;
    FACE_LOOP_END
}

/* NSSolver::ComputeCFL (CFD/NSSolver/src/NSSolver.cpp:595-668), CFLMethod 0. */
static double compute_cfl(const orc_solver *s, int step)
{
    if (step >= s->p.cfl_vary_step) return s->p.cfl_end;
    return s->p.cfl_start + (s->p.cfl_end - s->p.cfl_start) * step / s->p.cfl_vary_step;
}

/* LocalTimeStep (:5999-6050), LimitCFL (:6052-6118), ComputeMinTimeStep
 * (:5697-5714), ReduceMaxTimeStep (:5716-5743). dt is stored divided by volume. */
void orc_time_step(orc_solver *s, int step)
{
    int nC = s->nCell;
    double cfl = compute_cfl(s, step);
    double cflStart = s->p.cfl_start;
    double pLower = 0.001 * s->pTotal, pUpper = 0.02 * s->pTotal;
    CELL_LOOP
    for (int c = 0; c < nC; ++c) {
        double spec = s->specInv[c];
        if (s->p.viscous) spec += s->specVis[c];
        double cc = cfl;
        double p = s->q[IP][c];
        if (p < pLower) cc = cflStart;
        else if (p < pUpper) cc = (p - pUpper) * (cfl - cflStart) / (pUpper - pLower) + cfl;
        double sk = s->skew[c];
        if (sk < 0.1) cc *= 0.1;
        else if (sk < 10.0) cc *= (1.0 - exp(0.0 - sk));
        s->dt[c] = cc * s->vol[c] / spec;
    }
    double mn = LARGE;
    for (int c = 0; c < nC; ++c) mn = MIN(mn, s->dt[c]);
    if (s->p.ktmax > 0) {
        CELL_LOOP
        for (int c = 0; c < nC; ++c) s->dt[c] = MIN(s->dt[c], s->p.ktmax * mn);
    }
    CELL_LOOP
    for (int c = 0; c < nC; ++c) s->dt[c] /= s->vol[c];
}

/* ------------------------------------------------------------------ */
/* LU-SGS                                                              */
/* ------------------------------------------------------------------ */

/* MXDQ_STD, API/Physics/src/Jacobian.cpp:13-86 with
 * PerfectGas::ComputeTotalEnthalpyAndDH (PerfectGas.cpp:368-400). */
void orc_mxdq(const double *prim, double nxs, double nys, double nzs, double area,
              double gama, const double *dq, double radius, int ipn, double *f)
{
    double vgn = 0.0;
    double rm = prim[IR], um = prim[IU], vm = prim[IV], wm = prim[IW], pm = prim[IP];
    double c2 = ABS(gama * pm / rm);
    double cm = sqrt(c2);
    double cits = nxs * um + nys * vm + nzs * wm - vgn;
    double lmd1 = cits, lmd2 = cits + cm, lmd3 = cits - cm;
    double ipnrad = ipn * radius / MAX(SMALL, area);
    lmd1 = 0.5 * (lmd1 + ipnrad);
    lmd2 = 0.5 * (lmd2 + ipnrad);
    lmd3 = 0.5 * (lmd3 + ipnrad);
    double x1 = (lmd1 + lmd1 - lmd2 - lmd3) / (c2 + c2);
    double x2 = (lmd2 - lmd3) / (cm + cm);
    double dc = (cits + vgn) * dq[0] - nxs * dq[1] - nys * dq[2] - nzs * dq[3];
    double c2dc = c2 * dc;
    double v2 = um * um + vm * vm + wm * wm;
    double gama1 = gama - 1.0;
    double cc1 = 0.5 * gama1 * v2;
    double dh = -gama1 * (um * dq[1] + vm * dq[2] + wm * dq[3] - dq[4]);
    double enthalpy = (gama / gama1) * (pm / rm);
    dh += cc1 * dq[0];
    double hm = enthalpy + 0.5 * v2;
    double a = dh * x1 + dc * x2;
    double b = c2dc * x1 + dh * x2;
    f[0] = lmd1 * dq[0] - a;
    f[1] = lmd1 * dq[1] - um * a + nxs * b;
    f[2] = lmd1 * dq[2] - vm * a + nys * b;
    f[3] = lmd1 * dq[3] - wm * a + nzs * b;
    f[4] = lmd1 * dq[4] - hm * a + (cits + vgn) * b;
    for (int m = 0; m < 5; ++m) f[m] *= area;
}

/* RoeAveragedVariables (Flux_Inviscid.cpp:15440-15462) + GetRad (:15097-15116):
 * (|V.n| + c) * S at the Roe-averaged face state. */
static double lusgs_rad(const double *ql, double gl, const double *qr, double gr,
                        double nx, double ny, double nz, double area, double gama)
{
    double hl = gl / (gl - 1.0) * ql[IP] / ql[IR] + 0.5 * (ql[IU] * ql[IU] + ql[IV] * ql[IV] + ql[IW] * ql[IW]);
    double hr = gr / (gr - 1.0) * qr[IP] / qr[IR] + 0.5 * (qr[IU] * qr[IU] + qr[IV] * qr[IV] + qr[IW] * qr[IW]);
    double t0 = sqrt(qr[IR] / ql[IR]);
    double t1 = 1.0 / (1.0 + t0);
    double rm = sqrt(qr[IR] * ql[IR]);
    double um = (ql[IU] + qr[IU] * t0) * t1;
    double vm = (ql[IV] + qr[IV] * t0) * t1;
    double wm = (ql[IW] + qr[IW] * t0) * t1;
    double hm = (hl + hr * t0) * t1;
    double pm = rm * (hm - 0.5 * (um * um + vm * vm + wm * wm)) * (gl - 1.0) / gl;
    double vn = nx * um + ny * vm + nz * wm - 0.0;
    double eig = ABS(vn) + sqrt(gama * pm / rm);
    return eig * area;
}

double orc_lusgs_rad(const double *ql, const double *qr, double nx, double ny, double nz,
                     double area, double gamma)
{
    return lusgs_rad(ql, gamma, qr, gamma, nx, ny, nz, area, gamma);
}

/* One neighbour contribution of SolveLUSGSForward/Backward
 * (NSSolverUnstruct.cpp:15227-15305, 15582-15660). */
static void lusgs_face(const orc_solver *s, int f, int cell, double *df)
{
    int le = s->left[f], re = s->right[f];
    double nx = s->xfn[f], ny = s->yfn[f], nz = s->zfn[f], ar = s->area[f];
    if (re == cell) { int t = le; le = re; re = t; nx = -nx; ny = -ny; nz = -nz; }
    double ql[5], qr[5], dqn[5];
    for (int m = 0; m < 5; ++m) { ql[m] = s->q[m][le]; qr[m] = s->q[m][re]; dqn[m] = s->dq[m][re]; }
    double gama = s->gama[re];
    double rad = lusgs_rad(ql, s->gama[le], qr, s->gama[re], nx, ny, nz, ar, gama);
    orc_mxdq(qr, nx, ny, nz, ar, gama, dqn, rad, -1, df);
    if (s->p.viscous) {
        double dist = ABS(s->xfn[f] * (s->xcc[re] - s->xcc[le]) + s->yfn[f] * (s->ycc[re] - s->ycc[le]) +
                          s->zfn[f] * (s->zcc[re] - s->zcc[le]));
        double vl = 0.5 * (s->visl[le] + s->visl[re]);
        double vt = 0.5 * (s->vist[le] + s->vist[re]);
        double rho = 0.5 * (s->q[IR][le] + s->q[IR][re]);
        double vis = vl + vt;
        double sr = 2.0 * vis / (rho * dist * s->p.reynolds + SMALL);
        sr *= 0.5 * ar;
        for (int m = 0; m < 5; ++m) df[m] -= sr * dqn[m];
    }
}

/* SolveLUSGSForward, NSSolverUnstruct.cpp:15118-15444. */
/* One cell of the forward sweep; d2[m] receives (dq - dq_old)^2. */
static void lusgs_forward_cell(orc_solver *s, int c, double *d2)
{
    {
        double rhs0[5], dqOld[5], Ux[5], df[5];
        for (int m = 0; m < 5; ++m) {
            rhs0[m] = 0.0;
            dqOld[m] = s->dq[m][c];
            s->dq[m][c] = s->res[m][c];
            Ux[m] = s->dflux[m][c];
            s->dflux[m][c] = 0.0;
        }
        for (int k = s->c2fPtr[c]; k < s->c2fPtr[c + 1]; ++k) {
            int f = s->c2f[k];
            int le = s->left[f], re = s->right[f];
            if (le > c || re > c) continue;
            lusgs_face(s, f, c, df);
            for (int m = 0; m < 5; ++m) rhs0[m] += df[m];
        }
        double dqTry[5];
        for (int m = 0; m < 5; ++m) {
            dqTry[m] = (s->dq[m][c] - Ux[m] - rhs0[m]) / s->diag[c];
            s->dflux[m][c] += rhs0[m];
        }
        for (int m = 0; m < 5; ++m) {
            s->dq[m][c] = dqTry[m];
            double d = s->dq[m][c] - dqOld[m];
            d2[m] = d * d;
        }
    }
}

/* One cell of the backward sweep (SolveLUSGSBackward,
 * NSSolverUnstruct.cpp:15446-15764). */
static void lusgs_backward_cell(orc_solver *s, int c, double *d2)
{
    {
        double rhs0[5], dqOld[5], Lx[5], df[5];
        for (int m = 0; m < 5; ++m) {
            rhs0[m] = 0.0;
            dqOld[m] = s->dq[m][c];
            s->dq[m][c] = s->res[m][c];
            Lx[m] = s->dflux[m][c];
            s->dflux[m][c] = 0.0;
        }
        for (int k = s->c2fPtr[c]; k < s->c2fPtr[c + 1]; ++k) {
            int f = s->c2f[k];
            int le = s->left[f], re = s->right[f];
            if (le < c || re < c) continue;
            lusgs_face(s, f, c, df);
            for (int m = 0; m < 5; ++m) rhs0[m] += df[m];
        }
        double dqTry[5];
        for (int m = 0; m < 5; ++m) {
            dqTry[m] = (s->dq[m][c] - Lx[m] - rhs0[m]) / s->diag[c];
            s->dflux[m][c] += rhs0[m];
        }
        for (int m = 0; m < 5; ++m) {
            s->dq[m][c] = dqTry[m];
            double d = s->dq[m][c] - dqOld[m];
            d2[m] = d * d;
        }
    }
}

/* A sweep in cell order (forward: increasing, backward: decreasing), the
 * squared changes summed in that order.  The OpenMP build runs the cells of
 * each dependency level in parallel and sums the same terms in the same
 * order afterwards. */
static double lusgs_sweep(orc_solver *s, int forward)
{
    const int nC = s->nCell;
    double sweep = 0.0;
#ifdef _OPENMP
    struct orc_part *P = orc_partition(s);
    double *d2 = (double *)malloc(sizeof(double) * 5 * (size_t)nC);
    const int nl = forward ? P->nLvF : P->nLvB;
    const int *ptr = forward ? P->lvPtrF : P->lvPtrB, *cells = forward ? P->lvCellsF : P->lvCellsB;
    for (int l = 0; l < nl; ++l) {
#pragma omp parallel for schedule(static)
        for (int i = ptr[l]; i < ptr[l + 1]; ++i) {
            const int c = cells[i];
            if (forward) lusgs_forward_cell(s, c, d2 + 5 * (size_t)c);
            else lusgs_backward_cell(s, c, d2 + 5 * (size_t)c);
        }
    }
    for (int i = 0; i < nC; ++i) {
        const int c = forward ? i : nC - 1 - i;
        for (int m = 0; m < 5; ++m) sweep += d2[5 * (size_t)c + m];
    }
    free(d2);
#else
    for (int i = 0; i < nC; ++i) {
        const int c = forward ? i : nC - 1 - i;
        double d2[5];
        if (forward) lusgs_forward_cell(s, c, d2);
        else lusgs_backward_cell(s, c, d2);
        for (int m = 0; m < 5; ++m) sweep += d2[m];
    }
#endif
    return sweep;
}

/* PerfectGas::Primitive2Conservative (PerfectGas.cpp:284-317, non-porous). */
static void prim2cons(const double *prim, double gama, double *Q)
{
    double um = prim[IU], vm = prim[IV], wm = prim[IW], rho = prim[IR], p = prim[IP];
    double v2 = um * um + vm * vm + wm * wm;
    double em = (p / rho) / (gama - 1.0) + 0.5 * v2;
    Q[0] = prim[IR];
    Q[1] = prim[IR] * prim[IU];
    Q[2] = prim[IR] * prim[IV];
    Q[3] = prim[IR] * prim[IW];
    Q[4] = prim[IR] * em;
}

/* SetGhostDQLUSGS (:14767-14824) with SetSymmBCGhostDQLUSGS (:14826-14919),
 * SetWallBCGhostDQLUSGS (:14921-15000), SetExtraBCGhostDQLUSGS (:15002-15090). */
static void set_ghost_dq(orc_solver *s)
{
    CELL_LOOP
    for (int f = 0; f < s->nBFace; ++f) {
        int bc = s->bcType[f];
        int le = s->left[f], re = s->right[f];
        if (bc == ORC_BC_SYMMETRY || (is_wall(bc) && !s->p.viscous)) {
            double prim[5], Q[5];
            for (int m = 0; m < 5; ++m) prim[m] = s->q[m][le];
            prim2cons(prim, s->gama[le], Q);
            for (int m = 0; m < 5; ++m) Q[m] += s->dq[m][le];
            double rm = s->q[IR][re], um = s->q[IU][re], vm = s->q[IV][re], wm = s->q[IW][re], pm = s->q[IP][re];
            for (int m = 0; m < 5; ++m) s->dq[m][re] = s->dq[m][le];
            double nx = s->xfn[f], ny = s->yfn[f], nz = s->zfn[f];
            s->dq[IR][re] = s->dq[IR][le];
            double rvn = 2.0 * (nx * s->dq[1][le] + ny * s->dq[2][le] + nz * s->dq[3][le]);
            s->dq[1][re] = s->dq[1][le] - nx * rvn;
            s->dq[2][re] = s->dq[2][le] - ny * rvn;
            s->dq[3][re] = s->dq[3][le] - nz * rvn;
            rvn = 2.0 * (nx * Q[1] + ny * Q[2] + nz * Q[3]);
            double ru = Q[1] - nx * rvn, rv = Q[2] - ny * rvn, rw = Q[3] - nz * rvn;
            double pnew = Q[4] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) / Q[0];
            double gam1 = s->gama[le] - 1.0;
            s->dq[4][re] = pnew - pm / gam1 + 0.5 * ((ru * ru + rv * rv + rw * rw) / Q[0] - rm * (um * um + vm * vm + wm * wm));
        } else if (is_wall(bc) && s->p.viscous) {
            for (int m = 0; m < 5; ++m) s->dq[m][re] = s->dq[m][le];
            double ub = 0.0, vb = 0.0, wb = 0.0;
            s->dq[0][re] = s->dq[0][le];
            s->dq[1][re] = -s->dq[1][le] + 2.0 * ub * s->dq[0][le];
            s->dq[2][re] = -s->dq[2][le] + 2.0 * vb * s->dq[0][le];
            s->dq[3][re] = -s->dq[3][le] + 2.0 * wb * s->dq[0][le];
            s->dq[4][re] = s->dq[4][le] + ub * (s->dq[1][re] - s->dq[1][le]) + vb * (s->dq[2][re] - s->dq[2][le]) +
                           wb * (s->dq[3][re] - s->dq[3][le]);
        } else if (bc == ORC_BC_EXTRAPOLATION || bc == ORC_BC_OUTFLOW) {
            for (int m = 0; m < 5; ++m) s->dq[m][re] = s->dq[m][le];
        } else {
            for (int m = 0; m < 5; ++m) s->dq[m][re] = 0.0;
        }
    }
}

/* Diagonal (:15867-15960), steady. */
static void diagonal(orc_solver *s)
{
    CELL_LOOP
    for (int c = 0; c < s->nCell; ++c) {
        s->diag[c] = 1.0 / s->dt[c];
        s->diag[c] += s->specInv[c];
        if (s->p.viscous) s->diag[c] += s->specVis[c];
    }
}

/* SolutionFix (:14509-14585): average of admissible neighbours, read from
 * the partially updated field exactly as the in-place reference loop does. */
/* lower: the field the neighbours below c are read from (the in-place
 * update has already replaced them); neighbours above c come from s->q. */
static void solution_fix(const orc_solver *s, double *const *lower, double *prim, int c)
{
    double roo = s->primInf[IR], poo = s->primInf[IP];
    double dl = 1.0e-6 * roo, pl = 1.0e-6 * poo;
    for (int m = 0; m < 5; ++m) prim[m] = 0.0;
    double vs = 0.0;
    for (int k = s->c2fPtr[c]; k < s->c2fPtr[c + 1]; ++k) {
        int f = s->c2f[k];
        int le = s->left[f], re = s->right[f];
        int nb = le;
        if (c == le) nb = re;
        double *const *src = nb < c ? lower : s->q;
        if (src[IR][nb] < dl || src[IP][nb] < pl) continue;
        double vsur = 1.0;
        vs += vsur;
        for (int m = 0; m < 5; ++m) {
            double v = src[m][nb];
            if (m == IR || m == IP) v = ABS(v);
            prim[m] += v * vsur;
        }
    }
    if (vs > 0.0) {
        double o = 1.0 / vs;
        for (int m = 0; m < 5; ++m) prim[m] *= o;
    } else {
        for (int m = 0; m < 5; ++m) prim[m] = s->primInf[m];
    }
}

/* UpdateFlowField (:14587-14765), perfect gas, in place. */
/* Q += dq for one cell, converted back to primitives (out). */
static void update_cell(const orc_solver *s, int c, double *prim)
{
    {
        double Q[5];
        for (int m = 0; m < 5; ++m) prim[m] = s->q[m][c];
        double g = s->gama[c];
        prim2cons(prim, g, Q);
        for (int m = 0; m < 5; ++m) Q[m] += s->dq[m][c];
        /* Conservative2Primitive, PerfectGas.cpp:327-352 */
        if (Q[0] <= 0.0) {
            prim[IR] = Q[0];
        } else {
            double od = 1.0 / Q[0];
            double um = Q[1] * od, vm = Q[2] * od, wm = Q[3] * od;
            double ie = Q[4] - 0.5 * Q[0] * (um * um + vm * vm + wm * wm);
            prim[IR] = Q[0]; prim[IU] = um; prim[IV] = vm; prim[IW] = wm;
            prim[IP] = (g - 1.0) * ie;
        }
    }
}

/* UpdateFlowField: in cell order, each cell updated in place and repaired by
 * SolutionFix if its density or pressure falls below 1e-6 of freestream.
 * The OpenMP build updates every cell into a new field first, then repairs
 * the flagged cells in increasing order reading lower neighbours from the
 * new field and higher ones from the old (what the in-place loop sees). */
static int update_flow(orc_solver *s)
{
    double dl = 1.0e-6 * s->primInf[IR], pl = 1.0e-6 * s->primInf[IP];
    int nNeg = 0;
#ifdef _OPENMP
    const int nC = s->nCell;
    double *nw[5];
    for (int m = 0; m < 5; ++m) nw[m] = (double *)malloc(sizeof(double) * (size_t)nC);
    unsigned char *bad = (unsigned char *)malloc((size_t)nC);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nC; ++c) {
        double prim[5];
        update_cell(s, c, prim);
        for (int m = 0; m < 5; ++m) nw[m][c] = prim[m];
        bad[c] = prim[IR] < dl || prim[IP] < pl;
    }
    for (int c = 0; c < nC; ++c) {
        if (!bad[c]) continue;
        double prim[5];
        nNeg++;
        solution_fix(s, nw, prim, c);
        for (int m = 0; m < 5; ++m) nw[m][c] = prim[m];
    }
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nC; ++c)
        for (int m = 0; m < 5; ++m) s->q[m][c] = nw[m][c];
    for (int m = 0; m < 5; ++m) free(nw[m]);
    free(bad);
#else
    for (int c = 0; c < s->nCell; ++c) {
        double prim[5];
        update_cell(s, c, prim);
        if (prim[IR] < dl || prim[IP] < pl) {
            nNeg++;
            solution_fix(s, s->q, prim, c);
        }
        for (int m = 0; m < 5; ++m) s->q[m][c] = prim[m];
    }
#endif
    return nNeg;
}

/* Controller::LUSGS, CFD/Solver/src/Controller.cpp:1878-1961, with
 * LUSGSInitialization (:1732-1782). */
int orc_lusgs(orc_solver *s, double *sweep_normals)
{
    diagonal(s);
    int nS = s->p.n_lusgs_sweeps;
    for (int m = 0; m < 5; ++m)
        for (int c = 0; c < s->nTotal; ++c) {
            s->dflux[m][c] = 0.0;
            s->dq[m][c] = (nS == 1) ? s->res[m][c] : 0.0;
        }
    double init = 0.0, eps = 1.1;
    for (int it = 0; it < nS; ++it) {
        double sn = lusgs_sweep(s, 1);
        set_ghost_dq(s);
        sn += lusgs_sweep(s, 0);
        if (sweep_normals && it < 8) sweep_normals[it] = sn;
        if (it < 8) s->nsSweepNormals[it] = sn;
        s->nsSweepsDone = it + 1;
        if (it == 0) init = sn;
        else eps = sn / init;
        if (eps < s->p.lusgs_tol) break;
    }
    return update_flow(s);
}

/* The sweep pairs the last LU-SGS call made (NS: turb = 0; SA: turb = 1)
 * and the squared changes of the first 8 pairs. */
int orc_sweep_normals(const orc_solver *s, int turb, double *out8)
{
    for (int i = 0; i < 8; ++i) out8[i] = turb ? s->saSweepNormals[i] : s->nsSweepNormals[i];
    return turb ? s->saSweepsDone : s->nsSweepsDone;
}

/* ComputeResidualonGrid (API/Physics/src/Residual.cpp:534-617), unit grid
 * scale: per-equation sums of res^2 over real cells. */
void orc_residual_norm(const orc_solver *s, double *sum5)
{
    for (int m = 0; m < 5; ++m) {
        double acc = 0.0;
        for (int c = 0; c < s->nCell; ++c) {
            double r = s->res[m][c] * 1.0;
            acc += r * r;
        }
        sum5[m] = acc;
    }
}

/* Controller::Relaxation, LU-SGS branch (Controller.cpp:549-608).  Returns
 * the number of cells SolutionFix repaired; rms_hist[i] is the density
 * residual RMS of step first_step+i (ResidualGlobal, Residual.cpp:221-224). */
int orc_iterate(orc_solver *s, int first_step, int n_steps, double *rms_hist)
{
    int nNeg = 0;
    const int hookPhase = s->initPhase;
    for (int i = 0; i < n_steps; ++i) {
        s->initPhase = (first_step + i) <= s->p.flow_init_step;
        orc_residual(s);
        orc_spectrum(s);
        orc_time_step(s, first_step + i);
        if (rms_hist) {
            double sums[5];
            orc_residual_norm(s, sums);
            rms_hist[i] = sqrt(sums[0] / s->nCell);
        }
        nNeg += orc_lusgs(s, NULL);
        if (s->p.viscous == 3) orc_sa_relax(s);
    }
    s->initPhase = hookPhase;
    return nNeg;
}

/* ------------------------------------------------------------------ */
/* Wall forces                                                         */
/* ------------------------------------------------------------------ */

/* CorrectGradient, API/Physics/src/GradientOperation.cpp:21-31. */
static void correct_gradient(double fl, double fr, double *gx, double *gy, double *gz,
                             double dx, double dy, double dz, double ods)
{
    double fold = *gx * dx + *gy * dy + *gz * dz;
    double fnew = (fr - fl) * ods;
    double df = fnew - fold;
    *gx += df * dx;
    *gy += df * dy;
    *gz += df * dz;
}

/* NSSolverUnstruct::AirForceCoefParts over all solid-surface faces
 * (CFD/NSSolver/src/NSSolverUnstruct.cpp:16026-16250, partID = ALL_PART)
 * followed by Force::CompForce (API/Physics/src/Force.cpp:439-538).  The
 * boundary states, temperature, viscosity and gradients are first brought up
 * to date with the current q.  out[ORC_N_FORCE] is laid out as phl_forces. */
void orc_forces(orc_solver *s, double *out)
{
    orc_boundary(s);
    orc_gamma_t(s);
    if (s->p.viscous) orc_viscosity(s);
    if (s->p.gradient == 2) {  /* node values of the current state and its temperature */
        s->tFresh = 0;
        orc_node_values(s);
    }
    orc_gradient(s);

    const double poo = s->primInf[IP];
    const double oRe = 1.0 / s->p.reynolds;
    const double tx0 = s->p.force_ref_x, ty0 = s->p.force_ref_y, tz0 = s->p.force_ref_z;
    double CA_f = 0.0, CA_p = 0.0, CN_f = 0.0, CN_p = 0.0, CZ_f = 0.0, CZ_p = 0.0;
    double Cl_f = 0.0, Cl_p = 0.0, Cn_f = 0.0, Cn_p = 0.0, Cm_f = 0.0, Cm_p = 0.0;
    double cpx = 0.0, cpy = 0.0, cpz = 0.0;

    for (int f = 0; f < s->nBFace; ++f) {
        if (s->bcType[f] != ORC_BC_SOLID_SURFACE) continue;
        int le = s->left[f], re = s->right[f];
        double nx = s->xfn[f] * s->area[f];
        double ny = s->yfn[f] * s->area[f];
        double nz = s->zfn[f] * s->area[f];
        double cp = 2.0 * (s->q[IP][le] - poo);
        double dpx = nx * cp, dpy = ny * cp, dpz = nz * cp;
        CA_p += dpx; CN_p += dpy; CZ_p += dpz;
        double xc = s->xfc[f], yc = s->yfc[f], zc = s->zfc[f];
        if (s->p.viscous) {
            double dudx = s->dqdx[IU][le], dudy = s->dqdy[IU][le], dudz = s->dqdz[IU][le];
            double dvdx = s->dqdx[IV][le], dvdy = s->dqdy[IV][le], dvdz = s->dqdz[IV][le];
            double dwdx = s->dqdx[IW][le], dwdy = s->dqdy[IW][le], dwdz = s->dqdz[IW][le];
            double dx = s->xcc[re] - s->xcc[le];
            double dy = s->ycc[re] - s->ycc[le];
            double dz = s->zcc[re] - s->zcc[le];
            double ods = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
            dx = dx * ods; dy = dy * ods; dz = dz * ods;
            correct_gradient(s->q[IU][le], s->q[IU][re], &dudx, &dudy, &dudz, dx, dy, dz, ods);
            correct_gradient(s->q[IV][le], s->q[IV][re], &dvdx, &dvdy, &dvdz, dx, dy, dz, ods);
            correct_gradient(s->q[IW][le], s->q[IW][re], &dwdx, &dwdy, &dwdz, dx, dy, dz, ods);
            double vis = s->visl[le];
            const double c23 = 2.0 / 3.0;
            double txx = vis * c23 * (2.0 * dudx - dvdy - dwdz);
            double tyy = vis * c23 * (2.0 * dvdy - dwdz - dudx);
            double tzz = vis * c23 * (2.0 * dwdz - dudx - dvdy);
            double txy = vis * (dudy + dvdx);
            double txz = vis * (dudz + dwdx);
            double tyz = vis * (dvdz + dwdy);
            double fvsx = -2.0 * (nx * txx + ny * txy + nz * txz) * oRe;
            double fvsy = -2.0 * (nx * txy + ny * tyy + nz * tyz) * oRe;
            double fvsz = -2.0 * (nx * txz + ny * tyz + nz * tzz) * oRe;
            CA_f += fvsx; CN_f += fvsy; CZ_f += fvsz;
            Cl_f += (yc - ty0) * fvsz - (zc - tz0) * fvsy;
            Cn_f += (zc - tz0) * fvsx - (xc - tx0) * fvsz;
            Cm_f += (xc - tx0) * fvsy - (yc - ty0) * fvsx;
        }
        Cl_p += (yc - ty0) * dpz - (zc - tz0) * dpy;
        Cn_p += (zc - tz0) * dpx - (xc - tx0) * dpz;
        Cm_p += (xc - tx0) * dpy - (yc - ty0) * dpx;
        cpx += dpx; cpy += dpy; cpz += dpz;
    }

    double scf = 1.0 / s->p.force_ref_area;
    double vcm = 1.0 / (s->p.force_ref_area * s->p.force_ref_length);
    double sb = 1.0 / (s->p.force_ref_area * s->p.force_ref_span);
    double raw[15] = {CA_f * scf, CA_p * scf, CN_f * scf, CN_p * scf, CZ_f * scf, CZ_p * scf,
                      cpx * scf, cpy * scf, cpz * scf,
                      Cl_f * sb, Cl_p * sb, Cn_f * sb, Cn_p * sb, Cm_f * vcm, Cm_p * vcm};
    orc_force_coefficients(raw, s->p.alpha_deg, s->p.beta_deg, out);
}

/* Force::CompForce totals (API/Physics/src/Force.cpp:512-533) from the
 * scaled part sums raw[0..14]; out[0..14] = raw, out[15..21] = CL, CD,
 * side force, CD_pr, CD_sf, CD - CL^2/(pi AR) with AR = 9.5, xcp. */
void orc_force_coefficients(const double *raw, double alpha_deg, double beta_deg, double *out)
{
    double attack = alpha_deg * PI_ / 180.0, sideslip = beta_deg * PI_ / 180.0;
    double sina = sin(attack), cosa = cos(attack), sinb = sin(sideslip), cosb = cos(sideslip);
    for (int i = 0; i < 15; ++i) out[i] = raw[i];
    double cfx = raw[0] + raw[1], cfy = raw[2] + raw[3], cfz = raw[4] + raw[5];
    double cmz = raw[13] + raw[14];
    double cl = -sina * cfx + cosa * cfy;
    double cd = cosa * cosb * cfx + sina * cosb * cfy + sinb * cfz;
    double side = -cosa * sinb * cfx - sina * sinb * cfy + cosb * cfz;
    double cdp = cosa * cosb * raw[6] + sina * cosb * raw[7] + sinb * raw[8];
    const double AR = 9.5;
    out[15] = cl;
    out[16] = cd;
    out[17] = side;
    out[18] = cdp;
    out[19] = cd - cdp;
    out[20] = cd - (cl * cl) / (PI_ * AR);
    /* SIGN(one, cfy), Math_BasisFunction.h:150-155 */
    out[21] = (cfy < 0.0 ? -1.0 : 1.0) * cmz / (ABS(cfy) + SMALL);
}

/* ================================================================== */
/* Spalart-Allmaras one-equation model, NSMB formation (viscous == 3)  */
/* CFD/TurbSolver/src/TurbSolverUnstr.cpp, TurbSolver.cpp,             */
/* Param_TurbSolver.cpp; non-conservative variable (rhoFlag = 0).      */
/* ================================================================== */

/* PositiveSA, CFD/TurbSolver/include/TurbSolver.h:103-115 (Allmaras 2012). */
static double positive_sa(double nue, double visl, double rho)
{
    if (nue < 0.0) {
        double xsi = nue / (visl / rho);
        double xsi3 = xsi * xsi * xsi;
        return nue * (16.0 + xsi3) / (16.0 - xsi3);
    }
    return nue;
}

/* A 2-D grid: the Venkatakrishnan scale takes square roots of areas
 * (Limiter.cpp:1122-1131 with GetDim() = 2). */
void orc_set_dim(orc_solver *s, int dim) { s->dim = (dim == 2) ? 2 : 3; }

void orc_set_wall_distance(orc_solver *s, const double *wd)
{
    for (int c = 0; c < s->nCell; ++c) s->wdist[c] = wd[c];
}

/* Cell-centre to wall distance: the minimum over solid-surface faces of the
 * distance to the face, each face split into triangles (face centre, node i,
 * node i+1), the point-triangle distance by closest-point projection
 * (Ericson, Real-Time Collision Detection, 5.1.5).  Restates the geometry of
 * Pre_WalldistCompute::ComputeWallDist3DByProjectingAccurate
 * (API/PreProcess/src/Pre_WalldistCompute.cpp:613-728, 4355). */
static double tri_dist2(const double *p, const double *a, const double *b, const double *c)
{
    double ab[3], ac[3], ap[3], bp[3], cp[3], q[3];
    for (int i = 0; i < 3; ++i) { ab[i] = b[i] - a[i]; ac[i] = c[i] - a[i]; ap[i] = p[i] - a[i]; }
    double d1 = ab[0] * ap[0] + ab[1] * ap[1] + ab[2] * ap[2];
    double d2 = ac[0] * ap[0] + ac[1] * ap[1] + ac[2] * ap[2];
    if (d1 <= 0.0 && d2 <= 0.0) { for (int i = 0; i < 3; ++i) q[i] = a[i]; goto done; }
    for (int i = 0; i < 3; ++i) bp[i] = p[i] - b[i];
    double d3 = ab[0] * bp[0] + ab[1] * bp[1] + ab[2] * bp[2];
    double d4 = ac[0] * bp[0] + ac[1] * bp[1] + ac[2] * bp[2];
    if (d3 >= 0.0 && d4 <= d3) { for (int i = 0; i < 3; ++i) q[i] = b[i]; goto done; }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        double v = d1 / (d1 - d3);
        for (int i = 0; i < 3; ++i) q[i] = a[i] + v * ab[i];
        goto done;
    }
    for (int i = 0; i < 3; ++i) cp[i] = p[i] - c[i];
    double d5 = ab[0] * cp[0] + ab[1] * cp[1] + ab[2] * cp[2];
    double d6 = ac[0] * cp[0] + ac[1] * cp[1] + ac[2] * cp[2];
    if (d6 >= 0.0 && d5 <= d6) { for (int i = 0; i < 3; ++i) q[i] = c[i]; goto done; }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        double w = d2 / (d2 - d6);
        for (int i = 0; i < 3; ++i) q[i] = a[i] + w * ac[i];
        goto done;
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int i = 0; i < 3; ++i) q[i] = b[i] + w * (c[i] - b[i]);
        goto done;
    }
    {
        double den = 1.0 / (va + vb + vc);
        double v = vb * den, w = vc * den;
        for (int i = 0; i < 3; ++i) q[i] = a[i] + ab[i] * v + ac[i] * w;
    }
done:
    {
        double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
        return dx * dx + dy * dy + dz * dz;
    }
}

/* Pre_WalldistCompute::ComputeDistanceOfANodeToSegment
 * (API/PreProcess/src/Pre_WalldistCompute.cpp:4260-4352): squared distance to
 * the edge start-end whose centre is mid, by 30 bisections. */
static double seg_dist2(const double *p, const double *start, const double *mid, const double *end)
{
    double xmin = start[0], ymin = start[1], zmin = start[2];
    double xmax = end[0], ymax = end[1], zmax = end[2];
    double xh = mid[0], yh = mid[1], zh = mid[2];
    double dx = p[0] - xh, dy = p[1] - yh, dz = p[2] - zh;
    double distMid = dx * dx + dy * dy + dz * dz;
    dx = p[0] - xmin; dy = p[1] - ymin; dz = p[2] - zmin;
    double distStart = dx * dx + dy * dy + dz * dz;
    dx = p[0] - xmax; dy = p[1] - ymax; dz = p[2] - zmax;
    double distEnd = dx * dx + dy * dy + dz * dz;
    if (distStart < distMid) return distStart;
    if (distEnd < distMid) return distEnd;
    double distance = distMid;
    for (int iter = 0; iter < 30; ++iter) {
        double x1 = (xmin + xh) * 0.5, y1 = (ymin + yh) * 0.5, z1 = (zmin + zh) * 0.5;
        dx = p[0] - x1; dy = p[1] - y1; dz = p[2] - z1;
        double dist1 = dx * dx + dy * dy + dz * dz;
        double x2 = (xmax + xh) * 0.5, y2 = (ymax + yh) * 0.5, z2 = (zmax + zh) * 0.5;
        dx = p[0] - x2; dy = p[1] - y2; dz = p[2] - z2;
        double dist2 = dx * dx + dy * dy + dz * dz;
        if (distance < dist1 && distance < dist2) {
            xmin = x1; ymin = y1; zmin = z1;
            xmax = x2; ymax = y2; zmax = z2;
        } else if (dist1 < dist2) {
            xmax = xh; ymax = yh; zmax = zh;
            xh = x1; yh = y1; zh = z1;
            distance = dist1;
        } else {
            xmin = xh; ymin = yh; zmin = zh;
            xh = x2; yh = y2; zh = z2;
            distance = dist2;
        }
    }
    return distance;
}

/* ComputeWallDist2DByBinarySearching (Pre_WalldistCompute.cpp:359-507) and
 * PostWallDistance (:3913-3923): every wall edge taken from each of its nodes
 * in turn with the face centre as the first midpoint, edges of squared length
 * <= 1e-16 skipped; the square root of the minimum. */
void orc_wall_distance_2d(int nCell, const double *xcc, const double *ycc, const double *zcc, int nWall,
                          const int *wallFaceNodePtr, const int *wallFaceNodes, const double *wxfc,
                          const double *wyfc, const double *wzfc, const double *x, const double *y,
                          const double *z, double *out)
{
    for (int c = 0; c < nCell; ++c) {
        double p[3] = {xcc[c], ycc[c], zcc[c]};
        double best = LARGE;
        for (int f = 0; f < nWall; ++f) {
            double mid[3] = {wxfc[f], wyfc[f], wzfc[f]};
            int b0 = wallFaceNodePtr[f], n = wallFaceNodePtr[f + 1] - b0;
            for (int i = 0; i < n; ++i) {
                int n1 = wallFaceNodes[b0 + i], n2 = wallFaceNodes[b0 + (i + 1) % n];
                double a[3] = {x[n1], y[n1], z[n1]}, e[3] = {x[n2], y[n2], z[n2]};
                double dx = e[0] - a[0], dy = e[1] - a[1], dz = e[2] - a[2];
                if (dx * dx + dy * dy + dz * dz <= 1.0e-16) continue;
                double d2 = seg_dist2(p, a, mid, e);
                if (best > d2) best = d2;
            }
        }
        out[c] = sqrt(best);
    }
}

void orc_wall_distance(int nCell, const double *xcc, const double *ycc, const double *zcc, int nWall,
                       const int *wallFaceNodePtr, const int *wallFaceNodes, const double *wxfc,
                       const double *wyfc, const double *wzfc, const double *x, const double *y,
                       const double *z, double *out)
{
    for (int c = 0; c < nCell; ++c) {
        double p[3] = {xcc[c], ycc[c], zcc[c]};
        double best = LARGE;
        for (int f = 0; f < nWall; ++f) {
            double a[3] = {wxfc[f], wyfc[f], wzfc[f]};
            int b0 = wallFaceNodePtr[f], n = wallFaceNodePtr[f + 1] - b0;
            for (int i = 0; i < n; ++i) {
                int n1 = wallFaceNodes[b0 + i], n2 = wallFaceNodes[b0 + (i + 1) % n];
                double b[3] = {x[n1], y[n1], z[n1]}, cc[3] = {x[n2], y[n2], z[n2]};
                double d2 = tri_dist2(p, a, b, cc);
                if (d2 < best) best = d2;
            }
        }
        out[c] = sqrt(best);
    }
}

/* TurbSolverUnstr::Boundary (:5919): wall nu~_g = -nu~ (WallBCRegion :6116),
 * farfield by the sign of the Riemann normal velocity (FarfieldBCRegion
 * :6546), symmetry / outflow / extrapolation copy (OutflowBCRegion :6021).
 * The eddy viscosity of the ghost follows the same rule. */
void orc_sa_boundary(orc_solver *s)
{
    const double nuInf = s->p.sa_nu_inf, vistInf = s->p.sa_vist_inf;
    const double roo = s->primInf[IR], uoo = s->primInf[IU], voo = s->primInf[IV], woo = s->primInf[IW];
    const double poo = s->primInf[IP], g = s->p.gamma;
    CELL_LOOP
    for (int f = 0; f < s->nBFace; ++f) {
        int le = s->left[f], re = s->right[f], bc = s->bcType[f];
        if (is_wall(bc)) {
            s->vist[re] = -s->vist[le];
            s->nut[re] = -s->nut[le];
        } else if (bc == ORC_BC_FARFIELD) {
            double rin = s->q[IR][le], uin = s->q[IU][le], vin = s->q[IV][le], win = s->q[IW][le];
            double pin = s->q[IP][le];
            double vno = s->xfn[f] * uoo + s->yfn[f] * voo + s->zfn[f] * woo - 0.0;
            double vni = s->xfn[f] * uin + s->yfn[f] * vin + s->zfn[f] * win - 0.0;
            double vei = sqrt(uin * uin + vin * vin + win * win);
            double coo = sqrt(ABS(g * poo / roo));
            double cin = sqrt(ABS(g * pin / rin));
            int outflow;
            if (vei > cin) {
                outflow = vni >= 0.0;
            } else {
                double riemp = vni + 2.0 * cin / (g - 1.0);
                double riemm = vno - 2.0 * coo / (g - 1.0);
                double vnb = 0.5 * (riemp + riemm);
                outflow = vnb >= 0.0;
            }
            if (outflow) {
                s->vist[re] = s->vist[le];
                s->nut[re] = s->nut[le];
            } else {
                s->vist[re] = vistInf;
                s->nut[re] = nuInf;
            }
        } else {
            s->vist[re] = s->vist[le];
            s->nut[re] = s->nut[le];
        }
    }
}

/* TurbSolverUnstr::Viscosity, "1eq-sa" branch (:6910-6960): real cells. */
void orc_sa_viscosity(orc_solver *s)
{
    const double cv13 = 7.1 * 7.1 * 7.1;
    CELL_LOOP
    for (int c = 0; c < s->nCell; ++c) {
        double nue = s->nut[c];
        if (nue > 0.0) {
            double ld = nue * s->q[IR][c] / (s->visl[c] + SMALL);
            double ld3 = ld * ld * ld;
            double fv1 = ld3 / (ld3 + cv13);
            s->vist[c] = s->q[IR][c] * fv1 * nue;
        } else {
            s->vist[c] = 0.0;
        }
    }
}

/* GG-cell gradient of nu~ (TurbSolverUnstr::GetGradientField :8652 with
 * gradientName "ggcell"). */
/* TurbSolverUnstr::GetGradientField (:8652-8700), gradientName "ggcell".
 * The NSMB formation reads no gradient of nu~, so the other gradient forms
 * (ggnode through ComputeQTurbNodeValue :3480, lsq) are not restated: the
 * field stays zero for them. */
void orc_sa_gradient(orc_solver *s)
{
    if (s->p.gradient != 0) return;
    grad_ggcell(s, s->nut, s->gnx, s->gny, s->gnz);
}

/* InitSpectrum (:1025-1071, steady): spec = 1/(turbCFLScale dt + SMALL). */
static void sa_init_spectrum(orc_solver *s)
{
    CELL_LOOP
    for (int c = 0; c < s->nTotal; ++c) s->specT[c] = 0.0;
    CELL_LOOP
    for (int c = 0; c < s->nCell; ++c) s->specT[c] = 1.0 / (s->p.turb_cfl_scale * s->dt[c] + SMALL);
}

/* SourceFluxOneEquationOriginal (:2702-2916) at one cell, NSMB formation
 * (the diffusion source cb2 |grad nu~|^2 is not added), no transition
 * (ft2 = 0), length scale = wall distance (RANS).  Inputs: nu~, vorticity
 * magnitude, wall distance, density, laminar viscosity, 1/Re.  Outputs:
 * out[0] the source per unit volume, out[1] prde = (production - destruction)
 * / nu~, out[2] its linearisation prdep as the reference forms it, out[3] the
 * implicit diagonal term per unit volume (part1 + part2). */
static void sa_source_cell(double nueSA, double omega, double wdist, double rho, double visl, double oRe,
                           double *out)
{
    const double karm = 0.41, cb1 = 0.1355, cb2 = 0.622, cv1 = 7.1, cw2 = 0.3, cw3 = 2.0;
    const double sigma = 2.0 / 3.0;
    const double kap2 = karm * karm, oKap2 = 1.0 / kap2;
    const double cw1 = cb1 * oKap2 + (1.0 + cb2) / sigma;
    const double cv13 = cv1 * cv1 * cv1, cw36 = pow(cw3, 6), or6 = 1.0 / 6.0, ct3 = 1.2;
    const double saC2 = 0.7, saC3 = 0.9, cw1Kap2 = cw1 * kap2;
    double olam = rho / (visl + SMALL);
    double distance2 = wdist * wdist;
    double oDistance2 = 1.0 / distance2;
    double xsi = nueSA * olam + SMALL;
    double xsi3 = xsi * xsi * xsi;
    double fv1 = xsi3 / (xsi3 + cv13);
    double fv2 = 1.0 - xsi / (1.0 + xsi * fv1);
    double rs = nueSA * oKap2 * oDistance2 * oRe;
    double nuetRs = nueSA * rs;
    double sBar = rs * fv2;
    double std_;
    if (sBar >= -saC2 * omega) std_ = omega + sBar;
    else std_ = omega + omega * (saC2 * saC2 * omega + saC3 * sBar) / ((saC3 - 2 * saC2) * omega - sBar);
    double ostd = 1.0 / (std_ + SMALL);
    double r = rs * ostd;
    r = MIN(r, 10.0);
    double r5 = pow(r, 5);
    double gg = r + cw2 * (r5 * r - r);
    double g6 = pow(gg, 6);
    double fw = gg * pow((1.0 + cw36) / (g6 + cw36), or6);
    double ft2 = 0.0;
    double production = cb1 * (1.0 - ft2) * std_ * nueSA;
    double wallDest = (cw1Kap2 * fw - cb1 * ft2) * nuetRs;
    const double destrAlpha = 50.0;
    if (nueSA < 0.0) {
        production = cb1 * (1.0 - ct3) * omega * nueSA;
        wallDest = -destrAlpha * cw1Kap2 * nuetRs;
    }
    out[0] = production - wallDest;
    double dfv2dk = (3.0 * cv13 * pow(fv1 / (xsi + SMALL), 2) - 1.0) / pow(1.0 + xsi * fv1, 2);
    double dsdnu = oRe * oKap2 * oDistance2 * (fv2 + xsi * dfv2dk);
    double dfwdg = fw / (gg + SMALL) * (1.0 - g6 / (g6 + cw36));
    double dgdr = 1.0 + cw2 * (6.0 * r5 - 1.0);
    double drdnu = oRe * oKap2 * oDistance2 * ostd * (1.0 - nueSA * ostd * dsdnu);
    double dfwdnu = dfwdg * dgdr * drdnu;
    double prod = cb1 * (1.0 - ft2) * std_;
    double dest = (cw1 * fw - cb1 * ft2 * oKap2) * nueSA * oDistance2 * oRe;
    double prodp = cb1 * ((1.0 - ft2) * dsdnu);
    double destp = (cw1 * (fw + dfwdnu * nueSA) - cb1 * ft2 * oKap2) * oDistance2 * oRe;
    if (nueSA < 0.0) {
        prod = cb1 * (1.0 - ct3) * omega;
        dest = -destrAlpha * cw1 * nueSA * oDistance2 * oRe;
        prodp = cb1 * ((1.0 - ct3) * dsdnu);
        destp = -destrAlpha * cw1 * oDistance2 * oRe;
    }
    double prde = prod - dest;
    double prdep = prodp - destp;
    double part1 = -0.5 * (prde - ABS(prde));
    double part2 = -0.5 * (prdep - ABS(prdep)) * nueSA;
    out[1] = prde;
    out[2] = prdep;
    out[3] = part1 + part2;
}

void orc_sa_source_point(double nut, double omega, double wdist, double rho, double visl, double reynolds,
                         double *out)
{
    sa_source_cell(nut, omega, wdist, rho, visl, 1.0 / reynolds, out);
}

static void sa_source(orc_solver *s)
{
    const double oRe = 1.0 / s->p.reynolds;
    CELL_LOOP
    for (int c = 0; c < s->nCell; ++c) {
        double dudy = s->dqdy[IU][c], dudz = s->dqdz[IU][c];
        double dvdx = s->dqdx[IV][c], dvdz = s->dqdz[IV][c];
        double dwdx = s->dqdx[IW][c], dwdy = s->dqdy[IW][c];
        double a1 = dwdy - dvdz, a2 = dudz - dwdx, a3 = dvdx - dudy;
        double omega = sqrt(a1 * a1 + a2 * a2 + a3 * a3);
        double volume = s->vol[c];
        double o[4];
        sa_source_cell(s->nut[c], omega, s->wdist[c], s->q[IR][c], s->visl[c], oRe, o);
        s->resT[c] += (o[0]) * volume;
        s->specT[c] += o[3] * volume;
    }
}

/* InviscidFlux with first-order face values (GetQlQrTurbulence :4978,
 * GetQlQrNS :5065: boundary faces averaged, symmetry copies the left value)
 * and NNDFlux (:5721) in the non-conservative form; LoadFlux (:2205). */
static void sa_inviscid(orc_solver *s)
{
    FACE_LOOP(s, 0, s->nFace, f)
        int le = s->left[f], re = s->right[f];
        double qLv[5], qRv[5];
        for (int m = 0; m < 5; ++m) { qLv[m] = s->q[m][le]; qRv[m] = s->q[m][re]; }
        double tl = s->nut[le], tr = s->nut[re];
        if (f < s->nBFace) {
            for (int m = 0; m < 5; ++m) {
                double t = 0.5 * (qRv[m] + qLv[m]);
                qRv[m] = t; qLv[m] = t;
            }
            if (s->bcType[f] == ORC_BC_SYMMETRY) {
                tr = tl;
            } else {
                double t = 0.5 * (tr + tl);
                tr = t; tl = t;
            }
        }
        double vnL = s->xfn[f] * qLv[IU] + s->yfn[f] * qLv[IV] + s->zfn[f] * qLv[IW] - 0.0;
        double vnR = s->xfn[f] * qRv[IU] + s->yfn[f] * qRv[IV] + s->zfn[f] * qRv[IW] - 0.0;
        double vnLL = 0.5 * (vnL + ABS(vnL));
        double vnRR = 0.5 * (vnR - ABS(vnR));
        double fl = (vnLL * (1.0 * tl) + vnRR * (1.0 * tr)) * s->area[f];
        if (OWN(le)) s->resT[le] -= fl;
        if (f >= s->nBFace && OWN(re)) s->resT[re] += fl;
    FACE_LOOP_END
}

/* ViscousFlux: GetVisFaceValue mlt (:3893-4000, ONE_EQU) and ComputeVisflux
 * NSMB branch (:4438-4670) with turbSkewnessAngle. */
static void sa_viscous(orc_solver *s)
{
    const double cb2 = 0.622, sigma = 2.0 / 3.0, osigma = 1.0 / sigma;
    const double oRe = 1.0 / s->p.reynolds;
    const double turbSkew = s->p.turb_skew_angle;
    FACE_LOOP(s, 0, s->nFace, f)
        int le = s->left[f], re = s->right[f];
        double nueLeft = s->nut[le], visLeft = s->visl[le], rhoLeft = s->q[IR][le];
        double nueRight = s->nut[re], visRight = s->visl[re], rhoRight = s->q[IR][re];
        if (f < s->nBFace) {
            if (nueLeft < 0.0) {
                nueLeft = positive_sa(nueLeft, visLeft, rhoLeft);
                if (is_wall(s->bcType[f])) nueRight = -nueLeft;
            }
        } else {
            nueLeft = positive_sa(nueLeft, visLeft, rhoLeft);
            nueRight = positive_sa(nueRight, visRight, rhoRight);
        }
        double mlt = osigma * 0.5 * (visLeft / rhoLeft + visRight / rhoRight + nueLeft + nueRight);

        double nxs = s->xfn[f], nys = s->yfn[f], nzs = s->zfn[f];
        double dxL = s->xcc[le] - s->xfc[f], dyL = s->ycc[le] - s->yfc[f], dzL = s->zcc[le] - s->zfc[f];
        double dxR = s->xcc[re] - s->xfc[f], dyR = s->ycc[re] - s->yfc[f], dzR = s->zcc[re] - s->zfc[f];
        double dL = nxs * dxL + nys * dyL + nzs * dzL;
        double dR = nxs * dxR + nys * dyR + nzs * dzR;
        double t = -dL / (sqrt(dxL * dxL + dyL * dyL + dzL * dzL) + SMALL);
        t = MIN(t, 1.0); t = MAX(t, -1.0);
        double angle1 = asin(t) * 180.0 / PI_;
        t = dR / (sqrt(dxR * dxR + dyR * dyR + dzR * dzR) + SMALL);
        t = MIN(t, 1.0); t = MAX(t, -1.0);
        double angle2 = asin(t) * 180.0 / PI_;
        double f1 = s->nut[le], f2 = s->nut[re];
        double fMid = 0.5 * (f1 + f2);
        double dfdn = 0.0;
        if (angle1 > turbSkew && angle2 > turbSkew && ABS(dL) > TINY && ABS(dR) > TINY) {
            double dfd1 = (f1 - fMid) / dL;
            double dfd2 = (f2 - fMid) / dR;
            double dtmp = dL * dL + dR * dR;
            double weightL = dL * dL / dtmp;
            double weightR = 1.0 - weightL;
            dfdn = weightL * dfd1 + weightR * dfd2;
            if (f < s->nBFace && s->bcType[f] != ORC_BC_SYMMETRY) {
                double normalDist = (s->xcc[re] - s->xcc[le]) * nxs + (s->ycc[re] - s->ycc[le]) * nys +
                                    (s->zcc[re] - s->zcc[le]) * nzs;
                dfdn = (f2 - f1) / normalDist;
            }
        }
        double flux1 = -oRe * (1.0 + cb2) * mlt * dfdn * s->area[f];
        double flux2 = oRe * cb2 * osigma * dfdn * s->area[f];
        if (OWN(le)) {
            s->resT[le] -= flux1;
            s->resT[le] -= flux2 * (s->visl[le] / s->q[IR][le] + nueLeft);
        }
        if (f >= s->nBFace && OWN(re)) {
            s->resT[re] += flux1;
            s->resT[re] += flux2 * (s->visl[re] / s->q[IR][re] + nueRight);
        }
    FACE_LOOP_END
}

/* SpectrumRadiusOfOneEquation (:4690-4855), NSMB formation: convective
 * upwind Jacobians (rhoFlag = 0: density-weighted) and the diffusion
 * radius, accumulated into the spectrum and the face matrices. */
void orc_sa_spectrum(orc_solver *s)
{
    const double oRe = 1.0 / s->p.reynolds;
    const double sigma = 2.0 / 3.0, osigma = 1.0 / sigma, cb2 = 0.622;
    /* the face matrices are written by the thread that owns the left cell */
    FACE_LOOP(s, 0, s->nFace, f)
        int le = s->left[f], re = s->right[f];
        double rhoL = s->q[IR][le], uL = s->q[IU][le], vL = s->q[IV][le], wL = s->q[IW][le];
        double vnL = s->xfn[f] * uL + s->yfn[f] * vL + s->zfn[f] * wL - 0.0;
        double rcL = 0 * 1.0 + (1 - 0) * rhoL;
        double invJacobianL = vnL > 0.0 ? rcL * vnL * s->area[f] : 0.0;
        double rhoR = s->q[IR][re], uR = s->q[IU][re], vR = s->q[IV][re], wR = s->q[IW][re];
        double vnR = s->xfn[f] * uR + s->yfn[f] * vR + s->zfn[f] * wR - 0.0;
        double rcR = 0 * 1.0 + (1 - 0) * rhoR;
        double invJacobianR = vnR < 0.0 ? rcR * vnR * s->area[f] : 0.0;
        if (OWN(le)) s->specT[le] += invJacobianL;
        if (OWN(re)) s->specT[re] -= invJacobianR;
        if (OWN(le)) {
            s->matR[f] = invJacobianR;
            s->matL[f] = -invJacobianL;
        }
        double orL = 1.0 / (rhoL + SMALL), orR = 1.0 / (rhoR + SMALL);
        double nueL = positive_sa(s->nut[le], s->visl[le], rhoL);
        double nueR = positive_sa(s->nut[re], s->visl[re], rhoR);
        if (f < s->nBFace && is_wall(s->bcType[f])) nueR = -nueL;
        double nueffL = s->visl[le] * orL + ABS(nueL);
        double nueffR = s->visl[re] * orR + ABS(nueR);
        double nueff = 0.5 * (nueffL + nueffR);
        double volume = 0.5 * (s->vol[le] + s->vol[re]);
        double oVolume = 1.0 / volume;
        double ns2 = s->area[f] * s->area[f];
        double term = oRe * oVolume * (1.0 + cb2) * (osigma * nueff * ns2);
        if (OWN(le)) s->specT[le] += term;
        if (OWN(re)) s->specT[re] += term;
        if (OWN(le)) {
            s->matL[f] += -term;
            s->matR[f] += -term;
            s->matL[f] += oRe * oVolume * ns2 * (cb2 * osigma * nueffL);
            s->matR[f] += oRe * oVolume * ns2 * (cb2 * osigma * nueffR);
        }
    FACE_LOOP_END
}

/* TurbSolver::UpdateResiduals (TurbSolver.cpp:316-347) for the steady
 * one-equation case: boundary, spectrum initialisation, source, inviscid
 * and viscous fluxes (mod_turb_res = 0).  LoadResiduals leaves -0.0. */
void orc_sa_update_residuals(orc_solver *s)
{
    for (int c = 0; c < s->nTotal; ++c) s->resT[c] = -0.0;
    orc_sa_boundary(s);
    sa_init_spectrum(s);
    sa_source(s);
    sa_inviscid(s);
    sa_viscous(s);
}

/* Scalar LU-SGS: SolveLUSGSForward (:1400-1519), SetGhostDQLUSGS
 * (:1639-1731: wall 0; farfield, symmetry, inflow, outflow copy; others 0),
 * SolveLUSGSBackward (:1521-1637), UpdateFlowField (:1180-1255, RANS with
 * qTurbLimit = 0: no clipping), in Controller::LUSGS's loop of nLUSGSSweeps
 * sweep pairs (Controller.cpp:1878-1961): dq starts from the residual for one
 * pair and from zero for several, LUplusDQ from zero; the pairs stop once
 * the squared changes of a pair fall below LUSGSTolerance times those of the
 * first. */
/* One cell of the scalar forward (forward = 1) or backward sweep; returns
 * (dq - dq_old)^2. */
static double sa_sweep_cell(orc_solver *s, int c, int forward)
{
    double rhs0 = 0.0;
    const double dqOld = s->dqT[c];
    const double ux = s->dfluxT[c];  /* Ux (forward) or Lx (backward) */
    s->dfluxT[c] = 0.0;
    const double dqc = s->resT[c];
    for (int k = s->c2fPtr[c]; k < s->c2fPtr[c + 1]; ++k) {
        int f = s->c2f[k], le = s->left[f], re = s->right[f];
        if (forward ? (le > c || re > c) : (le < c || re < c)) continue;
        double df;
        if (re == c) { int t = le; le = re; re = t; df = s->matL[f] * s->dqT[re]; }
        else df = s->matR[f] * s->dqT[re];
        rhs0 += df;
    }
    s->dqT[c] = (dqc - ux - rhs0) / s->specT[c];
    s->dfluxT[c] += rhs0;
    const double d = s->dqT[c] - dqOld;
    return d * d;
}

static double sa_sweep(orc_solver *s, int forward)
{
    const int nC = s->nCell;
    double sweep = 0.0;
#ifdef _OPENMP
    struct orc_part *P = orc_partition(s);
    double *d2 = (double *)malloc(sizeof(double) * (size_t)nC);
    const int nl = forward ? P->nLvF : P->nLvB;
    const int *ptr = forward ? P->lvPtrF : P->lvPtrB, *cells = forward ? P->lvCellsF : P->lvCellsB;
    for (int l = 0; l < nl; ++l) {
#pragma omp parallel for schedule(static)
        for (int i = ptr[l]; i < ptr[l + 1]; ++i) d2[cells[i]] = sa_sweep_cell(s, cells[i], forward);
    }
    for (int i = 0; i < nC; ++i) sweep += d2[forward ? i : nC - 1 - i];
    free(d2);
#else
    for (int i = 0; i < nC; ++i) sweep += sa_sweep_cell(s, forward ? i : nC - 1 - i, forward);
#endif
    return sweep;
}

void orc_sa_lusgs(orc_solver *s)
{
    const int nC = s->nCell, nS = s->p.n_lusgs_sweeps;
    CELL_LOOP
    for (int c = 0; c < s->nTotal; ++c) { s->dqT[c] = (nS == 1) ? s->resT[c] : 0.0; s->dfluxT[c] = 0.0; }
    double init = 0.0, eps = 1.1;
    for (int it = 0; it < nS; ++it) {
        double sn = sa_sweep(s, 1);
        for (int f = 0; f < s->nBFace; ++f) {
            int bc = s->bcType[f], le = s->left[f], re = s->right[f];
            if (is_wall(bc)) s->dqT[re] = 0.0;
            else if (bc == ORC_BC_FARFIELD || bc == ORC_BC_SYMMETRY || bc == ORC_BC_INFLOW || bc == ORC_BC_OUTFLOW)
                s->dqT[re] = s->dqT[le];
            else s->dqT[re] = 0.0;
        }
        sn += sa_sweep(s, 0);
        if (it < 8) s->saSweepNormals[it] = sn;
        s->saSweepsDone = it + 1;
        if (it == 0) init = sn;
        else eps = sn / init;
        if (eps < s->p.lusgs_tol) break;
    }
    CELL_LOOP
    for (int c = 0; c < nC; ++c) s->nut[c] += s->dqT[c];
}

/* One relaxation of the turbulence solver after the NS one
 * (Controller::Relaxation with iSolver = TURB): SetBoundaryValueOfField,
 * UpdateIndependentField (eddy viscosity, gradient of nu~), the residual,
 * then LUSGS (Diagonal = SpectrumRadiusOfOneEquation, sweeps, update).
 * The NS gradients, laminar viscosity and time step are those of the NS
 * relaxation of the same step; q is the updated NS state. */
void orc_sa_relax(orc_solver *s)
{
    orc_sa_boundary(s);
    orc_sa_viscosity(s);
    orc_sa_gradient(s);
    orc_sa_update_residuals(s);
    orc_sa_spectrum(s);
    orc_sa_lusgs(s);
}

void orc_sa_face_matrices(const orc_solver *s, double *matL, double *matR)
{
    for (int f = 0; f < s->nFace; ++f) { matL[f] = s->matL[f]; matR[f] = s->matR[f]; }
}
