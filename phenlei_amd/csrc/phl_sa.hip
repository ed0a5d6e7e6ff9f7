// phl_sa.hip -- Spalart-Allmaras one-equation model (NSMB formation) on
// gfx950: wall distance, boundary, eddy viscosity, gradient of nu~, the
// residual (source, first-order upwind convection, diffusion), the spectrum
// with its face matrices, scalar LU-SGS sweeps and the update.
//
// Reference: CFD/TurbSolver/src/TurbSolverUnstr.cpp (Boundary :5919,
// Viscosity :6910, GetGradientField :8652, SourceFluxOneEquationOriginal
// :2702, NNDFlux :5721, GetVisFaceValue :3893, ComputeVisflux :4438,
// SpectrumRadiusOfOneEquation :4690, SolveLUSGSForward/Backward :1400/1521,
// SetGhostDQLUSGS :1639, UpdateFlowField :1180) and TurbSolver.cpp:316.
// Residuals are gathered per cell in the reference's accumulation order:
// source, then the convective faces in face order, then the diffusive ones.
#include "phl_internal.h"

namespace phl {

namespace {

constexpr double kSMALL = 1.0e-40;
constexpr double kTINY = 1.0e-40;
constexpr double kLARGE = 1.0e40;
constexpr double kPI = 3.141592653589793;
constexpr int kBC_WALL = 2, kBC_SYMM = 3, kBC_FAR = 4, kBC_INFLOW = 5, kBC_OUTFLOW = 6;
constexpr int kBlock = 256;

__device__ __forceinline__ double dMIN(double a, double b) { return (a < b) ? a : b; }
__device__ __forceinline__ double dMAX(double a, double b) { return (a > b) ? a : b; }
__device__ __forceinline__ double dABS(double a) { return (a < 0) ? -a : a; }

inline int nblk(size_t n, int b) { return (int)((n + b - 1) / b); }

// Constants of Param_TurbSolver / SourceFluxOneEquationOriginal.
constexpr double kKarm = 0.41, kCb1 = 0.1355, kCb2 = 0.622, kCv1 = 7.1, kCw2 = 0.3, kCw3 = 2.0;
constexpr double kSigma = 2.0 / 3.0;

__device__ __forceinline__ double positive_sa(double nue, double visl, double rho)
{
    if (nue < 0.0) {
        const double xsi = nue / (visl / rho);
        const double xsi3 = xsi * xsi * xsi;
        return nue * (16.0 + xsi3) / (16.0 - xsi3);
    }
    return nue;
}

// ---------------------------------------------------------- wall distance
// Squared distance from p to triangle (a, b, c) by closest-point projection
// (Ericson, Real-Time Collision Detection, 5.1.5).
__device__ double tri_dist2(const double *p, const double *a, const double *b, const double *c)
{
    double ab[3], ac[3], ap[3], q[3];
    for (int i = 0; i < 3; ++i) { ab[i] = b[i] - a[i]; ac[i] = c[i] - a[i]; ap[i] = p[i] - a[i]; }
    const double d1 = ab[0] * ap[0] + ab[1] * ap[1] + ab[2] * ap[2];
    const double d2 = ac[0] * ap[0] + ac[1] * ap[1] + ac[2] * ap[2];
    double bp[3], cp[3];
    for (int i = 0; i < 3; ++i) { bp[i] = p[i] - b[i]; cp[i] = p[i] - c[i]; }
    const double d3 = ab[0] * bp[0] + ab[1] * bp[1] + ab[2] * bp[2];
    const double d4 = ac[0] * bp[0] + ac[1] * bp[1] + ac[2] * bp[2];
    const double d5 = ab[0] * cp[0] + ab[1] * cp[1] + ab[2] * cp[2];
    const double d6 = ac[0] * cp[0] + ac[1] * cp[1] + ac[2] * cp[2];
    const double vc = d1 * d4 - d3 * d2;
    const double vb = d5 * d2 - d1 * d6;
    const double va = d3 * d6 - d5 * d4;
    if (d1 <= 0.0 && d2 <= 0.0) {
        for (int i = 0; i < 3; ++i) q[i] = a[i];
    } else if (d3 >= 0.0 && d4 <= d3) {
        for (int i = 0; i < 3; ++i) q[i] = b[i];
    } else if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        const double v = d1 / (d1 - d3);
        for (int i = 0; i < 3; ++i) q[i] = a[i] + v * ab[i];
    } else if (d6 >= 0.0 && d5 <= d6) {
        for (int i = 0; i < 3; ++i) q[i] = c[i];
    } else if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        const double w = d2 / (d2 - d6);
        for (int i = 0; i < 3; ++i) q[i] = a[i] + w * ac[i];
    } else if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int i = 0; i < 3; ++i) q[i] = b[i] + w * (c[i] - b[i]);
    } else {
        const double den = 1.0 / (va + vb + vc);
        const double v = vb * den, w = vc * den;
        for (int i = 0; i < 3; ++i) q[i] = a[i] + ab[i] * v + ac[i] * w;
    }
    const double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
    return dx * dx + dy * dy + dz * dz;
}

// Squared distance from p to a wall edge of a 2-D grid by the reference's
// bisection (Pre_WalldistCompute::ComputeDistanceOfANodeToSegment,
// API/PreProcess/src/Pre_WalldistCompute.cpp:4260-4352): the nearer end
// point if it beats the edge's centre; otherwise 30 halvings of the interval
// around the current nearest point.  e holds start, centre and end.
__device__ double seg_dist2(const double *p, const double *e)
{
    double lo[3] = {e[0], e[1], e[2]}, mid[3] = {e[3], e[4], e[5]}, hi[3] = {e[6], e[7], e[8]};
    auto d2 = [&](const double *q) {
        const double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
        return dx * dx + dy * dy + dz * dz;
    };
    const double dMid = d2(mid), dStart = d2(lo), dEnd = d2(hi);
    if (dStart < dMid) return dStart;
    if (dEnd < dMid) return dEnd;
    double dist = dMid;
    for (int it = 0; it < 30; ++it) {
        double a[3], b[3];
        for (int i = 0; i < 3; ++i) { a[i] = (lo[i] + mid[i]) * 0.5; b[i] = (hi[i] + mid[i]) * 0.5; }
        const double d1 = d2(a), dd2 = d2(b);
        if (dist < d1 && dist < dd2) {
            for (int i = 0; i < 3; ++i) { lo[i] = a[i]; hi[i] = b[i]; }
        } else if (d1 < dd2) {
            for (int i = 0; i < 3; ++i) { hi[i] = mid[i]; mid[i] = a[i]; }
            dist = d1;
        } else {
            for (int i = 0; i < 3; ++i) { lo[i] = mid[i]; mid[i] = b[i]; }
            dist = dd2;
        }
    }
    return dist;
}

// One thread per cell; the wall elements (9 doubles each: a triangle as face
// centre and two consecutive nodes, or a 2-D edge as start, centre, end) are
// staged through LDS in tiles.
constexpr int kTriTile = 256;
template <bool SEG>
__global__ void __launch_bounds__(kBlock) k_wall_distance(const double *__restrict__ xcc,
                                                          const double *__restrict__ ycc,
                                                          const double *__restrict__ zcc, int nCell,
                                                          const double *__restrict__ tri, int nTri,
                                                          double *__restrict__ wd)
{
    __shared__ double t[kTriTile][9];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    double p[3] = {0.0, 0.0, 0.0};
    if (c < nCell) { p[0] = xcc[c]; p[1] = ycc[c]; p[2] = zcc[c]; }
    double best = kLARGE;
    for (int t0 = 0; t0 < nTri; t0 += kTriTile) {
        __syncthreads();
        for (int e = threadIdx.x; e < kTriTile * 9; e += blockDim.x) {
            const int i = e / 9, j = e % 9;
            t[i][j] = (t0 + i < nTri) ? tri[(size_t)(t0 + i) * 9 + j] : 0.0;
        }
        __syncthreads();
        const int n = (nTri - t0 < kTriTile) ? nTri - t0 : kTriTile;
        for (int i = 0; i < n; ++i) {
            const double d2 = SEG ? seg_dist2(p, &t[i][0]) : tri_dist2(p, &t[i][0], &t[i][3], &t[i][6]);
            if (d2 < best) best = d2;
        }
    }
    if (c < nCell) wd[c] = sqrt(best);
}

// ------------------------------------------------------------- boundary
__global__ void __launch_bounds__(kBlock) k_sa_boundary(DevMesh M, Consts K, const double *__restrict__ q,
                                                        double *__restrict__ nut, double *__restrict__ vist)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= M.nBFace) return;
    const int nT = M.nTotal;
    const int le = M.fle[f], re = M.nCell + f;
    const int bc = M.bcType[f];
    if (bc == kBC_WALL) {
        vist[re] = -vist[le];
        nut[re] = -nut[le];
        return;
    }
    bool copy = true;
    if (bc == kBC_FAR) {
        const FaceRec fr = M.face[f];
        const double *pf = K.primInf;
        const double g = K.gamma;
        const double rin = q[le], uin = q[(size_t)nT + le], vin = q[(size_t)2 * nT + le];
        const double win = q[(size_t)3 * nT + le], pin = q[(size_t)4 * nT + le];
        const double vno = fr.nx * pf[1] + fr.ny * pf[2] + fr.nz * pf[3] - 0.0;
        const double vni = fr.nx * uin + fr.ny * vin + fr.nz * win - 0.0;
        const double vei = sqrt(uin * uin + vin * vin + win * win);
        const double coo = sqrt(dABS(g * pf[4] / pf[0]));
        const double cin = sqrt(dABS(g * pin / rin));
        if (vei > cin) {
            copy = vni >= 0.0;
        } else {
            const double riemp = vni + 2.0 * cin / (g - 1.0);
            const double riemm = vno - 2.0 * coo / (g - 1.0);
            copy = 0.5 * (riemp + riemm) >= 0.0;
        }
    }
    if (copy) {
        vist[re] = vist[le];
        nut[re] = nut[le];
    } else {
        vist[re] = K.saVistInf;
        nut[re] = K.saNuInf;
    }
}

// --------------------------------------------------- eddy viscosity, grad
__global__ void __launch_bounds__(kBlock) k_sa_viscosity(DevMesh M, const double *__restrict__ q,
                                                         const double *__restrict__ visl,
                                                         const double *__restrict__ nut, double *__restrict__ vist)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= M.nCell) return;
    const double cv13 = kCv1 * kCv1 * kCv1;
    const double nue = nut[c];
    double v = 0.0;
    if (nue > 0.0) {
        const double ld = nue * q[c] / (visl[c] + kSMALL);
        const double ld3 = ld * ld * ld;
        const double fv1 = ld3 / (ld3 + cv13);
        v = q[c] * fv1 * nue;
    }
    vist[c] = v;
}

// GG-cell gradient of nu~, faces in increasing order per cell (the
// reference's face loop order), ghost face values from the ghost cells.
__global__ void __launch_bounds__(kBlock) k_sa_gradient(DevMesh M, const double *__restrict__ nut,
                                                        double *__restrict__ gn)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int nC = M.nCell;
    if (c >= nC) return;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    const double vc = nut[c];
    for (int s = 0; s < M.ellW; ++s) {
        const int enc = M.c2fEnc[(size_t)s * nC + c];
        if (enc == 0) break;
        const bool isL = enc > 0;
        const int f = (isL ? enc : -enc) - 1;
        const int nb = M.c2n[(size_t)s * nC + c];
        const FaceRec fr = M.face[f];
        const double nx = fr.nx * fr.area, ny = fr.ny * fr.area, nz = fr.nz * fr.area;
        const double qf = 0.5 * (isL ? (vc + nut[nb]) : (nut[nb] + vc));
        if (isL) { gx += qf * nx; gy += qf * ny; gz += qf * nz; }
        else     { gx -= qf * nx; gy -= qf * ny; gz -= qf * nz; }
    }
    const double ovol = 1.0 / M.vol[c];
    gn[c] = gx * ovol;
    gn[(size_t)nC + c] = gy * ovol;
    gn[(size_t)2 * nC + c] = gz * ovol;
}

// -------------------------------------------------------------- residual
// Per real cell: spectrum initialisation, source (and its implicit part),
// convective and diffusive faces.  GT: storage type of the NS gradients.
// The cell's own part: the source term times the volume starts the residual
// (acc), the time-step term and the source's implicit part start the
// spectrum (spec).
template <typename GT>
__device__ __forceinline__ void sa_source_cell(const DevMesh &M, const Consts &K, const double *__restrict__ q,
                                               const double *__restrict__ visl, const GatherRec *__restrict__ rec,
                                               const double *__restrict__ nut, const double *__restrict__ dt, int c,
                                               double &acc, double &spec)
{
    const double oRe = K.oRe;
    // --- source (SourceFluxOneEquationOriginal, NSMB: no cb2 |grad|^2 term)
    const double kap2 = kKarm * kKarm, oKap2 = 1.0 / kap2;
    const double cw1 = kCb1 * oKap2 + (1.0 + kCb2) / kSigma;
    const double cv13 = kCv1 * kCv1 * kCv1, cw36 = pow(kCw3, 6), or6 = 1.0 / 6.0, ct3 = 1.2;
    const double saC2 = 0.7, saC3 = 0.9, cw1Kap2 = cw1 * kap2;
    const double nueSA = nut[c];
    // velocity gradients of the NS step (d/dx of variable m at [m], d/dy at [5+m], d/dz at [10+m])
    const GT *grad = reinterpret_cast<const GT *>(rec[c].g);
    const double dudy = double(grad[6]), dudz = double(grad[11]);
    const double dvdx = double(grad[2]), dvdz = double(grad[12]);
    const double dwdx = double(grad[3]), dwdy = double(grad[8]);
    const double a1 = dwdy - dvdz, a2 = dudz - dwdx, a3 = dvdx - dudy;
    const double omega = sqrt(a1 * a1 + a2 * a2 + a3 * a3);
    const double volume = M.vol[c];
    const double rhoC = q[c];
    const double olam = rhoC / (visl[c] + kSMALL);
    const double wdc = M.wallDist[c];
    const double distance2 = wdc * wdc;
    const double oDistance2 = 1.0 / distance2;
    const double xsi = nueSA * olam + kSMALL;
    const double xsi3 = xsi * xsi * xsi;
    const double fv1 = xsi3 / (xsi3 + cv13);
    const double fv2 = 1.0 - xsi / (1.0 + xsi * fv1);
    const double rs = nueSA * oKap2 * oDistance2 * oRe;
    const double nuetRs = nueSA * rs;
    const double sBar = rs * fv2;
    double stdv;
    if (sBar >= -saC2 * omega) stdv = omega + sBar;
    else stdv = omega + omega * (saC2 * saC2 * omega + saC3 * sBar) / ((saC3 - 2 * saC2) * omega - sBar);
    const double ostd = 1.0 / (stdv + kSMALL);
    double r = rs * ostd;
    r = dMIN(r, 10.0);
    const double r5 = pow(r, 5);
    const double g = r + kCw2 * (r5 * r - r);
    const double g6 = pow(g, 6);
    const double fw = g * pow((1.0 + cw36) / (g6 + cw36), or6);
    const double ft2 = 0.0;
    double production = kCb1 * (1.0 - ft2) * stdv * nueSA;
    double wallDest = (cw1Kap2 * fw - kCb1 * ft2) * nuetRs;
    const double destrAlpha = 50.0;
    if (nueSA < 0.0) {
        production = kCb1 * (1.0 - ct3) * omega * nueSA;
        wallDest = -destrAlpha * cw1Kap2 * nuetRs;
    }
    const double source = production - wallDest;
    acc = -0.0;
    acc += (source) * volume;
    const double dfv2dk = (3.0 * cv13 * pow(fv1 / (xsi + kSMALL), 2) - 1.0) / pow(1.0 + xsi * fv1, 2);
    const double dsdnu = oRe * oKap2 * oDistance2 * (fv2 + xsi * dfv2dk);
    const double dfwdg = fw / (g + kSMALL) * (1.0 - g6 / (g6 + cw36));
    const double dgdr = 1.0 + kCw2 * (6.0 * r5 - 1.0);
    const double drdnu = oRe * oKap2 * oDistance2 * ostd * (1.0 - nueSA * ostd * dsdnu);
    const double dfwdnu = dfwdg * dgdr * drdnu;
    double prod = kCb1 * (1.0 - ft2) * stdv;
    double dest = (cw1 * fw - kCb1 * ft2 * oKap2) * nueSA * oDistance2 * oRe;
    double prodp = kCb1 * ((1.0 - ft2) * dsdnu);
    double destp = (cw1 * (fw + dfwdnu * nueSA) - kCb1 * ft2 * oKap2) * oDistance2 * oRe;
    if (nueSA < 0.0) {
        prod = kCb1 * (1.0 - ct3) * omega;
        dest = -destrAlpha * cw1 * nueSA * oDistance2 * oRe;
        prodp = kCb1 * ((1.0 - ct3) * dsdnu);
        destp = -destrAlpha * cw1 * oDistance2 * oRe;
    }
    const double prde = prod - dest;
    const double prdep = prodp - destp;
    const double part1 = -0.5 * (prde - dABS(prde));
    const double part2 = -0.5 * (prdep - dABS(prdep)) * nueSA;
    spec = 0.0;
    spec = 1.0 / (K.turbCflScale * dt[c] + kSMALL);
    spec += (part1 + part2) * volume;
}

template <typename GT>
__device__ __forceinline__ void sa_residual_cell(const DevMesh &M, const Consts &K, const double *__restrict__ q,
                                                 const double *__restrict__ visl, const GatherRec *__restrict__ rec,
                                                 const double *__restrict__ nut, const double *__restrict__ dt,
                                                 int c, double &resOut, double &specOut)
{
    const int nC = M.nCell, nT = M.nTotal;
    const double oRe = K.oRe;
    double acc, spec;
    sa_source_cell<GT>(M, K, q, visl, rec, nut, dt, c, acc, spec);

    // --- convection: first-order upwind (NNDFlux, non-conservative form)
    for (int s = 0; s < M.ellW; ++s) {
        const int enc = M.c2fEnc[(size_t)s * nC + c];
        if (enc == 0) break;
        const bool isL = enc > 0;
        const int f = (isL ? enc : -enc) - 1;
        const int nb = M.c2n[(size_t)s * nC + c];
        const int le = isL ? c : nb, re = isL ? nb : c;
        const FaceRec fr = M.face[f];
        double uL = q[(size_t)nT + le], vL = q[(size_t)2 * nT + le], wL = q[(size_t)3 * nT + le];
        double uR = q[(size_t)nT + re], vR = q[(size_t)2 * nT + re], wR = q[(size_t)3 * nT + re];
        double tl = nut[le], tr = nut[re];
        if (f < M.nBFace) {
            double t = 0.5 * (uR + uL); uL = t; uR = t;
            t = 0.5 * (vR + vL); vL = t; vR = t;
            t = 0.5 * (wR + wL); wL = t; wR = t;
            if (M.bcType[f] == kBC_SYMM) {
                tr = tl;
            } else {
                t = 0.5 * (tr + tl);
                tr = t; tl = t;
            }
        }
        const double vnL = fr.nx * uL + fr.ny * vL + fr.nz * wL - 0.0;
        const double vnR = fr.nx * uR + fr.ny * vR + fr.nz * wR - 0.0;
        const double vnLL = 0.5 * (vnL + dABS(vnL));
        const double vnRR = 0.5 * (vnR - dABS(vnR));
        const double fl = (vnLL * (1.0 * tl) + vnRR * (1.0 * tr)) * fr.area;
        if (isL) acc -= fl;
        else acc += fl;
    }

    // --- diffusion: GetVisFaceValue mlt and ComputeVisflux, NSMB branch
    const double osigma = 1.0 / kSigma;
    for (int s = 0; s < M.ellW; ++s) {
        const int enc = M.c2fEnc[(size_t)s * nC + c];
        if (enc == 0) break;
        const bool isL = enc > 0;
        const int f = (isL ? enc : -enc) - 1;
        const int nb = M.c2n[(size_t)s * nC + c];
        const int le = isL ? c : nb, re = isL ? nb : c;
        const FaceRec fr = M.face[f];
        const bool bnd = f < M.nBFace;
        const int bc = bnd ? M.bcType[f] : 0;
        double nueLeft = nut[le], nueRight = nut[re];
        const double visLeft = visl[le], rhoLeft = q[le];
        const double visRight = visl[re], rhoRight = q[re];
        if (bnd) {
            if (nueLeft < 0.0) {
                nueLeft = positive_sa(nueLeft, visLeft, rhoLeft);
                if (bc == kBC_WALL) nueRight = -nueLeft;
            }
        } else {
            nueLeft = positive_sa(nueLeft, visLeft, rhoLeft);
            nueRight = positive_sa(nueRight, visRight, rhoRight);
        }
        const double mlt = osigma * 0.5 * (visLeft / rhoLeft + visRight / rhoRight + nueLeft + nueRight);
        const double nxs = fr.nx, nys = fr.ny, nzs = fr.nz;
        const double dxL = M.xcc[le] - fr.xc, dyL = M.ycc[le] - fr.yc, dzL = M.zcc[le] - fr.zc;
        const double dxR = M.xcc[re] - fr.xc, dyR = M.ycc[re] - fr.yc, dzR = M.zcc[re] - fr.zc;
        const double dL = nxs * dxL + nys * dyL + nzs * dzL;
        const double dR = nxs * dxR + nys * dyR + nzs * dzR;
        double dfdn = 0.0;
        if ((M.visFlag[f] & 4u) && dABS(dL) > kTINY && dABS(dR) > kTINY) {
            const double f1 = nut[le], f2 = nut[re];
            const double fMid = 0.5 * (f1 + f2);
            const double dfd1 = (f1 - fMid) / dL;
            const double dfd2 = (f2 - fMid) / dR;
            const double dtmp = dL * dL + dR * dR;
            const double weightL = dL * dL / dtmp;
            const double weightR = 1.0 - weightL;
            dfdn = weightL * dfd1 + weightR * dfd2;
            if (bnd && bc != kBC_SYMM) {
                const double normalDist = (M.xcc[re] - M.xcc[le]) * nxs + (M.ycc[re] - M.ycc[le]) * nys +
                                          (M.zcc[re] - M.zcc[le]) * nzs;
                dfdn = (f2 - f1) / normalDist;
            }
        }
        const double flux1 = -oRe * (1.0 + kCb2) * mlt * dfdn * fr.area;
        const double flux2 = oRe * kCb2 * osigma * dfdn * fr.area;
        if (isL) {
            acc -= flux1;
            acc -= flux2 * (visl[le] / q[le] + nueLeft);
        } else {
            acc += flux1;
            acc += flux2 * (visl[re] / q[re] + nueRight);
        }
    }
    resOut = acc;
    specOut = spec;
}

template <typename GT>
__global__ void __launch_bounds__(kBlock) k_sa_residual(DevMesh M, Consts K, const double *__restrict__ q,
                                                        const double *__restrict__ visl,
                                                        const GatherRec *__restrict__ grad,
                                                        const double *__restrict__ nut,
                                                        const double *__restrict__ dt,
                                                        double *__restrict__ resT, double *__restrict__ specT)
{
    const int c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (c >= M.nCell) return;
    double r, sp;
    sa_residual_cell<GT>(M, K, q, visl, grad, nut, dt, c, r, sp);
    resT[c] = r;
    specT[c] = sp;
}

// ------------------------------------------------------- spectrum, mats
// SpectrumRadiusOfOneEquation, NSMB: per cell in face order; the owner of a
// face (its left cell) writes the face matrices: mat[2f] is the entry a left
// cell applies to its right neighbour (matrixTurbulenceRight), mat[2f+1] the
// entry a right cell applies to its left neighbour (matrixTurbulenceLeft).
__device__ __forceinline__ void sa_face_spectrum(const DevMesh &M, const Consts &K, const double *__restrict__ q,
                                                 const double *__restrict__ visl, const double *__restrict__ nut,
                                                 int f, int le, int re, double &invL, double &invR, double &term,
                                                 double &coeffL, double &coeffR)
{
    const int nT = M.nTotal;
    const FaceRec fr = M.face[f];
    const double oRe = K.oRe, osigma = 1.0 / kSigma;
    const double rhoL = q[le], uL = q[(size_t)nT + le], vL = q[(size_t)2 * nT + le], wL = q[(size_t)3 * nT + le];
    const double rhoR = q[re], uR = q[(size_t)nT + re], vR = q[(size_t)2 * nT + re], wR = q[(size_t)3 * nT + re];
    const double vnL = fr.nx * uL + fr.ny * vL + fr.nz * wL - 0.0;
    const double vnR = fr.nx * uR + fr.ny * vR + fr.nz * wR - 0.0;
    const double rcL = 0 * 1.0 + (1 - 0) * rhoL, rcR = 0 * 1.0 + (1 - 0) * rhoR;
    invL = vnL > 0.0 ? rcL * vnL * fr.area : 0.0;
    invR = vnR < 0.0 ? rcR * vnR * fr.area : 0.0;
    const double orL = 1.0 / (rhoL + kSMALL), orR = 1.0 / (rhoR + kSMALL);
    const double nueL = positive_sa(nut[le], visl[le], rhoL);
    double nueR = positive_sa(nut[re], visl[re], rhoR);
    if (f < M.nBFace && M.bcType[f] == kBC_WALL) nueR = -nueL;
    const double nueffL = visl[le] * orL + dABS(nueL);
    const double nueffR = visl[re] * orR + dABS(nueR);
    const double nueff = 0.5 * (nueffL + nueffR);
    const double volume = 0.5 * (M.vol[le] + M.vol[re]);
    const double oVolume = 1.0 / volume;
    const double ns2 = fr.area * fr.area;
    term = oRe * oVolume * (1.0 + kCb2) * (osigma * nueff * ns2);
    coeffL = oRe * oVolume * ns2 * (kCb2 * osigma * nueffL);
    coeffR = oRe * oVolume * ns2 * (kCb2 * osigma * nueffR);
}

__device__ __forceinline__ void sa_spectrum_cell(const DevMesh &M, const Consts &K, const double *__restrict__ q,
                                                 const double *__restrict__ visl, const double *__restrict__ nut,
                                                 int c, double &spec, double *__restrict__ mat)
{
    const int nC = M.nCell;
    for (int s = 0; s < M.ellW; ++s) {
        const int enc = M.c2fEnc[(size_t)s * nC + c];
        if (enc == 0) break;
        const bool isL = enc > 0;
        const int f = (isL ? enc : -enc) - 1;
        const int nb = M.c2n[(size_t)s * nC + c];
        const int le = isL ? c : nb, re = isL ? nb : c;
        double invL, invR, term, coeffL, coeffR;
        sa_face_spectrum(M, K, q, visl, nut, f, le, re, invL, invR, term, coeffL, coeffR);
        if (isL) {
            spec += invL;
            spec += term;
            double mR = invR, mL = -invL;
            mL += -term;
            mR += -term;
            mL += coeffL;
            mR += coeffR;
            mat[(size_t)f * 2] = mR;
            mat[(size_t)f * 2 + 1] = mL;
        } else {
            spec -= invR;
            spec += term;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_sa_spectrum(DevMesh M, Consts K, const double *__restrict__ q,
                                                        const double *__restrict__ visl,
                                                        const double *__restrict__ nut, double *__restrict__ specT,
                                                        double *__restrict__ mat)
{
    const int c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (c >= M.nCell) return;
    double spec = specT[c];
    sa_spectrum_cell(M, K, q, visl, nut, c, spec, mat);
    specT[c] = spec;
}

// Residual and spectrum of a cell with at most kSaW faces in one pass over
// its faces: each face's states are loaded once for the convective,
// diffusive and spectral terms.  The sums keep the two-pass order (the
// source, every convective term, then every diffusive term; the spectrum in
// face order), the diffusive terms being held until the convective ones are
// in, so the values are those of sa_residual_cell followed by
// sa_spectrum_cell.
constexpr int kSaW = 6;
template <typename GT>
__device__ __forceinline__ void sa_residual_spectrum_cell(const DevMesh &M, const Consts &K,
                                                          const double *__restrict__ q,
                                                          const double *__restrict__ visl,
                                                          const GatherRec *__restrict__ rec,
                                                          const double *__restrict__ nut,
                                                          const double *__restrict__ dt, int c, double &resOut,
                                                          double &specOut, double *__restrict__ mat)
{
    const int nC = M.nCell, nT = M.nTotal;
    const double oRe = K.oRe, osigma = 1.0 / kSigma;
    double acc, spec;
    sa_source_cell<GT>(M, K, q, visl, rec, nut, dt, c, acc, spec);
    double d1[kSaW], d2[kSaW];
    bool left[kSaW];
    int n = 0;
#pragma unroll
    for (int s = 0; s < kSaW; ++s) {
        if (s >= M.ellW) break;
        const int enc = M.c2fEnc[(size_t)s * nC + c];
        if (enc == 0) break;
        n = s + 1;
        const bool isL = enc > 0;
        left[s] = isL;
        const int f = (isL ? enc : -enc) - 1;
        const int nb = M.c2n[(size_t)s * nC + c];
        const int le = isL ? c : nb, re = isL ? nb : c;
        const FaceRec fr = M.face[f];
        const bool bnd = f < M.nBFace;
        const int bc = bnd ? M.bcType[f] : 0;
        const double rhoL = q[le], uL0 = q[(size_t)nT + le], vL0 = q[(size_t)2 * nT + le],
                     wL0 = q[(size_t)3 * nT + le];
        const double rhoR = q[re], uR0 = q[(size_t)nT + re], vR0 = q[(size_t)2 * nT + re],
                     wR0 = q[(size_t)3 * nT + re];
        const double nutL = nut[le], nutR = nut[re], vislL = visl[le], vislR = visl[re];
        const double volL = M.vol[le], volR = M.vol[re];
        const double xL = M.xcc[le], yL = M.ycc[le], zL = M.zcc[le];
        const double xR = M.xcc[re], yR = M.ycc[re], zR = M.zcc[re];
        const unsigned vflag = M.visFlag[f];

        // convection: first-order upwind (NNDFlux, non-conservative form)
        {
            double uL = uL0, vL = vL0, wL = wL0, uR = uR0, vR = vR0, wR = wR0;
            double tl = nutL, tr = nutR;
            if (bnd) {
                double t = 0.5 * (uR + uL); uL = t; uR = t;
                t = 0.5 * (vR + vL); vL = t; vR = t;
                t = 0.5 * (wR + wL); wL = t; wR = t;
                if (bc == kBC_SYMM) {
                    tr = tl;
                } else {
                    t = 0.5 * (tr + tl);
                    tr = t; tl = t;
                }
            }
            const double vnL = fr.nx * uL + fr.ny * vL + fr.nz * wL - 0.0;
            const double vnR = fr.nx * uR + fr.ny * vR + fr.nz * wR - 0.0;
            const double vnLL = 0.5 * (vnL + dABS(vnL));
            const double vnRR = 0.5 * (vnR - dABS(vnR));
            const double fl = (vnLL * (1.0 * tl) + vnRR * (1.0 * tr)) * fr.area;
            if (isL) acc -= fl;
            else acc += fl;
        }

        // diffusion: GetVisFaceValue mlt and ComputeVisflux, NSMB branch
        {
            double nueLeft = nutL, nueRight = nutR;
            if (bnd) {
                if (nueLeft < 0.0) {
                    nueLeft = positive_sa(nueLeft, vislL, rhoL);
                    if (bc == kBC_WALL) nueRight = -nueLeft;
                }
            } else {
                nueLeft = positive_sa(nueLeft, vislL, rhoL);
                nueRight = positive_sa(nueRight, vislR, rhoR);
            }
            const double mlt = osigma * 0.5 * (vislL / rhoL + vislR / rhoR + nueLeft + nueRight);
            const double nxs = fr.nx, nys = fr.ny, nzs = fr.nz;
            const double dxL = xL - fr.xc, dyL = yL - fr.yc, dzL = zL - fr.zc;
            const double dxR = xR - fr.xc, dyR = yR - fr.yc, dzR = zR - fr.zc;
            const double dL = nxs * dxL + nys * dyL + nzs * dzL;
            const double dR = nxs * dxR + nys * dyR + nzs * dzR;
            double dfdn = 0.0;
            if ((vflag & 4u) && dABS(dL) > kTINY && dABS(dR) > kTINY) {
                const double f1 = nutL, f2 = nutR;
                const double fMid = 0.5 * (f1 + f2);
                const double dfd1 = (f1 - fMid) / dL;
                const double dfd2 = (f2 - fMid) / dR;
                const double dtmp = dL * dL + dR * dR;
                const double weightL = dL * dL / dtmp;
                const double weightR = 1.0 - weightL;
                dfdn = weightL * dfd1 + weightR * dfd2;
                if (bnd && bc != kBC_SYMM) {
                    const double normalDist = (xR - xL) * nxs + (yR - yL) * nys + (zR - zL) * nzs;
                    dfdn = (f2 - f1) / normalDist;
                }
            }
            d1[s] = -oRe * (1.0 + kCb2) * mlt * dfdn * fr.area;
            const double flux2 = oRe * kCb2 * osigma * dfdn * fr.area;
            d2[s] = isL ? flux2 * (vislL / rhoL + nueLeft) : flux2 * (vislR / rhoR + nueRight);
        }

        // spectrum and face matrices (SpectrumRadiusOfOneEquation, NSMB)
        {
            const double vnL = fr.nx * uL0 + fr.ny * vL0 + fr.nz * wL0 - 0.0;
            const double vnR = fr.nx * uR0 + fr.ny * vR0 + fr.nz * wR0 - 0.0;
            const double rcL = 0 * 1.0 + (1 - 0) * rhoL, rcR = 0 * 1.0 + (1 - 0) * rhoR;
            const double invL = vnL > 0.0 ? rcL * vnL * fr.area : 0.0;
            const double invR = vnR < 0.0 ? rcR * vnR * fr.area : 0.0;
            const double orL = 1.0 / (rhoL + kSMALL), orR = 1.0 / (rhoR + kSMALL);
            const double nueL = positive_sa(nutL, vislL, rhoL);
            double nueR = positive_sa(nutR, vislR, rhoR);
            if (bnd && bc == kBC_WALL) nueR = -nueL;
            const double nueffL = vislL * orL + dABS(nueL);
            const double nueffR = vislR * orR + dABS(nueR);
            const double nueff = 0.5 * (nueffL + nueffR);
            const double volume = 0.5 * (volL + volR);
            const double oVolume = 1.0 / volume;
            const double ns2 = fr.area * fr.area;
            const double term = oRe * oVolume * (1.0 + kCb2) * (osigma * nueff * ns2);
            if (isL) {
                const double coeffL = oRe * oVolume * ns2 * (kCb2 * osigma * nueffL);
                const double coeffR = oRe * oVolume * ns2 * (kCb2 * osigma * nueffR);
                spec += invL;
                spec += term;
                double mR = invR, mL = -invL;
                mL += -term;
                mR += -term;
                mL += coeffL;
                mR += coeffR;
                mat[(size_t)f * 2] = mR;
                mat[(size_t)f * 2 + 1] = mL;
            } else {
                spec -= invR;
                spec += term;
            }
        }
    }
#pragma unroll
    for (int s = 0; s < kSaW; ++s) {
        if (s >= n) break;
        if (left[s]) {
            acc -= d1[s];
            acc -= d2[s];
        } else {
            acc += d1[s];
            acc += d2[s];
        }
    }
    resOut = acc;
    specOut = spec;
}

// Residual and spectrum of one relaxation in one kernel: the one-pass form
// for cells of at most kSaW faces, otherwise the residual's passes followed
// by the spectrum's; the values are those of k_sa_residual followed by
// k_sa_spectrum.
template <typename GT>
__global__ void __launch_bounds__(kBlock) k_sa_residual_spectrum(DevMesh M, Consts K, const double *__restrict__ q,
                                                                 const double *__restrict__ visl,
                                                                 const GatherRec *__restrict__ grad,
                                                                 const double *__restrict__ nut,
                                                                 const double *__restrict__ dt,
                                                                 double *__restrict__ resT,
                                                                 double *__restrict__ specT, double *__restrict__ mat)
{
    const int c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (c >= M.nCell) return;
    double r, spec;
    if (M.ellW <= kSaW) {
        sa_residual_spectrum_cell<GT>(M, K, q, visl, grad, nut, dt, c, r, spec, mat);
        resT[c] = r;
        specT[c] = spec;
        return;
    }
    sa_residual_cell<GT>(M, K, q, visl, grad, nut, dt, c, r, spec);
    resT[c] = r;
    sa_spectrum_cell(M, K, q, visl, nut, c, spec, mat);
    specT[c] = spec;
}

// ------------------------------------------------------------ sweeps
// One position of a scalar sweep: the active faces of the cell (rows of the
// sweep order, 2f + side, packed at the front) against the neighbours' dq,
// in face order.  NW > 0: the slot count the lists are padded to, all loads
// issued before the sum; NW == 0: any width, read up to the first unused row.
template <int NW>
__device__ __forceinline__ void sa_sweep_position(const DevMesh &M, const SweepOrder &so, int p, bool useU,
                                                  const double *__restrict__ mat, const double *__restrict__ resT,
                                                  const double *__restrict__ specT, double *dq, double *dflux)
{
    const int P = M.nCell;
    const int c = so.cells[p];
    const double r = resT[c], D = specT[c];
    const double U = useU ? dflux[c] : 0.0;
    double rhs0 = 0.0;
    if constexpr (NW > 0) {
        int row[NW], nb[NW];
#pragma unroll
        for (int s = 0; s < NW; ++s) {
            row[s] = so.row[(size_t)s * P + p];
            nb[s] = so.nb[(size_t)s * P + p];
        }
        double a[NW], d[NW];
#pragma unroll
        for (int s = 0; s < NW; ++s) {
            const bool act = row[s] >= 0;
            a[s] = mat[act ? row[s] : 0];
            d[s] = dq[act ? nb[s] : 0];
        }
        // unused slots add +0.0, which leaves a sum started at +0.0 unchanged
#pragma unroll
        for (int s = 0; s < NW; ++s) rhs0 += (row[s] >= 0) ? a[s] * d[s] : 0.0;
    } else {
        for (int s = 0; s < so.Wpad; ++s) {
            const int rw = so.row[(size_t)s * P + p];
            if (rw < 0) break;
            rhs0 += mat[rw] * dq[so.nb[(size_t)s * P + p]];
        }
    }
    dq[c] = (r - U - rhs0) / D;
    dflux[c] = 0.0 + rhs0;
}

template <int NW>
__global__ void __launch_bounds__(64) k_sa_sweep_level(DevMesh M, SweepOrder so, int begin, int end, bool useU,
                                                      const double *__restrict__ mat,
                                                      const double *__restrict__ resT,
                                                      const double *__restrict__ specT, double *dq, double *dflux)
{
    const int p = begin + xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (p >= end) return;
    sa_sweep_position<NW>(M, so, p, useU, mat, resT, specT, dq, dflux);
}

// Runs of small levels walked by one workgroup, a barrier between levels.
constexpr int kSaBlock = 512;
template <int NW>
__global__ void __launch_bounds__(kSaBlock) k_sa_sweep_block(DevMesh M, SweepOrder so, const int *__restrict__ ptr,
                                                             int l0, int l1, bool useU,
                                                             const double *__restrict__ mat,
                                                             const double *__restrict__ resT,
                                                             const double *__restrict__ specT, double *dq,
                                                             double *dflux)
{
    for (int L = l0; L < l1; ++L) {
        const int b = ptr[L], e = ptr[L + 1];
        for (int p = b + (int)threadIdx.x; p < e; p += blockDim.x)
            sa_sweep_position<NW>(M, so, p, useU, mat, resT, specT, dq, dflux);
        __syncthreads();
    }
}

// SetGhostDQLUSGS: wall 0; farfield, symmetry, inflow, outflow copy; else 0.
// The ghost's value also goes to slot 5 of its NS increment row, where the
// dataflow NS sweeps that carry the SA sweep read a neighbour's nu~ increment.
__global__ void __launch_bounds__(kBlock) k_sa_ghost_dq(DevMesh M, double *__restrict__ dq,
                                                        double *__restrict__ rows)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= M.nBFace) return;
    const int bc = M.bcType[f];
    const bool copy = bc == kBC_FAR || bc == kBC_SYMM || bc == kBC_INFLOW || bc == kBC_OUTFLOW;
    const double v = copy ? dq[M.fle[f]] : 0.0;
    dq[M.nCell + f] = v;
    rows[(size_t)(M.nCell + f) * kRow + 5] = v;
}

__global__ void __launch_bounds__(kBlock) k_sa_update(DevMesh M, double *__restrict__ nut,
                                                      const double *__restrict__ dq)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= M.nCell) return;
    nut[c] += dq[c];
}

// dq = res and LUplusDQ = 0 before the first sweep (LUSGSInitialization).
__global__ void __launch_bounds__(kBlock) k_sa_sweep_init(int nTotal, int nCell, const double *__restrict__ resT,
                                                          double *__restrict__ dq, double *__restrict__ dflux)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nTotal) return;
    dq[i] = resT[i];
    if (i < nCell) dflux[i] = 0.0;
}

template <int NW>
void sa_sweep(const DevMesh &m, DevFlow &f, const Levels &lv, bool useU, hipStream_t s)
{
    const int small = kSaBlock;
    int L = 0;
    while (L < lv.nLevels) {
        const int sz = lv.hptr[L + 1] - lv.hptr[L];
        if (sz <= small) {
            int L1 = L + 1;
            while (L1 < lv.nLevels && lv.hptr[L1 + 1] - lv.hptr[L1] <= small) ++L1;
            hipLaunchKernelGGL(k_sa_sweep_block<NW>, dim3(1), dim3(kSaBlock), 0, s, m, lv.so, lv.ptr, L, L1, useU,
                               f.saMat, f.resT, f.specT, f.dqT, f.dfluxT);
            L = L1;
        } else {
            hipLaunchKernelGGL(k_sa_sweep_level<NW>, dim3(nblk(sz, 64)), dim3(64), 0, s, m, lv.so, lv.hptr[L],
                               lv.hptr[L + 1], useU, f.saMat, f.resT, f.specT, f.dqT, f.dfluxT);
            ++L;
        }
    }
}

void sa_sweep_dispatch(const DevMesh &m, DevFlow &f, const Levels &lv, bool useU, hipStream_t s)
{
    switch (lv.so.Wpad) {
    case 3: sa_sweep<3>(m, f, lv, useU, s); break;
    case 4: sa_sweep<4>(m, f, lv, useU, s); break;
    case 6: sa_sweep<6>(m, f, lv, useU, s); break;
    default: sa_sweep<0>(m, f, lv, useU, s); break;
    }
}

}  // namespace

void launch_wall_distance(const double *xcc, const double *ycc, const double *zcc, int nCell, const double *tri,
                          int nTri, bool edges, double *wd, hipStream_t s)
{
    if (edges)
        hipLaunchKernelGGL(k_wall_distance<true>, dim3(nblk(nCell, kBlock)), dim3(kBlock), 0, s, xcc, ycc, zcc, nCell,
                           tri, nTri, wd);
    else
        hipLaunchKernelGGL(k_wall_distance<false>, dim3(nblk(nCell, kBlock)), dim3(kBlock), 0, s, xcc, ycc, zcc, nCell,
                           tri, nTri, wd);
}

void launch_sa_boundary(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s)
{
    if (m.nBFace == 0) return;
    hipLaunchKernelGGL(k_sa_boundary, dim3(nblk(m.nBFace, kBlock)), dim3(kBlock), 0, s, m, k, f.q, f.nut, f.vist);
}

void launch_sa_viscosity(const DevMesh &m, const Consts &, DevFlow &f, hipStream_t s)
{
    hipLaunchKernelGGL(k_sa_viscosity, dim3(nblk(m.nCell, kBlock)), dim3(kBlock), 0, s, m, f.q, f.visl, f.nut,
                       f.vist);
}

void launch_sa_gradient(const DevMesh &m, const Consts &, DevFlow &f, hipStream_t s)
{
    hipLaunchKernelGGL(k_sa_gradient, dim3(nblk(m.nCell, kBlock)), dim3(kBlock), 0, s, m, f.nut, f.gradNut);
}

void launch_sa_residual(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s)
{
    const int nbc = nblk(m.nCell, kBlock);
    launch_sa_boundary(m, k, f, s);
    if (k.mixed)
        hipLaunchKernelGGL(k_sa_residual<float>, dim3(nbc), dim3(kBlock), 0, s, m, k, f.q, f.visl, f.rec, f.nut,
                           f.dt, f.resT, f.specT);
    else
        hipLaunchKernelGGL(k_sa_residual<double>, dim3(nbc), dim3(kBlock), 0, s, m, k, f.q, f.visl, f.rec, f.nut,
                           f.dt, f.resT, f.specT);
}

void launch_sa_spectrum(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s)
{
    hipLaunchKernelGGL(k_sa_spectrum, dim3(nblk(m.nCell, kBlock)), dim3(kBlock), 0, s, m, k, f.q, f.visl, f.nut,
                       f.specT, f.saMat);
}

void launch_sa_sweep_init(const DevMesh &m, DevFlow &f, hipStream_t s)
{
    hipLaunchKernelGGL(k_sa_sweep_init, dim3(nblk(m.nTotal, kBlock)), dim3(kBlock), 0, s, m.nTotal, m.nCell, f.resT,
                       f.dqT, f.dfluxT);
}

void launch_sa_ghost_dq(const DevMesh &m, DevFlow &f, hipStream_t s)
{
    if (m.nBFace > 0)
        hipLaunchKernelGGL(k_sa_ghost_dq, dim3(nblk(m.nBFace, kBlock)), dim3(kBlock), 0, s, m, f.dqT, f.dq);
}

void launch_sa_update(const DevMesh &m, DevFlow &f, hipStream_t s)
{
    hipLaunchKernelGGL(k_sa_update, dim3(nblk(m.nCell, kBlock)), dim3(kBlock), 0, s, m, f.nut, f.dqT);
}

void launch_sa_lusgs(const DevMesh &m, const Consts &k, DevFlow &f, const Levels &fwd, const Levels &bwd,
                     hipStream_t s, SweepLog *log)
{
    if (k.nSweeps <= 1 || !log) {
        launch_sa_sweep_init(m, f, s);
        sa_sweep_dispatch(m, f, fwd, true, s);
        launch_sa_ghost_dq(m, f, s);
        sa_sweep_dispatch(m, f, bwd, true, s);
        launch_sa_update(m, f, s);
        return;
    }
    // Controller::LUSGS with several pairs (Controller.cpp:1878-1961): dq~
    // and LUplusDQ start from zero; the squared change of dq~ over each pair
    // (both half-sweeps) against the first pair's decides, on the host, after
    // every pair
    // (an error of these calls stays the thread's last HIP error, which the
    // caller's launch check reports)
    (void)hipMemsetAsync(f.dqT, 0, sizeof(double) * m.nTotal, s);
    (void)hipMemsetAsync(f.dfluxT, 0, sizeof(double) * m.nCell, s);
    (void)hipMemsetAsync(f.dqTPrev, 0, sizeof(double) * m.nCell, s);
    double init = 0.0, eps = 1.1;
    log->done = 0;
    for (int it = 0; it < k.nSweeps; ++it) {
        sa_sweep_dispatch(m, f, fwd, true, s);
        launch_sweep_change(f.dqT, 1, 0, 1, m.nCell, f.dqTPrev, f.sweepParts, f.sweepNorm, s);
        launch_sa_ghost_dq(m, f, s);
        sa_sweep_dispatch(m, f, bwd, true, s);
        launch_sweep_change(f.dqT, 1, 0, 1, m.nCell, f.dqTPrev, f.sweepParts, f.sweepNorm + 1, s);
        (void)hipMemcpyAsync(log->host, f.sweepNorm, 2 * sizeof(double), hipMemcpyDeviceToHost, s);
        if (hipStreamSynchronize(s) != hipSuccess) break;  // reported by the caller's launch check
        const double sn = log->host[0] + log->host[1];
        if (it < 8) log->normals[it] = sn;
        log->done = it + 1;
        if (it == 0) init = sn;
        else eps = sn / init;
        if (eps < k.lusgsTol) break;
    }
    launch_sa_update(m, f, s);
}

// One relaxation of the turbulence model after the NS update (the turbulence
// solver's pass through Controller::Relaxation): boundary, eddy viscosity,
// gradient of nu~, the residual (TurbSolver::UpdateResiduals, which applies
// the boundary again), spectrum and face matrices (Diagonal), forward sweep,
// ghost dq, backward sweep, update.
// The gradient of nu~ feeds no term of the NSMB formation, so a relaxation
// leaves it out; the field read (and the standalone hook) computes it.
void launch_sa_pre(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s)
{
    launch_sa_boundary(m, k, f, s);
    launch_sa_viscosity(m, k, f, s);
    launch_sa_boundary(m, k, f, s);
    const int nbc = nblk(m.nCell, kBlock);
    if (k.mixed)
        hipLaunchKernelGGL(k_sa_residual_spectrum<float>, dim3(nbc), dim3(kBlock), 0, s, m, k, f.q, f.visl, f.rec,
                           f.nut, f.dt, f.resT, f.specT, f.saMat);
    else
        hipLaunchKernelGGL(k_sa_residual_spectrum<double>, dim3(nbc), dim3(kBlock), 0, s, m, k, f.q, f.visl, f.rec,
                           f.nut, f.dt, f.resT, f.specT, f.saMat);
}

void launch_sa_relax(const DevMesh &m, const Consts &k, DevFlow &f, const Levels &fwd, const Levels &bwd,
                     hipStream_t s, SweepLog *log)
{
    launch_sa_pre(m, k, f, s);
    launch_sa_lusgs(m, k, f, fwd, bwd, s, log);
}

}  // namespace phl
