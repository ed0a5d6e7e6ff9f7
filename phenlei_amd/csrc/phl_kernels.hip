// phl_kernels.hip -- gfx950 kernels of the residual + LU-SGS path (FP64).
//
// Every face loop of the reference (which scatters face contributions into
// both neighbour cells) is evaluated here as an owner-computes gather: one
// thread per cell walks its faces in increasing face order, recomputes each
// face quantity and adds it with the sign the reference's scatter would use.
// Because a cell receives its contributions in the same order as in the
// reference face loop, the per-cell sums round identically; there are no
// atomics on the residual.  Compile with -ffp-contract=off.
//
// Reference functions restated (CFD/NSSolver/src/NSSolverUnstruct.cpp unless
// noted): Boundary :13404, FarfieldBCRegion :21195, SymmetryBCRegion :19413,
// ViscousAdiabaticWallBCRegion :20988, OutflowBCRegion :19326;
// CompGradientGGCell Geo_UnstructGrid.cpp:3220; VencatLimiter7 Limiter.cpp:1091;
// ReConstructFaceValue :13991; BoundaryQlQrFix :13725; Roe_Scheme_ConservativeForm
// Flux_Inviscid.cpp:136; Flux_RoeEntropyCorrection Flux_RoeEntropyCorrection.cpp:93;
// LoadFlux :6435; ComputeVisflux :12977; SpectrumRadiusInviscid :5749;
// SpectrumRadiusViscous :5852; LocalTimeStep :5999; LimitCFL :6052;
// ReduceMaxTimeStep :5717; Diagonal :15867; SolveLUSGSForward :15118;
// SolveLUSGSBackward :15446; MXDQ_STD Jacobian.cpp:13; SetGhostDQLUSGS :14767;
// UpdateFlowField :14587; SolutionFix :14509.
#include "phl_internal.h"

#include <algorithm>
#include <mutex>

namespace phl {

namespace {

constexpr double kSMALL = 1.0e-40;
constexpr int kBcSymmetry = 3;  // SYMMETRY, Geo_SimpleBC.h:32-39
constexpr double kTINY = 1.0e-40;
constexpr int kBC_EXTRAP = 1, kBC_WALL = 2, kBC_SYMM = 3, kBC_FAR = 4, kBC_OUTFLOW = 6;
constexpr int kBlock = 256;

__device__ __forceinline__ double dMIN(double a, double b) { return (a < b) ? a : b; }
__device__ __forceinline__ double dMAX(double a, double b) { return (a > b) ? a : b; }
__device__ __forceinline__ double dABS(double a) { return (a < 0) ? -a : a; }
template <typename T> __device__ __forceinline__ T tMAX(T a, T b) { return (a > b) ? a : b; }
template <typename T> __device__ __forceinline__ T tMIN(T a, T b) { return (a < b) ? a : b; }
template <typename T> __device__ __forceinline__ T tABS(T a) { return (a < T(0)) ? -a : a; }

// The context scalars as seen from one cell: in a batched context the
// freestream-dependent ones come from the cell's case.
__device__ __forceinline__ Consts case_k(const Consts &K, const DevMesh &M, int cell)
{
    if (K.nCases <= 1) return K;
    Consts k = K;
    const CaseConsts &cc = K.cases[M.caseOf[cell]];
    k.R = cc.R;
    k.mach = cc.mach;
#pragma unroll
    for (int m = 0; m < 5; ++m) k.primInf[m] = cc.primInf[m];
    k.entropyCoef = cc.entropyCoef;
    k.epsTmp = cc.epsTmp;
    k.pLower = cc.pLower;
    k.pUpper = cc.pUpper;
    k.cp = cc.cp;
    k.densLimit = cc.densLimit;
    k.presLimit = cc.presLimit;
    return k;
}


template <typename T>
__device__ __forceinline__ T harten(T eig, T lim)
{
    T e = tABS(eig);
    if (e < lim) e = T(0.5) * (eig * eig + lim * lim) / lim;
    return e;
}

// Van Leer flux-vector splitting (Vanleer_Scheme, API/Physics/src/
// Flux_Inviscid.cpp:12697-12949): perfect gas, no preconditioning, static
// grid; the flux of the first-order initialisation iterations (flowInitStep,
// ComputeInviscidFlux NSSolverUnstruct.cpp:14229-14236).  Per unit area.
template <typename T>
__device__ void vanleer_flux(const T *qL, const T *qR, T xsn, T ysn, T zsn, T gama, T *flux)
{
    const T vb = T(0.0), one = T(1.0);
    const T rL = qL[0], uL = qL[1], vL = qL[2], wL = qL[3], pL = qL[4];
    const T rR = qR[0], uR = qR[1], vR = qR[2], wR = qR[3], pR = qR[4];
    const T v2L = uL * uL + vL * vL + wL * wL;
    const T v2R = uR * uR + vR * vR + wR * wR;
    const T hintL = (gama / (gama - one)) * (pL / rL);
    const T hintR = (gama / (gama - one)) * (pR / rR);
    const T HL = hintL + T(0.5) * v2L, HR = hintR + T(0.5) * v2R;
    const T vnL = xsn * uL + ysn * vL + zsn * wL - vb;
    const T vnR = xsn * uR + ysn * vR + zsn * wR - vb;
    const T rhoVnL = rL * vnL, rhoVnR = rR * vnR;
    const T gp = gama * pL / rL, gq = gama * pR / rR;
    const T cL = sqrt(gp < T(0.0) ? -gp : gp), cR = sqrt(gq < T(0.0) ? -gq : gq);
    const T machL = vnL / cL, machR = vnR / cR;
    T fL[5], fR[5];
    if (machL > one) {
        fL[0] = rhoVnL;
        fL[1] = rhoVnL * uL + xsn * pL;
        fL[2] = rhoVnL * vL + ysn * pL;
        fL[3] = rhoVnL * wL + zsn * pL;
        fL[4] = rhoVnL * HL + vb * pL;
    } else if (machL < -one) {
        for (int m = 0; m < 5; ++m) fL[m] = T(0.0);
    } else {
        const T fmass = T(0.25) * rL * cL * (machL + one) * (machL + one);
        const T tmp = (-vnL + T(2.0) * cL) / gama;
        fL[0] = fmass;
        fL[1] = fmass * (uL + xsn * tmp);
        fL[2] = fmass * (vL + ysn * tmp);
        fL[3] = fmass * (wL + zsn * tmp);
        fL[4] = fmass * (HL + vb * tmp);
    }
    if (machR < -one) {
        fR[0] = rhoVnR;
        fR[1] = rhoVnR * uR + xsn * pR;
        fR[2] = rhoVnR * vR + ysn * pR;
        fR[3] = rhoVnR * wR + zsn * pR;
        fR[4] = rhoVnR * HR + vb * pR;
    } else if (machR > one) {
        for (int m = 0; m < 5; ++m) fR[m] = T(0.0);
    } else {
        const T fmass = -T(0.25) * rR * cR * (machR - one) * (machR - one);
        const T tmp = (-vnR - T(2.0) * cR) / gama;
        fR[0] = fmass;
        fR[1] = fmass * (uR + xsn * tmp);
        fR[2] = fmass * (vR + ysn * tmp);
        fR[3] = fmass * (wR + zsn * tmp);
        fR[4] = fmass * (HR + vb * tmp);
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) flux[m] = fL[m] + fR[m];
}

// Roe flux per unit area, conservative form, static mesh, perfect gas.
// T is double for the FP64 form and float for the mixed form; the double
// instantiation performs exactly the operations of the reference.
template <typename T>
__device__ void roe_flux(const T *qLin, const T *qRin, T xsn, T ysn, T zsn, T gama_, int fixMethod, T ent,
                         T pcL, T pcR, T *flux)
{
    const T LITTLE = T(1.0E-10);
    const T gmL = gama_, gmR = gama_;
    T rL = qLin[0], uL = qLin[1], vL = qLin[2], wL = qLin[3], pL = qLin[4];
    T rR = qRin[0], uR = qRin[1], vR = qRin[2], wR = qRin[3], pR = qRin[4];
    rL = tMAX(rL, LITTLE); rR = tMAX(rR, LITTLE);
    pL = tMAX(pL, LITTLE); pR = tMAX(pR, LITTLE);
    const T v2L = uL * uL + vL * vL + wL * wL;
    const T v2R = uR * uR + vR * vR + wR * wR;
    const T hintL = (gmL / (gmL - T(1.0))) * (pL / rL);
    const T hintR = (gmR / (gmR - T(1.0))) * (pR / rR);
    const T hL = hintL + T(0.5) * v2L;
    const T hR = hintR + T(0.5) * v2R;
    const T vb = T(0.0);
    const T vnL = xsn * uL + ysn * vL + zsn * wL - vb;
    const T vnR = xsn * uR + ysn * vR + zsn * wR - vb;
    const T rvnL = rL * vnL, rvnR = rR * vnR;
    flux[0] = T(0.5) * (rvnL + rvnR);
    flux[1] = T(0.5) * (rvnL * uL + xsn * pL + rvnR * uR + xsn * pR);
    flux[2] = T(0.5) * (rvnL * vL + ysn * pL + rvnR * vR + ysn * pR);
    flux[3] = T(0.5) * (rvnL * wL + zsn * pL + rvnR * wR + zsn * pR);
    flux[4] = T(0.5) * (rvnL * hL + vb * pL + rvnR * hR + vb * pR);

    const T ratio = sqrt(rR / rL);
    const T coef = T(1.0) / (T(1.0) + ratio);
    const T rRoe = sqrt(rL * rR);
    const T uRoe = (uL + uR * ratio) * coef;
    const T vRoe = (vL + vR * ratio) * coef;
    const T wRoe = (wL + wR * ratio) * coef;
    const T pRoe = (pL + pR * ratio) * coef;
    const T gama = (gmL + gmR * ratio) * coef;
    const T gamm1 = gama - T(1.0);
    const T v2Roe = uRoe * uRoe + vRoe * vRoe + wRoe * wRoe;
    const T hRoe = gama / gamm1 * pRoe / rRoe + T(0.5) * v2Roe;
    const T vn = xsn * uRoe + ysn * vRoe + zsn * wRoe - vb;
    const T theta = vn + vb;
    const T c2 = gamm1 * (hRoe - T(0.5) * v2Roe);
    const T cm = sqrt(tABS(c2));
    T e1 = tABS(vn), e2 = tABS(vn + cm), e3 = tABS(vn - cm);
    const T EL = (T(1.0) / (gmL - T(1.0))) * (pL / rL) + T(0.5) * v2L;
    const T ER = (T(1.0) / (gmR - T(1.0))) * (pR / rR) + T(0.5) * v2R;
    const T dq0 = rR - rL;
    const T dq1 = rR * uR - rL * uL;
    const T dq2 = rR * vR - rL * vL;
    const T dq3 = rR * wR - rL * wL;
    const T dq4 = rR * ER - rL * EL;
    if (fixMethod == 1) {
        const T mx = tMAX(e2, e3);
        const T t0 = mx * ent, t1 = mx * ent;
        e1 = tMAX(t0, e1); e2 = tMAX(t1, e2); e3 = tMAX(t1, e3);
    } else if (fixMethod == 2) {
        // the pressure-sensor fix of Lin (JCP 1995): k1 0.25, k2 5, k3 15
        const T ps = tMAX(pcL, pcR);
        const T l1 = (tABS(vn) + cm) * T(15.0) * ps;
        const T l2 = (tABS(vn) + cm) * (T(0.25) + T(5.0) * ps);
        e1 = harten(e1, l1);
        e2 = harten(e2, l2);
        e3 = harten(e3, l2);
    } else if (fixMethod == 4) {
        const T mx = tMAX(e2, e3);
        const T ratio = tMIN(T(1.0), tABS(vn) / cm);
        const T t0 = tMAX(mx * ent * ratio, mx * ent / T(100.0));
        e1 = tMAX(t0, e1); e2 = tMAX(t0, e2); e3 = tMAX(t0, e3);
    } else if (fixMethod == 6) {
        // from the shock sensor, |V| of the Roe average and the sound speed
        const T absVel = sqrt(v2Roe);
        const T lim = (T(1.0) - tMIN(pcL, pcR)) * (absVel + cm) + T(0.0001) * cm;
        e1 = harten(e1, lim);
        e2 = harten(e2, lim);
        e3 = harten(e3, lim);
    } else if (fixMethod == 3) {
        const T me = tABS(vn) + cm;
        const T al = ent * me;
        const T cl = ent * me;
        e1 = harten(e1, cl);
        e2 = harten(e2, al);
        e3 = harten(e3, al);
    } else if (fixMethod == 5) {
        const T mx = tMAX(e2, e3);
        const T d = mx * ent;
        e1 += d; e2 += d; e3 += d;
    }
    const T xi1 = (T(2.0) * e1 - e2 - e3) / (T(2.0) * c2);
    const T xi2 = (e2 - e3) / (T(2.0) * cm);
    const T dc = theta * dq0 - xsn * dq1 - ysn * dq2 - zsn * dq3;
    const T c2dc = c2 * dc;
    const T dh = -gamm1 * (uRoe * dq1 + vRoe * dq2 + wRoe * dq3 - dq4) + T(0.5) * gamm1 * v2Roe * dq0;
    flux[0] -= T(0.5) * (e1 * dq0 - dh * xi1 - dc * xi2);
    flux[1] -= T(0.5) * (e1 * dq1 + (xsn * c2dc - uRoe * dh) * xi1 + (xsn * dh - uRoe * dc) * xi2);
    flux[2] -= T(0.5) * (e1 * dq2 + (ysn * c2dc - vRoe * dh) * xi1 + (ysn * dh - vRoe * dc) * xi2);
    flux[3] -= T(0.5) * (e1 * dq3 + (zsn * c2dc - wRoe * dh) * xi1 + (zsn * dh - wRoe * dc) * xi2);
    flux[4] -= T(0.5) * (e1 * dq4 + (theta * c2dc - hRoe * dh) * xi1 + (theta * dh - hRoe * dc) * xi2);
}

__device__ __forceinline__ double venfun(double x, double y, double eps)
{
    const double x2 = x * x;
    const double xy = x * y;
    const double y2 = y * y;
    return ((x2 + eps + 2.0 * xy) / (x2 + 2.0 * y2 + xy + eps));
}

__device__ __forceinline__ void load5(const double *__restrict__ a, int n, int i, double *v)
{
#pragma unroll
    for (int m = 0; m < 5; ++m) v[m] = a[(size_t)m * n + i];
}

// The first five values of row i of a [n][kRow] array (one 64-byte line).
__device__ __forceinline__ void load_row5(const double *a, int i, double *v)
{
    const double *r = a + (size_t)i * kRow;
    const double2 x = *reinterpret_cast<const double2 *>(r);
    const double2 y = *reinterpret_cast<const double2 *>(r + 2);
    v[0] = x.x; v[1] = x.y; v[2] = y.x; v[3] = y.y; v[4] = r[4];
}

__device__ __forceinline__ void store_row5(double *a, int i, const double *v)
{
    double *r = a + (size_t)i * kRow;
    *reinterpret_cast<double2 *>(r) = make_double2(v[0], v[1]);
    *reinterpret_cast<double2 *>(r + 2) = make_double2(v[2], v[3]);
    r[4] = v[4];
}

// ------------------------------------------------------------------ BCs
__global__ void __launch_bounds__(kBlock) k_boundary(DevMesh M, Consts K0, double *__restrict__ q,
                                                     GatherRec *__restrict__ rec)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= M.nBFace) return;
    const int nT = M.nTotal;
    const int le = M.fle[f], re = M.nCell + f;
    const Consts K = case_k(K0, M, le);
    double ql[5], qr[5];
    load5(q, nT, le, ql);
#pragma unroll
    for (int m = 0; m < 5; ++m) qr[m] = ql[m];
    const FaceRec fr = M.face[f];
    const int bc = M.bcType[f];
    const bool symm = (bc == kBC_SYMM) || (bc == kBC_WALL && !K.viscous);
    if (symm) {
        const double vn = fr.nx * ql[1] + fr.ny * ql[2] + fr.nz * ql[3] - 0.0;
        qr[1] = ql[1] - 2.0 * fr.nx * vn;
        qr[2] = ql[2] - 2.0 * fr.ny * vn;
        qr[3] = ql[3] - 2.0 * fr.nz * vn;
    } else if (bc == kBC_WALL) {
        qr[1] = -ql[1] + 2.0 * 0.0;
        qr[2] = -ql[2] + 2.0 * 0.0;
        qr[3] = -ql[3] + 2.0 * 0.0;
        if (K.twall > 0.0) {
            // isothermal wall (ViscousIsotropicWallBCRegion, NSSolverUnstruct.
            // cpp:20791-20915): the wall pressure, the temperature mirrored
            // about the wall's (at least a tenth of it), the density from both
            const double tL = (1.0 / K.R) * ql[4] / ql[0];
            const double pG = ql[4];
            double tG = 2.0 * K.twall - tL;
            if (tG <= K.twall * 0.1) tG = K.twall * 0.1;
            qr[0] = pG / (K.R * tG * 1.0);
            qr[4] = pG;
        }
    } else if (bc == kBC_FAR) {
        const double *pf = K.primInf;
        const double gl = K.gamma;
        const double rO = pf[0], uO = pf[1], vO = pf[2], wO = pf[3], pO = pf[4];
        const double cO = sqrt(dABS(K.gamma * pO / rO));
        const double rI = ql[0], uI = ql[1], vI = ql[2], wI = ql[3], pI = ql[4];
        const double nx = fr.nx, ny = fr.ny, nz = fr.nz;
        const double vnO = nx * uO + ny * vO + nz * wO - 0.0;
        const double vnI = nx * uI + ny * vI + nz * wI - 0.0;
        const double cI = sqrt(dABS(gl * pI / rI));
        const double gamm1 = gl - 1.0;
        const double velI = sqrt(uI * uI + vI * vI + wI * wI);
        if (velI > cI) {
            if (!(vnI >= 0.0)) {
#pragma unroll
                for (int m = 0; m < 5; ++m) qr[m] = pf[m];
            }
        } else {
            const double Rp = vnI + 2.0 * cI / gamm1;
            const double Rm = vnO - 2.0 * cO / gamm1;
            const double vnB = 0.5 * (Rp + Rm);
            const double cB = 0.25 * (Rp - Rm) * gamm1;
            double S, tx, ty, tz;
            if (vnB >= 0.0) {
                S = pI / pow(rI, gl);
                tx = uI - nx * vnI; ty = vI - ny * vnI; tz = wI - nz * vnI;
            } else {
                S = pO / pow(rO, gl);
                tx = uO - nx * vnO; ty = vO - ny * vnO; tz = wO - nz * vnO;
            }
            const double rB = pow((cB * cB / (S * gl)), 1.0 / gamm1);
            qr[0] = rB;
            qr[1] = tx + nx * vnB;
            qr[2] = ty + ny * vnB;
            qr[3] = tz + nz * vnB;
            qr[4] = cB * cB * rB / gl;
        }
    }
    // extrapolation / outflow: copy (already in qr)
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        q[(size_t)m * nT + re] = qr[m];
        rec[re].q[m] = qr[m];
    }
}

// Standalone gamma/T/viscosity hook (the fused kernels recompute T and mu).
__global__ void __launch_bounds__(kBlock) k_gamma_t(DevMesh M, Consts K0, const double *__restrict__ q,
                                                    double *__restrict__ t, double *__restrict__ visl)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M.nTotal) return;
    const Consts K = case_k(K0, M, i);
    const int nT = M.nTotal;
    const double tt = q[(size_t)4 * nT + i] / (K.R * q[i]);
    t[i] = tt;
    visl[i] = K.viscous ? tt * sqrt(tt) * (1.0 + K.tsuth) / (tt + K.tsuth) : 0.0;
}

// ------------------------------------------------------ LU-SGS operator
// The sweep term of a face, MXDQ_STD(q_nb, n, S, gamma, dq_nb, rad, -1) minus
// the viscous radius times dq_nb, is linear in dq_nb.  Its coefficients depend
// only on the two cell states, so they are evaluated once per face side in
// the (fully parallel) gradient kernel and the level-by-level sweep applies
// them to dq.  The split keeps every operation and its order, so results
// are identical to evaluating MXDQ_STD inside the sweep.
constexpr int kCoef = 16;  // doubles per face side
__device__ void mxdq(const double *prim, double nxs, double nys, double nzs, double area, double gama,
                     const double *dq, double radius, double *f)
{
    const double vgn = 0.0;
    const double rm = prim[0], um = prim[1], vm = prim[2], wm = prim[3], pm = prim[4];
    const double c2 = dABS(gama * pm / rm);
    const double cm = sqrt(c2);
    const double cits = nxs * um + nys * vm + nzs * wm - vgn;
    double lmd1 = cits, lmd2 = cits + cm, lmd3 = cits - cm;
    const double ipnrad = -1.0 * radius / dMAX(kSMALL, area);
    lmd1 = 0.5 * (lmd1 + ipnrad);
    lmd2 = 0.5 * (lmd2 + ipnrad);
    lmd3 = 0.5 * (lmd3 + ipnrad);
    const double x1 = (lmd1 + lmd1 - lmd2 - lmd3) / (c2 + c2);
    const double x2 = (lmd2 - lmd3) / (cm + cm);
    const double dc = (cits + vgn) * dq[0] - nxs * dq[1] - nys * dq[2] - nzs * dq[3];
    const double c2dc = c2 * dc;
    const double v2 = um * um + vm * vm + wm * wm;
    const double gama1 = gama - 1.0;
    const double cc1 = 0.5 * gama1 * v2;
    double dh = -gama1 * (um * dq[1] + vm * dq[2] + wm * dq[3] - dq[4]);
    const double enthalpy = (gama / gama1) * (pm / rm);
    dh += cc1 * dq[0];
    const double hm = enthalpy + 0.5 * v2;
    const double a = dh * x1 + dc * x2;
    const double b = c2dc * x1 + dh * x2;
    f[0] = lmd1 * dq[0] - a;
    f[1] = lmd1 * dq[1] - um * a + nxs * b;
    f[2] = lmd1 * dq[2] - vm * a + nys * b;
    f[3] = lmd1 * dq[3] - wm * a + nzs * b;
    f[4] = lmd1 * dq[4] - hm * a + (cits + vgn) * b;
#pragma unroll
    for (int m = 0; m < 5; ++m) f[m] *= area;
}

__device__ __forceinline__ double lusgs_rad(const double *ql, const double *qr, double nx, double ny,
                                            double nz, double area, double gama)
{
    const double gl = gama, gr = gama;
    const double hl = gl / (gl - 1.0) * ql[4] / ql[0] + 0.5 * (ql[1] * ql[1] + ql[2] * ql[2] + ql[3] * ql[3]);
    const double hr = gr / (gr - 1.0) * qr[4] / qr[0] + 0.5 * (qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]);
    const double t0 = sqrt(qr[0] / ql[0]);
    const double t1 = 1.0 / (1.0 + t0);
    const double rm = sqrt(qr[0] * ql[0]);
    const double um = (ql[1] + qr[1] * t0) * t1;
    const double vm = (ql[2] + qr[2] * t0) * t1;
    const double wm = (ql[3] + qr[3] * t0) * t1;
    const double hm = (hl + hr * t0) * t1;
    const double pm = rm * (hm - 0.5 * (um * um + vm * vm + wm * wm)) * (gl - 1.0) / gl;
    const double vn = nx * um + ny * vm + nz * wm - 0.0;
    const double eig = dABS(vn) + sqrt(gama * pm / rm);
    return eig * area;
}


__device__ __forceinline__ void sweep_coef(const Consts &K, const double *qc, const double *qn, const FaceRec &fr,
                                           bool isL, double dist, double vistC, double vistN, double *co)
{
    double nx = fr.nx, ny = fr.ny, nz = fr.nz;
    if (!isL) { nx = -nx; ny = -ny; nz = -nz; }
    const double gama = K.gamma;
    const double rad = lusgs_rad(qc, qn, nx, ny, nz, fr.area, gama);
    const double vgn = 0.0;
    const double rm = qn[0], um = qn[1], vm = qn[2], wm = qn[3], pm = qn[4];
    const double c2 = dABS(gama * pm / rm);
    const double cm = sqrt(c2);
    const double cits = nx * um + ny * vm + nz * wm - vgn;
    double lmd1 = cits, lmd2 = cits + cm, lmd3 = cits - cm;
    const double ipnrad = -1.0 * rad / dMAX(kSMALL, fr.area);
    lmd1 = 0.5 * (lmd1 + ipnrad);
    lmd2 = 0.5 * (lmd2 + ipnrad);
    lmd3 = 0.5 * (lmd3 + ipnrad);
    const double x1 = (lmd1 + lmd1 - lmd2 - lmd3) / (c2 + c2);
    const double x2 = (lmd2 - lmd3) / (cm + cm);
    const double v2 = um * um + vm * vm + wm * wm;
    const double gama1 = gama - 1.0;
    const double cc1 = 0.5 * gama1 * v2;
    const double enthalpy = (gama / gama1) * (pm / rm);
    const double hm = enthalpy + 0.5 * v2;
    double sr = 0.0;
    if (K.viscous) {
        const double tl = qc[4] / (K.R * qc[0]);
        const double tr = qn[4] / (K.R * qn[0]);
        const double vl = 0.5 * (tl * sqrt(tl) * (1.0 + K.tsuth) / (tl + K.tsuth) +
                                 tr * sqrt(tr) * (1.0 + K.tsuth) / (tr + K.tsuth));
        const double vt = 0.5 * (vistC + vistN);
        const double rho = 0.5 * (qc[0] + qn[0]);
        const double vis = vl + vt;
        sr = 2.0 * vis / (rho * dist * K.reynolds + kSMALL);
        sr *= 0.5 * fr.area;
    }
    co[0] = lmd1; co[1] = x1; co[2] = x2; co[3] = c2; co[4] = cits + vgn;
    co[5] = um; co[6] = vm; co[7] = wm; co[8] = hm; co[9] = cc1;
    co[10] = nx; co[11] = ny; co[12] = nz; co[13] = fr.area; co[14] = sr; co[15] = 0.0;
}

// VIS: 1 viscous, 0 inviscid, -1 read from K (the generic sweep).
template <int VIS = -1>
__device__ __forceinline__ void sweep_apply(const Consts &K, const double *co, const double *dq, double *f)
{
    const double gama1 = K.gamma - 1.0;
    const double dc = co[4] * dq[0] - co[10] * dq[1] - co[11] * dq[2] - co[12] * dq[3];
    const double c2dc = co[3] * dc;
    double dh = -gama1 * (co[5] * dq[1] + co[6] * dq[2] + co[7] * dq[3] - dq[4]);
    dh += co[9] * dq[0];
    const double a = dh * co[1] + dc * co[2];
    const double b = c2dc * co[1] + dh * co[2];
    f[0] = co[0] * dq[0] - a;
    f[1] = co[0] * dq[1] - co[5] * a + co[10] * b;
    f[2] = co[0] * dq[2] - co[6] * a + co[11] * b;
    f[3] = co[0] * dq[3] - co[7] * a + co[12] * b;
    f[4] = co[0] * dq[4] - co[8] * a + co[4] * b;
#pragma unroll
    for (int m = 0; m < 5; ++m) f[m] *= co[13];
    if (VIS == 1 || (VIS < 0 && K.viscous)) {
#pragma unroll
        for (int m = 0; m < 5; ++m) f[m] -= co[14] * dq[m];
    }
}

// ------------------------------------------------ node values (ggnode)
// ComputeNodeValue + ModifyNodeValue (NSSolverUnstruct.cpp:4734-5117), one
// thread per node: the inverse-distance sum of the node's contributors in
// the reference's order (a boundary face contributes the average of its cell
// and ghost, an interior node the real cells around it), divided by the
// weight sum.  t is the temperature array as the last gamma/T pass left it.
__global__ void __launch_bounds__(kBlock) k_node_values(DevMesh M, const double *__restrict__ q,
                                                        const double *__restrict__ t, double *__restrict__ nodeQ)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nN = M.nNode, nT = M.nTotal;
    if (i >= nN) return;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = M.nodeSrcPtr[i]; k < M.nodeSrcPtr[i + 1]; ++k) {
        const int2 src = M.nodeSrc[k];
        const double w = M.nodeW[k];
        if (src.y >= 0) {
#pragma unroll
            for (int m = 0; m < 5; ++m) acc[m] += 0.5 * (q[(size_t)m * nT + src.x] + q[(size_t)m * nT + src.y]) * w;
            acc[5] += 0.5 * (t[src.x] + t[src.y]) * w;
        } else {
#pragma unroll
            for (int m = 0; m < 5; ++m) acc[m] += q[(size_t)m * nT + src.x] * w;
            acc[5] += t[src.x] * w;
        }
    }
    const double ws = M.nodeWsum[i];
#pragma unroll
    for (int m = 0; m < 6; ++m) nodeQ[(size_t)m * nN + i] = ws > kSMALL ? acc[m] / ws : 0.0;
}

// Face means of the node values (CompGradientGGNode_NEW's face value,
// Geo_UnstructGrid.cpp:4238-4250,4268-4273): the node values added in the
// face's node order, divided by the node count.
__global__ void __launch_bounds__(kBlock) k_face_node_mean(DevMesh M, const double *__restrict__ nodeQ,
                                                           double *__restrict__ faceQ)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    const int nF = M.nFace, nN = M.nNode;
    if (f >= nF) return;
    const int b = M.fnPtr[f], e = M.fnPtr[f + 1];
    const double cnt = (double)(e - b);
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        double v = 0.0;
        for (int j = b; j < e; ++j) v += nodeQ[(size_t)m * nN + M.fnList[j]];
        faceQ[(size_t)m * nF + f] = v / cnt;
    }
}

// ------------------------------------------------ gradients + limiter
// GT is the storage type of the gradients: double, or float for the mixed
// form (computed in FP64, stored rounded to FP32).
template <typename GT, int BS>
__global__ void __launch_bounds__(BS) k_grad_lim(DevMesh M, Consts K0, const double *__restrict__ q,
                                                     GT *__restrict__ grad, GT *__restrict__ gradT,
                                                     double *__restrict__ lim, double *__restrict__ coef,
                                                     const double *__restrict__ vist, GatherRec *__restrict__ rec,
                                                     bool gradSoA, const double *__restrict__ faceQ)
{
    const int c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int nC = M.nCell, nT = M.nTotal;
    if (c >= nC) return;
    const Consts K = case_k(K0, M, c);
    double qc[5];
    load5(q, nT, c, qc);
    double gx[5], gy[5], gz[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) { gx[m] = 0.0; gy[m] = 0.0; gz[m] = 0.0; }
    double tgx = 0.0, tgy = 0.0, tgz = 0.0;
    const double tc = qc[4] / (K.R * qc[0]);
    double rmin = qc[0], rmax = qc[0], pmin = qc[4], pmax = qc[4];
    for (int s = 0; s < M.ellW; ++s) {
        const int enc = M.c2fEnc[(size_t)s * nC + c];
        if (enc == 0) break;
        const bool isL = enc > 0;
        const int f = (isL ? enc : -enc) - 1;
        const int nb = M.c2n[(size_t)s * nC + c];
        const FaceRec fr = M.face[f];
        double qn[5];
        load5(q, nT, nb, qn);
        if (K.gradient == 1) {
            // CompGradientLSQ (Geo_UnstructGrid.cpp:4027-4064): w (q_R - q_L) (x_R - x_L)
            // in the face's own orientation, summed in the cell's face order
            const double dx = isL ? M.xcc[nb] - M.xcc[c] : M.xcc[c] - M.xcc[nb];
            const double dy = isL ? M.ycc[nb] - M.ycc[c] : M.ycc[c] - M.ycc[nb];
            const double dz = isL ? M.zcc[nb] - M.zcc[c] : M.zcc[c] - M.zcc[nb];
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                const double dq = fr.lsqW * (isL ? qn[m] - qc[m] : qc[m] - qn[m]);
                gx[m] += dq * dx; gy[m] += dq * dy; gz[m] += dq * dz;
            }
            if (K.viscous) {
                const double tn = qn[4] / (K.R * qn[0]);
                const double dq = fr.lsqW * (isL ? tn - tc : tc - tn);
                tgx += dq * dx; tgy += dq * dy; tgz += dq * dz;
            }
        } else {
            // Green-Gauss: the face value is the average of the two cells
            // (ggcell), or for ggnode the mean of the face's node values on
            // interior and symmetry faces (other boundary faces keep the
            // average of the cell and its ghost)
            const double nx = fr.nx * fr.area, ny = fr.ny * fr.area, nz = fr.nz * fr.area;
            const bool nodeFace = K.gradient == 2 && (f >= M.nBFace || M.bcType[f] == kBcSymmetry);
            const int nF = M.nFace;
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                const double qfc = nodeFace ? faceQ[(size_t)m * nF + f] : 0.5 * (qc[m] + qn[m]);
                if (isL) { gx[m] += qfc * nx; gy[m] += qfc * ny; gz[m] += qfc * nz; }
                else     { gx[m] -= qfc * nx; gy[m] -= qfc * ny; gz[m] -= qfc * nz; }
            }
            if (K.viscous) {
                const double tn = qn[4] / (K.R * qn[0]);
                const double tfc = nodeFace ? faceQ[(size_t)5 * nF + f] : 0.5 * (tc + tn);
                if (isL) { tgx += tfc * nx; tgy += tfc * ny; tgz += tfc * nz; }
                else     { tgx -= tfc * nx; tgy -= tfc * ny; tgz -= tfc * nz; }
            }
        }
        if (nb < nC) {
            rmin = dMIN(rmin, qn[0]); rmax = dMAX(rmax, qn[0]);
            pmin = dMIN(pmin, qn[4]); pmax = dMAX(pmax, qn[4]);
        }
        // LU-SGS coefficients of this face seen from cell c (the dataflow
        // sweeps evaluate them themselves)
        if (!K.coefInSweep) {
            const double dist = K.viscous ? dABS(fr.nx * (M.xcc[nb] - M.xcc[c]) + fr.ny * (M.ycc[nb] - M.ycc[c]) +
                                                 fr.nz * (M.zcc[nb] - M.zcc[c]))
                                          : 0.0;
            double co[kCoef];
            sweep_coef(K, qc, qn, fr, isL, dist, vist[c], vist[nb], co);
            const size_t row = (size_t)f * 2 + (isL ? 0 : 1);
            if (K.coefF32) {
                float4 *dst = reinterpret_cast<float4 *>(reinterpret_cast<float *>(coef) + row * kCoef);
#pragma unroll
                for (int j = 0; j < kCoef / 4; ++j)
                    dst[j] = make_float4(float(co[4 * j]), float(co[4 * j + 1]), float(co[4 * j + 2]),
                                         float(co[4 * j + 3]));
            } else {
                double2 *dst = reinterpret_cast<double2 *>(coef + row * kCoef);
#pragma unroll
                for (int j = 0; j < kCoef / 2; ++j) dst[j] = make_double2(co[2 * j], co[2 * j + 1]);
            }
        }
    }
    const double ovol = 1.0 / M.vol[c];
    double lxx = 0.0, lyy = 0.0, lzz = 0.0, lxy = 0.0, lxz = 0.0, lyz = 0.0;
    if (K.gradient == 1) {
        lxx = M.lsqInv[c]; lyy = M.lsqInv[(size_t)nC + c]; lzz = M.lsqInv[(size_t)2 * nC + c];
        lxy = M.lsqInv[(size_t)3 * nC + c]; lxz = M.lsqInv[(size_t)4 * nC + c]; lyz = M.lsqInv[(size_t)5 * nC + c];
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            const double tx = gx[m], ty = gy[m], tz = gz[m];
            gx[m] = lxx * tx + lxy * ty + lxz * tz;
            gy[m] = lxy * tx + lyy * ty + lyz * tz;
            gz[m] = lxz * tx + lyz * ty + lzz * tz;
        }
    } else if (K.gradient == 2) {  // CompGradientGGNode_NEW divides by the volume
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            gx[m] /= M.vol[c]; gy[m] /= M.vol[c]; gz[m] /= M.vol[c];
        }
    } else {
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            gx[m] *= ovol; gy[m] *= ovol; gz[m] *= ovol;
        }
    }
    if (gradSoA) {  // for the force integration (the iteration reads the gather records)
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            grad[(size_t)m * nC + c] = GT(gx[m]);
            grad[(size_t)(5 + m) * nC + c] = GT(gy[m]);
            grad[(size_t)(10 + m) * nC + c] = GT(gz[m]);
        }
    }
    if (K.viscous) {
        if (K.gradient == 1) {
            const double tx = tgx, ty = tgy, tz = tgz;
            tgx = lxx * tx + lxy * ty + lxz * tz;
            tgy = lxy * tx + lyy * ty + lyz * tz;
            tgz = lxz * tx + lyz * ty + lzz * tz;
        } else if (K.gradient == 2) {
            tgx /= M.vol[c]; tgy /= M.vol[c]; tgz /= M.vol[c];
        } else {
            tgx *= ovol; tgy *= ovol; tgz *= ovol;
        }
        gradT[c] = GT(tgx); gradT[(size_t)nC + c] = GT(tgy); gradT[(size_t)2 * nC + c] = GT(tgz);
    }
    double l;
    if (K.limiter == 1) {
        l = 1.0;
    } else if (K.limiter == 2 || K.initPhase) {  // first order: not computed (ComputeLimiter :18195-18225)
        l = 0.0;
    } else {
        const double dminR = rmin - qc[0], dmaxR = rmax - qc[0];
        const double dminP = pmin - qc[4], dmaxP = pmax - qc[4];
        const double vf = M.venkatVf[c];
        double epsR = K.epsTmp * qc[0] * qc[0];
        epsR *= vf;
        double epsP = K.epsTmp * qc[4] * qc[4];
        epsP *= vf;
        l = 1.0;
        const double xc = M.xcc[c], yc = M.ycc[c], zc = M.zcc[c];
        for (int s = 0; s < M.ellW; ++s) {
            const int enc = M.c2fEnc[(size_t)s * nC + c];
            if (enc == 0) break;
            const int f = (enc > 0 ? enc : -enc) - 1;
            const FaceRec fr = M.face[f];
            const double dx = fr.xc - xc, dy = fr.yc - yc, dz = fr.zc - zc;
            const double dqR = gx[0] * dx + gy[0] * dy + gz[0] * dz;
            if (dqR > kSMALL) l = dMIN(l, venfun(dmaxR, dqR, epsR));
            else if (dqR < -kSMALL) l = dMIN(l, venfun(dminR, dqR, epsR));
            const double dqP = gx[4] * dx + gy[4] * dy + gz[4] * dz;
            if (dqP > kSMALL) l = dMIN(l, venfun(dmaxP, dqP, epsP));
            else if (dqP < -kSMALL) l = dMIN(l, venfun(dminP, dqP, epsP));
        }
        if (M.limZero[c]) l = 0.0;
    }
    lim[c] = l;
    if (rec) {  // the face loop's gather record of this cell (null: gradients for other readers only)
        GatherRec &r = rec[c];
        double2 *dq2 = reinterpret_cast<double2 *>(r.q);
        dq2[0] = make_double2(qc[0], qc[1]);
        dq2[1] = make_double2(qc[2], qc[3]);
        r.q[4] = qc[4];
        GT *g = reinterpret_cast<GT *>(r.g);
#pragma unroll
        for (int m = 0; m < 5; ++m) { g[m] = GT(gx[m]); g[5 + m] = GT(gy[m]); g[10 + m] = GT(gz[m]); }
        r.lim = l;
        if (K.gradient != 0) {  // a ghost takes its owner's gradient (lsq: Geo_UnstructGrid.cpp:4077-4084;
                                // ggnode: :4301-4314); ggcell leaves it zero
            for (int s = 0; s < M.ellW; ++s) {
                const int enc = M.c2fEnc[(size_t)s * nC + c];
                if (enc == 0) break;
                const int nb = M.c2n[(size_t)s * nC + c];
                if (nb < nC) continue;
                GT *gg = reinterpret_cast<GT *>(rec[nb].g);
#pragma unroll
                for (int m = 0; m < 15; ++m) gg[m] = g[m];
            }
        }
    }
}

// ------------------------------------------------- face fluxes (gather)
// One side of a face: the cell state (FP64) and the gradients in the face
// arithmetic type T (double, or float in the mixed form).
template <typename T>
struct CellSide {
    double q[5];
    T g[15];
    double lim, x, y, z;
};

template <typename T>
__device__ __forceinline__ void load_side(const DevMesh &M, const double *__restrict__ q, const T *__restrict__ grad,
                                          const double *__restrict__ lim, int i, bool needGrad, CellSide<T> &s)
{
    const int nT = M.nTotal, nC = M.nCell;
    load5(q, nT, i, s.q);
    s.x = M.xcc[i]; s.y = M.ycc[i]; s.z = M.zcc[i];
    s.lim = lim[i];
    if (needGrad && i < nC) {
#pragma unroll
        for (int m = 0; m < 15; ++m) s.g[m] = grad[(size_t)m * nC + i];
    } else {
#pragma unroll
        for (int m = 0; m < 15; ++m) s.g[m] = T(0.0);
    }
}

// State and centroid of cell or ghost i from its gather record (written for
// the step by k_grad_lim and k_boundary): the record's first 64 bytes.
__device__ __forceinline__ void load_rec_state(const GatherRec *__restrict__ rec, int i, double *qv, double *xv)
{
    const double2 *v = reinterpret_cast<const double2 *>(&rec[i]);
    const double2 a = v[0], b = v[1], c = v[2], d = v[3];
    qv[0] = a.x; qv[1] = a.y; qv[2] = b.x; qv[3] = b.y; qv[4] = c.x;
    xv[0] = c.y; xv[1] = d.x; xv[2] = d.y;
}

template <typename T>
__device__ __forceinline__ void load_side_rec(const GatherRec *__restrict__ rec, int i, bool needGrad, CellSide<T> &s)
{
    const GatherRec &r = rec[i];
    const double2 *v = reinterpret_cast<const double2 *>(&r);
    const double2 a = v[0], b = v[1], c = v[2], d = v[3];
    s.q[0] = a.x; s.q[1] = a.y; s.q[2] = b.x; s.q[3] = b.y; s.q[4] = c.x;
    s.x = c.y; s.y = d.x; s.z = d.y;
    s.lim = r.lim;
    const T *g = reinterpret_cast<const T *>(r.g);
#pragma unroll
    for (int m = 0; m < 15; ++m) s.g[m] = needGrad ? g[m] : T(0.0);
}

// MUSCL face value (reconmeth 1) with the first-order fallback on negative
// density or pressure.  The offset face - centroid is formed in FP64.
template <typename T>
__device__ __forceinline__ void reconstruct(const CellSide<T> &s, double lmin, const FaceRec &fr, T *out)
{
    T qt[5];
    const T dx = T(fr.xc - s.x), dy = T(fr.yc - s.y), dz = T(fr.zc - s.z);
    const T lm = T(lmin);
#pragma unroll
    for (int m = 0; m < 5; ++m) qt[m] = T(s.q[m]);
#pragma unroll
    for (int m = 0; m < 5; ++m) qt[m] += lm * (s.g[m] * dx + s.g[5 + m] * dy + s.g[10 + m] * dz);
    const bool ok = !(qt[0] <= T(0.0) || qt[4] <= T(0.0));
#pragma unroll
    for (int m = 0; m < 5; ++m) out[m] = ok ? qt[m] : T(s.q[m]);
}

// Viscous face flux (flux = -(1/Re) * A * fvis), one face, both sides given.
// Temperatures, geometry and the two-point differences along the normal are
// formed in FP64 (they are small differences of O(1) values); the Sutherland
// viscosity, the face-gradient assembly, the stresses and the flux are in T.
template <typename T>
__device__ __forceinline__ T sutherland(T t, T tsuth)
{
    return t * sqrt(t) * (T(1.0) + tsuth) / (t + tsuth);
}

template <typename T>
__device__ void viscous_face(const DevMesh &M, const Consts &K, int f, const FaceRec &fr, const CellSide<T> &L,
                             const CellSide<T> &R, const T *gTL, const T *gTR, bool bnd, bool wall, double vistL,
                             double vistR, double *fl)
{
    const T two3rd = T(2.0) / T(3.0);
    const double tLc = L.q[4] / (K.R * L.q[0]);
    const double tRc = R.q[4] / (K.R * R.q[0]);
    const T vlL = sutherland(T(tLc), T(K.tsuth));
    const T vlR = sutherland(T(tRc), T(K.tsuth));
    // FaceWeight (Geo_UnstructGrid.cpp:2420)
    const double dxL0 = fr.xc - L.x, dyL0 = fr.yc - L.y, dzL0 = fr.zc - L.z;
    const double dxR0 = fr.xc - R.x, dyR0 = fr.yc - R.y, dzR0 = fr.zc - R.z;
    const double d1 = sqrt(dxL0 * dxL0 + dyL0 * dyL0 + dzL0 * dzL0);
    const double d2 = sqrt(dxR0 * dxR0 + dyR0 * dyR0 + dzR0 * dzR0);
    const double dd = 1.0 / (d1 + d2 + kSMALL);
    double wL = d2 * dd, wR = d1 * dd;
    if (bnd) { wL = 1.0; wR = 0.0; }
    // GetVisFaceValue (:7126)
    const T mul = T(0.5) * (vlL + vlR);
    const T mut = T(0.5) * (T(vistL) + T(vistR));
    double primF[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) primF[m] = 0.5 * (L.q[m] + R.q[m]);
    const double tmF = 0.5 * (tLc + tRc);
    const T kcp = (mul * T(K.oprl) + mut * T(K.oprt)) * T(K.cp);
    if (wall) { primF[1] = 0.0; primF[2] = 0.0; primF[3] = 0.0; }
    // ComputeVisflux (:12977)
    const double nxs = fr.nx, nys = fr.ny, nzs = fr.nz;
    double t1x, t1y, t1z;
    if (dABS(nxs) > kSMALL) { t1x = nys; t1y = -nxs; t1z = 0.0; }
    else if (dABS(nys) > kSMALL) { t1x = -nys; t1y = nxs; t1z = 0.0; }
    else if (dABS(nzs) > kSMALL) { t1x = 0.0; t1y = -nzs; t1z = nys; }
    else {
#pragma unroll
        for (int m = 0; m < 5; ++m) fl[m] = 0.0;
        return;
    }
    const double on = 1.0 / sqrt(t1x * t1x + t1y * t1y + t1z * t1z);
    t1x *= on; t1y *= on; t1z *= on;
    const double t2x = nys * t1z - nzs * t1y;
    const double t2y = nzs * t1x - nxs * t1z;
    const double t2z = nxs * t1y - nys * t1x;
    const double dxL = L.x - fr.xc, dyL = L.y - fr.yc, dzL = L.z - fr.zc;
    const double dxR = R.x - fr.xc, dyR = R.y - fr.yc, dzR = R.z - fr.zc;
    const double dL = nxs * dxL + nys * dyL + nzs * dzL;
    const double dR = nxs * dxR + nys * dyR + nzs * dzR;
    const double dxnL = nxs * dL - dxL, dynL = nys * dL - dyL, dznL = nzs * dL - dzL;
    const double dxnR = nxs * dR - dxR, dynR = nys * dR - dyR, dznR = nzs * dR - dzR;
    const unsigned flag = M.visFlag[f];
    double f1[5], f2[5], fM[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) { f1[m] = L.q[m]; f2[m] = R.q[m]; }
    double tL = tLc, tR = tRc;
#pragma unroll
    for (int m = 0; m < 5; ++m) fM[m] = 0.5 * (f1[m] + f2[m]);
    double tM = 0.5 * (tL + tR);
    if (wall) {
        tM = tmF;
        fM[0] = fM[4] / (K.R * tM);
    }
    if (flag & 1u) {
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            f1[m] += double(L.g[m]) * dxnL + double(L.g[5 + m]) * dynL + double(L.g[10 + m]) * dznL;
            f2[m] += double(R.g[m]) * dxnR + double(R.g[5 + m]) * dynR + double(R.g[10 + m]) * dznR;
        }
        tL += double(gTL[0]) * dxnL + double(gTL[1]) * dynL + double(gTL[2]) * dznL;
        tR += double(gTR[0]) * dxnR + double(gTR[1]) * dynR + double(gTR[2]) * dznR;
        if (tL < kSMALL) tL = tLc;
        if (tR < kSMALL) tR = tRc;
#pragma unroll
        for (int m = 0; m < 5; ++m) fM[m] = primF[m];
        tM = tmF;
    }
    double dfdn[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double dtdn = 0.0;
    if ((flag & 2u) && dABS(dL) > kTINY && dABS(dR) > kTINY) {
        double dfd1[5], dfd2[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) dfd1[m] = (f1[m] - fM[m]) / dL;
#pragma unroll
        for (int m = 0; m < 5; ++m) dfd2[m] = (f2[m] - fM[m]) / dR;
        const double dtd1 = (tL - tM) / dL;
        const double dtd2 = (tR - tM) / dR;
        const double dtmp = dL * dL + dR * dR;
        double wl = dL * dL / dtmp;
        double wr = 1.0 - wl;
        if (bnd) { wl = 1.0; wr = 0.0; }
#pragma unroll
        for (int m = 0; m < 5; ++m) dfdn[m] = wl * dfd1[m] + wr * dfd2[m];
        dtdn = dtd1 * wl + dtd2 * wr;
    }
    const T tWL = T(wL), tWR = T(wR);
    const T n1 = T(nxs), n2 = T(nys), n3 = T(nzs);
    const T a1 = T(t1x), a2 = T(t1y), a3 = T(t1z), b1 = T(t2x), b2 = T(t2y), b3 = T(t2z);
    T dfdx[5], dfdy[5], dfdz[5], dft1[5], dft2[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        dfdx[m] = tWL * L.g[m] + tWR * R.g[m];
        dfdy[m] = tWL * L.g[5 + m] + tWR * R.g[5 + m];
        dfdz[m] = tWL * L.g[10 + m] + tWR * R.g[10 + m];
    }
#pragma unroll
    for (int m = 1; m <= 3; ++m) {
        dft1[m] = a1 * dfdx[m] + a2 * dfdy[m] + a3 * dfdz[m];
        dft2[m] = b1 * dfdx[m] + b2 * dfdy[m] + b3 * dfdz[m];
    }
#pragma unroll
    for (int m = 1; m <= 3; ++m) {
        const T dn = T(dfdn[m]);
        dfdx[m] = n1 * dn + a1 * dft1[m] + b1 * dft2[m];
        dfdy[m] = n2 * dn + a2 * dft1[m] + b2 * dft2[m];
        dfdz[m] = n3 * dn + a3 * dft1[m] + b3 * dft2[m];
    }
    const T dudx = dfdx[1], dudy = dfdy[1], dudz = dfdz[1];
    const T dvdx = dfdx[2], dvdy = dfdy[2], dvdz = dfdz[2];
    const T dwdx = dfdx[3], dwdy = dfdy[3], dwdz = dfdz[3];
    T qNorm = T(0.0);
    qNorm += kcp * T(dtdn);
    const T divv2p3 = two3rd * (dudx + dvdy + dwdz);
    const T vis = mul + mut;
    const T um = T(fM[1]), vm = T(fM[2]), wm = T(fM[3]);
    const T txx = vis * (T(2.0) * dudx - divv2p3);
    const T tyy = vis * (T(2.0) * dvdy - divv2p3);
    const T tzz = vis * (T(2.0) * dwdz - divv2p3);
    const T txy = vis * (dudy + dvdx);
    const T txz = vis * (dudz + dwdx);
    const T tyz = vis * (dvdz + dwdy);
    T fv[5];
    fv[0] = T(0.0);
    fv[1] = n1 * txx + n2 * txy + n3 * txz;
    fv[2] = n1 * txy + n2 * tyy + n3 * tyz;
    fv[3] = n1 * txz + n2 * tyz + n3 * tzz;
    fv[4] = um * fv[1] + vm * fv[2] + wm * fv[3] + qNorm;
    const T oRe = T(K.oRe), area = T(fr.area);
#pragma unroll
    for (int m = 0; m < 5; ++m) fl[m] = double(-oRe * area * fv[m]);
}

// One thread per cell: inviscid (and viscous) residual, spectral radii and
// the per-block partial sums of res^2.  SPEC_ONLY skips the fluxes.
// LocalTimeStep + LimitCFL for one cell: CFL*vol/spectral radius, CFL lowered
// in low-pressure cells and scaled by the precomputed skewness factor.
__device__ __forceinline__ double local_dt(const DevMesh &M, const Consts &K, int c, double p, double spec,
                                           double cfl)
{
    double cc = cfl;
    if (p < K.pLower) cc = K.cflStart;
    else if (p < K.pUpper) cc = (p - K.pUpper) * (cfl - K.cflStart) / (K.pUpper - K.pLower) + cfl;
    cc *= M.cflSkew[c];
    return cc * M.vol[c] / spec;
}

// One thread per cell: inviscid (and viscous) residual, spectral radii, and
// per-block partials of res^2 and, with FUSED_DT, of the local time step
// minimum (CFL of the current step read from cflArr[*stepCtr]).
// T is the face arithmetic type: double, or float in the mixed form (FP32
// gradients, reconstruction, Roe and viscous fluxes; FP64 accumulation).
template <bool SPEC_ONLY, bool FUSED_DT, typename T, int BS>
__global__ void __launch_bounds__(BS) k_flux(DevMesh M, Consts K0, const double *__restrict__ q,
                                                 const T *__restrict__ grad,
                                                 const T *__restrict__ gradT,
                                                 const double *__restrict__ lim, double *__restrict__ res,
                                                 double *__restrict__ specInv, double *__restrict__ specVis,
                                                 double *__restrict__ normPart, const double *__restrict__ cflArr,
                                                 const int *__restrict__ stepCtr, double *__restrict__ dtLoc,
                                                 double *__restrict__ dtMinPart, const double *__restrict__ vist,
                                                 const GatherRec *__restrict__ rec)
{
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int c = lb * blockDim.x + threadIdx.x;
    const int nC = M.nCell, nT = M.nTotal;
    (void)nT;
    double acc[5] = {-0.0, -0.0, -0.0, -0.0, -0.0};
    double dtl = 1.0e300;
    if (c < nC) {
        const Consts K = case_k(K0, M, c);
        const bool needG = !SPEC_ONLY && (K.limiter != 2 || K.viscous);
        CellSide<T> own;
        // the spectrum hook reads the SoA state directly (records are
        // refreshed by the boundary and gradient kernels of a residual)
        if (SPEC_ONLY) load_side(M, q, grad, lim, c, needG, own);
        else load_side_rec(rec, c, needG, own);
        double sInv = 0.0;
        for (int s = 0; s < M.ellW; ++s) {
            const int enc = M.c2fEnc[(size_t)s * nC + c];
            if (enc == 0) break;
            const bool isL = enc > 0;
            const int f = (isL ? enc : -enc) - 1;
            const int nb = M.c2n[(size_t)s * nC + c];
            const FaceRec fr = M.face[f];
            CellSide<T> oth;
            if (SPEC_ONLY) load_side(M, q, grad, lim, nb, false, oth);
            else load_side_rec(rec, nb, K.limiter != 2, oth);
            const CellSide<T> &L = isL ? own : oth;
            const CellSide<T> &R = isL ? oth : own;
            if (!SPEC_ONLY) {
                T qL[5], qR[5];
                if (K.limiter != 2 && !K.initPhase) {
                    const double lmin = dMIN(L.lim, R.lim);
                    reconstruct(L, lmin, fr, qL);
                    reconstruct(R, lmin, fr, qR);
                } else {
#pragma unroll
                    for (int m = 0; m < 5; ++m) { qL[m] = T(L.q[m]); qR[m] = T(R.q[m]); }
                }
                const T nx = T(fr.nx), ny = T(fr.ny), nz = T(fr.nz);
                if (f < M.nBFace) {
                    const int bc = M.bcType[f];
                    if (bc == kBC_WALL || bc == kBC_SYMM) {
#pragma unroll
                        for (int m = 0; m < 5; ++m) {
                            const T tmp = T(0.5 * (L.q[m] + R.q[m]));
                            qL[m] = tmp; qR[m] = tmp;
                        }
                        if (bc == kBC_SYMM) {
                            const T vn = T(2.0) * (nx * qL[1] + ny * qL[2] + nz * qL[3] - T(0.0));
                            qR[0] = qL[0];
                            qR[1] = qL[1] - vn * nx;
                            qR[2] = qL[2] - vn * ny;
                            qR[3] = qL[3] - vn * nz;
                            qR[4] = qL[4];
                        }
                        if (bc == kBC_WALL && K.viscous) {
                            qL[1] = T(0.0); qL[2] = T(0.0); qL[3] = T(0.0);
                            qR[1] = T(0.0); qR[2] = T(0.0); qR[3] = T(0.0);
                        }
                    }
                }
                T fl[5];
                if (K.initPhase) vanleer_flux<T>(qL, qR, nx, ny, nz, T(K.gamma), fl);
                else roe_flux<T>(qL, qR, nx, ny, nz, T(K.gamma), K.entropyFix, T(K.entropyCoef),
                                 K.rtem ? T(K.rtem[isL ? c : nb]) : T(0.0), K.rtem ? T(K.rtem[isL ? nb : c]) : T(0.0),
                                 fl);
                const T area = T(fr.area);
#pragma unroll
                for (int m = 0; m < 5; ++m) fl[m] *= area;
                if (isL) {
#pragma unroll
                    for (int m = 0; m < 5; ++m) acc[m] -= double(fl[m]);
                } else {
#pragma unroll
                    for (int m = 0; m < 5; ++m) acc[m] += double(fl[m]);
                }
            }
            // inviscid spectral radius at the face-averaged state
            const double vnl = fr.nx * L.q[1] + fr.ny * L.q[2] + fr.nz * L.q[3] - 0.0;
            const double vnr = fr.nx * R.q[1] + fr.ny * R.q[2] + fr.nz * R.q[3] - 0.0;
            const double g = 0.5 * (K.gamma + K.gamma);
            const double pm = 0.5 * (L.q[4] + R.q[4]);
            const double rm = 0.5 * (L.q[0] + R.q[0]);
            const double vn = 0.5 * (vnl + vnr);
            const double cm = sqrt(g * pm / rm);
            sInv += 0.5 * fr.area * (dABS(vn) + cm);
        }
        specInv[c] = sInv;
        double sVis = 0.0;
        if (K.viscous) {
            for (int s = 0; s < M.ellW; ++s) {
                const int enc = M.c2fEnc[(size_t)s * nC + c];
                if (enc == 0) break;
                const bool isL = enc > 0;
                const int f = (isL ? enc : -enc) - 1;
                const int nb = M.c2n[(size_t)s * nC + c];
                const FaceRec fr = M.face[f];
                CellSide<T> oth;
                if (SPEC_ONLY) load_side(M, q, grad, lim, nb, false, oth);
                else load_side_rec(rec, nb, true, oth);
                const CellSide<T> &L = isL ? own : oth;
                const CellSide<T> &R = isL ? oth : own;
                const bool bnd = f < M.nBFace;
                if (!SPEC_ONLY) {
                    const int li = isL ? c : nb, ri = isL ? nb : c;
                    T gTL[3] = {T(0.0), T(0.0), T(0.0)}, gTR[3] = {T(0.0), T(0.0), T(0.0)};
                    if (li < nC) { gTL[0] = gradT[li]; gTL[1] = gradT[(size_t)nC + li]; gTL[2] = gradT[(size_t)2 * nC + li]; }
                    if (ri < nC) { gTR[0] = gradT[ri]; gTR[1] = gradT[(size_t)nC + ri]; gTR[2] = gradT[(size_t)2 * nC + ri]; }
                    else if (K.gradient != 0) { gTR[0] = gTL[0]; gTR[1] = gTL[1]; gTR[2] = gTL[2]; }  // ghost: owner's
                    const bool wall = bnd && M.bcType[f] == kBC_WALL;
                    double fl[5];
                    viscous_face(M, K, f, fr, L, R, gTL, gTR, bnd, wall, vist[li], vist[ri], fl);
                    if (isL) {
#pragma unroll
                        for (int m = 0; m < 5; ++m) acc[m] -= fl[m];
                    } else {
#pragma unroll
                        for (int m = 0; m < 5; ++m) acc[m] += fl[m];
                    }
                }
                const double tl = L.q[4] / (K.R * L.q[0]);
                const double tr = R.q[4] / (K.R * R.q[0]);
                const double vl = 0.5 * (tl * sqrt(tl) * (1.0 + K.tsuth) / (tl + K.tsuth) +
                                         tr * sqrt(tr) * (1.0 + K.tsuth) / (tr + K.tsuth));
                const double vt = 0.5 * (vist[isL ? c : nb] + vist[isL ? nb : c]);
                const double dx = R.x - L.x, dy = R.y - L.y, dz = R.z - L.z;
                const double ds = dABS(fr.nx * dx + fr.ny * dy + fr.nz * dz);
                const double rho = 0.5 * (L.q[0] + R.q[0]);
                const double vis = vl + vt;
                double sr = 2.0 * vis / (rho * ds * K.reynolds + kSMALL);
                sr *= 0.5 * fr.area;
                sVis += sr;
            }
            specVis[c] = sVis;
        }
        if (FUSED_DT) {
            double spec = sInv;
            if (K.viscous) spec += sVis;
            dtl = local_dt(M, K, c, own.q[4], spec, cflArr[*stepCtr]);
            dtLoc[c] = dtl;
        }
        if (!SPEC_ONLY) {
#pragma unroll
            for (int m = 0; m < 5; ++m) res[(size_t)m * nC + c] = acc[m];
        }
    }
    if (SPEC_ONLY) return;
    // deterministic block partials: sum of res^2 and min of the local dt
    __shared__ double red[6][BS];
#pragma unroll
    for (int m = 0; m < 5; ++m) red[m][threadIdx.x] = (c < nC) ? acc[m] * acc[m] : 0.0;
    red[5][threadIdx.x] = dtl;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int m = 0; m < 5; ++m) red[m][threadIdx.x] += red[m][threadIdx.x + w];
            red[5][threadIdx.x] = dMIN(red[5][threadIdx.x], red[5][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int m = 0; m < 5; ++m) normPart[(size_t)lb * 5 + m] = red[m][0];
        if (FUSED_DT) dtMinPart[lb] = red[5][0];
    }
}

// Face-parallel form of the face loop: one thread per face evaluates the
// face's inviscid flux and spectral-radius term (and the viscous ones) once,
// from the two cells' gather records, and stores them; k_cell_gather then
// sums them per cell.  The per-face values are exactly what each owner
// computes in k_flux (they depend only on the face's left/right states), and
// the gather adds them in the same order, so the residuals are identical
// while every interior face is evaluated once instead of twice.
__device__ __forceinline__ void store6(double *dst, const double *v)
{
    double2 *d = reinterpret_cast<double2 *>(dst);
    d[0] = make_double2(v[0], v[1]);
    d[1] = make_double2(v[2], v[3]);
    d[2] = make_double2(v[4], v[5]);
}

// Compact face value (mixed form, 32 bytes): the five flux components are
// FP32 results widened to FP64, so they are stored as FP32 without loss; the
// spectral-radius term is FP64.
struct alignas(32) FaceValF32 {
    float f[6];  // f[5] unused
    double s;
};
__device__ __forceinline__ void store_compact(double *dst, const double *v)
{
    FaceValF32 *d = reinterpret_cast<FaceValF32 *>(dst);
    float4 *a = reinterpret_cast<float4 *>(d->f);
    a[0] = make_float4(float(v[0]), float(v[1]), float(v[2]), float(v[3]));
    reinterpret_cast<double2 *>(d)[1] = make_double2(__longlong_as_double((long long)__float_as_uint(float(v[4]))),
                                                     v[5]);
}
__device__ __forceinline__ void load_compact(const double *src, double *fl, double &spec)
{
    const float4 a = reinterpret_cast<const float4 *>(src)[0];
    const double2 b = reinterpret_cast<const double2 *>(src)[1];
    fl[0] = double(a.x); fl[1] = double(a.y); fl[2] = double(a.z); fl[3] = double(a.w);
    fl[4] = double(__uint_as_float((unsigned)__double_as_longlong(b.x)));
    spec = b.y;
}

// The stored values of face f: invOut[0..4] the area-scaled inviscid flux,
// invOut[5] the inviscid spectral-radius term; visOut[] likewise (viscous
// flow only).  Each group is stored as soon as it is known, as six doubles
// or (COMPACT, FP32 faces) as a FaceValF32.  PARTS selects the groups
// evaluated: bit 0 inviscid (InviscidFlux), bit 1 viscous (ViscousFlux).
template <typename T, bool COMPACT = false, int PARTS = 3>
__device__ __forceinline__ void face_values(const DevMesh &M, const Consts &K0, const GatherRec *__restrict__ rec,
                                            const T *__restrict__ gradT, const double *__restrict__ vist, int f,
                                            double *invOut, double *visOut)
{
    const int nC = M.nCell;
    const int le = M.fle[f], re = M.fre[f];
    const Consts K = case_k(K0, M, le);
    const FaceRec fr = M.face[f];
    const bool needG = K.limiter != 2 || K.viscous;
    CellSide<T> L, R;
    load_side_rec(rec, le, needG, L);
    load_side_rec(rec, re, needG, R);
    const bool bnd = f < M.nBFace;
    const int bc = bnd ? M.bcType[f] : 0;
    if constexpr ((PARTS & 1) != 0) {
        T qL[5], qR[5];
        if (K.limiter != 2 && !K.initPhase) {
            const double lmin = dMIN(L.lim, R.lim);
            reconstruct(L, lmin, fr, qL);
            reconstruct(R, lmin, fr, qR);
        } else {
#pragma unroll
            for (int m = 0; m < 5; ++m) { qL[m] = T(L.q[m]); qR[m] = T(R.q[m]); }
        }
        const T nx = T(fr.nx), ny = T(fr.ny), nz = T(fr.nz);
        if (bnd && (bc == kBC_WALL || bc == kBC_SYMM)) {
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                const T tmp = T(0.5 * (L.q[m] + R.q[m]));
                qL[m] = tmp; qR[m] = tmp;
            }
            if (bc == kBC_SYMM) {
                const T vn = T(2.0) * (nx * qL[1] + ny * qL[2] + nz * qL[3] - T(0.0));
                qR[0] = qL[0];
                qR[1] = qL[1] - vn * nx;
                qR[2] = qL[2] - vn * ny;
                qR[3] = qL[3] - vn * nz;
                qR[4] = qL[4];
            }
            if (bc == kBC_WALL && K.viscous) {
                qL[1] = T(0.0); qL[2] = T(0.0); qL[3] = T(0.0);
                qR[1] = T(0.0); qR[2] = T(0.0); qR[3] = T(0.0);
            }
        }
        T fl[5];
        if (K.initPhase) vanleer_flux<T>(qL, qR, nx, ny, nz, T(K.gamma), fl);
        else roe_flux<T>(qL, qR, nx, ny, nz, T(K.gamma), K.entropyFix, T(K.entropyCoef),
                         K.rtem ? T(K.rtem[le]) : T(0.0), K.rtem ? T(K.rtem[re]) : T(0.0), fl);
        const T area = T(fr.area);
        // inviscid spectral radius term at the face-averaged state
        const double vnl = fr.nx * L.q[1] + fr.ny * L.q[2] + fr.nz * L.q[3] - 0.0;
        const double vnr = fr.nx * R.q[1] + fr.ny * R.q[2] + fr.nz * R.q[3] - 0.0;
        const double g = 0.5 * (K.gamma + K.gamma);
        const double pm = 0.5 * (L.q[4] + R.q[4]);
        const double rm = 0.5 * (L.q[0] + R.q[0]);
        const double vn = 0.5 * (vnl + vnr);
        const double cm = sqrt(g * pm / rm);
        double inv[6];
#pragma unroll
        for (int m = 0; m < 5; ++m) inv[m] = double(fl[m] * area);
        inv[5] = 0.5 * fr.area * (dABS(vn) + cm);
        if constexpr (COMPACT) store_compact(invOut, inv);
        else store6(invOut, inv);
    }
    if ((PARTS & 2) != 0 && K.viscous) {
        T gTL[3] = {T(0.0), T(0.0), T(0.0)}, gTR[3] = {T(0.0), T(0.0), T(0.0)};
        if (le < nC) { gTL[0] = gradT[le]; gTL[1] = gradT[(size_t)nC + le]; gTL[2] = gradT[(size_t)2 * nC + le]; }
        if (re < nC) { gTR[0] = gradT[re]; gTR[1] = gradT[(size_t)nC + re]; gTR[2] = gradT[(size_t)2 * nC + re]; }
        else if (K.gradient != 0) { gTR[0] = gTL[0]; gTR[1] = gTL[1]; gTR[2] = gTL[2]; }  // ghost: owner's lsq gradient
        const bool wall = bnd && bc == kBC_WALL;
        double vis[6];
        viscous_face(M, K, f, fr, L, R, gTL, gTR, bnd, wall, vist[le], vist[re], vis);
        const double tl = L.q[4] / (K.R * L.q[0]);
        const double tr = R.q[4] / (K.R * R.q[0]);
        const double vl = 0.5 * (tl * sqrt(tl) * (1.0 + K.tsuth) / (tl + K.tsuth) +
                                 tr * sqrt(tr) * (1.0 + K.tsuth) / (tr + K.tsuth));
        const double vt = 0.5 * (vist[le] + vist[re]);
        const double dx = R.x - L.x, dy = R.y - L.y, dz = R.z - L.z;
        const double ds = dABS(fr.nx * dx + fr.ny * dy + fr.nz * dz);
        const double rho = 0.5 * (L.q[0] + R.q[0]);
        const double mu = vl + vt;
        double sr = 2.0 * mu / (rho * ds * K.reynolds + kSMALL);
        sr *= 0.5 * fr.area;
        vis[5] = sr;
        if constexpr (COMPACT) store_compact(visOut, vis);
        else store6(visOut, vis);
    }
}

template <typename T, int PARTS>
__global__ void __launch_bounds__(256) k_face_flux(DevMesh M, Consts K0, const GatherRec *__restrict__ rec,
                                                   const T *__restrict__ gradT, const double *__restrict__ vist,
                                                   double *__restrict__ faceInv, double *__restrict__ faceVis)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= M.nFace) return;
    // FP32 faces are stored compact (32 bytes a face and group)
    constexpr bool C = sizeof(T) == 4;
    constexpr int stride = C ? 4 : 6;
    face_values<T, C, PARTS>(M, K0, rec, gradT, vist, f, faceInv + (size_t)f * stride,
                             faceVis + (size_t)f * stride);
}

// The hook form of a face loop's LoadFlux (NSSolverUnstruct.cpp:6435): one
// thread per cell adds one group of stored face values (the inviscid ones of
// InviscidFlux, or the viscous ones of ViscousFlux) to the residual already
// in res, in the cell's face order.  LoadResiduals (res = -0.0), then the
// inviscid pass, then the viscous pass round exactly like k_cell_gather's
// single accumulation.
template <bool COMPACT>
__global__ void __launch_bounds__(kBlock) k_cell_add(DevMesh M, const double *__restrict__ fv,
                                                     double *__restrict__ res)
{
    const int c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int nC = M.nCell;
    if (c >= nC) return;
    double acc[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) acc[m] = res[(size_t)m * nC + c];
    for (int s = 0; s < M.ellW; ++s) {
        const int enc = M.c2fEnc[(size_t)s * nC + c];
        if (enc == 0) break;
        const bool isL = enc > 0;
        const int f = (isL ? enc : -enc) - 1;
        double fl[5], sp;
        if constexpr (COMPACT) {
            load_compact(fv + (size_t)f * 4, fl, sp);
        } else {
            const double2 *src = reinterpret_cast<const double2 *>(fv + (size_t)f * 6);
            const double2 a = src[0], b = src[1], e = src[2];
            fl[0] = a.x; fl[1] = a.y; fl[2] = b.x; fl[3] = b.y; fl[4] = e.x;
        }
        if (isL) {
#pragma unroll
            for (int m = 0; m < 5; ++m) acc[m] -= fl[m];
        } else {
#pragma unroll
            for (int m = 0; m < 5; ++m) acc[m] += fl[m];
        }
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) res[(size_t)m * nC + c] = acc[m];
}

// NSSolverUnstruct::CalPressureFactor (NSSolverUnstruct.cpp:6976-7045): the
// cell sensor of the pressure-sensor entropy fixes, one thread per real cell.
// Method 2: the pressures on the left and right sides of the cell's faces
// (ghost states on boundary faces) summed in face order, then
// |sumR - sumL| / |sumR + sumL| / nFaces.  Method 6: the pressure ratio
// max/min over the cell and its interior-face neighbours mapped through
// f = clip((2.6 - ratio) / 1.3, 0, 1) to (1 - cos(f pi)) / 2.  The device's
// cos may differ from the host library's in the last bit.
__global__ void __launch_bounds__(kBlock) k_pressure_factor(DevMesh M, int method, const double *__restrict__ q,
                                                            double *__restrict__ rtem)
{
    const int c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int nC = M.nCell;
    if (c >= nC) return;
    const double *p = q + (size_t)4 * M.nTotal;
    const double pc = p[c];
    double r;
    if (method == 2) {
        double sl = 0.0, sr = 0.0;
        int n = 0;
        for (int s = 0; s < M.ellW; ++s) {
            const int enc = M.c2fEnc[(size_t)s * nC + c];
            if (enc == 0) break;
            const double pn = p[M.c2n[(size_t)s * nC + c]];
            sl += enc > 0 ? pc : pn;
            sr += enc > 0 ? pn : pc;
            ++n;
        }
        r = 0.0;
        r += dABS(sr - sl) / dABS(sr + sl);
        r /= (double)n;
    } else {
        double lo = pc, hi = pc;
        for (int s = 0; s < M.ellW; ++s) {
            const int enc = M.c2fEnc[(size_t)s * nC + c];
            if (enc == 0) break;
            const int nb = M.c2n[(size_t)s * nC + c];
            if (nb >= nC) continue;  // a boundary face's ghost
            const double pn = p[nb];
            if (pn < lo) lo = pn;
            if (pn > hi) hi = pn;
        }
        const double ratio = hi / lo;
        const double fn = dMIN(1.0, dMAX(0.0, (2.6 - ratio) / (2.6 - 1.3)));
        const double cf = dMAX(-1.0, dMIN(cos(fn * 3.141592653589), 1.0));
        r = 0.5 * (1.0 - cf);
    }
    rtem[c] = r;
}

// SetGhostCellExceptInterface (Geo_UnstructGrid.cpp:1857-1881): a ghost
// takes its owner's pressure factor.
__global__ void __launch_bounds__(kBlock) k_ghost_copy(DevMesh M, double *__restrict__ v)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < M.nBFace) v[M.nCell + f] = v[M.fle[f]];
}

// LoadResiduals (NSSolverUnstruct.cpp:4116): res = -rhs, and rhs is zero on
// the single grid level, so every entry becomes -0.0.
__global__ void __launch_bounds__(kBlock) k_load_res(double *__restrict__ res, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) res[i] = -0.0;
}

// Block partials of res^2 over 256-cell blocks of the storage order, grouped
// and summed exactly as k_cell_gather<256> groups them, so a norm taken after
// the hook sequence equals the one of the fused residual evaluation.
__global__ void __launch_bounds__(256) k_res_partial(int nC, size_t ld, const double *__restrict__ res,
                                                     double *__restrict__ normPart)
{
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int c = lb * 256 + threadIdx.x;
    __shared__ double red[5][256];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const double v = c < nC ? res[(size_t)m * ld + c] : 0.0;
        red[m][threadIdx.x] = c < nC ? v * v : 0.0;
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int m = 0; m < 5; ++m) red[m][threadIdx.x] += red[m][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int m = 0; m < 5; ++m) normPart[(size_t)lb * 5 + m] = red[m][0];
    }
}

// Per cell: the face values summed in the cell's face order (every inviscid
// contribution, then every viscous one: the reference's LoadFlux order),
// spectral radii, local time step and the block partials of res^2 and min dt.
template <int BS, bool COMPACT>
__global__ void __launch_bounds__(BS) k_cell_gather(DevMesh M, Consts K0, const double *__restrict__ q,
                                                    const double *__restrict__ faceInv,
                                                    const double *__restrict__ faceVis, double *__restrict__ res,
                                                    double *__restrict__ specInv, double *__restrict__ specVis,
                                                    double *__restrict__ normPart, const double *__restrict__ cflArr,
                                                    const int *__restrict__ stepCtr, double *__restrict__ dtLoc,
                                                    double *__restrict__ dtMinPart, bool fusedDt)
{
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int c = lb * blockDim.x + threadIdx.x;
    const int nC = M.nCell;
    double acc[5] = {-0.0, -0.0, -0.0, -0.0, -0.0};
    double dtl = 1.0e300;
    if (c < nC) {
        const Consts K = case_k(K0, M, c);
        double sInv = 0.0, sVis = 0.0;
        for (int s = 0; s < M.ellW; ++s) {
            const int enc = M.c2fEnc[(size_t)s * nC + c];
            if (enc == 0) break;
            const bool isL = enc > 0;
            const int f = (isL ? enc : -enc) - 1;
            double fl[5], sp;
            if constexpr (COMPACT) {
                load_compact(faceInv + (size_t)f * 4, fl, sp);
            } else {
                const double2 *src = reinterpret_cast<const double2 *>(faceInv + (size_t)f * 6);
                const double2 a = src[0], b = src[1], e = src[2];
                fl[0] = a.x; fl[1] = a.y; fl[2] = b.x; fl[3] = b.y; fl[4] = e.x;
                sp = e.y;
            }
            if (isL) {
#pragma unroll
                for (int m = 0; m < 5; ++m) acc[m] -= fl[m];
            } else {
#pragma unroll
                for (int m = 0; m < 5; ++m) acc[m] += fl[m];
            }
            sInv += sp;
        }
        specInv[c] = sInv;
        if (K.viscous) {
            for (int s = 0; s < M.ellW; ++s) {
                const int enc = M.c2fEnc[(size_t)s * nC + c];
                if (enc == 0) break;
                const bool isL = enc > 0;
                const int f = (isL ? enc : -enc) - 1;
                double fl[5], sp;
                if constexpr (COMPACT) {
                    load_compact(faceVis + (size_t)f * 4, fl, sp);
                } else {
                    const double2 *src = reinterpret_cast<const double2 *>(faceVis + (size_t)f * 6);
                    const double2 a = src[0], b = src[1], e = src[2];
                    fl[0] = a.x; fl[1] = a.y; fl[2] = b.x; fl[3] = b.y; fl[4] = e.x;
                    sp = e.y;
                }
                if (isL) {
#pragma unroll
                    for (int m = 0; m < 5; ++m) acc[m] -= fl[m];
                } else {
#pragma unroll
                    for (int m = 0; m < 5; ++m) acc[m] += fl[m];
                }
                sVis += sp;
            }
            specVis[c] = sVis;
        }
        if (fusedDt) {
            double spec = sInv;
            if (K.viscous) spec += sVis;
            dtl = local_dt(M, K, c, q[(size_t)4 * M.nTotal + c], spec, cflArr[*stepCtr]);
            dtLoc[c] = dtl;
        }
#pragma unroll
        for (int m = 0; m < 5; ++m) res[(size_t)m * nC + c] = acc[m];
    }
    __shared__ double red[6][BS];
#pragma unroll
    for (int m = 0; m < 5; ++m) red[m][threadIdx.x] = (c < nC) ? acc[m] * acc[m] : 0.0;
    red[5][threadIdx.x] = dtl;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int m = 0; m < 5; ++m) red[m][threadIdx.x] += red[m][threadIdx.x + w];
            red[5][threadIdx.x] = dMIN(red[5][threadIdx.x], red[5][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int m = 0; m < 5; ++m) normPart[(size_t)lb * 5 + m] = red[m][0];
        if (fusedDt) dtMinPart[lb] = red[5][0];
    }
}

// Tile-fused face loop: workgroup t owns the cells [t*BS, (t+1)*BS) of the
// storage order.  It evaluates every face touching them (host-built list)
// into LDS, then each thread sums its cell's faces from LDS in the cell's
// face order, exactly like k_cell_gather sums the stored face values.  Faces
// between two tiles are evaluated by both (identical values); the face values
// never travel through HBM.
template <typename T, int BS>
__global__ void __launch_bounds__(BS) k_flux_tile(DevMesh M, Consts K0, const GatherRec *__restrict__ rec,
                                                  const T *__restrict__ gradT, const double *__restrict__ vist,
                                                  const int *__restrict__ tileFacePtr, const int *__restrict__ tileFaces,
                                                  const unsigned short *__restrict__ c2loc, int ldsStride,
                                                  const double *__restrict__ q, double *__restrict__ res,
                                                  double *__restrict__ specInv, double *__restrict__ specVis,
                                                  double *__restrict__ normPart, const double *__restrict__ cflArr,
                                                  const int *__restrict__ stepCtr, double *__restrict__ dtLoc,
                                                  double *__restrict__ dtMinPart, bool fusedDt)
{
    extern __shared__ double lds[];  // [ldsStride][6] inviscid, then [ldsStride][6] viscous
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int fb = tileFacePtr[lb], nf = tileFacePtr[lb + 1] - fb;
    double *li = lds, *lv = lds + (size_t)ldsStride * 6;
    for (int i = threadIdx.x; i < nf; i += BS)
        face_values<T>(M, K0, rec, gradT, vist, tileFaces[fb + i], li + (size_t)i * 6, lv + (size_t)i * 6);
    __syncthreads();
    const int c = lb * BS + threadIdx.x;
    const int nC = M.nCell;
    double acc[5] = {-0.0, -0.0, -0.0, -0.0, -0.0};
    double dtl = 1.0e300;
    if (c < nC) {
        const Consts K = case_k(K0, M, c);
        double sInv = 0.0, sVis = 0.0;
        for (int s = 0; s < M.ellW; ++s) {
            const int enc = M.c2fEnc[(size_t)s * nC + c];
            if (enc == 0) break;
            const double *v = li + (size_t)c2loc[(size_t)s * nC + c] * 6;
            if (enc > 0) {
#pragma unroll
                for (int m = 0; m < 5; ++m) acc[m] -= v[m];
            } else {
#pragma unroll
                for (int m = 0; m < 5; ++m) acc[m] += v[m];
            }
            sInv += v[5];
        }
        specInv[c] = sInv;
        if (K.viscous) {
            for (int s = 0; s < M.ellW; ++s) {
                const int enc = M.c2fEnc[(size_t)s * nC + c];
                if (enc == 0) break;
                const double *v = lv + (size_t)c2loc[(size_t)s * nC + c] * 6;
                if (enc > 0) {
#pragma unroll
                    for (int m = 0; m < 5; ++m) acc[m] -= v[m];
                } else {
#pragma unroll
                    for (int m = 0; m < 5; ++m) acc[m] += v[m];
                }
                sVis += v[5];
            }
            specVis[c] = sVis;
        }
        if (fusedDt) {
            double spec = sInv;
            if (K.viscous) spec += sVis;
            dtl = local_dt(M, K, c, q[(size_t)4 * M.nTotal + c], spec, cflArr[*stepCtr]);
            dtLoc[c] = dtl;
        }
#pragma unroll
        for (int m = 0; m < 5; ++m) res[(size_t)m * nC + c] = acc[m];
    }
    __syncthreads();  // the reduction below reuses the start of the LDS
    double *red = lds;  // [6][BS]
#pragma unroll
    for (int m = 0; m < 5; ++m) red[m * BS + threadIdx.x] = (c < nC) ? acc[m] * acc[m] : 0.0;
    red[5 * BS + threadIdx.x] = dtl;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int m = 0; m < 5; ++m) red[m * BS + threadIdx.x] += red[m * BS + threadIdx.x + w];
            red[5 * BS + threadIdx.x] = dMIN(red[5 * BS + threadIdx.x], red[5 * BS + threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int m = 0; m < 5; ++m) normPart[(size_t)lb * 5 + m] = red[m * BS];
        if (fusedDt) dtMinPart[lb] = red[5 * BS];
    }
}

// Residual sums of the step (row *stepCtr of normOut) and, optionally, the
// global minimum of the local time step (ComputeMinTimeStep; a minimum is
// independent of the reduction order).
// The partials are summed in a fixed order: thread t takes parts t, t+kBlock,
// ..., then a pairwise tree over the threads; red[m][0] holds sum m (and
// red[5][0] the minimum time step).
__device__ __forceinline__ void sum_parts(const double *__restrict__ part, const double *__restrict__ dtPart,
                                          int nPart, int withDt, double (*red)[kBlock])
{
    double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double mn = 1.0e300;
    for (int i = threadIdx.x; i < nPart; i += blockDim.x) {
#pragma unroll
        for (int m = 0; m < 5; ++m) a[m] += part[(size_t)i * 5 + m];
        if (withDt) mn = dMIN(mn, dtPart[i]);
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) red[m][threadIdx.x] = a[m];
    red[5][threadIdx.x] = mn;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int m = 0; m < 5; ++m) red[m][threadIdx.x] += red[m][threadIdx.x + w];
            red[5][threadIdx.x] = dMIN(red[5][threadIdx.x], red[5][threadIdx.x + w]);
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kBlock) k_norm_final(const double *__restrict__ part,
                                                       const double *__restrict__ dtPart, int nPart,
                                                       double *__restrict__ normOut,
                                                       const int *__restrict__ stepCtr,
                                                       double *__restrict__ dtMin, int withDt)
{
    __shared__ double red[6][kBlock];
    sum_parts(part, dtPart, nPart, withDt, red);
    if (threadIdx.x < 5) normOut[(size_t)(*stepCtr) * 5 + threadIdx.x] = red[threadIdx.x][0];
    if (threadIdx.x == 5 && withDt) *dtMin = red[5][0];
}

// The five sums alone, to a caller's array (the no-allocation entry point).
__global__ void __launch_bounds__(kBlock) k_norm_sum5(const double *__restrict__ part, int nPart,
                                                      double *__restrict__ sum5)
{
    __shared__ double red[6][kBlock];
    sum_parts(part, part, nPart, 0, red);
    if (threadIdx.x < 5) sum5[threadIdx.x] = red[threadIdx.x][0];
}

// The squared change of dq since the last call (the sweepNormal of
// Controller::LUSGS for nLUSGSSweeps > 1, Controller.cpp:1920-1950; the
// reference adds SQR(dq - dqOld) cell by cell in its sweeps): per block of
// kBlock real cells, over nv components (component m of cell c at
// dq[c * cs + m * ms]), and prev [nv][nC] takes the new dq.  k_sum_into then
// adds the block partials in a fixed order.
__global__ void __launch_bounds__(kBlock) k_sweep_change(const double *__restrict__ dq, size_t cs, size_t ms, int nv,
                                                         int nC, double *__restrict__ prev,
                                                         double *__restrict__ parts)
{
    __shared__ double red[kBlock];
    const int c = blockIdx.x * kBlock + (int)threadIdx.x;
    double acc = 0.0;
    if (c < nC)
        for (int m = 0; m < nv; ++m) {
            const double v = dq[(size_t)c * cs + (size_t)m * ms];
            const double d = v - prev[(size_t)m * nC + c];
            acc += d * d;
            prev[(size_t)m * nC + c] = v;
        }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(kBlock) k_sum_into(const double *__restrict__ parts, int n, double *__restrict__ out)
{
    __shared__ double red[kBlock];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += kBlock) acc += parts[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// Batched contexts: per-case residual sums of the step and minimum local time
// step in a fixed order.  Workgroup (case cs, slice j) reduces the j-th of
// kCaseSlices contiguous parts of the case's cells into casePart; then one
// thread per case adds its slices in order.
constexpr int kCaseSlices = 32;
__global__ void __launch_bounds__(kBlock) k_case_reduce(DevMesh M, const double *__restrict__ res,
                                                        const double *__restrict__ dtLoc,
                                                        double *__restrict__ casePart)
{
    __shared__ double red[6][kBlock];
    const int cs = blockIdx.x / kCaseSlices, j = blockIdx.x % kCaseSlices;
    const int cb = M.caseCellPtr[cs], ce = M.caseCellPtr[cs + 1];
    const int per = (ce - cb + kCaseSlices - 1) / kCaseSlices;
    const int b = min(ce, cb + j * per), e = min(ce, b + per);
    const int nC = M.nCell;
    double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double mn = 1.0e300;
    for (int c = b + (int)threadIdx.x; c < e; c += blockDim.x) {
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            const double r = res[(size_t)m * nC + c];
            a[m] += r * r;
        }
        mn = dMIN(mn, dtLoc[c]);
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) red[m][threadIdx.x] = a[m];
    red[5][threadIdx.x] = mn;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int m = 0; m < 5; ++m) red[m][threadIdx.x] += red[m][threadIdx.x + w];
            red[5][threadIdx.x] = dMIN(red[5][threadIdx.x], red[5][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) casePart[(size_t)blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void __launch_bounds__(64) k_case_final(const double *__restrict__ casePart, int nCases,
                                                   double *__restrict__ normOut, const int *__restrict__ stepCtr,
                                                   double *__restrict__ dtMin)
{
    const int cs = blockIdx.x * blockDim.x + threadIdx.x;
    if (cs >= nCases) return;
    double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double mn = 1.0e300;
    for (int j = 0; j < kCaseSlices; ++j) {
        const double *p = casePart + ((size_t)cs * kCaseSlices + j) * 6;
#pragma unroll
        for (int m = 0; m < 5; ++m) a[m] += p[m];
        mn = dMIN(mn, p[5]);
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) normOut[((size_t)(*stepCtr) * nCases + cs) * 5 + m] = a[m];
    dtMin[cs] = mn;
}

// ---------------------------------------------------------- time step
// Standalone TimeStep hook: local dt and block minima.
__global__ void __launch_bounds__(kBlock) k_time_step(DevMesh M, Consts K0, const double *__restrict__ q,
                                                      const double *__restrict__ specInv,
                                                      const double *__restrict__ specVis, double cfl,
                                                      double *__restrict__ dtLoc, double *__restrict__ dtMinPart)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    double dtl = 1.0e300;
    if (c < M.nCell) {
        const Consts K = case_k(K0, M, c);
        double spec = specInv[c];
        if (K.viscous) spec += specVis[c];
        dtl = local_dt(M, K, c, q[(size_t)4 * M.nTotal + c], spec, cfl);
        dtLoc[c] = dtl;
    }
    __shared__ double red[kBlock];
    red[threadIdx.x] = dtl;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = dMIN(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) dtMinPart[blockIdx.x] = red[0];
}

__global__ void k_min_final(const double *__restrict__ part, int nPart, double *__restrict__ out)
{
    __shared__ double red[kBlock];
    double mn = 1.0e300;
    for (int i = threadIdx.x; i < nPart; i += blockDim.x) mn = dMIN(mn, part[i]);
    red[threadIdx.x] = mn;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = dMIN(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// ReduceMaxTimeStep + Diagonal: dt = min(dt, ktmax*dtmin)/vol; D = 1/dt + radii.
__global__ void __launch_bounds__(kBlock) k_dt_diag(DevMesh M, Consts K, const double *__restrict__ dtLoc,
                                                    const double *__restrict__ specInv,
                                                    const double *__restrict__ specVis,
                                                    const double *__restrict__ dtMin,
                                                    double *__restrict__ dt, double *__restrict__ diag)
{
    const int c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (c >= M.nCell) return;
    const double mn = dtMin[K.nCases > 1 ? M.caseOf[c] : 0];
    double d = dtLoc[c];
    if (K.ktmax > 0) d = dMIN(d, K.ktmax * mn);
    d /= M.vol[c];
    dt[c] = d;
    double D = 1.0 / d;
    D += specInv[c];
    if (K.viscous) D += specVis[c];
    diag[c] = D;
}

__device__ __forceinline__ void sweep_finish(const DevMesh &M, int c, bool useU, const double *rhs0,
                                             const double *__restrict__ res, const double *__restrict__ diag,
                                             double *dq, double *dflux)
{
    const int nC = M.nCell;
    // every load before the first store, so they are issued together
    double r[5], U[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    load5(res, nC, c, r);
    const double D = diag[c];
    if (useU) load_row5(dflux, c, U);
    double v[5], w[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        v[m] = (r[m] - U[m] - rhs0[m]) / D;
        w[m] = 0.0 + rhs0[m];
    }
    store_row5(dq, c, v);
    store_row5(dflux, c, w);
}

__device__ __forceinline__ void load_coef(const double *__restrict__ coef, int f, bool isL, double *co)
{
    const double *p = coef + ((size_t)f * 2 + (isL ? 0 : 1)) * kCoef;
    const double2 *src = reinterpret_cast<const double2 *>(p);
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        const double2 v = src[j];
        co[2 * j] = v.x;
        co[2 * j + 1] = v.y;
    }
    co[14] = p[14];
}

// Sweep of one cell for meshes with at most kW faces per cell: the face list,
// the coefficients and the neighbours' dq are all loaded before any is used,
// so a level costs a few memory latencies.
constexpr int kW = 6;
constexpr int kSweepBlock = 512;
template <bool FWD>
__device__ __forceinline__ void sweep_cell_batched(const DevMesh &M, const Consts &K, int c, bool useU,
                                                   const double *__restrict__ coef, const double *__restrict__ res, const double *__restrict__ diag,
                                                   double *dq, double *dflux)
{
    const int nC = M.nCell;
    int enc[kW], nb[kW];
#pragma unroll
    for (int s = 0; s < kW; ++s) {
        enc[s] = (s < M.ellW) ? M.c2fEnc[(size_t)s * nC + c] : 0;
        nb[s] = (s < M.ellW) ? M.c2n[(size_t)s * nC + c] : 0;
    }
    // issue every load first (inactive slots read face 0 / cell 0 and are
    // masked), then accumulate in face order
    bool act[kW];
    double co[kW][15], dqn[kW][5];
#pragma unroll
    for (int s = 0; s < kW; ++s) {
        act[s] = enc[s] != 0 && (FWD ? (mesh_index(M, nb[s]) < mesh_index(M, c))
                                     : (mesh_index(M, nb[s]) > mesh_index(M, c)));
        const bool isL = enc[s] > 0;
        const int f = act[s] ? (isL ? enc[s] : -enc[s]) - 1 : 0;
        load_coef(coef, f, isL, co[s]);
        load_row5(dq, act[s] ? nb[s] : 0, dqn[s]);
    }
    double rhs0[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < kW; ++s) {
        if (!act[s]) continue;
        double df[5];
        sweep_apply(K, co[s], dqn[s], df);
#pragma unroll
        for (int m = 0; m < 5; ++m) rhs0[m] += df[m];
    }
    sweep_finish(M, c, useU, rhs0, res, diag, dq, dflux);
}

// Any number of faces per cell (polyhedral meshes): face by face.
template <bool FWD>
__device__ __forceinline__ void sweep_cell_generic(const DevMesh &M, const Consts &K, int c, bool useU,
                                                   const double *__restrict__ coef, const double *__restrict__ res, const double *__restrict__ diag,
                                                   double *dq, double *dflux)
{
    const int nC = M.nCell;
    double rhs0[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int s = 0; s < M.ellW; ++s) {
        const int enc = M.c2fEnc[(size_t)s * nC + c];
        if (enc == 0) break;
        const int nb = M.c2n[(size_t)s * nC + c];
        if (FWD ? (mesh_index(M, nb) > mesh_index(M, c)) : (mesh_index(M, nb) < mesh_index(M, c))) continue;
        const bool isL = enc > 0;
        const int f = (isL ? enc : -enc) - 1;
        double co[15], dqn[5], df[5];
        load_coef(coef, f, isL, co);
        load_row5(dq, nb, dqn);
        sweep_apply(K, co, dqn, df);
#pragma unroll
        for (int m = 0; m < 5; ++m) rhs0[m] += df[m];
    }
    sweep_finish(M, c, useU, rhs0, res, diag, dq, dflux);
}

template <bool FWD>
__device__ __forceinline__ void sweep_cell(const DevMesh &M, const Consts &K, int c, bool useU,
                                           const double *__restrict__ coef, const double *__restrict__ res, const double *__restrict__ diag, double *dq, double *dflux)
{
    if (M.ellW <= kW) sweep_cell_batched<FWD>(M, K, c, useU, coef, res, diag, dq, dflux);
    else sweep_cell_generic<FWD>(M, K, c, useU, coef, res, diag, dq, dflux);
}

// One level spread over the whole GPU (64-thread workgroups).
template <bool FWD>
__global__ void __launch_bounds__(64) k_sweep_level(DevMesh M, Consts K, const int *__restrict__ cells, int begin,
                                                    int end, bool useU, const double *__restrict__ coef,
                                                    const double *__restrict__ res, const double *__restrict__ diag,
                                                    double *dq, double *dflux)
{
    const int i = begin + xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= end) return;
    sweep_cell<FWD>(M, K, cells[i], useU, coef, res, diag, dq, dflux);
}

// One workgroup walks a run of small levels; the barrier between levels
// makes the previous level's dq visible to the whole workgroup.
template <bool FWD>
__global__ void __launch_bounds__(kSweepBlock) k_sweep_block(DevMesh M, Consts K, const int *__restrict__ ptr,
                                                      const int *__restrict__ cells, int l0, int l1, bool useU,
                                                      const double *__restrict__ coef, const double *__restrict__ res, const double *__restrict__ diag, double *dq, double *dflux)
{
    for (int L = l0; L < l1; ++L) {
        const int b = ptr[L], e = ptr[L + 1];
        for (int i = b + (int)threadIdx.x; i < e; i += blockDim.x)
            sweep_cell<FWD>(M, K, cells[i], useU, coef, res, diag, dq, dflux);
        __syncthreads();
    }
}

// Sweep-ordered variant.  For each direction the host lists, per position
// p of the level order, the cell and its active faces in face order (row =
// 2*face + side of the coefficient block, nb = the neighbour), so the loads
// of a position are independent of each other: one memory round for the
// lists, one for coefficients and neighbour dq, then the store.
// The whole 128-byte block is loaded so that no part of it is fetched later
// inside a branch.
template <typename CT>
__device__ __forceinline__ void load_coef_row(const double *__restrict__ coef, int row, double *co)
{
    if constexpr (sizeof(CT) == 4) {  // FP32 rows (mixed form), 64 bytes
        const float4 *src = reinterpret_cast<const float4 *>(reinterpret_cast<const float *>(coef) + (size_t)row * kCoef);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float4 v = src[j];
            co[4 * j] = v.x; co[4 * j + 1] = v.y; co[4 * j + 2] = v.z; co[4 * j + 3] = v.w;
        }
    } else {
        const double2 *src = reinterpret_cast<const double2 *>(coef + (size_t)row * kCoef);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const double2 v = src[j];
            co[2 * j] = v.x;
            co[2 * j + 1] = v.y;
        }
    }
}

template <int NW, int VIS, typename CT>
__device__ __forceinline__ void sweep_position(const DevMesh &M, const Consts &K, const SweepOrder &so, int p,
                                               bool useU, const double *__restrict__ coef,
                                               const double *__restrict__ res, const double *__restrict__ diag,
                                               double *dq, double *dflux)
{
    const int P = M.nCell;
    const int c = so.cells[p];
    int row[NW], nb[NW];
#pragma unroll
    for (int s = 0; s < NW; ++s) {  // the lists are padded to NW slots
        row[s] = so.row[(size_t)s * P + p];
        nb[s] = so.nb[(size_t)s * P + p];
    }
    // the cell's own operands first: they depend on c only
    double r[5], U[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    load5(res, P, c, r);
    const double D = diag[c];
    if (useU) load_row5(dflux, c, U);
    double co[NW][16], dqn[NW][5];
#pragma unroll
    for (int s = 0; s < NW; ++s) {
        const bool act = row[s] >= 0;
        load_coef_row<CT>(coef, act ? row[s] : 0, co[s]);
        load_row5(dq, act ? nb[s] : 0, dqn[s]);
    }
    // Unused slots add +0.0, which leaves the sum unchanged (it starts at
    // +0.0, so it is never -0.0); no branch lets the loads of a slot sink.
    double rhs0[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < NW; ++s) {
        double df[5];
        sweep_apply<VIS>(K, co[s], dqn[s], df);
        const bool act = row[s] >= 0;
#pragma unroll
        for (int m = 0; m < 5; ++m) rhs0[m] += act ? df[m] : 0.0;
    }
    double v[5], w[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        v[m] = (r[m] - U[m] - rhs0[m]) / D;
        w[m] = 0.0 + rhs0[m];
    }
    store_row5(dq, c, v);
    store_row5(dflux, c, w);
}

template <int NW, int VIS, typename CT>
__global__ void __launch_bounds__(64) k_sweep_level_so(DevMesh M, Consts K, SweepOrder so, int begin, int end,
                                                       bool useU, const double *__restrict__ coef,
                                                       const double *__restrict__ res, const double *__restrict__ diag, double *dq, double *dflux)
{
    const int p = begin + xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (p >= end) return;
    sweep_position<NW, VIS, CT>(M, K, so, p, useU, coef, res, diag, dq, dflux);
}

template <int NW, int VIS, typename CT>
__global__ void __launch_bounds__(kSweepBlock) k_sweep_block_so(DevMesh M, Consts K, SweepOrder so,
                                                               const int *__restrict__ ptr, int l0, int l1,
                                                               bool useU, const double *__restrict__ coef,
                                                               const double *__restrict__ res, const double *__restrict__ diag, double *dq,
                                                               double *dflux)
{
    for (int L = l0; L < l1; ++L) {
        const int b = ptr[L], e = ptr[L + 1];
        for (int p = b + (int)threadIdx.x; p < e; p += blockDim.x)
            sweep_position<NW, VIS, CT>(M, K, so, p, useU, coef, res, diag, dq, dflux);
        __syncthreads();
    }
}

// ------------------------------------------------- dataflow sweep
// One launch per direction instead of one per level (sweep_form 0 and 2).  Each
// 64-thread block (one wavefront) takes the next chunk of 64 positions of
// the padded level order from a ticket counter.  A chunk depends only on
// chunks of earlier levels, which were taken earlier by blocks that are
// running or done, so every wait ends.  The wave polls the flags of the
// chunks it depends on (one per lane), then sweeps its 64 cells exactly as
// sweep_position does and publishes its chunk.  dq and the flags are stored
// and loaded with device-scope (sc1) accesses, which are coherent across the
// XCDs' L2 caches, and the wave's dq stores complete (vmcnt 0) before its
// flag store is issued.
// ctl: [0] ticket, [1] epoch of the current sweep, [2] waves that gave up,
// [3] the wait bound in ticks of the 100 MHz wall clock (set by the host;
// 20,000,000 = 0.2 s by default).

// 128- and 64-bit payloads of the raw buffer loads and stores, and word 3
// of their buffer descriptor (untyped 32-bit data, as raw buffer accesses
// on gfx9-family parts take it); cache policy 16 is sc1 (device scope).
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v2i __attribute__((ext_vector_type(2)));
constexpr int kRawBufferWord3 = 0x00020000;
constexpr int kDeviceScope = 16;
__global__ void k_sweep_flow_prep(int *ctl)
{
    ctl[0] = 0;
    ctl[1] += 1;
}

__device__ __forceinline__ int flag_load(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double dq_load(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One chunk of the dataflow sweep, swept by one wavefront: false if the wait
// bound was exceeded (the chunk is then neither swept nor published).
// sa.on: the chunk's cells also take the same-direction step of the scalar
// Spalart-Allmaras sweep (the same lists and order as sa_sweep_position; its
// dq is published under the same flag).
template <int NW, int VIS, typename CT>
__device__ __forceinline__ bool flow_chunk(const DevMesh &M, const Consts &K, const FlowOrder &fo, int *ctl,
                                           int *chunkFlag, bool useU, const double *__restrict__ coef,
                                           const double *__restrict__ res, const double *__restrict__ diag,
                                           double *dq, double *dflux, const SaSweep &sa,
                                           const GatherRec *__restrict__ rec, const double *__restrict__ vist, int tk,
                                           int lane, int epoch, long long t0, long long waitTicks)
{
    const int nT = M.nTotal, nC = M.nCell, P = 64 * fo.nChunks;
    const int p = tk * 64 + lane;
    const int c = fo.cells[p];
    const bool live = c >= 0;
    const int cc = live ? c : 0;
    // Active slots come first in every list, so slots at or beyond the
    // chunk's widest cell are unused by all 64 cells: they are skipped
    // (a skipped slot would add +0.0 to a sum that is never -0.0).
    const int width = __builtin_amdgcn_readfirstlane(fo.width[tk]);
    // everything that does not depend on the neighbours' dq is loaded first
    int row[NW], nb[NW];
#pragma unroll
    for (int s = 0; s < NW; ++s) {
        row[s] = -1;
        nb[s] = 0;
        if (s < width) {
            row[s] = fo.row[(size_t)s * P + p];
            nb[s] = fo.nb[(size_t)s * P + p];
        }
    }
    double r[5];
    load5(res, nC, cc, r);
    const double D = diag[cc];
    double U[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (useU) load_row5(dflux, cc, U);
    double co[NW][16];
    if (K.coefInSweep) {
        // the face terms of k_grad_lim, evaluated here while the chunk waits:
        // sweep_coef on the same operands, rounded to CT as a stored row is.
        // States and centroids come from the gather records (the first 64
        // bytes of a record: one line per neighbour instead of eight).
        const Consts Kc = case_k(K, M, cc);
        double qc[5], xc[3];
        load_rec_state(rec, cc, qc, xc);
        const double vc = vist[cc];
#pragma unroll
        for (int s = 0; s < NW; ++s) {
            if (s >= width) break;
            const int rw = row[s] >= 0 ? row[s] : 0, n = row[s] >= 0 ? nb[s] : cc;
            const FaceRec fr = M.face[rw >> 1];
            const bool isL = (rw & 1) == 0;
            double qn[5], xn[3];
            load_rec_state(rec, n, qn, xn);
            const double dist = Kc.viscous ? dABS(fr.nx * (xn[0] - xc[0]) + fr.ny * (xn[1] - xc[1]) +
                                                  fr.nz * (xn[2] - xc[2]))
                                           : 0.0;
            sweep_coef(Kc, qc, qn, fr, isL, dist, vc, vist[n], co[s]);
            if constexpr (sizeof(CT) == 4) {
#pragma unroll
                for (int j = 0; j < 16; ++j) co[s][j] = double(float(co[s][j]));
            }
        }
    } else {
#pragma unroll
        for (int s = 0; s < NW; ++s)
            if (s < width) load_coef_row<CT>(coef, row[s] >= 0 ? row[s] : 0, co[s]);
    }
    double aT[NW], rT = 0.0, DT = 1.0, UT = 0.0;
    if (sa.on) {
#pragma unroll
        for (int s = 0; s < NW; ++s) aT[s] = s < width ? sa.mat[row[s] >= 0 ? row[s] : 0] : 0.0;
        rT = sa.res[cc];
        DT = sa.spec[cc];
        UT = sa.dflux[cc];
    }

    // wait for the chunks this one depends on, one flag per lane
    const int d0 = fo.depPtr[tk], d1 = fo.depPtr[tk + 1];
    for (int i0 = d0; i0 < d1; i0 += 64) {
        const int i = i0 + lane;
        bool ok = i >= d1;
        const int dc = ok ? 0 : fo.dep[i];
        for (int it = 1;; ++it) {
            if (!ok) ok = flag_load(&chunkFlag[dc]) == epoch;
            if (__all(ok)) break;
            // bounded wait: never reached by a correct schedule; a wave that
            // gives up makes every later waiter give up too
            if (wall_clock64() - t0 > waitTicks || ((it & 63) == 0 && flag_load(&ctl[2]) != 0)) return false;
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);  // the dq loads stay after the flag loads
    double dqn[NW][5], dT[NW];
#pragma unroll
    for (int s = 0; s < NW; ++s) {
        dT[s] = 0.0;
        if (s >= width) break;
        const int i = row[s] >= 0 ? nb[s] : 0;
        // the neighbour's row with device-scope (sc1) vector loads: three
        // requests to one line instead of five 8-byte loads; with the SA
        // sweep carried, the row's slot 5 holds the neighbour's nu~ increment
        {
            const __amdgpu_buffer_rsrc_t dqr =
                __builtin_amdgcn_make_buffer_rsrc(dq, 0, (int)((size_t)nT * kRow * sizeof(double)), kRawBufferWord3);
            const v4i a = __builtin_amdgcn_raw_buffer_load_b128(dqr, i * kRow * 8, 0, kDeviceScope);
            const v4i b = __builtin_amdgcn_raw_buffer_load_b128(dqr, i * kRow * 8 + 16, 0, kDeviceScope);
            const double2 da = __builtin_bit_cast(double2, a), db = __builtin_bit_cast(double2, b);
            dqn[s][0] = da.x; dqn[s][1] = da.y; dqn[s][2] = db.x; dqn[s][3] = db.y;
            if (sa.on) {
                const double2 dc = __builtin_bit_cast(
                    double2, __builtin_amdgcn_raw_buffer_load_b128(dqr, i * kRow * 8 + 32, 0, kDeviceScope));
                dqn[s][4] = dc.x;
                dT[s] = dc.y;
            } else {
                dqn[s][4] = __builtin_bit_cast(
                    double, __builtin_amdgcn_raw_buffer_load_b64(dqr, i * kRow * 8 + 32, 0, kDeviceScope));
            }
        }
    }
    double rhs0[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < NW; ++s) {
        if (s >= width) break;
        double df[5];
        sweep_apply<VIS>(K, co[s], dqn[s], df);
        const bool act = row[s] >= 0;
#pragma unroll
        for (int m = 0; m < 5; ++m) rhs0[m] += act ? df[m] : 0.0;
    }
    double rhsT = 0.0;
    if (sa.on) {
#pragma unroll
        for (int s = 0; s < NW; ++s)
            if (s < width) rhsT += (row[s] >= 0) ? aT[s] * dT[s] : 0.0;
    }
    if (live) {
        // the cell's row with device-scope (sc1) vector stores
        double v[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) v[m] = (r[m] - U[m] - rhs0[m]) / D;
        const __amdgpu_buffer_rsrc_t dqw =
            __builtin_amdgcn_make_buffer_rsrc(dq, 0, (int)((size_t)nT * kRow * sizeof(double)), kRawBufferWord3);
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4i, make_double2(v[0], v[1])), dqw, c * kRow * 8, 0,
                                               kDeviceScope);
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4i, make_double2(v[2], v[3])), dqw,
                                               c * kRow * 8 + 16, 0, kDeviceScope);
        if (sa.on)
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4i, make_double2(v[4], (rT - UT - rhsT) / DT)),
                                                   dqw, c * kRow * 8 + 32, 0, kDeviceScope);
        else
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2i, v[4]), dqw, c * kRow * 8 + 32, 0,
                                                  kDeviceScope);
    }
    // the wave's dq stores are complete before the flag that publishes them
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (lane == 0) __hip_atomic_store(&chunkFlag[tk], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // the off-diagonal sums are read only by the next sweep (another launch)
    if (live) {
        double w[5];
#pragma unroll
        for (int m = 0; m < 5; ++m) w[m] = 0.0 + rhs0[m];
        store_row5(dflux, c, w);
        if (sa.on) {
            sa.dflux[c] = 0.0 + rhsT;
            sa.dq[c] = (rT - UT - rhsT) / DT;  // for the ghost, update and field kernels
        }
    }
    return true;
}

// TICKET false (sweep_form 0): block b sweeps chunk b; (sweep_form 3) block b
// sweeps chunk blockChunk[b], the levels' chunks split into eighths over the
// XCDs (block ids still rise with the level).  Workgroups of a
// launch are dispatched to each XCD in increasing order, so when block b is
// resident every smaller block of its XCD has been dispatched; the smallest
// unfinished chunk then always has its dependencies done and proceeds.
// TICKET true (sweep_form 2): a block of 4 wavefronts takes 4 consecutive
// chunks from a ticket counter, so a chunk is only ever claimed by a running
// block after every smaller chunk, whatever the dispatch order (the counter
// is one device-scope atomic per block, which serialises large launches).
template <int NW, int VIS, typename CT, bool TICKET>
__global__ void __launch_bounds__(TICKET ? 256 : 64) k_sweep_flow(DevMesh M, Consts K, FlowOrder fo, int *ctl,
                                                                  int *chunkFlag, bool useU,
                                                                  const double *__restrict__ coef,
                                                                  const double *__restrict__ res, const double *__restrict__ diag, double *dq,
                                                                  double *dflux, SaSweep sa,
                                                                  const GatherRec *__restrict__ rec,
                                                                  const double *__restrict__ vist)
{
    const int lane = (int)threadIdx.x & 63;
    int tk;
    if constexpr (TICKET) {
        __shared__ int ticket;
        if (threadIdx.x == 0) ticket = __hip_atomic_fetch_add(&ctl[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        tk = __builtin_amdgcn_readfirstlane(ticket * 4 + ((int)threadIdx.x >> 6));
        if (tk >= fo.nChunks) return;
    } else if (K.sweepForm == 3) {
        tk = fo.blockChunk[blockIdx.x];
        if (tk < 0) return;
    } else {
        tk = blockIdx.x;
    }
    if (!flow_chunk<NW, VIS, CT>(M, K, fo, ctl, chunkFlag, useU, coef, res, diag, dq, dflux, sa, rec, vist, tk, lane,
                                 ctl[1], wall_clock64(), (long long)ctl[3]))
        if (lane == 0) atomicAdd(&ctl[2], 1);
}

__global__ void __launch_bounds__(kBlock) k_ghost_dq(DevMesh M, Consts K, const double *__restrict__ q,
                                                     double *__restrict__ dq)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= M.nBFace) return;
    const int nT = M.nTotal;
    const int le = M.fle[f], re = M.nCell + f;
    const int bc = M.bcType[f];
    double dl[5], dr[5];
    load_row5(dq, le, dl);
    if (bc == kBC_SYMM || (bc == kBC_WALL && !K.viscous)) {
        double prim[5], Q[5];
        load5(q, nT, le, prim);
        const double g = K.gamma;
        const double v2 = prim[1] * prim[1] + prim[2] * prim[2] + prim[3] * prim[3];
        const double em = (prim[4] / prim[0]) / (g - 1.0) + 0.5 * v2;
        Q[0] = prim[0]; Q[1] = prim[0] * prim[1]; Q[2] = prim[0] * prim[2]; Q[3] = prim[0] * prim[3];
        Q[4] = prim[0] * em;
#pragma unroll
        for (int m = 0; m < 5; ++m) Q[m] += dl[m];
        double qr[5];
        load5(q, nT, re, qr);
        const FaceRec fr = M.face[f];
        const double nx = fr.nx, ny = fr.ny, nz = fr.nz;
        dr[0] = dl[0];
        double rvn = 2.0 * (nx * dl[1] + ny * dl[2] + nz * dl[3]);
        dr[1] = dl[1] - nx * rvn;
        dr[2] = dl[2] - ny * rvn;
        dr[3] = dl[3] - nz * rvn;
        rvn = 2.0 * (nx * Q[1] + ny * Q[2] + nz * Q[3]);
        const double ru = Q[1] - nx * rvn, rv = Q[2] - ny * rvn, rw = Q[3] - nz * rvn;
        const double pnew = Q[4] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) / Q[0];
        const double gam1 = g - 1.0;
        dr[4] = pnew - qr[4] / gam1 +
                0.5 * ((ru * ru + rv * rv + rw * rw) / Q[0] - qr[0] * (qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]));
    } else if (bc == kBC_WALL) {
        dr[0] = dl[0];
        dr[1] = -dl[1] + 2.0 * 0.0 * dl[0];
        dr[2] = -dl[2] + 2.0 * 0.0 * dl[0];
        dr[3] = -dl[3] + 2.0 * 0.0 * dl[0];
        dr[4] = dl[4] + 0.0 * (dr[1] - dl[1]) + 0.0 * (dr[2] - dl[2]) + 0.0 * (dr[3] - dl[3]);
    } else if (bc == kBC_EXTRAP || bc == kBC_OUTFLOW) {
#pragma unroll
        for (int m = 0; m < 5; ++m) dr[m] = dl[m];
    } else {
#pragma unroll
        for (int m = 0; m < 5; ++m) dr[m] = 0.0;
    }
    store_row5(dq, re, dr);
}

// ------------------------------------------------------------- update
__global__ void __launch_bounds__(kBlock) k_update(DevMesh M, Consts K0, const double *__restrict__ q,
                                                   const double *__restrict__ dq, double *__restrict__ qNew,
                                                   unsigned char *__restrict__ negFlag, int *__restrict__ negCount,
                                                   int *__restrict__ negList, const unsigned char *__restrict__ frozen)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nT = M.nTotal;
    if (i >= nT) return;
    const Consts K = case_k(K0, M, i);
    double prim[5];
    load5(q, nT, i, prim);
    // a converged case of a batch keeps its state
    const bool hold = K.nCases > 1 && frozen[M.caseOf[i]];
    if (i < M.nCell && !hold) {
        const double g = K.gamma;
        const double v2 = prim[1] * prim[1] + prim[2] * prim[2] + prim[3] * prim[3];
        const double em = (prim[4] / prim[0]) / (g - 1.0) + 0.5 * v2;
        double Q[5];
        Q[0] = prim[0]; Q[1] = prim[0] * prim[1]; Q[2] = prim[0] * prim[2]; Q[3] = prim[0] * prim[3];
        Q[4] = prim[0] * em;
        double d5[5];
        load_row5(dq, i, d5);
#pragma unroll
        for (int m = 0; m < 5; ++m) Q[m] += d5[m];
        if (Q[0] <= 0.0) {
            prim[0] = Q[0];
        } else {
            const double od = 1.0 / Q[0];
            const double um = Q[1] * od, vm = Q[2] * od, wm = Q[3] * od;
            const double ie = Q[4] - 0.5 * Q[0] * (um * um + vm * vm + wm * wm);
            prim[0] = Q[0]; prim[1] = um; prim[2] = vm; prim[3] = wm;
            prim[4] = (g - 1.0) * ie;
        }
        if (prim[0] < K.densLimit || prim[4] < K.presLimit) {
            negFlag[i] = 1;
            const int slot = atomicAdd(negCount, 1);
            if (slot < kFixList) negList[slot] = M.logicalOf[i];
        }
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) qNew[(size_t)m * nT + i] = prim[m];
}

// SolutionFix (:14509-14585) of stored cell c (mesh index lc): the average of
// its admissible neighbours.  A neighbour below the cell is read from the
// updated field, one above from the old field, exactly as the reference's
// in-place update loop sees them.
__device__ void fix_cell(const DevMesh &M, const Consts &K0, const double *__restrict__ qOld, double *qNew, int c,
                         int lc)
{
    const int nC = M.nCell, nT = M.nTotal;
    const Consts K = case_k(K0, M, c);
    double prim[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double vs = 0.0;
    for (int s = 0; s < M.ellW; ++s) {
        const int enc = M.c2fEnc[(size_t)s * nC + c];
        if (enc == 0) break;
        const int nb = M.c2n[(size_t)s * nC + c];
        const double *src = (mesh_index(M, nb) < lc) ? qNew : qOld;
        const double rn = src[nb], pn = src[(size_t)4 * nT + nb];
        if (rn < K.densLimit || pn < K.presLimit) continue;
        vs += 1.0;
        for (int m = 0; m < 5; ++m) {
            double v = src[(size_t)m * nT + nb];
            if (m == 0 || m == 4) v = dABS(v);
            prim[m] += v * 1.0;
        }
    }
    if (vs > 0.0) {
        const double o = 1.0 / vs;
        for (int m = 0; m < 5; ++m) prim[m] *= o;
    } else {
        for (int m = 0; m < 5; ++m) prim[m] = K.primInf[m];
    }
    for (int m = 0; m < 5; ++m) qNew[(size_t)m * nT + c] = prim[m];
}

// Sequential SolutionFix of the flagged cells in increasing mesh index (the
// reference's loop order).  k_update listed up to kFixList flagged cells by
// mesh index; one workgroup sorts that list in LDS (ranks of distinct keys)
// and its first thread fixes the cells in order.  Beyond kFixList flagged
// cells the first thread walks the mesh order instead and stops after the
// last flagged cell.  With `advance` the step counter moves on and the fix
// count is accumulated.
__global__ void __launch_bounds__(1024) k_fix(DevMesh M, Consts K0, const double *__restrict__ qOld,
                                              double *__restrict__ qNew, unsigned char *__restrict__ negFlag,
                                              int *__restrict__ negCount, const int *__restrict__ negList,
                                              int *__restrict__ negTotal, int *__restrict__ stepCtr, int advance)
{
    __shared__ int keys[kFixList], order[kFixList];
    const int count = *negCount;  // the same value in every thread
    const bool listed = count > 0 && count <= kFixList;
    if (listed) {
        for (int i = threadIdx.x; i < count; i += blockDim.x) keys[i] = negList[i];
        __syncthreads();
        for (int i = threadIdx.x; i < count; i += blockDim.x) {
            const int v = keys[i];
            int r = 0;
            for (int j = 0; j < count; ++j) r += keys[j] < v;
            order[r] = v;
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (listed) {
        for (int k = 0; k < count; ++k) {
            const int lc = order[k], c = M.cellAt[lc];
            negFlag[c] = 0;
            fix_cell(M, K0, qOld, qNew, c, lc);
        }
    } else if (count > 0) {
        int left = count;
        for (int lc = 0; lc < M.nCell && left > 0; ++lc) {
            const int c = M.cellAt[lc];
            if (!negFlag[c]) continue;
            negFlag[c] = 0;
            --left;
            fix_cell(M, K0, qOld, qNew, c, lc);
        }
    }
    if (advance) {
        *negTotal += count;
        *negCount = 0;
        *stepCtr += 1;
    }
}

inline int nblk(int n, int b) { return (n + b - 1) / b; }

// Streaming copy used to measure achievable HBM bandwidth.
__global__ void __launch_bounds__(kBlock) k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                 size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Wall force and moment sums of AirForceCoefParts
// (NSSolverUnstruct.cpp:16026-16250): pressure and viscous stress on every
// solid-surface face, moments about the reference point.  One workgroup:
// thread t sums faces t, t+256, ... in increasing order, then a fixed tree
// reduction; the result does not depend on scheduling.  out[15] holds, in
// order, CA_f CA_p CN_f CN_p CZ_f CZ_p cpx cpy cpz Cl_f Cl_p Cn_f Cn_p Cm_f
// Cm_p, unscaled.
constexpr int kForceSums = 15;

__device__ __forceinline__ void correct_gradient(double fl, double fr, double &gx, double &gy, double &gz,
                                                 double dx, double dy, double dz, double ods)
{
    const double fold = gx * dx + gy * dy + gz * dz;
    const double fnew = (fr - fl) * ods;
    const double df = fnew - fold;
    gx += df * dx;
    gy += df * dy;
    gz += df * dz;
}

__global__ void __launch_bounds__(kBlock) k_forces(DevMesh M, Consts K, const double *__restrict__ q,
                                                   const double *__restrict__ grad,
                                                   const double *__restrict__ visl, double tx0, double ty0,
                                                   double tz0, double *__restrict__ out)
{
    __shared__ double red[kForceSums][kBlock];
    const int t = threadIdx.x;
    const int nT = M.nTotal, nC = M.nCell;
    const double poo = K.primInf[4];
    double acc[kForceSums];
#pragma unroll
    for (int j = 0; j < kForceSums; ++j) acc[j] = 0.0;
    for (int w = t; w < M.nWall; w += kBlock) {
        const int f = M.wallFaces[w];
        const int le = M.fle[f], re = nC + f;
        const FaceRec fr = M.face[f];
        const double nx = fr.nx * fr.area, ny = fr.ny * fr.area, nz = fr.nz * fr.area;
        const double cp = 2.0 * (q[(size_t)4 * nT + le] - poo);
        const double dpx = nx * cp, dpy = ny * cp, dpz = nz * cp;
        acc[1] += dpx; acc[3] += dpy; acc[5] += dpz;
        const double xc = fr.xc, yc = fr.yc, zc = fr.zc;
        if (K.viscous) {
            double g[9];  // du/d(x,y,z), dv/d(x,y,z), dw/d(x,y,z)
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                g[3 * m + 0] = grad[(size_t)(1 + m) * nC + le];
                g[3 * m + 1] = grad[(size_t)(6 + m) * nC + le];
                g[3 * m + 2] = grad[(size_t)(11 + m) * nC + le];
            }
            double dx = M.xcc[re] - M.xcc[le];
            double dy = M.ycc[re] - M.ycc[le];
            double dz = M.zcc[re] - M.zcc[le];
            const double ods = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
            dx = dx * ods; dy = dy * ods; dz = dz * ods;
#pragma unroll
            for (int m = 0; m < 3; ++m)
                correct_gradient(q[(size_t)(1 + m) * nT + le], q[(size_t)(1 + m) * nT + re], g[3 * m], g[3 * m + 1],
                                 g[3 * m + 2], dx, dy, dz, ods);
            const double vis = visl[le];
            const double c23 = 2.0 / 3.0;
            const double dudx = g[0], dudy = g[1], dudz = g[2];
            const double dvdx = g[3], dvdy = g[4], dvdz = g[5];
            const double dwdx = g[6], dwdy = g[7], dwdz = g[8];
            const double txx = vis * c23 * (2.0 * dudx - dvdy - dwdz);
            const double tyy = vis * c23 * (2.0 * dvdy - dwdz - dudx);
            const double tzz = vis * c23 * (2.0 * dwdz - dudx - dvdy);
            const double txy = vis * (dudy + dvdx);
            const double txz = vis * (dudz + dwdx);
            const double tyz = vis * (dvdz + dwdy);
            const double fvsx = -2.0 * (nx * txx + ny * txy + nz * txz) * K.oRe;
            const double fvsy = -2.0 * (nx * txy + ny * tyy + nz * tyz) * K.oRe;
            const double fvsz = -2.0 * (nx * txz + ny * tyz + nz * tzz) * K.oRe;
            acc[0] += fvsx; acc[2] += fvsy; acc[4] += fvsz;
            acc[9] += (yc - ty0) * fvsz - (zc - tz0) * fvsy;
            acc[11] += (zc - tz0) * fvsx - (xc - tx0) * fvsz;
            acc[13] += (xc - tx0) * fvsy - (yc - ty0) * fvsx;
        }
        acc[10] += (yc - ty0) * dpz - (zc - tz0) * dpy;
        acc[12] += (zc - tz0) * dpx - (xc - tx0) * dpz;
        acc[14] += (xc - tx0) * dpy - (yc - ty0) * dpx;
        acc[6] += dpx; acc[7] += dpy; acc[8] += dpz;
    }
#pragma unroll
    for (int j = 0; j < kForceSums; ++j) red[j][t] = acc[j];
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if (t < h) {
#pragma unroll
            for (int j = 0; j < kForceSums; ++j) red[j][t] += red[j][t + h];
        }
        __syncthreads();
    }
    if (t < kForceSums) out[t] = red[t][0];
}

}  // namespace

void launch_copy(const double *src, double *dst, size_t n, hipStream_t s)
{
    const size_t n2 = n / 2;
    hipLaunchKernelGGL(k_copy, dim3(256 * 32), dim3(kBlock), 0, s, (const double2 *)src, (double2 *)dst, n2);
}

// ------------------------------------------------------------ launchers
void launch_boundary(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s)
{
    if (m.nBFace == 0) return;
    hipLaunchKernelGGL(k_boundary, dim3(nblk(m.nBFace, kBlock)), dim3(kBlock), 0, s, m, k, f.q, f.rec);
}

void launch_gamma_t(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s)
{
    hipLaunchKernelGGL(k_gamma_t, dim3(nblk(m.nTotal, kBlock)), dim3(kBlock), 0, s, m, k, f.q, f.t, f.visl);
}

template <int BS>
static void launch_grad_lim_t(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s, bool fp32, bool soa)
{
    // the gather records hold the gradient type of the context's face loop
    GatherRec *rec = (fp32 == (k.mixed != 0)) ? f.rec : nullptr;
    if (k.gradient == 2)  // ggnode: face means of the node values in place
        hipLaunchKernelGGL(k_face_node_mean, dim3(nblk(m.nFace, kBlock)), dim3(kBlock), 0, s, m, f.nodeQ, f.faceQ);
    if (fp32)
        hipLaunchKernelGGL((k_grad_lim<float, BS>), dim3(nblk(m.nCell, BS)), dim3(BS), 0, s, m, k, f.q, f.gradF,
                           f.gradTF, f.lim, f.coef, f.vist, rec, soa, f.faceQ);
    else
        hipLaunchKernelGGL((k_grad_lim<double, BS>), dim3(nblk(m.nCell, BS)), dim3(BS), 0, s, m, k, f.q, f.grad,
                           f.gradT, f.lim, f.coef, f.vist, rec, soa, f.faceQ);
}

void launch_node_values(const DevMesh &m, DevFlow &f, hipStream_t s)
{
    hipLaunchKernelGGL(k_node_values, dim3(nblk(m.nNode, kBlock)), dim3(kBlock), 0, s, m, f.q, f.t, f.nodeQ);
}

void launch_grad_lim(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s, bool fp64Grad)
{
    const bool fp32 = k.mixed && !fp64Grad;
    switch (k.gradTile) {
    case 64: launch_grad_lim_t<64>(m, k, f, s, fp32, fp64Grad); break;
    case 128: launch_grad_lim_t<128>(m, k, f, s, fp32, fp64Grad); break;
    default: launch_grad_lim_t<256>(m, k, f, s, fp32, fp64Grad); break;
    }
}

// The gradients of the last residual evaluation, copied out of the gather
// records into the SoA arrays (grad[m * nC + c], m = 0..14) for a field read.
template <typename GT>
__global__ void __launch_bounds__(kBlock) k_rec_grad(int nC, const GatherRec *__restrict__ rec, GT *__restrict__ grad)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nC) return;
    const GT *g = reinterpret_cast<const GT *>(rec[c].g);
#pragma unroll
    for (int m = 0; m < 15; ++m) grad[(size_t)m * nC + c] = g[m];
}

void launch_rec_grad(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s)
{
    if (k.mixed)
        hipLaunchKernelGGL(k_rec_grad<float>, dim3(nblk(m.nCell, kBlock)), dim3(kBlock), 0, s, m.nCell, f.rec, f.gradF);
    else
        hipLaunchKernelGGL(k_rec_grad<double>, dim3(nblk(m.nCell, kBlock)), dim3(kBlock), 0, s, m.nCell, f.rec, f.grad);
}

int flux_blocks(const DevMesh &m, const Consts &k)
{
    return nblk(m.nCell, m.tileStride > 0 ? kTileCells : k.fluxTile);
}

template <bool SPEC_ONLY, bool FUSED_DT, typename T, int BS>
static void launch_flux_passes(const DevMesh &m, const Consts &k, DevFlow &f, const T *grad, const T *gradT,
                               hipStream_t s)
{
    const int nb = nblk(m.nCell, BS);
    hipLaunchKernelGGL((k_flux<SPEC_ONLY, FUSED_DT, T, BS>), dim3(nb), dim3(BS), 0, s, m, k, f.q, grad, gradT, f.lim,
                       f.res, f.specInv, f.specVis, f.normPart, f.cflArr, f.stepCtr, f.dtLoc, f.dtMinPart, f.vist,
                       f.rec);
}

template <bool FUSED_DT, int BS>
static void launch_flux_t(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s)
{
    if (m.tileStride > 0) {
        const size_t lds = std::max((size_t)m.tileStride * 6 * sizeof(double) * (k.viscous ? 2 : 1),
                                    (size_t)6 * kTileCells * sizeof(double));
        const int nb = nblk(m.nCell, kTileCells);
        if (k.mixed)
            hipLaunchKernelGGL((k_flux_tile<float, kTileCells>), dim3(nb), dim3(kTileCells), lds, s, m, k, f.rec,
                               f.gradTF, f.vist, m.tileFacePtr, m.tileFaces, m.c2loc, m.tileStride, f.q, f.res,
                               f.specInv, f.specVis, f.normPart, f.cflArr, f.stepCtr, f.dtLoc, f.dtMinPart, FUSED_DT);
        else
            hipLaunchKernelGGL((k_flux_tile<double, kTileCells>), dim3(nb), dim3(kTileCells), lds, s, m, k, f.rec,
                               f.gradT, f.vist, m.tileFacePtr, m.tileFaces, m.c2loc, m.tileStride, f.q, f.res,
                               f.specInv, f.specVis, f.normPart, f.cflArr, f.stepCtr, f.dtLoc, f.dtMinPart, FUSED_DT);
        return;
    }
    const int nbf = nblk(m.nFace, 256);
    if (k.mixed)
        hipLaunchKernelGGL((k_face_flux<float, 3>), dim3(nbf), dim3(256), 0, s, m, k, f.rec, f.gradTF, f.vist, f.faceInv,
                           f.faceVis);
    else
        hipLaunchKernelGGL((k_face_flux<double, 3>), dim3(nbf), dim3(256), 0, s, m, k, f.rec, f.gradT, f.vist, f.faceInv,
                           f.faceVis);
    if (k.mixed)  // the FP32 face kernel stored compact values
        hipLaunchKernelGGL((k_cell_gather<BS, true>), dim3(nblk(m.nCell, BS)), dim3(BS), 0, s, m, k, f.q, f.faceInv,
                           f.faceVis, f.res, f.specInv, f.specVis, f.normPart, f.cflArr, f.stepCtr, f.dtLoc,
                           f.dtMinPart, FUSED_DT);
    else
        hipLaunchKernelGGL((k_cell_gather<BS, false>), dim3(nblk(m.nCell, BS)), dim3(BS), 0, s, m, k, f.q, f.faceInv,
                           f.faceVis, f.res, f.specInv, f.specVis, f.normPart, f.cflArr, f.stepCtr, f.dtLoc,
                           f.dtMinPart, FUSED_DT);
}

void launch_flux(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s, bool fusedDt)
{
    if (!fusedDt) return launch_flux_t<false, 256>(m, k, f, s);  // the standalone residual hook
    switch (k.fluxTile) {
    case 64: launch_flux_t<true, 64>(m, k, f, s); break;
    case 128: launch_flux_t<true, 128>(m, k, f, s); break;
    default: launch_flux_t<true, 256>(m, k, f, s); break;
    }
}

void launch_pressure_factor(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s)
{
    if (!f.rtem || (k.entropyFix != 2 && k.entropyFix != 6) || m.nCell == 0) return;
    hipLaunchKernelGGL(k_pressure_factor, dim3(nblk(m.nCell, kBlock)), dim3(kBlock), 0, s, m, k.entropyFix, f.q,
                       f.rtem);
    if (m.nBFace > 0)
        hipLaunchKernelGGL(k_ghost_copy, dim3(nblk(m.nBFace, kBlock)), dim3(kBlock), 0, s, m, f.rtem);
}

void launch_load_residuals(const DevMesh &m, DevFlow &f, hipStream_t s)
{
    const int n = 5 * m.nCell;
    hipLaunchKernelGGL(k_load_res, dim3(nblk(n, kBlock)), dim3(kBlock), 0, s, f.res, n);
}

template <int PARTS>
static void launch_face_pass_t(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s)
{
    const int nbf = nblk(m.nFace, 256), nbc = nblk(m.nCell, kBlock);
    double *fv = PARTS == 1 ? f.faceInv : f.faceVis;
    if (k.mixed) {
        hipLaunchKernelGGL((k_face_flux<float, PARTS>), dim3(nbf), dim3(256), 0, s, m, k, f.rec, f.gradTF, f.vist,
                           f.faceInv, f.faceVis);
        hipLaunchKernelGGL(k_cell_add<true>, dim3(nbc), dim3(kBlock), 0, s, m, fv, f.res);
    } else {
        hipLaunchKernelGGL((k_face_flux<double, PARTS>), dim3(nbf), dim3(256), 0, s, m, k, f.rec, f.gradT, f.vist,
                           f.faceInv, f.faceVis);
        hipLaunchKernelGGL(k_cell_add<false>, dim3(nbc), dim3(kBlock), 0, s, m, fv, f.res);
    }
}

void launch_face_pass(const DevMesh &m, const Consts &k, DevFlow &f, int parts, hipStream_t s)
{
    if (parts & 1) launch_face_pass_t<1>(m, k, f, s);
    if ((parts & 2) && k.viscous) launch_face_pass_t<2>(m, k, f, s);
}

void launch_residual_norm(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s)
{
    if (k.nCases > 1) return launch_norm_final(m, k, f, false, s);  // per-case sums read res directly
    const int nb = nblk(m.nCell, 256);
    hipLaunchKernelGGL(k_res_partial, dim3(nb), dim3(256), 0, s, m.nCell, (size_t)m.nCell, f.res, f.normPart);
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(kBlock), 0, s, f.normPart, f.dtMinPart, nb, f.normOut, f.stepCtr,
                       f.dtMin, 0);
}

void launch_spectrum(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s)
{
    launch_flux_passes<true, false, double, 256>(m, k, f, f.grad, f.gradT, s);
}

void launch_norm_final(const DevMesh &m, const Consts &k, DevFlow &f, bool withDt, hipStream_t s)
{
    if (k.nCases > 1) {
        hipLaunchKernelGGL(k_case_reduce, dim3(k.nCases * kCaseSlices), dim3(kBlock), 0, s, m, f.res, f.dtLoc,
                           f.casePart);
        hipLaunchKernelGGL(k_case_final, dim3(nblk(k.nCases, 64)), dim3(64), 0, s, f.casePart, k.nCases, f.normOut,
                           f.stepCtr, f.dtMin);
        return;
    }
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(kBlock), 0, s, f.normPart, f.dtMinPart,
                       withDt ? flux_blocks(m, k) : nblk(m.nCell, 256), f.normOut,
                       f.stepCtr, f.dtMin, withDt ? 1 : 0);
}

void launch_time_step(const DevMesh &m, const Consts &k, DevFlow &f, double cfl, hipStream_t s)
{
    const int nb = nblk(m.nCell, kBlock);
    hipLaunchKernelGGL(k_time_step, dim3(nb), dim3(kBlock), 0, s, m, k, f.q, f.specInv, f.specVis, cfl, f.dtLoc,
                       f.dtMinPart);
    hipLaunchKernelGGL(k_min_final, dim3(1), dim3(kBlock), 0, s, f.dtMinPart, nb, f.dtMin);
}

void launch_dt_diag(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s)
{
    hipLaunchKernelGGL(k_dt_diag, dim3(nblk(m.nCell, kBlock)), dim3(kBlock), 0, s, m, k, f.dtLoc, f.specInv,
                       f.specVis, f.dtMin, f.dt, f.diag);
}

// Runs of levels with at most kSmallLevel cells are walked by one workgroup
// (barriers are cheap); larger levels are spread over the GPU.
static const int kSmallLevel = kSweepBlock;

template <int NW, int VIS, typename CT>
static void launch_sweep_so(const DevMesh &m, const Consts &k, DevFlow &f, const Levels &lv, bool useU,
                            bool withSA, hipStream_t s)
{
    if (lv.flow.nChunks > 0) {  // built unless sweep_form is 1
        SaSweep sa{};
        if (withSA) sa = SaSweep{true, f.saMat, f.resT, f.specT, f.dqT, f.dfluxT};
        hipLaunchKernelGGL(k_sweep_flow_prep, dim3(1), dim3(1), 0, s, f.sweepCtl);
        if (k.sweepForm == 2)
            hipLaunchKernelGGL((k_sweep_flow<NW, VIS, CT, true>), dim3(nblk(lv.flow.nChunks, 4)), dim3(256), 0, s, m,
                               k, lv.flow, f.sweepCtl, f.sweepFlag, useU, f.coef, f.res, f.diag, f.dq, f.dflux, sa,
                               f.rec, f.vist);
        else
            hipLaunchKernelGGL((k_sweep_flow<NW, VIS, CT, false>),
                               dim3(k.sweepForm == 3 ? lv.flow.nBlocks : lv.flow.nChunks), dim3(64), 0, s, m, k,
                               lv.flow, f.sweepCtl, f.sweepFlag, useU, f.coef, f.res, f.diag, f.dq, f.dflux, sa,
                               f.rec, f.vist);
        return;
    }
    int L = 0;
    while (L < lv.nLevels) {
        const int sz = lv.hptr[L + 1] - lv.hptr[L];
        if (sz <= kSmallLevel) {
            int L1 = L + 1;
            while (L1 < lv.nLevels && lv.hptr[L1 + 1] - lv.hptr[L1] <= kSmallLevel) ++L1;
            hipLaunchKernelGGL((k_sweep_block_so<NW, VIS, CT>), dim3(1), dim3(kSweepBlock), 0, s, m, k, lv.so, lv.ptr, L,
                               L1, useU, f.coef, f.res, f.diag, f.dq, f.dflux);
            L = L1;
        } else {
            hipLaunchKernelGGL((k_sweep_level_so<NW, VIS, CT>), dim3(nblk(sz, 64)), dim3(64), 0, s, m, k, lv.so,
                               lv.hptr[L], lv.hptr[L + 1], useU, f.coef, f.res, f.diag, f.dq, f.dflux);
            ++L;
        }
    }
}

void launch_sweep(const DevMesh &m, const Consts &k, DevFlow &f, const Levels &lv, bool forward,
                  bool firstSweep, hipStream_t s, bool withSA)
{
    const bool useU = !(forward && firstSweep);
    const bool v = k.viscous != 0;
    if (k.coefF32) {  // only set when both directions have at most 6 slots
        if (lv.so.Wpad == 3) return v ? launch_sweep_so<3, 1, float>(m, k, f, lv, useU, withSA, s) : launch_sweep_so<3, 0, float>(m, k, f, lv, useU, withSA, s);
        if (lv.so.Wpad == 4) return v ? launch_sweep_so<4, 1, float>(m, k, f, lv, useU, withSA, s) : launch_sweep_so<4, 0, float>(m, k, f, lv, useU, withSA, s);
        return v ? launch_sweep_so<6, 1, float>(m, k, f, lv, useU, withSA, s) : launch_sweep_so<6, 0, float>(m, k, f, lv, useU, withSA, s);
    }
    if (lv.so.Wpad == 3) return v ? launch_sweep_so<3, 1, double>(m, k, f, lv, useU, withSA, s) : launch_sweep_so<3, 0, double>(m, k, f, lv, useU, withSA, s);
    if (lv.so.Wpad == 4) return v ? launch_sweep_so<4, 1, double>(m, k, f, lv, useU, withSA, s) : launch_sweep_so<4, 0, double>(m, k, f, lv, useU, withSA, s);
    if (lv.so.Wpad == 6) return v ? launch_sweep_so<6, 1, double>(m, k, f, lv, useU, withSA, s) : launch_sweep_so<6, 0, double>(m, k, f, lv, useU, withSA, s);
    // wider polyhedra: face-by-face through the cell -> face lists
    int L = 0;
    while (L < lv.nLevels) {
        const int sz = lv.hptr[L + 1] - lv.hptr[L];
        if (sz <= kSmallLevel) {
            int L1 = L + 1;
            while (L1 < lv.nLevels && lv.hptr[L1 + 1] - lv.hptr[L1] <= kSmallLevel) ++L1;
            if (forward)
                hipLaunchKernelGGL(k_sweep_block<true>, dim3(1), dim3(kSweepBlock), 0, s, m, k, lv.ptr, lv.cells, L, L1,
                                   useU, f.coef, f.res, f.diag, f.dq, f.dflux);
            else
                hipLaunchKernelGGL(k_sweep_block<false>, dim3(1), dim3(kSweepBlock), 0, s, m, k, lv.ptr, lv.cells, L, L1,
                                   useU, f.coef, f.res, f.diag, f.dq, f.dflux);
            L = L1;
        } else {
            if (forward)
                hipLaunchKernelGGL(k_sweep_level<true>, dim3(nblk(sz, 64)), dim3(64), 0, s, m, k, lv.cells,
                                   lv.hptr[L], lv.hptr[L + 1], useU, f.coef, f.res, f.diag, f.dq, f.dflux);
            else
                hipLaunchKernelGGL(k_sweep_level<false>, dim3(nblk(sz, 64)), dim3(64), 0, s, m, k, lv.cells,
                                   lv.hptr[L], lv.hptr[L + 1], useU, f.coef, f.res, f.diag, f.dq, f.dflux);
            ++L;
        }
    }
}

void launch_sweep_change(const double *dq, size_t cs, size_t ms, int nv, int nC, double *prev, double *parts,
                         double *out, hipStream_t s)
{
    const int nb = nblk(nC, kBlock);
    hipLaunchKernelGGL(k_sweep_change, dim3(nb), dim3(kBlock), 0, s, dq, cs, ms, nv, nC, prev, parts);
    hipLaunchKernelGGL(k_sum_into, dim3(1), dim3(kBlock), 0, s, parts, nb, out);
}

void launch_ghost_dq(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s)
{
    if (m.nBFace == 0) return;
    hipLaunchKernelGGL(k_ghost_dq, dim3(nblk(m.nBFace, kBlock)), dim3(kBlock), 0, s, m, k, f.q, f.dq);
}

void launch_update(const DevMesh &m, const Consts &k, DevFlow &f, bool advance, hipStream_t s)
{
    hipLaunchKernelGGL(k_update, dim3(nblk(m.nTotal, kBlock)), dim3(kBlock), 0, s, m, k, f.q, f.dq, f.qAlt,
                       f.negFlag, f.negCount, f.negList, f.frozen);
    hipLaunchKernelGGL(k_fix, dim3(1), dim3(1024), 0, s, m, k, f.q, f.qAlt, f.negFlag, f.negCount, f.negList,
                       f.negTotal, f.stepCtr, advance ? 1 : 0);
}

void launch_forces(const DevMesh &m, const Consts &k, DevFlow &f, const double *torqueRef, hipStream_t s)
{
    hipLaunchKernelGGL(k_forces, dim3(1), dim3(kBlock), 0, s, m, k, f.q, f.grad, f.visl, torqueRef[0], torqueRef[1],
                       torqueRef[2], f.forceOut);
}

}  // namespace phl

// ---------------------------------------------------------------------------
// Residual sums without allocation: the block partials of the reduction live
// in this fixed device array instead of a context's buffer.  Every GPU the
// library runs on holds its own copy until the process ends; calls on one GPU
// share it, so calls that overlap on one GPU race.  Each call writes every
// partial it reads, so nothing has to be cleared between calls.
static __device__ __align__(256) double scratch[(size_t)phl::kNoAllocParts * 5];

namespace phl {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
double *g_scratchPtr[kMaxDevices] = {};
}  // namespace

// The device address of `scratch` on the current device, looked up once per
// device with that device current.  The array's name is only a host-side
// stand-in: kernels get this address.
hipError_t noalloc_scratch(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (!g_scratchPtr[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratchPtr[dev] = static_cast<double *>(ptr);
    }
    *out = g_scratchPtr[dev];
    return hipSuccess;
}

void launch_residual_norm_noalloc(const double *res, int nC, size_t ld, double *sum5, double *part, hipStream_t s)
{
    const int nb = nblk(nC, 256);
    hipLaunchKernelGGL(k_res_partial, dim3(nb), dim3(256), 0, s, nC, ld, res, part);
    hipLaunchKernelGGL(k_norm_sum5, dim3(1), dim3(kBlock), 0, s, part, nb, sum5);
}

}  // namespace phl
