// phl_internal.h -- device data layout shared by the kernels and the host
// driver.  Not part of the C-ABI.
//
// Cell-indexed arrays are in storage order (a locality order of the real
// cells, then the ghosts of boundary faces 0..nBFace-1); faces keep the
// mesh's numbering.
//
// Layout in HBM (FP64 plain form):
//   cell fields     SoA [m][nCell] or [m][nTotal] (ghosts after real cells)
//   face geometry   AoS FaceRec, 64 B per face (one gather touches one line)
//   cell->face map  ELL [W][nCell] (W = max faces per cell), faces in
//                   increasing face order so per-cell accumulation follows
//                   the reference's face-loop order exactly
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace phl {

struct FaceRec {
    double nx, ny, nz, area, xc, yc, zc;
    double lsqW;  // least-squares weight 1/(|x_R - x_L|^2 + SMALL) (gradient 1), else 0
};

// Read-only mesh data on the device.
struct DevMesh {
    int nCell, nBFace, nFace, nTotal, ellW;
    const FaceRec *face;        // [nFace]
    const int *bcType;          // [nBFace]
    const int *fle;             // [nFace] left cell
    const int *fre;             // [nFace] right cell (the ghost nCell+f of a boundary face)
    const int *c2fEnc;          // [W][nCell]: f+1 if the cell is left, -(f+1) if right, 0 pad
    const int *c2n;             // [W][nCell]: the other cell (ghost index nCell+f on boundaries)
    const double *xcc, *ycc, *zcc;  // [nTotal]
    const double *vol;          // [nTotal]
    const double *venkatVf;     // [nCell] (vol^(1/3))^3, Limiter.cpp:1126-1128
    const double *lsqInv;       // [6][nCell] least-squares inverse normal matrix xx, yy, zz, xy, xz, yz (gradient 1)
    const double *cflSkew;      // [nCell] skewness CFL factor, NSSolverUnstruct.cpp:6106-6116
    // Node-based Green-Gauss gradient (gradient 2): the contributors of each
    // node's inverse-distance average in the reference's accumulation order
    // (ComputeNodeValue, NSSolverUnstruct.cpp:4734-5081), as storage indices
    // (cell, -1) or, for a boundary face, (cell, its ghost), with their
    // weights and the weight sums added in the same order; the node lists of
    // the stored faces for the face means.
    int nNode;
    const int *nodeSrcPtr;      // [nNode+1]
    const int2 *nodeSrc;
    const double *nodeW;
    const double *nodeWsum;     // [nNode]
    const int *fnPtr, *fnList;  // [nFace+1], node lists per stored face
    const unsigned char *limZero;  // [nCell] 1 if next to a farfield-type face
    const double *limGhost;     // [nBFace] limiter value of the ghost cell
    const unsigned char *visFlag;  // [nFace] bit0: skewness correction, bit1: normal derivative
    const int *wallFaces;       // [nWall] solid-surface faces, increasing
    int nWall;
    const unsigned short *caseOf;  // [nTotal] case of each cell (batched contexts)
    const int *caseCellPtr;        // [nCases+1] cell range of each case
    const double *wallDist;        // [nCell] distance to the nearest solid surface (SA)
    // Cells are stored in a locality order; the sweeps and SolutionFix follow
    // the mesh's own numbering through these maps (ghost cells are not moved).
    const int *logicalOf;          // [nCell] mesh index of the cell stored at each position
    const int *cellAt;             // [nCell] storage position of each mesh cell
    // Tile-fused face loop (kTileCells cells per workgroup): the faces touching
    // each tile, and each cell's faces as slots of its tile's list (ELL order).
    int tileStride;                // most faces of any tile (0: the tile form is off)
    const int *tileFacePtr;        // [nTiles+1]
    const int *tileFaces;          // storage face ids, ascending within a tile
    const unsigned short *c2loc;   // [W][nCell]
};
constexpr int kTileCells = 128;
// Flagged cells of one update that SolutionFix sorts in LDS (more fall back
// to a walk over the mesh order).
constexpr int kFixList = 4096;
// Residual sums without allocation: 256-cell block partials of at most
// kNoAllocParts blocks fit the static scratch array (2.5 MiB per GPU), so
// arrays of up to 16,777,216 cells.
constexpr int kNoAllocParts = 65536;
constexpr int kNoAllocMaxCells = kNoAllocParts * 256;

// Logical block of hardware block b of nblk: workgroups are dispatched to
// the 8 XCDs round-robin (b % 8), so this gives XCD x the contiguous logical
// range of the x-th eighth; kernels that use it on storage-ordered cells keep
// a region of the mesh in one XCD's L2 from kernel to kernel.
__device__ __forceinline__ int xcd_block(int b, int nblk)
{
    const int q = nblk >> 3, r = nblk & 7;
    const int x = b & 7, i = b >> 3;
    return x * q + (x < r ? x : r) + i;
}

// Mesh index of a stored cell (ghost indices are their own: above every real cell).
__device__ __forceinline__ int mesh_index(const DevMesh &M, int i) { return i < M.nCell ? M.logicalOf[i] : i; }

// The scalars that depend on a case's freestream (Mach, angles) and mesh
// (mean volume).  A batched context holds one per case.
struct CaseConsts {
    double R, mach;
    double primInf[5];
    double entropyCoef, epsTmp, pLower, pUpper, cp, densLimit, presLimit;
};

// Scalars of the path, fixed per context.
struct Consts {
    double gamma, R, mach, reynolds, oRe;
    double primInf[5];
    double entropyCoef;     // RoeEntropyFixCoef1 == Coef2
    double epsTmp;          // K^3 / (avgVol^(1/3))^3
    double pLower, pUpper;  // 0.001, 0.02 of pTotal
    double cflStart, ktmax;
    double tsuth;           // 110.4 / T_ref
    double cp;              // 1/((gamma-1) M^2)
    double oprl, oprt;
    double densLimit, presLimit;  // 1e-6 freestream
    int viscous, limiter, entropyFix, nSweeps;
    int mixed;              // 1: FP32 gradients and face fluxes, FP64 accumulation
    int nCases;             // > 1: batched context, per-case scalars in cases[caseOf[cell]]
    int fluxTile, gradTile; // cells per workgroup of the flux and gradient kernels (64, 128, 256)
    int coefF32;            // mixed form: LU-SGS face coefficients stored in FP32 (sweep arithmetic stays FP64)
    int coefInSweep;        // dataflow sweeps evaluate the face coefficients themselves (none stored)
    int gradient;           // 0 Green-Gauss cell, 1 least squares, 2 Green-Gauss node
    int sweepForm;          // 0, 2, 3: dataflow sweep (launch order, ticket order, XCD groups); 1: per level
    int initPhase;          // isInIniting: first-order faces, Van Leer flux, limiter 0 (flowInitStep)
    double saNuInf, saVistInf, turbCflScale;  // SA: freestream nu~ and eddy viscosity, turbCFLScale
    double lusgsTol;        // LUSGSTolerance (nSweeps > 1)
    double twall;           // wallTemperature / refDimensionalTemperature; <= 0: adiabatic walls
    const CaseConsts *cases;
    const double *rtem;     // [nTotal] pressure factor of roeEntropyFixMethod 2 and 6, else null
};

// Host side of an LU-SGS call with several sweep pairs: a pinned buffer for
// the two half-sweep sums, and what the last call did.
struct SweepLog {
    double *host = nullptr;   // pinned [2]
    double normals[8] = {};   // squared dq change of the first 8 pairs
    int done = 0;             // pairs run by the last call
};

// What the face loop gathers of one cell, in one 192-byte record (two cache
// lines instead of one per scattered SoA value): state, centroid, then the
// gradients (FP64, or FP32 in the mixed form, filling the first 60 bytes of
// the slot), then the limiter.  Real cells are refreshed by the gradient
// kernel, ghost states by the boundary kernel; centroids, ghost gradients
// (zero) and ghost limiters never change.
struct alignas(64) GatherRec {
    double q[5];
    double x, y, z;
    double g[15];
    double lim;
};
static_assert(sizeof(GatherRec) == 192, "GatherRec is three 64-byte sectors");

// The LU-SGS increments dq and off-diagonal sums are per-cell rows of kRow
// doubles (one 64-byte line): a dataflow sweep reads a neighbour's five
// increments with two 16-byte and one 8-byte device-scope load from one
// line, and publishes its own with three such stores.
constexpr int kRow = 8;

// Mutable flow state on the device.
struct DevFlow {
    double *q;        // [5][nTotal] primitive variables (current)
    double *qAlt;     // [5][nTotal] update target, swapped with q
    double *res;      // [5][nCell]
    double *grad;     // [15][nCell]: d/dx of m at [m], d/dy at [5+m], d/dz at [10+m]
    double *gradT;    // [3][nCell] (viscous)
    float *gradF;     // [15][nCell] mixed form
    float *gradTF;    // [3][nCell] mixed form
    double *lim;      // [nTotal]
    double *rtem;     // [nTotal] pressure factor (entropy fix 2, 6; null otherwise)
    double *dq;       // [nTotal][kRow]: the five increments of a cell in one 64-byte line
    double *dflux;    // [nCell][kRow]: the off-diagonal sums of the last sweep
    double *casePart; // batched contexts: [nCases][32][6] partial residual sums and dt minima
    int *sweepFlag;   // [chunks] dataflow sweep: epoch of the sweep that last published the chunk's dq
    int *sweepCtl;    // [4] dataflow sweep: ticket, epoch, waves that gave up waiting
    double *specInv, *specVis, *dtLoc, *dt, *diag;  // [nCell]
    double *normPart;   // [nBlocksFlux][5]
    double *normOut;    // [maxSteps][nCases][5], row = *stepCtr
    double *dtMinPart;  // [nBlocksFlux] block minima of the local time step
    double *dtMin;      // [nCases] minimum local time step (ComputeMinTimeStep)
    int *stepCtr;       // step index within the current chunk (device side)
    double *cflArr;     // [maxSteps] CFL of each step of the chunk
    unsigned char *negFlag;  // [nCell]
    int *negCount;      // SolutionFix count of the last update
    int *negList;       // [kFixList] mesh indices of the flagged cells (the first kFixList of them)
    int *negTotal;      // running SolutionFix count
    double *coef;       // [nFace][2][16] LU-SGS face coefficients per side
    double *t, *visl;   // [nTotal], only for the standalone gamma/T hook
    double *forceOut;   // [15] wall force and moment sums
    unsigned char *frozen;  // [nCases] batched contexts: 1 once a case has converged
    double *vist;     // [nTotal] eddy viscosity (0 unless SA)
    double *nodeQ;    // gradient 2: [6][nNode] node values of the five primitives and T
    double *faceQ;    // gradient 2: [6][nFace] face means of the node values
    GatherRec *rec;   // [nTotal] face-loop gather records
    double *faceInv;  // [nFace][6] inviscid flux (area-scaled) and spectral-radius term of each face
    double *faceVis;  // [nFace][6] viscous flux and spectral-radius term (viscous flow only)
    // Spalart-Allmaras (viscous == 3), allocated only then
    double *nut;      // [nTotal] nu~ (non-conservative)
    double *resT;     // [nTotal] residual of nu~ (real cells written)
    double *specT;    // [nCell] diagonal of the scalar LU-SGS
    double *saMat;    // [2*nFace] face matrices: [2f] used by the left cell, [2f+1] by the right
    double *dqT;      // [nTotal]
    double *dfluxT;   // [nCell]
    double *gradNut;  // [3][nCell]
    // nLUSGSSweeps > 1, allocated only then: dq after the last half-sweep
    // (NS [5][nCell], SA [nCell]), block partials and the two half-sweep sums
    double *dqPrev, *dqTPrev, *sweepParts, *sweepNorm;
};

// LU-SGS level schedule.
// Sweep order of one direction, position-major: position p of the level
// order holds cell cells[p] and its active faces (lower neighbours forward,
// higher ones and ghosts backward) in increasing face order.
struct SweepOrder {
    int W;              // most active faces of any cell
    int Wpad;           // slots stored: 3, 4 or 6 (>= W), or W when W > 6
    const int *cells;   // [nCell]
    const int *row;     // [Wpad][nCell]: 2*face + coefficient side, -1 if unused
    const int *nb;      // [Wpad][nCell]: neighbour cell (0 if unused)
};

// Dataflow sweep order: the level order with every level padded to whole
// chunks of 64 positions (a chunk never spans two levels), so a chunk
// depends only on chunks of earlier levels.
struct FlowOrder {
    int nChunks;        // 0: not built
    const int *cells;   // [64*nChunks] storage index, -1 for padding
    const int *row;     // [Wpad][64*nChunks] as SweepOrder::row
    const int *nb;      // [Wpad][64*nChunks] as SweepOrder::nb
    const int *depPtr;  // [nChunks+1]
    const int *dep;     // the distinct earlier chunks holding a chunk's real neighbours
    const int *width;   // [nChunks] most active slots of any cell of the chunk (slots are packed first)
    // sweep_form 3: block i sweeps chunk blockChunk[i] (-1: nothing).  Each
    // level's chunks are split into 8 contiguous groups, group x going to the
    // blocks i = x (mod 8) of the level's block range, so that the k-th XCD
    // sweeps the k-th eighth (in storage order) of every level.
    const int *blockChunk;
    int nBlocks;
};

// The scalar Spalart-Allmaras sweep carried along by a dataflow NS sweep.
struct SaSweep {
    bool on;
    const double *mat, *res, *spec;
    double *dq, *dflux;
};

struct Levels {
    int nLevels;
    int *ptr;    // device [nLevels+1]
    int *cells;  // device [nCell]
    int *hptr;   // host copy of ptr
    SweepOrder so;
    FlowOrder flow;
};

// Launchers (phl_kernels.hip).  All are asynchronous on `s`.
void launch_boundary(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s);
void launch_gamma_t(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s);
// The gradients go to the gather records (the face loop and the SA source
// read them there).  fp64Grad: also store FP64 SoA gradients, even in the
// mixed form (the force integration).
void launch_grad_lim(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s, bool fp64Grad = false);
// Node values of the node-based gradient (ComputeNodeValue +
// ModifyNodeValue): from the current q and the temperature in f.t.
void launch_node_values(const DevMesh &m, DevFlow &f, hipStream_t s);
// Copy the gradients of the gather records to the SoA arrays of the
// context's precision (grad, or gradF in the mixed form).
void launch_rec_grad(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s);
void launch_flux(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s, bool fusedDt);
void launch_spectrum(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s);
// Hook form of the residual (Controller::UpdateResiduals split as the
// reference calls it): res = -0.0 (LoadResiduals), then face passes that add
// their face values to res (parts bit 0: inviscid, bit 1: viscous), then the
// residual sums of the current res into normOut row *stepCtr.
// Controller::PressureFactor: the cell sensor of entropy fixes 2 and 6.
void launch_pressure_factor(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s);
void launch_load_residuals(const DevMesh &m, DevFlow &f, hipStream_t s);
void launch_face_pass(const DevMesh &m, const Consts &k, DevFlow &f, int parts, hipStream_t s);
void launch_residual_norm(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s);
// Device address of the no-allocation scratch array on the current device.
hipError_t noalloc_scratch(double **out);
void launch_residual_norm_noalloc(const double *res, int nC, size_t ld, double *sum5, double *part, hipStream_t s);
// Workgroups of the fused flux launch (one partial-sum row each).
int flux_blocks(const DevMesh &m, const Consts &k);
void launch_norm_final(const DevMesh &m, const Consts &k, DevFlow &f, bool withDt, hipStream_t s);
void launch_time_step(const DevMesh &m, const Consts &k, DevFlow &f, double cfl, hipStream_t s);
void launch_dt_diag(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s);
// withSA: a dataflow sweep (flow order built) also takes the same
// direction of the pending Spalart-Allmaras sweep (launch_sa_sweep_init
// before the forward one, launch_sa_ghost_dq between, launch_sa_update after).
void launch_sweep(const DevMesh &m, const Consts &k, DevFlow &f, const Levels &lv, bool forward,
                  bool firstSweep, hipStream_t s, bool withSA = false);
void launch_ghost_dq(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s);
void launch_update(const DevMesh &m, const Consts &k, DevFlow &f, bool advance, hipStream_t s);
void launch_copy(const double *src, double *dst, size_t n, hipStream_t s);
void launch_forces(const DevMesh &m, const Consts &k, DevFlow &f, const double *torqueRef, hipStream_t s);

// Spalart-Allmaras (phl_sa.hip).  tri: [nTri][9] wall triangles (face centre
// and two consecutive face nodes).
void launch_sweep_change(const double *dq, size_t cs, size_t ms, int nv, int nC, double *prev, double *parts,
                         double *out, hipStream_t s);
void launch_wall_distance(const double *xcc, const double *ycc, const double *zcc, int nCell, const double *tri,
                          int nTri, bool edges, double *wd, hipStream_t s);
void launch_sa_boundary(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s);
void launch_sa_viscosity(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s);
void launch_sa_gradient(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s);
void launch_sa_residual(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s);
void launch_sa_spectrum(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s);
void launch_sa_sweep_init(const DevMesh &m, DevFlow &f, hipStream_t s);
void launch_sa_ghost_dq(const DevMesh &m, DevFlow &f, hipStream_t s);
void launch_sa_update(const DevMesh &m, DevFlow &f, hipStream_t s);
// One sweep pair (k.nSweeps == 1), or pairs until the squared change of a
// pair falls below k.lusgsTol times the first's (at most k.nSweeps; the
// stream is synchronised after each pair and `log` records the sums).
void launch_sa_lusgs(const DevMesh &m, const Consts &k, DevFlow &f, const Levels &fwd, const Levels &bwd,
                     hipStream_t s, SweepLog *log = nullptr);
// Everything of a relaxation before the sweeps: boundary, eddy viscosity,
// gradient, residual, spectrum and face matrices.  launch_sa_lusgs is the rest.
void launch_sa_pre(const DevMesh &m, const Consts &k, DevFlow &f, hipStream_t s);
void launch_sa_relax(const DevMesh &m, const Consts &k, DevFlow &f, const Levels &fwd, const Levels &bwd,
                     hipStream_t s, SweepLog *log = nullptr);

}  // namespace phl
