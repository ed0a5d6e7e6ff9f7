/*
 * phl.h -- C-ABI of phenlei_amd, the MI355X (gfx950) implementation of the
 * PHengLEI unstructured compressible residual + LU-SGS path.
 *
 * The entry points mirror the granularity of the reference's GPU call
 * surface (the CallGPU* hooks that NSSolverUnstruct early-returns into under
 * USE_CUDA, and the lifecycle calls of GPU/Manage).  Each declaration cites
 * the reference interface it replaces.  Plain pointers and sizes only.
 *
 * Arrays exchanged with the host are structure-of-arrays over the
 * reference's cell numbering: five-component fields are [5][nTotal] with
 * nTotal = nCell + nBFace, the ghost cell of boundary face f being nCell+f
 * (API/Geometry/src/Geo_UnstructGrid.cpp:2007).
 *
 * Errors: every call returns a phl_status; nothing calls exit().  The
 * reference instead prints and exits (GPU/Manage/src/cudaErrorHandle.cu:4-26).
 * phl_last_error() returns a thread-local message for the last failure.
 */
#ifndef PHL_H
#define PHL_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum phl_status {
    PHL_OK = 0,
    PHL_ERR_ARG = 1,         /* bad argument or inconsistent sizes */
    PHL_ERR_HIP = 2,         /* HIP runtime error */
    PHL_ERR_NUMERIC = 3,     /* NaN/Inf residual (reference: Residual.cpp:579-589 exits) */
    PHL_ERR_UNSUPPORTED = 4, /* option outside the implemented path */
    PHL_ERR_ALLOC = 5        /* host or device allocation failed */
} phl_status;

/* Solver parameters: the cfd_para.hypara keys on this path
 * (PHengLEI/examples/bin/cfd_para.hypara:327-612,1143-1145). */
typedef struct phl_params {
    double mach;            /* refMachNumber */
    double alpha_deg;       /* attackd */
    double beta_deg;        /* angleSlide */
    double gamma;           /* refGama (1.4) */
    double reynolds;        /* refReNumber per unit grid length */
    double t_ref;           /* refDimensionalTemperature [K] (288.15) */
    double prandtl;         /* prl (0.72) */
    double prandtl_turb;    /* prt (0.9) */
    double t_wall;          /* wallTemperature [K]; <= 0 adiabatic (-1), > 0 isothermal viscous walls */
    double venkat_k;        /* venkatCoeff (5.0) */
    double entropy_scale;   /* roeEntropyScale (1.0) */
    double cfl_start;       /* CFLStart (0.01) */
    double cfl_end;         /* CFLEnd (10.0) */
    double ktmax;           /* ktmax (1e10) */
    double lusgs_tol;       /* LUSGSTolerance (0.01) */
    double force_ref_length;/* forceReferenceLength (1.0) */
    double force_ref_area;  /* forceReferenceArea (1.0) */
    double force_ref_x, force_ref_y, force_ref_z; /* TorqueRefX/Y/Z (0) */
    double force_ref_span;  /* forceReferenceLengthSpanWise (1.0) */
    double sa_nu_inf;       /* muoo: freestream nu~ of the SA model (3.0) */
    double sa_vist_inf;     /* freeStreamViscosity: eddy viscosity of inflow ghosts (1e-3) */
    double turb_cfl_scale;  /* turbCFLScale (1.0) */
    double turb_skew_angle; /* turbSkewnessAngle, degrees (0.0) */
    int viscous;            /* viscousType: 0 inviscid, 1 laminar, 3 Spalart-Allmaras (1eq-sa) */
    int gradient;           /* gradientName: 0 "ggcell" (ghost gradient 0), 1 "lsq" (weights 1/|dx|^2,
                             * ghost takes its owner's gradient), 2 "ggnode" (the reference default,
                             * cfd_para.hypara:494: face values from inverse-distance node averages,
                             * ghost takes its owner's gradient); with the SA model the gradient of nu~
                             * (hook and field only; the NSMB source reads none) is ggcell only */
    int limiter;            /* uns_limiter_name: 0 "vencat" (ivencat=7), 1 "nolim", 2 "1st" */
    int entropy_fix;        /* roeEntropyFixMethod (3): 1 to 6, Flux_RoeEntropyCorrection.cpp:93-166 */
    int cfl_vary_step;      /* CFLVaryStep (500) */
    int n_lusgs_sweeps;     /* nLUSGSSweeps (1) */
    int precision;          /* 0 FP64; 1 mixed (FP32 gradients and face fluxes, FP64 accumulation/LU-SGS) */
    int storage_order;      /* device storage of the cells: 0 locality order (cells sorted along a
                             * Morton curve of their centroids, per case; default), 1 the mesh's order.
                             * Results do not depend on it: every sweep and accumulation follows the
                             * mesh's cell and face numbering either way. */
    int face_loop;          /* 0 (default): face kernel + cell gather, face values through HBM; 1:
                             * tile-fused (a workgroup evaluates the faces of its cells into LDS and sums
                             * them; faces between tiles are evaluated twice).  Identical results; the
                             * tile form measured slower on the hybrid meshes (1M-cell hybrid bump:
                             * 0.39 vs 0.32 ms). */
    int sweep_form;         /* LU-SGS sweeps (identical results: same arithmetic and summation order):
                             * 0 (default) dataflow, one launch per direction: chunks of 64 cells of
                             * the level order, each wavefront waiting on the flags of the chunks that
                             * hold its cells' swept neighbours, chunks taken in launch order;
                             * 1 one launch per dependency level;
                             * 2 dataflow with chunks taken from a ticket counter (independent of the
                             * workgroup dispatch order; slower on large meshes);
                             * 3 dataflow in launch order with each level's chunks split into eighths
                             * (storage order), the k-th eighth on the blocks of the k-th XCD.
                             * The dataflow waits are bounded (0.2 s, phl_set_sweep_wait): a wait that
                             * runs out makes phl_iterate/phl_lusgs fail with PHL_ERR_NUMERIC. */
    int flow_init_step;     /* flowInitStep: outer iterations 1..flow_init_step run first order with
                             * the Van Leer flux (isInIniting; Simulation.cpp:1190-1236, flowInitMethod 0).
                             * The reference's default is 100; phl_params_default sets 0. */
} phl_params;

/* Field identifiers for phl_get_field (same numbering as the oracle). */
enum {
    PHL_FIELD_Q = 0, PHL_FIELD_RES = 1, PHL_FIELD_DQDX = 2, PHL_FIELD_DQDY = 3,
    PHL_FIELD_DQDZ = 4, PHL_FIELD_LIMIT = 5, PHL_FIELD_DQ = 6, PHL_FIELD_SPEC_INV = 7,
    PHL_FIELD_SPEC_VIS = 8, PHL_FIELD_DT = 9, PHL_FIELD_DIAG = 10, PHL_FIELD_T = 11,
    PHL_FIELD_VISL = 12, PHL_FIELD_DTDX = 13, PHL_FIELD_DTDY = 14, PHL_FIELD_DTDZ = 15,
    PHL_FIELD_DELTA_FLUX = 16,
    /* Spalart-Allmaras (viscous = 3): eddy viscosity, nu~, its residual,
     * LU-SGS diagonal and increment, wall distance, gradient of nu~ */
    PHL_FIELD_VIST = 17, PHL_FIELD_NUT = 18, PHL_FIELD_RES_T = 19, PHL_FIELD_SPEC_T = 20,
    PHL_FIELD_DQ_T = 21, PHL_FIELD_WDIST = 22, PHL_FIELD_DNUDX = 23, PHL_FIELD_DNUDY = 24,
    PHL_FIELD_DNUDZ = 25,
    /* the pressure factor of roeEntropyFixMethod 2 and 6 (rtem, [nTotal];
     * zeros for the other methods) */
    PHL_FIELD_PRESSURE_FACTOR = 26
};

/* Boundary-condition codes (API/Geometry/include/Geo_SimpleBC.h:32-39). */
enum { PHL_BC_EXTRAPOLATION = 1, PHL_BC_SOLID_SURFACE = 2, PHL_BC_SYMMETRY = 3,
       PHL_BC_FARFIELD = 4, PHL_BC_INFLOW = 5, PHL_BC_OUTFLOW = 6 };

typedef struct phl_mesh phl_mesh; /* host-side unstructured grid + metrics */
typedef struct phl_ctx phl_ctx;   /* device-resident solver state for one zone */

const char *phl_last_error(void);
void phl_params_default(phl_params *p);

/* ---------------- meshes (host) ----------------------------------------
 * Replaces the grid read + UnstructGrid::ComputeMetrics step
 * (Main/src/Simulation.cpp:205-247; Geo_UnstructGrid.cpp:5695,5911,6545). */

/* Build from explicit connectivity.  Faces 0..nBFace-1 are boundary faces
 * (right[f] must be nCell+f), face normals point from left to right. */
phl_status phl_mesh_create(int nNode, const double *x, const double *y, const double *z,
                           int nFace, const int *faceNodePtr, const int *faceNodes,
                           const int *left, const int *right, int nBFace, const int *bcType,
                           int nCell, const int *cellNodePtr, const int *cellNodes,
                           phl_mesh **out);
/* Channel x in [-1.5,1.5], wall-fitted between y_w = h*exp(-25x^2) and 0.8,
 * z in [0,0.2]; nx*ny*nz hexahedra.  Inlet/outlet FARFIELD, lower wall
 * SOLID_SURFACE, top and sides SYMMETRY. */
phl_status phl_mesh_bump_hex(int nx, int ny, int nz, double bump, phl_mesh **out);
/* The same channel as a hybrid grid: n_prism prism layers over the
 * triangulated wall (first height h1, geometric growth up to the tet row
 * height) and n_tet rows of a seeded, jittered point lattice split into six
 * tetrahedra per lattice cell. */
phl_status phl_mesh_bump_hybrid(int nx, int nz, int n_prism, double h1, int n_tet, double jitter,
                                unsigned seed, double bump, phl_mesh **out);
/* Semi-span swept wing for RANS: root chord 1 on the symmetry plane z = 0,
 * taper 0.5, aspect ratio 4, quarter-chord sweep 25 deg, symmetric cubic
 * section (t/c 0.12, half-thickness 0.405 xi (1-xi)^2, closed over the last
 * 5% of the span), farfield 20 root chords away.  Lattice: n_chord cells on
 * the chord, 0.3 n_chord upstream, 0.6 n_chord downstream; n_span on the
 * span, 0.4 n_span beyond the tip; on each
 * side of the wing plane n_prism prism layers (first height h1, growth 1.2)
 * and n_tet rows of tetrahedra.  The chord and span stations are half
 * clustered (cosine at both chord edges, sine at the tip), half uniform: the
 * H-type lattice lines run to the farfield.  Cells: (1.9 n_chord)(1.4 n_span) x
 * 2 (2 n_prism + 6 n_tet); (80, 40, 30, 1e-5, 29) gives 3.98M.  Wall
 * SOLID_SURFACE, root plane SYMMETRY, outer boundary FARFIELD. */
phl_status phl_mesh_swept_wing(int n_chord, int n_span, int n_prism, double h1, int n_tet, phl_mesh **out);
/* Renumber the cells of a mesh into a new mesh.  method 0: greedy colouring
 * (cells listed colour by colour, each colour in the old order); method 1:
 * explicit permutation perm[new] = old.  The sweep order of LU-SGS is the cell
 * order (NSSolverUnstruct.cpp:15211,15541), so this selects the sweep order;
 * the reference's own renumbering step is API/PreProcess/src/Pre_Reorder.cpp. */
phl_status phl_mesh_renumber(const phl_mesh *in, int method, const int *perm, int *n_colours,
                             phl_mesh **out);
phl_status phl_mesh_sizes(const phl_mesh *m, int *nNode, int *nCell, int *nFace, int *nBFace,
                          int *nFaceNodes, int *nCellNodes);
phl_status phl_mesh_connectivity(const phl_mesh *m, double *x, double *y, double *z,
                                 int *faceNodePtr, int *faceNodes, int *left, int *right,
                                 int *bcType, int *cellNodePtr, int *cellNodes);
/* Face arrays are [nFace]; cell centroid and volume arrays are [nTotal]
 * (ghost centroid mirrored across the face); skew is [nCell] in degrees. */
phl_status phl_mesh_metrics(const phl_mesh *m, double *xfc, double *yfc, double *zfc,
                            double *xfn, double *yfn, double *zfn, double *area,
                            double *xcc, double *ycc, double *zcc, double *vol, double *skew);
void phl_mesh_destroy(phl_mesh *m);

/* The reference's partitioned grid file (.fts, HDF5; Region::ReadGrid,
 * API/Physics/src/Region.cpp:52, through IO_HDF5Read::ReadUnstructuredGrid,
 * API/PreProcess/src/Pre_HDF5File.cpp:1750-1900), read without the HDF5
 * library (superblock 0, symbol-table groups, contiguous datasets only).  The
 * zones are joined into one: cells zone by zone, nodes shared across zones
 * merged, the non-interface boundary faces first, then each zone's interior
 * faces, then one face per interface pair.  A grid of two-node faces is a 2-D
 * grid (the reference's 2-D metrics).  *n_zones receives the zone count. */
phl_status phl_read_fts(const char *path, phl_mesh **out, int *n_zones);
/* Its wall-distance file (.wdt: one dataset "Grid-<i>" of cell distances per
 * zone; IO_HDF5Read::ReadWallDist, Pre_HDF5File.cpp:909-975) in the joined
 * cell order; n_cell must match. */
phl_status phl_read_wdt(const char *path, int n_cell, double *wdist);
/* 2 for a 2-D grid (every face an edge), else 3. */
int phl_mesh_dim(const phl_mesh *m);

/* ---------------- output and restart files (host) ----------------------
 * Tecplot ASCII field file of a state, laid out like the reference's block
 * dump (Post_WriteTecplot::StoreUnsFieldVisualData, API/PostProcess/src/
 * Post_WriteTecplot.cpp:398-456, written by DumpToTecplotASCII::
 * DumpFlowByFaceTopo, API/PostProcess/src/Post_Visual.cpp:1801-1940): one
 * FEPolyhedron zone with node values of the reference's default visual
 * variables (density, u, v, w, pressure, temperature, Mach number, Cp;
 * cfd_para.hypara:790-791), interpolated from the cells as CompNodeVar does
 * (API/Geometry/src/Geo_UnstructGrid.cpp:8527-8603).  q is the primitive
 * state [5][nTotal] (phl_get_state), mach and gamma the freestream's. */
phl_status phl_write_tecplot(const phl_mesh *m, const double *q, double mach, double gamma, const char *title,
                             const char *path);
/* One zone's restart block of a steady run: int outnstep, then q[n_eq][n_total]
 * doubles, as NSSolverUnstruct::DumpRestart (CFD/NSSolver/src/
 * NSSolverUnstruct.cpp:1179-1197; n_eq = 5 primitives) and
 * TurbSolverUnstr::DumpRestart (CFD/TurbSolver/src/TurbSolverUnstr.cpp:
 * 435-456; n_eq = 1, nu~) write it.  phl_read_restart checks the file size
 * against n_eq and n_total (PHL_ERR_ARG when it does not match) and is the
 * InitFlowAsRestart counterpart (NSSolverUnstruct.cpp:1832): load q with
 * phl_set_state / phl_set_field. */
phl_status phl_write_restart(const char *path, int outnstep, int n_eq, int n_total, const double *q);
phl_status phl_read_restart(const char *path, int n_eq, int n_total, int *outnstep, double *q);

/* ---------------- batched form (no reference counterpart) ---------------
 * A batch is a disjoint union of meshes, one steady case each (a Mach /
 * angle / geometry sweep), solved together: every kernel launch of an
 * iteration covers all cases, and the level schedules of the LU-SGS sweeps
 * merge, so a level holds the cells of every case.  Each case keeps its own
 * connectivity, metrics and sweep order, so its state evolves exactly as in
 * a context of its own (its residual sums are accumulated in another order).
 * The reference runs one case per process (Simulation.cpp:1190). */

/* Concatenate n meshes into a batched mesh: nodes and cells in case order,
 * all boundary faces first (case by case), then all interior faces. */
phl_status phl_mesh_concat(const phl_mesh *const *parts, int n, phl_mesh **out);
/* Case ranges of a mesh: cell_ptr[n+1] (cells of case k are
 * cell_ptr[k]..cell_ptr[k+1]-1), bface_ptr[n+1] likewise for boundary faces
 * (and so ghost cells).  A plain mesh is one case.  Either array may be NULL. */
phl_status phl_mesh_cases(const phl_mesh *m, int *n_cases, int *cell_ptr, int *bface_ptr);

/* ---------------- device and context lifecycle ------------------------- */

/* GPUDeviceMultiGPU / GPUInfoQuery (GPU/Manage/src/GPUDeviceControl.cu:41-83). */
phl_status phl_device_info(int device, char *name, int name_len, int *n_cu, double *hbm_bytes);
/* GPUGeomInfoAllocCopy + GPUFlowVarMemAlloc + GPUFaceColorMain
 * (GPU/Kernels/src/BasicDeviceVariables.cu:459,2679; GPUFaceColor.cu:24):
 * uploads geometry, allocates flow fields, builds the LU-SGS level schedule. */
phl_status phl_create(const phl_mesh *mesh, const phl_params *p, int device, phl_ctx **out);
/* GPUMemoryFree (GPU/Kernels/include/BasicDeviceVariables.h:689). */
void phl_destroy(phl_ctx *ctx);
/* Batched context over a mesh from phl_mesh_concat with params[n_cases]: the
 * cases may differ only in mach, alpha_deg and beta_deg.  On a batched
 * context phl_iterate writes rms_hist[step * n_cases + case]; the standalone
 * phl_time_step and phl_lusgs hooks return PHL_ERR_UNSUPPORTED. */
phl_status phl_create_batch(const phl_mesh *mesh, const phl_params *params, int device, phl_ctx **out);
phl_status phl_batch_size(const phl_ctx *ctx, int *n_cases);
/* Freeze cases of a batched context (frozen[c] = 1: the update keeps case c's
 * state) from the next phl_iterate on; for drivers that judge convergence
 * themselves.  phl_solve_batch manages these flags on its own. */
phl_status phl_set_frozen(phl_ctx *ctx, const unsigned char *frozen);

/* ---------------- state transfer -------------------------------------- */
phl_status phl_set_state(phl_ctx *ctx, const double *q);          /* [5][nTotal] primitive */
/* CallGPUFlowVariablesCopyBack (GPU/Kernels/include/TemporaryOperations.h:33). */
phl_status phl_get_state(phl_ctx *ctx, double *q);
/* GPUResCopyBack (GPU/Kernels/include/TemporaryOperations.h:80). */
phl_status phl_get_residual(phl_ctx *ctx, double *res);
/* Any field of the path, laid out like the oracle's orc_get. */
phl_status phl_get_field(phl_ctx *ctx, int field, double *out);
/* Overwrite a field: Q ([5][nTotal], as phl_set_state), DT ([nCell]), VIST
 * ([nTotal]), NUT ([nTotal]) or WDIST ([nCell], e.g. a wall distance the
 * preprocessor computed; by default the context computes the exact distance
 * to the solid-surface faces when it is created). */
phl_status phl_set_field(phl_ctx *ctx, int field, const double *in);
/* Controller::InitFlow freestream fill (CFD/Solver/src/Controller.cpp:81-140). */
phl_status phl_init_freestream(phl_ctx *ctx);

/* ---------------- the path, one call per reference hook --------------- */
/* CallGPUBoundary (NSSolverUnstruct.cpp:13406). */
phl_status phl_boundary(phl_ctx *ctx);
/* CallGPUComputeGamaAndTemperature (:4668) and
 * CallGPUComputeViscousCoefficientWithoutChemical (:6724). */
phl_status phl_gamma_t(phl_ctx *ctx);
/* Controller::PressureFactor (Controller.cpp:1202-1234) ->
 * NSSolverUnstruct::CalPressureFactor (NSSolverUnstruct.cpp:6976-7045): the
 * cell sensor of roeEntropyFixMethod 2 and 6, which the Roe face flux reads;
 * nothing for the other methods.  UpdateIndependentField calls it after
 * gamma/T and before the gradient. */
phl_status phl_pressure_factor(phl_ctx *ctx);
/* NSSolverUnstruct::ComputeGradient -> GetGradientField (:18171,18244).
 * The same kernel also evaluates the limiter (both walk a cell's faces and
 * neighbours once) and refreshes the face loops' gather records. */
phl_status phl_gradient(phl_ctx *ctx);
/* gradient 2 ("ggnode"): NSSolverUnstruct::ComputeNodeValue + ModifyNodeValue
 * (:4734-5117; CallGPUComputeNodeValue :4738, CallGPUModifyNodeValue :5087),
 * which the reference runs in Controller::PostSolve after each update
 * (Controller.cpp:5151-5198): inverse-distance node averages of the state
 * and of the temperature array as the last phl_gamma_t left it.  phl_gradient
 * reads the node values; phl_iterate and phl_residual compute them right
 * after the boundary conditions, before the step's own phl_gamma_t.  After
 * phl_set_state / phl_init_freestream the first node pass takes the state's
 * own temperature (InitDependentVariables, NSSolver.cpp:761-794). */
phl_status phl_compute_node_value(phl_ctx *ctx);
/* The node values [6][nNode] (five primitives, then T) of the last node pass. */
phl_status phl_get_node_values(phl_ctx *ctx, double *out);
/* NSSolverUnstruct::ComputeLimiter (:18189).  Right after phl_gradient on an
 * unchanged state the limiter is already in place and this returns at once;
 * otherwise it runs the gradient kernel. */
phl_status phl_limiter(phl_ctx *ctx);
/* The steps of Controller::UpdateResiduals (CFD/Solver/src/Controller.cpp:
 * 1002-1083), one call per reference hook.  A host driver calls, per
 * iteration: phl_load_residuals, phl_boundary, phl_gamma_t (viscous),
 * phl_gradient, phl_limiter, then phl_update_residuals (or its two halves
 * phl_inviscid_flux and phl_viscous_flux), then phl_spectrum_radius,
 * phl_time_step and phl_lusgs.  The face passes read the states, gradients
 * and limiter that phl_boundary and phl_gradient left. */
/* NSSolverUnstruct::LoadResiduals (:4116, CallGPULoadResiduals :4118):
 * res = -rhs, with rhs = 0 on the single grid level. */
phl_status phl_load_residuals(phl_ctx *ctx);
/* NSSolverUnstruct::InviscidFlux (:6912): MUSCL reconstruction, BoundaryQlQrFix,
 * Roe, x area (CallGPUGetQlQr .. CallGPUinviscidScheme :13955-14466) and
 * LoadFlux (CallGPULoadFlux :6437), added to res. */
phl_status phl_inviscid_flux(phl_ctx *ctx);
/* NSSolverUnstruct::ViscousFlux (:7058; CallGPUGetVisFaceValue :7128,
 * CallGPUComputeVisflux :12979), added to res.  No-op for inviscid flow. */
phl_status phl_viscous_flux(phl_ctx *ctx);
/* NSSolver::UpdateResiduals (CFD/NSSolver/src/NSSolver.cpp:71-82): the
 * inviscid and viscous face loops (one face kernel for both), added to res. */
phl_status phl_update_residuals(phl_ctx *ctx);
/* Controller::LoadResiduals + Controller::UpdateResiduals fused
 * (Controller.cpp:1176,1002): res = -0.0, Boundary, the independent fields
 * (the laminar viscosity the turbulence hooks read; ggnode node values), the
 * gradient and limiter pass, both face loops, and the residual sums of the
 * result. */
phl_status phl_residual(phl_ctx *ctx);
/* CallGPUSpectrumRadiusInviscid / Viscous (:5751,5854). */
phl_status phl_spectrum_radius(phl_ctx *ctx);
/* CallGPULocalTimeStep + CallGPUComputeMinTimeStep + CallGPUReduceMaxTimeStep
 * (:6001,5705,5719); step is the outer iteration number (CFL ramp). */
phl_status phl_time_step(phl_ctx *ctx, int step);
/* Controller::LUSGS (CFD/Solver/src/Controller.cpp:1878): Diagonal, sweeps,
 * CallGPUUpdateFlowField.  n_fixed receives the SolutionFix count.  With
 * n_lusgs_sweeps > 1 the sweep pairs repeat from dq = 0 until a pair's
 * squared change of dq falls below lusgs_tol times the first pair's
 * (:1898-1951); the stream is synchronised after each pair, and phl_iterate
 * then launches its steps eagerly instead of replaying graphs. */
phl_status phl_lusgs(phl_ctx *ctx, int *n_fixed);
/* The sweep pairs the last NS (turb 0) or turbulence (turb 1) LU-SGS call
 * ran when n_lusgs_sweeps > 1 (the reference's sweepNormal of each pair, the
 * first 8 in normals8); *n_done is 0 for single-pair contexts.  No reference
 * counterpart: the reference keeps sweepNormal local to Controller::LUSGS. */
phl_status phl_sweep_normals(const phl_ctx *ctx, int turb, int *n_done, double *normals8);
/* The reference's isInIniting flag (GlobalDataBase "isInIniting", set for
 * the outer iterations 1..flowInitStep by Simulation::PreSolve,
 * Main/src/Simulation.cpp:1190-1236) for hook-by-hook callers: while on, the
 * face passes use first-order face values and the Van Leer flux
 * (NSSolverUnstruct.cpp:6923,14017,14229-14236) and the limiter is zero
 * (:18203).  phl_iterate sets it per step from flow_init_step itself. */
phl_status phl_set_init_phase(phl_ctx *ctx, int on);
/* ComputeResidualonGrid sums (API/Physics/src/Residual.cpp:534) of the
 * residual currently in res: sum5[m] = sum over real cells of res_m^2 (case 0
 * of a batched context).  Deterministic (fixed blocks, fixed order). */
phl_status phl_residual_norm(phl_ctx *ctx, double *sum5);
/* No reference counterpart.  For callers that cannot allocate device memory
 * (code under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the sums of phl_residual_norm over a caller-owned device array,
 * in the same fixed order (256-cell blocks, then the block sums), with no
 * context and no allocation.  The block partials live in a static device
 * array of 2.5 MiB that each GPU holds until the process ends.
 *   res    device pointer, rows [5][ld]; the first n_cell entries of each
 *          row are summed (the residual's real cells)
 *   sum5   device pointer to 5 doubles: sum over the cells of res_m^2
 *   stream a hipStream_t (NULL: the null stream); the work is only enqueued
 * n_cell above PHL_NOALLOC_MAX_CELLS (16,777,216) returns PHL_ERR_ARG.  All
 * calls on one GPU share the array: calls that overlap on one GPU (other
 * streams, other threads) race. */
#define PHL_NOALLOC_MAX_CELLS 16777216
phl_status phl_residual_norm_noalloc(const double *res, int n_cell, int ld, double *sum5, void *stream);

/* Wall force and moment coefficients: NSSolverUnstruct::AirForceCoefParts
 * over all solid-surface faces (:16026-16250) and Force::CompForce
 * (API/Physics/src/Force.cpp:439-533).  Boundary states, temperature,
 * viscosity and gradients are first refreshed from the current state.
 * coef[PHL_N_FORCE]: the part sums scaled as CompForce scales them (CA, CN,
 * CZ friction/pressure; pressure-only cpx/cpy/cpz; rolling, yawing and
 * pitching moments, friction/pressure), then the totals. */
enum {
    PHL_FORCE_CA_F = 0, PHL_FORCE_CA_P = 1, PHL_FORCE_CN_F = 2, PHL_FORCE_CN_P = 3,
    PHL_FORCE_CZ_F = 4, PHL_FORCE_CZ_P = 5, PHL_FORCE_CPX = 6, PHL_FORCE_CPY = 7, PHL_FORCE_CPZ = 8,
    PHL_FORCE_CMX_F = 9, PHL_FORCE_CMX_P = 10, PHL_FORCE_CMY_F = 11, PHL_FORCE_CMY_P = 12,
    PHL_FORCE_CMZ_F = 13, PHL_FORCE_CMZ_P = 14,
    PHL_FORCE_CL = 15, PHL_FORCE_CD = 16, PHL_FORCE_SIDE = 17, PHL_FORCE_CD_PR = 18,
    PHL_FORCE_CD_SF = 19, PHL_FORCE_CD_CL2PA = 20, PHL_FORCE_XCP = 21, PHL_N_FORCE = 22
};
phl_status phl_forces(phl_ctx *ctx, double *coef);
/* The same for case `c` of a batched context (phl_forces is case 0). */
phl_status phl_forces_case(phl_ctx *ctx, int c, double *coef);

/* ---------------- Spalart-Allmaras (viscous = 3) -----------------------
 * One-equation model, NSMB formation, non-conservative nu~, first-order
 * convection (CFD/TurbSolver/src/TurbSolverUnstr.cpp).  phl_iterate runs one
 * turbulence relaxation after each NS update, as Controller::Relaxation does
 * for the turbulence solver; the hooks below are its steps. */
/* TurbSolverUnstr::Boundary (:5919): ghost nu~ and eddy viscosity. */
phl_status phl_turb_boundary(phl_ctx *ctx);
/* TurbSolverUnstr::Viscosity (:6910), "1eq-sa": eddy viscosity of real cells. */
phl_status phl_turb_viscosity(phl_ctx *ctx);
/* TurbSolverUnstr::GetGradientField (:8652), ggcell gradient of nu~. */
phl_status phl_turb_gradient(phl_ctx *ctx);
/* TurbSolver::UpdateResiduals (TurbSolver.cpp:316): boundary, spectrum
 * initialisation from DT, SourceFluxOneEquationOriginal (:2702), first-order
 * NNDFlux (:5721) and the NSMB diffusion (ComputeVisflux :4438).  Reads the
 * NS gradients, the laminar viscosity (phl_gamma_t) and DT. */
phl_status phl_turb_update_residuals(phl_ctx *ctx);
/* SpectrumRadiusOfOneEquation (:4690): diagonal and face matrices. */
phl_status phl_turb_spectrum(phl_ctx *ctx);
/* SolveLUSGSForward / SetGhostDQLUSGS / SolveLUSGSBackward / UpdateFlowField
 * (:1400,1639,1521,1180): one sweep pair and the nu~ update. */
phl_status phl_turb_lusgs(phl_ctx *ctx);
/* One turbulence relaxation: boundary, viscosity, gradient, residuals,
 * spectrum, LU-SGS. */
phl_status phl_turb_relax(phl_ctx *ctx);

/* ---------------- host driver (Controller::Relaxation loop) ----------- */
/* n_steps outer iterations numbered first_step..; rms_hist[i] (may be NULL)
 * receives the density-residual RMS of each step (Residual.cpp:221-224). */
phl_status phl_iterate(phl_ctx *ctx, int first_step, int n_steps, double *rms_hist);
/* Iterate until R_k/R_1 <= drop or max_steps.  The residual is checked every
 * 32 steps, so the run ends on the first multiple of 32 steps at or after the
 * step that reached the drop; steps_done is the number of steps run (the
 * state returned is after that many), rms_hist[0..steps_done-1] the history. */
phl_status phl_solve(phl_ctx *ctx, int max_steps, double drop, int *steps_done, double *rms_hist);
/* Batched solve: each case iterates until its density residual RMS has
 * dropped by `drop` (then its state is frozen, from the end of the 32-step
 * chunk in which it converged) or max_steps.  steps_done[n_cases]: the step
 * at which each case reached the drop, or -(steps run) if it did not;
 * rms_hist[max_steps * n_cases] as phl_iterate. */
phl_status phl_solve_batch(phl_ctx *ctx, int max_steps, double drop, int *steps_done, double *rms_hist);
phl_status phl_synchronize(phl_ctx *ctx);
/* Number of dependency levels of the forward and backward LU-SGS sweeps
 * (cells of a level are updated in parallel; the levels run in order). */
phl_status phl_sweep_levels(const phl_ctx *ctx, int *fwd_levels, int *bwd_levels);
/* How phl_iterate / phl_solve launch (no reference counterpart: the
 * reference launches every kernel eagerly on the default stream).  mode 1
 * (default) replays hipGraphs of 8 captured steps (the steps of a call beyond
 * its last whole 8 launch eagerly); mode 0 launches eagerly; mode 2 launches
 * eagerly with the turbulence sweeps on a second stream.  Results are
 * identical; kernel timing always launches eagerly. */
phl_status phl_set_launch_mode(phl_ctx *ctx, int mode);
/* Autotuner (no reference counterpart): times the flux and gradient kernels
 * at workgroup tiles of 64, 128 and 256 cells on the current state and keeps
 * the fastest for this context.  Tiles only regroup the residual partial
 * sums.  ms_table (may be NULL) receives 6 mean times: flux at 64/128/256,
 * then gradient at 64/128/256. */
phl_status phl_autotune(phl_ctx *ctx, int *flux_tile, int *grad_tile, double *ms_table);
/* Switch between the FP64 form (0) and the mixed form (1: FP32 gradients,
 * reconstruction, Roe and viscous fluxes; FP64 residual accumulation,
 * spectral radii, time step, LU-SGS and norms) on a live context, e.g. to
 * finish a mixed-form solve in FP64.  The reference has FP64 only. */
phl_status phl_set_precision(phl_ctx *ctx, int precision);
phl_status phl_get_precision(const phl_ctx *ctx, int *precision);
/* Precision switch of phl_solve / phl_solve_batch (no reference counterpart):
 * a mixed context whose unconverged cases have all gone stall_steps steps
 * without halving their residual (the FP32 flux round-off floor, near a
 * 1e-5 drop) switches to FP64 and continues.  0 (default) never switches. */
phl_status phl_set_precision_switch(phl_ctx *ctx, int stall_steps);

/* ---------------- learned warm start (no reference counterpart) -------
 * A mesh message-passing network predicting a case's flow field from cell
 * features (position and case parameters), trained on converged fields the
 * solver generates; the prediction is loaded with phl_set_state and the
 * solver's own residual judges it.  Hidden width 128, FP32, dense layers on
 * MFMA.  Parameters, flat: W0 [8][128], b0 [128], A [L][128][128],
 * B [L][128][128], c [L][128], Wo [128][5], bo [5].  Features X [N][8],
 * targets / outputs [N][5], row-major. */
typedef struct phl_mpnn phl_mpnn;
phl_status phl_mpnn_create(int layers, unsigned seed, int device, phl_mpnn **out);
void phl_mpnn_destroy(phl_mpnn *net);
phl_status phl_mpnn_n_params(const phl_mpnn *net, int *count);
phl_status phl_mpnn_set_params(phl_mpnn *net, const float *flat);   /* also resets Adam */
phl_status phl_mpnn_get_params(phl_mpnn *net, float *flat);
/* Graph: nbr[n_nodes][width] neighbour lists, -1 padded; or the interior
 * face neighbours of a mesh's cells.  The lists must be symmetric with equal
 * multiplicity (j appears in i's list as often as i in j's, as face relations
 * are): the backward pass applies the transpose of the neighbour mean as a
 * gather over each node's own list.  Other lists return PHL_ERR_ARG. */
phl_status phl_mpnn_set_graph(phl_mpnn *net, int n_nodes, int width, const int *nbr);
phl_status phl_mpnn_set_graph_mesh(phl_mpnn *net, const phl_mesh *mesh);
/* Features (and targets, may be NULL) of the current graph, kept on the device. */
phl_status phl_mpnn_set_data(phl_mpnn *net, const float *X, const float *T);
phl_status phl_mpnn_forward(phl_mpnn *net, float *Y);
/* Precision of phl_mpnn_forward's hidden layers: 0 FP32 (default); 1 BF16
 * products h A_l + m B_l (h, m and the weights rounded to BF16, FP32
 * accumulation on v_mfma_f32_32x32x16_bf16; the input and output layers, the
 * neighbour means, biases and residual adds stay FP32).  Training and
 * phl_mpnn_gradient always run in FP32. */
phl_status phl_mpnn_set_forward_precision(phl_mpnn *net, int bf16);
/* Loss (mean of squared errors) and its gradient, flat like the parameters. */
phl_status phl_mpnn_gradient(phl_mpnn *net, double *loss, float *grad);
/* steps of Adam (lr, beta 0.9 / 0.999, eps 1e-8); loss_hist[steps] may be NULL. */
phl_status phl_mpnn_train(phl_mpnn *net, int steps, float lr, double *loss_hist);

/* ---------------- measurement (GPU/Manage/src/TimeMeasure.cu:15-67) ----
 * With timing on, phl_iterate records HIP events on the context's stream
 * around each kernel class of the iteration.  Classes, in launch order:
 * 0 boundary, 1 gradient+limiter, 2 face flux (+spectral radius, local time
 * step), 3 residual norm + minimum time step, 4 dt/diagonal, 5 forward sweep,
 * 6 ghost dq, 7 backward sweep, 8 update, 9 turbulence relaxation (SA; empty
 * otherwise).  ms_total[k] is the summed time of class k over n_steps[0]
 * timed steps. */
#define PHL_N_KERNEL_CLASSES 10
phl_status phl_set_kernel_timing(phl_ctx *ctx, int enable);
phl_status phl_get_kernel_timing(phl_ctx *ctx, double *ms_total, int *n_steps);
/* Achievable HBM bandwidth: a streaming copy of `bytes` bytes, best of
 * `iters` launches, in GB/s (reported beside the vendor peak). */
phl_status phl_hbm_copy(int device, double bytes, int iters, double *gbps);
/* Launch statistics (no reference counterpart): graph replays and eagerly
 * launched steps of phl_iterate since the context was created. */
phl_status phl_launch_counts(const phl_ctx *ctx, long long *graph_replays, long long *eager_steps);
/* Wait bound of the dataflow sweeps, in ticks of the 100 MHz wall clock (no
 * reference counterpart; default 20,000,000 = 0.2 s).  A wave that waits
 * longer gives up, every later waiter gives up at once, every kernel still
 * exits, and phl_iterate / phl_lusgs return PHL_ERR_NUMERIC; the state is then
 * invalid until phl_set_state or phl_init_freestream replaces it (which also
 * clears the give-up count).  0 makes every wave that would wait give up. */
phl_status phl_set_sweep_wait(phl_ctx *ctx, int ticks);

#ifdef __cplusplus
}
#endif
#endif
