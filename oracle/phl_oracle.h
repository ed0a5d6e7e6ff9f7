/*
 * phl_oracle.h -- CPU oracle for the unstructured residual + LU-SGS path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, sequential restatement of the
 * reference solver (PHengLEI NSSolverUnstruct, FP64, one zone) used as the
 * parity checker for the HIP implementation.  Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load it; the product library never
 * links, calls or falls back to it.
 *
 * Every function cites the reference file:line it restates.  Loops keep the
 * reference's structure (face loops that scatter into cells, sequential
 * LU-SGS in cell-index order) so that a gather-based GPU implementation is
 * checked against the reference's own evaluation order.
 */
#ifndef PHL_ORACLE_H
#define PHL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Parameter block.  Field order and types are identical to phl_params in
 * include/phl.h so one ctypes structure serves both. */
typedef struct orc_params {
    double mach;            /* refMachNumber */
    double alpha_deg;       /* attackd */
    double beta_deg;        /* angleSlide */
    double gamma;           /* refGama */
    double reynolds;        /* refReNumber, per unit grid length */
    double t_ref;           /* refDimensionalTemperature [K] */
    double prandtl;         /* prl */
    double prandtl_turb;    /* prt */
    double t_wall;          /* wallTemperature [K], <= 0 adiabatic */
    double venkat_k;        /* venkatCoeff */
    double entropy_scale;   /* roeEntropyScale */
    double cfl_start;       /* CFLStart */
    double cfl_end;         /* CFLEnd */
    double ktmax;           /* ktmax */
    double lusgs_tol;       /* LUSGSTolerance */
    double force_ref_length;
    double force_ref_area;
    double force_ref_x, force_ref_y, force_ref_z;
    double force_ref_span;
    double sa_nu_inf, sa_vist_inf, turb_cfl_scale, turb_skew_angle;
    int viscous;            /* 0 inviscid, 1 laminar, 3 Spalart-Allmaras */
    int gradient;           /* 0 ggcell, 1 lsq, 2 ggnode */
    int limiter;            /* 0 vencat (ivencat=7), 1 nolim, 2 first order */
    int entropy_fix;        /* roeEntropyFixMethod (1..6) */
    int cfl_vary_step;      /* CFLVaryStep */
    int n_lusgs_sweeps;     /* nLUSGSSweeps */
    int precision;          /* 0 fp64, 1 mixed (ignored by the oracle) */
    int storage_order;  /* device storage order of the product; the oracle ignores it */
    int face_loop;      /* face-loop form of the product; the oracle ignores it */
    int sweep_form;     /* sweep launch form of the product; the oracle ignores it */
    int flow_init_step; /* flowInitStep: iterations 1..flow_init_step first order with Van Leer */
} orc_params;

/* Boundary-condition codes (API/Geometry/include/Geo_SimpleBC.h:32-39). */
enum { ORC_BC_EXTRAPOLATION = 1, ORC_BC_SOLID_SURFACE = 2, ORC_BC_SYMMETRY = 3,
       ORC_BC_FARFIELD = 4, ORC_BC_INFLOW = 5, ORC_BC_OUTFLOW = 6 };

/* Field identifiers for orc_get / orc_set. */
enum { ORC_Q = 0, ORC_RES = 1, ORC_DQDX = 2, ORC_DQDY = 3, ORC_DQDZ = 4,
       ORC_LIMIT = 5, ORC_DQ = 6, ORC_SPEC_INV = 7, ORC_SPEC_VIS = 8,
       ORC_DT = 9, ORC_DIAG = 10, ORC_T = 11, ORC_VISL = 12, ORC_DTDX = 13,
       ORC_DTDY = 14, ORC_DTDZ = 15, ORC_DELTA_FLUX = 16,
       /* Spalart-Allmaras (viscous == 3) */
       ORC_VIST = 17, ORC_NUT = 18, ORC_RES_T = 19, ORC_SPEC_T = 20, ORC_DQ_T = 21,
       ORC_WDIST = 22, ORC_DNUDX = 23, ORC_DNUDY = 24, ORC_DNUDZ = 25,
       /* pressure factor of roeEntropyFixMethod 2 and 6 (rtem) */
       ORC_PRESSURE_FACTOR = 26 };

typedef struct orc_solver orc_solver;

/* ---- geometry (restates API/Geometry/src/Geo_UnstructGrid.cpp) ---- */
void orc_face_metrics_2d(int nFace, const int *fp, const int *fn, const double *x, const double *y, const double *z,
                         double *xfc, double *yfc, double *zfc, double *xfn, double *yfn, double *zfn, double *area);
void orc_cell_metrics_2d(int nCell, int nBFace, int nFace, const int *left, const int *right,
                         const double *xfc, const double *yfc, const double *zfc,
                         const double *xfn, const double *yfn, const double *zfn, const double *area,
                         double *xcc, double *ycc, double *zcc, double *vol);
void orc_face_metrics(int nFace, const int *faceNodePtr, const int *faceNodes,
                      const double *x, const double *y, const double *z,
                      double *xfc, double *yfc, double *zfc,
                      double *xfn, double *yfn, double *zfn, double *area);
void orc_cell_metrics(int nCell, int nBFace, int nFace, const int *left, const int *right,
                      const int *faceNodePtr, const int *faceNodes,
                      const int *cellNodePtr, const int *cellNodes,
                      const double *x, const double *y, const double *z,
                      const double *xfc, const double *yfc, const double *zfc,
                      const double *xfn, const double *yfn, const double *zfn,
                      const double *area,
                      double *xcc, double *ycc, double *zcc, double *vol);
void orc_skewness(int nCell, int nBFace, int nFace, const int *left, const int *right,
                  const double *xfc, const double *yfc, const double *zfc,
                  const double *xfn, const double *yfn, const double *zfn, const double *area,
                  const double *xcc, const double *ycc, const double *zcc, double *skew);

/* ---- solver state ---- */
orc_solver *orc_create(const orc_params *p, int nCell, int nBFace, int nFace,
                       const int *left, const int *right, const int *bcType,
                       const double *xfc, const double *yfc, const double *zfc,
                       const double *xfn, const double *yfn, const double *zfn,
                       const double *area, const double *xcc, const double *ycc,
                       const double *zcc, const double *vol, const double *skew);
void orc_destroy(orc_solver *s);
int  orc_get(orc_solver *s, int field, double *out);
int  orc_set(orc_solver *s, int field, const double *in);
void orc_freestream(const orc_solver *s, double *prim5);
void orc_init_freestream(orc_solver *s);

/* ---- node-based Green-Gauss gradient (gradient == 2, "ggnode") ---- */
/* Node coordinates and the face/cell node lists (required for gradient 2). */
void orc_set_nodes(orc_solver *s, int nNode, const double *x, const double *y, const double *z,
                   const int *faceNodePtr, const int *faceNodes, const int *cellNodePtr, const int *cellNodes);
/* ComputeNodeValue + ModifyNodeValue (NSSolverUnstruct.cpp:4734-5117); the
 * residual evaluation calls it after the boundary conditions. */
void orc_node_values(orc_solver *s);
void orc_get_node_values(const orc_solver *s, double *out6);  /* [6][nNode]: 5 primitives, T */

/* ---- one iteration, piece by piece (NSSolverUnstruct / Controller) ---- */
void orc_boundary(orc_solver *s);
void orc_gamma_t(orc_solver *s);
void orc_pressure_factor(orc_solver *s);
void orc_viscosity(orc_solver *s);
void orc_gradient(orc_solver *s);
void orc_limiter(orc_solver *s);
void orc_zero_residual(orc_solver *s);
void orc_inviscid_flux(orc_solver *s);
void orc_viscous_flux(orc_solver *s);
void orc_update_residuals(orc_solver *s);          /* NSSolver::UpdateResiduals: both face loops */
void orc_residual(orc_solver *s);                  /* LoadResiduals + Controller::UpdateResiduals */
void orc_spectrum(orc_solver *s);
void orc_time_step(orc_solver *s, int step);
int  orc_lusgs(orc_solver *s, double *sweep_normals); /* returns #negative cells fixed */
/* sweep pairs of the last NS (turb 0) or SA (turb 1) LU-SGS call; the
 * squared changes of its first 8 pairs in out8 */
int orc_sweep_normals(const orc_solver *s, int turb, double *out8);
/* isInIniting for the hook-by-hook calls (orc_iterate sets it per step) */
void orc_set_init_phase(orc_solver *s, int on);
void orc_vanleer_flux(const double *qL, const double *qR, double xsn, double ysn, double zsn, double gama,
                      double *flux);
void orc_residual_norm(const orc_solver *s, double *sum5);
int  orc_iterate(orc_solver *s, int first_step, int n_steps, double *rms_hist);
/* ---- Spalart-Allmaras (TurbSolverUnstr) ---- */
void orc_set_wall_distance(orc_solver *s, const double *wd);
void orc_set_dim(orc_solver *s, int dim);
void orc_wall_distance(int nCell, const double *xcc, const double *ycc, const double *zcc, int nWall,
                       const int *wallFaceNodePtr, const int *wallFaceNodes, const double *wxfc,
                       const double *wyfc, const double *wzfc, const double *x, const double *y,
                       const double *z, double *out);
void orc_wall_distance_2d(int nCell, const double *xcc, const double *ycc, const double *zcc, int nWall,
                          const int *wallFaceNodePtr, const int *wallFaceNodes, const double *wxfc,
                          const double *wyfc, const double *wzfc, const double *x, const double *y,
                          const double *z, double *out);
void orc_sa_boundary(orc_solver *s);
void orc_sa_viscosity(orc_solver *s);
void orc_sa_gradient(orc_solver *s);
void orc_sa_update_residuals(orc_solver *s);
void orc_sa_spectrum(orc_solver *s);
void orc_sa_lusgs(orc_solver *s);
void orc_sa_relax(orc_solver *s);   /* one turbulence relaxation after the NS one */
void orc_sa_face_matrices(const orc_solver *s, double *matL, double *matR);
/* SA source at one point: out[0] source per volume, out[1] (P - D)/nu~,
 * out[2] its linearisation, out[3] the implicit diagonal term per volume. */
void orc_sa_source_point(double nut, double omega, double wdist, double rho, double visl, double reynolds,
                         double *out);
/* MUSCL face states of face f (after orc_gradient / orc_limiter), with the
 * positivity fallback and BoundaryQlQrFix applied. */
void orc_face_states(const orc_solver *s, int f, double *qL, double *qR);
#define ORC_N_FORCE 22
void orc_forces(orc_solver *s, double *out);        /* AirForceCoefParts + CompForce */
void orc_force_coefficients(const double *raw, double alpha_deg, double beta_deg, double *out);

/* ---- point functions, exposed for closed-form tests ---- */
void   orc_roe_flux(const double *qL, const double *qR, double nx, double ny, double nz,
                    double gamma, int fix_method, double acoustic, double convect, double *flux);
/* with the pressure factors of the face's cells (roeEntropyFixMethod 2, 6) */
void   orc_roe_flux_pc(const double *qL, const double *qR, double nx, double ny, double nz, double gamma,
                       int fix_method, double acoustic, double convect, double pcL, double pcR, double *flux);
double orc_harten(double eig, double limit);
double orc_venfun(double x, double y, double eps);
void   orc_mxdq(const double *prim, double nx, double ny, double nz, double area,
                double gamma, const double *dq, double radius, int ipn, double *f);
double orc_lusgs_rad(const double *ql, const double *qr, double nx, double ny, double nz,
                     double area, double gamma);
double orc_entropy_coef(double mach, double scale);
double orc_ptotal(double mach, double gamma);

#ifdef __cplusplus
}
#endif
#endif
