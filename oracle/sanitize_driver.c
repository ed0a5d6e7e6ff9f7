/*
 * sanitize_driver.c -- TEST INFRASTRUCTURE ONLY: drives the CPU oracle on a
 * small hexahedral box under AddressSanitizer and UndefinedBehaviorSanitizer
 * (tests/test_oracle_sanitize.py builds and runs it).  Every gradient form
 * (ggcell, lsq, ggnode), limiter and flow model (inviscid, laminar,
 * Spalart-Allmaras), with the Roe entropy fixes 1-6 spread over the runs,
 * takes a few iterations, the forces and the node values;
 * the sanitizers report any out-of-bounds access, leak or undefined
 * operation.
 */
#include "phl_oracle.h"

#include <stdio.h>
#include <stdlib.h>

#define NX 5
#define NY 3
#define NZ 2

static int node_id(int i, int j, int k) { return i + (NX + 1) * (j + (NY + 1) * k); }
static int cell_id(int i, int j, int k) { return i + NX * (j + NY * k); }

typedef struct {
    int n[4], left, right, bc;
} face_t;

static int add_face(face_t *fs, int nf, int a, int b, int c, int d, int left, int right, int bc, int flip)
{
    face_t f;
    if (flip) { f.n[0] = a; f.n[1] = d; f.n[2] = c; f.n[3] = b; }
    else      { f.n[0] = a; f.n[1] = b; f.n[2] = c; f.n[3] = d; }
    f.left = left; f.right = right; f.bc = bc;
    fs[nf] = f;
    return nf + 1;
}

int main(void)
{
    const int nNode = (NX + 1) * (NY + 1) * (NZ + 1), nCell = NX * NY * NZ;
    double *x = malloc(sizeof(double) * nNode), *y = malloc(sizeof(double) * nNode), *z = malloc(sizeof(double) * nNode);
    for (int k = 0; k <= NZ; ++k)
        for (int j = 0; j <= NY; ++j)
            for (int i = 0; i <= NX; ++i) {
                int n = node_id(i, j, k);
                x[n] = -1.0 + 2.0 * i / NX + 0.03 * j;  /* sheared a little */
                y[n] = 0.8 * j / NY;
                z[n] = 0.2 * k / NZ;
            }
    const int maxF = 3 * (NX + 1) * (NY + 1) * (NZ + 1);
    face_t *bf = malloc(sizeof(face_t) * maxF), *inf = malloc(sizeof(face_t) * maxF);
    int nb = 0, ni = 0;
    /* x faces: nodes (i,j,k) (i,j+1,k) (i,j+1,k+1) (i,j,k+1) give +x */
    for (int k = 0; k < NZ; ++k)
        for (int j = 0; j < NY; ++j)
            for (int i = 0; i <= NX; ++i) {
                int a = node_id(i, j, k), b = node_id(i, j + 1, k), c = node_id(i, j + 1, k + 1), d = node_id(i, j, k + 1);
                if (i == 0) nb = add_face(bf, nb, a, b, c, d, cell_id(0, j, k), -1, ORC_BC_FARFIELD, 1);
                else if (i == NX) nb = add_face(bf, nb, a, b, c, d, cell_id(NX - 1, j, k), -1, ORC_BC_FARFIELD, 0);
                else ni = add_face(inf, ni, a, b, c, d, cell_id(i - 1, j, k), cell_id(i, j, k), 0, 0);
            }
    /* y faces: (i,j,k) (i,j,k+1) (i+1,j,k+1) (i+1,j,k) give +y */
    for (int k = 0; k < NZ; ++k)
        for (int j = 0; j <= NY; ++j)
            for (int i = 0; i < NX; ++i) {
                int a = node_id(i, j, k), b = node_id(i, j, k + 1), c = node_id(i + 1, j, k + 1), d = node_id(i + 1, j, k);
                if (j == 0) nb = add_face(bf, nb, a, b, c, d, cell_id(i, 0, k), -1, ORC_BC_SOLID_SURFACE, 1);
                else if (j == NY) nb = add_face(bf, nb, a, b, c, d, cell_id(i, NY - 1, k), -1, ORC_BC_SYMMETRY, 0);
                else ni = add_face(inf, ni, a, b, c, d, cell_id(i, j - 1, k), cell_id(i, j, k), 0, 0);
            }
    /* z faces: (i,j,k) (i+1,j,k) (i+1,j+1,k) (i,j+1,k) give +z */
    for (int k = 0; k <= NZ; ++k)
        for (int j = 0; j < NY; ++j)
            for (int i = 0; i < NX; ++i) {
                int a = node_id(i, j, k), b = node_id(i + 1, j, k), c = node_id(i + 1, j + 1, k), d = node_id(i, j + 1, k);
                if (k == 0) nb = add_face(bf, nb, a, b, c, d, cell_id(i, j, 0), -1, ORC_BC_SYMMETRY, 1);
                else if (k == NZ) nb = add_face(bf, nb, a, b, c, d, cell_id(i, j, NZ - 1), -1, ORC_BC_SYMMETRY, 0);
                else ni = add_face(inf, ni, a, b, c, d, cell_id(i, j, k - 1), cell_id(i, j, k), 0, 0);
            }
    const int nFace = nb + ni, nTotal = nCell + nb;
    int *fnp = malloc(sizeof(int) * (nFace + 1)), *fn = malloc(sizeof(int) * 4 * nFace);
    int *left = malloc(sizeof(int) * nFace), *right = malloc(sizeof(int) * nFace), *bc = malloc(sizeof(int) * nb);
    for (int f = 0; f < nFace; ++f) {
        const face_t *s = f < nb ? &bf[f] : &inf[f - nb];
        fnp[f] = 4 * f;
        for (int q = 0; q < 4; ++q) fn[4 * f + q] = s->n[q];
        left[f] = s->left;
        right[f] = f < nb ? nCell + f : s->right;
        if (f < nb) bc[f] = s->bc;
    }
    fnp[nFace] = 4 * nFace;
    int *cnp = malloc(sizeof(int) * (nCell + 1)), *cn = malloc(sizeof(int) * 8 * nCell);
    for (int k = 0; k < NZ; ++k)
        for (int j = 0; j < NY; ++j)
            for (int i = 0; i < NX; ++i) {
                int c = cell_id(i, j, k), *o = cn + 8 * c;
                o[0] = node_id(i, j, k); o[1] = node_id(i + 1, j, k); o[2] = node_id(i + 1, j + 1, k);
                o[3] = node_id(i, j + 1, k); o[4] = node_id(i, j, k + 1); o[5] = node_id(i + 1, j, k + 1);
                o[6] = node_id(i + 1, j + 1, k + 1); o[7] = node_id(i, j + 1, k + 1);
                cnp[c] = 8 * c;
            }
    cnp[nCell] = 8 * nCell;

    double *xfc = malloc(sizeof(double) * nFace), *yfc = malloc(sizeof(double) * nFace), *zfc = malloc(sizeof(double) * nFace);
    double *xfn = malloc(sizeof(double) * nFace), *yfn = malloc(sizeof(double) * nFace), *zfn = malloc(sizeof(double) * nFace);
    double *area = malloc(sizeof(double) * nFace);
    double *xcc = malloc(sizeof(double) * nTotal), *ycc = malloc(sizeof(double) * nTotal), *zcc = malloc(sizeof(double) * nTotal);
    double *vol = malloc(sizeof(double) * nTotal), *skew = malloc(sizeof(double) * nCell);
    orc_face_metrics(nFace, fnp, fn, x, y, z, xfc, yfc, zfc, xfn, yfn, zfn, area);
    orc_cell_metrics(nCell, nb, nFace, left, right, fnp, fn, cnp, cn, x, y, z, xfc, yfc, zfc, xfn, yfn, zfn, area,
                     xcc, ycc, zcc, vol);
    orc_skewness(nCell, nb, nFace, left, right, xfc, yfc, zfc, xfn, yfn, zfn, area, xcc, ycc, zcc, skew);
    for (int c = 0; c < nCell; ++c)
        if (!(vol[c] > 0.0)) { fprintf(stderr, "non-positive volume in cell %d\n", c); return 2; }

    double *wd = malloc(sizeof(double) * nCell), *nodes6 = malloc(sizeof(double) * 6 * nNode);
    double forces[ORC_N_FORCE], hist[3];
    for (int c = 0; c < nCell; ++c) wd[c] = ycc[c];
    const int viscous[3] = {0, 1, 3};
    for (int g = 0; g < 3; ++g)
        for (int v = 0; v < 3; ++v)
            for (int lim = 0; lim < 3; ++lim) {
                orc_params p = {0};
                p.mach = 0.5; p.alpha_deg = 2.0; p.gamma = 1.4; p.reynolds = 1e5; p.t_ref = 288.15;
                p.prandtl = 0.72; p.prandtl_turb = 0.9; p.t_wall = -1.0; p.venkat_k = 5.0; p.entropy_scale = 1.0;
                p.cfl_start = 5.0; p.cfl_end = 5.0; p.ktmax = 1e10; p.lusgs_tol = 0.01;
                p.force_ref_length = 1.0; p.force_ref_area = 1.0; p.force_ref_span = 1.0;
                p.sa_nu_inf = 3.0; p.sa_vist_inf = 1e-3; p.turb_cfl_scale = 1.0;
                p.viscous = viscous[v]; p.gradient = g; p.limiter = lim;
                p.entropy_fix = 1 + (9 * g + 3 * v + lim) % 6;  /* every Roe entropy fix */
                p.cfl_vary_step = 500; p.n_lusgs_sweeps = 1;
                orc_solver *s = orc_create(&p, nCell, nb, nFace, left, right, bc, xfc, yfc, zfc, xfn, yfn, zfn, area,
                                           xcc, ycc, zcc, vol, skew);
                orc_set_nodes(s, nNode, x, y, z, fnp, fn, cnp, cn);
                if (p.viscous == 3) orc_set_wall_distance(s, wd);
                orc_init_freestream(s);
                orc_iterate(s, 1, 3, hist);
                orc_forces(s, forces);
                orc_get_node_values(s, nodes6);
                for (int i = 0; i < 3; ++i)
                    if (!(hist[i] == hist[i])) { fprintf(stderr, "NaN residual (g%d v%d l%d)\n", g, v, lim); return 3; }
                orc_destroy(s);
            }
    free(x); free(y); free(z); free(bf); free(inf); free(fnp); free(fn); free(left); free(right); free(bc);
    free(cnp); free(cn); free(xfc); free(yfc); free(zfc); free(xfn); free(yfn); free(zfn); free(area);
    free(xcc); free(ycc); free(zcc); free(vol); free(skew); free(wd); free(nodes6);
    printf("sanitized oracle runs: ok\n");
    return 0;
}
