// phl_mesh.h -- host-side unstructured grid for the residual + LU-SGS path.
//
// Connectivity follows the reference's UnstructGrid conventions: boundary
// faces first, ghost cell of boundary face f = nCell + f, face normal from
// left to right cell (API/Geometry/src/Geo_UnstructGrid.cpp:2007,5911-6157).
#pragma once

#include <string>
#include <vector>

namespace phl {

struct Mesh {
    int nNode = 0, nCell = 0, nFace = 0, nBFace = 0;
    std::vector<double> x, y, z;
    std::vector<int> faceNodePtr, faceNodes;  // CSR, nFace+1
    std::vector<int> cellNodePtr, cellNodes;  // CSR, nCell+1
    std::vector<int> left, right, bcType;     // bcType sized nBFace

    // Metrics (UnstructGrid::ComputeMetrics).
    std::vector<double> xfc, yfc, zfc, xfn, yfn, zfn, area;  // [nFace]
    std::vector<double> xcc, ycc, zcc, vol;                  // [nTotal]
    std::vector<double> skew;                                // [nCell], degrees
    double avgVol = 0.0;

    // Cases of a concatenated (batched) mesh: case k owns cells
    // [caseCellPtr[k], caseCellPtr[k+1]) and boundary faces
    // [caseBFacePtr[k], caseBFacePtr[k+1]); avgVol of each case.  Empty for
    // a single-case mesh.
    std::vector<int> caseCellPtr, caseBFacePtr;
    std::vector<double> caseAvgVol;

    // 2 for a planar grid (every face an edge of two nodes, cells polygons,
    // z = 0): metrics by the reference's 2-D routines
    // (UnstructGrid::ComputeMetrics, Geo_UnstructGrid.cpp:5695-5709), and the
    // Venkatakrishnan scale takes square roots of areas.
    int dim = 3;

    int nTotal() const { return nCell + nBFace; }
    int nCases() const { return caseCellPtr.empty() ? 1 : (int)caseCellPtr.size() - 1; }
    void computeMetrics();  // by dim, then skewness and mean volume
    void computeMetrics2D();
    void computeMetrics3D();
    void computeSkewness();
};

// Disjoint union of meshes, one case each: cells and nodes concatenated in
// case order, all boundary faces first (case by case), then all interior
// faces (case by case).  Each cell keeps its faces in the same relative
// order, and the metrics are copied, so every case sees exactly the
// connectivity, metrics and sweep order of its own mesh.
bool concatMeshes(const std::vector<const Mesh *> &parts, Mesh &out, std::string &err);

// One cell of a generic element mesh, before faces are derived.
enum CellKind { TET = 0, PYRAMID = 1, PRISM = 2, HEX = 3 };

// Derives faces from cells (shared faces matched by node set), orients each
// face from the lower-numbered to the higher-numbered cell (boundary faces
// outward), orders boundary faces first, then computes metrics.  classify(f)
// returns the BC code of a boundary face given its centroid.
struct ElementMesh {
    std::vector<double> x, y, z;
    std::vector<int> kind;       // CellKind per cell
    std::vector<int> cellNodes;  // 4/5/6/8 nodes per cell, VTK local order
    std::vector<int> cellNodePtr;
};
typedef int (*BcClassifier)(double xc, double yc, double zc, const void *user);
bool buildMesh(const ElementMesh &em, BcClassifier classify, const void *user, Mesh &out,
               std::string &err);

bool makeBumpHex(int nx, int ny, int nz, double bump, Mesh &out, std::string &err);

// Hybrid bump channel: nPrism prism layers over the triangulated lower wall
// (first height h1, geometric growth chosen so the last layer matches the tet
// spacing), then a tetrahedral fill of nyTet rows: a seeded jittered point
// lattice (interior points moved by up to `jitter` of the local spacing) split
// into six tetrahedra per lattice cell.
bool makeBumpHybrid(int nx, int nz, int nPrism, double h1, int nyTet, double jitter, unsigned seed,
                    double bump, Mesh &out, std::string &err);

// Semi-span swept wing in a box: root chord 1 on the symmetry plane z = 0,
// taper 0.5, aspect ratio 4, quarter-chord sweep 25 degrees, symmetric cubic
// section of 12% thickness; farfield 20 root chords away.  A lattice of
// nChord cells on the chord (plus 0.3 nChord upstream and 0.6 nChord
// downstream) by nSpan cells on the span (plus 0.4 nSpan beyond the tip);
// on each side of the wing plane nPrism prism layers (first height h1,
// growth 1.2) and nTet rows of tetrahedra.
bool makeSweptWing(int nChord, int nSpan, int nPrism, double h1, int nTet, Mesh &out, std::string &err);

// Renumber cells: new cell i is old cell perm[i].  Faces are re-derived in
// the reference's order (boundary first, then interior by owner) and the
// metrics recomputed.  The LU-SGS sweep order of the reference is the cell
// order, so a renumbering is a choice of sweep order.
bool renumberCells(const Mesh &in, const std::vector<int> &perm, Mesh &out, std::string &err);
// Greedy face-neighbour colouring in cell order; returns the permutation that
// lists cells by (colour, old index) and the number of colours.
std::vector<int> colourOrder(const Mesh &m, int &nColours);

}  // namespace phl

// The C-ABI's opaque mesh handle.
struct phl_mesh {
    phl::Mesh m;
};

namespace phl {
// Message returned by phl_last_error (thread-local).
void set_last_error(const std::string &msg);
}  // namespace phl
