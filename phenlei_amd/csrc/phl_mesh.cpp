// phl_mesh.cpp -- synthetic grid generation, face derivation and metrics.
//
// Metrics restate the reference's UnstructGrid routines so that the device
// solver sees the same geometry the reference would compute from the same
// nodes:
//   face centroid/normal/area  Geo_UnstructGrid.cpp:5911-6157
//   cell centroid/volume       Geo_UnstructGrid.cpp:6545-6820
//   cell skewness              Geo_UnstructGrid.cpp:2206-2300
//   mean volume                API/Physics/src/Region.cpp:3203
#include "phl_mesh.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>

namespace phl {

namespace {
const double kSmall = 1.0e-40;
const double kTiny = 1.0e-40;
const double kPi = 3.141592653589793;

inline double dmin(double a, double b) { return (a < b) ? a : b; }
}  // namespace

// Planar metrics (Geo_UnstructGrid.cpp: ComputeFaceNormal2D :6158-6187,
// ComputeFaceCenter2D :6370-6390, ComputeCellCenterVol2D :7058-7190): the
// edge p1->p2 has the normal (y2 - y1, x1 - x2)/length, its midpoint as
// centre and its length as area; cell centroid and area from the divergence
// theorem over the edges (boundary faces first), ghosts mirrored as in 3-D.
void Mesh::computeMetrics2D()
{
    const int nT = nTotal();
    xfc.assign(nFace, 0.0); yfc.assign(nFace, 0.0); zfc.assign(nFace, 0.0);
    xfn.assign(nFace, 0.0); yfn.assign(nFace, 0.0); zfn.assign(nFace, 0.0);
    area.assign(nFace, 0.0);
    xcc.assign(nT, 0.0); ycc.assign(nT, 0.0); zcc.assign(nT, 0.0); vol.assign(nT, 0.0);
    skew.assign(nCell, 0.0);
    for (int f = 0; f < nFace; ++f) {
        const int p1 = faceNodes[faceNodePtr[f]], p2 = faceNodes[faceNodePtr[f] + 1];
        xfn[f] = y[p2] - y[p1];
        yfn[f] = x[p1] - x[p2];
        zfn[f] = 0.0;
        area[f] = std::sqrt(xfn[f] * xfn[f] + yfn[f] * yfn[f] + zfn[f] * zfn[f]);
        const double oa = 1.0 / (area[f] + kSmall);
        xfn[f] *= oa; yfn[f] *= oa; zfn[f] *= oa;
    }
    for (int f = 0; f < nFace; ++f) {
        const int p1 = faceNodes[faceNodePtr[f]], p2 = faceNodes[faceNodePtr[f] + 1];
        xfc[f] = 0.5 * (x[p1] + x[p2]);
        yfc[f] = 0.5 * (y[p1] + y[p2]);
        zfc[f] = 0.5 * (z[p1] + z[p2]);
    }
    for (int f = 0; f < nFace; ++f) {
        const int le = left[f];
        const double dot = (xfc[f] * xfn[f] + yfc[f] * yfn[f] + zfc[f] * zfn[f]) * area[f];
        xcc[le] += xfc[f] * dot; ycc[le] += yfc[f] * dot; zcc[le] += zfc[f] * dot; vol[le] += dot;
        if (f >= nBFace) {
            const int re = right[f];
            xcc[re] -= xfc[f] * dot; ycc[re] -= yfc[f] * dot; zcc[re] -= zfc[f] * dot; vol[re] -= dot;
        }
    }
    for (int c = 0; c < nCell; ++c) {
        const double tmp = 1.0 / (1.5 * vol[c] + kSmall);
        xcc[c] *= tmp; ycc[c] *= tmp; zcc[c] *= tmp;
        vol[c] *= 0.5;
    }
    for (int f = 0; f < nBFace; ++f) {
        const int le = left[f], re = nCell + f;
        if (area[f] > kSmall) {
            const double tmp = 2.0 * ((xcc[le] - xfc[f]) * xfn[f] + (ycc[le] - yfc[f]) * yfn[f] +
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
}

void Mesh::computeMetrics()
{
    if (dim == 2) computeMetrics2D();
    else computeMetrics3D();
    computeSkewness();
}

void Mesh::computeMetrics3D()
{
    const int nT = nTotal();
    xfc.assign(nFace, 0.0); yfc.assign(nFace, 0.0); zfc.assign(nFace, 0.0);
    xfn.assign(nFace, 0.0); yfn.assign(nFace, 0.0); zfn.assign(nFace, 0.0);
    area.assign(nFace, 0.0);
    xcc.assign(nT, 0.0); ycc.assign(nT, 0.0); zcc.assign(nT, 0.0); vol.assign(nT, 0.0);
    skew.assign(nCell, 0.0);

    // Face centroid: node average, then four area-weighted refinements,
    // then normal and area from the triangle fan about the centroid.
    for (int f = 0; f < nFace; ++f) {
        const int b = faceNodePtr[f], nn = faceNodePtr[f + 1] - b;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int j = 0; j < nn; ++j) {
            const int n1 = faceNodes[b + j];
            sx += x[n1]; sy += y[n1]; sz += z[n1];
        }
        const double coef = 1.0 / nn;
        xfc[f] = sx * coef; yfc[f] = sy * coef; zfc[f] = sz * coef;
    }
    for (int cyc = 0; cyc <= 4; ++cyc) {
        const bool last = (cyc == 4);
        for (int f = 0; f < nFace; ++f) {
            const int b = faceNodePtr[f], nn = faceNodePtr[f + 1] - b;
            double xt = 0.0, yt = 0.0, zt = 0.0, at = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
            const double cx = xfc[f], cy = yfc[f], cz = zfc[f];
            for (int j = 0; j < nn; ++j) {
                const int p1 = faceNodes[b + j];
                const int p2 = faceNodes[b + (j + 1) % nn];
                const double dx1 = x[p1] - cx, dy1 = y[p1] - cy, dz1 = z[p1] - cz;
                const double dx2 = x[p2] - cx, dy2 = y[p2] - cy, dz2 = z[p2] - cz;
                const double anx = dy1 * dz2 - dy2 * dz1;
                const double any = dz1 * dx2 - dz2 * dx1;
                const double anz = dx1 * dy2 - dx2 * dy1;
                const double a = std::sqrt(anx * anx + any * any + anz * anz);
                const double x00 = x[p1] + x[p2] + cx;
                const double y00 = y[p1] + y[p2] + cy;
                const double z00 = z[p1] + z[p2] + cz;
                at += a;
                xt += a * x00; yt += a * y00; zt += a * z00;
                sx += anx; sy += any; sz += anz;
            }
            if (at > kSmall) {
                const double a = 1.0 / at / 3.0;
                xfc[f] = xt * a; yfc[f] = yt * a; zfc[f] = zt * a;
            }
            if (last) {
                const double a = std::sqrt(sx * sx + sy * sy + sz * sz);
                area[f] = a * 0.5;
                const double oa = 1.0 / (a + kSmall);
                xfn[f] = sx * oa; yfn[f] = sy * oa; zfn[f] = sz * oa;
            }
        }
    }

    // Cell centroid: node average, then the pyramid-volume-weighted centroid
    // unless it falls outside a face plane (the reference's 'mark' test).
    for (int c = 0; c < nCell; ++c) {
        const int b = cellNodePtr[c], nn = cellNodePtr[c + 1] - b;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int j = 0; j < nn; ++j) {
            const int n2 = cellNodes[b + j];
            sx += x[n2]; sy += y[n2]; sz += z[n2];
        }
        xcc[c] = sx / nn; ycc[c] = sy / nn; zcc[c] = sz / nn;
    }
    std::vector<double> xt(nCell, 0.0), yt(nCell, 0.0), zt(nCell, 0.0);
    for (int f = 0; f < nFace; ++f) {
        const int b = faceNodePtr[f], nn = faceNodePtr[f + 1] - b;
        const int le = left[f], re = right[f];
        const bool interior = f >= nBFace;
        for (int j = 0; j < nn; ++j) {
            const int p2 = faceNodes[b + j];
            const int p3 = faceNodes[b + (j + 1) % nn];
            const double x21 = x[p2] - xfc[f], y21 = y[p2] - yfc[f], z21 = z[p2] - zfc[f];
            const double x31 = x[p3] - xfc[f], y31 = y[p3] - yfc[f], z31 = z[p3] - zfc[f];
            const double nx = y21 * z31 - y31 * z21;
            const double ny = z21 * x31 - z31 * x21;
            const double nz = x21 * y31 - x31 * y21;
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
        const double tmp = 1.0 / (4.0 * vol[c] + kSmall);
        xt[c] *= tmp; yt[c] *= tmp; zt[c] *= tmp;
        vol[c] /= 6.0;
    }
    std::vector<char> mark(nCell, 1);
    for (int f = 0; f < nFace; ++f) {
        const int le = left[f];
        double dx = xt[le] - xfc[f], dy = yt[le] - yfc[f], dz = zt[le] - zfc[f];
        double tmp = dx * xfn[f] + dy * yfn[f] + dz * zfn[f];
        if (tmp > 0.0) mark[le] = 0;
        if (f >= nBFace) {
            const int re = right[f];
            dx = xt[re] - xfc[f]; dy = yt[re] - yfc[f]; dz = zt[re] - zfc[f];
            tmp = dx * xfn[f] + dy * yfn[f] + dz * zfn[f];
            if (tmp < 0.0) mark[re] = 0;
        }
    }
    for (int c = 0; c < nCell; ++c)
        if (mark[c]) { xcc[c] = xt[c]; ycc[c] = yt[c]; zcc[c] = zt[c]; }
    for (int f = 0; f < nBFace; ++f) {
        const int le = left[f], re = nCell + f;
        if (area[f] > kSmall) {
            const double tmp = 2.0 * ((xcc[le] - xfc[f]) * xfn[f] + (ycc[le] - yfc[f]) * yfn[f] +
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
}

// Skewness: minimum complement angle between the face normal and the
// face-to-centroid lines, interior faces first, then boundary faces; and the
// mean volume.
void Mesh::computeSkewness()
{
    skew.assign(nCell, 180.0);
    for (int f = nBFace; f < nFace; ++f) {
        if (area[f] <= kTiny) continue;
        const int le = left[f], re = right[f];
        const double x1 = xfc[f] - xcc[le], y1 = yfc[f] - ycc[le], z1 = zfc[f] - zcc[le];
        const double d1 = std::sqrt(x1 * x1 + y1 * y1 + z1 * z1);
        double p1 = (xfn[f] * x1 + yfn[f] * y1 + zfn[f] * z1) / (d1 + kTiny);
        if (p1 > 1.0) p1 = 1.0;
        if (p1 < -1.0) p1 = -1.0;
        const double x2 = xcc[re] - xfc[f], y2 = ycc[re] - yfc[f], z2 = zcc[re] - zfc[f];
        const double d2 = std::sqrt(x2 * x2 + y2 * y2 + z2 * z2);
        double p2 = (xfn[f] * x2 + yfn[f] * y2 + zfn[f] * z2) / (d2 + kTiny);
        if (p2 > 1.0) p2 = 1.0;
        if (p2 < -1.0) p2 = -1.0;
        const double angle = dmin(std::asin(p1), std::asin(p2)) * 180.0 / kPi;
        skew[le] = dmin(skew[le], angle);
        skew[re] = dmin(skew[re], angle);
    }
    for (int f = 0; f < nBFace; ++f) {
        if (area[f] <= kTiny) continue;
        const int le = left[f];
        const double x1 = xfc[f] - xcc[le], y1 = yfc[f] - ycc[le], z1 = zfc[f] - zcc[le];
        const double d1 = std::sqrt(x1 * x1 + y1 * y1 + z1 * z1);
        double p1 = (xfn[f] * x1 + yfn[f] * y1 + zfn[f] * z1) / (d1 + kTiny);
        if (p1 > 1.0) p1 = 1.0;
        if (p1 < -1.0) p1 = -1.0;
        const double angle = std::asin(p1) * 180.0 / kPi;
        skew[le] = dmin(skew[le], angle);
    }

    double tv = 0.0;
    for (int c = 0; c < nCell; ++c) tv += vol[c];
    avgVol = tv / nCell;
}

namespace {

// Local faces of each element kind (VTK node order), each listed cyclically.
const int kTetFaces[4][4] = {{0, 1, 2, -1}, {0, 1, 3, -1}, {1, 2, 3, -1}, {0, 2, 3, -1}};
const int kPyrFaces[5][4] = {{0, 1, 2, 3}, {0, 1, 4, -1}, {1, 2, 4, -1}, {2, 3, 4, -1}, {3, 0, 4, -1}};
const int kPriFaces[5][4] = {{0, 1, 2, -1}, {3, 4, 5, -1}, {0, 1, 4, 3}, {1, 2, 5, 4}, {2, 0, 3, 5}};
const int kHexFaces[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4},
                             {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}};

int nLocalFaces(int kind) { return kind == TET ? 4 : (kind == HEX ? 6 : 5); }
const int *localFace(int kind, int i)
{
    switch (kind) {
    case TET: return kTetFaces[i];
    case PYRAMID: return kPyrFaces[i];
    case PRISM: return kPriFaces[i];
    default: return kHexFaces[i];
    }
}

struct FaceRef {
    int key[4];  // sorted node ids, -1 padded (sorts first)
    int cell;
    int local;
};

bool keyLess(const FaceRef &a, const FaceRef &b)
{
    for (int i = 0; i < 4; ++i)
        if (a.key[i] != b.key[i]) return a.key[i] < b.key[i];
    return a.cell < b.cell;
}
bool keyEq(const FaceRef &a, const FaceRef &b)
{
    return a.key[0] == b.key[0] && a.key[1] == b.key[1] && a.key[2] == b.key[2] && a.key[3] == b.key[3];
}

}  // namespace

bool buildMesh(const ElementMesh &em, BcClassifier classify, const void *user, Mesh &out,
               std::string &err)
{
    const int nCell = (int)em.kind.size();
    const int nNode = (int)em.x.size();
    if (nCell <= 0 || (int)em.cellNodePtr.size() != nCell + 1) {
        err = "buildMesh: inconsistent cell arrays";
        return false;
    }
    std::vector<FaceRef> refs;
    refs.reserve((size_t)nCell * 6);
    for (int c = 0; c < nCell; ++c) {
        const int *cn = &em.cellNodes[em.cellNodePtr[c]];
        for (int lf = 0; lf < nLocalFaces(em.kind[c]); ++lf) {
            const int *loc = localFace(em.kind[c], lf);
            FaceRef r;
            for (int j = 0; j < 4; ++j) r.key[j] = (loc[j] >= 0) ? cn[loc[j]] : -1;
            std::sort(r.key, r.key + 4);
            r.cell = c;
            r.local = lf;
            refs.push_back(r);
        }
    }
    std::sort(refs.begin(), refs.end(), keyLess);

    struct Face { int le, re, local, localRe, bc; };
    std::vector<Face> bfaces, ifaces;
    for (size_t i = 0; i < refs.size();) {
        if (i + 1 < refs.size() && keyEq(refs[i], refs[i + 1])) {
            if (i + 2 < refs.size() && keyEq(refs[i], refs[i + 2])) {
                err = "buildMesh: face shared by more than two cells";
                return false;
            }
            ifaces.push_back({refs[i].cell, refs[i + 1].cell, refs[i].local, refs[i + 1].local, 0});
            i += 2;
        } else {
            bfaces.push_back({refs[i].cell, -1, refs[i].local, -1, 0});
            i += 1;
        }
    }

    // Face node lists taken from the owner's local face, then oriented so the
    // area vector points away from the owner's node-average centroid.
    auto faceNodesOf = [&](const Face &fc, std::vector<int> &nodes) {
        const int *cn = &em.cellNodes[em.cellNodePtr[fc.le]];
        const int *loc = localFace(em.kind[fc.le], fc.local);
        nodes.clear();
        for (int j = 0; j < 4; ++j)
            if (loc[j] >= 0) nodes.push_back(cn[loc[j]]);
    };
    auto centroid = [&](int c, double &cx, double &cy, double &cz) {
        const int b = em.cellNodePtr[c], nn = em.cellNodePtr[c + 1] - b;
        cx = cy = cz = 0.0;
        for (int j = 0; j < nn; ++j) {
            const int n = em.cellNodes[b + j];
            cx += em.x[n]; cy += em.y[n]; cz += em.z[n];
        }
        cx /= nn; cy /= nn; cz /= nn;
    };
    auto orient = [&](int owner, std::vector<int> &nodes, double &fx, double &fy, double &fz) {
        const int nn = (int)nodes.size();
        fx = fy = fz = 0.0;
        for (int j = 0; j < nn; ++j) { fx += em.x[nodes[j]]; fy += em.y[nodes[j]]; fz += em.z[nodes[j]]; }
        fx /= nn; fy /= nn; fz /= nn;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int j = 0; j < nn; ++j) {
            const int p1 = nodes[j], p2 = nodes[(j + 1) % nn];
            const double dx1 = em.x[p1] - fx, dy1 = em.y[p1] - fy, dz1 = em.z[p1] - fz;
            const double dx2 = em.x[p2] - fx, dy2 = em.y[p2] - fy, dz2 = em.z[p2] - fz;
            sx += dy1 * dz2 - dy2 * dz1;
            sy += dz1 * dx2 - dz2 * dx1;
            sz += dx1 * dy2 - dx2 * dy1;
        }
        double cx, cy, cz;
        centroid(owner, cx, cy, cz);
        if (sx * (fx - cx) + sy * (fy - cy) + sz * (fz - cz) < 0.0) std::reverse(nodes.begin(), nodes.end());
    };

    std::vector<int> nodes;
    for (Face &bf : bfaces) {
        faceNodesOf(bf, nodes);
        double fx, fy, fz;
        orient(bf.le, nodes, fx, fy, fz);
        bf.bc = classify(fx, fy, fz, user);
    }
    std::stable_sort(bfaces.begin(), bfaces.end(), [](const Face &a, const Face &b) {
        if (a.bc != b.bc) return a.bc < b.bc;
        return a.le < b.le;
    });
    for (Face &f : ifaces)
        if (f.le > f.re) {  // owner = lower-numbered cell
            std::swap(f.le, f.re);
            std::swap(f.local, f.localRe);
        }
    std::sort(ifaces.begin(), ifaces.end(), [](const Face &a, const Face &b) {
        if (a.le != b.le) return a.le < b.le;
        return a.re < b.re;
    });

    Mesh &m = out;
    m = Mesh();
    m.nNode = nNode;
    m.nCell = nCell;
    m.nBFace = (int)bfaces.size();
    m.nFace = m.nBFace + (int)ifaces.size();
    m.x = em.x; m.y = em.y; m.z = em.z;
    m.cellNodePtr = em.cellNodePtr;
    m.cellNodes = em.cellNodes;
    m.faceNodePtr.assign(1, 0);
    m.left.resize(m.nFace);
    m.right.resize(m.nFace);
    m.bcType.resize(m.nBFace);
    for (int f = 0; f < m.nFace; ++f) {
        const bool bnd = f < m.nBFace;
        Face fc = bnd ? bfaces[f] : ifaces[f - m.nBFace];
        faceNodesOf(fc, nodes);
        double fx, fy, fz;
        orient(fc.le, nodes, fx, fy, fz);
        for (int n : nodes) m.faceNodes.push_back(n);
        m.faceNodePtr.push_back((int)m.faceNodes.size());
        m.left[f] = fc.le;
        m.right[f] = bnd ? nCell + f : fc.re;
        if (bnd) m.bcType[f] = fc.bc;
    }
    m.computeMetrics();
    for (int c = 0; c < nCell; ++c) {
        if (!(m.vol[c] > 0.0)) {
            err = "buildMesh: non-positive cell volume";
            return false;
        }
    }
    return true;
}

namespace {
struct BumpGeom {
    double x0, x1, ytop, zmax;
};
int classifyBump(double xc, double yc, double zc, const void *user)
{
    const BumpGeom *g = (const BumpGeom *)user;
    const double tol = 1e-9;
    if (std::fabs(xc - g->x0) < tol || std::fabs(xc - g->x1) < tol) return 4;  // FARFIELD
    if (std::fabs(zc) < tol || std::fabs(zc - g->zmax) < tol) return 3;        // SYMMETRY
    if (std::fabs(yc - g->ytop) < tol) return 3;                               // SYMMETRY (top)
    return 2;                                                                  // SOLID_SURFACE
}
}  // namespace

bool makeBumpHex(int nx, int ny, int nz, double bump, Mesh &out, std::string &err)
{
    if (nx < 1 || ny < 1 || nz < 1) {
        err = "makeBumpHex: bad dimensions";
        return false;
    }
    const double x0 = -1.5, x1 = 1.5, ytop = 0.8, zmax = 0.2;
    ElementMesh em;
    const int NX = nx + 1, NY = ny + 1, NZ = nz + 1;
    em.x.resize((size_t)NX * NY * NZ);
    em.y.resize(em.x.size());
    em.z.resize(em.x.size());
    for (int k = 0; k < NZ; ++k)
        for (int j = 0; j < NY; ++j)
            for (int i = 0; i < NX; ++i) {
                const size_t n = (size_t)i + (size_t)NX * ((size_t)j + (size_t)NY * k);
                const double xx = x0 + (x1 - x0) * i / nx;
                const double yw = bump * std::exp(-25.0 * xx * xx);
                em.x[n] = xx;
                em.y[n] = yw + (ytop - yw) * j / ny;
                em.z[n] = zmax * k / nz;
            }
    const size_t nC = (size_t)nx * ny * nz;
    em.kind.assign(nC, HEX);
    em.cellNodePtr.resize(nC + 1);
    em.cellNodes.resize(nC * 8);
    auto nid = [&](int i, int j, int k) { return i + NX * (j + NY * k); };
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const size_t c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                int *cn = &em.cellNodes[c * 8];
                cn[0] = nid(i, j, k);         cn[1] = nid(i + 1, j, k);
                cn[2] = nid(i + 1, j + 1, k); cn[3] = nid(i, j + 1, k);
                cn[4] = nid(i, j, k + 1);     cn[5] = nid(i + 1, j, k + 1);
                cn[6] = nid(i + 1, j + 1, k + 1); cn[7] = nid(i, j + 1, k + 1);
                em.cellNodePtr[c] = (int)(c * 8);
            }
    em.cellNodePtr[nC] = (int)(nC * 8);
    BumpGeom g{x0, x1, ytop, zmax};
    return buildMesh(em, classifyBump, &g, out, err);
}

bool makeBumpHybrid(int nx, int nz, int nPrism, double h1, int nyTet, double jitter, unsigned seed,
                    double bump, Mesh &out, std::string &err)
{
    if (nx < 1 || nz < 1 || nPrism < 1 || nyTet < 1 || !(h1 > 0.0) || jitter < 0.0 || jitter > 0.3) {
        err = "makeBumpHybrid: bad arguments";
        return false;
    }
    const double x0 = -1.5, x1 = 1.5, ytop = 0.8, zmax = 0.2;
    const int NX = nx + 1, NZ = nz + 1;
    // growth ratio so that the last prism layer is close to the tet row height
    auto yw = [&](double xx) { return bump * std::exp(-25.0 * xx * xx); };
    double r = 1.2;
    {
        // solve h1 * r^(nPrism-1) = (ytop - thickness(r)) / nyTet by bisection
        double lo = 1.0 + 1e-9, hi = 3.0;
        for (int it = 0; it < 200; ++it) {
            const double mid = 0.5 * (lo + hi);
            const double thick = h1 * (std::pow(mid, nPrism) - 1.0) / (mid - 1.0);
            const double last = h1 * std::pow(mid, nPrism - 1);
            const double target = (ytop - thick) / nyTet;
            if (last < target) lo = mid; else hi = mid;
        }
        r = 0.5 * (lo + hi);
    }
    std::vector<double> d(nPrism + 1, 0.0);
    for (int l = 1; l <= nPrism; ++l) d[l] = d[l - 1] + h1 * std::pow(r, l - 1);
    if (!(d[nPrism] < 0.5 * (ytop - bump))) {
        err = "makeBumpHybrid: prism layers too thick";
        return false;
    }
    ElementMesh em;
    auto nPrismNode = [&](int l, int i, int k) { return (l * NZ + k) * NX + i; };
    const int prismNodes = (nPrism + 1) * NX * NZ;
    auto nTetNode = [&](int j, int i, int k) {
        return j == 0 ? nPrismNode(nPrism, i, k) : prismNodes + ((j - 1) * NZ + k) * NX + i;
    };
    const size_t nNode = (size_t)prismNodes + (size_t)nyTet * NX * NZ;
    em.x.resize(nNode); em.y.resize(nNode); em.z.resize(nNode);
    for (int l = 0; l <= nPrism; ++l)
        for (int k = 0; k < NZ; ++k)
            for (int i = 0; i < NX; ++i) {
                const int n = nPrismNode(l, i, k);
                const double xx = x0 + (x1 - x0) * i / nx;
                em.x[n] = xx; em.y[n] = yw(xx) + d[l]; em.z[n] = zmax * k / nz;
            }
    uint64_t st = 0x9E3779B97F4A7C15ull ^ (uint64_t)seed;
    auto urand = [&]() {  // xorshift64*, uniform in [-1, 1)
        st ^= st >> 12; st ^= st << 25; st ^= st >> 27;
        return (double)((st * 2685821657736338717ull) >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    };
    const double dx = (x1 - x0) / nx, dz = zmax / nz;
    for (int j = 1; j <= nyTet; ++j)
        for (int k = 0; k < NZ; ++k)
            for (int i = 0; i < NX; ++i) {
                const int n = nTetNode(j, i, k);
                double xx = x0 + (x1 - x0) * i / nx;
                const double yb = yw(xx) + d[nPrism];
                const double dy = (ytop - yb) / nyTet;
                double yy = yb + (ytop - yb) * j / nyTet;
                double zz = zmax * k / nz;
                const double jx = urand(), jy = urand(), jz = urand();
                if (j < nyTet && i > 0 && i < nx && k > 0 && k < nz) {
                    xx += jitter * dx * jx; yy += jitter * dy * jy; zz += jitter * dz * jz;
                }
                em.x[n] = xx; em.y[n] = yy; em.z[n] = zz;
            }
    // cells, numbered column by column (k outer, then i), from the wall up
    em.cellNodePtr.assign(1, 0);
    auto addCell = [&](int kind, std::initializer_list<int> nodes) {
        em.kind.push_back(kind);
        for (int n : nodes) em.cellNodes.push_back(n);
        em.cellNodePtr.push_back((int)em.cellNodes.size());
    };
    for (int k = 0; k < nz; ++k)
        for (int i = 0; i < nx; ++i) {
            for (int l = 0; l < nPrism; ++l) {
                // wall quad split along (i,k)-(i+1,k+1), matching the tet split above
                const int a0 = nPrismNode(l, i, k), b0 = nPrismNode(l, i + 1, k);
                const int c0 = nPrismNode(l, i + 1, k + 1), e0 = nPrismNode(l, i, k + 1);
                const int a1 = nPrismNode(l + 1, i, k), b1 = nPrismNode(l + 1, i + 1, k);
                const int c1 = nPrismNode(l + 1, i + 1, k + 1), e1 = nPrismNode(l + 1, i, k + 1);
                addCell(PRISM, {a0, b0, c0, a1, b1, c1});
                addCell(PRISM, {a0, c0, e0, a1, c1, e1});
            }
            for (int j = 0; j < nyTet; ++j) {
                // lattice cell corners v[dx][dy][dz]; six tetrahedra along the
                // (0,0,0)-(1,1,1) diagonal, one per ordering of the axes
                int v[2][2][2];
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b)
                        for (int c = 0; c < 2; ++c) v[a][b][c] = nTetNode(j + b, i + a, k + c);
                static const int perms[6][3] = {{0, 2, 1}, {2, 0, 1}, {0, 1, 2}, {1, 0, 2}, {1, 2, 0}, {2, 1, 0}};
                for (const auto &pm : perms) {
                    int p[3] = {0, 0, 0};
                    int nodes[4];
                    nodes[0] = v[0][0][0];
                    for (int s2 = 0; s2 < 3; ++s2) {
                        p[pm[s2]] = 1;
                        nodes[s2 + 1] = v[p[0]][p[1]][p[2]];
                    }
                    addCell(TET, {nodes[0], nodes[1], nodes[2], nodes[3]});
                }
            }
        }
    BumpGeom g{x0, x1, ytop, zmax};
    return buildMesh(em, classifyBump, &g, out, err);
}

namespace {
// Planform of the swept wing: root chord cr at z = 0 (a symmetry plane),
// linear taper to the tip at z = b, quarter-chord line swept by tan25.
struct WingGeom {
    double b, cr, taper, tanLE, far;
    double zc(double z) const { return z < b ? z : b; }
    double xLE(double z) const { return tanLE * zc(z); }
    double chord(double z) const { return cr * (1.0 - (1.0 - taper) * zc(z) / b); }
};

int classifyWing(double xc, double yc, double zc, const void *user)
{
    const WingGeom *g = (const WingGeom *)user;
    const double tol = 1e-9;
    if (std::fabs(zc) < tol) return 3;  // SYMMETRY: the root plane
    if (std::fabs(yc) < 0.5 && zc < g->b + tol && xc > g->xLE(zc) - tol && xc < g->xLE(zc) + g->chord(zc) + tol)
        return 2;  // SOLID_SURFACE
    return 4;      // FARFIELD
}

// Ratio r of a geometric sequence of n spacings, the first s1 * r, summing
// to `total` (bisection; 1 < r < 4).
double growthFor(double s1, int n, double total)
{
    double lo = 1.0 + 1e-12, hi = 4.0;
    for (int it = 0; it < 200; ++it) {
        const double r = 0.5 * (lo + hi);
        double sum = 0.0, sp = s1;
        for (int m = 0; m < n; ++m) { sp *= r; sum += sp; }
        if (sum < total) lo = r; else hi = r;
    }
    return 0.5 * (lo + hi);
}
}  // namespace

bool makeSweptWing(int nChord, int nSpan, int nPrism, double h1, int nTet, Mesh &out, std::string &err)
{
    if (nChord < 4 || nSpan < 2 || nPrism < 1 || nTet < 1 || !(h1 > 0.0)) {
        err = "makeSweptWing: bad arguments";
        return false;
    }
    // root chord 1, taper 0.5, aspect ratio 4 (semi-span (1 + taper) = 1.5),
    // quarter-chord sweep 25 degrees, farfield 20 root chords away
    WingGeom g;
    g.cr = 1.0;
    g.taper = 0.5;
    g.b = 0.25 * 4.0 * g.cr * (1.0 + g.taper);
    const double kPi = 3.141592653589793;
    g.tanLE = std::tan(25.0 * kPi / 180.0) + 0.25 * g.cr * (1.0 - g.taper) / g.b;
    g.far = 20.0;
    const double growth = 1.2, tau = 0.12, tipRound = 0.05 * g.b;
    const int nUp = std::max(2, (3 * nChord + 5) / 10), nDown = std::max(2, (6 * nChord + 5) / 10);
    const int nTip = std::max(2, (4 * nSpan + 5) / 10);
    const int NI = nUp + nChord + nDown, NK = nSpan + nTip, NJ = 2 * (nPrism + nTet), jw = nPrism + nTet;

    // Spanwise stations: half sine clustering towards the tip, half uniform,
    // then geometric.  The lattice lines run to the farfield (an H-type
    // grid), so a strong clustering at the wing's edges would leave slivers of
    // extreme aspect ratio in the tetrahedral rows far from the wing.
    std::vector<double> zs(NK + 1);
    for (int k = 0; k <= nSpan; ++k)
        zs[k] = g.b * (0.5 * std::sin(0.5 * kPi * k / nSpan) + 0.5 * (double)k / nSpan);
    {
        const double s1 = zs[nSpan] - zs[nSpan - 1];
        const double r = growthFor(s1, nTip, g.far - g.b);
        double sp = s1, z = g.b;
        for (int m = 1; m <= nTip; ++m) { sp *= r; z += sp; zs[nSpan + m] = z; }
        zs[NK] = g.far;
    }
    // wall-normal distances: nPrism layers growing by 1.2 from h1, then
    // nTet rows growing geometrically to the farfield
    std::vector<double> eta(nPrism + nTet + 1, 0.0);
    {
        double sp = h1;
        for (int l = 1; l <= nPrism; ++l) { eta[l] = eta[l - 1] + sp; sp *= growth; }
        const double last = eta[nPrism] - eta[nPrism - 1];
        if (!(eta[nPrism] < 0.1 * g.far)) {
            err = "makeSweptWing: prism layers too thick";
            return false;
        }
        const double r = growthFor(last, nTet, g.far - eta[nPrism]);
        sp = last;
        for (int l = nPrism + 1; l <= nPrism + nTet; ++l) { sp *= r; eta[l] = eta[l - 1] + sp; }
        eta[nPrism + nTet] = g.far;
    }
    // chordwise stations per span station: upstream geometric, cosine on the
    // chord, downstream geometric; xi[i] is the chord fraction (-1 off the wing)
    std::vector<double> xs((size_t)(NI + 1) * (NK + 1)), xiOf(NI + 1, -1.0);
    // chord fractions: half cosine clustering at both edges, half uniform
    for (int i = 0; i <= nChord; ++i)
        xiOf[nUp + i] = 0.25 * (1.0 - std::cos(kPi * i / nChord)) + 0.5 * (double)i / nChord;
    for (int k = 0; k <= NK; ++k) {
        const double c = g.chord(zs[k]), xle = g.xLE(zs[k]);
        double *row = &xs[(size_t)k * (NI + 1)];
        for (int i = 0; i <= nChord; ++i) row[nUp + i] = xle + c * xiOf[nUp + i];
        const double s1 = c * xiOf[nUp + 1];
        const double r1 = growthFor(s1, nUp - 1, g.far - s1);
        double sp = s1, d = s1;
        for (int m = 2; m <= nUp; ++m) { sp *= r1; d += sp; row[nUp - m] = xle - d; }
        row[nUp - 1] = xle - s1;
        row[0] = xle - g.far;
        const double s2 = c * (1.0 - xiOf[nUp + nChord - 1]);
        const double r2 = growthFor(s2, nDown - 1, g.far - s2);
        sp = s2; d = s2;
        row[nUp + nChord + 1] = xle + c + s2;
        for (int m = 2; m <= nDown; ++m) { sp *= r2; d += sp; row[nUp + nChord + m] = xle + c + d; }
        row[NI] = xle + c + g.far;
    }
    // half-thickness: cubic section t/c = 0.12 (maximum at 1/3 chord,
    // t = 6.75 * 0.06 * xi (1 - xi)^2), closed parabolically over the last
    // 5% of the span so that the surface ends in the tip edge
    auto thickness = [&](int i, int k) {
        const double xi = xiOf[i], z = zs[k];
        if (!(xi > 0.0 && xi < 1.0) || !(z < g.b)) return 0.0;
        double t = g.chord(z) * 6.75 * (0.5 * tau) * xi * (1.0 - xi) * (1.0 - xi);
        if (z > g.b - tipRound) {
            const double s = (z - (g.b - tipRound)) / tipRound;
            t *= 1.0 - s * s;
        }
        return t;
    };

    // nodes: the (i, j, k) lattice; on the wing the j = jw plane has a
    // second (lower-surface) copy of every node where the thickness is > 0
    ElementMesh em;
    const size_t nBase = (size_t)(NI + 1) * (NJ + 1) * (NK + 1);
    auto base = [&](int i, int j, int k) { return (int)(((size_t)k * (NJ + 1) + j) * (NI + 1) + i); };
    std::vector<int> lower((size_t)(NI + 1) * (NK + 1));
    em.x.resize(nBase); em.y.resize(nBase); em.z.resize(nBase);
    for (int k = 0; k <= NK; ++k)
        for (int i = 0; i <= NI; ++i) {
            const double t = thickness(i, k), x = xs[(size_t)k * (NI + 1) + i];
            for (int j = 0; j <= NJ; ++j) {
                const int n = base(i, j, k);
                const double e = eta[j >= jw ? j - jw : jw - j];
                const double yy = e + t * (1.0 - e / g.far);
                em.x[n] = x;
                em.y[n] = j >= jw ? yy : -yy;
                em.z[n] = zs[k];
            }
            int lo = base(i, jw, k);
            if (t > 0.0) {
                lo = (int)em.x.size();
                em.x.push_back(x);
                em.y.push_back(-t);
                em.z.push_back(zs[k]);
            }
            lower[(size_t)k * (NI + 1) + i] = lo;
        }
    if (em.x.size() > (size_t)0x7fffffff) {
        err = "makeSweptWing: too many nodes";
        return false;
    }
    // cells column by column (k outer, then i), bottom to top; near the
    // slit plane two prisms per lattice cell, beyond it six tetrahedra along
    // the (0,0,0)-(1,1,1) diagonal (both split every shared face alike)
    em.cellNodePtr.assign(1, 0);
    const size_t nCells = (size_t)NI * NK * 2 * (2 * nPrism + 6 * nTet);
    em.kind.reserve(nCells);
    em.cellNodes.reserve(nCells * 5);
    auto addCell = [&](int kind, std::initializer_list<int> nodes) {
        em.kind.push_back(kind);
        for (int n : nodes) em.cellNodes.push_back(n);
        em.cellNodePtr.push_back((int)em.cellNodes.size());
    };
    for (int k = 0; k < NK; ++k)
        for (int i = 0; i < NI; ++i)
            for (int j = 0; j < NJ; ++j) {
                const bool below = j < jw;
                int v[2][2][2];
                for (int a = 0; a < 2; ++a)
                    for (int b2 = 0; b2 < 2; ++b2)
                        for (int c = 0; c < 2; ++c) {
                            const int jj = j + b2;
                            v[a][b2][c] = (jj == jw && below) ? lower[(size_t)(k + c) * (NI + 1) + i + a]
                                                               : base(i + a, jj, k + c);
                        }
                const int layer = below ? jw - 1 - j : j - jw;
                if (layer < nPrism) {
                    addCell(PRISM, {v[0][0][0], v[1][0][0], v[1][0][1], v[0][1][0], v[1][1][0], v[1][1][1]});
                    addCell(PRISM, {v[0][0][0], v[1][0][1], v[0][0][1], v[0][1][0], v[1][1][1], v[0][1][1]});
                } else {
                    static const int perms[6][3] = {{0, 2, 1}, {2, 0, 1}, {0, 1, 2}, {1, 0, 2}, {1, 2, 0}, {2, 1, 0}};
                    for (const auto &pm : perms) {
                        int p[3] = {0, 0, 0};
                        int nodes[4];
                        nodes[0] = v[0][0][0];
                        for (int s2 = 0; s2 < 3; ++s2) {
                            p[pm[s2]] = 1;
                            nodes[s2 + 1] = v[p[0]][p[1]][p[2]];
                        }
                        addCell(TET, {nodes[0], nodes[1], nodes[2], nodes[3]});
                    }
                }
            }
    return buildMesh(em, classifyWing, &g, out, err);
}

bool renumberCells(const Mesh &in, const std::vector<int> &perm, Mesh &out, std::string &err)
{
    const int nC = in.nCell;
    if ((int)perm.size() != nC) {
        err = "renumberCells: permutation size";
        return false;
    }
    std::vector<int> inv(nC, -1);
    for (int i = 0; i < nC; ++i) {
        if (perm[i] < 0 || perm[i] >= nC || inv[perm[i]] >= 0) {
            err = "renumberCells: not a permutation";
            return false;
        }
        inv[perm[i]] = i;
    }
    struct F { int le, re, bc, old; bool flip; };
    std::vector<F> bf, inf;
    for (int f = 0; f < in.nFace; ++f) {
        if (f < in.nBFace) {
            bf.push_back({inv[in.left[f]], -1, in.bcType[f], f, false});
        } else {
            int a = inv[in.left[f]], b = inv[in.right[f]];
            const bool flip = a > b;
            if (flip) std::swap(a, b);
            inf.push_back({a, b, 0, f, flip});
        }
    }
    std::stable_sort(bf.begin(), bf.end(), [](const F &x, const F &y) {
        if (x.bc != y.bc) return x.bc < y.bc;
        return x.le < y.le;
    });
    std::sort(inf.begin(), inf.end(), [](const F &x, const F &y) {
        if (x.le != y.le) return x.le < y.le;
        return x.re < y.re;
    });
    Mesh m;
    m.nNode = in.nNode; m.nCell = nC; m.nBFace = in.nBFace; m.nFace = in.nFace;
    m.dim = in.dim;
    m.x = in.x; m.y = in.y; m.z = in.z;
    m.cellNodePtr.assign(1, 0);
    for (int i = 0; i < nC; ++i) {
        const int o = perm[i];
        for (int k = in.cellNodePtr[o]; k < in.cellNodePtr[o + 1]; ++k) m.cellNodes.push_back(in.cellNodes[k]);
        m.cellNodePtr.push_back((int)m.cellNodes.size());
    }
    m.faceNodePtr.assign(1, 0);
    m.left.resize(m.nFace); m.right.resize(m.nFace); m.bcType.resize(m.nBFace);
    for (int f = 0; f < m.nFace; ++f) {
        const bool bnd = f < m.nBFace;
        const F &fc = bnd ? bf[f] : inf[f - m.nBFace];
        const int b = in.faceNodePtr[fc.old], e = in.faceNodePtr[fc.old + 1];
        std::vector<int> nodes(in.faceNodes.begin() + b, in.faceNodes.begin() + e);
        if (fc.flip) std::reverse(nodes.begin(), nodes.end());
        for (int n : nodes) m.faceNodes.push_back(n);
        m.faceNodePtr.push_back((int)m.faceNodes.size());
        m.left[f] = fc.le;
        m.right[f] = bnd ? nC + f : fc.re;
        if (bnd) m.bcType[f] = fc.bc;
    }
    m.computeMetrics();
    out = std::move(m);
    return true;
}

std::vector<int> colourOrder(const Mesh &m, int &nColours)
{
    const int nC = m.nCell;
    std::vector<int> ptr(nC + 1, 0);
    for (int f = m.nBFace; f < m.nFace; ++f) { ptr[m.left[f] + 1]++; ptr[m.right[f] + 1]++; }
    for (int i = 0; i < nC; ++i) ptr[i + 1] += ptr[i];
    std::vector<int> nb(ptr[nC]), fill(ptr.begin(), ptr.end() - 1);
    for (int f = m.nBFace; f < m.nFace; ++f) {
        nb[fill[m.left[f]]++] = m.right[f];
        nb[fill[m.right[f]]++] = m.left[f];
    }
    std::vector<int> colour(nC, -1);
    nColours = 0;
    std::vector<int> used;
    for (int c = 0; c < nC; ++c) {
        used.assign(nColours + 1, 0);
        for (int k = ptr[c]; k < ptr[c + 1]; ++k)
            if (colour[nb[k]] >= 0) used[colour[nb[k]]] = 1;
        int col = 0;
        while (used[col]) ++col;
        colour[c] = col;
        nColours = std::max(nColours, col + 1);
    }
    std::vector<int> perm(nC);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return colour[a] < colour[b]; });
    return perm;
}

bool concatMeshes(const std::vector<const Mesh *> &parts, Mesh &out, std::string &err)
{
    if (parts.empty()) { err = "no meshes to concatenate"; return false; }
    Mesh m;
    long nNode = 0, nCell = 0, nFace = 0, nB = 0;
    for (const Mesh *p : parts) {
        if (!p->caseCellPtr.empty()) { err = "cannot concatenate batched meshes"; return false; }
        if (p->dim != parts[0]->dim) { err = "cannot concatenate 2-D and 3-D meshes"; return false; }
        nNode += p->nNode; nCell += p->nCell; nFace += p->nFace; nB += p->nBFace;
    }
    if (nCell + nB > 2147483000L || parts.size() > 65535) { err = "batch too large"; return false; }
    m.nNode = (int)nNode; m.nCell = (int)nCell; m.nFace = (int)nFace; m.nBFace = (int)nB;
    m.dim = parts[0]->dim;
    const int nT = m.nCell + m.nBFace;
    m.left.resize(nFace); m.right.resize(nFace); m.bcType.resize(nB);
    m.faceNodePtr.assign(nFace + 1, 0);
    m.cellNodePtr.assign(m.nCell + 1, 0);
    m.xfc.resize(nFace); m.yfc.resize(nFace); m.zfc.resize(nFace);
    m.xfn.resize(nFace); m.yfn.resize(nFace); m.zfn.resize(nFace); m.area.resize(nFace);
    m.xcc.resize(nT); m.ycc.resize(nT); m.zcc.resize(nT); m.vol.resize(nT);
    m.skew.resize(m.nCell);
    // face index in the union of each part's face f
    std::vector<std::vector<int>> fmap(parts.size());
    int bOff = 0, iOff = m.nBFace, cOff = 0, nOff = 0;
    m.caseCellPtr.push_back(0);
    m.caseBFacePtr.push_back(0);
    for (size_t k = 0; k < parts.size(); ++k) {
        const Mesh &p = *parts[k];
        fmap[k].resize(p.nFace);
        for (int f = 0; f < p.nFace; ++f) fmap[k][f] = f < p.nBFace ? bOff + f : iOff + (f - p.nBFace);
        bOff += p.nBFace; iOff += p.nFace - p.nBFace;
        cOff += p.nCell; nOff += p.nNode;
        m.caseCellPtr.push_back(cOff);
        m.caseBFacePtr.push_back(bOff);
        m.caseAvgVol.push_back(p.avgVol);
    }
    // face node counts, then nodes
    for (size_t k = 0; k < parts.size(); ++k) {
        const Mesh &p = *parts[k];
        for (int f = 0; f < p.nFace; ++f) m.faceNodePtr[fmap[k][f] + 1] = p.faceNodePtr[f + 1] - p.faceNodePtr[f];
    }
    for (long f = 0; f < nFace; ++f) m.faceNodePtr[f + 1] += m.faceNodePtr[f];
    m.faceNodes.resize(m.faceNodePtr[nFace]);
    cOff = 0; nOff = 0;
    int cnOff = 0;
    for (size_t k = 0; k < parts.size(); ++k) {
        const Mesh &p = *parts[k];
        for (int i = 0; i < p.nNode; ++i) { m.x.push_back(p.x[i]); m.y.push_back(p.y[i]); m.z.push_back(p.z[i]); }
        for (int f = 0; f < p.nFace; ++f) {
            const int g = fmap[k][f];
            int o = m.faceNodePtr[g];
            for (int j = p.faceNodePtr[f]; j < p.faceNodePtr[f + 1]; ++j) m.faceNodes[o++] = p.faceNodes[j] + nOff;
            m.left[g] = p.left[f] + cOff;
            m.right[g] = f < p.nBFace ? m.nCell + g : p.right[f] + cOff;
            if (f < p.nBFace) m.bcType[g] = p.bcType[f];
            m.xfc[g] = p.xfc[f]; m.yfc[g] = p.yfc[f]; m.zfc[g] = p.zfc[f];
            m.xfn[g] = p.xfn[f]; m.yfn[g] = p.yfn[f]; m.zfn[g] = p.zfn[f]; m.area[g] = p.area[f];
            if (f < p.nBFace) {  // ghost cell metrics
                const int gs = p.nCell + f, gd = m.nCell + g;
                m.xcc[gd] = p.xcc[gs]; m.ycc[gd] = p.ycc[gs]; m.zcc[gd] = p.zcc[gs]; m.vol[gd] = p.vol[gs];
            }
        }
        for (int c = 0; c < p.nCell; ++c) {
            m.xcc[cOff + c] = p.xcc[c]; m.ycc[cOff + c] = p.ycc[c]; m.zcc[cOff + c] = p.zcc[c];
            m.vol[cOff + c] = p.vol[c]; m.skew[cOff + c] = p.skew[c];
            m.cellNodePtr[cOff + c + 1] = p.cellNodePtr[c + 1] - p.cellNodePtr[c];
        }
        for (int j = 0; j < p.cellNodePtr[p.nCell]; ++j) m.cellNodes.push_back(p.cellNodes[j] + nOff);
        cOff += p.nCell; nOff += p.nNode; cnOff += p.cellNodePtr[p.nCell];
    }
    for (int c = 0; c < m.nCell; ++c) m.cellNodePtr[c + 1] += m.cellNodePtr[c];
    double vs = 0.0;
    for (int c = 0; c < m.nCell; ++c) vs += m.vol[c];
    m.avgVol = vs / m.nCell;
    out = std::move(m);
    return true;
}

}  // namespace phl
